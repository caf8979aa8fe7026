"""rcv_loss_noalloc against rcv_loss through the C ABI: the same loss and
gradient bit for bit, on two calls in a row (the second with a smaller
input), for both loss kinds."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import announce
from recursive_control_variates_amd import _native as N

pytestmark = pytest.mark.gpu


def _call(name, x, y, kind):
    loss = torch.zeros(1, device="cuda")
    grad = torch.empty_like(x)
    N.call(name, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), x.numel(), kind,
           C.c_void_p(loss.data_ptr()), C.c_void_p(grad.data_ptr()), C.c_void_p(N.stream_handle()))
    torch.cuda.synchronize()
    return loss.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("kind", [N.LOSS_L1, N.LOSS_L2])
def test_loss_noalloc_equals_loss(cuda, kind):
    announce("rcv_loss_noalloc vs rcv_loss", kind)
    rng = np.random.default_rng(kind)
    for n in (256 * 1024 * 3, 1000):       # second call: smaller input
        x = torch.from_numpy(rng.uniform(0, 1, n).astype(np.float32)).cuda()
        y = torch.from_numpy(rng.uniform(0, 1, n).astype(np.float32)).cuda()
        l0, g0 = _call("rcv_loss", x, y, kind)
        l1, g1 = _call("rcv_loss_noalloc", x, y, kind)
        assert np.array_equal(l0.view(np.uint32), l1.view(np.uint32)), (n, l0, l1)
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32)), n
        d = x.double().cpu().numpy() - y.double().cpu().numpy()
        expect = np.abs(d).mean() if kind == N.LOSS_L1 else (d * d).mean()
        assert abs(float(l1[0]) - expect) <= 1e-5 * expect


def test_loss_noalloc_rejects_bad_arguments(cuda):
    x = torch.zeros(10, device="cuda")
    loss = torch.zeros(1, device="cuda")
    with pytest.raises(N.RcvError):
        N.call("rcv_loss_noalloc", C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), 0, N.LOSS_L1,
               C.c_void_p(loss.data_ptr()), None, C.c_void_p(N.stream_handle()))
    with pytest.raises(N.RcvError):
        N.call("rcv_loss_noalloc", C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), 10, 7,
               C.c_void_p(loss.data_ptr()), None, C.c_void_p(N.stream_handle()))
