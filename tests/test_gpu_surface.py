"""GPU parity of the two-state surface integrator (K1): BVH traversal vs the
oracle's brute-force intersection on the Cornell box workload, closed forms
and the two-state invariants."""
import numpy as np
import pytest
import torch

import render_ref as R
from conftest import announce
from helpers import floor_light_scene, furnace_box_scene
import recursive_control_variates_amd as rcv
from recursive_control_variates_amd import scenes

pytestmark = pytest.mark.gpu
PATH = 0


def test_cornell_matches_oracle(cuda):
    announce("cornell vs oracle fp32")
    tex_old = scenes.texture_target(64, 0)
    tex_new = np.full((64, 64, 3), 0.5, np.float32)
    sc = rcv.load_dict(scenes.cornell_scene_dict(tex_old, film_res=32, spp=8))
    p = rcv.traverse(sc)
    p[scenes.TEXTURE_KEY] = torch.from_numpy(tex_new).cuda()
    p.update()
    integ = rcv.load_dict({"type": "twostateprb", "max_depth": 4})
    H, F = integ.render_twostates(sc, 0, 17, 8)
    Ho, Fo = R.render(sc.desc, PATH, 4, 5, 17, 8, params_old=[tex_old], params_new=[tex_new], precision=32)
    for g, o in ((H, Ho), (F, Fo)):
        g = g.cpu().numpy()
        close = np.abs(g - o) <= 1e-4 + 1e-3 * np.abs(o)
        assert close.mean() > 0.98, close.mean()
        assert abs(g.mean() - o.mean()) < 2e-3 * o.mean()


def test_cornell_equal_states_close(cuda):
    # f_old/pdf and the sampled weight differ only by rounding
    announce("cornell H ~ F")
    sc = rcv.load_dict(scenes.cornell_scene_dict(scenes.texture_target(), film_res=24, spp=4))
    integ = rcv.load_dict({"type": "twostateprb", "max_depth": 4})
    H, F = integ.render_twostates(sc, 0, 3, 4)
    np.testing.assert_allclose(H.cpu().numpy(), F.cpu().numpy(), rtol=1e-5, atol=1e-6)
    plain = rcv.load_dict({"type": "prb", "max_depth": 4})
    assert torch.equal(plain.render(sc, 0, 3, 4), F)


@pytest.mark.parametrize("max_depth", [1, 3])
def test_furnace_closed_form(cuda, max_depth):
    announce("furnace", max_depth)
    sc = rcv.load_dict(furnace_box_scene(albedo=0.5, w=16, h=16, spp=64))
    integ = rcv.load_dict({"type": "prb", "max_depth": max_depth, "rr_depth": 100})
    imgs = np.stack([integ.render(sc, 0, 10 + s, 64).cpu().numpy() for s in range(8)])
    expect = sum(0.5 ** k for k in range(max_depth))
    se = imgs.mean(axis=(1, 2, 3)).std(ddof=1) / np.sqrt(8) + 1e-9
    assert abs(imgs.mean() - expect) < 5 * se + 1e-6, (imgs.mean(), expect)


def test_direct_light_vs_oracle_fp64(cuda):
    announce("direct light vs fp64")
    sc = rcv.load_dict(floor_light_scene(w=12, h=12, spp=256))
    integ = rcv.load_dict({"type": "prb", "max_depth": 2})
    runs = np.stack([integ.render(sc, 0, 50 + s, 256).cpu().numpy() for s in range(8)])
    o, _, v, _ = R.render(sc.desc, PATH, 2, 5, 1, 8192, precision=64, with_variance=True)
    se = np.sqrt(runs.var(0, ddof=1) / 8 + v / 8192) + 1e-7
    z = (runs.mean(0) - o) / se
    assert np.mean(np.abs(z) > 3) < 0.02 and np.abs(z).max() < 6
