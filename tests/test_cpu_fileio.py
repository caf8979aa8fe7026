"""On-disk formats (recursive_control_variates_amd.fileio): OpenEXR bytes
against a file laid out by hand from the OpenEXR file-layout document, PNG
decoded by an independent decoder, Mitsuba .vol bytes against the volume
format's header description, and scene descriptions loading grids from
files. CPU only."""
import os
import struct

import numpy as np
import pytest

from recursive_control_variates_amd import fileio, scene_desc
from helpers import volume_box_scene


def _attr(name, tname, payload):
    return name.encode() + b"\x00" + tname.encode() + b"\x00" + struct.pack("<i", len(payload)) + payload


def test_exr_bytes_match_hand_layout(tmp_path):
    # 2 x 1 RGB image; pixels (1, 2, 3) and (4, 5, 6)
    img = np.array([[[1, 2, 3], [4, 5, 6]]], np.float32)
    p = tmp_path / "a.exr"
    fileio.write_exr(str(p), img)
    ch = b""
    for name in (b"B", b"G", b"R"):                     # channel list sorted by name
        ch += name + b"\x00" + struct.pack("<i", 2) + b"\x00\x00\x00\x00" + struct.pack("<ii", 1, 1)
    ch += b"\x00"
    header = (b"\x76\x2f\x31\x01" + b"\x02\x00\x00\x00"
              + _attr("channels", "chlist", ch)
              + _attr("compression", "compression", b"\x00")
              + _attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, 1, 0))
              + _attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, 1, 0))
              + _attr("lineOrder", "lineOrder", b"\x00")
              + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
              + b"\x00")
    offset = len(header) + 8
    line = struct.pack("<ii", 0, 24) + struct.pack("<6f", 3, 6, 2, 5, 1, 4)   # B plane, G plane, R plane
    assert p.read_bytes() == header + struct.pack("<Q", offset) + line


@pytest.mark.parametrize("c", [1, 3, 4])
def test_exr_round_trip(tmp_path, c):
    rng = np.random.default_rng(c)
    img = rng.normal(size=(7, 5, c)).astype(np.float32)
    img[0, 0, 0] = np.inf
    p = str(tmp_path / "b.exr")
    fileio.write_exr(p, img)
    back, names = fileio.read_exr(p)
    expect_names = {1: ["Y"], 3: ["B", "G", "R"], 4: ["A", "B", "G", "R"]}[c]
    assert names == expect_names
    order = {1: [0], 3: [2, 1, 0], 4: [3, 2, 1, 0]}[c]
    assert np.array_equal(back, img[..., order])


def test_png_decodes_to_srgb(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    lin = np.array([[[0.0, 0.5, 1.0], [0.001, 0.2, 2.0]],
                    [[0.04, 0.0031308, -1.0], [0.75, 0.9, 0.1]]], np.float32)
    p = str(tmp_path / "c.png")
    fileio.write_png(p, lin)
    got = np.asarray(Image.open(p).convert("RGB"))
    x = np.clip(lin.astype(np.float64), 0, 1)
    srgb = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    assert np.array_equal(got, np.round(srgb * 255).astype(np.uint8))


def test_vol_bytes_and_round_trip(tmp_path):
    g = np.arange(2 * 3 * 4 * 1, dtype=np.float32).reshape(2, 3, 4, 1)     # (z, y, x, c)
    p = str(tmp_path / "d.vol")
    fileio.write_vol(p, g, (0.0, 0.5, -1.0), (1.0, 1.5, 2.0))
    b = open(p, "rb").read()
    assert b[:4] == b"VOL\x03"
    assert struct.unpack_from("<5i", b, 4) == (1, 4, 3, 2, 1)               # float32, x, y, z res, channels
    assert struct.unpack_from("<6f", b, 24) == (0.0, 0.5, -1.0, 1.0, 1.5, 2.0)
    # data: x fastest, then y, then z
    vals = np.frombuffer(b, "<f4", 24, 48)
    assert vals[1] == g[0, 0, 1, 0] and vals[4] == g[0, 1, 0, 0] and vals[12] == g[1, 0, 0, 0]
    back, lo, hi = fileio.read_vol(p)
    assert np.array_equal(back, g) and lo == (0.0, 0.5, -1.0) and hi == (1.0, 1.5, 2.0)


def test_scene_loads_grid_and_bitmap_files(tmp_path):
    rng = np.random.default_rng(3)
    g = rng.uniform(0, 1, (4, 5, 6, 1)).astype(np.float32)
    vp = str(tmp_path / "sigma.vol")
    fileio.write_vol(vp, g, (0, 0, 0), (2, 2, 2))
    d_data = scene_desc.describe(volume_box_scene(g))
    d = volume_box_scene(g)
    d["box"]["interior"]["old"]["sigma_t"] = {"type": "gridvolume", "filename": vp}
    d["box"]["interior"]["new"] = d["box"]["interior"]["incoming"] = d["box"]["interior"]["old"]
    d_file = scene_desc.describe(d)
    assert np.array_equal(d_file.params[0].init, d_data.params[0].init)
    assert d_file.media[0].bbox_max == d_data.media[0].bbox_max == (1.0, 1.0, 1.0)   # placed by to_world
    d["box"]["interior"]["old"]["sigma_t"]["use_grid_bbox"] = True
    assert scene_desc.describe(d).media[0].bbox_max == (2.0, 2.0, 2.0)              # the file's box
    tex = rng.uniform(0, 1, (3, 4, 3)).astype(np.float32)
    ep = str(tmp_path / "tex.exr")
    fileio.write_exr(ep, tex)
    back = scene_desc._bitmap_data({"type": "bitmap", "filename": ep}, "t")
    assert np.array_equal(back, tex)                   # R, G, B by channel name


def test_save_result_npz(tmp_path):
    r = {"loss": np.ones(3, np.float32), "var": np.zeros(3, np.float32), "runtime": np.zeros((5, 2), np.float32),
         "final": {"x": 1}}
    path = fileio.save_result(str(tmp_path), r)
    with np.load(path) as z:
        assert sorted(z.files) == ["loss", "runtime", "var"]
