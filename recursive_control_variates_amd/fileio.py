"""On-disk formats the reference reads and writes through Mitsuba (utils.py:8-14,
45-57; run_experiment.py:33), written from the published format
descriptions so that outputs of this build open in the same tools:

* OpenEXR, single-part scanline files with NO_COMPRESSION and FLOAT (32-bit)
  channels (``write_exr`` / ``read_exr``). Images are (H, W, C) float32;
  channel names default to R, G, B (A for a fourth, Y for one channel).
* PNG, 8-bit sRGB (``write_png``), the tone-mapped copy of an image that
  ``mi.util.convert_to_bitmap`` produces: linear values clamped to [0, 1],
  sRGB transfer curve, rounded to 8 bits.
* Mitsuba ``.vol`` grid volumes, version 3 with float32 data (``write_vol`` /
  ``read_vol``): grids are (D, H, W, C) float32 with an axis-aligned box.
* ``result.npz`` of an optimisation run (``save_result``).
"""
from __future__ import annotations

import os
import struct
import zlib
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

# --------------------------------------------------------------------- EXR
_EXR_MAGIC = 20000630
_EXR_FLOAT = 2


def _exr_attr(name: str, type_name: str, payload: bytes) -> bytes:
    return name.encode() + b"\0" + type_name.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def _default_channels(c: int) -> Sequence[str]:
    return {1: ("Y",), 3: ("R", "G", "B"), 4: ("R", "G", "B", "A")}.get(c, tuple(f"C{i}" for i in range(c)))


def write_exr(path: str, img, channels: Optional[Sequence[str]] = None) -> None:
    """Writes an (H, W) or (H, W, C) image as a scanline OpenEXR file with
    uncompressed 32-bit float channels (one scanline per block)."""
    a = np.asarray(img.detach().cpu() if hasattr(img, "detach") else img, dtype=np.float32)
    if a.ndim == 2:
        a = a[..., None]
    if a.ndim != 3:
        raise ValueError(f"write_exr: expected (H, W[, C]), got shape {a.shape}")
    h, w, c = a.shape
    names = list(channels) if channels is not None else list(_default_channels(c))
    if len(names) != c:
        raise ValueError("write_exr: one channel name per channel")
    order = sorted(range(c), key=lambda i: names[i])        # channels are stored sorted by name
    chlist = b"".join(names[i].encode() + b"\0" + struct.pack("<iB3xii", _EXR_FLOAT, 0, 1, 1) for i in order)
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = b"".join([
        struct.pack("<ii", _EXR_MAGIC, 2),
        _exr_attr("channels", "chlist", chlist + b"\0"),
        _exr_attr("compression", "compression", b"\0"),
        _exr_attr("dataWindow", "box2i", box),
        _exr_attr("displayWindow", "box2i", box),
        _exr_attr("lineOrder", "lineOrder", b"\0"),
        _exr_attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        _exr_attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)),
        _exr_attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
        b"\0",
    ])
    line_bytes = w * c * 4
    first = len(header) + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * (8 + line_bytes) for y in range(h)])
    planar = np.ascontiguousarray(a[:, :, order].transpose(0, 2, 1))    # (H, C, W): per line, channel by channel
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(header)
        f.write(offsets)
        for y in range(h):
            f.write(struct.pack("<ii", y, line_bytes))
            f.write(planar[y].astype("<f4").tobytes())


def read_exr(path: str) -> Tuple[np.ndarray, Sequence[str]]:
    """Reads a scanline OpenEXR file with uncompressed FLOAT channels written
    by ``write_exr`` or any writer using the same options. Returns the (H, W,
    C) float32 image (channels in the file's name order) and the names."""
    with open(path, "rb") as f:
        data = f.read()
    magic, version = struct.unpack_from("<ii", data, 0)
    if magic != _EXR_MAGIC:
        raise ValueError(f"{path}: not an OpenEXR file")
    if version & 0xFFFFFF00:
        raise ValueError(f"{path}: only single-part scanline files are supported")
    pos = 8
    attrs: Dict[str, Tuple[str, bytes]] = {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        end = data.index(b"\0", pos)
        tname = data[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        attrs[name] = (tname, data[pos:pos + size])
        pos += size
    pos += 1
    if attrs["compression"][1][0] != 0:
        raise ValueError(f"{path}: compressed EXR files are not supported")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    names, types = [], []
    cl = attrs["channels"][1]
    q = 0
    while cl[q] != 0:
        end = cl.index(b"\0", q)
        names.append(cl[q:end].decode())
        (ptype,) = struct.unpack_from("<i", cl, end + 1)
        types.append(ptype)
        q = end + 1 + 16
    if any(t != _EXR_FLOAT for t in types):
        raise ValueError(f"{path}: only FLOAT channels are supported")
    c = len(names)
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    out = np.empty((h, c, w), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        out[y - y0] = np.frombuffer(data, "<f4", c * w, off + 8).reshape(c, w)
    return out.transpose(0, 2, 1).copy(), names


def read_exr_rgb(path: str) -> np.ndarray:
    """An OpenEXR image as (H, W, 3) float32 in R, G, B order (a single Y
    channel is repeated)."""
    img, names = read_exr(path)
    if all(n in names for n in ("R", "G", "B")):
        return np.ascontiguousarray(img[..., [names.index("R"), names.index("G"), names.index("B")]])
    if names == ["Y"]:
        return np.repeat(img, 3, axis=2)
    raise ValueError(f"{path}: expected R, G, B or Y channels, found {names}")


# --------------------------------------------------------------------- PNG
def _srgb(x: np.ndarray) -> np.ndarray:
    x = np.clip(x, 0.0, 1.0)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)


def write_png(path: str, img) -> None:
    """8-bit sRGB PNG of a linear (H, W, 3) or (H, W) image."""
    a = np.asarray(img.detach().cpu() if hasattr(img, "detach") else img, dtype=np.float64)
    if a.ndim == 3 and a.shape[2] == 1:
        a = a[..., 0]
    color = 2 if a.ndim == 3 else 0
    u8 = np.round(_srgb(a) * 255.0).astype(np.uint8)
    h, w = u8.shape[:2]
    rows = u8.reshape(h, -1)
    raw = b"".join(b"\0" + rows[y].tobytes() for y in range(h))

    def chunk(tag: bytes, payload: bytes) -> bytes:
        return struct.pack(">I", len(payload)) + tag + payload + struct.pack(">I", zlib.crc32(tag + payload))

    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def save_img(img, output_dir: str, img_name: str, it: int, save_png: bool = False) -> str:
    """utils.py:8-14: ``output_dir/img_name/{it:04d}.exr`` (+ the PNG copy)."""
    filename = os.path.join(output_dir, img_name, f"{it:04d}.exr")
    write_exr(filename, img)
    if save_png:
        write_png(filename.replace(".exr", ".png"), img)
    return filename


# --------------------------------------------------------------------- VOL
def write_vol(path: str, grid, bbox_min=(0.0, 0.0, 0.0), bbox_max=(1.0, 1.0, 1.0)) -> None:
    """Mitsuba grid volume, version 3, float32 encoding: 'VOL', 3, encoding
    1, x/y/z resolution, channels, bounding box (min xyz, max xyz), then the
    data with x fastest, then y, then z, channels interleaved."""
    a = np.asarray(grid.detach().cpu() if hasattr(grid, "detach") else grid, dtype=np.float32)
    if a.ndim == 3:
        a = a[..., None]
    if a.ndim != 4:
        raise ValueError(f"write_vol: expected (D, H, W[, C]), got shape {a.shape}")
    d, h, w, c = a.shape
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]))
        f.write(struct.pack("<iiiii", 1, w, h, d, c))
        f.write(struct.pack("<6f", *bbox_min, *bbox_max))
        f.write(np.ascontiguousarray(a).astype("<f4").tobytes())


def read_vol(path: str) -> Tuple[np.ndarray, Tuple[float, ...], Tuple[float, ...]]:
    """Reads a float32 Mitsuba grid volume; returns ((D, H, W, C) float32,
    bbox min, bbox max)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:3] != b"VOL" or data[3] != 3:
        raise ValueError(f"{path}: not a version-3 Mitsuba volume")
    enc, w, h, d, c = struct.unpack_from("<iiiii", data, 4)
    if enc != 1:
        raise ValueError(f"{path}: only float32 volumes (encoding 1) are supported, got {enc}")
    box = struct.unpack_from("<6f", data, 24)
    arr = np.frombuffer(data, "<f4", d * h * w * c, 48).reshape(d, h, w, c).astype(np.float32)
    return arr, tuple(box[:3]), tuple(box[3:])


# -------------------------------------------------------------- run results
def save_result(output_dir: str, result: Dict) -> str:
    """run_experiment.py:33: the array entries of a run_opt result as
    ``output_dir/result.npz``."""
    os.makedirs(output_dir, exist_ok=True)
    arrays = {k: np.asarray(v) for k, v in result.items() if isinstance(v, np.ndarray)}
    path = os.path.join(output_dir, "result.npz")
    np.savez(path, **arrays)
    return path
