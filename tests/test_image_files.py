"""Image files (CPU): the EXR and PFM writers of the host library (sol::Image::save, reference
src/image.cpp:26-37 and src/formats/exr.cpp:48-100), read back by a reader written here from the
OpenEXR file layout (magic, version, attributes, offset table, scanline blocks)."""
import struct

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and version & 0xFF == 2 and version >> 8 == 0   # single-part scanline
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        tend = data.index(b"\0", end + 1)
        typ = data[end + 1:tend].decode()
        (size,) = struct.unpack_from("<i", data, tend + 1)
        attrs[name] = (typ, data[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    chans, raw = [], attrs["channels"][1]
    q = 0
    while raw[q] != 0:
        end = raw.index(b"\0", q)
        ptype, linear, xs, ys = struct.unpack_from("<iB3xii", raw, end + 1)
        chans.append((raw[q:end].decode(), ptype, xs, ys))
        q = end + 1 + 16
    assert attrs["compression"][1] == b"\0" and attrs["lineOrder"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    out = {name: np.zeros((h, w), np.float32) for name, *_ in chans}
    for y, off in enumerate(offsets):
        yy, size = struct.unpack_from("<ii", data, off)
        assert yy == y and size == 4 * w * len(chans)
        row = np.frombuffer(data, "<f4", count=w * len(chans), offset=off + 8).reshape(len(chans), w)
        for (name, ptype, xs, ys), r in zip(chans, row):
            assert ptype == 2 and xs == 1 and ys == 1      # FLOAT, full resolution
            out[name][y] = r
    return [c[0] for c in chans], out


def test_exr_holds_the_image_in_b_g_r_channels(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.standard_normal((3, 7, 11)).astype(np.float32)
    img[0, 0, 0], img[1, 6, 10], img[2, 3, 4] = np.inf, -0.0, 1e-40     # bit patterns survive
    for fmt, name in (("auto", "a.exr"), ("exr", "b.exr")):
        sol_amd.save_image(tmp_path / name, img, fmt)
        names, ch = read_exr(tmp_path / name)
        assert names == ["B", "G", "R"]
        for i, c in enumerate("RGB"):
            assert ch[c].tobytes() == img[i].tobytes()


@pytest.mark.parametrize("n,names", [(1, ["Channel_0"]), (2, ["Channel_0", "Channel_1"]),
                                     (4, ["A", "B", "G", "R"])])
def test_exr_channel_names_follow_sol(tmp_path, n, names):
    img = np.arange(n * 2 * 3, dtype=np.float32).reshape(n, 2, 3)
    sol_amd.save_image(tmp_path / "x.exr", img)
    got, ch = read_exr(tmp_path / "x.exr")
    assert got == names
    order = {1: ["Channel_0"], 2: ["Channel_0", "Channel_1"], 4: ["R", "G", "B", "A"]}[n]
    for i, c in enumerate(order):
        np.testing.assert_array_equal(ch[c], img[i])


def test_pfm_and_unsupported_formats(tmp_path):
    img = np.random.default_rng(1).random((3, 4, 5)).astype(np.float32)
    sol_amd.save_image(tmp_path / "x.pfm", img, "pfm")
    with open(tmp_path / "x.pfm", "rb") as f:
        assert f.readline() == b"PF\n" and f.readline().split() == [b"5", b"4"] and float(f.readline()) < 0
        rows = np.frombuffer(f.read(), "<f4").reshape(4, 5, 3)[::-1]
    np.testing.assert_array_equal(rows.transpose(2, 0, 1), img)
    for fmt in ("png", "jpeg", "tiff"):
        with pytest.raises(F.SolError):
            sol_amd.save_image(tmp_path / f"x.{fmt}", img, fmt)
    with pytest.raises(F.SolError):
        sol_amd.save_image(tmp_path / "y.pfm", img[:2], "pfm")
