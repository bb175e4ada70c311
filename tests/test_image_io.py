"""Frame and flow files of the VBM3Ddenoising command line (CPU): PNG read
and write, JPEG read, .flo read, through libvbm3d_host.so's C entry
points, checked against PIL's codecs and numpy-written files.  Samples
load as their integer values (0-255, or 0-65535 for 16-bit PNG), planar
[c][y][x], with the reference's rule that three equal channels collapse to
one (src/Utilities/LibImages.cpp:80-87); PNG output stores v + 0.5 clamped
to [0, 255] and truncated, iio's float-to-uint8 rule
(src/Utilities/iio.c:667, 736)."""
import ctypes as C
import os

import numpy as np
import pytest
from PIL import Image

from conftest import ROOT

HOST = os.path.join(ROOT, "vbm3d_amd", "libvbm3d_host.so")


@pytest.fixture(scope="module")
def io():
    L = C.CDLL(HOST)
    up = C.POINTER(C.c_uint)
    fp = C.POINTER(C.c_float)
    L.vbm3d_io_last_error.restype = C.c_char_p
    L.vbm3d_io_load_frame.argtypes = [C.c_char_p, fp, C.c_size_t, up, up, up]
    L.vbm3d_io_save_frame.argtypes = [C.c_char_p, fp, C.c_uint, C.c_uint, C.c_uint]
    L.vbm3d_io_load_flow.argtypes = [C.c_char_p, fp, C.c_size_t, up, up]
    return L


def load(io, path):
    w, h, c = C.c_uint(), C.c_uint(), C.c_uint()
    assert io.vbm3d_io_load_frame(str(path).encode(), None, 0, w, h, c) == 0, io.vbm3d_io_last_error()
    out = np.empty((c.value, h.value, w.value), np.float32)
    assert io.vbm3d_io_load_frame(str(path).encode(), out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                  w, h, c) == 0
    return out


def save(io, path, planar):
    a = np.ascontiguousarray(planar, np.float32)
    c, h, w = a.shape
    return io.vbm3d_io_save_frame(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_float)), w, h, c)


@pytest.mark.parametrize("mode", ["L", "RGB", "RGBA", "P", "I;16"])
def test_png_read_matches_pil(io, tmp_path, mode):
    rng = np.random.default_rng(1)
    if mode == "I;16":
        arr = rng.integers(0, 65536, (21, 33), dtype=np.uint16)
        Image.fromarray(arr).save(tmp_path / "a.png")
        expect = arr.astype(np.float32)[None]
    else:
        arr = rng.integers(0, 256, (21, 33, 4), dtype=np.uint8)
        im = Image.fromarray(arr, "RGBA").convert(mode)
        im.save(tmp_path / "a.png")
        ref = np.asarray(Image.open(tmp_path / "a.png").convert("RGB" if mode != "L" else "L"), np.float32)
        expect = ref[None] if ref.ndim == 2 else np.transpose(ref, (2, 0, 1))
    got = load(io, tmp_path / "a.png")
    np.testing.assert_array_equal(got, expect)


def test_png_grey_rgb_collapses_to_one_channel(io, tmp_path):
    g = np.random.default_rng(2).integers(0, 256, (10, 12), dtype=np.uint8)
    Image.fromarray(np.stack([g, g, g], -1), "RGB").save(tmp_path / "g.png")
    got = load(io, tmp_path / "g.png")
    assert got.shape == (1, 10, 12)
    np.testing.assert_array_equal(got[0], g.astype(np.float32))


def test_png_write_rounds_and_clamps_like_iio(io, tmp_path):
    v = np.array([[[-3.0, 0.2, 0.5, 1.49, 1.5, 254.4, 254.6, 300.0]]], np.float32)
    assert save(io, tmp_path / "o.png", v) == 0
    back = np.asarray(Image.open(tmp_path / "o.png"), np.float32)
    np.testing.assert_array_equal(back[0], [0, 0, 1, 1, 2, 254, 255, 255])
    rgb = np.random.default_rng(3).uniform(0, 255, (3, 9, 7)).astype(np.float32)
    assert save(io, tmp_path / "c.png", rgb) == 0
    back = np.transpose(np.asarray(Image.open(tmp_path / "c.png"), np.float32), (2, 0, 1))
    np.testing.assert_array_equal(back, np.floor(rgb + 0.5))


@pytest.mark.parametrize("mode", ["L", "RGB"])
def test_jpeg_read_close_to_pil(io, tmp_path, mode):
    """Both decoders follow the JPEG standard's IDCT within its accuracy,
    so they may differ by a few levels (4:4:4 chroma, so that the two
    libraries' chroma upsampling filters do not enter); a flat image
    decodes exactly."""
    rng = np.random.default_rng(4)
    arr = (rng.uniform(0, 1, (32, 48, 3)) * 60 + np.linspace(0, 180, 48)[None, :, None]).astype(np.uint8)
    Image.fromarray(arr, "RGB").convert(mode).save(tmp_path / "a.jpg", quality=95, subsampling=0)
    ref = np.asarray(Image.open(tmp_path / "a.jpg"), np.float32)
    expect = ref[None] if ref.ndim == 2 else np.transpose(ref, (2, 0, 1))
    got = load(io, tmp_path / "a.jpg")
    assert got.shape == expect.shape
    assert np.abs(got - expect).max() <= 3 and np.abs(got - expect).mean() < 0.5
    Image.fromarray(np.full((16, 16), 77, np.uint8), "L").save(tmp_path / "f.jpg", quality=90)
    np.testing.assert_array_equal(load(io, tmp_path / "f.jpg"), np.full((1, 16, 16), 77, np.float32))


def test_flo_read(io, tmp_path):
    """Middlebury .flo: tag 202021.25, width, height, then (u, v) pairs."""
    rng = np.random.default_rng(5)
    uv = rng.uniform(-5, 5, (13, 17, 2)).astype(np.float32)
    with open(tmp_path / "a.flo", "wb") as f:
        np.array([202021.25], "<f4").tofile(f)
        np.array([17, 13], "<i4").tofile(f)
        uv.astype("<f4").tofile(f)
    w, h = C.c_uint(), C.c_uint()
    out = np.empty((2, 13, 17), np.float32)
    assert io.vbm3d_io_load_flow(str(tmp_path / "a.flo").encode(), out.ctypes.data_as(C.POINTER(C.c_float)),
                                 out.size, w, h) == 0
    assert (w.value, h.value) == (17, 13)
    np.testing.assert_array_equal(out, np.transpose(uv, (2, 0, 1)))
    with open(tmp_path / "bad.flo", "wb") as f:
        np.array([1.0], "<f4").tofile(f)
        np.array([17, 13], "<i4").tofile(f)
    assert io.vbm3d_io_load_flow(str(tmp_path / "bad.flo").encode(), None, 0, w, h) == -1
    assert b"not a .flo file" in io.vbm3d_io_last_error()


def test_unknown_format_is_refused(io, tmp_path):
    (tmp_path / "a.bmp").write_bytes(b"BM")
    w, h, c = C.c_uint(), C.c_uint(), C.c_uint()
    assert io.vbm3d_io_load_frame(str(tmp_path / "a.bmp").encode(), None, 0, w, h, c) == -1
    assert b"unsupported image format" in io.vbm3d_io_last_error()
