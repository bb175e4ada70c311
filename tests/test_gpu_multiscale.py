"""GPU parity of the multiscale pyramid (pyramid.hip) and the multiscale
denoising pipeline with the oracle.

The Gaussian and Lanczos pyramids are float convolutions in the
reference's order: bit-exact.  The DCT pyramid's transforms are FP64 matrix
products (rocBLAS) rounded once to float where the oracle sums the same
products in double sequentially: the two can differ by one float ulp
(tolerance 1e-3 on the 0-255 scale).  The pipeline is checked in its
parts: the noisy levels against the oracle's decomposition, the denoised
levels bit for bit against the oracle's run_vbm3d on the GPU's own noisy
levels, the result against the oracle's recomposition of the GPU's
denoised levels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def img(ch, rows, cols, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:rows, 0:cols]
    base = 120 + 60 * np.sin(x / 5.0) * np.cos(y / 7.0)
    return np.stack([base + rng.normal(0, 10, (rows, cols)) for _ in range(ch)]).astype(np.float32)


SHAPES = [(1, 40, 56), (3, 37, 51), (3, 64, 96)]


@pytest.mark.parametrize("type_", [0, 1, 2], ids=["dct", "gauss", "lanczos"])
@pytest.mark.parametrize("shape", SHAPES, ids=["g40x56", "rgb37x51", "rgb64x96"])
def test_decompose_recompose(gpu_ctx, type_, shape):
    import vbm3d_amd as V
    from oracle import multiscale as M
    a = img(*shape, seed=sum(shape) + type_)
    lev = V.pyr_decompose(a, 3, type_, ctx=gpu_ctx)
    olev = M.decompose(a, 3, type_)
    for g, o in zip(lev, olev):
        if type_ == 0:
            np.testing.assert_allclose(g, o, rtol=0, atol=1e-3)
        else:
            assert bits_equal(g, o)
    # recompose from the same (oracle) levels on both sides, modified so
    # that the coarse levels matter
    mod = [l + (3.0 * i) for i, l in enumerate(olev)]
    for factor in (0.5, 0.7):
        r = V.pyr_recompose(mod, type_, factor, ctx=gpu_ctx)
        o = M.recompose(mod, type_, factor)
        if type_ == 0:
            np.testing.assert_allclose(r, o, rtol=0, atol=1e-3)
        else:
            assert bits_equal(r, o)


def test_merge_coarse(gpu_ctx):
    import vbm3d_amd as V
    from oracle import multiscale as M
    a = img(3, 48, 72, 11)
    coarse = M.decompose(a, 2, 0)[1] + 4.0
    np.testing.assert_allclose(V.pyr_merge_coarse(a, coarse, 0.8, ctx=gpu_ctx), M.merge_coarse(a, coarse, 0.8),
                               rtol=0, atol=1e-3)
    with pytest.raises(V.VBM3DError):
        V.pyr_merge_coarse(a, coarse, 2.5, ctx=gpu_ctx)


def _level_sigma(sigma, lvl):
    return float(np.floor(sigma / 2 ** lvl * 100 + 1e-9) / 100)


@pytest.mark.parametrize("type_", [0, 1, 2], ids=["dct", "gauss", "lanczos"])
def test_multiscale_denoise_pipeline(gpu_ctx, type_):
    import vbm3d_amd as V
    from oracle import multiscale as M
    from oracle import oracle as O
    from vbm3d_amd import synth
    sigma, levels = 25.0, 3
    clean, noisy = synth.noisy_clip(96, 64, 6, 3, sigma, 40 + type_)
    out, nlev, dlev = V.multiscale_denoise(noisy, sigma, levels, type_, 0.7, return_levels=True, ctx=gpu_ctx)
    F = noisy.shape[0]
    # the noisy levels: the oracle's decomposition of every frame
    for t in range(F):
        ol = M.decompose(noisy[t], levels, type_)
        for i in range(levels):
            if type_ == 0:
                np.testing.assert_allclose(nlev[i][t], ol[i], rtol=0, atol=1e-3)
            else:
                assert bits_equal(nlev[i][t], ol[i])
    # the denoised levels: the oracle's run_vbm3d on those levels, with the
    # level's sigma and its default parameters
    for i in range(levels):
        sl = _level_sigma(sigma, i)
        p1, p2 = V.initialize_parameters_1(sl), V.initialize_parameters_2(sl)
        _, of, _ = O.run_vbm3d(sl, nlev[i], p1, p2, True, 1)
        assert bits_equal(dlev[i], of), (i, np.abs(dlev[i] - of).max())
    # the result: the oracle's recomposition of the denoised levels
    for t in range(F):
        o = M.recompose([dlev[i][t] for i in range(levels)], type_, 0.7)
        if type_ == 0:
            np.testing.assert_allclose(out[t], o, rtol=0, atol=1e-3)
        else:
            assert bits_equal(out[t], o)
    assert synth.psnr(clean, out) > synth.psnr(clean, noisy) + 3


def test_pyramid_command_lines(gpu_ctx, tmp_path):
    """bin/decompose, bin/recompose and bin/merge_coarse on float TIFF
    files: the same levels and images as the API (the multiscale script's
    invocations, src/utils/VBM3Ddenoising_multiscale.sh:45-75)."""
    import os
    import subprocess
    import ctypes as C
    import vbm3d_amd as V
    from conftest import ROOT
    a = img(3, 48, 64, 21)
    io = C.CDLL(os.path.join(ROOT, "vbm3d_amd", "libvbm3d_host.so"))
    io.vbm3d_io_save_frame.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_uint]
    io.vbm3d_io_load_frame.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_uint),
                                       C.POINTER(C.c_uint), C.POINTER(C.c_uint)]

    def save(path, x):
        x = np.ascontiguousarray(x, np.float32)
        assert io.vbm3d_io_save_frame(str(path).encode(), x.ctypes.data_as(C.POINTER(C.c_float)), x.shape[2],
                                      x.shape[1], x.shape[0]) == 0

    def load(path):
        w, h, c = C.c_uint(), C.c_uint(), C.c_uint()
        assert io.vbm3d_io_load_frame(str(path).encode(), None, 0, w, h, c) == 0
        out = np.empty((c.value, h.value, w.value), np.float32)
        assert io.vbm3d_io_load_frame(str(path).encode(), out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                      w, h, c) == 0
        return out

    bindir = os.path.join(ROOT, "vbm3d_amd", "bin")
    save(tmp_path / "in.tiff", a)
    for t in (0, 1):
        r = subprocess.run([os.path.join(bindir, "decompose"), str(tmp_path / "in.tiff"), str(tmp_path / f"lev{t}_"),
                            "3", ".tiff", "-t", str(t)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        api = V.pyr_decompose(a, 3, t, ctx=gpu_ctx)
        for i in range(3):
            assert bits_equal(load(tmp_path / f"lev{t}_{i}.tiff"), api[i])
        r = subprocess.run([os.path.join(bindir, "recompose"), str(tmp_path / f"lev{t}_"), "3", ".tiff",
                            str(tmp_path / f"out{t}.tiff"), "-c", "0.7", "-t", str(t)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert bits_equal(load(tmp_path / f"out{t}.tiff"), V.pyr_recompose(api, t, 0.7, ctx=gpu_ctx))
    r = subprocess.run([os.path.join(bindir, "merge_coarse"), str(tmp_path / "in.tiff"), str(tmp_path / "lev0_1.tiff"),
                        str(tmp_path / "m.tiff")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    coarse = V.pyr_decompose(a, 2, 0, ctx=gpu_ctx)[1]
    assert bits_equal(load(tmp_path / "m.tiff"), V.pyr_merge_coarse(a, coarse, 0.8, ctx=gpu_ctx))
    r = subprocess.run([os.path.join(bindir, "decompose"), str(tmp_path / "in.tiff")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "Usage" in r.stderr
