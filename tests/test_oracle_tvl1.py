"""Pins of the TV-L1 flow oracle (oracle/tvl1_oracle.cpp), CPU only.

Three independent checks:
  * against the reference's own library, compiled from
    tvl1flow/tvl1flow_lib.c into oracle/_ref (``make -C oracle ref``),
    bit for bit, building block by building block and for whole flows
    (skipped where that library is not built);
  * against tests/golden/tvl1_flow.npz, flows the reference library
    wrote (tests/golden/make_tvl1_golden.py), bit for bit;
  * against what the mathematics fixes: the discrete divergence is minus
    the adjoint of the forward gradient (Chambolle 2004, the identity the
    reference's edge cases are built for), the Gaussian keeps constants,
    the bicubic cell interpolates its samples and reproduces linear ramps,
    and the flow of a translated image is that translation.
"""
import os

import numpy as np
import pytest

from oracle import tvl1 as T

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "tvl1_flow.npz")
needs_ref = pytest.mark.skipif(not T.have_ref(), reason="reference TV-L1 library not built (make -C oracle ref)")


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def pair(w, h, seed, sigma=10.0):
    from vbm3d_amd import synth
    _, noisy = synth.noisy_clip(w, h, 2, 1, sigma, seed)
    return noisy[0, 0].copy(), noisy[1, 0].copy()


@needs_ref
def test_building_blocks_match_reference():
    rng = np.random.default_rng(0)
    I0, I1 = pair(61, 43, 1)
    for s in (0.8, 0.6 * np.sqrt(3.0), 2.5):
        assert bits_equal(T.gaussian(I0, s), T.gaussian(I0, s, ref=True)), s
    for f in (0.5, 0.7, 0.33):
        assert bits_equal(T.zoom_out(I0, f), T.zoom_out(I0, f, ref=True)), f
    for nxx, nyy in ((122, 86), (97, 61), (61, 44)):
        assert bits_equal(T.zoom_in(I0, nxx, nyy), T.zoom_in(I0, nxx, nyy, ref=True))
    assert bits_equal(T.divergence(I0, I1), T.divergence(I0, I1, ref=True))
    for a, b in zip(T.centered_gradient(I0), T.centered_gradient(I0, ref=True)):
        assert bits_equal(a, b)
    for a, b in zip(T.forward_gradient(I0), T.forward_gradient(I0, ref=True)):
        assert bits_equal(a, b)
    for a, b in zip(T.normalize(I0, I1 * 3 - 40), T.normalize(I0, I1 * 3 - 40, ref=True)):
        assert bits_equal(a, b)
    # samples inside, on the edges, outside, at negative positions
    pts = np.concatenate([rng.uniform(-4, 66, (600, 2)), rng.integers(-2, 63, (100, 2))]).astype(np.float32)
    for bo in (False, True):
        got = [T.bicubic(I0, x, y, bo) for x, y in pts]
        ref = [T.bicubic(I0, x, y, bo, ref=True) for x, y in pts]
        assert bits_equal(got, ref)


FLOW_CASES = {
    "default": (64, 48, 2, {}),
    "eps0": (40, 32, 3, {"epsilon": 0.0, "warps": 2, "nscales": 2}),
    "fscale1": (72, 56, 4, {"fscale": 1}),
    "z07-w3": (80, 40, 5, {"zfactor": 0.7, "warps": 3}),
    "odd": (37, 29, 6, {"tau": 0.2, "lam": 0.3, "theta": 0.5}),
}


@needs_ref
@pytest.mark.parametrize("case", list(FLOW_CASES), ids=list(FLOW_CASES))
def test_flow_matches_reference(case):
    w, h, seed, kw = FLOW_CASES[case]
    I0, I1 = pair(w, h, seed)
    u, v, it = T.flow(I0, I1, **kw)
    ur, vr, _ = T.flow(I0, I1, ref=True, **kw)
    assert bits_equal(u, ur) and bits_equal(v, vr)
    assert it.min() >= 1 and it.max() <= 300


def test_golden_flows():
    g = np.load(GOLDEN)
    names = sorted({k.split("/")[0] for k in g.files})
    assert len(names) >= 4
    for n in names:
        fscale, zf, warps, eps, nsc = g[f"{n}/params"]
        u, v, _ = T.flow(g[f"{n}/I0"], g[f"{n}/I1"], fscale=int(fscale), zfactor=float(zf), warps=int(warps),
                         epsilon=float(eps), nscales=int(nsc))
        assert bits_equal(u, g[f"{n}/u"]) and bits_equal(v, g[f"{n}/v"]), n


def test_divergence_is_minus_adjoint_of_gradient():
    """<grad u, p> = -<u, div p> for every u and p, edges included: a
    wrong sign, index or edge term breaks it."""
    rng = np.random.default_rng(1)
    for nx, ny in ((9, 7), (16, 16), (33, 5)):
        u = rng.standard_normal((ny, nx)).astype(np.float32)
        p1 = rng.standard_normal((ny, nx)).astype(np.float32)
        p2 = rng.standard_normal((ny, nx)).astype(np.float32)
        ux, uy = T.forward_gradient(u)
        lhs = np.sum(ux.astype(np.float64) * p1 + uy.astype(np.float64) * p2)
        rhs = -np.sum(u.astype(np.float64) * T.divergence(p1, p2))
        assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + nx * ny), (nx, ny, lhs, rhs)


def test_gradients_of_ramps():
    y, x = np.mgrid[0:20, 0:30].astype(np.float32)
    f = 2.0 * x - 3.0 * y
    dx, dy = T.centered_gradient(f)
    np.testing.assert_allclose(dx[:, 1:-1], 2.0)
    np.testing.assert_allclose(dy[1:-1, :], -3.0)
    np.testing.assert_allclose(dx[:, [0, -1]], 1.0)   # one-sided, halved
    np.testing.assert_allclose(dy[[0, -1], :], -1.5)
    fx, fy = T.forward_gradient(f)
    np.testing.assert_allclose(fx[:, :-1], 2.0)
    np.testing.assert_allclose(fx[:, -1], 0.0)
    np.testing.assert_allclose(fy[:-1, :], -3.0)
    np.testing.assert_allclose(fy[-1, :], 0.0)


def test_gaussian_keeps_constants_and_is_symmetric():
    c = np.full((30, 40), 77.0, np.float32)
    np.testing.assert_allclose(T.gaussian(c, 0.8), 77.0, rtol=1e-6)
    np.testing.assert_allclose(T.gaussian(c, 1.7), 77.0, rtol=1e-6)
    imp = np.zeros((31, 31), np.float32)
    imp[15, 15] = 1.0
    g = T.gaussian(imp, 1.2)
    np.testing.assert_allclose(g, g.T, atol=1e-9)
    np.testing.assert_allclose(g, g[::-1, ::-1], atol=1e-9)
    assert abs(g.sum() - 1.0) < 1e-5
    # the taps follow exp(-i^2 / (2 sigma^2)) up to normalisation
    row = g[15, 15:21] / g[15, 15]
    np.testing.assert_allclose(row, np.exp(-np.arange(6) ** 2 / (2 * 1.2 ** 2)), rtol=1e-4)


def test_bicubic_interpolates_and_reproduces_ramps():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 255, (12, 14)).astype(np.float32)
    for yy in range(1, 10):
        for xx in range(1, 12):
            assert T.bicubic(img, float(xx), float(yy), False) == img[yy, xx]
    y, x = np.mgrid[0:12, 0:14].astype(np.float32)
    ramp = (1.5 * x + 0.25 * y + 3.0).astype(np.float32)
    for xx, yy in rng.uniform(1.2, 9.8, (50, 2)):
        assert abs(T.bicubic(ramp, float(xx), float(yy), False) - (1.5 * np.float32(xx) + 0.25 * np.float32(yy) + 3)) < 1e-4
    # border_out: a sample whose support leaves the image reads as zero
    assert T.bicubic(img, 12.6, 5.0, True) == 0.0
    assert T.bicubic(img, 12.6, 5.0, False) != 0.0


def test_flow_of_a_translation():
    """A smooth random image and its copy shifted by (1.25, -0.5) pixels
    (Fourier shift): the recovered flow is that shift in the interior."""
    rng = np.random.default_rng(3)
    n = 96
    base = T.gaussian(rng.uniform(0, 255, (n, n)).astype(np.float32), 2.0)
    dx, dy = 1.25, -0.5
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    F = np.fft.fft2(base.astype(np.float64))
    # I1(x) = I0(x - d): the flow from I0 to I1 is +d
    shifted = np.real(np.fft.ifft2(F * np.exp(-2j * np.pi * (kx * dx + ky * dy)))).astype(np.float32)
    u, v, _ = T.flow(base, shifted)
    c = slice(16, n - 16)
    assert abs(np.median(u[c, c]) - dx) < 0.05, np.median(u[c, c])
    assert abs(np.median(v[c, c]) - dy) < 0.05, np.median(v[c, c])


def test_level_cap():
    """main.c:139-146: 1 + log(|diag| / 16) / log(1 / zfactor) levels."""
    for nx, ny, z in ((1920, 1080, 0.5), (176, 144, 0.5), (640, 360, 0.7), (33, 17, 0.5)):
        expect = int(np.float32(1 + np.log(np.hypot(nx, ny) / 16.0) / np.log(float(np.float32(1) / np.float32(z)))))
        assert T.nscales(nx, ny, z, 100) == min(100, expect)
    assert T.nscales(1920, 1080, 0.5, 3) == 3
