"""Pins the oracle's building blocks against closed forms and independent
derivations (the reference cannot be built here and ships no tests)."""
import numpy as np
import pytest
import scipy.fft

import fp64_model as M


def _bior_1d_from_oracle(oracle):
    """Recover the 8x8 analysis / synthesis matrices from the oracle's 2D
    transforms applied to impulses."""
    b00 = 0.353553390593274
    BF = np.zeros((8, 8))
    BI = np.zeros((8, 8))
    for s in range(8):
        x = np.zeros((8, 8), np.float32)
        x[0, s] = 1.0
        out = oracle.bior_fwd(x).reshape(8, 8).astype(np.float64)  # out[xf][yf]
        BF[:, s] = out[:, 0] / b00
        c = np.zeros(64, np.float32)
        c[s] = 1.0  # xf = 0, yf = s
        BI[:, s] = oracle.bior_inv(c).reshape(8, 8)[:, 0].astype(np.float64) / b00
    return BF, BI


def test_bior_tables_are_the_bior15_filter_bank(oracle):
    """Analysis table = 3-level periodised bior1.5 DWT with unit-norm rows;
    synthesis table = the periodised inverse DWT (not re-normalised), so the
    reference's 2D Bior round trip is NOT the identity."""
    BF, BI = _bior_1d_from_oracle(oracle)
    MF, MI = M.bior_matrices()
    np.testing.assert_allclose(BF, MF, atol=2e-7)
    np.testing.assert_allclose(BI, MI, atol=2e-7)
    err = np.abs(BI @ BF - np.eye(8)).max()
    assert 0.01 < err < 0.02


def test_bior_2d_layout(oracle):
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 255, (8, 8)).astype(np.float32)
    MF, MI = M.bior_matrices()
    coef = MF @ x.astype(np.float64) @ MF.T           # [yf][xf]
    out = oracle.bior_fwd(x).reshape(8, 8)           # [xf][yf]
    np.testing.assert_allclose(out, coef.T, rtol=1e-5, atol=1e-3)
    back = oracle.bior_inv(out.reshape(-1)).reshape(8, 8)
    np.testing.assert_allclose(back, MI @ coef @ MI.T, rtol=1e-5, atol=1e-3)
    # constant patch -> DC only: 8 * value
    c = np.full((8, 8), 3.0, np.float32)
    o = oracle.bior_fwd(c)
    assert abs(o[0] - 24.0) < 1e-4 and np.abs(o[1:]).max() < 1e-4


@pytest.mark.parametrize("k", [4, 5, 6, 7, 8])
def test_dct_is_orthonormal_dct2(oracle, k):
    rng = np.random.default_rng(k)
    x = rng.uniform(0, 255, (k, k)).astype(np.float32)
    ref = scipy.fft.dctn(x.astype(np.float64), type=2, norm="ortho")
    got = oracle.dct_fwd(x).reshape(k, k)
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-4)
    back = oracle.dct_inv(got.reshape(-1), k).reshape(k, k)
    np.testing.assert_allclose(back, scipy.fft.idctn(got.astype(np.float64), type=2, norm="ortho"), atol=2e-4)
    np.testing.assert_allclose(back, x, atol=3e-4)


@pytest.mark.parametrize("k", [4, 7, 8])
def test_dct_norm_tables(oracle, k):
    nrm, inv = oracle.dct_norm(k)
    c = 0.5 / k
    ref = np.full((k, k), c)
    ref[0, :] *= np.sqrt(0.5)
    ref[:, 0] *= np.sqrt(0.5)
    np.testing.assert_allclose(nrm, ref, rtol=1e-6)
    refi = np.ones((k, k))
    refi[0, :] *= np.sqrt(2)
    refi[:, 0] *= np.sqrt(2)
    np.testing.assert_allclose(inv, refi, rtol=1e-6)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
def test_haar_orthonormal(oracle, n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n).astype(np.float32) * 50
    Hm = M.haar_matrix(n)
    np.testing.assert_allclose(Hm @ Hm.T, np.eye(n), atol=1e-12)
    f = oracle.haar_fwd(v)
    np.testing.assert_allclose(f, Hm @ v.astype(np.float64), rtol=1e-6, atol=1e-4)
    np.testing.assert_allclose(oracle.haar_inv(f), v, atol=1e-4)
    assert abs(np.sum(f.astype(np.float64) ** 2) - np.sum(v.astype(np.float64) ** 2)) < 1e-3 * np.sum(v ** 2)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_hadamard(oracle, n):
    Hm = np.stack([oracle.hadamard(np.eye(n, dtype=np.float32)[i]) for i in range(n)], axis=1)
    assert set(np.unique(Hm)) == {-1.0, 1.0}
    np.testing.assert_array_equal(Hm @ Hm.T, n * np.eye(n))


@pytest.mark.parametrize("k", [4, 5, 6, 7, 8, 12])
def test_kaiser_tables(oracle, k):
    np.testing.assert_allclose(oracle.kaiser(k), M.kaiser2d(k), atol=1e-7)


def test_sigma_tables(oracle):
    assert oracle.sigma_table(20.0, 1, 1)[0] == np.float32(20.0)
    for cs, ycbcr in ((1, True), (0, False)):
        np.testing.assert_allclose(oracle.sigma_table(20.0, 3, cs), M.sigma_rows(20.0, 3, ycbcr), rtol=1e-6)
    # the quirk: colour_space=true selects the YCbCr row (not the YUV noise)
    assert abs(oracle.sigma_table(1.0, 3, 1)[1] - 0.622995) < 1e-5
    with pytest.raises(ValueError):
        oracle.sigma_table(20.0, 2, 1)


def test_colour_transform(oracle):
    rng = np.random.default_rng(1)
    v = rng.uniform(0, 255, (2, 3, 5, 7)).astype(np.float32)
    f = oracle.color(v, True)
    np.testing.assert_allclose(f, np.einsum("ij,fjhw->fihw", M.YUV, v.astype(np.float64)), atol=1e-3)
    b = oracle.color(f, False)
    np.testing.assert_allclose(b, np.einsum("ij,fjhw->fihw", M.YUV_INV, f.astype(np.float64)), atol=1e-3)
    # the published pair is only approximately inverse
    assert np.abs(b - v).max() < 0.05
    g = rng.uniform(0, 255, (2, 1, 5, 7)).astype(np.float32)
    np.testing.assert_array_equal(oracle.color(g, True), g)


def test_psnr_definition(oracle):
    from vbm3d_amd import synth
    rng = np.random.default_rng(2)
    a = rng.uniform(-20, 280, 5000).astype(np.float32)
    b = rng.uniform(-20, 280, 5000).astype(np.float32)
    p, r = oracle.psnr(a, b)
    d = np.clip(a, 0, 255).astype(np.float64) - np.clip(b, 0, 255)
    rm = np.sqrt(np.mean(d ** 2))
    assert abs(r - rm) < 1e-4 and abs(p - 20 * np.log10(255 / rm)) < 1e-4
    assert abs(synth.psnr(a.reshape(1, 1, 50, 100), b.reshape(1, 1, 50, 100)) - p) < 1e-4
