"""Pins of the multiscale pyramid oracle (oracle/multiscale_oracle.cpp),
CPU only.  The reference's multiscaler needs FFTW and iio and does not
build here, so the pins are what the mathematics fixes:
  * the forward and inverse DCT equal scipy's unnormalised DCT-II / DCT-III
    (FFTW's REDFT10 / REDFT01) with the 1 / (4 rows cols) normalisation;
  * decompose then recompose returns the image: exactly for the Gaussian
    and Lanczos pyramids (recompose subtracts the same downsampling that
    built the next level), to rounding for the DCT pyramid (a level's
    normalised DCT is the full DCT's low-frequency block);
  * the blur and the resamplers keep constant images constant;
  * merge_coarse with the image's own next level leaves it unchanged.
"""
import numpy as np
import pytest
import scipy.fft as F

from oracle import multiscale as M


def img(ch, rows, cols, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:rows, 0:cols]
    base = 120 + 60 * np.sin(x / 5.0) * np.cos(y / 7.0)
    return np.stack([base + rng.normal(0, 10, (rows, cols)) for _ in range(ch)]).astype(np.float32)


@pytest.mark.parametrize("shape", [(1, 16, 16), (3, 37, 50), (1, 9, 64)])
def test_dct_is_redft10_and_redft01(shape):
    a = img(*shape, seed=1)
    rows, cols = shape[1:]
    d = M.dct2(a)
    np.testing.assert_allclose(d, F.dctn(a.astype(np.float64), type=2, axes=(1, 2)) / (4 * rows * cols),
                               rtol=0, atol=1e-4)
    np.testing.assert_allclose(M.idct2(d), F.dctn(d.astype(np.float64), type=3, axes=(1, 2)), rtol=0, atol=2e-4)
    np.testing.assert_allclose(M.idct2(d), a, rtol=0, atol=1e-3)


@pytest.mark.parametrize("type_", [1, 2])
@pytest.mark.parametrize("shape", [(1, 40, 56), (3, 37, 51)])
def test_gaussian_and_lanczos_round_trip_is_exact(type_, shape):
    a = img(*shape, seed=2)
    lev = M.decompose(a, 3, type_)
    assert [l.shape[1:] for l in lev] == [(shape[1] >> i, shape[2] >> i) for i in range(3)]
    np.testing.assert_array_equal(M.recompose(lev, type_, 0.7), a)


def test_dct_round_trip():
    a = img(3, 37, 50, seed=3)
    for factor in (0.5, 0.7, 1.0):
        np.testing.assert_allclose(M.recompose(M.decompose(a, 3, 0), 0, factor), a, rtol=0, atol=2e-3)


def test_dct_levels_keep_the_mean_and_low_frequencies():
    a = img(1, 32, 48, seed=4)
    lev = M.decompose(a, 3, 0)
    for l in lev:
        assert abs(l.mean() - a.mean()) < 1e-3
    # a level's normalised DCT is the full DCT's top-left block
    d = M.dct2(a)
    np.testing.assert_allclose(M.dct2(lev[1]), d[:, :16, :24], rtol=0, atol=1e-4)


def test_constants_survive_blur_and_resampling():
    c = np.full((1, 20, 30), 77.0, np.float32)
    for lz in (False, True):
        np.testing.assert_allclose(M.down(c, lz), 77.0, rtol=1e-6)
        np.testing.assert_allclose(M.up(c[:, :10, :15], 20, 30, lz), 77.0, rtol=1e-6)
        np.testing.assert_allclose(M.up(c[:, :10, :15], 21, 31, lz), 77.0, rtol=1e-6)
    np.testing.assert_allclose(M.gblur(c, 0.7), 77.0, rtol=1e-6)
    a = img(1, 12, 12, seed=5)
    np.testing.assert_array_equal(M.gblur(a, 0.4), a)  # below 0.41: no blur


def test_merge_with_own_next_level_is_identity():
    a = img(3, 40, 60, seed=6)
    coarse = M.decompose(a, 2, 0)[1]
    np.testing.assert_allclose(M.merge_coarse(a, coarse, 0.8), a, rtol=0, atol=2e-3)
    # a different coarse image does change the low frequencies
    other = coarse + 5.0
    assert np.abs(M.merge_coarse(a, other, 0.8) - a).mean() > 1.0


def test_upsampling_of_uneven_targets_fills_every_sample():
    a = img(1, 10, 15, seed=7)
    for lz in (False, True):
        u = M.up(a, 21, 31, lz)
        assert np.isfinite(u).all() and (u != 0).all()
