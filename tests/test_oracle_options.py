"""The oracle's non-default options against the independent FP64 model:
temporal patches (kt > 1) with the temporal DCT, motion-compensated
trajectories (mc), group sizes N that are not powers of two or exceed 16,
and 12x12 patches with the 12-point Kaiser window.  Each comparison runs on
tiny clips whose decisions are far from ties (margins checked), and each
option is shown to matter (the model without it is farther off)."""
import numpy as np
import pytest
import scipy.fft

import fp64_model as M
from vbm3d_amd import synth


@pytest.mark.parametrize("kt", [2, 3, 4, 5])
def test_temporal_dct_is_the_orthonormal_dct(oracle, kt):
    """REDFT10 times 1/sqrt(2 kt), DC times 1/sqrt(2) (src/vbm3d.cpp:1866-1910)
    is the orthonormal DCT-II; the inverse (:1912-1956) undoes it."""
    rng = np.random.default_rng(kt)
    v = (rng.standard_normal(kt) * 40).astype(np.float32)
    f = oracle.temporal_fwd(v)
    np.testing.assert_allclose(f, scipy.fft.dct(v.astype(np.float64), type=2, norm="ortho"), rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(oracle.temporal_inv(f), v, atol=1e-4)


def _run(oracle, noisy, sigma, p1, p2, ff=None, bf=None):
    b32, f32, _ = oracle.run_vbm3d(sigma, noisy, p1, p2, True, 1, fflow=ff, bflow=bf)
    b64, f64, marg = M.run(sigma, noisy, p1.as_dict(), p2.as_dict(), True, ff=ff, bf=bf)
    assert marg.rank > 1e-3 and marg.tau > 1e-3 and marg.thr > 1e-3 and marg.traj > 1e-3, vars(marg)
    return np.abs(b32 - b64).max(), np.abs(f32 - f64).max(), b32, f32


def _motion_flow(F, H, W, seed):
    """Smooth flows (a global drift plus a slow spatial wave), so that
    trajectories move by whole pixels over a few slices."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    ff = np.empty((F - 1, 2, H, W), np.float32)
    for t in range(F - 1):
        ff[t, 0] = 0.8 + 0.5 * np.sin(xx / 7.0 + t) + rng.uniform(-0.05, 0.05, (H, W))
        ff[t, 1] = -0.6 + 0.5 * np.cos(yy / 5.0 + t) + rng.uniform(-0.05, 0.05, (H, W))
    return ff, -ff


@pytest.mark.parametrize("kt,seed,dct1", [(2, 15, False), (3, 2, True)])
def test_temporal_patches(oracle, kt, seed, dct1):
    """kt-frame patches: distances over all slices, 2D transforms per slice,
    the temporal DCT (not with Bior, src/vbm3d.cpp:573, 818), and
    aggregation into every slice's frame."""
    sigma = 20.0
    clean, noisy = synth.noisy_clip(32, 24, 5, 1, sigma, seed)
    p1 = oracle.params_1(sigma, kt=kt, T_2D=0 if dct1 else -1)
    p2 = oracle.params_2(sigma, kt=kt)
    eb, ef, _, f32 = _run(oracle, noisy, sigma, p1, p2)
    assert eb < 1e-3 and ef < 1e-3, (eb, ef)
    assert synth.psnr(clean, f32) > synth.psnr(clean, noisy) + 3


def test_motion_compensated_trajectories(oracle):
    """mc: every slice of a patch follows the running sum of the forward
    flow (FLOAT_TRAJECTORIES, src/vbm3d.cpp:378-425) in the distance, the
    transforms and the aggregation; without mc the answer differs."""
    sigma = 20.0
    W, H, F = 32, 24, 5
    _, noisy = synth.noisy_clip(W, H, F, 1, sigma, 1)
    ff, bf = _motion_flow(F, H, W, 1)
    p1 = oracle.params_1(sigma, kt=3)
    p2 = oracle.params_2(sigma, kt=3)
    p1.mc = p2.mc = 1
    eb, ef, b32, f32 = _run(oracle, noisy, sigma, p1, p2, ff, bf)
    assert eb < 1e-3 and ef < 1e-3, (eb, ef)
    d1, d2 = p1.as_dict(), p2.as_dict()
    d1["mc"] = d2["mc"] = 0
    _, f_nomc, _ = M.run(sigma, noisy, d1, d2, True, ff=ff, bf=bf)
    assert np.abs(f32 - f_nomc).max() > 1.0


@pytest.mark.parametrize("N,seed", [(6, 4), (12, 14), (32, 34)])
def test_group_sizes_beyond_powers_of_two(oracle, N, seed):
    """N keeps min(N, candidates) sorted matches and truncates the group to
    a power of two (src/vbm3d.cpp:1169-1184): N = 6 and 12 (not powers of
    two) and N = 32 (more than the fast kernels hold)."""
    sigma = 20.0
    _, noisy = synth.noisy_clip(32, 24, 5, 1, sigma, seed)
    nb = 4 if N == 32 else -1  # enough candidates per frame for 32-member groups
    p1 = oracle.params_1(sigma, N=N, Nb=nb)
    p2 = oracle.params_2(sigma, N=N, Nb=nb)
    eb, ef, _, _ = _run(oracle, noisy, sigma, p1, p2)
    assert eb < 1e-3 and ef < 1e-3, (eb, ef)


def test_patch_size_12(oracle):
    """12x12 patches in both steps (DCT, the 12-point Kaiser window of
    src/vbm3d.cpp:1820-1836)."""
    sigma = 30.0
    clean, noisy = synth.noisy_clip(40, 32, 4, 1, sigma, 0)
    p1 = oracle.params_1(sigma, k=12)
    p2 = oracle.params_2(sigma, k=12)
    assert p1.T_2D == 0 and p2.T_2D == 0
    eb, ef, _, f32 = _run(oracle, noisy, sigma, p1, p2)
    assert eb < 1e-3 and ef < 1e-3, (eb, ef)
    assert synth.psnr(clean, f32) > synth.psnr(clean, noisy) + 3
