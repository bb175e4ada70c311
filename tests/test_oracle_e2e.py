"""End to end: the FP32 oracle against the independent FP64 model on tiny
clips whose decisions are far from ties, plus invariants."""
import numpy as np
import pytest

import fp64_model as M
from vbm3d_amd import synth


# seeds whose ranking, tau and threshold decisions all have margins > 1e-3
@pytest.mark.parametrize("C,seed,sigma", [(1, 14, 20.0), (3, 10, 20.0), (1, 14, 40.0), (3, 20, 40.0)])
def test_oracle_matches_fp64_model(oracle, C, seed, sigma):
    clean, noisy = synth.noisy_clip(32, 24, 4, C, sigma, seed)
    p1, p2 = oracle.params_1(sigma), oracle.params_2(sigma)
    b32, f32, _ = oracle.run_vbm3d(sigma, noisy, p1, p2, True, 1)
    b64, f64, marg = M.run(sigma, noisy, p1.as_dict(), p2.as_dict(), True)
    # decisions must be unambiguous for the comparison to be meaningful
    assert marg.rank > 1e-3 and marg.tau > 1e-3 and marg.thr > 1e-3, vars(marg)
    assert np.abs(b32 - b64).max() < 1e-3
    assert np.abs(f32 - f64).max() < 1e-3
    assert synth.psnr(clean, f32) > synth.psnr(clean, noisy) + 3


def test_constant_video_is_preserved(oracle):
    v = np.full((5, 3, 24, 32), 77.0, np.float32)
    b, f, _ = oracle.run_vbm3d(10.0, v, oracle.params_1(10.0), oracle.params_2(10.0), True, 1)
    assert np.abs(b - 77.0).max() < 0.02 and np.abs(f - 77.0).max() < 0.02


def test_fp64_replay_per_step_error(oracle):
    """The FP32 step against its FP64 re-evaluation with the same groups and
    threshold decisions: the per-step arithmetic error is tiny."""
    _, noisy = synth.noisy_clip(48, 40, 5, 3, 20.0, 5)
    yuv = oracle.color(noisy, True)
    p1 = oracle.params_1(20.0)
    n32, d32 = oracle.step(1, 20.0, yuv, None, p1)
    n64, d64 = oracle.step_fp64(1, 20.0, yuv, None, p1)
    b32 = n32 / d32
    b64 = n64 / d64
    assert np.abs(b32 - b64).max() < 1e-3
    p2 = oracle.params_2(20.0)
    m32, e32 = oracle.step(2, 20.0, yuv, b32, p2)
    m64, e64 = oracle.step_fp64(2, 20.0, yuv, b32, p2)
    assert np.abs(m32 / e32 - m64 / e64).max() < 1e-3


def test_threaded_grid_changes_output_only_through_the_grid(oracle):
    """nb > 1 adds duplicated reference patches near tile borders; the result
    stays a valid denoising and differs from nb = 1 (the reference's output
    depends on its thread count)."""
    clean, noisy = synth.noisy_clip(96, 64, 5, 1, 20.0, 6)
    p1, p2 = oracle.params_1(20.0), oracle.params_2(20.0)
    _, f1, _ = oracle.run_vbm3d(20.0, noisy, p1, p2, True, 1)
    _, f4, _ = oracle.run_vbm3d(20.0, noisy, p1, p2, True, 4)
    assert not np.array_equal(f1, f4)
    assert abs(synth.psnr(clean, f1) - synth.psnr(clean, f4)) < 0.2
