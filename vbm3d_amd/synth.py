"""Deterministic synthetic video clips for parity tests and benchmarks.

A clip is the sum of six oriented sinusoids plus twelve constant-intensity
rectangles moving at integer velocities, clipped to [0, 255], then noised
with float32 AWGN (no clipping after the noise, as the reference's
VideoUtils::addNoise, src/Utilities/LibVideoT.hpp:541-564).

Fixed generator choices:
  * RNG: numpy PCG64 ``default_rng(seed)``; geometry draws first, then the
    noise from the same stream.
  * Geometry (sinusoid orientation, period, phase; rectangle size, start,
    velocity) is shared by all channels; amplitudes, rectangle intensities
    and the channel offset are drawn per channel.  Gray clips use offset
    128; RGB clips draw each channel's offset from U[64, 192].
  * Sinusoids: orientation U[0, pi), period U[8, 64] px, amplitude
    U[10, 30], phase U[0, 2 pi); static over time.
  * Rectangles: sides integer U[10, 60] px, top-left start anywhere that
    keeps at least one pixel inside the frame, per-axis integer velocity
    U{-vmax..vmax} px/frame, intensity U[0, 255]; painted in draw order
    over the background.
Layout: float32 [F][C][H][W], the reference's planar Video layout
(src/Utilities/LibVideoT.hpp:88-92).
"""
from __future__ import annotations

import numpy as np


def make_clip(width: int, height: int, frames: int, channels: int, seed: int,
              vmax: int = 2, n_sin: int = 6, n_rect: int = 12):
    """Return the clean clip (float32 [F][C][H][W]) and the RNG to noise it."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, np.pi, n_sin)
    period = rng.uniform(8.0, 64.0, n_sin)
    phase = rng.uniform(0.0, 2.0 * np.pi, n_sin)
    amp = rng.uniform(10.0, 30.0, (channels, n_sin))
    if channels == 1:
        offset = np.array([128.0])
    else:
        offset = rng.uniform(64.0, 192.0, channels)
    rw = rng.integers(10, 61, n_rect)
    rh = rng.integers(10, 61, n_rect)
    rx = np.array([rng.integers(-w + 1, width) for w in rw])
    ry = np.array([rng.integers(-h + 1, height) for h in rh])
    vx = rng.integers(-vmax, vmax + 1, n_rect)
    vy = rng.integers(-vmax, vmax + 1, n_rect)
    inten = rng.uniform(0.0, 255.0, (channels, n_rect))

    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    bg = np.empty((channels, height, width), np.float64)
    for c in range(channels):
        acc = np.full((height, width), offset[c])
        for s in range(n_sin):
            u = xx * np.cos(theta[s]) + yy * np.sin(theta[s])
            acc += amp[c, s] * np.sin(2.0 * np.pi * u / period[s] + phase[s])
        bg[c] = acc

    clean = np.empty((frames, channels, height, width), np.float32)
    for t in range(frames):
        fr = bg.copy()
        for r in range(n_rect):
            x0 = rx[r] + vx[r] * t
            y0 = ry[r] + vy[r] * t
            xa, xb = max(0, x0), min(width, x0 + rw[r])
            ya, yb = max(0, y0), min(height, y0 + rh[r])
            if xa < xb and ya < yb:
                fr[:, ya:yb, xa:xb] = inten[:, r, None, None]
        clean[t] = np.clip(fr, 0.0, 255.0).astype(np.float32)
    return clean, rng


def noisy_clip(width: int, height: int, frames: int, channels: int, sigma: float,
               seed: int, vmax: int = 2):
    """Clean clip and its float32 AWGN-noised version (sigma on 0-255)."""
    clean, rng = make_clip(width, height, frames, channels, seed, vmax=vmax)
    noise = rng.standard_normal(clean.shape, dtype=np.float32)
    noisy = clean + np.float32(sigma) * noise
    return clean, noisy.astype(np.float32)


BATCH_SIZES = ((320, 240), (640, 360), (960, 540), (1280, 720))
BATCH_SIGMAS = (10.0, 20.0, 30.0, 50.0)


def batch_spec(seed: int):
    """(W, H, F, C, sigma) of the mixed-clip batch's clip `seed`: from
    default_rng(seed), in this order, F = integers(8, 25), C = 1 or 3
    (integers(0, 2) -> grey, RGB), the size index integers(0, 4) into
    BATCH_SIZES and the sigma index integers(0, 4) into BATCH_SIGMAS.  The
    clip itself is noisy_clip(W, H, F, C, sigma, seed)."""
    rng = np.random.default_rng(seed)
    F = int(rng.integers(8, 25))
    C = 3 if int(rng.integers(0, 2)) else 1
    W, H = BATCH_SIZES[int(rng.integers(0, 4))]
    sigma = BATCH_SIGMAS[int(rng.integers(0, 4))]
    return W, H, F, C, sigma


def psnr(ref: np.ndarray, est: np.ndarray) -> float:
    """PSNR as VideoUtils::computePSNR (src/Utilities/LibVideoT.hpp:577-597):
    both clipped to [0, 255], squared float differences summed in double,
    RMSE through sqrtf, 20*log10f(255/RMSE)."""
    a = np.clip(ref.astype(np.float32), 0, 255)
    b = np.clip(est.astype(np.float32), 0, 255)
    d = (a - b).astype(np.float32)
    sq = (d * d).astype(np.float32).astype(np.float64)
    rmse = np.float32(np.sqrt(np.float32(sq.sum() / np.float32(a.size))))
    return float(np.float32(20.0) * np.log10(np.float32(255.0) / rmse, dtype=np.float32))
