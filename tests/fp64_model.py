"""Independent FP64 model of the VBM3D path, for pinning the FP32 oracle.

Written from the reference's algorithm (src/vbm3d.cpp) with closed-form
transforms instead of the reference's tables:
  * 2D DCT: scipy.fft.dctn / idctn with norm="ortho" (what FFTW REDFT10 /
    REDFT01 times coef_norm(_inv) compute, src/vbm3d.cpp:1843-1863);
  * Bior1.5: derived here from the published bior1.5 filter bank (3-level
    periodised DWT; analysis rows normalised to unit norm, synthesis not),
    which reproduces the reference's MATLAB tables
    (src/lib_transforms.cpp:141-148, 281-288);
  * Haar along the group: the orthonormal Haar matrix;
  * Kaiser window: Kaiser(beta = 2) rounded to 4 decimals;
  * colour: the BT.601 YUV matrices.
Only the single-thread grid (nb = 1), zero flow and kt = 1 are modelled.
Every discrete decision (candidate ranking, tau cut, hard threshold) also
reports its margin so that tests can require decisions far from a tie.
"""
from __future__ import annotations

import numpy as np
import scipy.fft

S2 = np.sqrt(2.0)
BIOR_DEC_LO = np.array([3, -3, -22, 22, 128, 128, 22, -22, -3, 3]) / (128 * S2)
BIOR_DEC_HI = np.array([0, 0, 0, 0, -1, 1, 0, 0, 0, 0]) / S2
BIOR_REC_LO = np.array([0, 0, 0, 0, 1, 1, 0, 0, 0, 0]) / S2
BIOR_REC_HI = np.array([3, 3, -22, -22, 128, -128, 22, 22, -3, -3]) / (128 * S2)


def _dwt_matrix(phase):
    def level(sig):
        n = len(sig)
        lo = np.array([sum(BIOR_DEC_LO[m] * sig[(2 * k + phase - m) % n] for m in range(10)) for k in range(n // 2)])
        hi = np.array([sum(BIOR_DEC_HI[m] * sig[(2 * k + phase - m) % n] for m in range(10)) for k in range(n // 2)])
        return lo, hi
    M = np.zeros((8, 8))
    for s in range(8):
        e = np.zeros(8)
        e[s] = 1
        a1, d1 = level(e)
        a2, d2 = level(a1)
        a3, d3 = level(a2)
        M[:, s] = np.concatenate([a3, d3, d2, d1])
    return M


def _idwt_matrix(phase):
    def level(a, d):
        n = 2 * len(a)
        out = np.zeros(n)
        for k in range(len(a)):
            for m in range(10):
                out[(2 * k + phase + m) % n] += BIOR_REC_LO[m] * a[k] + BIOR_REC_HI[m] * d[k]
        return out
    S = np.zeros((8, 8))
    for c in range(8):
        e = np.zeros(8)
        e[c] = 1
        x = level(level(level(e[0:1], e[1:2]), e[2:4]), e[4:8])
        S[:, c] = x
    return S


def bior_matrices(phase_fwd=-11, phase_inv=-12):
    """(analysis BF[j][s], synthesis BI[s][j]) of the 8-point Bior1.5 basis."""
    M = _dwt_matrix(phase_fwd)
    return M / np.linalg.norm(M, axis=1, keepdims=True), _idwt_matrix(phase_inv)


def haar_matrix(n):
    """Orthonormal Haar analysis matrix with the reference's coefficient
    order [L_J, H_J, H_{J-1} (2), ..., H_1 (n/2)] (src/lib_transforms.cpp:453-476)."""
    if n == 1:
        return np.ones((1, 1))
    h = n // 2
    lo = np.zeros((h, n))
    hi = np.zeros((h, n))
    for i in range(h):
        lo[i, 2 * i] = lo[i, 2 * i + 1] = 1 / S2
        hi[i, 2 * i], hi[i, 2 * i + 1] = 1 / S2, -1 / S2
    return np.vstack([haar_matrix(h) @ lo, hi])


def kaiser2d(k):
    if k in (4, 6, 7, 8, 12):
        w = np.kaiser(k, 2.0)
        return np.round(np.outer(w, w), 4)
    return np.ones((k, k))


YUV = np.array([[0.299, 0.587, 0.114], [-0.14713, -0.28886, 0.436], [0.615, -0.51498, -0.10001]])
YUV_INV = np.array([[1.0, 0.0, 1.13983], [1.0, -0.39465, -0.5806], [1.0, 2.03211, 0.0]])


def sigma_rows(sigma, C, ycbcr=True):
    if C == 1:
        return np.array([sigma])
    rows = [[0.299, 0.587, 0.114]]
    rows += [[0.169, 0.331, 0.5], [0.5, 0.419, 0.081]] if ycbcr else [[0.14713, 0.28886, 0.436], [0.615, 0.51498, 0.10001]]
    return np.array([np.sqrt(np.sum(np.square(r))) * sigma for r in rows])


def axis_grid(end, step):
    g = list(range(0, end + 1, step))
    if g[-1] < end:
        g.append(end)
    return g


def pow2_floor(n):
    r = 1
    while r * 2 <= n:
        r *= 2
    return r


class Margins:
    def __init__(self):
        self.rank = np.inf    # smallest gap between consecutive ranked distances that matters
        self.tau = np.inf     # smallest |distance - tau| among group candidates
        self.thr = np.inf     # smallest ||coefficient| - threshold|

    def gaps(self, d, keep):
        d = np.sort(np.asarray(d))
        upto = min(keep + 1, len(d))
        if upto >= 2:
            self.rank = min(self.rank, float(np.min(np.diff(d[:upto]))))


def block_match(vid, prm, rx, ry, rt, marg: Margins):
    """computeSimilarPatches (src/vbm3d.cpp:1082-1197) in FP64, zero flow."""
    F, C, H, W = vid.shape
    k, Nb = prm["k"], prm["Nb"]
    ref = vid[rt, :, ry:ry + k, rx:rx + k]
    seen = set()

    def local(cx, cy, ct, s):
        h = (s - 1) // 2
        lst = []
        for qy in range(max(0, cy - h), min(H - k, cy + h) + 1):
            for qx in range(max(0, cx - h), min(W - k, cx + h) + 1):
                if (qx, qy, ct) in seen:
                    continue
                seen.add((qx, qy, ct))
                d = np.sum((ref - vid[ct, :, qy:qy + k, qx:qx + k]) ** 2) / (k * k * C)
                if qx == rx and qy == ry:
                    d -= prm["d"]
                lst.append((d, (qx, qy, ct)))
        marg.gaps([e[0] for e in lst], min(Nb, len(lst)))
        lst.sort(key=lambda e: e[0])
        return lst[:Nb]

    best = local(rx, ry, rt, prm["Ns"])
    pre = post = [e[1] for e in best]
    for direction, chain in ((1, post), (-1, pre)):
        nst = (min(rt + prm["Nf"], F - 1) - rt) if direction > 0 else (rt - max(rt - prm["Nf"], 0))
        for _ in range(nst):
            fb = []
            for (x, y, t) in chain:
                fb += local(x, y, t + direction, prm["Npr"])
            marg.gaps([e[0] for e in fb], min(Nb, len(fb)))
            fb.sort(key=lambda e: e[0])
            fb = fb[:Nb]
            chain = [e[1] for e in fb]
            best += fb
    marg.gaps([e[0] for e in best], min(prm["N"], len(best)))
    best.sort(key=lambda e: e[0])
    best = best[:prm["N"]]
    d = [e[0] for e in best]
    marg.tau = min(marg.tau, min(abs(x - prm["tau"]) for x in d))
    j = len(d) - 1
    while d[j] > prm["tau"] and j > 0:
        j -= 1
    return [e[1] for e in best], pow2_floor(j + 1)


def _fwd2d(patch, prm, BF):
    if prm["T_2D"] == 1:
        return BF @ patch @ BF.T          # [vertical freq][horizontal freq]
    return scipy.fft.dctn(patch, type=2, norm="ortho")


def _inv2d(coef, prm, BI):
    if prm["T_2D"] == 1:
        return BI @ coef @ BI.T
    return scipy.fft.idctn(coef, type=2, norm="ortho")


def step(stepno, noisy, basic, prm, sig, marg: Margins):
    """One step (nb = 1): returns (num, den) in FP64."""
    F, C, H, W = noisy.shape
    k = prm["k"]
    BF, BI = bior_matrices()
    kai = kaiser2d(k)
    num = np.zeros(noisy.shape)
    den = np.zeros(noisy.shape)
    match = noisy if stepno == 1 else basic
    for rt in range(F):
        for ry in axis_grid(H - k, prm["p"]):
            for rx in axis_grid(W - k, prm["p"]):
                members, nsx = block_match(match, prm, rx, ry, rt, marg)
                members = members[:nsx]
                Hm = haar_matrix(nsx)
                for c in range(C):
                    z = np.stack([_fwd2d(noisy[t, c, y:y + k, x:x + k], prm, BF) for (x, y, t) in members])
                    zh = np.tensordot(Hm, z, axes=(1, 0))  # along the group
                    if stepno == 1:
                        thr = prm["lambda3D"] * sig[c]
                        keep = np.abs(zh) > thr
                        keep[0, 0, 0] = True
                        marg.thr = min(marg.thr, float(np.min(np.abs(np.abs(zh) - thr))))
                        est_h = np.where(keep, zh, 0.0)
                        w = 1.0 / (sig[c] ** 2 * keep.sum())
                    else:
                        e = np.stack([_fwd2d(basic[t, c, y:y + k, x:x + k], prm, BF) for (x, y, t) in members])
                        eh = np.tensordot(Hm, e, axes=(1, 0))
                        v = eh ** 2 / (eh ** 2 + sig[c] ** 2)
                        est_h = zh * v
                        w = 1.0 / (sig[c] ** 2 * np.sum(v ** 2))
                    est = np.tensordot(Hm.T, est_h, axes=(1, 0))
                    for n, (x, y, t) in enumerate(members):
                        num[t, c, y:y + k, x:x + k] += kai * w * _inv2d(est[n], prm, BI)
                        den[t, c, y:y + k, x:x + k] += kai * w
    return num, den


def run(sigma, noisy, p1, p2, color=True, marg: Margins | None = None):
    """run_vbm3d (nb = 1, zero flow) in FP64.  Returns (basic, final)."""
    marg = marg or Margins()
    v = np.asarray(noisy, np.float64)
    F, C, H, W = v.shape
    if color and C == 3:
        v = np.einsum("ij,fjhw->fihw", YUV, v)
    sig = sigma_rows(sigma, C, ycbcr=bool(color))
    num, den = step(1, v, None, p1, sig, marg)
    basic = np.where(den == 0, v, num / np.where(den == 0, 1, den))
    num, den = step(2, v, basic, p2, sig, marg)
    final = np.where(den == 0, basic, num / np.where(den == 0, 1, den))
    if color and C == 3:
        basic = np.einsum("ij,fjhw->fihw", YUV_INV, basic)
        final = np.einsum("ij,fjhw->fihw", YUV_INV, final)
    return basic, final, marg
