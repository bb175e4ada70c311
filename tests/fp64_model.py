"""Independent FP64 model of the VBM3D path, for pinning the FP32 oracle.

Written from the reference's algorithm (src/vbm3d.cpp) with closed-form
transforms instead of the reference's tables:
  * 2D DCT: scipy.fft.dctn / idctn with norm="ortho" (what FFTW REDFT10 /
    REDFT01 times coef_norm(_inv) compute, src/vbm3d.cpp:1843-1863);
  * temporal DCT (kt > 1): scipy.fft.dct / idct along the patch's frames
    with norm="ortho" (REDFT10 times 1/sqrt(2 kt) and 1/sqrt(2) on the DC,
    src/vbm3d.cpp:1866-1956);
  * Bior1.5: derived here from the published bior1.5 filter bank (3-level
    periodised DWT; analysis rows normalised to unit norm, synthesis not),
    which reproduces the reference's MATLAB tables
    (src/lib_transforms.cpp:141-148, 281-288);
  * Haar along the group: the orthonormal Haar matrix;
  * Hadamard along the group: scipy.linalg.hadamard (Sylvester) with its
    rows in bit-reversed order, which is the ordering the reference's
    recursion produces (src/lib_transforms.cpp:410-440);
  * Kaiser window: Kaiser(beta = 2) rounded to 4 decimals;
  * colour: the BT.601 YUV matrices.
Modelled: any reference grid (nb = 1, or the nb-crop tiling with the
crop-order merge and the den == 0 fallbacks, src/vbm3d.cpp:213-354),
forward and backward flows in the predictive search (x projected first,
y read at the projected x, :1128-1131, 1152-1155), kt >= 1 with or without
motion-compensated trajectories (patch_trajectory, :378-425), DCT or Bior
2D transforms and Haar or Hadamard along the group.
Every discrete decision (candidate ranking, tau cut, hard threshold) also
reports its margin so that tests can require decisions far from a tie.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft
import scipy.linalg

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


def _bitrev(i, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def hadamard_matrix(n):
    """The unnormalised Walsh-Hadamard matrix in the reference's output
    order: Sylvester's matrix with row i taken from row bitrev(i)."""
    if n == 1:
        return np.ones((1, 1))
    bits = n.bit_length() - 1
    Hs = scipy.linalg.hadamard(n).astype(np.float64)
    return Hs[[_bitrev(i, bits) for i in range(n)]]


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


def axis_grid(end, step, begin=0):
    g = list(range(begin, end + 1, step))
    if g[-1] < end:
        g.append(end)
    return g


def pow2_floor(n):
    r = 1
    while r * 2 <= n:
        r *= 2
    return r


def round_half_away(v):
    """std::round: halves go away from zero (numpy's round goes to even)."""
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


# Tiles of the reference's OpenMP grid (determineFactor, listed for the
# thread counts it accepts): nb -> (columns, rows).
TILES = {2: (2, 1), 4: (2, 2), 8: (4, 2), 16: (8, 2), 32: (16, 2)}


def crop_grids(W, H, k, p, nb, n=16):
    """Per crop (in processing order): absolute reference rows and columns
    (src/vbm3d.cpp:471-489 over subDivideTight's crops,
    src/Utilities/LibVideoT.hpp:903-961)."""
    if nb == 1:
        return [(axis_grid(H - k, p), axis_grid(W - k, p))]
    nW, nH = TILES[nb]
    w, h = -(-W // nW), -(-H // nH)
    out = []
    for ty in range(nH):
        for tx in range(nW):
            ox, oy = max(0, tx * w - 2 * n), max(0, ty * h - 2 * n)
            ex, ey = min(W, (tx + 1) * w + 2 * n), min(H, (ty + 1) * h + 2 * n)
            rows = axis_grid(ey - oy - k if ey == H else ey - oy - k - n, p, 0 if oy == 0 else n)
            cols = axis_grid(ex - ox - k if ex == W else ex - ox - k - n, p, 0 if ox == 0 else n)
            out.append(([oy + r for r in rows], [ox + c for c in cols]))
    return out


class Margins:
    def __init__(self):
        self.rank = np.inf    # smallest gap between consecutive ranked distances that matters
        self.tau = np.inf     # smallest |distance - tau| among group candidates
        self.thr = np.inf     # smallest ||coefficient| - threshold|
        self.traj = np.inf    # smallest distance of a rounded flow sum from a half integer

    def gaps(self, d, keep):
        d = np.sort(np.asarray(d))
        upto = min(keep + 1, len(d))
        if upto >= 2:
            self.rank = min(self.rank, float(np.min(np.diff(d[:upto]))))


def _flow(fl, x, y, t, c):
    return 0.0 if fl is None else float(fl[t, c, y, x])


def trajectory(px, py, pt, prm, ff, W, H, marg=None):
    """patch_trajectory (src/vbm3d.cpp:378-425) with FLOAT_TRAJECTORIES:
    the spatial position of each of the kt slices.  With mc, the next
    position is the running sum of the flow read at the current position,
    rounded half away from zero and clamped to the frame."""
    k, kt = prm["k"], prm["kt"]
    out = []
    fx, fy = float(px), float(py)
    for ht in range(kt):
        out.append((px, py, pt + ht))
        if prm["mc"] and ht < kt - 1:
            fx += _flow(ff, px + k // 2, py + k // 2, pt + ht, 0)
            fy += _flow(ff, px + k // 2, py + k // 2, pt + ht, 1)
            if marg is not None:
                marg.traj = min(marg.traj, abs(abs(fx - math.floor(fx)) - 0.5), abs(abs(fy - math.floor(fy)) - 0.5))
            px = min(W - k, max(0, round_half_away(fx)))
            py = min(H - k, max(0, round_half_away(fy)))
    return out


def patch(vid, x, y, t, prm, ff):
    """The k x k x kt patch at (x, y, t) along its trajectory: [kt][C][k][k]."""
    F, C, H, W = vid.shape
    k = prm["k"]
    return np.stack([vid[tt, :, yy:yy + k, xx:xx + k] for (xx, yy, tt) in trajectory(x, y, t, prm, ff, W, H)])


def block_match(vid, prm, rx, ry, rt, marg: Margins, ff=None, bf=None, y_at_old_x=False):
    """computeSimilarPatches (src/vbm3d.cpp:1082-1197) in FP64.
    y_at_old_x plants a slip (the flow's y read before x is projected) for
    tests that must tell the two rules apart."""
    F, C, H, W = vid.shape
    k, kt, Nb = prm["k"], prm["kt"], prm["Nb"]
    ref = patch(vid, rx, ry, rt, prm, ff)
    seen = set()

    def local(cx, cy, ct, s):
        h = (s - 1) // 2
        lst = []
        for qy in range(max(0, cy - h), min(H - k, cy + h) + 1):
            for qx in range(max(0, cx - h), min(W - k, cx + h) + 1):
                if (qx, qy, ct) in seen:
                    continue
                seen.add((qx, qy, ct))
                d = np.sum((ref - patch(vid, qx, qy, ct, prm, ff)) ** 2) / (k * k * kt * C)
                if qx == rx and qy == ry:
                    d -= prm["d"]
                lst.append((d, (qx, qy, ct)))
        marg.gaps([e[0] for e in lst], min(Nb, len(lst)))
        lst.sort(key=lambda e: e[0])
        return lst[:Nb]

    def project(m, direction):
        x, y, t = m
        fl, ft = (ff, t) if direction > 0 else (bf, t - 1)
        nx = min(max(x + round_half_away(_flow(fl, x + k // 2, y + k // 2, ft, 0)), 0), W - k)
        xr = x if y_at_old_x else nx
        y = min(max(y + round_half_away(_flow(fl, xr + k // 2, y + k // 2, ft, 1)), 0), H - k)
        return nx, y, t + direction

    best = local(rx, ry, rt, prm["Ns"])
    first = [e[1] for e in best]
    for direction in (1, -1):
        chain = list(first)
        nst = (min(rt + prm["Nf"], F - kt) - rt) if direction > 0 else (rt - max(rt - prm["Nf"], 0))
        for _ in range(nst):
            fb = []
            for m in chain:
                x, y, t = project(m, direction)
                fb += local(x, y, t, prm["Npr"])
            marg.gaps([e[0] for e in fb], min(Nb, len(fb)))
            fb.sort(key=lambda e: e[0])
            fb = fb[:Nb]
            # a frame with fewer than Nb picks keeps the older matches in
            # the remaining places of the chain (:1140-1145, :1164-1168)
            for i, e in enumerate(fb):
                chain[i] = e[1]
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


def _fwd2d(p2, prm, BF):
    if prm["T_2D"] == 1:
        return BF @ p2 @ BF.T          # [vertical freq][horizontal freq]
    return scipy.fft.dctn(p2, type=2, norm="ortho")


def _inv2d(coef, prm, BI):
    if prm["T_2D"] == 1:
        return BI @ coef @ BI.T
    return scipy.fft.idctn(coef, type=2, norm="ortho")


def _fwd3d(p3, prm, BF):
    """[kt][k][k] patch -> coefficients: 2D per slice, then the temporal
    DCT when kt > 1 (not with Bior, src/vbm3d.cpp:573, 818)."""
    c = np.stack([_fwd2d(s, prm, BF) for s in p3])
    if prm["kt"] > 1 and prm["T_2D"] != 1:
        c = scipy.fft.dct(c, type=2, norm="ortho", axis=0)
    return c


def _inv3d(c, prm, BI):
    if prm["kt"] > 1 and prm["T_2D"] != 1:
        c = scipy.fft.idct(c, type=2, norm="ortho", axis=0)
    return np.stack([_inv2d(s, prm, BI) for s in c])


def step(stepno, noisy, basic, prm, sig, marg: Margins, ff=None, bf=None, nb=1):
    """One step: returns (num, den) in FP64, the crops' sums merged."""
    F, C, H, W = noisy.shape
    k, kt = prm["k"], prm["kt"]
    BF, BI = bior_matrices()
    kai = kaiser2d(k)
    num = np.zeros(noisy.shape)
    den = np.zeros(noisy.shape)
    match = noisy if stepno == 1 else basic
    had = prm["T_3D"] == 2
    for rows, cols in crop_grids(W, H, k, prm["p"], nb, prm.get("n", 16)):
        for rt in range(F - kt + 1):
            for ry in rows:
                for rx in cols:
                    members, nsx = block_match(match, prm, rx, ry, rt, marg, ff, bf)
                    members = members[:nsx]
                    T3 = hadamard_matrix(nsx) if had else haar_matrix(nsx)
                    T3i = T3 / nsx if had else T3.T
                    for c in range(C):
                        # [n][kt][k][k]
                        z = np.stack([_fwd3d(patch(noisy, x, y, t, prm, ff)[:, c], prm, BF) for (x, y, t) in members])
                        zh = np.tensordot(T3, z, axes=(1, 0))  # along the group
                        if stepno == 1:
                            thr = prm["lambda3D"] * sig[c] * (np.sqrt(nsx) if had else 1.0)
                            keep = np.abs(zh) > thr
                            keep[0, 0, 0, 0] = True
                            marg.thr = min(marg.thr, float(np.min(np.abs(np.abs(zh) - thr))))
                            est_h = np.where(keep, zh, 0.0)
                            w = 1.0 / (sig[c] ** 2 * keep.sum())
                        else:
                            e = np.stack([_fwd3d(patch(basic, x, y, t, prm, ff)[:, c], prm, BF) for (x, y, t) in members])
                            eh = np.tensordot(T3, e, axes=(1, 0))
                            e2 = eh ** 2 / nsx if had else eh ** 2
                            v = e2 / (e2 + sig[c] ** 2)
                            est_h = zh * v
                            sv = np.sum(v ** 2)
                            w = (1.0 / (sig[c] ** 2 * sv) if sv > 0 else 1.0) if had else 1.0 / (sig[c] ** 2 * sv)
                        est = np.tensordot(T3i, est_h, axes=(1, 0))
                        for n, (x, y, t) in enumerate(members):
                            sl = _inv3d(est[n], prm, BI)
                            for ht, (xx, yy, tt) in enumerate(trajectory(x, y, t, prm, ff, W, H)):
                                num[tt, c, yy:yy + k, xx:xx + k] += kai * w * sl[ht]
                                den[tt, c, yy:yy + k, xx:xx + k] += kai * w
    return num, den


def run(sigma, noisy, p1, p2, color=True, marg: Margins | None = None, ff=None, bf=None, nb=1):
    """run_vbm3d in FP64.  Returns (basic, final, margins)."""
    marg = marg or Margins()
    v = np.asarray(noisy, np.float64)
    F, C, H, W = v.shape
    if color and C == 3:
        v = np.einsum("ij,fjhw->fihw", YUV, v)
    sig = sigma_rows(sigma, C, ycbcr=bool(color))
    num, den = step(1, v, None, p1, sig, marg, ff, bf, nb)
    basic = np.where(den == 0, v, num / np.where(den == 0, 1, den))
    num, den = step(2, v, basic, p2, sig, marg, ff, bf, nb)
    # den == 0: the basic estimate with one thread (:200), the noisy video
    # with the tiled grid (:342-343)
    final = np.where(den == 0, basic if nb == 1 else v, num / np.where(den == 0, 1, den))
    if color and C == 3:
        basic = np.einsum("ij,fjhw->fihw", YUV_INV, basic)
        final = np.einsum("ij,fjhw->fihw", YUV_INV, final)
    return basic, final, marg
