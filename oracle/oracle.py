"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NONE = -1


class OrcParams(C.Structure):
    """POD mirror of Parameters (src/parameters.h:21-58)."""
    _fields_ = [("k", C.c_uint), ("kt", C.c_uint), ("p", C.c_uint), ("N", C.c_uint),
                ("tau", C.c_float), ("d", C.c_float),
                ("Nf", C.c_uint), ("Ns", C.c_uint), ("Npr", C.c_uint), ("Nb", C.c_uint),
                ("T_2D", C.c_uint), ("T_3D", C.c_uint), ("lambda3D", C.c_float),
                ("mc", C.c_int), ("n", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def as_orc(prm) -> OrcParams:
    """Copy any parameter record with the same field names into OrcParams."""
    if isinstance(prm, OrcParams):
        return prm
    o = OrcParams()
    for f, _ in OrcParams._fields_:
        setattr(o, f, getattr(prm, f))
    return o


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        pp = C.POINTER(OrcParams)
        L.orc_init_params_1.argtypes = [pp] + [C.c_int] * 9 + [C.c_float, C.c_float, C.c_uint, C.c_uint, C.c_float]
        L.orc_init_params_2.argtypes = [pp] + [C.c_int] * 9 + [C.c_float, C.c_uint, C.c_uint, C.c_float]
        L.orc_grid.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int, C.c_int, up, up, up]
        L.orc_grid.restype = C.c_long
        L.orc_block_match.argtypes = [fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, fp, fp, pp,
                                      C.c_long, up, up, fp, up, up]
        L.orc_run_vbm3d.argtypes = [C.c_float, fp, fp, fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                    pp, pp, C.c_int, C.c_int, C.c_int, fp, fp]
        L.orc_step.argtypes = [C.c_int, C.c_float, fp, fp, fp, fp, C.c_uint, C.c_uint, C.c_uint,
                               C.c_uint, pp, C.c_int, C.c_int, C.c_int, fp, fp]
        L.orc_step_fp64.argtypes = [C.c_int, C.c_float, fp, fp, fp, fp, C.c_uint, C.c_uint, C.c_uint,
                                    C.c_uint, pp, C.c_int, dp, dp]
        L.orc_step2_adjoint_fp64.argtypes = [C.c_float, fp, fp, fp, fp, C.c_uint, C.c_uint, C.c_uint,
                                             C.c_uint, pp, C.c_int, C.c_int, dp, dp, dp]
        L.orc_bior_fwd.argtypes = [fp, fp]
        L.orc_bior_inv.argtypes = [fp]
        L.orc_dct_fwd.argtypes = [C.c_uint, fp, fp]
        L.orc_dct_inv.argtypes = [C.c_uint, fp]
        L.orc_temporal_fwd.argtypes = [C.c_uint, fp]
        L.orc_temporal_inv.argtypes = [C.c_uint, fp]
        L.orc_haar_fwd.argtypes = [fp, C.c_uint]
        L.orc_haar_inv.argtypes = [fp, C.c_uint]
        L.orc_hadamard.argtypes = [fp, C.c_uint]
        L.orc_kaiser.argtypes = [C.c_uint, fp]
        L.orc_dct_norm.argtypes = [C.c_uint, fp, fp]
        L.orc_sigma_table.argtypes = [C.c_float, C.c_uint, C.c_uint, fp]
        L.orc_color.argtypes = [fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int]
        L.orc_partial_sort.argtypes = [fp, up, C.c_uint, C.c_uint]
        L.orc_psnr.argtypes = [fp, fp, C.c_size_t, dp, dp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_stats.argtypes = [C.c_int, C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def params_1(sigma, k=NONE, kt=NONE, Nf=NONE, Ns=NONE, Npr=NONE, Nb=NONE, p=NONE, N=NONE,
             d=NONE, tau=NONE, lambda3D=NONE, T_2D=NONE, T_3D=NONE) -> OrcParams:
    o = OrcParams()
    lib().orc_init_params_1(C.byref(o), k, kt, Nf, Ns, Npr, Nb, p, N, d, float(tau), float(lambda3D),
                            T_2D & 0xFFFFFFFF, T_3D & 0xFFFFFFFF, float(sigma))
    return o


def params_2(sigma, k=NONE, kt=NONE, Nf=NONE, Ns=NONE, Npr=NONE, Nb=NONE, p=NONE, N=NONE,
             d=NONE, tau=NONE, T_2D=NONE, T_3D=NONE) -> OrcParams:
    o = OrcParams()
    lib().orc_init_params_2(C.byref(o), k, kt, Nf, Ns, Npr, Nb, p, N, d, float(tau),
                            T_2D & 0xFFFFFFFF, T_3D & 0xFFFFFFFF, float(sigma))
    return o


def grid(W, H, k, p, n=16, nb=1):
    cnt = lib().orc_grid(W, H, k, p, n, nb, None, None, None)
    if cnt < 0:
        raise ValueError("invalid grid")
    xs = np.empty(cnt, np.uint32)
    ys = np.empty(cnt, np.uint32)
    cr = np.empty(cnt, np.uint32)
    lib().orc_grid(W, H, k, p, n, nb, _u(xs), _u(ys), _u(cr))
    return xs, ys, cr


def block_match(vid, prm, ref_idx, fflow=None, bflow=None):
    prm = as_orc(prm)
    vid = np.ascontiguousarray(vid, np.float32)
    F, Cc, H, W = vid.shape
    ref_idx = np.ascontiguousarray(ref_idx, np.uint32)
    n = len(ref_idx)
    N = prm.N
    idx = np.empty((n, N), np.uint32)
    dist = np.empty((n, N), np.float32)
    nsim = np.empty(n, np.uint32)
    nsx = np.empty(n, np.uint32)
    ff = None if fflow is None else _f(np.ascontiguousarray(fflow, np.float32))
    bf = None if bflow is None else _f(np.ascontiguousarray(bflow, np.float32))
    _chk(lib().orc_block_match(_f(vid), W, H, F, Cc, ff, bf, C.byref(prm), n, _u(ref_idx),
                               _u(idx), _f(dist), _u(nsim), _u(nsx)))
    return idx, dist, nsim, nsx


def run_vbm3d(sigma, noisy, p1, p2, color_space=True, nb=1, nthreads=0, basic_in=None,
              fflow=None, bflow=None):
    """Returns (basic, denoised, noisy_after) -- noisy_after is the input
    after the reference's forward+inverse colour round trip."""
    p1, p2 = as_orc(p1), as_orc(p2)
    nz = np.array(noisy, np.float32, copy=True, order="C")
    F, Cc, H, W = nz.shape
    basic = np.zeros_like(nz) if basic_in is None else np.array(basic_in, np.float32, copy=True, order="C")
    den = np.empty_like(nz)
    ff = None if fflow is None else _f(np.ascontiguousarray(fflow, np.float32))
    bf = None if bflow is None else _f(np.ascontiguousarray(bflow, np.float32))
    _chk(lib().orc_run_vbm3d(float(sigma), _f(nz), ff, bf, W, H, F, Cc, C.byref(p1), C.byref(p2),
                             int(bool(color_space)), nb, nthreads, _f(basic), _f(den)))
    return basic, den, nz


def step(stepno, sigma, noisy, basic, prm, color_space=True, nb=1, nthreads=0):
    prm = as_orc(prm)
    nz = np.ascontiguousarray(noisy, np.float32)
    F, Cc, H, W = nz.shape
    bs = None if basic is None else _f(np.ascontiguousarray(basic, np.float32))
    num = np.empty_like(nz)
    den = np.empty_like(nz)
    _chk(lib().orc_step(stepno, float(sigma), _f(nz), bs, None, None, W, H, F, Cc, C.byref(prm),
                        int(bool(color_space)), nb, nthreads, _f(num), _f(den)))
    return num, den


STAT_NAMES = ("refs", "candidates", "members", "near_tie_refs", "near_tau_refs",
              "near_threshold_coefs", "threshold_coefs")


def last_stats(stepno):
    """Counters of the last step / run for step 1 or 2 (orc_last_stats)."""
    out = (C.c_longlong * 7)()
    lib().orc_last_stats(stepno, out)
    return dict(zip(STAT_NAMES, list(out)))


def step_fp64(stepno, sigma, noisy, basic, prm, color_space=True):
    prm = as_orc(prm)
    nz = np.ascontiguousarray(noisy, np.float32)
    F, Cc, H, W = nz.shape
    bs = None if basic is None else _f(np.ascontiguousarray(basic, np.float32))
    num = np.empty(nz.shape, np.float64)
    den = np.empty(nz.shape, np.float64)
    _chk(lib().orc_step_fp64(stepno, float(sigma), _f(nz), bs, None, None, W, H, F, Cc, C.byref(prm),
                             int(bool(color_space)), _d(num), _d(den)))
    return num, den


def step2_adjoint_fp64(sigma, noisy, basic, prm, grad_out, color_space=True, nb=1):
    """(grad_noisy, final) in float64: the transpose of step 2's Jacobian
    (groups, Wiener gains and weights frozen by the basic estimate) applied
    to grad_out, and the FP64 forward value."""
    prm = as_orc(prm)
    nz = np.ascontiguousarray(noisy, np.float32)
    bs = np.ascontiguousarray(basic, np.float32)
    F, Cc, H, W = nz.shape
    g = np.ascontiguousarray(grad_out, np.float64)
    gz = np.empty(nz.shape, np.float64)
    fin = np.empty(nz.shape, np.float64)
    _chk(lib().orc_step2_adjoint_fp64(float(sigma), _f(nz), _f(bs), None, None, W, H, F, Cc, C.byref(prm),
                                      int(bool(color_space)), nb, _d(g), _d(gz), _d(fin)))
    return gz, fin


def bior_fwd(patch):
    x = np.ascontiguousarray(patch, np.float32).reshape(64)
    o = np.empty(64, np.float32)
    lib().orc_bior_fwd(_f(x), _f(o))
    return o


def bior_inv(coef):
    x = np.array(coef, np.float32, copy=True).reshape(64)
    lib().orc_bior_inv(_f(x))
    return x


def dct_fwd(patch):
    k = patch.shape[0]
    x = np.ascontiguousarray(patch, np.float32).reshape(-1)
    o = np.empty(k * k, np.float32)
    lib().orc_dct_fwd(k, _f(x), _f(o))
    return o


def dct_inv(coef, k):
    x = np.array(coef, np.float32, copy=True).reshape(-1)
    lib().orc_dct_inv(k, _f(x))
    return x


def temporal_fwd(v):
    x = np.array(v, np.float32, copy=True)
    lib().orc_temporal_fwd(len(x), _f(x))
    return x


def temporal_inv(v):
    x = np.array(v, np.float32, copy=True)
    lib().orc_temporal_inv(len(x), _f(x))
    return x


def haar_fwd(v):
    x = np.array(v, np.float32, copy=True)
    lib().orc_haar_fwd(_f(x), len(x))
    return x


def haar_inv(v):
    x = np.array(v, np.float32, copy=True)
    lib().orc_haar_inv(_f(x), len(x))
    return x


def hadamard(v):
    x = np.array(v, np.float32, copy=True)
    lib().orc_hadamard(_f(x), len(x))
    return x


def kaiser(k):
    o = np.empty(k * k, np.float32)
    lib().orc_kaiser(k, _f(o))
    return o.reshape(k, k)


def dct_norm(k):
    a = np.empty(k * k, np.float32)
    b = np.empty(k * k, np.float32)
    lib().orc_dct_norm(k, _f(a), _f(b))
    return a.reshape(k, k), b.reshape(k, k)


def sigma_table(sigma, C_, cs):
    o = np.zeros(3, np.float32)
    rc = lib().orc_sigma_table(float(sigma), C_, cs, _f(o))
    if rc:
        raise ValueError("sigma table")
    return o[:C_]


def color(vid, forward):
    x = np.array(vid, np.float32, copy=True, order="C")
    F, Cc, H, W = x.shape
    lib().orc_color(_f(x), W, H, F, Cc, int(bool(forward)))
    return x


def partial_sort(dist, idx, mid):
    d = np.array(dist, np.float32, copy=True)
    i = np.array(idx, np.uint32, copy=True)
    lib().orc_partial_sort(_f(d), _u(i), len(d), mid)
    return d, i


def psnr(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    p = C.c_double()
    r = C.c_double()
    lib().orc_psnr(_f(a), _f(b), a.size, C.byref(p), C.byref(r))
    return p.value, r.value
