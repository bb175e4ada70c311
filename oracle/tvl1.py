"""ctypes bindings of the TV-L1 flow checkers.

TEST INFRASTRUCTURE ONLY (imported by tests/): the oracle's restatement
(liboracle.so, tvl1_oracle.cpp) and, when it has been built, the
reference's own library compiled from its sources (_ref/libtvl1_ref.so,
``make -C oracle ref``).  Images are float32 [ny][nx]; flows come back as
two float32 [ny][nx] planes (u: x component, v: y component).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import oracle as _orc

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(HERE, "_ref", "libtvl1_ref.so")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _f(a):
    return a.ctypes.data_as(_fp)


def _orc_lib():
    L = _orc.lib()
    if not getattr(L, "_tvl1_bound", False):
        L.orc_tvl1_scale.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                     C.c_int, C.c_float, _ip]
        L.orc_tvl1_multiscale.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                          C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, _ip]
        L.orc_tvl1_normalize.argtypes = [_fp, _fp, _fp, _fp, C.c_int]
        L.orc_tvl1_nscales.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int]
        L.orc_tvl1_gaussian.argtypes = [_fp, C.c_int, C.c_int, C.c_double]
        L.orc_tvl1_bicubic.argtypes = [_fp, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int]
        L.orc_tvl1_bicubic.restype = C.c_float
        L.orc_tvl1_zoom_out.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_float]
        L.orc_tvl1_zoom_in.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_tvl1_divergence.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L.orc_tvl1_centered_gradient.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L.orc_tvl1_forward_gradient.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L._tvl1_bound = True
    return L


_ref = None


def have_ref() -> bool:
    return os.path.exists(REF_PATH)


def _ref_lib():
    """The reference's own tvl1flow_lib.c (its exported names)."""
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.Dual_TVL1_optic_flow.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           C.c_int, C.c_float, C.c_bool]
        L.Dual_TVL1_optic_flow_multiscale.argtypes = [_fp, _fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float,
                                                      C.c_float, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float,
                                                      C.c_bool]
        L.gaussian.argtypes = [_fp, C.c_int, C.c_int, C.c_double]
        L.bicubic_interpolation_at.argtypes = [_fp, C.c_float, C.c_float, C.c_int, C.c_int, C.c_bool]
        L.bicubic_interpolation_at.restype = C.c_float
        L.zoom_out.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_float]
        L.zoom_in.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.divergence.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L.centered_gradient.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L.forward_gradient.argtypes = [_fp, _fp, _fp, C.c_int, C.c_int]
        L.image_normalization.argtypes = [_fp, _fp, _fp, _fp, C.c_int]
        _ref = L
    return _ref


def _c32(a):
    return np.ascontiguousarray(a, np.float32)


def nscales(nx, ny, zfactor=0.5, requested=100):
    return _orc_lib().orc_tvl1_nscales(nx, ny, zfactor, requested)


def flow(I0, I1, tau=0.25, lam=0.15, theta=0.3, nscales=100, fscale=0, zfactor=0.5, warps=5, epsilon=0.01,
         cap_scales=True, ref=False):
    """The multiscale flow from I0 to I1, as the reference's command line
    computes it (tvl1flow/main.c:139-160) when cap_scales.  ref=True runs
    the reference's own library instead of the restatement.  Returns
    (u, v, iterations per solved level and warp) -- the counts only from
    the restatement."""
    I0, I1 = _c32(I0), _c32(I1)
    ny, nx = I0.shape
    if cap_scales:
        nscales = min(nscales, _orc_lib().orc_tvl1_nscales(nx, ny, zfactor, nscales))
        fscale = min(fscale, nscales)
    u = np.zeros((ny, nx), np.float32)
    v = np.zeros((ny, nx), np.float32)
    if ref:
        _ref_lib().Dual_TVL1_optic_flow_multiscale(_f(I0.copy()), _f(I1.copy()), _f(u), _f(v), nx, ny, tau, lam,
                                                   theta, nscales, fscale, zfactor, warps, epsilon, False)
        return u, v, None
    iters = np.zeros(max(nscales, 1) * warps, np.int32)
    _orc_lib().orc_tvl1_multiscale(_f(I0), _f(I1), _f(u), _f(v), nx, ny, tau, lam, theta, nscales, fscale,
                                   zfactor, warps, epsilon, iters.ctypes.data_as(_ip))
    return u, v, iters[: (nscales - fscale) * warps].reshape(nscales - fscale, warps)


def scale(I0, I1, u, v, tau=0.25, lam=0.15, theta=0.3, warps=5, epsilon=0.01, ref=False):
    """One level (Dual_TVL1_optic_flow) from the initial flow (u, v);
    returns the updated (u, v) and the iteration count per warp."""
    I0, I1 = _c32(I0), _c32(I1)
    ny, nx = I0.shape
    u, v = _c32(u).copy(), _c32(v).copy()
    if ref:
        _ref_lib().Dual_TVL1_optic_flow(_f(I0), _f(I1), _f(u), _f(v), nx, ny, tau, lam, theta, warps, epsilon,
                                        False)
        return u, v, None
    it = np.zeros(warps, np.int32)
    _orc_lib().orc_tvl1_scale(_f(I0), _f(I1), _f(u), _f(v), nx, ny, tau, lam, theta, warps, epsilon,
                              it.ctypes.data_as(_ip))
    return u, v, it


def gaussian(I, sigma, ref=False):
    I = _c32(I).copy()
    ny, nx = I.shape
    (_ref_lib().gaussian if ref else _orc_lib().orc_tvl1_gaussian)(_f(I), nx, ny, sigma)
    return I


def bicubic(I, uu, vv, border_out, ref=False):
    I = _c32(I)
    ny, nx = I.shape
    if ref:
        return _ref_lib().bicubic_interpolation_at(_f(I), uu, vv, nx, ny, bool(border_out))
    return _orc_lib().orc_tvl1_bicubic(_f(I), uu, vv, nx, ny, int(bool(border_out)))


def zoom_out(I, factor, ref=False):
    I = _c32(I)
    ny, nx = I.shape
    nxx, nyy = int(np.float32(nx) * np.float32(factor) + 0.5), int(np.float32(ny) * np.float32(factor) + 0.5)
    out = np.zeros((nyy, nxx), np.float32)
    (_ref_lib().zoom_out if ref else _orc_lib().orc_tvl1_zoom_out)(_f(I), _f(out), nx, ny, factor)
    return out


def zoom_in(I, nxx, nyy, ref=False):
    I = _c32(I)
    ny, nx = I.shape
    out = np.zeros((nyy, nxx), np.float32)
    (_ref_lib().zoom_in if ref else _orc_lib().orc_tvl1_zoom_in)(_f(I), _f(out), nx, ny, nxx, nyy)
    return out


def divergence(v1, v2, ref=False):
    v1, v2 = _c32(v1), _c32(v2)
    ny, nx = v1.shape
    out = np.zeros_like(v1)
    (_ref_lib().divergence if ref else _orc_lib().orc_tvl1_divergence)(_f(v1), _f(v2), _f(out), nx, ny)
    return out


def centered_gradient(I, ref=False):
    I = _c32(I)
    ny, nx = I.shape
    dx, dy = np.zeros_like(I), np.zeros_like(I)
    (_ref_lib().centered_gradient if ref else _orc_lib().orc_tvl1_centered_gradient)(_f(I), _f(dx), _f(dy), nx, ny)
    return dx, dy


def forward_gradient(I, ref=False):
    I = _c32(I)
    ny, nx = I.shape
    dx, dy = np.zeros_like(I), np.zeros_like(I)
    (_ref_lib().forward_gradient if ref else _orc_lib().orc_tvl1_forward_gradient)(_f(I), _f(dx), _f(dy), nx, ny)
    return dx, dy


def normalize(I0, I1, ref=False):
    I0, I1 = _c32(I0), _c32(I1)
    a, b = np.zeros_like(I0), np.zeros_like(I1)
    (_ref_lib().image_normalization if ref else _orc_lib().orc_tvl1_normalize)(_f(I0), _f(I1), _f(a), _f(b),
                                                                               I0.size)
    return a, b
