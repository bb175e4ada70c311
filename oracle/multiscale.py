"""ctypes binding of the multiscale pyramid oracle (multiscale_oracle.cpp).

TEST INFRASTRUCTURE ONLY (imported by tests/).  Images are planar float32
[c][rows][cols]; a pyramid is returned as a list of levels, level i of
size rows / 2^i x cols / 2^i (integer halving).  type: 0 DCT, 1 Gaussian,
2 Lanczos.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as _orc

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _lib():
    L = _orc.lib()
    if not getattr(L, "_pyr_bound", False):
        L.orc_pyr_sizes.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
        L.orc_pyr_decompose.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp]
        L.orc_pyr_recompose.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _fp]
        L.orc_pyr_merge_coarse.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, C.c_int, C.c_float, _fp]
        for f in ("orc_pyr_dct2", "orc_pyr_idct2"):
            getattr(L, f).argtypes = [_fp, C.c_int, C.c_int, C.c_int]
        L.orc_pyr_gblur.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_float]
        L.orc_pyr_down.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, _fp]
        L.orc_pyr_up.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp]
        L._pyr_bound = True
    return L


def _f(a):
    return a.ctypes.data_as(_fp)


def _img(a):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim == 2:
        a = a[None]
    return a


def sizes(rows, cols, levels):
    lr = np.zeros(levels, np.int32)
    lc = np.zeros(levels, np.int32)
    _lib().orc_pyr_sizes(rows, cols, levels, lr.ctypes.data_as(_ip), lc.ctypes.data_as(_ip))
    return list(zip(lr.tolist(), lc.tolist()))


def decompose(img, levels, type_=0):
    img = _img(img)
    ch, rows, cols = img.shape
    sz = sizes(rows, cols, levels)
    out = np.zeros(sum(ch * r * c for r, c in sz), np.float32)
    _lib().orc_pyr_decompose(_f(img), rows, cols, ch, levels, type_, _f(out))
    res, o = [], 0
    for r, c in sz:
        res.append(out[o:o + ch * r * c].reshape(ch, r, c).copy())
        o += ch * r * c
    return res


def recompose(levels, type_=0, factor=0.8):
    ch, rows, cols = levels[0].shape
    flat = np.concatenate([np.ascontiguousarray(l, np.float32).ravel() for l in levels])
    out = np.zeros((ch, rows, cols), np.float32)
    _lib().orc_pyr_recompose(_f(flat), rows, cols, ch, len(levels), type_, factor, _f(out))
    return out


def merge_coarse(fine, coarse, factor=0.8):
    fine, coarse = _img(fine), _img(coarse)
    ch, rows, cols = fine.shape
    out = np.zeros_like(fine)
    _lib().orc_pyr_merge_coarse(_f(fine), rows, cols, _f(coarse), coarse.shape[1], coarse.shape[2], ch, factor,
                                _f(out))
    return out


def dct2(img):
    a = _img(img).copy()
    _lib().orc_pyr_dct2(_f(a), a.shape[1], a.shape[2], a.shape[0])
    return a


def idct2(img):
    a = _img(img).copy()
    _lib().orc_pyr_idct2(_f(a), a.shape[1], a.shape[2], a.shape[0])
    return a


def gblur(img, sigma):
    a = _img(img).copy()
    _lib().orc_pyr_gblur(_f(a), a.shape[1], a.shape[2], a.shape[0], sigma)
    return a


def down(img, lanczos=False):
    a = _img(img)
    ch, rows, cols = a.shape
    out = np.zeros((ch, rows // 2, cols // 2), np.float32)
    _lib().orc_pyr_down(_f(a), rows, cols, ch, int(lanczos), _f(out))
    return out


def up(img, sx, sy, lanczos=False):
    a = _img(img)
    ch, rows, cols = a.shape
    out = np.zeros((ch, sx, sy), np.float32)
    _lib().orc_pyr_up(_f(a), rows, cols, ch, sx, sy, int(lanczos), _f(out))
    return out
