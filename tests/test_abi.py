"""CPU-side checks of the C ABI library: it loads, exports every symbol
declared in include/vbm3d_hip.h, and its host-side logic (parameter
defaults, reference grid, group selection) equals the oracle's."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "vbm3d_hip.h")).read()
    return sorted(set(re.findall(r"\b(vbm3d_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    import vbm3d_amd as V
    L = V.lib()
    syms = _declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s


def _param_cases():
    return [20.0, 10.0, 30.0, 30.5, 40.0, 50.0]


@pytest.mark.parametrize("sigma", _param_cases())
def test_default_params_match_oracle(sigma, oracle):
    import vbm3d_amd as V
    for a, b in [(V.initialize_parameters_1(sigma), oracle.params_1(sigma)),
                 (V.initialize_parameters_2(sigma), oracle.params_2(sigma))]:
        da, db = a.as_dict(), b.as_dict()
        assert da == db


def test_default_params_overrides(oracle):
    import vbm3d_amd as V
    a = V.initialize_parameters_1(20.0, k=6, p=3, d=3, tau=100.0, lambda3D=2.0, T_2D=1)
    b = oracle.params_1(20.0, k=6, p=3, d=3, tau=100.0, lambda3D=2.0, T_2D=1)
    assert a.as_dict() == b.as_dict()
    assert a.T_2D == 0  # BIOR is only kept for k == 8 (src/main.cpp:113-119)
    a = V.initialize_parameters_2(40.0, k=8, T_2D=1, T_3D=2)
    b = oracle.params_2(40.0, k=8, T_2D=1, T_3D=2)
    assert a.as_dict() == b.as_dict()


@pytest.mark.parametrize("W,H,k,p", [(176, 144, 8, 6), (176, 144, 7, 3), (960, 540, 8, 6),
                                     (1920, 1080, 8, 4), (640, 360, 7, 3), (33, 20, 8, 6)])
@pytest.mark.parametrize("nb", [1, 2, 4, 8, 16, 32])
def test_grid_matches_oracle(W, H, k, p, nb, oracle):
    import vbm3d_amd as V
    try:
        ref = oracle.grid(W, H, k, p, 16, nb)
    except ValueError:
        with pytest.raises(ValueError):
            V.grid(W, H, k, p, 16, nb)
        return
    got = V.grid(W, H, k, p, 16, nb)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_selection_matches_std_partial_sort(oracle):
    """The kernels' heap select equals std::partial_sort, ties included."""
    import vbm3d_amd as V
    rng = np.random.default_rng(0)
    for trial in range(3000):
        n = int(rng.integers(1, 70))
        mid = int(rng.integers(1, n + 1))
        levels = int(rng.integers(1, 6))
        d = rng.integers(0, levels, n).astype(np.float32)
        if trial % 3 == 0:
            d = rng.standard_normal(n).astype(np.float32)
        idx = rng.permutation(1000)[:n].astype(np.uint32)
        d1, i1 = V.api.selftest_partial_sort(d, idx, mid)
        d2, i2 = oracle.partial_sort(d, idx, mid)
        np.testing.assert_array_equal(i1[:mid], i2[:mid])
        np.testing.assert_array_equal(d1[:mid], d2[:mid])
