"""Predictive block matching of the oracle against an independent FP64
brute force (tests/fp64_model.py) on tiny tie-free clips, plus invariants."""
import numpy as np
import pytest

import fp64_model as M
from vbm3d_amd import synth


@pytest.mark.parametrize("C,seed,which", [(1, 0, 1), (1, 1, 2), (3, 2, 1), (3, 3, 2)])
def test_bm_matches_fp64_brute_force(oracle, C, seed, which):
    clean, noisy = synth.noisy_clip(32, 24, 5, C, 25.0, seed)
    prm = oracle.params_1(25.0) if which == 1 else oracle.params_2(25.0)
    d = prm.as_dict()
    F, _, H, W = noisy.shape
    refs, expect = [], []
    checked = 0
    for t in range(F):
        for y in M.axis_grid(H - prm.k, prm.p):
            for x in M.axis_grid(W - prm.k, prm.p):
                marg = M.Margins()
                members, nsx = M.block_match(noisy.astype(np.float64), d, x, y, t, marg)
                if marg.rank < 1e-2 or marg.tau < 1e-2:
                    continue  # near-tie: FP32 and FP64 may legitimately differ
                refs.append(t * W * H * C + y * W + x)
                expect.append(([mx + my * W + mt * W * H * C for (mx, my, mt) in members], nsx))
                checked += 1
    assert checked > 0.9 * F * len(M.axis_grid(H - prm.k, prm.p)) * len(M.axis_grid(W - prm.k, prm.p))
    idx, dist, nsim, nsx = oracle.block_match(noisy, prm, np.array(refs, np.uint32))
    for r, (mem, ns) in enumerate(expect):
        assert nsx[r] == ns
        np.testing.assert_array_equal(idx[r, :len(mem)], mem)


def test_bm_invariants(oracle):
    _, noisy = synth.noisy_clip(40, 32, 6, 1, 20.0, 9)
    prm = oracle.params_1(20.0)
    xs, ys, _ = oracle.grid(40, 32, prm.k, prm.p)
    refs = np.concatenate([t * 40 * 32 + ys * 40 + xs for t in range(6)]).astype(np.uint32)
    idx, dist, nsim, nsx = oracle.block_match(noisy, prm, refs)
    assert np.all(idx[:, 0] == refs)                  # the reference patch ranks first (distance -d)
    assert np.all(np.abs(dist[:, 0] + prm.d) < 1e-3)
    for r in range(len(refs)):
        d = dist[r, :nsim[r]]
        assert np.all(np.diff(d) >= 0)               # ascending
        assert nsx[r] in (1, 2, 4, 8) and nsx[r] <= nsim[r]
        assert np.all(d[1:nsx[r]] <= prm.tau)        # the group stays within tau
        assert len(set(idx[r, :nsim[r]])) == nsim[r]  # no patch twice
    # per-frame quota: at most Nb candidates from each frame other than... the
    # reference frame contributes Nb, every other frame at most Nb
    W, H = 40, 32
    for r in range(len(refs)):
        frames = idx[r, :nsim[r]] // (W * H)
        assert np.bincount(frames).max() <= prm.Nb
