"""GPU parity of the general kernels (kernels_gen.hip) against the oracle,
bit for bit: the parameter sets the fast kernels leave out -- temporal
patches (kt > 1) with the temporal DCT, motion-compensated trajectories,
patch sizes 9..16 (the 12-point Kaiser window), group sizes N that are
not powers of two or exceed 16, search windows up to 15 x 15, up to 16
matches per frame, and predictive windows smaller than Nb."""
import numpy as np
import pytest

from vbm3d_amd import synth

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bits_equal(a, b, what):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    neq = bits(a) != bits(b)
    if neq.any():
        i = np.argwhere(neq)[0]
        raise AssertionError(f"{what}: {neq.sum()} of {a.size} differ; first at {tuple(i)}: "
                             f"{a[tuple(i)]!r} vs {b[tuple(i)]!r}, max|d|={np.abs(a - b).max()}")


def all_refs(W, H, F, C, k, p, kt=1):
    import vbm3d_amd as V
    xs, ys, _ = V.grid(W, H, k, p, 16, 1)
    per = (ys.astype(np.int64) * W + xs).astype(np.int64)
    return np.concatenate([t * W * H * C + per for t in range(F - kt + 1)]).astype(np.uint32)


def motion_flow(F, H, W, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    ff = np.empty((F - 1, 2, H, W), np.float32)
    for t in range(F - 1):
        ff[t, 0] = 0.8 + 0.5 * np.sin(xx / 7.0 + t) + rng.uniform(-0.05, 0.05, (H, W))
        ff[t, 1] = -0.6 + 0.5 * np.cos(yy / 5.0 + t) + rng.uniform(-0.05, 0.05, (H, W))
    return ff, -ff


def _bm(V, oracle, ctx, vid, prm, ff=None, bf=None):
    F, C, H, W = vid.shape
    refs = all_refs(W, H, F, C, prm.k, prm.p, prm.kt)
    gi, gd, gs, gn = V.block_match(vid, prm, refs, ff, bf, ctx=ctx)
    oi, od, os_, on = oracle.block_match(vid, prm, refs, ff, bf)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(gn, on)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(np.where(np.isnan(gd), 0, gd), np.where(np.isnan(od), 0, od), "distances")
    return gn


BM_CASES = {
    "ns11-npr9-nb12": dict(Ns=11, Npr=9, Nb=12, Nf=2, N=16),
    "ns15-npr15-nb3": dict(Ns=15, Npr=15, Nb=3, Nf=2),
    "npr1-nb2": dict(Npr=1, Nb=2),          # predictive windows of one candidate: fewer picks than Nb
    "npr3-nb16": dict(Ns=9, Npr=3, Nb=16, Nf=1, N=32),
    "n12": dict(N=12),
    "k12": dict(k=12),
    "kt2": dict(kt=2),
    "kt3-n6": dict(kt=3, N=6),
}


@pytest.mark.parametrize("case", list(BM_CASES), ids=list(BM_CASES))
def test_bm_general(oracle, gpu_ctx, case):
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(56, 44, 6, 3, 20.0, 23)
    yuv = oracle.color(noisy, True)
    for mk in (V.initialize_parameters_1, V.initialize_parameters_2):
        prm = mk(20.0, **BM_CASES[case])
        _bm(V, oracle, gpu_ctx, yuv, prm)


def test_bm_mc_trajectories(oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(48, 40, 6, 1, 20.0, 24)
    ff, bf = motion_flow(6, 40, 48, 3)
    for mk in (V.initialize_parameters_1, V.initialize_parameters_2):
        prm = mk(20.0, kt=3)
        prm.mc = 1
        _bm(V, oracle, gpu_ctx, noisy, prm, ff, bf)


def _run_both(V, oracle, ctx, noisy, sigma, p1, p2, nb=1, ff=None, bf=None):
    nz = noisy.copy()
    gb, gf = np.zeros_like(nz), np.zeros_like(nz)
    V.run_vbm3d(sigma, nz, ff, bf, gb, gf, p1, p2, True, nb, ctx=ctx)
    ob, of, onz = oracle.run_vbm3d(sigma, noisy, p1, p2, True, nb, fflow=ff, bflow=bf)
    assert_bits_equal(gb, ob, "basic")
    assert_bits_equal(gf, of, "final")
    assert_bits_equal(nz, onz, "noisy after the colour round trip")
    return gb, gf


RUN_CASES = {
    "kt2": (dict(kt=2), dict(kt=2), 3),
    "kt3-dct": (dict(kt=3, T_2D=0), dict(kt=3), 1),
    "kt2-hadamard": (dict(kt=2, T_3D=2), dict(kt=2, T_3D=2), 1),
    "k12": (dict(k=12), dict(k=12), 1),
    "k10-k9": (dict(k=10), dict(k=9), 3),
    "n6-n12": (dict(N=6), dict(N=12), 3),
    "n32": (dict(N=32, Nb=4), dict(N=32, Nb=4), 1),
    "npr1-nb2": (dict(Npr=1), dict(Npr=1), 3),
    "ns11-nb12": (dict(Ns=11, Npr=9, Nb=12, Nf=2, N=16), dict(Ns=11, Npr=9, Nb=12, Nf=2, N=16), 1),
}


@pytest.mark.parametrize("case", list(RUN_CASES), ids=list(RUN_CASES))
def test_run_general(oracle, gpu_ctx, case):
    """Both steps end to end through the host-buffer entry."""
    import vbm3d_amd as V
    o1, o2, C = RUN_CASES[case]
    sigma = 20.0
    clean, noisy = synth.noisy_clip(64, 48, 7, C, sigma, 31)
    p1, p2 = V.initialize_parameters_1(sigma, **o1), V.initialize_parameters_2(sigma, **o2)
    gb, gf = _run_both(V, oracle, gpu_ctx, noisy, sigma, p1, p2)
    assert synth.psnr(clean, gf) > synth.psnr(clean, noisy) + 3


def test_run_mc_with_flows_and_tiled_grid(oracle, gpu_ctx):
    """kt = 3 with motion-compensated trajectories and flows, on the
    single-thread grid and on the 2-crop grid."""
    import vbm3d_amd as V
    W, H, F = 80, 40, 6
    _, noisy = synth.noisy_clip(W, H, F, 1, 20.0, 32)
    ff, bf = motion_flow(F, H, W, 5)
    p1, p2 = V.initialize_parameters_1(20.0, kt=3), V.initialize_parameters_2(20.0, kt=3)
    p1.mc = p2.mc = 1
    for nb in (1, 2):
        _run_both(V, oracle, gpu_ctx, noisy, 20.0, p1, p2, nb=nb, ff=ff, bf=bf)


def test_run_kt_chunked_ring(oracle, monkeypatch):
    """kt = 2 with a tiny ring budget: one-frame chunks, a ring of
    2 Nf + kt frames; the result does not change."""
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(48, 40, 9, 1, 20.0, 33)
    monkeypatch.setenv("VBM3D_EST_BUDGET_MB", "1")
    p1, p2 = V.initialize_parameters_1(20.0, kt=2), V.initialize_parameters_2(20.0, kt=2)
    with V.Context(0) as ctx:
        _run_both(V, oracle, ctx, noisy, 20.0, p1, p2)


def test_general_limits_refused(gpu_ctx):
    """Outside the general kernels' bounds the call fails with a message."""
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(48, 40, 4, 1, 20.0, 34)
    for bad in (dict(k=17), dict(kt=9), dict(N=33), dict(Nb=17), dict(Ns=17), dict(k=16, kt=2)):
        p1 = V.initialize_parameters_1(20.0, **bad)
        with pytest.raises(V.VBM3DError):
            V.denoise(noisy, 20.0, prms_1=p1, ctx=gpu_ctx)


def test_batch_refuses_general_parameters():
    import torch
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(48, 40, 4, 1, 20.0, 35)
    it = V.BatchItem(20.0, torch.from_numpy(noisy).to("cuda:0"), prms_1=V.initialize_parameters_1(20.0, kt=2))
    with pytest.raises(V.VBM3DError, match="batched form"):
        V.run_batch([it])
