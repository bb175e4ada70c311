"""GPU parity: libvbm3d_hip.so (through the C ABI) against the CPU oracle.

The bar is bit-exact equality: block-matching indices, distances, group
sizes, the aggregation sums and the basic / final estimates, on the same
float32 inputs.  Sizes are chosen so the oracle finishes in seconds.
"""
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


def all_refs(W, H, F, C, k, p, nb=1):
    import vbm3d_amd as V
    xs, ys, _ = V.grid(W, H, k, p, 16, nb)
    per = (ys.astype(np.int64) * W + xs).astype(np.int64)
    return np.concatenate([t * W * H * C + per for t in range(F)]).astype(np.uint32)


@pytest.fixture(scope="module")
def cfg1():
    clean, noisy = synth.noisy_clip(176, 144, 10, 1, 20.0, 0)
    return clean, noisy


@pytest.fixture(scope="module")
def small_rgb():
    clean, noisy = synth.noisy_clip(96, 72, 7, 3, 20.0, 11)
    return clean, noisy


def _bm_compare(vid, prm, refs, oracle, ctx, fflow=None, bflow=None):
    import vbm3d_amd as V
    gi, gd, gs, gn = V.block_match(vid, prm, refs, fflow, bflow, ctx=ctx)
    oi, od, os_, on = oracle.block_match(vid, prm, refs, fflow, bflow)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(gn, on)
    np.testing.assert_array_equal(gi, oi)
    assert_bits_equal(np.where(np.isnan(gd), 0, gd), np.where(np.isnan(od), 0, od), "distances")
    return gn


def test_bm_config1_step1(cfg1, oracle, gpu_ctx):
    _, noisy = cfg1
    prm = oracle.params_1(20.0)
    refs = all_refs(176, 144, 10, 1, prm.k, prm.p)
    import vbm3d_amd as V
    nsx = _bm_compare(noisy, V.initialize_parameters_1(20.0), refs, oracle, gpu_ctx)
    assert set(np.unique(nsx)) <= {1, 2, 4, 8}


def test_bm_config1_step2_params(cfg1, oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = cfg1
    prm = V.initialize_parameters_2(20.0)
    refs = all_refs(176, 144, 10, 1, prm.k, prm.p)
    _bm_compare(noisy, prm, refs, oracle, gpu_ctx)


def test_bm_rgb_colour(small_rgb, oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = small_rgb
    yuv = oracle.color(noisy, True)
    for prm in (V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0),
                V.initialize_parameters_2(40.0)):
        refs = all_refs(96, 72, 7, 3, prm.k, prm.p)
        _bm_compare(yuv, prm, refs, oracle, gpu_ctx)


# search shapes other than the default (Ns 7, Npr 5, Nb 2), which runs a
# kernel compiled with those sizes fixed: the paired form with run-time
# sizes (Nb 1, Npr 3), and the general form (Nb > 2, or 7x7 windows)
SHAPES = [dict(Ns=7, Npr=5, Nb=1, Nf=4), dict(Ns=5, Npr=3, Nb=2, Nf=2), dict(Ns=3, Npr=3, Nb=2, Nf=1),
          dict(Ns=7, Npr=7, Nb=2, Nf=3), dict(Ns=7, Npr=5, Nb=3, Nf=3), dict(Ns=8, Npr=8, Nb=4, Nf=2)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda d: "ns{Ns}-npr{Npr}-nb{Nb}-nf{Nf}".format(**d))
def test_bm_search_shapes(oracle, gpu_ctx, shape):
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(64, 48, 7, 3, 20.0, 21)
    yuv = oracle.color(noisy, True)
    for prm in (V.initialize_parameters_1(20.0, **shape), V.initialize_parameters_2(20.0, **shape)):
        refs = all_refs(64, 48, 7, 3, prm.k, prm.p)
        _bm_compare(yuv, prm, refs, oracle, gpu_ctx)


def test_run_search_shape(small_rgb, oracle, gpu_ctx):
    """Both steps end to end with a non-default search shape."""
    import vbm3d_amd as V
    _, noisy = small_rgb
    shape = dict(Ns=5, Npr=3, Nb=1, Nf=2)
    _run_both(noisy, 20.0, oracle, gpu_ctx, p1=V.initialize_parameters_1(20.0, **shape),
              p2=V.initialize_parameters_2(20.0, **shape))


def test_bm_ties(oracle, gpu_ctx):
    """Integer-valued clip with 3 levels: distances tie everywhere, so the
    heap-select tie order of std::partial_sort decides."""
    import vbm3d_amd as V
    rng = np.random.default_rng(5)
    vid = rng.integers(0, 3, (6, 1, 40, 48)).astype(np.float32)
    for prm in (V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)):
        refs = all_refs(48, 40, 6, 1, prm.k, prm.p)
        _bm_compare(vid, prm, refs, oracle, gpu_ctx)
    flat = np.full((5, 1, 24, 24), 7.0, np.float32)
    prm = V.initialize_parameters_1(20.0)
    _bm_compare(flat, prm, all_refs(24, 24, 5, 1, prm.k, prm.p), oracle, gpu_ctx)


def test_bm_with_flow(oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(64, 48, 6, 1, 20.0, 3)
    rng = np.random.default_rng(9)
    ff = rng.uniform(-2.6, 2.6, (5, 2, 48, 64)).astype(np.float32)
    bf = rng.uniform(-2.6, 2.6, (5, 2, 48, 64)).astype(np.float32)
    prm = V.initialize_parameters_1(20.0)
    refs = all_refs(64, 48, 6, 1, prm.k, prm.p)
    _bm_compare(noisy, prm, refs, oracle, gpu_ctx, ff, bf)


@pytest.mark.parametrize("nb", [1, 4])
def test_step1_sums_gray(cfg1, oracle, gpu_ctx, nb):
    import vbm3d_amd as V
    _, noisy = cfg1
    prm = V.initialize_parameters_1(20.0)
    gn, gd = V.step(1, 20.0, noisy, None, prm, True, nb, ctx=gpu_ctx)
    on, od = oracle.step(1, 20.0, noisy, None, prm, True, nb)
    assert_bits_equal(gd, od, "step-1 denominator")
    assert_bits_equal(gn, on, "step-1 numerator")


def test_step_sums_rgb(small_rgb, oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = small_rgb
    yuv = oracle.color(noisy, True)
    p1 = V.initialize_parameters_1(20.0)
    gn, gd = V.step(1, 20.0, yuv, None, p1, True, 1, ctx=gpu_ctx)
    on, od = oracle.step(1, 20.0, yuv, None, p1, True, 1)
    assert_bits_equal(gd, od, "step-1 denominator")
    assert_bits_equal(gn, on, "step-1 numerator")
    basic = np.where(od == 0, yuv, on / np.where(od == 0, 1, od)).astype(np.float32)
    p2 = V.initialize_parameters_2(20.0)
    gn, gd = V.step(2, 20.0, yuv, basic, p2, True, 1, ctx=gpu_ctx)
    on, od = oracle.step(2, 20.0, yuv, basic, p2, True, 1)
    assert_bits_equal(gd, od, "step-2 denominator")
    assert_bits_equal(gn, on, "step-2 numerator")


def _run_both(noisy, sigma, oracle, ctx, nb=1, color=True, p1=None, p2=None):
    import vbm3d_amd as V
    p1 = p1 or V.initialize_parameters_1(sigma)
    p2 = p2 or V.initialize_parameters_2(sigma)
    nz = noisy.copy()
    gb = np.zeros_like(nz)
    gf = np.zeros_like(nz)
    V.run_vbm3d(sigma, nz, None, None, gb, gf, p1, p2, color, nb, ctx=ctx)
    ob, of, onz = oracle.run_vbm3d(sigma, noisy, p1, p2, color, nb)
    assert_bits_equal(gb, ob, "basic")
    assert_bits_equal(gf, of, "final")
    assert_bits_equal(nz, onz, "noisy after the colour round trip")
    return gb, gf


def test_run_config1(cfg1, oracle, gpu_ctx):
    clean, noisy = cfg1
    gb, gf = _run_both(noisy, 20.0, oracle, gpu_ctx)
    assert synth.psnr(clean, gf) > synth.psnr(clean, gb) > synth.psnr(clean, noisy) + 8


@pytest.mark.parametrize("nb", [2, 8])
def test_run_config1_threaded_grid(cfg1, oracle, gpu_ctx, nb):
    _, noisy = cfg1
    _run_both(noisy, 20.0, oracle, gpu_ctx, nb=nb)


@pytest.mark.parametrize("sigma", [20.0, 40.0])
def test_run_rgb(small_rgb, oracle, gpu_ctx, sigma):
    clean, _ = small_rgb
    rng = np.random.default_rng(int(sigma))
    noisy = (clean + np.float32(sigma) * rng.standard_normal(clean.shape, dtype=np.float32)).astype(np.float32)
    gb, gf = _run_both(noisy, sigma, oracle, gpu_ctx)
    assert synth.psnr(clean, gf) > synth.psnr(clean, noisy) + 6


def test_run_rgb_no_colour_and_hadamard(small_rgb, oracle, gpu_ctx):
    import vbm3d_amd as V
    _, noisy = small_rgb
    p1 = V.initialize_parameters_1(20.0, T_3D=V.HADAMARD)
    p2 = V.initialize_parameters_2(20.0, T_3D=V.HADAMARD)
    _run_both(noisy, 20.0, oracle, gpu_ctx, color=False, p1=p1, p2=p2)
    p1 = V.initialize_parameters_1(20.0, T_2D=V.DCT)
    p2 = V.initialize_parameters_2(40.0, T_2D=V.BIOR)
    _run_both(noisy, 20.0, oracle, gpu_ctx, p1=p1, p2=p2)


@pytest.mark.parametrize("rgb", [False, True])
def test_run_group_size_16(cfg1, small_rgb, oracle, gpu_ctx, rgb):
    """N = 16, the largest group the kernels hold: most groups have 16
    members, so the filter's 16-member Haar, the aggregation's 16-member
    scan and its largest chunk list all run.  On the 7-frame RGB clip
    Nb = 4 gives enough candidates (and the general matching form)."""
    import vbm3d_amd as V
    _, noisy = small_rgb if rgb else cfg1
    nb = 4 if rgb else V.NONE
    p1 = V.initialize_parameters_1(20.0, N=16, Nb=nb)
    p2 = V.initialize_parameters_2(20.0, N=16, Nb=nb)
    _run_both(noisy, 20.0, oracle, gpu_ctx, p1=p1, p2=p2)
    vid = oracle.color(noisy, True) if rgb else noisy
    F, C, H, W = noisy.shape
    _, _, _, nsx = V.block_match(vid, p1, all_refs(W, H, F, C, p1.k, p1.p), ctx=gpu_ctx)
    assert (nsx == 16).mean() > 0.3


@pytest.mark.parametrize("shape", [(9, 9, 1, 1), (12, 9, 2, 3), (33, 17, 3, 1), (40, 8, 5, 3), (9, 40, 1, 3)])
def test_run_edge_sizes(oracle, gpu_ctx, shape):
    """Frames one pixel larger than the step-1 patch, a frame exactly one
    patch high, one-frame clips, and sizes that leave partial aggregation
    tiles on both axes: the oracle's bits."""
    W, H, F, C = shape
    _, noisy = synth.noisy_clip(W, H, F, C, 20.0, 5)
    _run_both(noisy, 20.0, oracle, gpu_ctx)


def test_run_rejects_window_smaller_than_nb(oracle, gpu_ctx):
    """An 8x8 frame leaves one step-1 candidate for Nb = 2 matches, which
    the reference's partial_sort cannot take; both paths refuse it."""
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(8, 8, 2, 1, 20.0, 5)
    with pytest.raises(V.VBM3DError, match="smaller than Nb"):
        V.denoise(noisy, 20.0, ctx=gpu_ctx)
    with pytest.raises(RuntimeError, match="smaller than Nb"):
        oracle.run_vbm3d(20.0, noisy, oracle.params_1(20.0), oracle.params_2(20.0), True, 1)


@pytest.mark.parametrize("colour", [True, False])
def test_run_skipped_steps(small_rgb, oracle, gpu_ctx, colour):
    """The host-buffer entry with step 1 skipped (k = 0: the basic estimate
    is an input, and takes the colour round trip) and with step 2 skipped
    (the final estimate is the basic one): every returned video, the noisy
    one after its round trip included, equals the oracle's."""
    import vbm3d_amd as V
    _, noisy = small_rgb
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    skip1, skip2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    skip1.k = 0
    skip2.k = 0
    given = noisy[::-1].copy()  # any video of the right shape serves as the basic estimate
    for q1, q2, basic_in in [(skip1, p2, given), (p1, skip2, None)]:
        nz = noisy.copy()
        gb = np.zeros_like(nz) if basic_in is None else basic_in.copy()
        gf = np.zeros_like(nz)
        V.run_vbm3d(20.0, nz, None, None, gb, gf, q1, q2, colour, 1, ctx=gpu_ctx)
        ob, of, onz = oracle.run_vbm3d(20.0, noisy, q1, q2, colour, 1, basic_in=basic_in)
        assert_bits_equal(gb, ob, "basic")
        assert_bits_equal(gf, of, "final")
        assert_bits_equal(nz, onz, "noisy after the colour round trip")


def test_run_with_flow(oracle, gpu_ctx):
    """The whole path with forward and backward flows, through the
    host-buffer entry: 20 RGB frames, so that step 1 runs in several chunks
    while the upload is still arriving, and the aggregation's reach grows
    with the largest flow."""
    import vbm3d_amd as V
    W, H, F = 80, 56, 20
    _, noisy = synth.noisy_clip(W, H, F, 3, 20.0, 4)
    rng = np.random.default_rng(12)
    ff = rng.uniform(-3.4, 3.4, (F - 1, 2, H, W)).astype(np.float32)
    bf = rng.uniform(-3.4, 3.4, (F - 1, 2, H, W)).astype(np.float32)
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    nz = noisy.copy()
    gb = np.zeros_like(nz)
    gf = np.zeros_like(nz)
    V.run_vbm3d(20.0, nz, ff, bf, gb, gf, p1, p2, True, 1, ctx=gpu_ctx)
    ob, of, onz = oracle.run_vbm3d(20.0, noisy, p1, p2, True, 1, fflow=ff, bflow=bf)
    assert_bits_equal(gb, ob, "basic")
    assert_bits_equal(gf, of, "final")
    assert_bits_equal(nz, onz, "noisy after the colour round trip")


def test_run_chunked_ring(cfg1, oracle, gpu_ctx, monkeypatch):
    """A tiny estimate budget forces one-frame chunks and a ring of
    2 Nf + 1 frames; the result must not change."""
    import vbm3d_amd as V
    _, noisy = cfg1
    monkeypatch.setenv("VBM3D_EST_BUDGET_MB", "1")
    with V.Context(0) as ctx:
        _run_both(noisy, 20.0, oracle, ctx)


@pytest.mark.parametrize("frames", [1, 3])
def test_run_per_frame_chunks(small_rgb, oracle, frames):
    """vbm3d_set_chunk_frames: the plain per-frame path (one launch per
    kernel and frame) and 3-frame chunks give the oracle's bits."""
    import vbm3d_amd as V
    _, noisy = small_rgb
    with V.Context(0) as ctx:
        ctx.set_chunk_frames(frames)
        _run_both(noisy, 20.0, oracle, ctx)


def test_run_config2_threaded_grid(oracle, gpu_ctx):
    """30 RGB frames 960x540, sigma 20, seed 1, against the oracle run with
    the OpenMP grid of 16 threads (ref_grid_nb = 16), bit for bit."""
    import vbm3d_amd as V
    clean, noisy = synth.noisy_clip(960, 540, 30, 3, 20.0, 1)
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    gb, gf = V.denoise(noisy, 20.0, True, 16, ctx=gpu_ctx)
    ob, of, _ = oracle.run_vbm3d(20.0, noisy, p1, p2, True, 16, 16)
    assert_bits_equal(gb, ob, "basic")
    assert_bits_equal(gf, of, "final")
    assert synth.psnr(clean, gf) > synth.psnr(clean, noisy) + 10


def test_step_error_against_fp64(cfg1, oracle, gpu_ctx):
    """Per-step error of the GPU path against the FP64 replay of the same
    step: within 1e-3 on the 0-255 scale."""
    import vbm3d_amd as V
    _, noisy = cfg1
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    n1, d1 = V.step(1, 20.0, noisy, None, p1, ctx=gpu_ctx)
    basic = np.where(d1 == 0, noisy, n1 / np.where(d1 == 0, 1, d1)).astype(np.float32)
    m1, e1 = oracle.step_fp64(1, 20.0, noisy, None, p1)
    err1 = np.abs(basic - np.where(e1 == 0, noisy, m1 / np.where(e1 == 0, 1, e1))).max()
    n2, d2 = V.step(2, 20.0, noisy, basic, p2, ctx=gpu_ctx)
    final = np.where(d2 == 0, basic, n2 / np.where(d2 == 0, 1, d2))
    m2, e2 = oracle.step_fp64(2, 20.0, noisy, basic, p2)
    err2 = np.abs(final - np.where(e2 == 0, basic, m2 / np.where(e2 == 0, 1, e2))).max()
    print(f"max |GPU - FP64|: step 1 {err1:.3g}, step 2 {err2:.3g}")
    assert err1 < 1e-3 and err2 < 1e-3
