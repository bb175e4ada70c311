"""GPU parity at the benchmark configurations and on counters.

* 1080p RGB, sigma 40, seed 2, velocities up to 4 (the headline clip at
  its own noise level): the first 10 frames bit for bit against the
  oracle run with the grid of a 16-thread reference run; the whole
  60-frame clip through both entry points (device buffers and host
  buffers) with size-independent properties.
* The 64-clip mixed batch (seeds 100..163): the batched form bit-identical
  to the per-clip path on every clip, and three clips against the oracle.
* The kernel's count of evaluated candidates equals the oracle's count of
  patchDistance calls.
* The den == 0 fallback and a flow value far outside the frame.
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


CFG3 = dict(width=1920, height=1080, frames=60, channels=3, sigma=40.0, seed=2, vmax=4)


def test_config3_first_frames_vs_oracle(oracle, gpu_ctx):
    """Frames 0..9 of the 1080p sigma-40 clip: basic and final estimates
    bit-identical to the oracle (both with the 16-thread grid)."""
    import vbm3d_amd as V
    c = CFG3
    clean, noisy = synth.noisy_clip(c["width"], c["height"], 10, 3, c["sigma"], c["seed"], vmax=c["vmax"])
    gb, gf = V.denoise(noisy, c["sigma"], True, 16, ctx=gpu_ctx)
    p1, p2 = V.initialize_parameters_1(c["sigma"]), V.initialize_parameters_2(c["sigma"])
    ob, of, _ = oracle.run_vbm3d(c["sigma"], noisy, p1, p2, True, 16, 16)
    assert_bits_equal(gb, ob, "basic")
    assert_bits_equal(gf, of, "final")
    assert abs(synth.psnr(clean, gf) - synth.psnr(clean, of)) == 0.0
    assert synth.psnr(clean, gf) > synth.psnr(clean, noisy) + 12


def test_config3_full_clip_properties(gpu_ctx):
    """All 60 frames: the 10-frame clip is the full clip's first 10 frames
    (the generator draws the noise in order); the device-buffer entry and
    the host-buffer entry give the same bits, and the filter gains more
    than 14 dB."""
    import torch
    import vbm3d_amd as V
    c = CFG3
    clean, noisy = synth.noisy_clip(c["width"], c["height"], c["frames"], 3, c["sigma"], c["seed"], vmax=c["vmax"])
    _, first = synth.noisy_clip(c["width"], c["height"], 10, 3, c["sigma"], c["seed"], vmax=c["vmax"])
    assert np.array_equal(bits(first), bits(noisy[:10]))
    del first
    p1, p2 = V.initialize_parameters_1(c["sigma"]), V.initialize_parameters_2(c["sigma"])
    nz = noisy.copy()
    hb = np.empty_like(nz)
    hf = np.empty_like(nz)
    V.run_vbm3d(c["sigma"], nz, None, None, hb, hf, p1, p2, True, 1, ctx=gpu_ctx)
    dev = torch.device("cuda:0")
    dz = torch.from_numpy(noisy).to(dev)
    db, df = torch.empty_like(dz), torch.empty_like(dz)
    V.run_device(c["sigma"], dz, db, df, p1, p2, True, 1, ctx=gpu_ctx)
    torch.cuda.synchronize()
    assert np.array_equal(bits(db.cpu().numpy()), bits(hb))
    assert np.array_equal(bits(df.cpu().numpy()), bits(hf))
    assert np.array_equal(bits(dz.cpu().numpy()), bits(nz))
    gain = synth.psnr(clean, hf) - synth.psnr(clean, noisy)
    print(f"1080p sigma 40: {synth.psnr(clean, noisy):.3f} -> {synth.psnr(clean, hf):.3f} dB")
    assert gain > 14.0


def test_config4_batch_vs_per_clip_and_oracle(oracle):
    """The 64 mixed clips (seeds 100..163, synth.batch_spec): the batched
    form (graph-captured, default launch shape) equals the per-clip path
    on every clip; the three smallest clips equal the oracle."""
    import torch
    import vbm3d_amd as V
    dev = torch.device("cuda:0")
    specs, host = [], []
    for seed in range(100, 164):
        W, H, F, C, s = synth.batch_spec(seed)
        specs.append((W, H, F, C, s, seed))
        host.append(synth.noisy_clip(W, H, F, C, s, seed)[1])
    with V.Context(0) as ctx:
        per = []
        for (W, H, F, C, s, seed), nz in zip(specs, host):
            z = torch.from_numpy(nz).to(dev)
            b, f = torch.empty_like(z), torch.empty_like(z)
            V.run_device(s, z, b, f, V.initialize_parameters_1(s), V.initialize_parameters_2(s), True, 1, ctx=ctx)
            per.append((z, b, f))
        items = [V.BatchItem(s, torch.from_numpy(nz).to(dev)) for (W, H, F, C, s, seed), nz in zip(specs, host)]
        V.run_batch(items, 1, None, ctx=ctx)
        for it, nz in zip(items, host):  # a graph replay on restored inputs
            it.noisy.copy_(torch.from_numpy(nz).to(dev))
        V.run_batch(items, 1, None, ctx=ctx)
        torch.cuda.synchronize()
        for i, (it, (z, b, f)) in enumerate(zip(items, per)):
            assert torch.equal(it.basic, b), f"clip {i} basic"
            assert torch.equal(it.denoised, f), f"clip {i} final"
            assert torch.equal(it.noisy, z), f"clip {i} noisy"
    order = sorted(range(len(specs)), key=lambda i: specs[i][0] * specs[i][1] * specs[i][2] * specs[i][3])
    for i in order[:3]:
        W, H, F, C, s, seed = specs[i]
        ob, of, _ = oracle.run_vbm3d(s, host[i], V.initialize_parameters_1(s), V.initialize_parameters_2(s),
                                     True, 1)
        assert_bits_equal(per[i][1].cpu().numpy(), ob, f"clip {seed} basic")
        assert_bits_equal(per[i][2].cpu().numpy(), of, f"clip {seed} final")


@pytest.mark.parametrize("case", ["config1", "config2_first4_nb16"])
def test_candidate_count_equals_oracle(oracle, case):
    """Candidates evaluated per step (the roofline's work count) equal the
    oracle's patchDistance calls on the same clip and grid."""
    import torch
    import vbm3d_amd as V
    from vbm3d_amd import api
    if case == "config1":
        _, noisy = synth.noisy_clip(176, 144, 10, 1, 20.0, 0)
        nb = 1
    else:
        _, noisy = synth.noisy_clip(960, 540, 4, 3, 20.0, 1)
        nb = 16
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    oracle.run_vbm3d(20.0, noisy, p1, p2, True, nb, 16)
    expect = [oracle.last_stats(st) for st in (1, 2)]
    with V.Context(0) as ctx:
        dev = torch.device("cuda:0")
        z = torch.from_numpy(noisy).to(dev)
        b, f = torch.empty_like(z), torch.empty_like(z)
        api.profile(2, ctx)
        V.run_device(20.0, z, b, f, p1, p2, True, nb, ctx=ctx)
        torch.cuda.synchronize()
        got = [api.profile_read("block_match", ctx, st)[2] for st in (1, 2)]
        api.profile(0, ctx)
    for st in (0, 1):
        print(f"step {st + 1}: GPU {got[st]:.0f} candidates, oracle {expect[st]['candidates']}, "
              f"near-tie refs {expect[st]['near_tie_refs']} of {expect[st]['refs']}")
        assert got[st] == expect[st]["candidates"]


@pytest.mark.parametrize("nb", [1, 2])
def test_zero_denominator_falls_back_to_noisy(oracle, gpu_ctx, nb):
    """With sigma = 1e20 every group weight 1/(sigma^2 count) is 0 (sigma^2
    overflows to inf), so the step-1 denominator is 0 everywhere and the
    basic estimate is the noisy video (src/vbm3d.cpp:168-170, 277), on the
    single-thread and the tiled grid alike."""
    import vbm3d_amd as V
    _, noisy = synth.noisy_clip(80, 72, 4, 1, 20.0, 3)
    p1 = V.initialize_parameters_1(1e20)
    p2 = V.initialize_parameters_2(1e20)
    p2.k = 0
    nz = noisy.copy()
    gb, gf = np.zeros_like(nz), np.zeros_like(nz)
    V.run_vbm3d(1e20, nz, None, None, gb, gf, p1, p2, True, nb, ctx=gpu_ctx)
    ob, of, _ = oracle.run_vbm3d(1e20, noisy, p1, p2, True, nb)
    assert_bits_equal(gb, noisy, "basic = noisy")
    assert_bits_equal(gb, ob, "basic vs oracle")
    assert_bits_equal(gf, of, "final vs oracle")


def test_far_flow_value_clamps_like_the_frame_edge(oracle, gpu_ctx):
    """A flow sample of 1e9 projects the window onto the frame edge, as a
    sample equal to W does; the aggregation's reach stays bounded, so the
    output equals the oracle's and the run with the sample set to W."""
    import vbm3d_amd as V
    W, H, F = 64, 48, 6
    _, noisy = synth.noisy_clip(W, H, F, 1, 20.0, 8)
    rng = np.random.default_rng(3)
    ff = rng.uniform(-2.4, 2.4, (F - 1, 2, H, W)).astype(np.float32)
    bf = rng.uniform(-2.4, 2.4, (F - 1, 2, H, W)).astype(np.float32)
    ff[1, 0, 20:30, 20:30] = 1e9
    ff_w = ff.copy()
    ff_w[1, 0, 20:30, 20:30] = W
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    outs = []
    for f in (ff, ff_w):
        nz = noisy.copy()
        gb, gf = np.zeros_like(nz), np.zeros_like(nz)
        V.run_vbm3d(20.0, nz, f, bf, gb, gf, p1, p2, True, 1, ctx=gpu_ctx)
        outs.append((gb, gf))
    ob, of, _ = oracle.run_vbm3d(20.0, noisy, p1, p2, True, 1, fflow=ff, bflow=bf)
    assert_bits_equal(outs[0][1], of, "final vs oracle")
    assert_bits_equal(outs[0][0], outs[1][0], "basic: 1e9 vs W")
    assert_bits_equal(outs[0][1], outs[1][1], "final: 1e9 vs W")
