"""Batched form (vbm3d_run_batch): many clips of uneven size, channels and
sigma through persistent work-queue launches and a captured HIP graph.
Every clip's basic, final and colour-round-tripped noisy video must be
bit-identical to the per-clip path (vbm3d_run_device), for every launch
configuration and on every graph replay."""
import numpy as np
import pytest

from vbm3d_amd import synth

pytestmark = pytest.mark.gpu

# (W, H, F, C, sigma, seed): k = 7 and k = 8 step 2, grey and RGB, odd sizes
SPECS = [(64, 48, 6, 1, 10.0, 1), (72, 56, 9, 3, 20.0, 2), (40, 40, 5, 1, 30.0, 3),
         (96, 64, 7, 3, 50.0, 4), (50, 38, 4, 1, 50.0, 5), (33, 29, 3, 3, 20.0, 6)]


@pytest.fixture(scope="module")
def env():
    import torch
    import vbm3d_amd as V
    dev = torch.device("cuda:0")
    clips = [synth.noisy_clip(W, H, F, C, s, seed)[1] for (W, H, F, C, s, seed) in SPECS]
    return torch, V, dev, clips


def _per_clip(env, nb=1):
    torch, V, dev, clips = env
    out = []
    for (W, H, F, C, s, seed), nz in zip(SPECS, clips):
        z = torch.from_numpy(nz).to(dev)
        b, f = torch.empty_like(z), torch.empty_like(z)
        V.run_device(s, z, b, f, V.initialize_parameters_1(s), V.initialize_parameters_2(s), True, nb)
        out.append((z.cpu().numpy(), b.cpu().numpy(), f.cpu().numpy()))
    return out


def _items(env):
    torch, V, dev, clips = env
    return [V.BatchItem(s, torch.from_numpy(nz).to(dev)) for (W, H, F, C, s, seed), nz in zip(SPECS, clips)]


def _restore(env, items):
    torch, V, dev, clips = env
    for it, nz in zip(items, clips):
        it.noisy.copy_(torch.from_numpy(nz).to(dev))


def _same(items, ref, what):
    for i, (it, (z, b, f)) in enumerate(zip(items, ref)):
        for name, got, exp in (("noisy", it.noisy, z), ("basic", it.basic, b), ("final", it.denoised, f)):
            g = got.cpu().numpy()
            assert np.array_equal(g.view(np.uint32), exp.view(np.uint32)), \
                f"{what}: clip {i} {name} differs (max |d| {np.abs(g - exp).max()})"


@pytest.fixture(scope="module")
def reference(env):
    return _per_clip(env)


@pytest.mark.parametrize("cfg", [
    dict(chunk_frames=0, waves_per_group=0, grab=1, groups_per_cu=8, use_graph=0, persistent=0),
    dict(chunk_frames=1, waves_per_group=1, grab=4, groups_per_cu=2, use_graph=0, persistent=1),
    dict(chunk_frames=2, waves_per_group=2, grab=2, groups_per_cu=16, use_graph=0, persistent=1),
    dict(chunk_frames=3, waves_per_group=4, grab=8, groups_per_cu=1, use_graph=0, persistent=1),
    dict(chunk_frames=1, waves_per_group=2, grab=1, groups_per_cu=1, use_graph=0, persistent=0),
])
def test_batch_matches_per_clip(env, reference, cfg):
    torch, V, dev, clips = env
    items = _items(env)
    V.run_batch(items, 1, cfg)
    _same(items, reference, str(cfg))


def test_batch_graph_replays(env, reference):
    """Capture once, then replay with eager work and synchronisation
    between the replays; every replay equals the per-clip result."""
    torch, V, dev, clips = env
    items = _items(env)
    cfg = dict(chunk_frames=2, waves_per_group=0, grab=2, groups_per_cu=4, use_graph=1, persistent=1)
    for rep in range(4):
        _restore(env, items)
        for it in items:
            it.basic.fill_(float(rep))  # overwritten outputs
            it.denoised.fill_(-1.0)
        junk = torch.randn(1 << 20, device=dev).sum()  # unrelated eager work
        torch.cuda.synchronize()
        V.run_batch(items, 1, cfg)
        _same(items, reference, f"replay {rep}")
        assert torch.isfinite(junk)


def test_batch_threaded_grid(env):
    torch, V, dev, clips = env
    ref = _per_clip(env, nb=4)
    items = _items(env)
    V.run_batch(items, 4, dict(chunk_frames=0, waves_per_group=0, grab=1, groups_per_cu=8, use_graph=0,
                               persistent=0))
    _same(items, ref, "nb=4")


def test_autotune_leaves_inputs_and_matches(env, reference):
    torch, V, dev, clips = env
    items = _items(env)
    best = V.autotune_batch(items)
    print("autotuned:", best)
    for it, nz in zip(items, clips):
        assert np.array_equal(it.noisy.cpu().numpy().view(np.uint32), nz.view(np.uint32))
    # 0 = as many waves per group as fit (the default, kept when no trial beats it)
    assert 0 <= best["waves_per_group"] <= 4 and best["grab"] >= 1 and best["groups_per_cu"] >= 1
    V.run_batch(items)  # the tuned configuration, with a graph
    _same(items, reference, "autotuned")


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_mixed_search_shapes(env, persistent):
    """Clips of one kernel class with the default search shape and with
    another share a launch, which then runs the run-time sized kernel for
    all of them; each clip still equals its own per-clip run (where the
    default shape runs the kernel compiled with those sizes)."""
    torch, V, dev, clips = env
    shape = dict(Ns=5, Npr=3, Nb=1, Nf=2)
    picks = [1, 5]  # two RGB clips at sigma 20 (one kernel class per step); the first keeps the default shape
    items, ref = [], []
    for j, i in enumerate(picks):
        W, H, F, C, s, seed = SPECS[i]
        p1 = V.initialize_parameters_1(s, **(shape if j else {}))
        p2 = V.initialize_parameters_2(s, **(shape if j else {}))
        z = torch.from_numpy(clips[i]).to(dev)
        b, f = torch.empty_like(z), torch.empty_like(z)
        V.run_device(s, z, b, f, p1, p2, True, 1)
        ref.append((z.cpu().numpy(), b.cpu().numpy(), f.cpu().numpy()))
        items.append(V.BatchItem(s, torch.from_numpy(clips[i]).to(dev), p1, p2))
    V.run_batch(items, 1, dict(chunk_frames=0, waves_per_group=0, grab=2, groups_per_cu=4, use_graph=0,
                                persistent=persistent))
    _same(items, ref, f"mixed shapes, persistent={persistent}")
