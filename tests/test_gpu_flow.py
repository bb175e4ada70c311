"""GPU parity of the TV-L1 flow (flow.hip) with the oracle, which the CPU
tests pin bit for bit to the reference's own tvl1flow_lib.c.

Every pixel kernel evaluates the reference's expressions in its order and
precision, so a flow whose warps stop after the same iteration counts is
bit-identical.  The one difference allowed by design: the stopping
test's mean squared change is summed in double on the GPU and in float
by the reference, so near epsilon^2 a warp may stop one iteration apart;
the counts are compared, and bit-exactness is required whenever they
agree (with epsilon = 0 the counts are fixed at 300)."""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden", "tvl1_flow.npz")


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def pair(w, h, seed, sigma=10.0):
    from vbm3d_amd import synth
    _, noisy = synth.noisy_clip(w, h, 2, 1, sigma, seed)
    return noisy[0, 0].copy(), noisy[1, 0].copy()


def gpu_flow(I0, I1, ctx, tau=0.25, lam=0.15, theta=0.3, nscales=100, fscale=0, zfactor=0.5, warps=5,
             epsilon=0.01):
    import vbm3d_amd as V
    p = V.tvl1_params(tau=tau, lam=lam, theta=theta, nscales=nscales, fscale=fscale, zfactor=zfactor,
                      nwarps=warps, epsilon=epsilon)
    return V.tvl1_flow(I0, I1, p, ctx=ctx)


CASES = {
    "default-64x48": (64, 48, 2, {}),
    "eps0-40x32": (40, 32, 3, {"epsilon": 0.0, "warps": 2, "nscales": 2}),
    "fscale1-72x56": (72, 56, 4, {"fscale": 1}),
    "z07-w3-80x40": (80, 40, 5, {"zfactor": 0.7, "warps": 3}),
    "odd-37x29": (37, 29, 6, {"tau": 0.2, "lam": 0.3, "theta": 0.5}),
    "default-160x120": (160, 120, 7, {}),
}


@pytest.mark.parametrize("case", list(CASES), ids=list(CASES))
def test_flow_matches_oracle(gpu_ctx, case):
    from oracle import tvl1 as T
    w, h, seed, kw = CASES[case]
    I0, I1 = pair(w, h, seed)
    u, v, it = gpu_flow(I0, I1, gpu_ctx, **kw)
    okw = dict(kw)
    ou, ov, oit = T.flow(I0, I1, **okw)
    np.testing.assert_array_equal(it, oit)
    assert bits_equal(u, ou) and bits_equal(v, ov), (np.abs(u - ou).max(), np.abs(v - ov).max())


def test_golden_flows(gpu_ctx):
    g = np.load(GOLDEN)
    for n in sorted({k.split("/")[0] for k in g.files}):
        fscale, zf, warps, eps, nsc = g[f"{n}/params"]
        u, v, _ = gpu_flow(g[f"{n}/I0"], g[f"{n}/I1"], gpu_ctx, fscale=int(fscale), zfactor=float(zf),
                           warps=int(warps), epsilon=float(eps), nscales=int(nsc))
        assert bits_equal(u, g[f"{n}/u"]) and bits_equal(v, g[f"{n}/v"]), n


def test_device_entry_equals_host_entry(gpu_ctx):
    import torch
    import vbm3d_amd as V
    I0, I1 = pair(96, 64, 8)
    u, v, it = gpu_flow(I0, I1, gpu_ctx)
    p = V.tvl1_params()
    p.nscales = V.tvl1_nscales(96, 64)
    d0 = torch.from_numpy(I0).to("cuda:0")
    d1 = torch.from_numpy(I1).to("cuda:0")
    du = torch.empty_like(d0)
    dv = torch.empty_like(d0)
    torch.cuda.synchronize()
    it2 = V.tvl1_flow_device(d0, d1, du, dv, p, ctx=gpu_ctx)
    np.testing.assert_array_equal(it, it2)
    assert bits_equal(du.cpu().numpy(), u) and bits_equal(dv.cpu().numpy(), v)


def test_larger_pair_against_oracle(gpu_ctx):
    """640x360 (seven levels): iteration counts equal in every warp, and
    then the flows are bit-identical; otherwise within 0.05 px."""
    from oracle import tvl1 as T
    I0, I1 = pair(640, 360, 9, sigma=20.0)
    u, v, it = gpu_flow(I0, I1, gpu_ctx)
    ou, ov, oit = T.flow(I0, I1)
    assert it.shape == oit.shape
    if np.array_equal(it, oit):
        assert bits_equal(u, ou) and bits_equal(v, ov), (np.abs(u - ou).max(), np.abs(v - ov).max())
    else:
        assert np.abs(it - oit).max() <= 1
        assert np.abs(u - ou).max() < 0.05 and np.abs(v - ov).max() < 0.05


def test_translation_recovered(gpu_ctx):
    from oracle import tvl1 as T
    rng = np.random.default_rng(3)
    n = 128
    base = T.gaussian(rng.uniform(0, 255, (n, n)).astype(np.float32), 2.0)
    k = np.fft.fftfreq(n)
    ph = np.exp(-2j * np.pi * (k[None, :] * -0.75 + k[:, None] * 2.0))
    shifted = np.real(np.fft.ifft2(np.fft.fft2(base.astype(np.float64)) * ph)).astype(np.float32)
    u, v, _ = gpu_flow(base, shifted, gpu_ctx)
    c = slice(20, n - 20)
    assert abs(np.median(u[c, c]) + 0.75) < 0.05 and abs(np.median(v[c, c]) - 2.0) < 0.05


def test_invalid_arguments_refused(gpu_ctx):
    import vbm3d_amd as V
    I0, I1 = pair(48, 40, 10)
    with pytest.raises(V.VBM3DError):
        V.tvl1_flow(I0[:1, :], I1[:1, :], ctx=gpu_ctx)          # one row
    with pytest.raises(V.VBM3DError):
        V.tvl1_flow(I0, I1, V.tvl1_params(zfactor=1.5), ctx=gpu_ctx)
    with pytest.raises(V.VBM3DError):
        V.tvl1_flow(I0[:5, :5], I1[:5, :5], V.tvl1_params(nscales=1), ctx=gpu_ctx)  # Gaussian as wide as the image
    with pytest.raises(ValueError):
        V.tvl1_flow(I0, I1[:, :-1], ctx=gpu_ctx)


def test_tvl1flow_command_line(gpu_ctx, tmp_path):
    """bin/tvl1flow on two PNG frames writes the .flo the API computes
    (grey frames: the file's samples are the planes)."""
    import subprocess
    from PIL import Image
    import vbm3d_amd as V
    I0, I1 = pair(64, 48, 11, sigma=0.5)
    a = np.clip(np.rint(I0), 0, 255).astype(np.uint8)
    b = np.clip(np.rint(I1), 0, 255).astype(np.uint8)
    Image.fromarray(a).save(tmp_path / "a.png")
    Image.fromarray(b).save(tmp_path / "b.png")
    exe = os.path.join(ROOT, "vbm3d_amd", "bin", "tvl1flow")
    r = subprocess.run([exe, str(tmp_path / "a.png"), str(tmp_path / "b.png"), str(tmp_path / "f.flo")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "f.flo", np.float32)
    assert raw[0] == np.float32(202021.25)
    dims = raw[1:3].view(np.int32)
    assert tuple(dims) == (64, 48)
    uv = raw[3:].reshape(48, 64, 2)
    u, v, _ = V.tvl1_flow(a.astype(np.float32), b.astype(np.float32), ctx=gpu_ctx)
    assert bits_equal(uv[..., 0], u) and bits_equal(uv[..., 1], v)
