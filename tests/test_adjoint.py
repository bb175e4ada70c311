"""Step-2 adjoint (groups, Wiener gains and weights frozen by the basic
estimate) and the Monte-Carlo SURE divergence built on it.

With the basic estimate fixed, step 2 is affine in the noisy video, so its
Jacobian J is a fixed matrix.  The FP64 oracle adjoint is pinned by the
dot-product identity <J x, y> = <x, J^T y> against the oracle's own FP64
forward step; the GPU adjoint is checked against that FP64 gradient, the
same identity on the GPU, central finite differences of the GPU forward
through a torch.autograd.Function, and an exact trace.
"""
import numpy as np
import pytest

from vbm3d_amd import synth


def _basic(oracle, sigma, noisy, color):
    b, _, _ = oracle.run_vbm3d(sigma, noisy, oracle.params_1(sigma), oracle.params_2(sigma), color, 1)
    return b


# (C, step-2 overrides, nb): DCT 7x7 (default), RGB, BIOR 8x8, Hadamard, tiled grid
CASES = [(1, {}, 1), (3, {}, 1), (1, {"k": 8, "T_2D": 1}, 1), (1, {"T_3D": 2}, 1), (1, {}, 2)]


@pytest.mark.parametrize("C,over,nb", CASES)
def test_oracle_adjoint_dot_product(oracle, C, over, nb):
    sigma = 25.0
    _, noisy = synth.noisy_clip(40, 32, 4, C, sigma, 3)
    basic = _basic(oracle, sigma, noisy, False)
    p2 = oracle.params_2(sigma, **over)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(noisy.shape) * 10).astype(np.float32)
    y = rng.standard_normal(noisy.shape)
    gz, fx = oracle.step2_adjoint_fp64(sigma, x, basic, p2, y, False, nb)
    _, f0 = oracle.step2_adjoint_fp64(sigma, np.zeros_like(x), basic, p2, y, False, nb)
    lhs, rhs = np.sum((fx - f0) * y), np.sum(x.astype(np.float64) * gz)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_oracle_adjoint_forward_is_the_fp64_step(oracle):
    """The adjoint's forward value is the FP64 replay of step 2."""
    sigma = 20.0
    _, noisy = synth.noisy_clip(40, 32, 4, 1, sigma, 5)
    basic = _basic(oracle, sigma, noisy, False)
    p2 = oracle.params_2(sigma)
    _, fin = oracle.step2_adjoint_fp64(sigma, noisy, basic, p2, np.zeros(noisy.shape), False, 1)
    num, den = oracle.step_fp64(2, sigma, noisy, basic, p2, False)
    np.testing.assert_allclose(fin, np.where(den == 0, basic, num / np.where(den == 0, 1, den)), rtol=1e-12)


def _exact_trace_fp64(oracle, sigma, noisy, basic, p2):
    """tr(J) from the FP64 forward applied to every basis vector."""
    z0 = np.zeros_like(noisy)
    _, f0 = oracle.step2_adjoint_fp64(sigma, z0, basic, p2, np.zeros(noisy.shape), False, 1)
    tr = 0.0
    flat = z0.reshape(-1)
    for i in range(flat.size):
        flat[i] = 1.0
        _, fi = oracle.step2_adjoint_fp64(sigma, z0, basic, p2, np.zeros(noisy.shape), False, 1)
        tr += fi.reshape(-1)[i] - f0.reshape(-1)[i]
        flat[i] = 0.0
    return tr


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("C,over,nb", CASES)
def test_gpu_adjoint_matches_fp64(oracle, gpu_ctx, torch_dev, C, over, nb):
    import vbm3d_amd as V
    torch, dev = torch_dev
    sigma = 25.0
    _, noisy = synth.noisy_clip(40, 32, 4, C, sigma, 3)
    basic = _basic(oracle, sigma, noisy, False)
    p2 = oracle.params_2(sigma, **over)
    vp2 = V.initialize_parameters_2(sigma)
    for f, _ in type(p2)._fields_:
        setattr(vp2, f, getattr(p2, f))
    y = np.random.default_rng(1).standard_normal(noisy.shape).astype(np.float32)
    ref, fin64 = oracle.step2_adjoint_fp64(sigma, noisy, basic, p2, y, False, nb)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    fin = torch.empty_like(t(noisy))
    g = V.step2_adjoint(sigma, t(noisy), t(basic), t(y), vp2, False, nb, final_t=fin, ctx=gpu_ctx).cpu().numpy()
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max()
    print(f"C={C} {over} nb={nb}: max|J^T y| {scale:.4g}, max err vs FP64 {err:.3g}")
    assert err <= 1e-4 * scale
    # the forward value is step 2 of the oracle run, to FP32 accuracy
    assert np.abs(fin.cpu().numpy() - fin64).max() < 1e-3


class Step2(object):
    """torch.autograd.Function over step 2 with the basic estimate fixed."""

    @staticmethod
    def make(V, torch, sigma, basic, p2, ctx):
        skip1 = V.initialize_parameters_1(sigma)
        skip1.k = 0
        class F(torch.autograd.Function):
            @staticmethod
            def forward(fctx, z):
                # step 2 alone: step 1 skipped (k = 0), the basic estimate given
                fctx.save_for_backward(z)
                out = torch.empty_like(z)
                V.run_device(sigma, z.detach().clone(), basic.clone(), out, skip1, p2, False, 1, ctx=ctx)
                return out

            @staticmethod
            def backward(fctx, g):
                (z,) = fctx.saved_tensors
                return V.step2_adjoint(sigma, z.contiguous(), basic, g.contiguous(), p2, False, 1, ctx=ctx)
        return F.apply


@pytest.fixture(scope="module")
def cfg5(torch_dev):
    """32 grey frames 640x360, sigma 25, seed 7, and its basic estimate."""
    import vbm3d_amd as V
    torch, dev = torch_dev
    sigma = 25.0
    clean, noisy = synth.noisy_clip(640, 360, 32, 1, sigma, 7)
    z = torch.from_numpy(noisy).to(dev)
    work, basic, final = z.clone(), torch.empty_like(z), torch.empty_like(z)
    V.run_device(sigma, work, basic, final, V.initialize_parameters_1(sigma), V.initialize_parameters_2(sigma),
                 False, 1)
    torch.cuda.synchronize()
    return sigma, clean, noisy, z, basic, final


@pytest.mark.gpu
def test_gpu_adjoint_dot_product_and_autograd_fd(gpu_ctx, torch_dev, cfg5):
    """On the 640x360x32 clip: <J x, y> = <x, J^T y>, then the autograd
    gradient of L = <y, f(z)> against central differences at 256 random
    pixels.  With the basic estimate fixed, f is affine, so the difference
    quotient is the same at every base point; only FP32 output rounding
    differs.  Around the noisy clip (outputs near 100, ulp 7.6e-6) the
    rounding of the thousands of outputs a pixel reaches limits the
    quotient to about 1e-3 relative even at eps = 1; around z = 0 the
    outputs are of the size of the step, so eps = 1e-3 resolves the
    gradient to FP32 accuracy."""
    import vbm3d_amd as V
    torch, dev = torch_dev
    sigma, _, noisy, z, basic, final = cfg5
    p2 = V.initialize_parameters_2(sigma)
    f = Step2.make(V, torch, sigma, basic, p2, gpu_ctx)
    gen = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(z.shape, generator=gen).to(dev)
    y = torch.randn(z.shape, generator=gen).to(dev)
    zero = torch.zeros_like(z)
    lhs = float(((f(x) - f(zero)).double() * y.double()).sum())
    jty = V.step2_adjoint(sigma, x, basic, y, p2, False, 1, ctx=gpu_ctx)
    rhs = float((x.double() * jty.double()).sum())
    print(f"dot product: {lhs:.10g} vs {rhs:.10g}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)

    zz = z.clone().requires_grad_(True)
    loss = (f(zz).double() * y.double()).sum()
    loss.backward()
    grad = zz.grad.double()
    rng = np.random.default_rng(5)
    idx = rng.choice(z.numel(), 256, replace=False)
    flat = z.reshape(-1)

    def fd(eps, base):
        out = np.empty(256)
        for j, i in enumerate(idx):
            zp, zm = base.clone(), base.clone()
            zp[i] += eps
            zm[i] -= eps
            h = float(zp[i].double() - zm[i].double())
            lp = float((f(zp.view_as(z)).double() * y.double()).sum())
            lm = float((f(zm.view_as(z)).double() * y.double()).sum())
            out[j] = (lp - lm) / h
        return out

    g_adj = grad.reshape(-1)[torch.from_numpy(idx).to(dev)].cpu().numpy()
    g3 = fd(1e-3, torch.zeros_like(flat))
    rel3 = np.linalg.norm(g_adj - g3) / np.linalg.norm(g3)
    g1 = fd(1.0, flat)
    rel1 = np.linalg.norm(g_adj - g1) / np.linalg.norm(g1)
    print(f"FD eps=1e-3 at 0: rel err {rel3:.3g} (max abs {np.abs(g_adj - g3).max():.3g}); "
          f"eps=1 at the noisy clip: rel err {rel1:.3g}; |grad| {np.linalg.norm(g3):.4g}")
    assert rel3 < 1e-4
    assert rel1 < 1e-2


@pytest.mark.gpu
def test_gpu_divergence_probe_and_exact_trace(oracle, gpu_ctx, torch_dev):
    """Per probe, b . J^T b on the GPU equals the FP64 value; the mean over
    Rademacher probes converges to tr(J) computed exactly from basis
    vectors on a small clip."""
    import vbm3d_amd as V
    torch, dev = torch_dev
    sigma = 25.0
    _, noisy = synth.noisy_clip(16, 16, 3, 1, sigma, 9)
    basic = _basic(oracle, sigma, noisy, False)
    p2o = oracle.params_2(sigma)
    p2 = V.initialize_parameters_2(sigma)
    tr = _exact_trace_fp64(oracle, sigma, noisy, basic, p2o)
    zt, bt = torch.from_numpy(noisy).to(dev), torch.from_numpy(basic).to(dev)
    rng = np.random.default_rng(11)
    est = []
    for k in range(200):
        b = (rng.integers(0, 2, noisy.shape) * 2 - 1).astype(np.float32)
        jtb = V.step2_adjoint(sigma, zt, bt, torch.from_numpy(b).to(dev), p2, False, 1, ctx=gpu_ctx)
        v = float((jtb.double().cpu().numpy() * b).sum())
        if k < 3:
            ref, _ = oracle.step2_adjoint_fp64(sigma, noisy, basic, p2o, b.astype(np.float64), False, 1)
            assert abs(v - float((ref * b).sum())) <= 1e-4 * abs(tr)
        est.append(v)
    est = np.array(est)
    se = est.std(ddof=1) / np.sqrt(len(est))
    print(f"exact tr(J) {tr:.6g}; Hutchinson mean {est.mean():.6g} +- {se:.3g}")
    assert abs(est.mean() - tr) < 4 * se


@pytest.mark.gpu
def test_gpu_sure_sweep(gpu_ctx, torch_dev, cfg5):
    """Monte-Carlo SURE (one Rademacher probe per run) across 8 sigma
    settings on the sigma-25 clip.  "full": the adjoint's step-2
    divergence plus the pilot's share by a central difference through step
    1; SURE then matches the true MSE and picks the same sigma without the
    clean frames.  "frozen" (adjoint only, pilot from the same noise) and
    "pilot" (only the Wiener step's sigma swept over a fixed sigma-25
    pilot) miss the pilot's dependence on the noise and are reported for
    comparison."""
    import vbm3d_amd as V
    torch, dev = torch_dev
    sigma_n, clean, noisy, z, basic, _ = cfg5
    sigmas = (10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0, 45.0)
    res = {}
    for mode in ("pilot", "frozen", "full"):
        rows = []
        for s in sigmas:
            sv, div, fin = V.sure(sigma_n, z, V.initialize_parameters_1(s), V.initialize_parameters_2(s), sigma=s,
                                  basic_t=basic if mode == "pilot" else None, probe_seed=int(s),
                                  pilot_eps=0.01 if mode == "full" else 0.0, ctx=gpu_ctx)
            mse = float(((fin.double().cpu().numpy() - clean.astype(np.float64)) ** 2).mean())
            rows.append((s, sv, mse, div / z.numel()))
            print(f"{mode:6s} sigma {s:4.0f}: SURE {sv:9.3f}  MSE {mse:9.3f}  div/n {div / z.numel():.4f}")
        res[mode] = np.array(rows)
    for r in res.values():
        assert np.all((r[:, 3] > 0) & (r[:, 3] < 1))
        assert np.all(np.diff(r[:, 3]) < 0)  # more smoothing, smaller divergence
    for m, r in res.items():
        print(m, "corr(SURE, MSE) =", np.corrcoef(r[:, 1], r[:, 2])[0, 1])
    r = res["full"]
    assert np.all(np.abs(r[:, 1] - r[:, 2]) <= 0.1 * r[:, 2] + 1.0)
    assert int(np.argmin(r[:, 1])) == int(np.argmin(r[:, 2]))
