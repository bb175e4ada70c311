"""Python mirror of the reference's denoising interface over the C ABI.

Names, argument meaning and error behaviour follow the reference:
``Parameters`` (src/parameters.h:21-58), ``initialize_parameters_1/_2``
(src/main.cpp:35-207) and ``run_vbm3d`` (src/vbm3d.cpp:87-376), with
numpy arrays in the reference's planar [F][C][H][W] layout in place of
``Video<float>``.  Everything runs through libvbm3d_hip.so on the GPU;
there is no CPU fallback: a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvbm3d_hip.so")

NONE = -1
DCT, BIOR, HADAMARD, HAAR = 0, 1, 2, 3


class VBM3DError(RuntimeError):
    """Error returned through the C ABI (code < 0)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"vbm3d error {code}: {msg}")
        self.code = code


class Parameters(C.Structure):
    """struct Parameters (src/parameters.h:21-58), field for field."""
    _fields_ = [("k", C.c_uint), ("kt", C.c_uint), ("p", C.c_uint), ("N", C.c_uint),
                ("tau", C.c_float), ("d", C.c_float),
                ("Nf", C.c_uint), ("Ns", C.c_uint), ("Npr", C.c_uint), ("Nb", C.c_uint),
                ("T_2D", C.c_uint), ("T_3D", C.c_uint), ("lambda3D", C.c_float),
                ("mc", C.c_int), ("n", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

    def copy(self):
        o = Parameters()
        C.memmove(C.byref(o), C.byref(self), C.sizeof(Parameters))
        return o


class Clip(C.Structure):
    """POD mirror of vbm3d_clip: one clip of a batch (device pointers)."""
    _fields_ = [("noisy", C.c_void_p), ("basic", C.c_void_p), ("denoised", C.c_void_p),
                ("fflow", C.c_void_p), ("bflow", C.c_void_p),
                ("W", C.c_uint), ("H", C.c_uint), ("F", C.c_uint), ("C", C.c_uint),
                ("sigma", C.c_float), ("p1", Parameters), ("p2", Parameters), ("color_space", C.c_int)]


class BatchConfig(C.Structure):
    """POD mirror of vbm3d_batch_config."""
    _fields_ = [("chunk_frames", C.c_uint), ("waves_per_group", C.c_int), ("grab", C.c_int),
                ("groups_per_cu", C.c_int), ("use_graph", C.c_int), ("persistent", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_lock = threading.Lock()


def lib():
    """Load libvbm3d_hip.so (built by vbm3d_amd.build); raises if absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: run `python -m vbm3d_amd.build` first")
            # PyTorch carries its own HIP runtime.  Loaded first, it satisfies
            # this library's HIP dependency, so the process holds one runtime;
            # two runtimes in one process cannot both open the GPU.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            L = C.CDLL(LIB_PATH)
            fp, up, pp = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.POINTER(Parameters)
            L.vbm3d_create.argtypes = [C.c_int]
            L.vbm3d_create.restype = C.c_void_p
            L.vbm3d_destroy.argtypes = [C.c_void_p]
            L.vbm3d_last_error.restype = C.c_char_p
            L.vbm3d_default_params_1.argtypes = [pp] + [C.c_int] * 9 + [C.c_float, C.c_float, C.c_uint, C.c_uint, C.c_float]
            L.vbm3d_default_params_2.argtypes = [pp] + [C.c_int] * 9 + [C.c_float, C.c_uint, C.c_uint, C.c_float]
            L.vbm3d_run.argtypes = [C.c_void_p, C.c_float, fp, fp, fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                    pp, pp, C.c_int, C.c_int, fp, fp]
            L.vbm3d_run_device.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_uint, C.c_uint, C.c_uint, C.c_uint, pp, pp, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
            L.vbm3d_step2_adjoint.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_uint, pp, C.c_int,
                                              C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            L.vbm3d_run_batch.argtypes = [C.c_void_p, C.POINTER(Clip), C.c_int, C.c_int,
                                          C.POINTER(BatchConfig), C.c_void_p]
            L.vbm3d_autotune_batch.argtypes = [C.c_void_p, C.POINTER(Clip), C.c_int, C.c_int,
                                               C.POINTER(BatchConfig), C.c_void_p]
            L.vbm3d_block_match.argtypes = [C.c_void_p, fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, fp, fp, pp,
                                            C.c_long, up, up, fp, up, up]
            L.vbm3d_step.argtypes = [C.c_void_p, C.c_int, C.c_float, fp, fp, fp, fp, C.c_uint, C.c_uint,
                                     C.c_uint, C.c_uint, pp, C.c_int, C.c_int, fp, fp]
            L.vbm3d_grid.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_int, C.c_int, up, up, up]
            L.vbm3d_grid.restype = C.c_long
            L.vbm3d_selftest_partial_sort.argtypes = [fp, up, C.c_uint, C.c_uint]
            L.vbm3d_profile.argtypes = [C.c_void_p, C.c_int]
            L.vbm3d_set_chunk_frames.argtypes = [C.c_void_p, C.c_uint]
            L.vbm3d_profile_read.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_long), C.POINTER(C.c_double)]
            _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise VBM3DError(rc, lib().vbm3d_last_error().decode())


def _f(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def initialize_parameters_1(sigma, k=NONE, kt=NONE, Nf=NONE, Ns=NONE, Npr=NONE, Nb=NONE, p=NONE,
                            N=NONE, d=NONE, tau=NONE, lambda3D=NONE, T_2D=NONE, T_3D=NONE) -> Parameters:
    """Step-1 defaults (initializeParameters_1, src/main.cpp:35-125)."""
    o = Parameters()
    lib().vbm3d_default_params_1(C.byref(o), k, kt, Nf, Ns, Npr, Nb, p, N, d, float(tau), float(lambda3D),
                                 T_2D & 0xFFFFFFFF, T_3D & 0xFFFFFFFF, float(sigma))
    return o


def initialize_parameters_2(sigma, k=NONE, kt=NONE, Nf=NONE, Ns=NONE, Npr=NONE, Nb=NONE, p=NONE,
                            N=NONE, d=NONE, tau=NONE, T_2D=NONE, T_3D=NONE) -> Parameters:
    """Step-2 defaults (initializeParameters_2, src/main.cpp:127-207)."""
    o = Parameters()
    lib().vbm3d_default_params_2(C.byref(o), k, kt, Nf, Ns, Npr, Nb, p, N, d, float(tau),
                                 T_2D & 0xFFFFFFFF, T_3D & 0xFFFFFFFF, float(sigma))
    return o


class Context:
    """A device context (owns device buffers and a HIP stream)."""

    def __init__(self, device: int = 0):
        self.h = lib().vbm3d_create(device)
        if not self.h:
            raise VBM3DError(-2, lib().vbm3d_last_error().decode())

    def set_chunk_frames(self, frames: int):
        """Frames of reference patches per chunk for the per-clip entries:
        0 (default) = as many as the memory budget allows, 1 = one launch
        per kernel and frame (the plain per-frame path).  Results do not
        depend on it."""
        _check(lib().vbm3d_set_chunk_frames(self.h, int(frames)))

    def close(self):
        if self.h:
            lib().vbm3d_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _video(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or a.ndim != 4 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous float32 array [F][C][H][W]")
    return a


def _flow(fl, name, F, H, W):
    """A flow as the C ABI reads it: float32 [F'][2][H][W] with F' >= F - 1
    (src/main.cpp:353-369), or None (zero flow).  A one-frame clip has no
    predictive search, so its flows are never read."""
    if fl is None or F <= 1:
        return None
    a = np.ascontiguousarray(fl, np.float32)
    if a.ndim != 4 or a.shape[1:] != (2, H, W) or a.shape[0] < F - 1:
        raise ValueError(f"{name} must have shape ({F - 1}, 2, {H}, {W}), got {a.shape}")
    return a


def run_vbm3d(sigma, vid_noisy, fflow, bflow, vid_basic, vid_denoised, prms_1, prms_2,
              color_space=True, ref_grid_nb=1, ctx: Context | None = None) -> int:
    """run_vbm3d (src/vbm3d.cpp:87-376).  vid_noisy is mutated by the
    colour round trip like the reference; vid_basic / vid_denoised are
    filled in place (vid_basic is read when prms_1.k == 0).  Returns 0
    (EXIT_SUCCESS); raises VBM3DError on failure."""
    _video(vid_noisy, "vid_noisy")
    _video(vid_basic, "vid_basic")
    _video(vid_denoised, "vid_denoised")
    F, Cc, H, W = vid_noisy.shape
    if vid_basic.shape != vid_noisy.shape or vid_denoised.shape != vid_noisy.shape:
        raise ValueError("vid_basic / vid_denoised must have vid_noisy's shape")
    ff = _flow(fflow, "fflow", F, H, W)
    bf = _flow(bflow, "bflow", F, H, W)
    ctx = ctx or default_context()
    _check(lib().vbm3d_run(ctx.h, float(sigma), _f(vid_noisy), _f(ff), _f(bf), W, H, F, Cc,
                           C.byref(prms_1), C.byref(prms_2), int(bool(color_space)), int(ref_grid_nb),
                           _f(vid_basic), _f(vid_denoised)))
    return 0


def denoise(noisy, sigma, color_space=True, ref_grid_nb=1, prms_1=None, prms_2=None, ctx=None):
    """Convenience: returns (basic, final) for a float32 [F][C][H][W] clip
    without touching the input."""
    nz = np.array(noisy, np.float32, copy=True, order="C")
    basic = np.zeros_like(nz)
    final = np.zeros_like(nz)
    p1 = prms_1 or initialize_parameters_1(sigma)
    p2 = prms_2 or initialize_parameters_2(sigma)
    run_vbm3d(sigma, nz, None, None, basic, final, p1, p2, color_space, ref_grid_nb, ctx)
    return basic, final


def run_device(sigma, noisy_t, basic_t, final_t, prms_1, prms_2, color_space=True, ref_grid_nb=1,
               fflow_t=None, bflow_t=None, stream=None, ctx: Context | None = None) -> int:
    """Device-resident run on torch CUDA tensors (float32, contiguous,
    [F][C][H][W]); noisy_t is mutated like the reference's vid_noisy."""
    F, Cc, H, W = noisy_t.shape
    ctx = ctx or default_context()
    sp = None
    if stream is not None:
        sp = C.c_void_p(int(stream.cuda_stream))
    _check(lib().vbm3d_run_device(ctx.h, float(sigma), C.c_void_p(noisy_t.data_ptr()),
                                  None if fflow_t is None else C.c_void_p(fflow_t.data_ptr()),
                                  None if bflow_t is None else C.c_void_p(bflow_t.data_ptr()),
                                  W, H, F, Cc, C.byref(prms_1), C.byref(prms_2), int(bool(color_space)),
                                  int(ref_grid_nb), C.c_void_p(basic_t.data_ptr()),
                                  C.c_void_p(final_t.data_ptr()), sp))
    return 0


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def step2_adjoint(sigma, noisy_t, basic_t, grad_t, prms_2, color_space=True, ref_grid_nb=1,
                  fflow_t=None, bflow_t=None, final_t=None, stream=None, ctx: Context | None = None):
    """J^T grad_t for step 2 with the basic estimate fixed (groups, Wiener
    gains and weights frozen), on torch CUDA tensors [F][C][H][W] in the
    colour-transformed space; the step-2 output goes to final_t if given.
    Returns the gradient with respect to noisy_t (a new tensor)."""
    F, Cc, H, W = noisy_t.shape
    for t in (noisy_t, basic_t, grad_t):
        if tuple(t.shape) != (F, Cc, H, W) or not t.is_contiguous():
            raise ValueError("noisy, basic and grad must be contiguous and of one shape")
    ctx = ctx or default_context()
    out = noisy_t.new_empty(noisy_t.shape)
    sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
    _check(lib().vbm3d_step2_adjoint(ctx.h, float(sigma), _ptr(noisy_t), _ptr(basic_t), _ptr(fflow_t),
                                     _ptr(bflow_t), W, H, F, Cc, C.byref(prms_2), int(bool(color_space)),
                                     int(ref_grid_nb), _ptr(grad_t), _ptr(out), _ptr(final_t), sp))
    return out


def sure(sigma_noise, noisy_t, prms_1, prms_2, sigma=None, basic_t=None, probe_seed=0, ref_grid_nb=1,
         pilot_eps=0.01, stream=None, ctx: Context | None = None):
    """Monte-Carlo SURE of the denoiser on a grey clip without clean frames:
        SURE = |f(y) - y|^2 / n - s^2 + 2 s^2 div f / n,   s = sigma_noise,
    with div f estimated by one Rademacher probe b as b . J^T b, J the
    step-2 Jacobian with the basic estimate frozen (the hand-written
    adjoint).  When the pilot is computed here from y, its own share of
    the divergence -- y -> step 1 -> basic -> Wiener gains -> output -- is
    added as a central difference along b with step pilot_eps*sigma_noise
    (0 leaves it out).  `sigma` is the denoiser's parameter (default
    sigma_noise).  With basic_t given, step 1 is skipped and that pilot is
    used as a constant.  Returns (sure, divergence, final)."""
    import torch
    if noisy_t.shape[1] != 1:
        raise ValueError("sure() takes grey clips (the divergence is taken in the filter's space)")
    sigma = float(sigma_noise if sigma is None else sigma)
    y = noisy_t.contiguous()
    work = y.clone()
    final = torch.empty_like(y)
    if basic_t is None:
        basic = torch.empty_like(y)
        p1 = prms_1
    else:
        basic = basic_t.contiguous().clone()
        p1 = prms_1.copy()
        p1.k = 0
    run_device(sigma, work, basic, final, p1, prms_2, False, ref_grid_nb, stream=stream, ctx=ctx)
    g = torch.Generator(device="cpu").manual_seed(int(probe_seed))
    b = (torch.randint(0, 2, tuple(y.shape), generator=g, dtype=torch.int8).to(torch.float32) * 2 - 1).to(y.device)
    jtb = step2_adjoint(sigma, y, basic, b, prms_2, False, ref_grid_nb, stream=stream, ctx=ctx)
    n = y.numel()
    div = float((b.double() * jtb.double()).sum())
    if basic_t is None and pilot_eps:
        # the pilot's share of the divergence, b . (d final / d basic)(d basic / d y) b,
        # as a central difference through step 1 and the Wiener gains
        skip2 = prms_2.copy()
        skip2.k = 0
        skip1 = prms_1.copy()
        skip1.k = 0
        eps = float(pilot_eps) * float(sigma_noise)
        outs = []
        for sgn in (1.0, -1.0):
            yp = (y + sgn * eps * b).contiguous()
            bp, fp = torch.empty_like(y), torch.empty_like(y)
            run_device(sigma, yp.clone(), bp, fp, prms_1, skip2, False, ref_grid_nb, stream=stream, ctx=ctx)
            f2 = torch.empty_like(y)
            run_device(sigma, y.clone(), bp, f2, skip1, prms_2, False, ref_grid_nb, stream=stream, ctx=ctx)
            outs.append(f2)
        div += float((b.double() * (outs[0].double() - outs[1].double())).sum()) / (2.0 * eps)
    s2 = float(sigma_noise) ** 2
    res = float(((final.double() - y.double()) ** 2).sum()) / n
    return res - s2 + 2.0 * s2 * div / n, div, final


class BatchItem:
    """One clip of a batch: torch CUDA tensors [F][C][H][W] (float32,
    contiguous) and its parameters.  `basic` and `denoised` are allocated
    when not given; `noisy` is mutated like vid_noisy of run_vbm3d."""

    def __init__(self, sigma, noisy, prms_1=None, prms_2=None, color_space=True, basic=None,
                 denoised=None, fflow=None, bflow=None):
        import torch
        self.sigma = float(sigma)
        self.noisy = noisy
        self.basic = basic if basic is not None else torch.empty_like(noisy)
        self.denoised = denoised if denoised is not None else torch.empty_like(noisy)
        self.fflow, self.bflow = fflow, bflow
        self.p1 = prms_1 if prms_1 is not None else initialize_parameters_1(sigma)
        self.p2 = prms_2 if prms_2 is not None else initialize_parameters_2(sigma)
        self.color_space = bool(color_space)

    def as_clip(self) -> Clip:
        for t in (self.noisy, self.basic, self.denoised):
            if not t.is_contiguous() or t.shape != self.noisy.shape:
                raise ValueError("batch buffers must be contiguous and of one shape")
        F, Cc, H, W = self.noisy.shape
        return Clip(_ptr(self.noisy).value, _ptr(self.basic).value, _ptr(self.denoised).value,
                    None if self.fflow is None else self.fflow.data_ptr(),
                    None if self.bflow is None else self.bflow.data_ptr(),
                    W, H, F, Cc, self.sigma, self.p1, self.p2, int(self.color_space))


def _clips(items):
    arr = (Clip * len(items))()
    for i, it in enumerate(items):
        arr[i] = it.as_clip()
    return arr


def run_batch(items, ref_grid_nb=1, config: BatchConfig | dict | None = None, stream=None,
              ctx: Context | None = None):
    """Denoise a list of BatchItem together (vbm3d_run_batch): persistent
    work-queue launches over all clips, captured in a HIP graph and
    replayed on later identical calls.  config None uses the context's
    autotuned configuration (or the defaults)."""
    ctx = ctx or default_context()
    cfg = None
    if config is not None:
        cfg = config if isinstance(config, BatchConfig) else BatchConfig(**config)
    sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
    arr = _clips(items)
    _check(lib().vbm3d_run_batch(ctx.h, arr, len(items), int(ref_grid_nb),
                                 None if cfg is None else C.byref(cfg), sp))


def autotune_batch(items, ref_grid_nb=1, stream=None, ctx: Context | None = None) -> dict:
    """Time launch shapes of the batched form on these clips; the context
    keeps the fastest for run_batch(config=None).  Inputs are left as
    they were."""
    ctx = ctx or default_context()
    best = BatchConfig()
    sp = None if stream is None else C.c_void_p(int(stream.cuda_stream))
    _check(lib().vbm3d_autotune_batch(ctx.h, _clips(items), len(items), int(ref_grid_nb), C.byref(best), sp))
    return best.as_dict()


def block_match(vid, prm, ref_idx, fflow=None, bflow=None, ctx=None):
    """computeSimilarPatches (src/vbm3d.cpp:1082-1197) for a list of
    reference indices t*W*H*C + y*W + x.  Returns (idx[n,N], dist[n,N],
    nsim[n], nsx[n])."""
    vid = np.ascontiguousarray(vid, np.float32)
    F, Cc, H, W = vid.shape
    ref_idx = np.ascontiguousarray(ref_idx, np.uint32)
    n = len(ref_idx)
    N = prm.N
    idx = np.empty((n, N), np.uint32)
    dist = np.empty((n, N), np.float32)
    nsim = np.empty(n, np.uint32)
    nsx = np.empty(n, np.uint32)
    ff = _flow(fflow, "fflow", F, H, W)
    bf = _flow(bflow, "bflow", F, H, W)
    ctx = ctx or default_context()
    _check(lib().vbm3d_block_match(ctx.h, _f(vid), W, H, F, Cc, _f(ff), _f(bf), C.byref(prm), n,
                                   _u(ref_idx), _u(idx), _f(dist), _u(nsim), _u(nsx)))
    return idx, dist, nsim, nsx


def step(stepno, sigma, noisy, basic, prm, color_space=True, ref_grid_nb=1, fflow=None, bflow=None,
         ctx=None):
    """One step on a colour-transformed video; returns (num, den)."""
    nz = np.ascontiguousarray(noisy, np.float32)
    F, Cc, H, W = nz.shape
    bs = None if basic is None else np.ascontiguousarray(basic, np.float32)
    ff = _flow(fflow, "fflow", F, H, W)
    bf = _flow(bflow, "bflow", F, H, W)
    num = np.empty_like(nz)
    den = np.empty_like(nz)
    ctx = ctx or default_context()
    _check(lib().vbm3d_step(ctx.h, stepno, float(sigma), _f(nz), _f(bs), _f(ff), _f(bf), W, H, F, Cc,
                            C.byref(prm), int(bool(color_space)), int(ref_grid_nb), _f(num), _f(den)))
    return num, den


def grid(W, H, k, p, n=16, nb=1):
    """Reference-patch grid of one frame: (xs, ys, crop) in processing order."""
    cnt = lib().vbm3d_grid(W, H, k, p, n, nb, None, None, None)
    if cnt < 0:
        raise ValueError("invalid grid")
    xs = np.empty(cnt, np.uint32)
    ys = np.empty(cnt, np.uint32)
    cr = np.empty(cnt, np.uint32)
    lib().vbm3d_grid(W, H, k, p, n, nb, _u(xs), _u(ys), _u(cr))
    return xs, ys, cr


def selftest_partial_sort(dist, idx, mid):
    d = np.array(dist, np.float32, copy=True)
    i = np.array(idx, np.uint32, copy=True)
    lib().vbm3d_selftest_partial_sort(_f(d), _u(i), len(d), mid)
    return d, i


KERNELS = {"block_match": 0, "group_filter": 1, "aggregate": 2, "colour": 3, "bm_pixels": 4}


def profile(mode, ctx=None):
    """Reset kernel profiling: 0 off, 1 (or True) event timing, 2 timing
    plus work counters (device atomics: use on an untimed run)."""
    ctx = ctx or default_context()
    _check(lib().vbm3d_profile(ctx.h, int(mode)))


def profile_read(kernel: str, ctx=None, step: int = 0):
    """(total_ms, launches, work) for one kernel class since profile(True);
    step 1 or 2 restricts the figures to that step, 0 sums both."""
    ctx = ctx or default_context()
    if step not in (0, 1, 2):
        raise ValueError("step must be 0, 1 or 2")
    ms, n, w = C.c_double(), C.c_long(), C.c_double()
    code = KERNELS[kernel] + 8 * step
    _check(lib().vbm3d_profile_read(ctx.h, code, C.byref(ms), C.byref(n), C.byref(w)))
    return ms.value, n.value, w.value


class TVL1Params(C.Structure):
    """POD mirror of vbm3d_tvl1_params: the TV-L1 flow's settings, in the
    order tvl1flow/main.c reads them from its command line."""
    _fields_ = [("tau", C.c_float), ("lambda_", C.c_float), ("theta", C.c_float), ("nscales", C.c_int),
                ("fscale", C.c_int), ("zfactor", C.c_float), ("nwarps", C.c_int), ("epsilon", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _bind_tvl1(L):
    if getattr(L, "_tvl1_bound", False):
        return L
    fp, tp, ip = C.POINTER(C.c_float), C.POINTER(TVL1Params), C.POINTER(C.c_int)
    L.vbm3d_tvl1_default_params.argtypes = [tp]
    L.vbm3d_tvl1_nscales.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int]
    L.vbm3d_tvl1_flow.argtypes = [C.c_void_p, fp, fp, fp, fp, C.c_int, C.c_int, tp, ip]
    L.vbm3d_tvl1_flow_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                         C.c_int, tp, ip]
    L._tvl1_bound = True
    return L


def tvl1_params(tau=None, lam=None, theta=None, nscales=None, fscale=None, zfactor=None, nwarps=None,
                epsilon=None) -> TVL1Params:
    """The flow command line's defaults (tvl1flow/main.c:20-30), with any
    value given here replacing its default."""
    p = TVL1Params()
    _bind_tvl1(lib()).vbm3d_tvl1_default_params(C.byref(p))
    for name, val in (("tau", tau), ("lambda_", lam), ("theta", theta), ("nscales", nscales), ("fscale", fscale),
                      ("zfactor", zfactor), ("nwarps", nwarps), ("epsilon", epsilon)):
        if val is not None:
            setattr(p, name, val)
    return p


def tvl1_nscales(nx, ny, zfactor=0.5, nscales=100) -> int:
    """Pyramid levels as the reference's command line caps them
    (tvl1flow/main.c:139-146)."""
    return _bind_tvl1(lib()).vbm3d_tvl1_nscales(nx, ny, zfactor, nscales)


def tvl1_flow(I0, I1, params: TVL1Params | None = None, cap_scales=True, ctx=None):
    """TV-L1 optical flow from grey image I0 to I1 ([ny][nx] float32) on
    the GPU (Dual_TVL1_optic_flow_multiscale, tvl1flow/tvl1flow_lib.c:
    384-485).  cap_scales applies the command line's level cap.  Returns
    (u, v, iterations [solved level, coarsest first][warp])."""
    L = _bind_tvl1(lib())
    I0 = np.ascontiguousarray(I0, np.float32)
    I1 = np.ascontiguousarray(I1, np.float32)
    if I0.ndim != 2 or I0.shape != I1.shape:
        raise ValueError(f"I0 and I1 must be equal 2-D arrays, got {I0.shape} and {I1.shape}")
    ny, nx = I0.shape
    p = tvl1_params() if params is None else params
    if cap_scales:
        p = TVL1Params(**{f: getattr(p, f) for f, _ in TVL1Params._fields_})
        p.nscales = L.vbm3d_tvl1_nscales(nx, ny, p.zfactor, p.nscales)
        p.fscale = min(p.fscale, p.nscales)
    u = np.zeros((ny, nx), np.float32)
    v = np.zeros((ny, nx), np.float32)
    nl = max(p.nscales - p.fscale, 0)
    iters = np.zeros(max(nl * p.nwarps, 1), np.int32)
    c = ctx or default_context()
    _check(L.vbm3d_tvl1_flow(c.h, _f(I0), _f(I1), _f(u), _f(v), nx, ny, C.byref(p),
                             iters.ctypes.data_as(C.POINTER(C.c_int))))
    return u, v, iters[: nl * p.nwarps].reshape(nl, p.nwarps)


def tvl1_flow_device(I0_t, I1_t, u_t, v_t, params: TVL1Params | None = None, ctx=None):
    """The same on device tensors ([ny][nx] float32, contiguous); u_t and
    v_t receive the flow.  Returns the iteration counts."""
    L = _bind_tvl1(lib())
    ny, nx = I0_t.shape
    for t in (I1_t, u_t, v_t):
        if tuple(t.shape) != (ny, nx):
            raise ValueError("all four tensors must have I0's shape")
    p = tvl1_params() if params is None else params
    nl = max(p.nscales - p.fscale, 0)
    iters = np.zeros(max(nl * p.nwarps, 1), np.int32)
    c = ctx or default_context()
    _check(L.vbm3d_tvl1_flow_device(c.h, _ptr(I0_t), _ptr(I1_t), _ptr(u_t), _ptr(v_t), nx, ny, C.byref(p),
                                    iters.ctypes.data_as(C.POINTER(C.c_int))))
    return iters[: nl * p.nwarps].reshape(nl, p.nwarps)


def bench_transforms(K=7, mode=0, npatch=1 << 22, reps=32):
    """The transform micro-benchmark (vbm3d_bench_transforms): k x k DCT
    round trips on the VALU (mode 0, the denoiser's exact transforms) or
    the matrix cores (mode 1, FP32 MFMA).  Returns (ms per launch, 2D
    transforms per second, largest round-trip error)."""
    L = lib()
    if not getattr(L, "_xb_bound", False):
        L.vbm3d_bench_transforms.argtypes = [C.c_int, C.c_int, C.c_long, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
        L._xb_bound = True
    ms, err = C.c_double(), C.c_double()
    _check(L.vbm3d_bench_transforms(K, mode, npatch, reps, C.byref(ms), C.byref(err)))
    return ms.value, 2.0 * npatch * reps / (ms.value * 1e-3), err.value


# ---- multiscale pyramid (multiscaler/) and multiscale denoising
PYR_DCT, PYR_GAUSSIAN, PYR_LANCZOS = 0, 1, 2


def _bind_pyr(L):
    if getattr(L, "_pyr_bound", False):
        return L
    fp, ip, pp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(Parameters)
    L.vbm3d_pyr_levels.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip]
    L.vbm3d_pyr_decompose.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp]
    L.vbm3d_pyr_recompose.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, fp]
    L.vbm3d_pyr_merge_coarse.argtypes = [C.c_void_p, fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, C.c_int, C.c_float,
                                         fp]
    L.vbm3d_multiscale_denoise.argtypes = [C.c_void_p, C.c_float, fp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, pp, pp,
                                           C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, fp, fp, fp]
    L._pyr_bound = True
    return L


def _planar(img):
    a = np.ascontiguousarray(img, np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("an image is [c][rows][cols] (or [rows][cols])")
    return a


def pyr_levels(rows, cols, levels):
    """Level sizes: each level halves the previous (decompose.cpp:44-53)."""
    L = _bind_pyr(lib())
    lr = np.zeros(levels, np.int32)
    lc = np.zeros(levels, np.int32)
    _check(L.vbm3d_pyr_levels(rows, cols, levels, lr.ctypes.data_as(C.POINTER(C.c_int)),
                              lc.ctypes.data_as(C.POINTER(C.c_int))))
    return list(zip(lr.tolist(), lc.tolist()))


def pyr_decompose(img, levels, type_=PYR_DCT, ctx=None):
    """multiscaler's decompose on the GPU: the list of levels of one image."""
    L = _bind_pyr(lib())
    a = _planar(img)
    ch, rows, cols = a.shape
    sz = pyr_levels(rows, cols, levels)
    out = np.zeros(sum(ch * r * c for r, c in sz), np.float32)
    c = ctx or default_context()
    _check(L.vbm3d_pyr_decompose(c.h, _f(a), rows, cols, ch, levels, type_, _f(out)))
    res, o = [], 0
    for r, cc in sz:
        res.append(out[o:o + ch * r * cc].reshape(ch, r, cc).copy())
        o += ch * r * cc
    return res


def pyr_recompose(levels, type_=PYR_DCT, factor=0.8, ctx=None):
    """multiscaler's recompose on the GPU (-c factor, -t type)."""
    L = _bind_pyr(lib())
    lev = [_planar(l) for l in levels]
    ch, rows, cols = lev[0].shape
    if [tuple(l.shape[1:]) for l in lev] != pyr_levels(rows, cols, len(lev)) or any(l.shape[0] != ch for l in lev):
        raise ValueError("levels must halve the previous one's size and share the channel count")
    flat = np.concatenate([l.ravel() for l in lev])
    out = np.zeros((ch, rows, cols), np.float32)
    c = ctx or default_context()
    _check(L.vbm3d_pyr_recompose(c.h, _f(flat), rows, cols, ch, len(lev), type_, factor, _f(out)))
    return out


def pyr_merge_coarse(fine, coarse, factor=0.8, ctx=None):
    """multiscaler's merge_coarse on the GPU."""
    L = _bind_pyr(lib())
    f, g = _planar(fine), _planar(coarse)
    out = np.zeros_like(f)
    c = ctx or default_context()
    _check(L.vbm3d_pyr_merge_coarse(c.h, _f(f), f.shape[1], f.shape[2], _f(g), g.shape[1], g.shape[2], f.shape[0],
                                    factor, _f(out)))
    return out


def multiscale_denoise(noisy, sigma, levels=3, type_=PYR_DCT, factor=0.7, prms_1=None, prms_2=None,
                       color_space=True, ref_grid_nb=1, return_levels=False, ctx=None):
    """The reference's multiscale denoising (src/utils/VBM3Ddenoising_
    multiscale.sh, defaults: 3 levels, DCT pyramid, recomposition ratio
    0.7) of a [F][C][H][W] float32 video.  Returns the denoised video, and
    with return_levels also the noisy and denoised levels (lists of
    [F][C][rows_i][cols_i])."""
    L = _bind_pyr(lib())
    v = _video(np.ascontiguousarray(noisy, np.float32), "noisy")
    F, Ch, H, W = v.shape
    out = np.zeros_like(v)
    sz = pyr_levels(H, W, levels)
    nl = sum(F * Ch * r * cc for r, cc in sz)
    nlev = np.zeros(nl, np.float32) if return_levels else None
    dlev = np.zeros(nl, np.float32) if return_levels else None
    c = ctx or default_context()
    _check(L.vbm3d_multiscale_denoise(c.h, float(sigma), _f(v), W, H, F, Ch,
                                      None if prms_1 is None else C.byref(prms_1),
                                      None if prms_2 is None else C.byref(prms_2), levels, type_, factor,
                                      int(bool(color_space)), ref_grid_nb, _f(out), _f(nlev), _f(dlev)))
    if not return_levels:
        return out

    def split(flat):
        res, o = [], 0
        for r, cc in sz:
            res.append(flat[o:o + F * Ch * r * cc].reshape(F, Ch, r, cc).copy())
            o += F * Ch * r * cc
        return res
    return out, split(nlev), split(dlev)
