"""vbm3d_run_device_noalloc through the C ABI: the entry that allocates
nothing (its rings live in the library's fixed scratch array) against
vbm3d_run_device on the same inputs, bit for bit, on two calls in a row,
the second with a smaller clip; and the clear refusal of a clip whose
rings do not fit."""
import ctypes as C

import numpy as np
import pytest

from vbm3d_amd import synth

pytestmark = pytest.mark.gpu


def _noalloc(V, torch, sigma, z, b, f, p1, p2, nb=1):
    L = V.lib()
    pp = C.POINTER(V.Parameters)
    L.vbm3d_run_device_noalloc.argtypes = [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                           C.c_uint, C.c_uint, pp, pp, C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
    L.vbm3d_run_device_noalloc.restype = C.c_int
    sp = C.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    F, Cc, H, W = z.shape
    return L.vbm3d_run_device_noalloc(float(sigma), C.c_void_p(z.data_ptr()), None, None, W, H, F, Cc,
                                      C.byref(p1), C.byref(p2), 1, nb, C.c_void_p(b.data_ptr()),
                                      C.c_void_p(f.data_ptr()), sp)


def _pair(V, torch, noisy, sigma, nb):
    dev = torch.device("cuda:0")
    p1, p2 = V.initialize_parameters_1(sigma), V.initialize_parameters_2(sigma)
    outs = []
    for fn in ("noalloc", "device"):
        z = torch.from_numpy(noisy).to(dev)
        b, f = torch.zeros_like(z), torch.zeros_like(z)
        if fn == "noalloc":
            rc = _noalloc(V, torch, sigma, z, b, f, p1, p2, nb)
            assert rc == 0, V.lib().vbm3d_last_error().decode()
        else:
            V.run_device(sigma, z, b, f, p1, p2, True, nb)
        torch.cuda.synchronize()
        outs.append((z.cpu().numpy(), b.cpu().numpy(), f.cpu().numpy()))
    for name, a, r in zip(("noisy", "basic", "final"), outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), f"{name} differs (max |d| {np.abs(a - r).max()})"


def test_noalloc_matches_run_device_twice():
    """A 176x144 RGB clip of 12 frames (chunked rings: more frames than fit
    at once), then a smaller 96x72 grey clip of 5 frames on the tiled grid:
    the scratch keeps the first call's contents, which the second must not
    depend on."""
    import torch
    import vbm3d_amd as V
    _, big = synth.noisy_clip(176, 144, 12, 3, 20.0, 41)
    _pair(V, torch, big, 20.0, 1)
    _, small = synth.noisy_clip(96, 72, 5, 1, 40.0, 42)
    _pair(V, torch, small, 40.0, 2)


def test_noalloc_refuses_a_clip_that_does_not_fit():
    import torch
    import vbm3d_amd as V
    dev = torch.device("cuda:0")
    _, nz = synth.noisy_clip(640, 360, 10, 3, 20.0, 43)
    z = torch.from_numpy(nz).to(dev)
    b, f = torch.zeros_like(z), torch.zeros_like(z)
    p1, p2 = V.initialize_parameters_1(20.0), V.initialize_parameters_2(20.0)
    rc = _noalloc(V, torch, 20.0, z, b, f, p1, p2)
    assert rc == -3  # VBM3D_ERR_UNSUPPORTED
    assert b"scratch" in V.lib().vbm3d_last_error()
    assert torch.count_nonzero(b).item() == 0  # nothing was enqueued
