"""Writes tvl1_flow.npz: input pairs and the flows the reference's own
TV-L1 library computes for them (tvl1flow/tvl1flow_lib.c, built from its
sources by ``make -C oracle ref`` into oracle/_ref/libtvl1_ref.so and run
sequentially).  The fixture lets the oracle and the GPU path be checked
against the reference where that library is not built.

    python tests/golden/make_tvl1_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import tvl1  # noqa: E402
from vbm3d_amd import synth  # noqa: E402

# (name, width, height, seed, sigma, parameters)
CASES = [
    ("default", 48, 40, 11, 0.0, {}),
    ("noisy", 56, 36, 12, 20.0, {}),
    ("fscale1_z07", 64, 48, 13, 10.0, {"fscale": 1, "zfactor": 0.7, "warps": 3}),
    ("eps0_2warps", 40, 32, 14, 5.0, {"epsilon": 0.0, "warps": 2, "nscales": 2}),
]


def main():
    if not tvl1.have_ref():
        raise SystemExit("build the reference library first: make -C oracle ref")
    out = {}
    for name, w, h, seed, sigma, kw in CASES:
        clean, noisy = synth.noisy_clip(w, h, 2, 1, max(sigma, 1e-6), seed)
        src = noisy if sigma > 0 else clean
        I0, I1 = src[0, 0].copy(), src[1, 0].copy()
        args = dict(kw)
        eps = args.pop("epsilon", 0.01)
        u, v, _ = tvl1.flow(I0, I1, epsilon=eps, ref=True, **args)
        out[f"{name}/I0"] = I0
        out[f"{name}/I1"] = I1
        out[f"{name}/u"] = u
        out[f"{name}/v"] = v
        out[f"{name}/params"] = np.array([kw.get("fscale", 0), kw.get("zfactor", 0.5), kw.get("warps", 5), eps,
                                          kw.get("nscales", 100)], np.float64)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tvl1_flow.npz"), **out)


if __name__ == "__main__":
    main()
