"""vbm3d_amd -- MI355X-native VBM3D collaborative filter.

The denoising path of tehret/vbm3d (run_vbm3d, src/vbm3d.cpp:87-376) and
its TV-L1 flow companion (tvl1flow/tvl1flow_lib.c) on hand-written gfx950
HIP kernels behind a C ABI (include/vbm3d_hip.h, libvbm3d_hip.so).  The
Python API mirrors the reference's interface.
"""
from .api import (BIOR, DCT, HAAR, HADAMARD, NONE, BatchConfig, BatchItem, Context, Parameters, VBM3DError,
                  autotune_batch, block_match, denoise, grid, initialize_parameters_1, initialize_parameters_2,
                  TVL1Params, lib, run_batch, run_device, run_vbm3d, step, step2_adjoint, sure, tvl1_flow,
                  tvl1_flow_device, tvl1_nscales, tvl1_params, PYR_DCT, PYR_GAUSSIAN, PYR_LANCZOS, pyr_levels,
                  pyr_decompose, pyr_recompose, pyr_merge_coarse, multiscale_denoise)

__all__ = ["BIOR", "DCT", "HAAR", "HADAMARD", "NONE", "BatchConfig", "BatchItem", "Context", "Parameters",
           "VBM3DError", "autotune_batch", "run_batch",
           "block_match", "denoise", "grid", "initialize_parameters_1", "initialize_parameters_2",
           "lib", "run_device", "run_vbm3d", "step", "step2_adjoint", "sure",
           "TVL1Params", "tvl1_flow", "tvl1_flow_device", "tvl1_nscales", "tvl1_params",
           "PYR_DCT", "PYR_GAUSSIAN", "PYR_LANCZOS", "pyr_levels", "pyr_decompose", "pyr_recompose",
           "pyr_merge_coarse", "multiscale_denoise"]
