/*
 * vbm3d.h -- drop-in C++ entry point of the denoising path.
 *
 * run_vbm3d keeps the reference's signature and semantics
 * (src/vbm3d.h:25-35, src/vbm3d.cpp:87-376): vid_noisy is colour-
 * transformed in place and back, vid_basic / vid_denoised are resized to
 * vid_noisy's size and filled, vid_basic is read when prms_1.k == 0, and
 * the return value is EXIT_SUCCESS.  The work runs on the GPU through
 * libvbm3d_hip.so; unlike the reference, a failure (unsupported
 * parameters, no GPU) returns EXIT_FAILURE with a message on stderr
 * instead of silently skipping a step.  No FFTW type appears here.
 */
#ifndef VBM3D_AMD_VBM3D_H
#define VBM3D_AMD_VBM3D_H

#include "parameters.h"
#include "video.h"

int run_vbm3d(
    const float sigma
,   Video<float> &vid_noisy
,   Video<float> &fflow
,   Video<float> &bflow
,   Video<float> &vid_basic
,   Video<float> &vid_denoised
,   const Parameters& prms_1
,   const Parameters& prms_2
,   const bool color_space
);

/* Reference-patch grid emulated by run_vbm3d: 1 (default) is the grid of
 * the reference built without OpenMP; 2..32 (power of two) reproduces the
 * grid and merge order of an OpenMP build running that many threads
 * (src/vbm3d.cpp:117-128).  The environment variable VBM3D_REF_GRID_NB
 * sets the initial value. */
void vbm3d_set_ref_grid_nb(int nb);
int vbm3d_get_ref_grid_nb();

/* Device used by run_vbm3d (default 0). */
void vbm3d_set_device(int device);

#endif
