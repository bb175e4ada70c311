/*
 * vbm3d_oracle.h -- C API of the CPU oracle for the VBM3D collaborative
 * filter.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (vbm3d_amd, libvbm3d_hip.so) never links, loads or calls
 * it.
 *
 * The oracle restates tehret/vbm3d's run_vbm3d path (src/vbm3d.cpp:87-376)
 * in FP32 with the reference's operation order (sequential float sums,
 * no FMA contraction, the same float/double mixing).  Each function in
 * vbm3d_oracle.cpp cites the reference file:line it follows.
 */
#ifndef VBM3D_ORACLE_H
#define VBM3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* POD mirror of Parameters (src/parameters.h:21-58). */
typedef struct orc_params {
    unsigned k, kt, p, N;
    float tau, d;
    unsigned Nf, Ns, Npr, Nb;
    unsigned T_2D, T_3D;
    float lambda3D;
    int mc;
    int n;
} orc_params;

/* Default rules of initializeParameters_1/_2 (src/main.cpp:35-207).
 * Negative arguments mean "use the default" (NONE = -1). */
void orc_init_params_1(orc_params* p, int k, int kt, int Nf, int Ns, int Npr,
                       int Nb, int pstep, int N, int d, float tau,
                       float lambda3D, unsigned T_2D, unsigned T_3D, float sigma);
void orc_init_params_2(orc_params* p, int k, int kt, int Nf, int Ns, int Npr,
                       int Nb, int pstep, int N, int d, float tau,
                       unsigned T_2D, unsigned T_3D, float sigma);

/* Reference-patch grid of one frame, in processing order (crop, row,
 * col).  Returns the count; when xs/ys/crop are non-NULL they receive the
 * absolute positions and the crop id of each entry.  nb = 1 is the
 * single-thread grid; nb = 2..32 (power of two) emulates the OpenMP
 * tiling.  Returns -1 on invalid sizes. */
long orc_grid(unsigned W, unsigned H, unsigned k, unsigned p, int n, int nb,
              unsigned* xs, unsigned* ys, unsigned* crop);

/* Predictive block matching (computeSimilarPatches) for nrefs reference
 * patches given by index t*W*H*C + y*W + x.  Outputs per ref: N indices,
 * N distances (entries past nsim are left as 0xFFFFFFFF / NaN), nsim
 * (number of sorted candidates) and nsx (group size).  fflow/bflow may be
 * NULL (zero flow); otherwise [Ff][2][H][W] with Ff >= F-1. */
int orc_block_match(const float* vid, unsigned W, unsigned H, unsigned F,
                    unsigned C, const float* fflow, const float* bflow,
                    const orc_params* prm, long nrefs, const uint32_t* ref_idx,
                    uint32_t* idx_out, float* dist_out, uint32_t* nsim_out,
                    uint32_t* nsx_out);

/* Whole path (run_vbm3d).  noisy is read and, like the reference, mutated
 * by the forward and inverse colour transforms.  basic must hold a basic
 * estimate when p1->k == 0.  nthreads <= 0 means "one thread per crop".
 * Returns 0, or -1 with a message in orc_last_error(). */
int orc_run_vbm3d(float sigma, float* noisy, const float* fflow,
                  const float* bflow, unsigned W, unsigned H, unsigned F,
                  unsigned C, const orc_params* p1, const orc_params* p2,
                  int color_space, int nb, int nthreads, float* basic,
                  float* denoised);

/* One step on an already colour-transformed video; num/den are the merged
 * aggregation buffers (the reference's per-thread buffers summed in thread
 * order when nb > 1).  step = 1 uses `noisy` for matching and filtering;
 * step = 2 matches on `basic` and filters noisy + basic. */
int orc_step(int step, float sigma, const float* noisy, const float* basic,
             const float* fflow, const float* bflow, unsigned W, unsigned H,
             unsigned F, unsigned C, const orc_params* prm, int color_space,
             int nb, int nthreads, float* num, float* den);

/* Counters of the last orc_step (or orc_run_vbm3d) for step 1 or 2:
 * out[0] reference patches, [1] candidates evaluated (patchDistance
 * calls), [2] group members, [3] references with a block-matching
 * ranking decision within 1e-5 (relative) of a tie, [4] references with a
 * group candidate within 1e-5 of tau, [5] hard-threshold coefficients
 * within 1e-5 of the threshold, [6] coefficients tested against it. */
void orc_last_stats(int step, long long* out);

/* FP64 re-evaluation of one step with the FP32 run's decisions frozen
 * (same groups, same hard-threshold keep/zero decisions).  Used to report
 * the FP32 path's per-step error.  Produces num/den in double. */
int orc_step_fp64(int step, float sigma, const float* noisy, const float* basic,
                  const float* fflow, const float* bflow, unsigned W,
                  unsigned H, unsigned F, unsigned C, const orc_params* prm,
                  int color_space, double* num, double* den);

/* Adjoint of step 2 with the basic estimate fixed (so the groups, Wiener
 * gains and weights are frozen), in FP64: grad_noisy = J^T grad_out,
 * where J is the Jacobian of the step-2 output num/den (with the den == 0
 * fallback) with respect to the colour-transformed noisy video.
 * final_out (may be NULL) receives the FP64 forward value.  Runs
 * single-threaded: meant for small clips. */
int orc_step2_adjoint_fp64(float sigma, const float* noisy, const float* basic,
                           const float* fflow, const float* bflow, unsigned W,
                           unsigned H, unsigned F, unsigned C, const orc_params* prm,
                           int color_space, int nb, const double* grad_out,
                           double* grad_noisy, double* final_out);

/* Building blocks, exposed for the oracle's own pinning tests. */
void orc_bior_fwd(const float* patch64, float* out64);
void orc_bior_inv(float* sig64);
void orc_dct_fwd(unsigned k, const float* patch, float* out);
void orc_dct_inv(unsigned k, float* sig);
/* temporal_transform / temporal_inv_transform (src/vbm3d.cpp:1866-1956)
 * of one coefficient's kt (<= 8) values, in place. */
void orc_temporal_fwd(unsigned kt, float* v);
void orc_temporal_inv(unsigned kt, float* v);
void orc_haar_fwd(float* v, unsigned n);
void orc_haar_inv(float* v, unsigned n);
void orc_hadamard(float* v, unsigned n);
void orc_kaiser(unsigned k, float* out);
void orc_dct_norm(unsigned k, float* norm, float* norm_inv);
int  orc_sigma_table(float sigma, unsigned C, unsigned color_space, float* out);
void orc_color(float* vid, unsigned W, unsigned H, unsigned F, unsigned C,
               int forward);
/* std::partial_sort on (dist, idx) pairs with the reference comparator
 * (src/Utilities/Utilities.cpp:40-45); in place. */
void orc_partial_sort(float* dist, uint32_t* idx, unsigned len, unsigned mid);
/* VideoUtils::computePSNR (src/Utilities/LibVideoT.hpp:577-597). */
void orc_psnr(const float* a, const float* b, size_t n, double* psnr,
              double* rmse);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
