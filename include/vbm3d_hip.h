/*
 * vbm3d_hip.h -- C ABI of libvbm3d_hip.so, the MI355X (gfx950) VBM3D
 * collaborative filter.
 *
 * Plain C types only: pointers, sizes, a POD parameter block.  No
 * exception crosses this boundary; every entry point returns VBM3D_OK (0)
 * or a negative error code, with a message in vbm3d_last_error() (thread
 * local).  Videos are float32, planar [F][C][H][W] (the reference's
 * VideoSize layout, src/Utilities/LibVideoT.hpp:88-92); flows are
 * [F-1 or more][2][H][W] or NULL for zero flow (src/main.cpp:353-369).
 *
 * Citations are relative to the reference repository (tehret/vbm3d).
 */
#ifndef VBM3D_HIP_H
#define VBM3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBM3D_OK 0
#define VBM3D_ERR_ARG (-1)         /* invalid argument or size */
#define VBM3D_ERR_HIP (-2)         /* HIP runtime error */
#define VBM3D_ERR_UNSUPPORTED (-3) /* parameters outside what the kernels implement */
#define VBM3D_ERR_NOMEM (-4)       /* device allocation failed */

/* Transform codes, as accepted by the reference CLI (src/main.cpp:27-30). */
#define VBM3D_DCT 0
#define VBM3D_BIOR 1
#define VBM3D_HADAMARD 2
#define VBM3D_HAAR 3

/* POD mirror of struct Parameters (src/parameters.h:21-58), field for
 * field and in the same order. */
typedef struct vbm3d_params {
    unsigned k;      /* patch size */
    unsigned kt;     /* temporal patch size (1 only) */
    unsigned p;      /* step between reference patches */
    unsigned N;      /* maximum group size */
    float tau;       /* distance threshold */
    float d;         /* bias toward the reference position */
    unsigned Nf;     /* frames searched before and after */
    unsigned Ns;     /* search window in the reference frame */
    unsigned Npr;    /* predictive search window */
    unsigned Nb;     /* matches kept per frame */
    unsigned T_2D;   /* VBM3D_DCT or VBM3D_BIOR */
    unsigned T_3D;   /* VBM3D_HADAMARD or VBM3D_HAAR */
    float lambda3D;  /* hard threshold (step 1) */
    int mc;          /* motion-compensated 3D patches (no effect at kt 1) */
    int n;           /* tile border of the multi-threaded grid (16) */
} vbm3d_params;

typedef struct vbm3d_ctx vbm3d_ctx;

/* Context owning device memory and a HIP stream on `device`.  Replaces the
 * per-call FFTW plan setup and per-thread buffers of run_vbm3d
 * (src/vbm3d.cpp:135-160, 216-217).  NULL on failure. */
vbm3d_ctx* vbm3d_create(int device);
void vbm3d_destroy(vbm3d_ctx* ctx);
const char* vbm3d_last_error(void);

/* Default rules of initializeParameters_1 (src/main.cpp:35-125);
 * negative arguments mean "not given" (the CLI's NONE = -1).  T_2D/T_3D
 * take (unsigned)-1 for "not given".  lambda3D, tau: negative = default. */
void vbm3d_default_params_1(vbm3d_params* p, int k, int kt, int Nf, int Ns,
                            int Npr, int Nb, int pstep, int N, int d,
                            float tau, float lambda3D, unsigned T_2D,
                            unsigned T_3D, float sigma);
/* Default rules of initializeParameters_2 (src/main.cpp:127-207). */
void vbm3d_default_params_2(vbm3d_params* p, int k, int kt, int Nf, int Ns,
                            int Npr, int Nb, int pstep, int N, int d,
                            float tau, unsigned T_2D, unsigned T_3D,
                            float sigma);

/* Whole path on host buffers; replaces run_vbm3d (src/vbm3d.cpp:87-376,
 * declared src/vbm3d.h:25-35).  Semantics kept from the reference:
 *   - `noisy` is mutated by the forward then inverse colour transform;
 *   - `basic` is an input when p1->k == 0 (step 1 skipped), else output;
 *   - p2->k == 0 copies basic to denoised;
 *   - ref_grid_nb = 1 gives the single-thread reference grid, 2..32
 *     (power of two) the grid and merge order of the reference run with
 *     that many OpenMP threads (src/vbm3d.cpp:117-128, 213-354).
 * The noisy video is uploaded in chunks on a second stream of the context
 * while step 1 runs on the frames already there, and the noisy and basic
 * videos are copied back on that stream while the later steps compute;
 * the call returns when all three videos are in host memory. */
int vbm3d_run(vbm3d_ctx* ctx, float sigma, float* noisy, const float* fflow,
              const float* bflow, unsigned W, unsigned H, unsigned F,
              unsigned C, const vbm3d_params* p1, const vbm3d_params* p2,
              int color_space, int ref_grid_nb, float* basic,
              float* denoised);

/* Same on device-resident buffers, enqueued on `stream` (a hipStream_t,
 * NULL = the context's stream) and synchronised before return.  d_noisy
 * is mutated like `noisy` above. */
int vbm3d_run_device(vbm3d_ctx* ctx, float sigma, float* d_noisy,
                     const float* d_fflow, const float* d_bflow, unsigned W,
                     unsigned H, unsigned F, unsigned C,
                     const vbm3d_params* p1, const vbm3d_params* p2,
                     int color_space, int ref_grid_nb, float* d_basic,
                     float* d_denoised, void* stream);

/* vbm3d_run_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as vbm3d_run_device, without a context: runs on the current
 * device, enqueued on `stream` (a hipStream_t; NULL = the null stream) and
 * synchronised before return.  It calls no hipMalloc, hipFree or other
 * allocating routine: the per-patch rings, transform tables and the flow
 * maximum live in a 256 MiB device array of the library, which each GPU
 * the library runs on holds from the first call on it until the process
 * ends.  Clips whose rings at the smallest chunking (2 Nf + 1 frames of
 * reference patches, N k^2 C + N + C + 1 words each) do not fit are
 * refused with VBM3D_ERR_UNSUPPORTED before anything is enqueued; at the
 * default parameters that admits e.g. 320x240 grey or 176x144 RGB clips
 * of any length.  Calls that overlap on one GPU share that GPU's array and
 * race: serialise them. */
int vbm3d_run_device_noalloc(float sigma, float* d_noisy, const float* d_fflow,
                             const float* d_bflow, unsigned W, unsigned H,
                             unsigned F, unsigned C, const vbm3d_params* p1,
                             const vbm3d_params* p2, int color_space,
                             int ref_grid_nb, float* d_basic,
                             float* d_denoised, void* stream);

/* Predictive block matching only (computeSimilarPatches,
 * src/vbm3d.cpp:1082-1197) for nrefs reference patches given by index
 * t*W*H*C + y*W + x, on host buffers.  Per ref: N sorted indices and
 * distances (entries past nsim are 0xFFFFFFFF / NaN), nsim and the group
 * size nsx. */
int vbm3d_block_match(vbm3d_ctx* ctx, const float* vid, unsigned W,
                      unsigned H, unsigned F, unsigned C, const float* fflow,
                      const float* bflow, const vbm3d_params* prm,
                      long nrefs, const uint32_t* ref_idx, uint32_t* idx_out,
                      float* dist_out, uint32_t* nsim_out, uint32_t* nsx_out);

/* One step on an already colour-transformed video, returning the merged
 * aggregation buffers (vbm3d_1st_step src/vbm3d.cpp:447-654 when step = 1,
 * vbm3d_2nd_step :676-928 when step = 2, plus the thread-order merge
 * :267-280 / :333-346 when ref_grid_nb > 1).  Host buffers. */
int vbm3d_step(vbm3d_ctx* ctx, int step, float sigma, const float* noisy,
               const float* basic, const float* fflow, const float* bflow,
               unsigned W, unsigned H, unsigned F, unsigned C,
               const vbm3d_params* prm, int color_space, int ref_grid_nb,
               float* num, float* den);

/* Adjoint of step 2 (no counterpart in the reference): with the basic
 * estimate fixed, the groups, Wiener gains and aggregation weights are
 * frozen and step 2 is linear in the noisy video.  Computes
 * d_grad_noisy = J^T d_grad_out, J the Jacobian of the step-2 output
 * (num/den with the den == 0 fallback of vbm3d_run) with respect to the
 * colour-transformed noisy video.  The groups are recomputed from
 * d_basic (block matching runs twice) rather than stored.  d_final, if
 * not NULL, receives the forward output.  Device buffers, [F][C][H][W];
 * enqueued on `stream` (NULL = the context's stream) and synchronised. */
int vbm3d_step2_adjoint(vbm3d_ctx* ctx, float sigma, const float* d_noisy,
                        const float* d_basic, const float* d_fflow,
                        const float* d_bflow, unsigned W, unsigned H,
                        unsigned F, unsigned C, const vbm3d_params* p2,
                        int color_space, int ref_grid_nb,
                        const float* d_grad_out, float* d_grad_noisy,
                        float* d_final, void* stream);

/* One clip of a batch (no counterpart in the reference, which denoises
 * one clip per call).  Device buffers, planar [F][C][H][W]; the same
 * semantics per clip as vbm3d_run_device: `noisy` is mutated by the
 * colour round trip, `basic` is an input when p1.k == 0. */
typedef struct vbm3d_clip {
    float* noisy;
    float* basic;
    float* denoised;
    const float* fflow; /* or NULL */
    const float* bflow; /* or NULL */
    unsigned W, H, F, C;
    float sigma;
    vbm3d_params p1, p2;
    int color_space;
} vbm3d_clip;

/* Launch shape of the batched form.  Any choice gives bit-identical
 * results; only the speed differs. */
typedef struct vbm3d_batch_config {
    unsigned chunk_frames;  /* reference frames per chunk; 0 = as many as the memory budget allows */
    int waves_per_group;    /* waves per workgroup of the matching and filtering launches, 1..4; 0 = most that fit */
    int grab;               /* reference patches a wave takes from the work queue at a time, >= 1 */
    int groups_per_cu;      /* persistent workgroups per compute unit, >= 1 */
    int use_graph;          /* capture the whole batch in a HIP graph and replay it on later identical calls */
    int persistent;         /* 1: matching, filtering and aggregation run as persistent workgroups that take
                               patches (output tiles) from the global work queue; 0: one wave per patch and
                               one workgroup per output tile, in workgroup dispatch order */
} vbm3d_batch_config;

/* Denoises `nclips` clips of any sizes and parameters together: per step,
 * per chunk of frames and per kernel class, one launch over the
 * reference patches (or output tiles) of all clips, ordered as a global
 * work queue; with `persistent`, resident workgroups take them from the
 * queue through an atomic counter.  With use_graph the whole sequence is captured once
 * in a HIP graph and replayed on later calls with the same clips,
 * buffers and configuration (not when a clip has flows).  cfg NULL means
 * the configuration vbm3d_autotune_batch last chose for this context, or
 * the defaults.  Each clip's output is bit-identical to vbm3d_run_device
 * on that clip.  Enqueued after the work on `stream` (NULL = the
 * context's stream); synchronised before return. */
int vbm3d_run_batch(vbm3d_ctx* ctx, const vbm3d_clip* clips, int nclips,
                    int ref_grid_nb, const vbm3d_batch_config* cfg,
                    void* stream);

/* Times candidate configurations of the batched form on these clips
 * (coordinate search over persistent, waves_per_group, grab and
 * groups_per_cu) and
 * returns the fastest in *best; the context keeps it for later
 * vbm3d_run_batch calls with cfg NULL.  The clips' buffers are restored
 * after each trial, so inputs are left as they were. */
int vbm3d_autotune_batch(vbm3d_ctx* ctx, const vbm3d_clip* clips,
                         int nclips, int ref_grid_nb,
                         vbm3d_batch_config* best, void* stream);

/* Reference-patch grid of one frame in processing order (crop, row, col);
 * ind_initialize + subDivideTight (src/Utilities/Utilities.cpp:400-415,
 * src/Utilities/LibVideoT.hpp:903-961).  Returns the count, or -1. */
long vbm3d_grid(unsigned W, unsigned H, unsigned k, unsigned p, int n,
                int nb, unsigned* xs, unsigned* ys, unsigned* crop);

/* Frames of reference patches per chunk for the per-clip entries
 * (vbm3d_run, vbm3d_run_device, vbm3d_step): 0, the default, takes as
 * many as the memory budget allows; 1 launches every kernel once per
 * frame, the plain per-frame path.  Results do not depend on it.  No
 * counterpart in the reference, whose loops run per patch on the CPU. */
int vbm3d_set_chunk_frames(vbm3d_ctx* ctx, unsigned frames);

/* Kernel timing and work counters (a diagnostic with no counterpart in
 * the reference, for roofline measurements).  vbm3d_profile(ctx, 1)
 * resets the counters and brackets every later launch with HIP events on
 * the stream it runs on; vbm3d_profile(ctx, 2) does the same and also
 * counts the work done (an extra store per patch and a reduction launch
 * after each block-matching launch, so not for timed runs);
 * vbm3d_profile(ctx, 0) stops.  vbm3d_profile_read
 * returns, for kernel 0 (block matching), 1 (group filtering),
 * 2 (aggregation) or 3 (colour transform): total milliseconds, number of
 * launches and the work done -- candidate patches evaluated (0), group
 * members filtered (1), output samples produced (2).  kernel 4 reads the
 * block-matching pixel comparisons (candidates * k*k*C; no time).  Adding
 * 8 (step 1) or 16 (step 2) to the code restricts the figures to one
 * step; the plain code sums both. */
int vbm3d_profile(vbm3d_ctx* ctx, int enable);
int vbm3d_profile_read(vbm3d_ctx* ctx, int kernel, double* ms, long* launches,
                       double* work);

/* ---- TV-L1 optical flow (the flow companion the reference ships for
 * -fflow / -bflow: tvl1flow/tvl1flow_lib.c, driven by tvl1flow/main.c
 * and src/utils/VBM3Ddenoising_OF.sh).  Images are grey float [ny][nx];
 * the flow is returned as u (x component) and v (y component), each
 * [ny][nx], the planes the reference writes to its flow file. */
typedef struct vbm3d_tvl1_params {
    float tau;      /* time step of the dual update */
    float lambda;   /* weight of the data term */
    float theta;    /* coupling of the primal and auxiliary flows */
    int nscales;    /* pyramid levels (capped by the caller: vbm3d_tvl1_nscales) */
    int fscale;     /* finest level solved; finer levels are only upsampled */
    float zfactor;  /* downsampling factor between levels, in (0, 1) */
    int nwarps;     /* warps per level */
    float epsilon;  /* stopping threshold on the RMS flow change */
} vbm3d_tvl1_params;

/* The command line's defaults (tvl1flow/main.c:20-30). */
void vbm3d_tvl1_default_params(vbm3d_tvl1_params* p);

/* The level count the reference's command line uses: `nscales`, capped so
 * that the coarsest level stays about 16 pixels across (main.c:139-146). */
int vbm3d_tvl1_nscales(int nx, int ny, float zfactor, int nscales);

/* Dual_TVL1_optic_flow_multiscale (tvl1flow_lib.c:384-485) on the GPU:
 * normalisation, pre-smoothing, pyramid, and per level `nwarps` warps of
 * at most 300 primal-dual iterations.  iters (optional) receives the
 * iterations of each solved level (coarsest first) and warp:
 * (nscales - fscale) * nwarps ints.  The _device form takes device
 * buffers and runs on the context's stream; the plain form copies host
 * buffers.  Both return VBM3D_OK or an error code. */
int vbm3d_tvl1_flow(vbm3d_ctx* ctx, const float* I0, const float* I1, float* u,
                    float* v, int nx, int ny, const vbm3d_tvl1_params* p,
                    int* iters);
int vbm3d_tvl1_flow_device(vbm3d_ctx* ctx, const float* d_I0, const float* d_I1,
                           float* d_u, float* d_v, int nx, int ny,
                           const vbm3d_tvl1_params* p, int* iters);

/* ---- Multiscale pyramid (multiscaler/: decompose.cpp, recompose.cpp,
 * merge_coarse.cpp, the functions of multiscaler.cpp) and the multiscale
 * denoising it serves (src/utils/VBM3Ddenoising_multiscale.sh).  Images
 * are planar float [c][rows][cols] host buffers, c = 1 or 3; a pyramid of
 * `levels` levels is stored level after level, level i of
 * rows_i x cols_i, each half the previous (integer division).  type: 0 the
 * DCT pyramid (full-frame DCT, low frequencies kept), 1 Gaussian,
 * 2 Lanczos; factor is the recomposition ratio (-c). */

/* Level sizes (decompose.cpp:44-53). */
int vbm3d_pyr_levels(int rows, int cols, int levels, int* lrows, int* lcols);

/* decompose: the levels of one image. */
int vbm3d_pyr_decompose(vbm3d_ctx* ctx, const float* img, int rows, int cols,
                        int ch, int levels, int type, float* out);

/* recompose: one image from its levels; type 0 copies the rows_i * factor
 * x cols_i * factor lowest DCT frequencies of every coarser level into the
 * finest level's, types 1 and 2 add back blurred, upsampled differences. */
int vbm3d_pyr_recompose(vbm3d_ctx* ctx, const float* levels, int rows,
                        int cols, int ch, int nlevels, int type, float factor,
                        float* out);

/* merge_coarse: the coarse image's lowest DCT frequencies replace the
 * fine image's. */
int vbm3d_pyr_merge_coarse(vbm3d_ctx* ctx, const float* fine, int rows,
                           int cols, const float* coarse, int crows, int ccols,
                           int ch, float factor, float* out);

/* The multiscale denoising of a video [F][C][H][W]: every frame into its
 * levels, every level denoised as a video by run_vbm3d with sigma / 2^level
 * truncated to two decimals (the multiscale shell script's `bc`) and default
 * parameters for that sigma (or p1 / p2 when not NULL), every frame
 * recomposed.  noisy_levels / denoised_levels (optional) receive the
 * levels, level after level, each [F][C][rows_i][cols_i]. */
int vbm3d_multiscale_denoise(vbm3d_ctx* ctx, float sigma, const float* noisy,
                             unsigned W, unsigned H, unsigned F, unsigned C,
                             const vbm3d_params* p1, const vbm3d_params* p2,
                             int levels, int type, float factor,
                             int color_space, int ref_grid_nb, float* denoised,
                             float* noisy_levels, float* denoised_levels);

/* Micro-benchmark of the group filter's k x k DCT (k = 7 or 8), with no
 * counterpart in the reference: `reps` forward + inverse round trips over
 * `npatch` random patches kept on chip, on the vector ALUs with the
 * denoiser's own tables and operation order (mode 0) or on the matrix
 * cores, FP32 MFMA with the orthonormal DCT matrix (mode 1), on the
 * current device.  Returns the milliseconds of one launch (median of 3
 * after a warm-up) and the largest |round trip - input| after one round
 * trip.  It decides whether MFMA pays for these transforms. */
int vbm3d_bench_transforms(int K, int mode, long npatch, int reps, double* ms,
                           double* max_err);

/* Host-side check of the group-selection routine the kernels use: sorts
 * the first `mid` of `len` (dist, idx) pairs like std::partial_sort with
 * the reference comparator (src/Utilities/Utilities.cpp:40-45). */
void vbm3d_selftest_partial_sort(float* dist, uint32_t* idx, unsigned len,
                                 unsigned mid);

#ifdef __cplusplus
}
#endif
#endif
