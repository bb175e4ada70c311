// kernels.h -- device-side argument blocks and launch helpers shared by
// the HIP kernels (kernels_bm.hip, kernels_filter.hip, kernels_agg.hip)
// and the host orchestration (capi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vb {

constexpr int kMaxN = 16;       // largest group size the kernels hold in registers
constexpr int kMaxWin = 8;      // largest Ns / Npr (window <= 64 candidates)
constexpr int kMaxBest = 64;    // Nb * (2 Nf + 1) bound
constexpr int kMaxSlots = 512;  // Nb * Npr^2 bound
constexpr int kAggTileX = 32;   // aggregation tile width (pixels)
constexpr int kAggTileY = 16;   // aggregation tile height
constexpr int kAggThreads = kAggTileX * kAggTileY;  // one thread per output pixel

// Exact unsigned 32-bit division by a run-time constant d >= 1 with a
// multiply-high and two shifts (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, round-up variant):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0), t = mulhi(m, n).
struct Div32 {
    uint32_t m, s1, s2;
};

__host__ __device__ inline Div32 make_div32(uint32_t d) {
    uint32_t l = 0;
    while (((uint64_t)1 << l) < d) l++;
    const uint64_t m = (((((uint64_t)1) << l) - d) << 32) / d + 1;
    return Div32{(uint32_t)m, l ? 1u : 0u, l ? l - 1 : 0u};
}

__host__ __device__ inline uint32_t div32(uint32_t n, Div32 v) {
    const uint32_t t = (uint32_t)(((uint64_t)v.m * n) >> 32);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}

__host__ __device__ inline int pow2ceil(int v) {
    int r = 1;
    while (r < v) r *= 2;
    return r;
}

// Transform tables of one step, uploaded once per step.
struct Tabs {
    float biorF[64];   // analysis rows [j][s]
    float biorI[64];   // synthesis [s][j]
    float dctF[64];    // REDFT10 matrix [v][q]
    float dctI[64];    // REDFT01 matrix [q][v]
    float nrm[64];     // coef_norm
    float inv[64];     // coef_norm_inv
    float kaiser[64];  // aggregation window
    float dct_out;     // 1/(2k)
};

// Per-wavefront LDS layout of k_bm, in 4-byte words.  Tiles use an odd
// pitch so that lanes reading neighbouring candidates hit different banks.
struct BMLay {
    int th0, tp0, th1, tp1;
    int ws;      // words from one predictive window's area to the next
    bool piped;  // the paired form: both chains of a frame in shared passes
    int ref, tile, sd, si, sv, cd, ci, bd, bi, fd, fi, od, oi, scr, pp, win, total;
};

__host__ __device__ inline BMLay bm_layout(int K, int C, int Ns, int Npr, int Nb, int Nf) {
    BMLay L;
    int o = 0;
    L.th0 = Ns + K - 1;
    L.tp0 = L.th0 | 1;
    L.th1 = Npr + K - 1;
    L.tp1 = L.th1 | 1;
    // predictive-stage slots: window w owns [w * P2p, (w + 1) * P2p),
    // P2p = Npr^2 rounded up to a power of two.  The paired form (at most
    // 2 matches per frame, one chain's slots within 64, at most 12 tile
    // rows) gives chain h the slots [64 h, 64 h + 64), reads the reference
    // patch through scalar loads and needs no validity array.
    const int P2p = pow2ceil(Npr * Npr);
    L.piped = Nb <= 2 && Nb * P2p <= 64 && L.th1 <= 12;
    // The paired form's lanes hold the valid slots of up to four windows in
    // list order, each reading its window's area at (row * tp1 + column).
    // Pitch 13 and window areas 5 words apart modulo 32 spread those
    // addresses over the LDS banks: in a model of typical window overlaps,
    // about 36 % fewer bank cycles per read than pitch th1 | 1 with
    // back-to-back areas.  (The smallest such padding keeps four 4-wave
    // workgroups of the 8x8 RGB form within a CU's LDS.)
    if (L.piped) L.tp1 = 13;
    L.ws = L.th1 * L.tp1 * C;
    if (L.piped) L.ws += ((5 - L.ws) % 32 + 32) % 32;
    const int S = L.piped ? 128 : ((2 * Nb * P2p + 127) / 128) * 128;
    const int t0 = L.th0 * L.tp0 * C, t1 = 2 * Nb * L.ws;
    int t = t0 > t1 ? t0 : t1;
    t = t > 128 ? t : 128;  // the slow-path scratch (64 entries) reuses the tile
    const int nbest = (Nb * (2 * Nf + 1) + 3) & ~3;  // picks of all frames
    L.ref = o; o += L.piped ? 0 : K * K * C;
    L.tile = o; o += t;
    L.scr = L.tile;
    L.sd = o; o += S;
    L.si = o; o += S;
    L.sv = o; o += L.piped ? 0 : S;
    L.cd = o; o += 64;
    L.ci = o; o += 64;  // directly after cd: the paired form's slot list spans both
    L.fd = o; o += L.piped ? 8 : 64;
    L.fi = o; o += L.piped ? 8 : 64;
    L.od = o; o += 16;
    L.oi = o; o += 16;
    L.pp = o; o += 16;
    L.win = o; o += L.piped ? 20 : 80;
    L.bd = o; o += nbest;  // last: the only arrays whose place depends on Nf
    L.bi = o; o += nbest;
    L.total = o;
    return L;
}

struct BMArgs {
    const float* vid;  // matching video [F][C][H][W]
    const float* ff;   // forward flow [F-1][2][H][W] or null
    const float* bf;   // backward flow or null
    unsigned W, H, F, C, wh, whc;
    unsigned nrf;             // reference patches per frame (grid mode)
    const uint32_t* ref_xy;   // grid mode: x | y << 16 per in-frame ref
    const uint32_t* ref_list; // list mode: global patch indices (overrides grid)
    long r_begin, r_end;      // global ref ids t*nrf + i (or list positions)
    unsigned ring;            // slot = (t % ring) * nrf + i (grid mode)
    int Ns, Npr, Nb, Nf, N, kt;
    float d, tau;
    uint32_t* idx_out;   // [slot][N]
    float* dist_out;     // [slot][N] or null
    uint32_t* nsim_out;  // [slot] or null
    uint32_t* nsx_out;   // [slot]
    Div32 dW, dWHC;       // division by W and by W*H*C
    uint32_t* count_out;  // null, or per patch of the launch, at 2*(r - r_begin):
                          // candidates evaluated, group size
    int band;             // grid mode: workgroups in band order (band_order)
};

struct FArgs {
    const float* noisy;
    const float* basic;  // step 2 only
    unsigned W, H, F, C, wh, whc;
    unsigned nrf, ring;
    long r_begin, r_end;
    const uint32_t* idx;
    const uint32_t* nsx;
    int N, T2D, T3D;
    float lambda;
    float sig[3];
    const Tabs* tabs;
    float* est;  // [slot][C][N][K*K]
    float* wgt;  // [slot][C]
    Div32 dW, dWHC;      // division by W and by W*H*C
    const float* gy;   // step-2 adjoint: gradient with respect to the output
    const float* den;  // step-2 adjoint: the forward pass's denominator
};

struct AArgs {
    unsigned W, H, F, C, wh, whc;
    int K, N, Nf, kt;
    unsigned nrf, ring, Fr;
    int ncrop;
    const int* crop_meta;       // per crop: offset in frame list, nrows, ncols, rows_at, cols_at
    const uint16_t* grid_pos;   // concatenated sorted row and column positions
    const uint32_t* idx;
    const uint32_t* nsx;
    const float* est;
    const float* wgt;
    const Tabs* tabs;
    int R;                      // largest member-to-reference displacement
    int R0, Rs;                 // a member k frames away lies within min(R, R0 + k Rs)
    // optional, per crop: the grid rows (ia, ib, -, -) of [tile row][k],
    // then the grid columns (ja, jb, m, s1 | s2 << 8) of [tile column][k],
    // k = frame distance 0..Nf, (m, s1, s2) the division by jb - ja (or 1);
    // null: the kernel searches the grid itself
    const int4* boxtab;
    int nty, ntx;               // tile rows and columns of a frame
    Div32 divW;                 // division by W
    unsigned t_begin;           // first output frame of this launch
    const float* fallback;      // used where den == 0
    float* out;                 // out = den == 0 ? fallback : num / den
    float* num_out;             // optional raw sums
    float* den_out;             // optional
    int raw;                    // 1: plain sums of est (no Kaiser window, no weight)
};

// Global work queue of a batched (persistent) launch.
struct WorkQueue {
    unsigned long long* counter;  // zeroed before the launch
    const long* item_begin;       // [njobs + 1] first item of each job
    int njobs;
    long total;                   // item_begin[njobs]
    int grab;                     // items taken per atomic
};

// ---------------------------------------------------------------------
// General path (kernels_gen.hip): the parameter sets the fast kernels do
// not implement, within these bounds.
// ---------------------------------------------------------------------
constexpr int kGenMaxK = 16;      // patch size
constexpr int kGenMaxKt = 8;      // temporal patch size
constexpr int kGenMaxK2t = 256;   // k^2 kt, coefficients per member and channel
constexpr int kGenMaxN = 32;      // group size
constexpr int kGenMaxNb = 16;     // matches per frame
constexpr int kGenMaxWin = 15;    // Ns, Npr
constexpr int kGenMaxNf = 8;      // frames searched each way
constexpr int kGenAggScan = 128;  // reference patches per aggregation scan chunk
constexpr int kGenAggTileX = 32;  // general aggregation tile (one thread per pixel)
constexpr int kGenAggTileY = 16;

// Transform tables of one step of the general path (k up to 16, kt up
// to 8), uploaded once per step.
struct TabsG {
    float dctF[256];   // REDFT10 [v][q], row stride k
    float dctI[256];   // REDFT01 [q][v]
    float nrm[256];    // coef_norm
    float inv[256];    // coef_norm_inv
    float kaiser[256];
    float biorF[64], biorI[64];
    float tF[64], tI[64];  // temporal REDFT10 [u][t], REDFT01 [t][u], row stride kt
    float dct_out, tnorm;  // 1/(2k), 1/sqrt(2 kt)
};

struct GenArgs {
    const float* vid;    // matching video (noisy in step 1, basic in step 2)
    const float* noisy;  // filtering: noisy and basic videos
    const float* basic;
    const float* ff;     // forward flow [F-1][2][H][W] or null (also the trajectories' flow)
    const float* bf;     // backward flow or null
    unsigned W, H, F, C, wh, whc;
    unsigned nrf, ring;
    const uint32_t* ref_xy;    // x | y << 16 per in-frame reference patch
    const uint32_t* ref_list;  // list mode (matching only): global patch indices, slot = r - r_begin
    long r_begin, r_end;       // global ref ids t*nrf + i (or list positions)
    int K, kt, mc, Ns, Npr, Nb, Nf, N;
    float d, tau;
    int T2D, T3D;
    float lambda;
    float sig[3];
    uint32_t* idx_out;   // [slot][N]
    float* dist_out;     // [slot][N] or null
    uint32_t* nsim_out;  // [slot] or null
    uint32_t* nsx_out;   // [slot]
    uint32_t* count_out; // null, or per patch: candidates evaluated, group size
    const TabsG* tabs;
    float* est;          // [slot][N][C][kt][k^2]
    float* wgt;          // [slot][C]
    uint32_t* mpos;      // [slot][N][kt] slice corners x | y << 16, or null (kt == 1)
};

// Per-wavefront LDS layout of k_bm_gen, in 4-byte words.
struct GenBMLay {
    int win, wd, wi, fd, fi, bd, bi, od, oi, chain, rtr, scr, total;
};

__host__ __device__ inline GenBMLay gen_bm_layout(int Nb, int Nf, int Ns, int Npr) {
    GenBMLay L;
    const int ws = (Ns > Npr ? Ns : Npr), wmax = ws * ws;     // window list
    const int nwin = 1 + 2 * Nf * Nb;                          // windows of one reference patch
    const int flist = Nb * Nb, blist = Nb * (2 * Nf + 1);     // frame and best lists
    int smax = wmax > flist ? wmax : flist;
    smax = smax > blist ? smax : blist;
    smax = smax > 64 ? smax : 64;
    int o = 0;
    L.win = o; o += 5 * nwin;
    L.wd = o; o += wmax;
    L.wi = o; o += wmax;
    L.fd = o; o += flist;
    L.fi = o; o += flist;
    L.bd = o; o += blist;
    L.bi = o; o += blist;
    L.od = o; o += kGenMaxN;
    L.oi = o; o += kGenMaxN;
    L.chain = o; o += 32;
    L.rtr = o; o += 2 * kGenMaxKt;
    o = (o + 1) & ~1;  // 8-byte alignment of the heap entries
    L.scr = o; o += 2 * smax;
    L.total = o;
    return L;
}

// LDS words of k_filter_gen (one wavefront): the group(s) [k^2 kt][N],
// three k x k scratch tiles and the members' slice corners.
__host__ __device__ inline int gen_filter_words(int K, int kt, int N, int step) {
    return K * K * kt * N * (step == 2 ? 2 : 1) + 3 * K * K + N * kt;
}

// Launchers (kernels_*.hip).  Return hipSuccess or the launch error.
// 1: the search shape the kernels also hold as compile-time constants
// (Ns 7, Npr 5, Nb 2: the default parameters of both steps); 0 otherwise.
inline int bm_shape(const BMArgs& a) { return a.Ns == 7 && a.Npr == 5 && a.Nb == 2 ? 1 : 0; }

hipError_t launch_bm(int K, const BMArgs& a, hipStream_t s);
hipError_t launch_filter(int K, int step, const FArgs& a, hipStream_t s);
hipError_t launch_aggregate(int C, const AArgs& a, unsigned frames, hipStream_t s);

// Transform micro-benchmark (xform_bench.hip): the k x k DCT on the vector
// ALUs with the denoiser's tables (mode 0) or on the matrix cores with the
// orthonormal matrix M (mode 1); `reps` forward + inverse round trips.
struct XTabs {
    float F[64], I[64];      // REDFT10 / REDFT01 tables (Tabs::dctF, dctI)
    float nrm[64], inv[64];  // coef_norm and coef_norm_inv
    float dct_out;           // 1 / (2k)
    float M[8][8];           // orthonormal DCT-II matrix, zero-padded to 8x8
};
hipError_t launch_xform_bench(int K, int mode, const float* in, float* out, long npatch, int reps, const XTabs* t,
                              hipStream_t s);
hipError_t launch_color(float* v, unsigned W, unsigned H, unsigned F, bool forward,
                        hipStream_t s);
hipError_t launch_filter_adjoint(int K, const FArgs& a, hipStream_t s);
hipError_t launch_bm_queue(int K, int C, const BMArgs* jobs, const WorkQueue& q, int words, int waves,
                           int blocks, int shape, hipStream_t s);
hipError_t launch_filter_queue(int K, int C, int step, const FArgs* jobs, const WorkQueue& q, int words,
                               int waves, int blocks, hipStream_t s);
hipError_t launch_bm_multi(int K, int C, const BMArgs* jobs, const WorkQueue& q, int words, int waves,
                           int shape, hipStream_t s);
hipError_t launch_filter_multi(int K, int C, int step, const FArgs* jobs, const WorkQueue& q, int words,
                               int waves, hipStream_t s);
hipError_t launch_aggregate_queue(int C, const AArgs* jobs, const WorkQueue& q, int blocks, int maxN, hipStream_t s);
hipError_t launch_aggregate_multi(int C, const AArgs* jobs, const WorkQueue& q, int maxN, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
int bm_lds_words(int K, int C, int Ns, int Npr, int Nb, int Nf);  // per wave, 4-byte words
int filter_words(int K, int C, int N, int step);
hipError_t launch_adjoint_fallback(float* gz, const float* gy, const float* den, size_t n, hipStream_t s);
hipError_t launch_sum_counts(const uint32_t* counts, long n, unsigned long long* stats, hipStream_t s);
hipError_t launch_flow_absmax(const float* fl, size_t n, unsigned* out, hipStream_t s);
hipError_t launch_bm_gen(const GenArgs& a, hipStream_t s);
hipError_t launch_filter_gen(int step, const GenArgs& a, hipStream_t s);
hipError_t launch_aggregate_gen(int C, const AArgs& a, const uint32_t* mpos, const TabsG* tabs, unsigned frames,
                                hipStream_t s);

}  // namespace vb
