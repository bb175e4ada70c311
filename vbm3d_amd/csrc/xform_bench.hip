// xform_bench.hip -- CDNA4 (gfx950) micro-benchmark of the group
// filter's 2D transforms: the k x k DCT forward and inverse (k = 7, 8) on
// the vector ALUs, exactly as k_filter computes them, against the same
// transforms on the matrix cores (FP32 MFMA).  It answers one question:
// do MFMA instructions beat VALU for these 8x8-sized products?  Nothing
// in the denoising path calls it.
//
// Both kernels keep a wavefront's patches on chip and apply `reps`
// forward + inverse round trips, so the figure is transform throughput,
// not memory traffic; the result is written once so that no work is
// dropped.  One "transform" is a 2D DCT of one k x k patch (2 k^3
// multiply-adds as separable products).
//
// VALU (k_xform_valu): lane = (patch, row), 64 / k patches per wavefront,
// the parity-folded passes of k_filter with LDS transposes and the
// coef_norm scaling; bit-exact with the denoiser's transforms.
//
// MFMA (k_xform_mfma): lane = 8 p + r holds row r of patch p (8 patches
// per wavefront, k padded to 8 with zeros).  One separable pass is
// 16 v_mfma_f32_4x4x1f32 (16 4x4 blocks each): block b = lane / 4 takes
// rows 4 (b % 2) .. + 3 of patch b / 2, the A operand is the lane's own
// row value at step q, the B operand a table of the transform matrix.
// After a pass a lane holds 4 rows of two columns; one exchange with
// lane ^ 4 gives it a whole column, i.e. the transposed layout the next
// pass consumes, so four passes make a forward + inverse round trip with
// no LDS.  FP32 products and sums in the matrix core's order: not the
// reference's rounding.
#include "kernels_common.h"

#pragma clang fp contract(off)

namespace vb {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int K>
__global__ __launch_bounds__(256) void k_xform_valu(const float* __restrict__ in, float* out, long npatch, int reps,
                                                   const XTabs* __restrict__ t) {
    constexpr int K2 = K * K, PB = 64 / K, XS = xstride(K), XP = K + 1;
    __shared__ float lds[4][PB * XS + PB * K2];
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat* TF = (const cfloat*)t->F;
    const cfloat* TI = (const cfloat*)t->I;
    const cfloat* NRM = (const cfloat*)t->nrm;
    const cfloat* INV = (const cfloat*)t->inv;
    const float dct_out = t->dct_out;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* X = lds[w];
    float* G = lds[w] + PB * XS;
    const long base = ((long)blockIdx.x * 4 + w) * PB;
    const int nl = lane / K, rr = lane - nl * K;
    const bool act = lane < PB * K && base + nl < npatch;
    float x[K];
#pragma unroll
    for (int q = 0; q < K; q++) x[q] = act ? in[(base + nl) * K2 + rr * K + q] : 0.f;
    for (int it = 0; it < reps; it++) {
        float y[K], col[K];
        dct_fwd_pass<K>(x, y, TF);
#pragma unroll
        for (int v = 0; v < K; v++)
            if (act) X[nl * XS + v * XP + rr] = y[v];
        wave_sync();
#pragma unroll
        for (int p = 0; p < K; p++) col[p] = act ? X[nl * XS + rr * XP + p] : 0.f;
        dct_fwd_pass<K>(col, y, TF);
#pragma unroll
        for (int u = 0; u < K; u++)
            if (act) G[nl * K2 + u * K + rr] = y[u] * NRM[u * K + rr];
        wave_sync();
#pragma unroll
        for (int s2 = 0; s2 < K; s2++) col[s2] = act ? G[nl * K2 + rr * K + s2] * INV[rr * K + s2] : 0.f;
        dct_inv_pass<K>(col, y, TI);
        wave_sync();
#pragma unroll
        for (int v = 0; v < K; v++)
            if (act) X[nl * XS + v * XP + rr] = y[v];
        wave_sync();
#pragma unroll
        for (int p = 0; p < K; p++) col[p] = act ? X[nl * XS + rr * XP + p] : 0.f;
        dct_inv_pass<K>(col, y, TI);
#pragma unroll
        for (int u = 0; u < K; u++) x[u] = dct_out * y[u];
        wave_sync();
    }
    // a round trip leaves lane rr with column rr of the patch; after an
    // even number of them the lane holds row rr again
    if (act)
#pragma unroll
        for (int q = 0; q < K; q++) out[(base + nl) * K2 + ((reps & 1) ? q * K + rr : rr * K + q)] = x[q];
}

// One separable pass of 8 patches: lane 8 p + r holds v[0..7], row r of
// patch p; returns acc0[i] = out[4 s + i][j], acc1[i] = out[4 s + i][4 + j]
// with out[row][c] = sum_q v_row[q] * B[q][c], s = (lane / 4) % 2,
// j = lane % 4; tab[cb][q] = B[q][4 cb + j].
__device__ __forceinline__ void mfma_pass(const float v[8], const float tab[2][8], f4& acc0, f4& acc1) {
    acc0 = f4{0.f, 0.f, 0.f, 0.f};
    acc1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        acc0 = __builtin_amdgcn_mfma_f32_4x4x1f32(v[q], tab[0][q], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_4x4x1f32(v[q], tab[1][q], acc1, 0, 0, 0);
    }
}

// From the pass layout to whole columns: lane 8 p + c ends with column c
// of its patch in v[0..7].
__device__ __forceinline__ void mfma_exchange(const f4& acc0, const f4& acc1, float v[8], int lane) {
    const bool s = (lane >> 2) & 1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float send = s ? acc0[i] : acc1[i];
        const float recv = __shfl_xor(send, 4, 64);
        v[i] = s ? recv : acc0[i];
        v[4 + i] = s ? acc1[i] : recv;
    }
}

template <int K>
__global__ __launch_bounds__(256) void k_xform_mfma(const float* __restrict__ in, float* out, long npatch, int reps,
                                                   const XTabs* __restrict__ t) {
    constexpr int K2 = K * K;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long base = ((long)blockIdx.x * 4 + w) * 8;
    const int p = lane >> 3, r = lane & 7, j = lane & 3;
    const bool act = r < K && base + p < npatch;
    // forward: out = M X M^T, B[q][c] = M[c][q]; inverse: B[q][c] = M[q][c]
    // flat indices into the 8x8 table (row-major): written as t->M[a][b]
    // these loads compiled wrongly with this toolchain (rows 5-7 of the
    // forward table came from row 0 of the inverse's loads)
    const float* Mf = &t->M[0][0];
    float tf[2][8], ti[2][8];
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int q = 0; q < 8; q++) {
            tf[cb][q] = Mf[(4 * cb + j) * 8 + q];
            ti[cb][q] = Mf[q * 8 + 4 * cb + j];
        }
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = act && q < K ? in[(base + p) * K2 + r * K + q] : 0.f;
    f4 a0, a1;
    for (int it = 0; it < reps; it++) {
        mfma_pass(v, tf, a0, a1);
        mfma_exchange(a0, a1, v, lane);
        mfma_pass(v, tf, a0, a1);
        mfma_exchange(a0, a1, v, lane);
        mfma_pass(v, ti, a0, a1);
        mfma_exchange(a0, a1, v, lane);
        mfma_pass(v, ti, a0, a1);
        mfma_exchange(a0, a1, v, lane);
    }
    if (act)
#pragma unroll
        for (int q = 0; q < K; q++) out[(base + p) * K2 + r * K + q] = v[q];
}

}  // namespace

hipError_t launch_xform_bench(int K, int mode, const float* in, float* out, long npatch, int reps, const XTabs* t,
                              hipStream_t s) {
    if ((K != 7 && K != 8) || npatch <= 0 || reps <= 0 || (mode != 0 && mode != 1)) return hipErrorInvalidValue;
    const long per_wave = mode == 0 ? 64 / K : 8;
    const long waves = (npatch + per_wave - 1) / per_wave;
    const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
    if (mode == 0) {
        if (K == 7) hipLaunchKernelGGL(k_xform_valu<7>, grid, blk, 0, s, in, out, npatch, reps, t);
        else hipLaunchKernelGGL(k_xform_valu<8>, grid, blk, 0, s, in, out, npatch, reps, t);
    } else {
        if (K == 7) hipLaunchKernelGGL(k_xform_mfma<7>, grid, blk, 0, s, in, out, npatch, reps, t);
        else hipLaunchKernelGGL(k_xform_mfma<8>, grid, blk, 0, s, in, out, npatch, reps, t);
    }
    return hipGetLastError();
}

}  // namespace vb
