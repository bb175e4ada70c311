// kernels_filter.hip -- CDNA4 (gfx950) group filtering of the VBM3D
// collaborative filter (2D transforms, transform along the group,
// shrinkage, inverse transforms) and its step-2 adjoint: one wavefront per
// reference patch; lane = coefficient position, the group dimension lives
// in registers.
//
// Every kernel reproduces the reference's FP32 operation order so that the
// GPU result is bit-identical to the sequential CPU restatement:
//   * no FMA contraction (file-wide pragma + -ffp-contract=off);
//   * float sums accumulated sequentially in the reference's index order;
//   * IEEE-correct float division (hipcc's default);
//   * the Haar butterflies multiply a float sum by a double constant and
//     round once to float, as src/lib_transforms.cpp:468-469 does.
#include "kernels_common.h"

#pragma clang fp contract(off)

namespace vb {
namespace {

// ---------------------------------------------------------------------
// Group filtering.
// ---------------------------------------------------------------------
// Column of (member n, coefficient kk) in row kk of the Wiener v^2 buffer:
// n rotated by kk / (64 / NS), so that the 64 lanes' stores of one n fall
// in 64 distinct banks while each row stays contiguous.
template <int NS>
__host__ __device__ constexpr int v2_col(int kk, int n) {
    return (n + (kk >> (NS == 1 ? 6 : NS == 2 ? 5 : NS == 4 ? 4 : NS == 8 ? 3 : 2))) & (NS - 1);
}

// 1D transform along the group and shrinkage for one channel
// (ht_filtering_haar/_hadamard src/vbm3d.cpp:1405-1521,
// wiener_filtering_haar/_hadamard :1539-1666).  Lane kk holds coefficient
// kk of every group member: zg/eg are [n][64].  Step 1 returns the
// aggregation weight (valid in every lane); step 2 stores value^2 of
// (n, kk) in row kk of v2, at column v2_col for the weight sum done
// afterwards in the reference's flat order n + kk*nSx (wiener_sum).
template <int K, int NS, int STEP>
__device__ float shrink(float* zg, float* eg, float* v2, int lane, bool had, float lambda,
                        float sig) {
    constexpr int K2 = K * K;
    const bool act = lane < K2;
    float z[NS], e[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
        z[n] = act ? zg[n * K2 + lane] : 0.f;
        if (STEP == 2) e[n] = act ? eg[n * K2 + lane] : 0.f;
    }
    float weight = 0.f;
    if (STEP == 1) {
        t3_fwd<NS>(z, had);
        const float thr = had ? lambda * sig * sqrtf((float)NS) : lambda * sig;
        int cnt = 0;
#pragma unroll
        for (int n = 0; n < NS; n++) {
            const bool keep = act && ((n == 0 && lane == 0) || fabsf(z[n]) > thr);
            if (keep) cnt++;
            else z[n] = 0.f;
        }
        cnt = wave_sum_i32(cnt);
        t3_inv<NS>(z, had);
        if (had) {
            const float coef = 1.0f / (float)NS;
#pragma unroll
            for (int n = 0; n < NS; n++) z[n] *= coef;
        }
        weight = 1.f / (sig * sig * (float)cnt);
#pragma unroll
        for (int n = 0; n < NS; n++)
            if (act) zg[n * K2 + lane] = z[n];
    } else {
        t3_fwd<NS>(z, had);
        t3_fwd<NS>(e, had);
        const float s2 = sig * sig;
        const float coef = 1.0f / (float)NS;
#pragma unroll
        for (int n = 0; n < NS; n++) {
            float value = e[n] * e[n];
            if (had) value = value * coef;
            value /= (value + s2);
            e[n] = had ? z[n] * value * coef : z[n] * value;
            if (act) v2[lane * NS + v2_col<NS>(lane, n)] = value * value;
        }
        t3_inv<NS>(e, had);
#pragma unroll
        for (int n = 0; n < NS; n++)
            if (act) eg[n * K2 + lane] = e[n];
    }
    return weight;
}

// Sequential float sum of channel c's v^2 over the flat index
// f = n + kk*nSx (src/vbm3d.cpp:1651).  Rows are NS words, read 16 bytes
// at a time; the rotation undone here is v2_col's.
template <int K, int NS>
__device__ __forceinline__ float wiener_sum(const float* vc) {
    constexpr int K2 = K * K;
    float acc = 0.f;
#pragma unroll
    for (int kk = 0; kk < K2; kk++) {
        float x[NS];
        if constexpr (NS >= 4) {
#pragma unroll
            for (int j = 0; j < NS / 4; j++) {
                const float4 q = *(const float4*)(vc + kk * NS + 4 * j);
                x[4 * j] = q.x;
                x[4 * j + 1] = q.y;
                x[4 * j + 2] = q.z;
                x[4 * j + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < NS; j++) x[j] = vc[kk * NS + j];
        }
#pragma unroll
        for (int n = 0; n < NS; n++) acc = acc + x[v2_col<NS>(kk, n)];
    }
    return acc;
}

// Per-wavefront LDS layout of k_filter (4-byte words): transposes, zg,
// eg, then one v^2 buffer per channel, each region 16-byte aligned.
__host__ __device__ constexpr int round4(int v) { return (v + 3) & ~3; }
__host__ __device__ inline int filter_lds_words(int K, int C, int N, int step) {
    const int pb = 64 / K;
    int w = round4(pb * xstride(K)) + round4(N * K * K);
    if (step == 2) w += round4(N * K * K) + C * round4(K * K * N);
    return w;
}

// Channel sigma without a dynamically indexed array (a private array
// would live in scratch memory).
template <class A>
__device__ __forceinline__ float sig_of(const A& a, int c) {
    return c == 0 ? a.sig[0] : (c == 1 ? a.sig[1] : a.sig[2]);
}

template <int K, int C, int STEP, class A>
__device__ __forceinline__ void filter_one(const A& a, const long r, float* X, const int lane) {
    constexpr int K2 = K * K;
    constexpr int PB = 64 / K;           // patches per lane batch
    constexpr int XS = xstride(K), XP = K + 1;
    // X: [PB][XS] transposes, then the group buffers
    const int gw = round4(a.N * K2);               // group buffer words
    float* zg = X + round4(PB * XS);               // [N][K2]
    float* eg = zg + gw;                           // [N][K2] (step 2)
    float* v2 = eg + gw;                           // [C][gw] (step 2)
    const unsigned t = (unsigned)(r / a.nrf), i = (unsigned)(r - (long)t * a.nrf);
    const unsigned slot = __builtin_amdgcn_readfirstlane((t % a.ring) * a.nrf + i);
    const unsigned W = a.W, wh = a.wh, whc = a.whc;
    const int nsx = (int)__builtin_amdgcn_readfirstlane(a.nsx[slot]);
    const bool bior = a.T2D == 1;
    const bool had = a.T3D == 2;
    const uint32_t* mem = a.idx + (size_t)slot * a.N;
    const Div32 dW = {a.dW.m, a.dW.s1, a.dW.s2}, dWHC = {a.dWHC.m, a.dWHC.s1, a.dWHC.s2};
    typedef const __attribute__((address_space(4))) float cfloat;
    const cfloat* TF = (const cfloat*)(bior ? a.tabs->biorF : a.tabs->dctF);
    const cfloat* TI = (const cfloat*)(bior ? a.tabs->biorI : a.tabs->dctI);
    const cfloat* NRM = (const cfloat*)a.tabs->nrm;
    const cfloat* INV = (const cfloat*)a.tabs->inv;
    const float dct_out = a.tabs->dct_out;
    const int nl = lane / K, rr = lane - nl * K;    // batch patch, row / column index
    const bool lact = lane < PB * K;
    float wsel = 0.f;  // step 1: lane c ends with channel c's weight
    // mb0: offset of row rr of member nl, the lane's row in the first lane
    // batch (its index is read whether or not the member exists, so that
    // the load does not wait for the group size; every slot holds N
    // entries).  The first batch's row of each (channel, source) pair is
    // fetched one pair ahead of its use.
    auto decode = [&](unsigned mi) {
        const unsigned pt = div32(mi, dWHC);
        const unsigned py = div32(mi - pt * whc, dW), px = mi - pt * whc - py * W;
        return pt * whc + (py + rr) * W + px;
    };
    const unsigned mb0 = decode(mem[min(nl, a.N - 1)]);
    const bool act0 = lact && nl < nsx;
    float xn[K];
    auto fetch = [&](int c, int g) {
        if (act0) {
            const float* row = (g == 0 ? a.noisy : a.basic) + mb0 + c * wh;
#pragma unroll
            for (int q = 0; q < K; q++) xn[q] = row[q];
        }
    };
    fetch(0, 0);

    for (int c = 0; c < C; c++) {
        // ---- forward 2D transforms of all members (bior_2d_forward
        //      src/lib_transforms.cpp:125-181; DCT src/vbm3d.cpp:1245-1302)
        for (int g = 0; g < (STEP == 2 ? 2 : 1); g++) {
            const float* src = g == 0 ? a.noisy : a.basic;
            float* dst = g == 0 ? zg : eg;
            for (int nb0 = 0; nb0 < nsx; nb0 += PB) {
                const int n = nb0 + nl;
                const bool act = lact && n < nsx;
                if (act) {
                    float x[K];
                    if (nb0 == 0) {
                        // the first batch's row was fetched ahead; fetch the
                        // next (channel, source) row while this one is used
#pragma unroll
                        for (int q = 0; q < K; q++) x[q] = xn[q];
                        if (STEP == 2 && g == 0) fetch(c, 1);
                        else if (c + 1 < C) fetch(c + 1, 0);
                    } else {
                        const float* row = src + decode(mem[n]) + c * wh;
#pragma unroll
                        for (int q = 0; q < K; q++) x[q] = row[q];
                    }
                    // pass 1 along the row: t[v] = sum_q x[q] M[v][q]
                    if (bior) {
#pragma unroll
                        for (int v = 0; v < K; v++) {
                            float m = 0.f;
#pragma unroll
                            for (int q = 0; q < K; q++) m = m + x[q] * TF[v * K + q];
                            X[nl * XS + v * XP + rr] = m;  // stored transposed
                        }
                    } else {
                        float y[K];
                        dct_fwd_pass<K>(x, y, TF);
#pragma unroll
                        for (int v = 0; v < K; v++) X[nl * XS + v * XP + rr] = y[v];
                    }
                }
                wave_sync();
                if (act) {
                    // pass 2 down column rr of t
                    float col[K];
#pragma unroll
                    for (int p = 0; p < K; p++) col[p] = X[nl * XS + rr * XP + p];
                    // BIOR keeps out[i*8+j] = (column i, output j); the DCT
                    // writes out[u*K+v] scaled by coef_norm
                    if (bior) {
#pragma unroll
                        for (int u = 0; u < K; u++) {
                            float m = 0.f;
#pragma unroll
                            for (int p = 0; p < K; p++) m = m + col[p] * TF[u * K + p];
                            dst[n * K2 + rr * 8 + u] = m;
                        }
                    } else {
                        float y[K];
                        dct_fwd_pass<K>(col, y, TF);
#pragma unroll
                        for (int u = 0; u < K; u++) dst[n * K2 + u * K + rr] = y[u] * NRM[u * K + rr];
                    }
                }
                wave_sync();
            }
        }
        // ---- 1D transform along the group, shrinkage, inverse
        float* v2c = v2 + c * gw;
        const float sgc = sig_of(a, c);
        float wc;
        switch (nsx) {
            case 1: wc = shrink<K, 1, STEP>(zg, eg, v2c, lane, had, a.lambda, sgc); break;
            case 2: wc = shrink<K, 2, STEP>(zg, eg, v2c, lane, had, a.lambda, sgc); break;
            case 4: wc = shrink<K, 4, STEP>(zg, eg, v2c, lane, had, a.lambda, sgc); break;
            case 8: wc = shrink<K, 8, STEP>(zg, eg, v2c, lane, had, a.lambda, sgc); break;
            default: wc = shrink<K, 16, STEP>(zg, eg, v2c, lane, had, a.lambda, sgc); break;
        }
        if (lane == c) wsel = wc;
        wave_sync();
        // ---- inverse 2D transforms (bior_2d_inverse src/lib_transforms.cpp:
        //      270-321; dct_2d_inv src/vbm3d.cpp:1678-1713), store [n][pixel].
        //      A member's outputs replace its coefficients in LDS (pass 1
        //      of its batch has read them all), and the group's planes then
        //      leave in contiguous 256-byte stores.
        float* est = STEP == 2 ? eg : zg;
        float* dstg = a.est + (((size_t)slot * C + c) * a.N) * K2;
        for (int nb0 = 0; nb0 < nsx; nb0 += PB) {
            const int n = nb0 + nl;
            const bool act = lact && n < nsx;
            if (act) {
                float x[K];
#pragma unroll
                for (int s2 = 0; s2 < K; s2++)
                    x[s2] = bior ? est[n * K2 + s2 * 8 + rr] : est[n * K2 + rr * K + s2] * INV[rr * K + s2];
                if (bior) {
#pragma unroll
                    for (int v = 0; v < K; v++) {
                        float m = 0.f;
#pragma unroll
                        for (int q = 0; q < K; q++) m = m + x[q] * TI[v * K + q];
                        X[nl * XS + v * XP + rr] = m;
                    }
                } else {
                    float y[K];
                    dct_inv_pass<K>(x, y, TI);
#pragma unroll
                    for (int v = 0; v < K; v++) X[nl * XS + v * XP + rr] = y[v];
                }
            }
            wave_sync();
            if (act) {
                float col[K];
#pragma unroll
                for (int p = 0; p < K; p++) col[p] = X[nl * XS + rr * XP + p];
                // BIOR: pixel (row u, column rr); DCT: pixel (row u, column
                // rr) after the 1/(2k) factor
                if (bior) {
#pragma unroll
                    for (int u = 0; u < K; u++) {
                        float m = 0.f;
#pragma unroll
                        for (int p = 0; p < K; p++) m = m + col[p] * TI[u * K + p];
                        est[n * K2 + u * K + rr] = m;
                    }
                } else {
                    float y[K];
                    dct_inv_pass<K>(col, y, TI);
#pragma unroll
                    for (int u = 0; u < K; u++) est[n * K2 + u * K + rr] = dct_out * y[u];
                }
            }
            wave_sync();
        }
        for (int j = lane; j < nsx * K2; j += 64) dstg[j] = est[j];
        wave_sync();  // read before the next channel's transforms refill it
    }
    if (STEP == 2) {
        // Wiener weights: sequential float sums over the flat index
        // (src/vbm3d.cpp:1651), one lane per channel
        wave_sync();
        if (lane < C) {
            const float* vc = v2 + lane * gw;
            float acc;
            switch (nsx) {
                case 1: acc = wiener_sum<K, 1>(vc); break;
                case 2: acc = wiener_sum<K, 2>(vc); break;
                case 4: acc = wiener_sum<K, 4>(vc); break;
                case 8: acc = wiener_sum<K, 8>(vc); break;
                default: acc = wiener_sum<K, 16>(vc); break;
            }
            const float sg = sig_of(a, lane), s2 = sg * sg;
            float wgt;
            if (had) wgt = acc > 0.0f ? 1.0f / (s2 * acc) : 1.0f;
            else wgt = 1.f / (s2 * acc);
            a.wgt[(size_t)slot * C + lane] = wgt;
        }
    } else if (lane < C) {
        a.wgt[(size_t)slot * C + lane] = wsel;
    }
}

template <int K, int C, int STEP>
__global__ __launch_bounds__(256) void k_filter(FArgs a) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const long r = a.r_begin + (long)xcd_order(blockIdx.x, gridDim.x) * nw + w;
    if (r >= a.r_end) return;
    filter_one<K, C, STEP>(a, r, smem + (size_t)w * filter_lds_words(K, C, a.N, STEP), lane);
}

// One item per wave over a batch: global wave g takes item g, so the
// hardware's workgroup dispatch order is the queue.
template <int K, int C, int STEP>
__global__ __launch_bounds__(256) void k_filter_multi(const FArgs* __restrict__ jobs, WorkQueue q, int words) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const long g = uniform_long((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (g >= q.total) return;
    const int j = queue_job(q, g);
    typedef const __attribute__((address_space(4))) FArgs cFA;
    cFA& a = ((cFA*)jobs)[j];
    filter_one<K, C, STEP>(a, a.r_begin + (g - uniform_long(q.item_begin[j])),
                           smem + (size_t)(threadIdx.x >> 6) * words, lane);
}

template <int K, int C, int STEP>
__global__ __launch_bounds__(256) void k_filter_queue(const FArgs* __restrict__ jobs, WorkQueue q, int words) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    float* X = smem + (size_t)(threadIdx.x >> 6) * words;
    typedef const __attribute__((address_space(4))) FArgs cFA;
    int j = 0;
    long jb = 0, je = 0;
    long it0 = queue_take(q, lane);
    while (it0 < q.total) {
        const unsigned long long nxt = queue_issue(q, lane);
        const long it1 = min(q.total, it0 + (long)q.grab);
        for (long it = it0; it < it1; it++) {
            queue_seek(q, it, j, jb, je);
            cFA& a = ((cFA*)jobs)[j];  // scalar loads of the job
            filter_one<K, C, STEP>(a, a.r_begin + (it - jb), X, lane);
        }
        it0 = queue_value(nxt);
    }
}

// ---------------------------------------------------------------------
// Adjoint of step 2 with the groups, Wiener gains and weights frozen by
// the basic estimate.  Step 2 is then linear in the noisy video:
//   final = D^-1 sum_g Agg_g (kaiser w_g) G T3^-1 V_g T3 F Gather_g z,
// so its transpose gathers u = grad / den, applies the transposes of the
// inverse 2D transform G, of the group operator (T3^-1 V T3)^T, which is
// T3^-1 V T3 again for the orthonormal Haar and for the symmetric
// Hadamard, and of the forward 2D transform F, scaled by w_g.  The
// per-member results go to `est`; the aggregation kernel in raw mode then
// sums them over every member patch covering a pixel.  Lane = coefficient
// (or pixel) index; these are plain FP32 evaluations of the transposed
// operators, not bit-replays of a reference computation.
// ---------------------------------------------------------------------
// Forward 2D transform of the K x K patch `in` (LDS, row-major) as the
// forward pass computes it; returns coefficient `lane`.
template <int K>
__device__ float adj_fwd2d(const float* in, float* tmp, int lane, bool bior, const Tabs& T) {
    constexpr int K2 = K * K;
    const bool act = lane < K2;
    const int a = lane / K, b = lane - (lane / K) * K;
    float o = 0.f;
    if (bior) {
        if (act) {
            float m = 0.f;
            for (int s = 0; s < 8; s++) m = m + in[a * 8 + s] * T.biorF[b * 8 + s];
            tmp[lane] = m;
        }
        wave_sync();
        if (act)
            for (int s = 0; s < 8; s++) o = o + tmp[s * 8 + a] * T.biorF[b * 8 + s];
    } else {
        // the forward filter's folded passes, so that the Wiener gains
        // taken from these coefficients are the forward pass's bits
        if (act) tmp[lane] = dct_fwd_one<K>(in + a * K, 1, b, T.dctF);
        wave_sync();
        if (act) o = dct_fwd_one<K>(tmp + b, K, a, T.dctF) * T.nrm[lane];
    }
    wave_sync();
    return o;
}

// Transpose of the inverse 2D transform applied to the pixel tile `in`;
// returns coefficient `lane`.  DCT: G(c) = dct_out M (c o inv) M^T with
// M[x][y] = dctI[x*K+y], so G^T(r) = inv o dct_out (M^T r M).  BIOR:
// G(c)[u][m] = sum S[u][p] c[q][p] S[m][q] with S = biorI, so
// G^T(r)[q][p] = sum r[u][m] S[u][p] S[m][q].
template <int K>
__device__ float adj_inv2d_T(const float* in, float* tmp, int lane, bool bior, const Tabs& T) {
    constexpr int K2 = K * K;
    const bool act = lane < K2;
    const int a = lane / K, b = lane - (lane / K) * K;
    float o = 0.f;
    if (bior) {
        if (act) {  // t[u][q] = sum_m r[u][m] S[m][q]
            float m = 0.f;
            for (int mm = 0; mm < 8; mm++) m = m + in[a * 8 + mm] * T.biorI[mm * 8 + b];
            tmp[lane] = m;
        }
        wave_sync();
        if (act)  // c'[q = a][p = b] = sum_u S[u][p] t[u][q]
            for (int u = 0; u < 8; u++) o = o + T.biorI[u * 8 + b] * tmp[u * 8 + a];
    } else {
        if (act) {  // t[u][b] = sum_m r[u][m] M[m][b]
            float m = 0.f;
            for (int mm = 0; mm < K; mm++) m = m + in[a * K + mm] * T.dctI[mm * K + b];
            tmp[lane] = m;
        }
        wave_sync();
        if (act) {  // c'[a][b] = inv dct_out sum_u M[u][a] t[u][b]
            for (int u = 0; u < K; u++) o = o + T.dctI[u * K + a] * tmp[u * K + b];
            o = T.inv[lane] * T.dct_out * o;
        }
    }
    wave_sync();
    return o;
}

// Transpose of the forward 2D transform applied to the coefficient tile
// `in`; returns pixel `lane`.  DCT: F(x) = nrm o (A x A^T), A = dctF, so
// F^T(c) = A^T (nrm o c) A.  BIOR: F(x) = B x^T B^T, B = biorF, so
// F^T(c)[s][q] = sum c[i][j] B[j][s] B[i][q].
template <int K>
__device__ float adj_fwd2d_T(const float* in, float* tmp, int lane, bool bior, const Tabs& T) {
    constexpr int K2 = K * K;
    const bool act = lane < K2;
    const int a = lane / K, b = lane - (lane / K) * K;
    float o = 0.f;
    if (bior) {
        if (act) {  // t[i][s] = sum_j c[i][j] B[j][s]
            float m = 0.f;
            for (int j = 0; j < 8; j++) m = m + in[a * 8 + j] * T.biorF[j * 8 + b];
            tmp[lane] = m;
        }
        wave_sync();
        if (act)  // y[s = a][q = b] = sum_i t[i][s] B[i][q]
            for (int i = 0; i < 8; i++) o = o + tmp[i * 8 + a] * T.biorF[i * 8 + b];
    } else {
        if (act) {  // t[u][q] = sum_v nrm[u][v] c[u][v] A[v][q]
            float m = 0.f;
            for (int v = 0; v < K; v++) m = m + T.nrm[a * K + v] * in[a * K + v] * T.dctF[v * K + b];
            tmp[lane] = m;
        }
        wave_sync();
        if (act)  // y[p = a][q = b] = sum_u A[u][p] t[u][q]
            for (int u = 0; u < K; u++) o = o + T.dctF[u * K + a] * tmp[u * K + b];
    }
    wave_sync();
    return o;
}

// Group operator of one channel, transposed: Wiener gains from the basic
// coefficients eg, then T3^-1 V T3 applied to zg in place; v^2 to v2 at
// the flat index n + kk*nSx.
template <int K, int NS>
__device__ void adj_group(float* zg, const float* eg, float* v2, int lane, bool had, float sig) {
    constexpr int K2 = K * K;
    const bool act = lane < K2;
    float z[NS], e[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
        z[n] = act ? zg[n * K2 + lane] : 0.f;
        e[n] = act ? eg[n * K2 + lane] : 0.f;
    }
    t3_fwd<NS>(e, had);
    t3_fwd<NS>(z, had);
    const float s2 = sig * sig;
    const float coef = 1.0f / (float)NS;
#pragma unroll
    for (int n = 0; n < NS; n++) {
        float value = e[n] * e[n];
        if (had) value = value * coef;
        value /= (value + s2);
        z[n] = had ? z[n] * value * coef : z[n] * value;
        if (act) v2[lane * NS + n] = value * value;
    }
    t3_inv<NS>(z, had);
#pragma unroll
    for (int n = 0; n < NS; n++)
        if (act) zg[n * K2 + lane] = z[n];
}

__host__ __device__ inline int adjoint_lds_words(int K, int N) { return 2 * K * K + 3 * N * K * K; }

template <int K, int C>
__global__ __launch_bounds__(256) void k_filter_adjoint(FArgs a) {
    constexpr int K2 = K * K;
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const long r = a.r_begin + (long)blockIdx.x * nw + w;
    if (r >= a.r_end) return;
    float* P = smem + (size_t)w * adjoint_lds_words(K, a.N);  // [K2] tile
    float* Tm = P + K2;                                         // [K2] scratch
    float* zg = Tm + K2;                                        // [N][K2]
    float* eg = zg + a.N * K2;                                  // [N][K2]
    float* v2 = eg + a.N * K2;                                  // [K2*N]
    const unsigned t = (unsigned)(r / a.nrf), i = (unsigned)(r - (long)t * a.nrf);
    const unsigned slot = __builtin_amdgcn_readfirstlane((t % a.ring) * a.nrf + i);
    const unsigned W = a.W, wh = a.wh, whc = a.whc;
    const int nsx = (int)__builtin_amdgcn_readfirstlane(a.nsx[slot]);
    const bool bior = a.T2D == 1, had = a.T3D == 2;
    const Tabs& T = *a.tabs;
    const uint32_t* mem = a.idx + (size_t)slot * a.N;
    const bool act = lane < K2;
    const int pr = lane / K, pc = lane - (lane / K) * K;

    for (int c = 0; c < C; c++) {
        for (int n = 0; n < nsx; n++) {
            const unsigned mi = mem[n];
            const unsigned px = mi % W, py = (mi % wh) / W, pt = mi / whc;
            const size_t o = (size_t)pt * whc + (size_t)c * wh + (size_t)(py + pr) * W + px + pc;
            if (act) P[lane] = a.basic[o];
            wave_sync();
            const float ev = adj_fwd2d<K>(P, Tm, lane, bior, T);
            if (act) {
                eg[n * K2 + lane] = ev;
                const float dn = a.den[o];
                P[lane] = T.kaiser[lane] * (dn != 0.f ? a.gy[o] / dn : 0.f);
            }
            wave_sync();
            const float zv = adj_inv2d_T<K>(P, Tm, lane, bior, T);
            if (act) zg[n * K2 + lane] = zv;
        }
        wave_sync();
        switch (nsx) {
            case 1: adj_group<K, 1>(zg, eg, v2, lane, had, a.sig[c]); break;
            case 2: adj_group<K, 2>(zg, eg, v2, lane, had, a.sig[c]); break;
            case 4: adj_group<K, 4>(zg, eg, v2, lane, had, a.sig[c]); break;
            case 8: adj_group<K, 8>(zg, eg, v2, lane, had, a.sig[c]); break;
            default: adj_group<K, 16>(zg, eg, v2, lane, had, a.sig[c]); break;
        }
        wave_sync();
        // the group's aggregation weight, summed in the forward pass's order
        float acc = 0.f;
        if (lane == 0)
            for (int f = 0; f < K2 * nsx; f++) acc = acc + v2[f];
        acc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), 0));
        const float s2 = a.sig[c] * a.sig[c];
        const float wg = had ? (acc > 0.0f ? 1.0f / (s2 * acc) : 1.0f) : 1.f / (s2 * acc);
        float* dst = a.est + (((size_t)slot * C + c) * a.N) * K2;
        for (int n = 0; n < nsx; n++) {
            if (act) P[lane] = zg[n * K2 + lane];
            wave_sync();
            const float y = adj_fwd2d_T<K>(P, Tm, lane, bior, T);
            if (act) dst[(size_t)n * K2 + lane] = wg * y;
        }
        wave_sync();
    }
}

__global__ void k_adjoint_fallback(float* gz, const float* gy, const float* den, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && den[i] == 0.f) gz[i] = gz[i] + gy[i];
}

}  // namespace

// ---------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------
template <int K, int C, int STEP>
static hipError_t filter_k(const FArgs& a, hipStream_t s) {
    const long n = a.r_end - a.r_begin;
    if (n <= 0) return hipSuccess;
    const size_t per_wave = (size_t)filter_lds_words(K, C, a.N, STEP) * 4;
    int waves = 4;
    while (waves > 1 && per_wave * waves > 64 * 1024) waves--;
    if (per_wave > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_filter<K, C, STEP>), dim3((unsigned)((n + waves - 1) / waves)),
                       dim3(64 * waves), per_wave * waves, s, a);
    return hipGetLastError();
}

template <int K>
static hipError_t filter_kc(int step, const FArgs& a, hipStream_t s) {
    if (a.C == 1) return step == 1 ? filter_k<K, 1, 1>(a, s) : filter_k<K, 1, 2>(a, s);
    return step == 1 ? filter_k<K, 3, 1>(a, s) : filter_k<K, 3, 2>(a, s);
}

hipError_t launch_filter(int K, int step, const FArgs& a, hipStream_t s) {
    switch (K) {
        case 4: return filter_kc<4>(step, a, s);
        case 5: return filter_kc<5>(step, a, s);
        case 6: return filter_kc<6>(step, a, s);
        case 7: return filter_kc<7>(step, a, s);
        case 8: return filter_kc<8>(step, a, s);
        default: return hipErrorInvalidValue;
    }
}

template <int K>
static hipError_t filter_adj_k(const FArgs& a, hipStream_t s) {
    const long n = a.r_end - a.r_begin;
    if (n <= 0) return hipSuccess;
    const size_t per_wave = (size_t)adjoint_lds_words(K, a.N) * 4;
    int waves = 4;
    while (waves > 1 && per_wave * waves > 64 * 1024) waves--;
    if (per_wave > 64 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + waves - 1) / waves)), block(64 * waves);
    if (a.C == 1) hipLaunchKernelGGL((k_filter_adjoint<K, 1>), grid, block, per_wave * waves, s, a);
    else hipLaunchKernelGGL((k_filter_adjoint<K, 3>), grid, block, per_wave * waves, s, a);
    return hipGetLastError();
}

hipError_t launch_filter_adjoint(int K, const FArgs& a, hipStream_t s) {
    switch (K) {
        case 4: return filter_adj_k<4>(a, s);
        case 5: return filter_adj_k<5>(a, s);
        case 6: return filter_adj_k<6>(a, s);
        case 7: return filter_adj_k<7>(a, s);
        case 8: return filter_adj_k<8>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint_fallback(float* gz, const float* gy, const float* den, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_fallback, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gz, gy, den, n);
    return hipGetLastError();
}

template <int K, int C, int STEP>
static hipError_t filter_queue_k(const FArgs* jobs, const WorkQueue& q, int words, int waves, int blocks,
                                 hipStream_t s) {
    hipLaunchKernelGGL((k_filter_queue<K, C, STEP>), dim3(blocks), dim3(64 * waves), (size_t)words * 4 * waves, s,
                       jobs, q, words);
    return hipGetLastError();
}

template <int K>
static hipError_t filter_queue_kc(int C, int step, const FArgs* jobs, const WorkQueue& q, int words, int waves,
                                  int blocks, hipStream_t s) {
    if (C == 1)
        return step == 1 ? filter_queue_k<K, 1, 1>(jobs, q, words, waves, blocks, s)
                         : filter_queue_k<K, 1, 2>(jobs, q, words, waves, blocks, s);
    return step == 1 ? filter_queue_k<K, 3, 1>(jobs, q, words, waves, blocks, s)
                     : filter_queue_k<K, 3, 2>(jobs, q, words, waves, blocks, s);
}

template <int K, int C, int STEP>
static hipError_t filter_multi_k(const FArgs* jobs, const WorkQueue& q, int words, int waves, hipStream_t s) {
    const long blocks = (q.total + waves - 1) / waves;
    hipLaunchKernelGGL((k_filter_multi<K, C, STEP>), dim3((unsigned)blocks), dim3(64 * waves),
                       (size_t)words * 4 * waves, s, jobs, q, words);
    return hipGetLastError();
}

template <int K>
static hipError_t filter_multi_kc(int C, int step, const FArgs* jobs, const WorkQueue& q, int words, int waves,
                                  hipStream_t s) {
    if (C == 1)
        return step == 1 ? filter_multi_k<K, 1, 1>(jobs, q, words, waves, s)
                         : filter_multi_k<K, 1, 2>(jobs, q, words, waves, s);
    return step == 1 ? filter_multi_k<K, 3, 1>(jobs, q, words, waves, s)
                     : filter_multi_k<K, 3, 2>(jobs, q, words, waves, s);
}

hipError_t launch_filter_multi(int K, int C, int step, const FArgs* jobs, const WorkQueue& q, int words,
                               int waves, hipStream_t s) {
    if (q.total <= 0) return hipSuccess;
    if ((size_t)words * 4 * waves > 64 * 1024 || waves < 1 || waves > 4) return hipErrorInvalidValue;
    switch (K) {
        case 4: return filter_multi_kc<4>(C, step, jobs, q, words, waves, s);
        case 5: return filter_multi_kc<5>(C, step, jobs, q, words, waves, s);
        case 6: return filter_multi_kc<6>(C, step, jobs, q, words, waves, s);
        case 7: return filter_multi_kc<7>(C, step, jobs, q, words, waves, s);
        case 8: return filter_multi_kc<8>(C, step, jobs, q, words, waves, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_filter_queue(int K, int C, int step, const FArgs* jobs, const WorkQueue& q, int words,
                               int waves, int blocks, hipStream_t s) {
    if (q.total <= 0 || blocks <= 0) return hipSuccess;
    if ((size_t)words * 4 * waves > 64 * 1024 || waves < 1 || waves > 4) return hipErrorInvalidValue;
    switch (K) {
        case 4: return filter_queue_kc<4>(C, step, jobs, q, words, waves, blocks, s);
        case 5: return filter_queue_kc<5>(C, step, jobs, q, words, waves, blocks, s);
        case 6: return filter_queue_kc<6>(C, step, jobs, q, words, waves, blocks, s);
        case 7: return filter_queue_kc<7>(C, step, jobs, q, words, waves, blocks, s);
        case 8: return filter_queue_kc<8>(C, step, jobs, q, words, waves, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

int filter_words(int K, int C, int N, int step) { return filter_lds_words(K, C, N, step); }

}  // namespace vb
