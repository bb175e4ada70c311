// kernels_agg.hip -- CDNA4 (gfx950) ordered aggregation of the VBM3D
// collaborative filter: one 32x16 output tile per workgroup; contributions
// are gathered in the reference's (crop, frame, row, column, n) order, so
// every pixel's sum is formed in the same order as the CPU and no atomics
// are used.  Also the colour transform and the small utility kernels.
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
// Aggregation (src/vbm3d.cpp:612-650, 860-924) with the division and the
// den == 0 fallback (:168-170, :198-202, :267-280, :333-346).
// ---------------------------------------------------------------------
__device__ __forceinline__ int lower_bound_u16(const uint16_t* a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One 32x16 output tile of one frame per workgroup (eight wavefronts).
// The reference patches whose groups can reach the tile are scanned in
// processing order (crop, frame, row, column), one per thread at a time;
// the group members that overlap the tile are compacted into an ordered
// list, and each wavefront then accumulates, for its 8x8 block, the
// members that overlap the block in list order: every output sample sees
// its contributions in exactly the reference's order.
// ONE: the grid is a single crop (nb = 1); the per-crop sums are then the
// totals, and no crop-order merge is kept live across the scan.
template <int C, bool RAW, int NM, bool ONE, class A>
__device__ __forceinline__ void agg_tile(const A& a, const int bx, const int by, const unsigned bz, const int nl) {
    typedef const __attribute__((address_space(4))) float cfloat;
    constexpr int TX = kAggTileX, TY = kAggTileY, NT = kAggThreads;
    constexpr int QS = 8;             // a wavefront's block edge
    constexpr int NWV = NT / 64, WX = TX / QS;
    static_assert(TX % QS == 0 && TY % QS == 0 && NT % 64 == 0, "tile of 8x8 blocks");
    // the ordered list of one NT-patch chunk (at most nl >= N members
    // each): entry = {position x | y << 16, slot * kMaxN + n}, one 8-byte
    // LDS read per entry
    extern __shared__ uint2 agg_lists[];
    uint2* L_e = agg_lists;
    __shared__ float s_kai[kMaxWin * kMaxWin * 4];
    __shared__ int s_w[NWV];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int x0 = bx * TX, y0 = by * TY;
    const unsigned to = a.t_begin + bz;
    const int qx0 = x0 + (wv % WX) * QS, qy0 = y0 + (wv / WX) * QS;  // this wave's block
    const int px = qx0 + (lane & (QS - 1)), py = qy0 + lane / QS;
    const bool inside = px < (int)a.W && py < (int)a.H;
    const int K = a.K, K2 = K * K, N = a.N;
    const unsigned W = a.W, wh = a.wh, whc = a.whc;
    for (int i = tid; i < K2; i += NT) s_kai[i] = a.tabs->kaiser[i];
    float tn[C], td[C];
#pragma unroll
    for (int c = 0; c < C; c++) tn[c] = td[c] = 0.f;
    const int trlo = max(0, (int)to - a.Nf), trhi = min((int)a.Fr - 1, (int)to + a.Nf);
    const int nfr = trhi - trlo + 1;
    // a member at q = to*whc + y*W + x overlaps the tile's rows iff
    // q - to*whc lies in [ylo*W, yhi*W)
    const unsigned tbase = to * whc;
    const Div32 dW = {a.divW.m, a.divW.s1, a.divW.s2};
    const unsigned rmin = (unsigned)max(0, y0 - K + 1) * W, rmax = (unsigned)min((int)a.H, y0 + TY) * W;

    // per frame of a crop's box: {first index, columns, first row, first
    // column} of the grid patches whose members can reach the tile, and the
    // division by the column count
    __shared__ int4 s_box[2 * kGenMaxNf + 1];
    __shared__ int s_bnd[6 * (2 * kGenMaxNf + 1)];
    __shared__ Div32 s_div[2 * kGenMaxNf + 1];
    __shared__ int s_total;

    for (int ci = 0; ci < (ONE ? 1 : a.ncrop); ci++) {
        const int* meta = a.crop_meta + ci * 5;
        const int off = meta[0], nrows = meta[1], ncols = meta[2];
        const uint16_t* rows = a.grid_pos + meta[3];
        const uint16_t* cols = a.grid_pos + meta[4];
        // A member k frames away from its reference patch lies within
        // min(R, R0 + k Rs) of it, so the box narrows towards the output
        // frame.  The frames' boxes are concatenated in frame order: the
        // scan below still visits the patches in processing order (crop,
        // frame, row, column).
        __syncthreads();  // the previous crop's table is no longer read
        if (a.boxtab) {
            // the host's table of the same bounds: one load per bound, the
            // columns' division with them
            if (tid < 2 * nfr) {
                const int q = tid >> 1, col = tid & 1;
                const int k = abs(trlo + q - (int)to);
                const size_t tl = (size_t)ci * (a.nty + a.ntx) + (col ? a.nty + bx : by);
                const int4 v = a.boxtab[tl * (a.Nf + 1) + k];
                s_bnd[6 * q + 2 * col] = v.x;
                s_bnd[6 * q + 2 * col + 1] = v.y;
                if (col) {
                    s_bnd[6 * q + 4] = v.z;
                    s_bnd[6 * q + 5] = v.w;
                }
            }
        } else if (tid < 4 * nfr) {
            // the four searches of a frame run in four threads (each is a
            // chain of dependent loads)
            const int q = tid >> 2, which = tid & 3;
            const int k = abs(trlo + q - (int)to);
            const int Rk = (int)min((long)a.R, (long)a.R0 + (long)k * a.Rs);
            const bool row = which < 2;
            const int lo = row ? y0 - Rk - (K - 1) : x0 - Rk - (K - 1);
            const int hi = row ? y0 + TY + Rk : x0 + TX + Rk;
            s_bnd[6 * q + which] = lower_bound_u16(row ? rows : cols, row ? nrows : ncols, (which & 1) ? hi : lo);
        }
        __syncthreads();
        // wavefront 0: lane q < nfr assembles frame q's box; the boxes'
        // first indices are an exclusive scan of their patch counts
        if (wv == 0) {
            int cnt = 0, nc = 0, ia = 0, ja = 0;
            if (lane < nfr) {
                ia = s_bnd[6 * lane];
                ja = s_bnd[6 * lane + 2];
                const int nr = max(s_bnd[6 * lane + 1] - ia, 0);
                nc = max(s_bnd[6 * lane + 3] - ja, 0);
                cnt = nr * nc;
                if (a.boxtab) {
                    const unsigned sh = (unsigned)s_bnd[6 * lane + 5];
                    s_div[lane] = Div32{(uint32_t)s_bnd[6 * lane + 4], sh & 0xFFu, sh >> 8};
                } else {
                    s_div[lane] = make_div32(nc > 0 ? nc : 1);
                }
            }
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const int y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            if (lane < nfr) s_box[lane] = make_int4(incl - cnt, nc, ia, ja);
            if (lane == nfr - 1) s_total = incl;
        }
        __syncthreads();
        const long total = s_total;
        float pn[C], pd[C];
#pragma unroll
        for (int c = 0; c < C; c++) pn[c] = pd[c] = 0.f;

        // The scan data of a chunk (slot, group size, member indices, up
        // to NM >= N of them) is loaded one chunk ahead: the loads for chunk
        // cb + NT fly while chunk cb is accumulated.
        uint32_t slot_n = 0, mi_n[NM];
        int ns_n = 0;
        auto fetch = [&](long cb) {
            const unsigned e = (unsigned)(cb + tid);
            slot_n = 0;
            ns_n = 0;
#pragma unroll
            for (int n = 0; n < NM; n++) mi_n[n] = 0u;
            if (e < total) {
                // the frame whose box holds index e: the last one starting
                // at or before it (empty boxes start where the next does)
                int q = 0;
#pragma unroll
                for (int st = 16; st >= 1; st >>= 1)
                    if (q + st < nfr && (int)e >= s_box[q + st].x) q += st;
                const int4 bx = s_box[q];
                const Div32 dq = s_div[q];
                const unsigned rest = e - (unsigned)bx.x;
                const unsigned ii = div32(rest, dq), jj = rest - ii * (unsigned)bx.y;
                const unsigned tr = trlo + q;
                slot_n = (tr % a.ring) * a.nrf + off + (bx.z + ii) * ncols + (bx.w + jj);
                ns_n = (int)a.nsx[slot_n];
                const uint32_t* mem = a.idx + (size_t)slot_n * N;
#pragma unroll
                for (int n = 0; n < NM; n++) mi_n[n] = n < N ? mem[n] : 0u;
            }
        };
        if (total > 0) fetch(0);
        for (long cb = 0; cb < total; cb += NT) {
            const uint32_t slot = slot_n;
            const int ns = ns_n;
            uint32_t mi[NM];
#pragma unroll
            for (int n = 0; n < NM; n++) mi[n] = mi_n[n];
            uint32_t hits = 0;
#pragma unroll
            for (int n = 0; n < NM; n++) {
                if (n >= ns) break;
                const unsigned r = mi[n] - tbase;
                if (r < rmin || r >= rmax) continue;
                const int qy = (int)div32(r, dW), qx = (int)(r - (unsigned)qy * W);
                if (qx <= x0 + TX - 1 && qx + K - 1 >= x0) hits |= 1u << n;
            }
            // ordered compaction: exclusive block scan of the hit counts
            const int cnt = __popc(hits);
            int x = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) s_w[wv] = x;
            __syncthreads();
            int base = 0, tot = 0;
#pragma unroll 1
            for (int q = 0; q < NWV; q++) {
                const int sq = s_w[q];
                base += q < wv ? sq : 0;
                tot += sq;
            }
            int o = base + x - cnt;
#pragma unroll
            for (int n = 0; n < NM; n++)
                if ((hits >> n) & 1u) {
                    const unsigned r = mi[n] - tbase;
                    const unsigned qy = div32(r, dW), qx = r - qy * W;
                    L_e[o] = make_uint2(qx | (qy << 16), slot * kMaxN + n);
                    o++;
                }
            __syncthreads();
            if (cb + NT < total) fetch(cb + NT);
            // this wave: the list entries overlapping its 8x8 block, in order
            for (int j0 = 0; j0 < tot; j0 += 64) {
                bool ov = false;
                if (j0 + lane < tot) {
                    const uint32_t pos = L_e[j0 + lane].x;
                    const int ex = (int)(pos & 0xFFFFu), ey = (int)(pos >> 16);
                    ov = ex <= qx0 + QS - 1 && ex + K - 1 >= qx0 && ey <= qy0 + QS - 1 && ey + K - 1 >= qy0;
                }
                uint64_t m = __ballot(ov);
                while (m) {
                    // up to 4 entries at a time: every load is issued before
                    // the first sum, and the sums then run in list order.
                    // The entry is the same for every lane: its position,
                    // slot and member are decoded on the scalar unit, its
                    // weights arrive through scalar loads.
                    constexpr int U = 4;
                    int jj[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        jj[u] = m ? j0 + (int)__builtin_ctzll(m) : -1;
                        m &= m - 1;
                    }
                    // straight-line loads: an absent entry repeats the
                    // first one, a lane outside a patch reads the patch's
                    // sample nearest to it (within the rows and columns the
                    // patch's own lanes read, so no extra cache line); only
                    // the sums below are masked
                    float ev[U][C], kai[U], w[U][C];
                    bool on[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int jv = jj[u] >= 0 ? jj[u] : jj[0];
                        const uint2 ent = L_e[jv];
                        const uint32_t pos = __builtin_amdgcn_readfirstlane(ent.x);
                        const uint32_t sn = __builtin_amdgcn_readfirstlane(ent.y);
                        const uint32_t sl = sn / kMaxN, n = sn % kMaxN;
                        const cfloat* wb = (const cfloat*)(a.wgt + (size_t)sl * C);
#pragma unroll
                        for (int c = 0; c < C; c++) w[u][c] = RAW ? 0.f : wb[c];
                        const int dx = px - (int)(pos & 0xFFFFu), dy = py - (int)(pos >> 16);
                        on[u] = (jj[u] >= 0) & inside & ((unsigned)dx < (unsigned)K) & ((unsigned)dy < (unsigned)K);
                        // clamp to [0, K-1] in one instruction per axis
                        int cx, cy;
                        asm("v_med3_i32 %0, %1, 0, %2" : "=v"(cx) : "v"(dx), "s"(K - 1));
                        asm("v_med3_i32 %0, %1, 0, %2" : "=v"(cy) : "v"(dy), "s"(K - 1));
                        const unsigned kk = __umul24((unsigned)cy, (unsigned)K) + (unsigned)cx;
                        kai[u] = s_kai[kk];
                        // wave-uniform base, 32-bit lane byte offset; the
                        // plane index slot * C * N + n fits 32 bits (the
                        // host refuses larger rings)
                        const unsigned pl = sl * (unsigned)(C * N) + n;
                        const char* eb = (const char*)a.est + (size_t)pl * (unsigned)(K2 * 4);
#pragma unroll
                        for (int c = 0; c < C; c++)
                            ev[u][c] = *(const float*)(eb + (size_t)c * N * K2 * 4 + kk * 4u);
                    }
                    // keeps the loads above from being sunk to their sums
                    // (the weights' scalar loads included)
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int c = 0; c < C; c++) asm volatile("" ::"s"(w[u][c]));
                    asm volatile("" ::: "memory");
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (!on[u]) continue;
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            if (RAW) {
                                pn[c] = pn[c] + ev[u][c];
                            } else {
                                const float kw = kai[u] * w[u][c];
                                pn[c] = pn[c] + kw * ev[u][c];
                                pd[c] = pd[c] + kw;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            if (ONE) {
                tn[c] = pn[c];
                td[c] = pd[c];
            } else {
                tn[c] = tn[c] + pn[c];
                td[c] = td[c] + pd[c];
            }
        }
    }
    if (inside) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            const size_t o = (size_t)to * whc + (size_t)c * wh + (size_t)py * W + px;
            if (a.num_out) a.num_out[o] = tn[c];
            if (a.den_out) a.den_out[o] = td[c];
            if (a.out) a.out[o] = td[c] == 0.f ? a.fallback[o] : tn[c] / td[c];
        }
    }
}

// Occupancy floor of 8 waves per SIMD (64 registers, two of them spilled):
// the kernel waits on memory more than it computes, and measured faster
// than at its unconstrained 7 waves.
template <int C, bool RAW, int NM>
__global__ __launch_bounds__(kAggThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_aggregate(AArgs a) {
    if (a.ncrop == 1) agg_tile<C, RAW, NM, true>(a, blockIdx.x, blockIdx.y, blockIdx.z, a.N);
    else agg_tile<C, RAW, NM, false>(a, blockIdx.x, blockIdx.y, blockIdx.z, a.N);
}

// Batched form with one workgroup per tile: workgroup b takes item b, so
// the hardware's dispatch order is the queue; item -> (job, frame, tile
// row, tile column).  The same occupancy floor as k_aggregate.
template <int C, int NM>
__global__ __launch_bounds__(kAggThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_aggregate_multi(
    const AArgs* __restrict__ jobs, WorkQueue q, int nl) {
    const long it = (long)blockIdx.x;
    if (it >= q.total) return;  // workgroup-uniform
    const int j = queue_job(q, it);
    typedef const __attribute__((address_space(4))) AArgs cAA;
    cAA& a = ((cAA*)jobs)[j];  // scalar loads of the job
    const long local = it - uniform_long(q.item_begin[j]);
    const long tx = (a.W + kAggTileX - 1) / kAggTileX, ty = (a.H + kAggTileY - 1) / kAggTileY;
    const long bz = local / (tx * ty), rem = local - bz * tx * ty;
    if (a.ncrop == 1) agg_tile<C, false, NM, true>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
    else agg_tile<C, false, NM, false>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
}

// Persistent form: one tile per take (the workgroup is the unit); item ->
// (job, frame, tile row, tile column).
template <int C, int NM>
__global__ __launch_bounds__(kAggThreads) void k_aggregate_queue(const AArgs* __restrict__ jobs, WorkQueue q, int nl) {
    __shared__ long s_item;
    for (;;) {
        __syncthreads();  // every thread has read the previous item
        if (threadIdx.x == 0) s_item = (long)atomicAdd(q.counter, 1ull);
        __syncthreads();
        const long it = s_item;
        if (it >= q.total) break;
        const int j = queue_job(q, it);
        typedef const __attribute__((address_space(4))) AArgs cAA;
        cAA& a = ((cAA*)jobs)[j];  // scalar loads of the job
        const long local = it - q.item_begin[j];
        const long tx = (a.W + kAggTileX - 1) / kAggTileX, ty = (a.H + kAggTileY - 1) / kAggTileY;
        const long bz = local / (tx * ty), rem = local - bz * tx * ty;
        if (a.ncrop == 1) agg_tile<C, false, NM, true>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
        else agg_tile<C, false, NM, false>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
    }
}

// ---------------------------------------------------------------------
// Colour transform, 3 channels (src/Utilities/LibVideoT.hpp:777-887).
// ---------------------------------------------------------------------
__global__ void k_color(float* v, unsigned wh, unsigned F, int forward) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)wh * F) return;
    const size_t f = i / wh, k = i - f * wh;
    float* r = v + f * 3 * wh;
    const float R = r[k], G = r[wh + k], B = r[2 * wh + k];
    float o0, o1, o2;
    if (forward) {
        o0 = 0.299f * R + 0.587f * G + 0.114f * B;
        o1 = -0.14713f * R - 0.28886f * G + 0.436f * B;
        o2 = 0.615f * R - 0.51498f * G - 0.10001f * B;
    } else {
        o0 = R + 1.13983f * B;
        o1 = R - 0.39465f * G - 0.5806f * B;
        o2 = R + 2.03211f * G;
    }
    r[k] = o0;
    r[wh + k] = o1;
    r[2 * wh + k] = o2;
}

// Clears the work-queue counters of a batch (a kernel rather than a
// memset, so that it can sit inside a captured graph).
__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Work counters of one block-matching launch: stats[0] += candidates,
// stats[1] += group members (one atomic per block).
__global__ void k_sum_counts(const uint32_t* cnt, long n, unsigned long long* stats) {
    __shared__ unsigned long long part[2][4];
    unsigned long long c0 = 0, c1 = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        c0 += cnt[2 * i];
        c1 += cnt[2 * i + 1];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        c0 += (unsigned long long)__shfl_xor((long long)c0, o, 64);
        c1 += (unsigned long long)__shfl_xor((long long)c1, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { part[0][w] = c0; part[1][w] = c1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s0 = 0, s1 = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) { s0 += part[0][k]; s1 += part[1][k]; }
        atomicAdd(&stats[0], s0);
        atomicAdd(&stats[1], s1);
    }
}

__global__ void k_flow_absmax(const float* fl, size_t n, unsigned* out) {
    unsigned m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        m = max(m, (unsigned)fminf(fabsf(roundf(fl[i])), 65535.f));  // NaN and huge values -> 65535
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}


}  // namespace

// ---------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------
hipError_t launch_aggregate_multi(int C, const AArgs* jobs, const WorkQueue& q, int maxN, hipStream_t s) {
    if (q.total <= 0) return hipSuccess;
    if (maxN < 1 || maxN > kMaxN || q.total > 0x7FFFFFFFL) return hipErrorInvalidValue;
    const size_t lds = (size_t)kAggThreads * maxN * sizeof(uint2);
    const dim3 grid((unsigned)q.total), blk(kAggThreads);
    if (maxN <= 8) {
        if (C == 1) hipLaunchKernelGGL((k_aggregate_multi<1, 8>), grid, blk, lds, s, jobs, q, maxN);
        else hipLaunchKernelGGL((k_aggregate_multi<3, 8>), grid, blk, lds, s, jobs, q, maxN);
    } else {
        if (C == 1) hipLaunchKernelGGL((k_aggregate_multi<1, 16>), grid, blk, lds, s, jobs, q, maxN);
        else hipLaunchKernelGGL((k_aggregate_multi<3, 16>), grid, blk, lds, s, jobs, q, maxN);
    }
    return hipGetLastError();
}

hipError_t launch_aggregate_queue(int C, const AArgs* jobs, const WorkQueue& q, int blocks, int maxN, hipStream_t s) {
    if (q.total <= 0 || blocks <= 0) return hipSuccess;
    if (maxN < 1 || maxN > kMaxN) return hipErrorInvalidValue;
    const size_t lds = (size_t)kAggThreads * maxN * sizeof(uint2);
    const dim3 blk(kAggThreads);
    // members held in registers by the scan: 8, or 16 when some job has N = 16
    if (maxN <= 8) {
        if (C == 1) hipLaunchKernelGGL((k_aggregate_queue<1, 8>), dim3(blocks), blk, lds, s, jobs, q, maxN);
        else hipLaunchKernelGGL((k_aggregate_queue<3, 8>), dim3(blocks), blk, lds, s, jobs, q, maxN);
    } else {
        if (C == 1) hipLaunchKernelGGL((k_aggregate_queue<1, 16>), dim3(blocks), blk, lds, s, jobs, q, maxN);
        else hipLaunchKernelGGL((k_aggregate_queue<3, 16>), dim3(blocks), blk, lds, s, jobs, q, maxN);
    }
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_aggregate(int C, const AArgs& a, unsigned frames, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    dim3 grid((a.W + kAggTileX - 1) / kAggTileX, (a.H + kAggTileY - 1) / kAggTileY, frames);
    if (a.N < 1 || a.N > kMaxN) return hipErrorInvalidValue;
    const size_t lds = (size_t)kAggThreads * a.N * sizeof(uint2);  // the chunk's ordered list
    const dim3 blk(kAggThreads);
    // members held in registers by the scan: 8, or 16 for N = 16
#define VB_AGG(NM)                                                                         \
    if (a.raw) {                                                                           \
        if (C == 1) hipLaunchKernelGGL((k_aggregate<1, true, NM>), grid, blk, lds, s, a);  \
        else hipLaunchKernelGGL((k_aggregate<3, true, NM>), grid, blk, lds, s, a);         \
    } else {                                                                               \
        if (C == 1) hipLaunchKernelGGL((k_aggregate<1, false, NM>), grid, blk, lds, s, a); \
        else hipLaunchKernelGGL((k_aggregate<3, false, NM>), grid, blk, lds, s, a);        \
    }
    if (a.N <= 8) {
        VB_AGG(8)
    } else {
        VB_AGG(16)
    }
#undef VB_AGG
    return hipGetLastError();
}

hipError_t launch_color(float* v, unsigned W, unsigned H, unsigned F, bool forward,
                        hipStream_t s) {
    const size_t n = (size_t)W * H * F;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_color, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v,
                       W * H, F, forward ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_sum_counts(const uint32_t* counts, long n, unsigned long long* stats, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_counts, dim3(256), dim3(256), 0, s, counts, n, stats);
    return hipGetLastError();
}

hipError_t launch_flow_absmax(const float* fl, size_t n, unsigned* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flow_absmax, dim3(1024), dim3(256), 0, s, fl, n, out);
    return hipGetLastError();
}

}  // namespace vb
