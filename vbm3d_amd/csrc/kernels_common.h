// kernels_common.h -- wavefront helpers shared by the HIP kernels of the
// VBM3D collaborative filter (kernels_bm.hip, kernels_filter.hip,
// kernels_agg.hip).  Device-only, internal linkage: every translation unit
// gets its own copy.
#pragma once
#include "kernels.h"
#include "select.h"

#pragma clang fp contract(off)

namespace vb {
namespace {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups of a launch are dealt to the 8 XCDs in turn (workgroup b runs
// on XCD b % 8), and every XCD has its own L2.  xcd_order maps b to a
// logical index such that XCD x takes one contiguous range of the logical
// order, where neighbouring patches read the same data.  It is a bijection
// on [0, n).  Measured: it speeds up the group filter; the aggregation
// runs faster in the plain order, where all XCDs work on neighbouring data
// at the same time and share it in the last-level cache (also faster than
// band_order below, though that reads 14 % fewer bytes).
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned n) {
    const unsigned q = n >> 3, r = n & 7, x = b & 7, j = b >> 3;
    return x * q + min(x, r) + j;
}

// Band order: the logical order is cut into super-blocks of 8 s
// workgroups (about one frame of reference patches when s is an eighth of
// a frame's workgroups); within each, XCD x takes the x-th run of s
// workgroups, so that all XCDs work on the same frame at once, each on
// its own band of neighbouring patches.  A last partial super-block is
// split as xcd_order does.  A bijection on [0, n).  Block matching in this
// order takes the time of the plain order with a fifth of its L2-miss
// bytes; one contiguous range per XCD (xcd_order) was slower, its XCDs
// ending at different times.
__device__ __forceinline__ unsigned band_order(unsigned b, unsigned n, unsigned s) {
    const unsigned q = s * 8, full = n / q * q;
    if (b >= full) return full + xcd_order(b - full, n - full);
    const unsigned x = b & 7, k = b >> 3, sb = k / s;
    return sb * q + x * s + (k - sb * s);
}

// A global work queue: items [0, total) of a batched launch, taken `grab`
// at a time through an atomic counter (zeroed before the launch); job j
// owns items [item_begin[j], item_begin[j+1]).
// queue_issue sends the atomic (lane 0 holds the reply); queue_value makes
// it wave-uniform.  Issuing the next take before working on the current
// items hides the atomic's round trip.
__device__ __forceinline__ unsigned long long queue_issue(const WorkQueue& q, int lane) {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(q.counter, (unsigned long long)q.grab);
    return v;
}

__device__ __forceinline__ long queue_value(unsigned long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ long uniform_long(long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ long queue_take(const WorkQueue& q, int lane) {
    return queue_value(queue_issue(q, lane));
}

// The job that owns item `it`, starting from the cached job j with items
// [jb, je): a search only when `it` has left that range.
__device__ __forceinline__ void queue_seek(const WorkQueue& q, long it, int& j, long& jb, long& je);

__device__ __forceinline__ int queue_job(const WorkQueue& q, long it) {
    int lo = 0, hi = q.njobs - 1;  // largest j with item_begin[j] <= it
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.item_begin[mid] <= it) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);  // wave-uniform: scalar loads of the job follow
}

__device__ __forceinline__ void queue_seek(const WorkQueue& q, long it, int& j, long& jb, long& je) {
    if (it >= jb && it < je) return;
    j = queue_job(q, it);
    jb = uniform_long(q.item_begin[j]);
    je = uniform_long(q.item_begin[j + 1]);
}

// Monotone map float -> uint32 (for finite inputs); -0 and +0 map alike so
// that equal floats give equal keys.
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned u = __float_as_uint(f + 0.0f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Wave-wide reductions on the DPP network: quad swaps, half-row and row
// mirrors, then the row-15 / lane-31 broadcasts; lane 63 ends with the
// result, which readlane makes wave-uniform.
// The "old" operand is the operation's identity (all ones for an unsigned
// min, 0 for a sum): lanes a step does not write keep their value, and the
// compiler folds each move into the min or add as a DPP operand.
#define VB_DPPMIN(v, ctrl, rmask) __builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)(v), ctrl, rmask, 0xF, false)
#define VB_DPP0(v, ctrl, rmask) __builtin_amdgcn_update_dpp(0, (int)(v), ctrl, rmask, 0xF, false)

__device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
    v = min(v, (unsigned)VB_DPPMIN(v, 0xB1, 0xF));   // quad_perm [1,0,3,2]
    v = min(v, (unsigned)VB_DPPMIN(v, 0x4E, 0xF));   // quad_perm [2,3,0,1]
    v = min(v, (unsigned)VB_DPPMIN(v, 0x141, 0xF));  // row_half_mirror
    v = min(v, (unsigned)VB_DPPMIN(v, 0x140, 0xF));  // row_mirror
    v = min(v, (unsigned)VB_DPPMIN(v, 0x142, 0xA));  // row_bcast15
    v = min(v, (unsigned)VB_DPPMIN(v, 0x143, 0xC));  // row_bcast31
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ int wave_sum_i32(int v) {
    v += VB_DPP0(v, 0xB1, 0xF);
    v += VB_DPP0(v, 0x4E, 0xF);
    v += VB_DPP0(v, 0x141, 0xF);
    v += VB_DPP0(v, 0x140, 0xF);
    v += VB_DPP0(v, 0x142, 0xA);
    v += VB_DPP0(v, 0x143, 0xC);
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long m, int lane) {
    return (unsigned)__popcll(m & ((1ull << lane) - 1ull));
}

// floor(n / d) for 0 <= n < 4096, 1 <= d <= 64: the reciprocal's relative
// error (about 1e-7) stays far below the 0.5/d margin of n + 0.5.
__device__ __forceinline__ int small_div(int n, int d) {
    return (int)(((float)n + 0.5f) * __builtin_amdgcn_rcpf((float)d));
}

// The `mid` smallest of list (ld, li)[0..m), m <= 64, in std::partial_sort
// order, written to (od, oi)[0..mid).  Fast path: repeated wave arg-min; a
// value equal to another remaining value at ranks < mid means the heap's
// tie handling decides, and lane 0 then replays the heap select exactly.
__device__ void wave_select(const float* ld, const uint32_t* li, int m, int mid,
                            float* od, uint32_t* oi, VbEnt* scr, int lane) {
    if (mid > 0) {
        unsigned key = lane < m ? fkey(ld[lane]) : 0xFFFFFFFFu;
        bool tie = false;
        for (int j = 0; j < mid; j++) {
            const unsigned mk = wave_min_u32(key);
            const unsigned long long bal = __ballot(key == mk);
            if (__popcll(bal) != 1) { tie = true; break; }
            const int pos = __ffsll((long long)bal) - 1;
            if (lane == 0) { od[j] = ld[pos]; oi[j] = li[pos]; }
            if (lane == pos) key = 0xFFFFFFFFu;
        }
        if (tie) {
            wave_sync();
            if (lane == 0) {
                for (int q = 0; q < m; q++) { scr[q].d = ld[q]; scr[q].i = li[q]; }
                vb_partial_sort(scr, m, mid);
                for (int q = 0; q < mid; q++) { od[q] = scr[q].d; oi[q] = scr[q].i; }
            }
        }
    }
    wave_sync();
}

// The same selection over a sparse list: lane l holds entry l when `valid`
// (bal = ballot of valid); the list is the valid entries in lane order.
// Only the tie path compacts them (into cd, ci) for the exact replay.
__device__ void wave_select_sparse(const float* ld, const uint32_t* li, bool valid, unsigned long long bal,
                                   int mid, float* od, uint32_t* oi, float* cd, uint32_t* ci, VbEnt* scr,
                                   int lane) {
    if (mid > 0) {
        unsigned key = valid ? fkey(ld[lane]) : 0xFFFFFFFFu;
        bool tie = false;
        for (int j = 0; j < mid; j++) {
            const unsigned mk = wave_min_u32(key);
            const unsigned long long b = __ballot(key == mk);
            if (__popcll(b) != 1) { tie = true; break; }
            const int pos = __ffsll((long long)b) - 1;
            if (lane == 0) { od[j] = ld[pos]; oi[j] = li[pos]; }
            if (lane == pos) key = 0xFFFFFFFFu;
        }
        if (tie) {
            if (valid) {
                const unsigned p = lanes_below(bal, lane);
                cd[p] = ld[lane];
                ci[p] = li[lane];
            }
            wave_sync();
            const int m = __popcll(bal);
            if (lane == 0) {
                for (int q = 0; q < m; q++) { scr[q].d = cd[q]; scr[q].i = ci[q]; }
                vb_partial_sort(scr, m, mid);
                for (int q = 0; q < mid; q++) { od[q] = scr[q].d; oi[q] = scr[q].i; }
            }
        }
    }
    wave_sync();
}

typedef __attribute__((address_space(3))) void lds_void;

// LDS byte address of a pointer into the wave's LDS, made wave-uniform.
__device__ __forceinline__ unsigned lds_addr(const float* p) {
    return __builtin_amdgcn_readfirstlane((unsigned)(size_t)(const lds_void*)p);
}

// Copies one 4-byte word per active lane from src (per lane) to the LDS
// byte address lds + 4*lane (lds wave-uniform) without staging it in
// registers (global_load_lds_dword).  The caller waits for vmcnt before
// reading it.
__device__ __forceinline__ void lds_dma(const void* src, unsigned lds) {
    __builtin_amdgcn_global_load_lds(src, (lds_void*)(size_t)lds, 4, 0, 0);
}

// ---------------------------------------------------------------------
// Transforms along the group (registers, NS members).
// ---------------------------------------------------------------------
__device__ __forceinline__ float haar_scale(float s) {
    return (float)((double)s * 0.7071067811865475);  // SQRT2_INV, src/vbm3d.cpp:33
}

// haar_forward (src/lib_transforms.cpp:453-476) on registers.
template <int NS>
__device__ __forceinline__ void haar_fwd(float* v) {
    float tmp[NS > 1 ? NS : 1];
#pragma unroll
    for (int k = NS / 2; k >= 1; k /= 2) {
#pragma unroll
        for (int i = 0; i < k; i++) {
            tmp[i] = haar_scale(v[2 * i] + v[2 * i + 1]);
            tmp[i + k] = haar_scale(v[2 * i] - v[2 * i + 1]);
        }
#pragma unroll
        for (int i = 0; i < 2 * k; i++) v[i] = tmp[i];
    }
}

// haar_inverse (src/lib_transforms.cpp:489-512) on registers.
template <int NS>
__device__ __forceinline__ void haar_inv(float* v) {
    float tmp[NS > 1 ? NS : 1];
#pragma unroll
    for (int k = 1; k * 2 <= NS; k *= 2) {
#pragma unroll
        for (int i = 0; i < k; i++) {
            tmp[2 * i] = haar_scale(v[i] + v[i + k]);
            tmp[2 * i + 1] = haar_scale(v[i] - v[i + k]);
        }
#pragma unroll
        for (int i = 0; i < 2 * k; i++) v[i] = tmp[i];
    }
}

// hadamard_transform, unnormalised (src/lib_transforms.cpp:410-440).
template <int NS>
__device__ __forceinline__ void hadamard(float* v) {
    if constexpr (NS == 2) {
        const float a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (NS > 2) {
        constexpr int h = NS / 2;
        float tmp[h];
#pragma unroll
        for (int k = 0; k < h; k++) {
            const float a = v[2 * k], b = v[2 * k + 1];
            v[k] = a + b;
            tmp[k] = a - b;
        }
#pragma unroll
        for (int k = 0; k < h; k++) v[h + k] = tmp[k];
        hadamard<h>(v);
        hadamard<h>(v + h);
    }
}

template <int NS>
__device__ __forceinline__ void t3_fwd(float* v, bool had) {
    if (had) hadamard<NS>(v);
    else haar_fwd<NS>(v);
}
template <int NS>
__device__ __forceinline__ void t3_inv(float* v, bool had) {
    if (had) hadamard<NS>(v);
    else haar_inv<NS>(v);
}

// One 1D pass of REDFT10 on K values (K >= 2): row v of the matrix is
// even (v even) or odd (v odd) about the centre, so each output takes the
// folded inputs x[q] + x[K-1-q] or x[q] - x[K-1-q], q < K/2, plus the
// centre sample for odd K and even v; sums in q order starting from the
// first product, as the oracle's dct_pass_fwd.  About half the
// multiplications of the matrix product.
template <int K, class TAB>
__device__ __forceinline__ void dct_fwd_pass(const float* x, float* y, const TAB* F) {
    constexpr int H = K / 2;
    static_assert(H >= 1, "a pass of at least two samples");
    float s[H], d[H];
#pragma unroll
    for (int q = 0; q < H; q++) {
        s[q] = x[q] + x[K - 1 - q];
        d[q] = x[q] - x[K - 1 - q];
    }
#pragma unroll
    for (int v = 0; v < K; v++) {
        float m;
        if ((v & 1) == 0) {
            m = s[0] * F[v * K];
#pragma unroll
            for (int q = 1; q < H; q++) m = m + s[q] * F[v * K + q];
            if (K & 1) m = m + x[H] * F[v * K + H];
        } else {
            m = d[0] * F[v * K];
#pragma unroll
            for (int q = 1; q < H; q++) m = m + d[q] * F[v * K + q];
        }
        y[v] = m;
    }
}

// Output v of dct_fwd_pass on x[0], x[stride], ... (the same folding and
// the same order of operations, so the two agree bit for bit).
template <int K, class TAB>
__device__ __forceinline__ float dct_fwd_one(const float* x, int stride, int v, const TAB* F) {
    constexpr int H = K / 2;
    static_assert(H >= 1, "a pass of at least two samples");
    float m;
    if ((v & 1) == 0) {
        m = (x[0] + x[(K - 1) * stride]) * F[v * K];
#pragma unroll
        for (int q = 1; q < H; q++) m = m + (x[q * stride] + x[(K - 1 - q) * stride]) * F[v * K + q];
        if (K & 1) m = m + x[H * stride] * F[v * K + H];
    } else {
        m = (x[0] - x[(K - 1) * stride]) * F[v * K];
#pragma unroll
        for (int q = 1; q < H; q++) m = m + (x[q * stride] - x[(K - 1 - q) * stride]) * F[v * K + q];
    }
    return m;
}

// One 1D pass of REDFT01 (I[o][i] = table[o*K + i], K >= 2): rows o and
// K-1-o differ only in the sign of the odd-i terms, so E (even i) and O
// (odd i) give y[o] = E + O and y[K-1-o] = E - O; the centre row of odd K
// is E.  Sums in i order starting from the first product, as the oracle's
// dct_pass_inv.
template <int K, class TAB>
__device__ __forceinline__ void dct_inv_pass(const float* z, float* y, const TAB* I) {
    constexpr int H = K / 2;
    static_assert(H >= 1, "a pass of at least two samples");
#pragma unroll
    for (int o = 0; o < H; o++) {
        float e = z[0] * I[o * K], od = z[1] * I[o * K + 1];
#pragma unroll
        for (int i = 2; i < K; i += 2) e = e + z[i] * I[o * K + i];
#pragma unroll
        for (int i = 3; i < K; i += 2) od = od + z[i] * I[o * K + i];
        y[o] = e + od;
        y[K - 1 - o] = e - od;
    }
    if (K & 1) {
        float e = z[0] * I[H * K];
#pragma unroll
        for (int i = 2; i < K; i += 2) e = e + z[i] * I[H * K + i];
        y[H] = e;
    }
}

// Transposition buffer: patch n of a lane batch is a K x (K+1) tile at
// n*xstride(K) (row pitch K+1).  The strides keep both the row-wise write and
// the column-wise read of the 64 lanes (n, r) = n*K + r bank-conflict free
// at K = 8 and within 2-way at K = 7.
__host__ __device__ constexpr int xstride(int K) { return (K & 1) ? K * (K + 1) + 1 : K * (K + 1); }

}  // namespace
}  // namespace vb
