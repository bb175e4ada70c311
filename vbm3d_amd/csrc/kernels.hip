// kernels.hip -- CDNA4 (gfx950) kernels of the VBM3D collaborative filter.
//
// Every kernel reproduces the reference's FP32 operation order so that the
// GPU result is bit-identical to the sequential CPU restatement:
//   * no FMA contraction (file-wide pragma + -ffp-contract=off);
//   * float sums accumulated sequentially in the reference's index order;
//   * IEEE-correct float division (hipcc's default);
//   * the Haar butterflies multiply a float sum by a double constant and
//     round once to float, as src/lib_transforms.cpp:468-469 does.
// Work decomposition (one wavefront = 64 lanes):
//   k_bm         one wavefront per reference patch; lanes evaluate the
//                candidates of a search window in parallel, selection is a
//                wave-wide arg-min with a tie check (exact heap-select
//                fallback in one lane when distances tie);
//   k_filter     one wavefront per reference patch; lane = coefficient
//                position, the group dimension lives in registers;
//   k_aggregate  one 32x16 output tile per workgroup; contributions are
//                gathered in the reference's (crop, frame, row, column, n)
//                order, so every pixel's sum is formed in the same order as
//                the CPU and no atomics are used.
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
// on [0, n).  Measured: it speeds up the group filter; block matching and
// aggregation run faster in the plain order, where all XCDs work on
// neighbouring data at the same time and share it in the last-level cache.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned n) {
    const unsigned q = n >> 3, r = n & 7, x = b & 7, j = b >> 3;
    return x * q + min(x, r) + j;
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
// Predictive block matching: computeSimilarPatches / localSearch /
// patchDistance (src/vbm3d.cpp:930-1197).
//
// One wavefront per reference patch.  Each search stage (the reference
// frame, then every forward and backward frame) first copies the pixel
// area its windows cover into a per-wave LDS tile; every lane then forms
// one candidate's sum of squared differences from LDS against the
// reference patch (also in LDS), in the reference's (c, y, x) order.
// ---------------------------------------------------------------------
template <int K, int C, int SH, class A>
__device__ __forceinline__ void bm_one(const A& a, const long r, float* base, const int lane) {
    // SH = 1: the default search shape (Ns 7, Npr 5, Nb 2) fixed at compile
    // time, so that tile and window offsets become instruction immediates
    constexpr bool FIX = SH == 1;
    const int Ns = FIX ? 7 : a.Ns, Npr = FIX ? 5 : a.Npr, Nb = FIX ? 2 : a.Nb;
    const BMLay L = bm_layout(K, C, Ns, Npr, Nb, a.Nf);
    float* tile = base + L.tile;
    float* s_d = base + L.sd;
    uint32_t* s_i = (uint32_t*)(base + L.si);
    uint32_t* s_v = (uint32_t*)(base + L.sv);
    float* c_d = base + L.cd;
    uint32_t* c_i = (uint32_t*)(base + L.ci);
    float* b_d = base + L.bd;
    uint32_t* b_i = (uint32_t*)(base + L.bi);
    float* f_d = base + L.fd;
    uint32_t* f_i = (uint32_t*)(base + L.fi);
    float* o_d = base + L.od;
    uint32_t* o_i = (uint32_t*)(base + L.oi);
    VbEnt* scr = (VbEnt*)(base + L.scr);
    uint32_t* pp = (uint32_t*)(base + L.pp);  // [0..8) backward chain, [8..16) forward
    int* win = (int*)(base + L.win);          // per window: x0, y0, wx, wy, frame

    const unsigned W = a.W, H = a.H, wh = a.wh, whc = a.whc;
    const float* __restrict__ vid = a.vid;
    unsigned ridx, slot;
    if (a.ref_list) {
        ridx = a.ref_list[r];
        slot = (unsigned)(r - a.r_begin);
    } else {
        const unsigned t = (unsigned)(r / a.nrf), i = (unsigned)(r - (long)t * a.nrf);
        const unsigned xy = a.ref_xy[i];
        ridx = t * whc + (xy >> 16) * W + (xy & 0xFFFFu);
        slot = (t % a.ring) * a.nrf + i;
    }
    ridx = __builtin_amdgcn_readfirstlane(ridx);
    slot = __builtin_amdgcn_readfirstlane(slot);
    const Div32 dW = {a.dW.m, a.dW.s1, a.dW.s2}, dWHC = {a.dWHC.m, a.dWHC.s1, a.dWHC.s2};
    const unsigned rt = div32(ridx, dWHC), ry = div32(ridx - rt * whc, dW), rx = ridx - rt * whc - ry * W;
    const float norm = (float)(unsigned)(K * K * a.kt * C);
    constexpr int K2 = K * K;

    // The reference patch is wave-uniform: read it through the constant
    // address space so it arrives in scalar registers (SMEM loads) and
    // costs neither LDS bandwidth nor vector issue slots.
    typedef const __attribute__((address_space(4))) float cfloat;
    float* refp = base + L.ref;  // the reference patch, [c][y][x] (general form only)
    for (int i = lane; i < (L.piped ? 0 : K2 * C); i += 64) {
        const int c = i / K2, e = i - c * K2;
        refp[i] = vid[ridx + c * wh + (e / K) * W + e % K];
    }

    // SSD of the reference patch against the candidate whose top-left
    // pixel sits at t[0] of a tile with pitch tp and channel stride cs.
    auto ssd = [&](const float* t, int tp, int cs) -> float {
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int y = 0; y < K; ++y) {
                cfloat* row = (cfloat*)(vid + ridx + c * wh + y * W);
#pragma unroll
                for (int x = 0; x < K; ++x) {
                    const float dif = row[x] - t[c * cs + y * tp + x];
                    acc = acc + dif * dif;
                }
            }
        return acc;
    };
    // two candidates at once: independent accumulators, shared reference
    // rows; each sum keeps its own sequential order
    auto ssd2 = [&](const float* t, const float* u, int tp, int cs, float& ra, float& rb) {
        float acc = 0.f, acd = 0.f;
#pragma unroll 1
        for (int c = 0; c < C; ++c)
#pragma unroll 2
            for (int y = 0; y < K; ++y) {
#pragma unroll
                for (int x = 0; x < K; ++x) {
                    const float r = refp[(c * K + y) * K + x];
                    const float dif = r - t[c * cs + y * tp + x];
                    const float dig = r - u[c * cs + y * tp + x];
                    acc = acc + dif * dif;
                    acd = acd + dig * dig;
                }
            }
        ra = acc;
        rb = acd;
    };
    // copy the (wx+K-1) x (wy+K-1) area at (x0, y0) of frame tf into a
    // tile; lanes cover 4 rows of 16 columns (tile rows are at most 15 wide)
    const int lrow = lane >> 4, lcol = lane & 15;
    auto load_tile = [&](float* t, int tp, int cs, int x0, int y0, int wx, int wy, unsigned tf) {
        const int aw = wx + K - 1, ah = wy + K - 1;
        const unsigned fb = tf * whc + y0 * W + x0;
        if (lcol < aw)
            for (int c = 0; c < C; c++)
                for (int ty = lrow; ty < ah; ty += 4)
                    t[c * cs + ty * tp + lcol] = vid[fb + c * wh + ty * W + lcol];
    };

    // -- search in the reference frame (src/vbm3d.cpp:1114-1119)
    int nbest;
    unsigned n_eval = 0;
    {
        const int h = (Ns - 1) / 2;
        const int x0 = max(0, (int)rx - h), y0 = max(0, (int)ry - h);
        const int x1 = min((int)W - K, (int)rx + h), y1 = min((int)H - K, (int)ry + h);
        const int wx = x1 - x0 + 1, wy = y1 - y0 + 1, m = wx * wy;
        const int cs = L.tp0 * L.th0;
        {
            // the (wx+K-1) x (wy+K-1) area, memory to LDS: word e = k*64 +
            // lane of a channel's area is its row e / tp0, column e % tp0
            const int aw = wx + K - 1, ah = wy + K - 1;
            const unsigned fb = rt * whc + (unsigned)y0 * W + (unsigned)x0;
            const unsigned tl0 = lds_addr(tile);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k * 64 >= cs) break;
                const int e = k * 64 + lane;
                const int er = small_div(e, L.tp0), ec = e - er * L.tp0;
                if (er < ah && ec < aw) {
                    const unsigned go = ((unsigned)er * W + (unsigned)ec) * 4u;
#pragma unroll
                    for (int c = 0; c < C; c++)
                        lds_dma((const char*)(vid + fb + c * wh) + go, tl0 + (unsigned)((c * cs + k * 64) * 4));
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        wave_sync();
        if (lane < m) {
            const int lx = lane % wx, ly = lane / wx;
            const float s = ssd(tile + ly * L.tp0 + lx, L.tp0, cs) / norm;
            const unsigned qx = x0 + lx, qy = y0 + ly;
            c_d[lane] = s - ((qx == rx && qy == ry) ? a.d : 0.f);
            c_i[lane] = rt * whc + qy * W + qx;
        }
        wave_sync();
        n_eval += m;
        const int nb = min(Nb, m);
        wave_select(c_d, c_i, m, nb, o_d, o_i, scr, lane);
        if (lane < nb) {
            b_d[lane] = o_d[lane];
            b_i[lane] = o_i[lane];
            pp[lane] = o_i[lane];
            pp[8 + lane] = o_i[lane];
        }
        nbest = nb;
        wave_sync();
    }

    // -- predictive search (src/vbm3d.cpp:1121-1167).  The forward and
    //    backward chains only depend on themselves, so forward frame f and
    //    backward frame f are searched in the same pass; their picks are
    //    stored where the reference's sequential order puts them (all
    //    forward frames, then all backward frames).
    const int P2 = Npr * Npr, hs = (Npr - 1) / 2;
    const int P2p = pow2ceil(P2), lg2 = __builtin_ctz(P2p);
    const int cs1 = L.tp1 * L.th1;
    const int nfw = (int)(min(rt + (unsigned)a.Nf, a.F - (unsigned)a.kt) - rt);
    const int nbw = (int)rt - max((int)rt - a.Nf, 0);
    const int nst = max(nfw, nbw);
    // Paired form (at most 2 matches per frame, windows of at most 64
    // slots per chain, at most 12 tile rows): frame f of the forward chain
    // and frame f of the backward chain are searched together.  Chain h
    // owns slots h*64 + m*P2p + l; the valid slots of both chains are
    // listed in slot order and evaluated 64 per pass, so that the lanes
    // are not left idle by windows smaller than the wave.
    const bool piped = L.piped;
    if (piped) {
        const int nfr[2] = {nfw, nbw};
        uint32_t* lst = (uint32_t*)c_d;  // c_d and c_i, adjacent: 128 entries
        // windows of frame f of both chains: flow projection of the chain's
        // current matches (x first, then y at the updated x), clamped
        // Npr x Npr box in the next frame; lanes < 2 Nb
        auto windows = [&](int f) {
            if (lane < 2 * Nb) {
                const int ch = lane >= Nb ? 1 : 0, m = lane - ch * Nb;
                if (f < nfr[ch]) {
                    const int dir = ch == 0 ? 1 : -1;
                    const float* fl = ch == 0 ? a.ff : a.bf;
                    const unsigned mi = pp[(ch == 0 ? 8 : 0) + m];
                    const unsigned pt = div32(mi, dWHC);
                    unsigned py = div32(mi - pt * whc, dW), px = mi - pt * whc - py * W;
                    const unsigned ft = ch == 0 ? pt : pt - 1;
                    const unsigned hk = K / 2;
                    const float fx = fl ? fl[ft * 2 * wh + (py + hk) * W + px + hk] : 0.f;
                    px = (unsigned)min(max((int)((float)px + roundf(fx)), 0), (int)(W - K));
                    const float fy = fl ? fl[ft * 2 * wh + wh + (py + hk) * W + px + hk] : 0.f;
                    py = (unsigned)min(max((int)((float)py + roundf(fy)), 0), (int)(H - K));
                    const int x0 = max(0, (int)px - hs), y0 = max(0, (int)py - hs);
                    win[lane * 5 + 0] = x0;
                    win[lane * 5 + 1] = y0;
                    win[lane * 5 + 2] = min((int)W - K, (int)px + hs) - x0 + 1;
                    win[lane * 5 + 3] = min((int)H - K, (int)py + hs) - y0 + 1;
                    win[lane * 5 + 4] = (int)(pt + dir);
                }
            }
            wave_sync();
        };
        // a (window, channel) area is th1 rows of pitch tp1 (at most 192
        // words); word e = k*64 + lane sits at row e / tp1, column e % tp1
        int er[3], ec[3];
        unsigned goff[3];  // byte offset in the frame
#pragma unroll
        for (int k = 0; k < 3; k++) {
            er[k] = small_div(k * 64 + lane, L.tp1);
            ec[k] = k * 64 + lane - er[k] * L.tp1;
            goff[k] = ((unsigned)er[k] * W + (unsigned)ec[k]) * 4u;
        }
        const unsigned tile_lds = lds_addr(tile);
        // slot lane of a chain: window m = lane / P2p, position l = ly*Npr + lx
        // of the Npr x Npr grid (the valid positions of a clipped window
        // keep its raster order)
        const int sm = lane >> lg2, sl = lane & (P2p - 1);
        const int sly = small_div(sl, Npr), slx = sl - sly * Npr;
        if (nst > 0) windows(0);
        for (int f = 0; f < nst; f++) {
            const bool act[2] = {f < nfw, f < nbw};
            // the areas of all windows of the frame, from memory straight
            // into the tile: a wave-uniform row base plus the lane's byte
            // offset.  The empty asm keeps the LDS addresses from being
            // hoisted out of the frame loop.
            {
                const int ne = L.th1 * L.tp1;
                unsigned tl = tile_lds;
                asm volatile("" : "+s"(tl));
#pragma unroll
                for (int ch = 0; ch < 2; ch++)
#pragma unroll
                    for (int m = 0; m < 2; m++) {
                        if (!act[ch] || m >= Nb) continue;
                        const int wn = ch * Nb + m;
                        const int aw = __builtin_amdgcn_readfirstlane(win[wn * 5 + 2]) + K - 1;
                        const int ah = __builtin_amdgcn_readfirstlane(win[wn * 5 + 3]) + K - 1;
                        const unsigned fb = (unsigned)__builtin_amdgcn_readfirstlane(win[wn * 5 + 4]) * whc +
                                            (unsigned)__builtin_amdgcn_readfirstlane(win[wn * 5 + 1]) * W +
                                            (unsigned)__builtin_amdgcn_readfirstlane(win[wn * 5]);
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            if (k * 64 >= ne) break;
                            if (er[k] < ah && ec[k] < aw) {
#pragma unroll
                                for (int c = 0; c < C; c++)
                                    lds_dma((const char*)(vid + fb + c * wh) + goff[k],
                                            tl + (unsigned)(((wn * C + c) * cs1 + k * 64) * 4));
                            }
                        }
                    }
                __builtin_amdgcn_s_waitcnt(0);
            }
            wave_sync();
            // slot lane of chain h; a valid slot keeps its tile offset (bit
            // 31: the reference's own position) in s_d until its distance
            // replaces it
            unsigned long long bal[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                // the chain's windows 0 and 1 (window 0 again when Nb = 1),
                // read at wave-uniform addresses; lanes of window 1 select
                // its values, and a candidate inside window 0 was seen
                const int* wa = win + (h * Nb) * 5;
                const int* wb = win + (h * Nb + (Nb > 1 ? 1 : 0)) * 5;
                const int ax0 = wa[0], ay0 = wa[1], awx = wa[2], awy = wa[3];
                const int bx0 = wb[0], by0 = wb[1], bwx = wb[2], bwy = wb[3];
                const bool m1 = sm == 1;
                const int x0 = m1 ? bx0 : ax0, y0 = m1 ? by0 : ay0;
                const int wx = m1 ? bwx : awx, wy = m1 ? bwy : awy;
                const int qx = x0 + slx, qy = y0 + sly;
                const bool seen = m1 & (qx >= ax0) & (qx < ax0 + awx) & (qy >= ay0) & (qy < ay0 + awy);
                const bool valid = act[h] & (sm < Nb) & (sl < P2) & (slx < wx) & (sly < wy) & !seen;
                if (valid) {
                    const int wn = h * Nb + sm;
                    const unsigned toff = (unsigned)(wn * C * cs1 + sly * L.tp1 + slx);
                    const bool self = (unsigned)qx == rx && (unsigned)qy == ry;
                    s_d[h * 64 + lane] = __uint_as_float(toff | (self ? 0x80000000u : 0u));
                    s_i[h * 64 + lane] = (unsigned)(m1 ? wb[4] : wa[4]) * whc + qy * W + qx;
                }
                bal[h] = __ballot(valid);
                if (valid) lst[(h ? __popcll(bal[0]) : 0) + lanes_below(bal[h], lane)] = (unsigned)(h * 64 + lane);
            }
            wave_sync();
            const int nv = __popcll(bal[0]) + __popcll(bal[1]);
            for (int e0 = 0; e0 < nv; e0 += 64) {
                if (e0 + lane < nv) {
                    const unsigned sl = lst[e0 + lane];
                    const unsigned tb = __float_as_uint(s_d[sl]);
                    const float sv = ssd(tile + (tb & 0x7FFFFFFFu), L.tp1, cs1);
                    s_d[sl] = sv / norm - ((tb >> 31) ? a.d : 0.f);
                }
            }
            wave_sync();
            for (int ch = 0; ch < 2; ch++) {
                if (!act[ch]) continue;
                // The reference takes the top Nb of each window, then the top
                // Nb of those picks.  When each of the chain's Nb smallest
                // keys is held by one slot only, that equals the Nb smallest
                // over all of the chain's slots (the window ranks cannot
                // matter), found with Nb wave minimums; a repeated key falls
                // through to the exact selection below.
                {
                    const unsigned MX = 0xFFFFFFFFu;
                    unsigned key = ((bal[ch] >> lane) & 1ull) ? fkey(s_d[ch * 64 + lane]) : MX;
                    const unsigned k1 = wave_min_u32(key);
                    const unsigned long long e1 = __ballot(key == k1);
                    const int p1 = __ffsll((long long)e1) - 1;
                    bool unique = __popcll(e1) == 1;
                    int p2 = p1;
                    if (Nb == 2) {
                        if (lane == p1) key = MX;
                        const unsigned k2 = wave_min_u32(key);
                        const unsigned long long e2 = __ballot(key == k2);
                        unique = unique && __popcll(e2) == 1;
                        p2 = __ffsll((long long)e2) - 1;
                    }
                    if (unique) {
                        n_eval += __popcll(bal[ch]);
                        const int at = Nb * (1 + (ch == 0 ? f : nfw + f));
                        if (lane < Nb) {
                            const int sl = ch * 64 + (lane == 0 ? p1 : p2);
                            const uint32_t iv = s_i[sl];
                            pp[(ch == 0 ? 8 : 0) + lane] = iv;
                            b_d[at + lane] = s_d[sl];
                            b_i[at + lane] = iv;
                        }
                        wave_sync();
                        continue;
                    }
                }
                // top Nb of each window's list, appended in window order
                int nfb = 0;
                for (int m = 0; m < Nb; m++) {
                    const int sb = ch * 64 + m * P2p;
                    const unsigned long long wmask = P2p >= 64 ? ~0ull : ((1ull << P2p) - 1ull);
                    const unsigned long long bm = (bal[ch] >> (m * P2p)) & wmask;
                    const bool v = (bm >> lane) & 1ull;
                    const int mm = __popcll(bm);
                    n_eval += mm;
                    const int nb = min(Nb, mm);
                    wave_select_sparse(s_d + sb, s_i + sb, v, bm, nb, o_d, o_i, c_d, c_i, scr, lane);
                    if (lane < nb) {
                        f_d[nfb + lane] = o_d[lane];
                        f_i[nfb + lane] = o_i[lane];
                    }
                    nfb += nb;
                    wave_sync();
                }
                wave_select(f_d, f_i, nfb, Nb, o_d, o_i, scr, lane);
                const int at = Nb * (1 + (ch == 0 ? f : nfw + f));
                if (lane < Nb) {
                    pp[(ch == 0 ? 8 : 0) + lane] = o_i[lane];
                    b_d[at + lane] = o_d[lane];
                    b_i[at + lane] = o_i[lane];
                }
                wave_sync();
            }
            if (f + 1 < nst) windows(f + 1);
        }
    }
    for (int f = 0; f < (piped ? 0 : nst); f++) {
        const bool act[2] = {f < nfw, f < nbw};  // [0] forward, [1] backward
        // windows: flow projection of each temp match (x first, then y at
        // the updated x), clamped Npr x Npr box in the next frame
        if (lane < 2 * Nb) {
            const int ch = lane / Nb, m = lane - ch * Nb;
            if (act[ch]) {
                const int dir = ch == 0 ? 1 : -1;
                const float* fl = ch == 0 ? a.ff : a.bf;
                const unsigned mi = pp[(ch == 0 ? 8 : 0) + m];
                const unsigned pt = div32(mi, dWHC);
                unsigned py = div32(mi - pt * whc, dW), px = mi - pt * whc - py * W;
                const unsigned ft = ch == 0 ? pt : pt - 1;
                const unsigned hk = K / 2;
                const float fx = fl ? fl[ft * 2 * wh + (py + hk) * W + px + hk] : 0.f;
                px = (unsigned)min(max((int)((float)px + roundf(fx)), 0), (int)(W - K));
                const float fy = fl ? fl[ft * 2 * wh + wh + (py + hk) * W + px + hk] : 0.f;
                py = (unsigned)min(max((int)((float)py + roundf(fy)), 0), (int)(H - K));
                const int x0 = max(0, (int)px - hs), y0 = max(0, (int)py - hs);
                win[lane * 5 + 0] = x0;
                win[lane * 5 + 1] = y0;
                win[lane * 5 + 2] = min((int)W - K, (int)px + hs) - x0 + 1;
                win[lane * 5 + 3] = min((int)H - K, (int)py + hs) - y0 + 1;
                win[lane * 5 + 4] = (int)(pt + dir);
            }
        }
        wave_sync();
        for (int wn = 0; wn < 2 * Nb; wn++)
            if (act[wn / Nb])
                load_tile(tile + wn * C * cs1, L.tp1, cs1, win[wn * 5], win[wn * 5 + 1], win[wn * 5 + 2],
                          win[wn * 5 + 3], (unsigned)win[wn * 5 + 4]);
        wave_sync();
        // slots: chain ch, window m, candidate l -> (ch * Nb + m) * P2p + l;
        // lane evaluates slots b0 + lane and b0 + 64 + lane together
        const int nslots = 2 * Nb * P2p;
        for (int b0 = 0; b0 < nslots; b0 += 128) {
            bool valid[2] = {false, false};
            int toff[2] = {0, 0}, qxs[2] = {0, 0}, qys[2] = {0, 0};
            unsigned tfs[2] = {0, 0};
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int s = b0 + 64 * h + lane;
                if (s < nslots) {
                    const int wn = s >> lg2, l = s & (P2p - 1), ch = wn >= Nb ? 1 : 0, m = wn - ch * Nb;
                    if (act[ch]) {
                        const int x0 = win[wn * 5], y0 = win[wn * 5 + 1], wx = win[wn * 5 + 2], wy = win[wn * 5 + 3];
                        if (l < wx * wy) {
                            const int ly = small_div(l, wx), lx = l - ly * wx;
                            const int qx = x0 + lx, qy = y0 + ly;
                            // already seen = inside an earlier window of this frame
                            bool seen = false;
                            for (int m2 = 0; m2 < m; m2++) {
                                const int* w2 = win + (ch * Nb + m2) * 5;
                                seen |= qx >= w2[0] && qx < w2[0] + w2[2] && qy >= w2[1] && qy < w2[1] + w2[3];
                            }
                            valid[h] = !seen;
                            toff[h] = wn * C * cs1 + ly * L.tp1 + lx;
                            qxs[h] = qx;
                            qys[h] = qy;
                            tfs[h] = (unsigned)win[wn * 5 + 4];
                        }
                    }
                }
            }
            if (valid[0] || valid[1]) {
                float sa, sb;
                ssd2(tile + (valid[0] ? toff[0] : toff[1]), tile + (valid[1] ? toff[1] : toff[0]), L.tp1, cs1, sa, sb);
                const float sv[2] = {sa, sb};
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int s = b0 + 64 * h + lane;
                    if (valid[h]) {
                        s_d[s] = sv[h] / norm - (((unsigned)qxs[h] == rx && (unsigned)qys[h] == ry) ? a.d : 0.f);
                        s_i[s] = tfs[h] * whc + qys[h] * W + qxs[h];
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int s = b0 + 64 * h + lane;
                if (s < nslots) s_v[s] = valid[h];
            }
        }
        wave_sync();
        for (int ch = 0; ch < 2; ch++) {
            if (!act[ch]) continue;
            // top Nb of each window's list, appended in window order
            int nfb = 0;
            for (int m = 0; m < Nb; m++) {
                const int sb = (ch * Nb + m) * P2p;
                const bool v = lane < P2p && s_v[sb + lane];
                const unsigned long long bal = __ballot(v);
                const int mm = __popcll(bal);
                n_eval += mm;
                const int nb = min(Nb, mm);
                // candidates in window raster order; the list is only
                // compacted when the exact heap replay is needed
                wave_select_sparse(s_d + sb, s_i + sb, v, bal, nb, o_d, o_i, c_d, c_i, scr, lane);
                if (lane < nb) {
                    f_d[nfb + lane] = o_d[lane];
                    f_i[nfb + lane] = o_i[lane];
                }
                nfb += nb;
                wave_sync();
            }
            // each frame contributes exactly Nb picks (window 0 alone holds
            // Nb candidates), so the slot of frame f in `best` is fixed
            wave_select(f_d, f_i, nfb, Nb, o_d, o_i, scr, lane);
            const int at = Nb * (1 + (ch == 0 ? f : nfw + f));
            if (lane < Nb) {
                pp[(ch == 0 ? 8 : 0) + lane] = o_i[lane];
                b_d[at + lane] = o_d[lane];
                b_i[at + lane] = o_i[lane];
            }
            wave_sync();
        }
    }
    nbest = Nb * (1 + nfw + nbw);

    // -- N best overall, tau cut, power-of-two group size
    //    (src/vbm3d.cpp:1169-1184)
    const int nsim = min(a.N, nbest);
    wave_select(b_d, b_i, nbest, nsim, o_d, o_i, scr, lane);
    int it = nsim - 1;
    while (o_d[it] > a.tau && it > 0) it--;
    unsigned nsx = 1;
    while (nsx * 2 <= (unsigned)(it + 1)) nsx *= 2;
    if (lane < a.N) {
        a.idx_out[(size_t)slot * a.N + lane] = lane < nsim ? o_i[lane] : 0xFFFFFFFFu;
        if (a.dist_out)
            a.dist_out[(size_t)slot * a.N + lane] = lane < nsim ? o_d[lane] : __builtin_nanf("");
    }
    if (lane == 0) {
        a.nsx_out[slot] = nsx;
        if (a.nsim_out) a.nsim_out[slot] = nsim;
        if (a.count_out) {
            a.count_out[2 * (r - a.r_begin)] = (uint32_t)n_eval;
            a.count_out[2 * (r - a.r_begin) + 1] = nsx;
        }
    }
}

// Occupancy floor per instantiation (waves per SIMD the register
// allocator must leave room for), measured: with three channels the kernel
// runs fastest held to 5 waves (96 registers, a few spills; 4 and 6 are
// slower), with one channel unconstrained.
template <int K, int C>
constexpr int bm_waves_floor() { return C == 3 ? 5 : 1; }

template <int K, int C, int SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(bm_waves_floor<K, C>()))) void k_bm(BMArgs a) {
    extern __shared__ float smem[];
    const BMLay L = SH == 1 ? bm_layout(K, C, 7, 5, 2, a.Nf) : bm_layout(K, C, a.Ns, a.Npr, a.Nb, a.Nf);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const long r = a.r_begin + (long)blockIdx.x * nw + w;
    if (r >= a.r_end) return;  // wave-uniform
    bm_one<K, C, SH>(a, r, smem + (size_t)w * L.total, lane);
}

template <int K, int C, int SH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(bm_waves_floor<K, C>()))) void k_bm_multi(const BMArgs* __restrict__ jobs, WorkQueue q, int words) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const long g = uniform_long((long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (g >= q.total) return;
    const int j = queue_job(q, g);
    typedef const __attribute__((address_space(4))) BMArgs cBM;
    cBM& a = ((cBM*)jobs)[j];
    bm_one<K, C, SH>(a, a.r_begin + (g - uniform_long(q.item_begin[j])), smem + (size_t)(threadIdx.x >> 6) * words, lane);
}

// Persistent form over a batch: every wave takes `grab` items at a time
// from the queue counter until the queue is empty; item -> (job, patch).
template <int K, int C, int SH>
__global__ __launch_bounds__(256) void k_bm_queue(const BMArgs* __restrict__ jobs, WorkQueue q, int words) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    float* base = smem + (size_t)(threadIdx.x >> 6) * words;
    typedef const __attribute__((address_space(4))) BMArgs cBM;
    int j = 0;
    long jb = 0, je = 0;
    long it0 = queue_take(q, lane);
    while (it0 < q.total) {
        const unsigned long long nxt = queue_issue(q, lane);
        const long it1 = min(q.total, it0 + (long)q.grab);
        for (long it = it0; it < it1; it++) {
            queue_seek(q, it, j, jb, je);
            cBM& a = ((cBM*)jobs)[j];  // scalar loads of the job
            bm_one<K, C, SH>(a, a.r_begin + (it - jb), base, lane);
        }
        it0 = queue_value(nxt);
    }
}

// ---------------------------------------------------------------------
// Group filtering.
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

// One 1D pass of REDFT10 on K values: row v of the matrix is even (v
// even) or odd (v odd) about the centre, so each output takes the folded
// inputs x[q] + x[K-1-q] or x[q] - x[K-1-q], q < K/2, plus the centre
// sample for odd K and even v; sums in q order, as the oracle's
// dct_pass_fwd.  About half the multiplications of the matrix product.
template <int K, class TAB>
__device__ __forceinline__ void dct_fwd_pass(const float* x, float* y, const TAB* F) {
    constexpr int H = K / 2;
    float s[H], d[H];
#pragma unroll
    for (int q = 0; q < H; q++) {
        s[q] = x[q] + x[K - 1 - q];
        d[q] = x[q] - x[K - 1 - q];
    }
#pragma unroll
    for (int v = 0; v < K; v++) {
        float m = 0.f;
        if ((v & 1) == 0) {
#pragma unroll
            for (int q = 0; q < H; q++) m = m + s[q] * F[v * K + q];
            if (K & 1) m = m + x[H] * F[v * K + H];
        } else {
#pragma unroll
            for (int q = 0; q < H; q++) m = m + d[q] * F[v * K + q];
        }
        y[v] = m;
    }
}

// Output v of dct_fwd_pass on x[0], x[stride], ... (the same folding and
// the same order of operations, so the two agree bit for bit).
template <int K, class TAB>
__device__ __forceinline__ float dct_fwd_one(const float* x, int stride, int v, const TAB* F) {
    constexpr int H = K / 2;
    float m = 0.f;
    if ((v & 1) == 0) {
#pragma unroll
        for (int q = 0; q < H; q++) m = m + (x[q * stride] + x[(K - 1 - q) * stride]) * F[v * K + q];
        if (K & 1) m = m + x[H * stride] * F[v * K + H];
    } else {
#pragma unroll
        for (int q = 0; q < H; q++) m = m + (x[q * stride] - x[(K - 1 - q) * stride]) * F[v * K + q];
    }
    return m;
}

// One 1D pass of REDFT01 (I[o][i] = table[o*K + i]): rows o and K-1-o
// differ only in the sign of the odd-i terms, so E (even i) and O (odd i)
// give y[o] = E + O and y[K-1-o] = E - O; the centre row of odd K is E.
// Sums in i order, as the oracle's dct_pass_inv.
template <int K, class TAB>
__device__ __forceinline__ void dct_inv_pass(const float* z, float* y, const TAB* I) {
    constexpr int H = K / 2;
#pragma unroll
    for (int o = 0; o < H; o++) {
        float e = 0.f, od = 0.f;
#pragma unroll
        for (int i = 0; i < K; i += 2) e = e + z[i] * I[o * K + i];
#pragma unroll
        for (int i = 1; i < K; i += 2) od = od + z[i] * I[o * K + i];
        y[o] = e + od;
        y[K - 1 - o] = e - od;
    }
    if (K & 1) {
        float e = 0.f;
#pragma unroll
        for (int i = 0; i < K; i += 2) e = e + z[i] * I[H * K + i];
        y[H] = e;
    }
}

// Transposition buffer: patch n of a lane batch is a K x (K+1) tile at
// n*xstride(K) (row pitch K+1).  The strides keep both the row-wise write and
// the column-wise read of the 64 lanes (n, r) = n*K + r bank-conflict free
// at K = 8 and within 2-way at K = 7.
__host__ __device__ constexpr int xstride(int K) { return (K & 1) ? K * (K + 1) + 1 : K * (K + 1); }

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
    // row rr of member nl of the first lane batch: its base offset, and the
    // row of the next (channel, source) pair fetched ahead of its use
    const bool act0 = lact && nl < nsx;
    unsigned mb = 0;
    if (act0) {
        const unsigned mi = mem[nl];
        const unsigned pt = div32(mi, dWHC);
        const unsigned py = div32(mi - pt * whc, dW), px = mi - pt * whc - py * W;
        mb = pt * whc + (py + rr) * W + px;
    }
    float xn[K];
    auto fetch = [&](int c, int g) {
        if (act0) {
            const float* row = (g == 0 ? a.noisy : a.basic) + mb + c * wh;
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
                        const unsigned mi = mem[n];
                        const unsigned pt = div32(mi, dWHC);
                        const unsigned py = div32(mi - pt * whc, dW), px = mi - pt * whc - py * W;
                        const float* row = src + pt * whc + c * wh + (py + rr) * W + px;
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
        //      270-321; dct_2d_inv src/vbm3d.cpp:1678-1713), store [n][pixel]
        const float* est = STEP == 2 ? eg : zg;
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
                        dstg[(size_t)n * K2 + u * K + rr] = m;
                    }
                } else {
                    float y[K];
                    dct_inv_pass<K>(col, y, TI);
#pragma unroll
                    for (int u = 0; u < K; u++) dstg[(size_t)n * K2 + u * K + rr] = dct_out * y[u];
                }
            }
            wave_sync();
        }
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
template <int C, bool RAW, int NM, class A>
__device__ __forceinline__ void agg_tile(const A& a, const int bx, const int by, const unsigned bz, const int nl) {
    typedef const __attribute__((address_space(4))) float cfloat;
    constexpr int TX = kAggTileX, TY = kAggTileY, NT = kAggThreads;
    constexpr int QS = 8;             // a wavefront's block edge
    constexpr int NWV = NT / 64, WX = TX / QS;
    static_assert(TX % QS == 0 && TY % QS == 0 && NT % 64 == 0, "tile of 8x8 blocks");
    // the ordered list of one NT-patch chunk: at most nl >= N members each
    extern __shared__ uint32_t agg_lists[];
    uint32_t* L_sn = agg_lists;
    uint32_t* L_pos = agg_lists + NT * nl;
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
    const int rlo = y0 - a.R - (K - 1), rhi = y0 + TY - 1 + a.R;
    const int clo = x0 - a.R - (K - 1), chi = x0 + TX - 1 + a.R;
    // a member at q = to*whc + y*W + x overlaps the tile's rows iff
    // q - to*whc lies in [ylo*W, yhi*W)
    const unsigned tbase = to * whc;
    const Div32 dW = {a.divW.m, a.divW.s1, a.divW.s2};
    const unsigned rmin = (unsigned)max(0, y0 - K + 1) * W, rmax = (unsigned)min((int)a.H, y0 + TY) * W;

    for (int ci = 0; ci < a.ncrop; ci++) {
        const int* meta = a.crop_meta + ci * 5;
        const int off = meta[0], nrows = meta[1], ncols = meta[2];
        const uint16_t* rows = a.grid_pos + meta[3];
        const uint16_t* cols = a.grid_pos + meta[4];
        const int ia = lower_bound_u16(rows, nrows, rlo), ib = lower_bound_u16(rows, nrows, rhi + 1);
        const int ja = lower_bound_u16(cols, ncols, clo), jb = lower_bound_u16(cols, ncols, chi + 1);
        const int nr = ib - ia, nc = jb - ja, ntr = trhi - trlo + 1;
        const long total = (nr > 0 && nc > 0 && ntr > 0) ? (long)ntr * nr * nc : 0;
        const Div32 dnc = make_div32(nc > 0 ? nc : 1), dnr = make_div32(nr > 0 ? nr : 1);
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
                const unsigned rest = div32(e, dnc), jj = e - rest * nc;
                const unsigned tq = div32(rest, dnr), ii = rest - tq * nr;
                const unsigned tr = trlo + tq;
                slot_n = (tr % a.ring) * a.nrf + off + (ia + ii) * ncols + (ja + jj);
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
                    L_sn[o] = slot * kMaxN + n;
                    L_pos[o] = qx | (qy << 16);
                    o++;
                }
            __syncthreads();
            if (cb + NT < total) fetch(cb + NT);
            // this wave: the list entries overlapping its 8x8 block, in order
            for (int j0 = 0; j0 < tot; j0 += 64) {
                bool ov = false;
                if (j0 + lane < tot) {
                    const uint32_t pos = L_pos[j0 + lane];
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
                    // first word; only the sums below are masked
                    float ev[U][C], kai[U], w[U][C];
                    bool on[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int jv = jj[u] >= 0 ? jj[u] : jj[0];
                        const uint32_t pos = __builtin_amdgcn_readfirstlane(L_pos[jv]);
                        const uint32_t sn = __builtin_amdgcn_readfirstlane(L_sn[jv]);
                        const uint32_t sl = sn / kMaxN, n = sn % kMaxN;
                        const cfloat* wb = (const cfloat*)(a.wgt + (size_t)sl * C);
#pragma unroll
                        for (int c = 0; c < C; c++) w[u][c] = RAW ? 0.f : wb[c];
                        const int dx = px - (int)(pos & 0xFFFFu), dy = py - (int)(pos >> 16);
                        on[u] = (jj[u] >= 0) & inside & ((unsigned)dx < (unsigned)K) & ((unsigned)dy < (unsigned)K);
                        const unsigned kk = on[u] ? (unsigned)(dy * K + dx) : 0u;
                        kai[u] = s_kai[kk];
                        // wave-uniform base, 32-bit lane byte offset
                        const char* eb = (const char*)(a.est + ((size_t)sl * C * N + n) * K2);
#pragma unroll
                        for (int c = 0; c < C; c++)
                            ev[u][c] = *(const float*)(eb + (size_t)c * N * K2 * 4 + kk * 4u);
                    }
                    // keeps the loads above from being sunk to their sums
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
            if (a.ncrop == 1) {
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
    agg_tile<C, RAW, NM>(a, blockIdx.x, blockIdx.y, blockIdx.z, a.N);
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
    agg_tile<C, false, NM>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
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
        agg_tile<C, false, NM>(a, (int)(rem % tx), (int)(rem / tx), (unsigned)bz, nl);
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
        m = max(m, (unsigned)fabsf(roundf(fl[i])));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace

// ---------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------
template <int K, int C>
static hipError_t bm_k(const BMArgs& a, hipStream_t s) {
    const long n = a.r_end - a.r_begin;
    if (n <= 0) return hipSuccess;
    const BMLay L = bm_layout(K, C, a.Ns, a.Npr, a.Nb, a.Nf);
    const size_t per_wave = (size_t)L.total * 4;
    int waves = 4;
    while (waves > 1 && per_wave * waves > 64 * 1024) waves--;
    if (per_wave > 64 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + waves - 1) / waves)), block(64 * waves);
    if (bm_shape(a) == 1) hipLaunchKernelGGL((k_bm<K, C, 1>), grid, block, per_wave * waves, s, a);
    else hipLaunchKernelGGL((k_bm<K, C, 0>), grid, block, per_wave * waves, s, a);
    return hipGetLastError();
}

hipError_t launch_bm(int K, const BMArgs& a, hipStream_t s) {
#define VB_BM(KK)                                           \
    case KK:                                                \
        return a.C == 1 ? bm_k<KK, 1>(a, s) : bm_k<KK, 3>(a, s);
    switch (K) {
        VB_BM(4) VB_BM(5) VB_BM(6) VB_BM(7) VB_BM(8)
        default: return hipErrorInvalidValue;
    }
#undef VB_BM
}

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

// ---- persistent launches over a batch (jobs and queue arrays in device memory)
template <int K, int C>
static hipError_t bm_queue_k(const BMArgs* jobs, const WorkQueue& q, int words, int waves, int blocks, int shape,
                             hipStream_t s) {
    const size_t lds = (size_t)words * 4 * waves;
    if (shape == 1) hipLaunchKernelGGL((k_bm_queue<K, C, 1>), dim3(blocks), dim3(64 * waves), lds, s, jobs, q, words);
    else hipLaunchKernelGGL((k_bm_queue<K, C, 0>), dim3(blocks), dim3(64 * waves), lds, s, jobs, q, words);
    return hipGetLastError();
}

template <int K, int C>
static hipError_t bm_multi_k(const BMArgs* jobs, const WorkQueue& q, int words, int waves, int shape,
                             hipStream_t s) {
    const dim3 grid((unsigned)((q.total + waves - 1) / waves)), block(64 * waves);
    const size_t lds = (size_t)words * 4 * waves;
    if (shape == 1) hipLaunchKernelGGL((k_bm_multi<K, C, 1>), grid, block, lds, s, jobs, q, words);
    else hipLaunchKernelGGL((k_bm_multi<K, C, 0>), grid, block, lds, s, jobs, q, words);
    return hipGetLastError();
}

hipError_t launch_bm_multi(int K, int C, const BMArgs* jobs, const WorkQueue& q, int words, int waves, int shape,
                           hipStream_t s) {
    if (q.total <= 0) return hipSuccess;
    if ((size_t)words * 4 * waves > 64 * 1024 || waves < 1 || waves > 4) return hipErrorInvalidValue;
#define VB_BMM(KK)                                                                               \
    case KK:                                                                                     \
        return C == 1 ? bm_multi_k<KK, 1>(jobs, q, words, waves, shape, s)                       \
                      : bm_multi_k<KK, 3>(jobs, q, words, waves, shape, s);
    switch (K) {
        VB_BMM(4) VB_BMM(5) VB_BMM(6) VB_BMM(7) VB_BMM(8)
        default: return hipErrorInvalidValue;
    }
#undef VB_BMM
}

hipError_t launch_bm_queue(int K, int C, const BMArgs* jobs, const WorkQueue& q, int words, int waves,
                           int blocks, int shape, hipStream_t s) {
    if (q.total <= 0 || blocks <= 0) return hipSuccess;
    if ((size_t)words * 4 * waves > 64 * 1024 || waves < 1 || waves > 4) return hipErrorInvalidValue;
#define VB_BMQ(KK)                                                                               \
    case KK:                                                                                     \
        return C == 1 ? bm_queue_k<KK, 1>(jobs, q, words, waves, blocks, shape, s)               \
                      : bm_queue_k<KK, 3>(jobs, q, words, waves, blocks, shape, s);
    switch (K) {
        VB_BMQ(4) VB_BMQ(5) VB_BMQ(6) VB_BMQ(7) VB_BMQ(8)
        default: return hipErrorInvalidValue;
    }
#undef VB_BMQ
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

hipError_t launch_aggregate_multi(int C, const AArgs* jobs, const WorkQueue& q, int maxN, hipStream_t s) {
    if (q.total <= 0) return hipSuccess;
    if (maxN < 1 || maxN > kMaxN || q.total > 0x7FFFFFFFL) return hipErrorInvalidValue;
    const size_t lds = (size_t)2 * kAggThreads * maxN * sizeof(uint32_t);
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
    const size_t lds = (size_t)2 * kAggThreads * maxN * sizeof(uint32_t);
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

int bm_lds_words(int K, int C, int Ns, int Npr, int Nb, int Nf) { return bm_layout(K, C, Ns, Npr, Nb, Nf).total; }
int filter_words(int K, int C, int N, int step) { return filter_lds_words(K, C, N, step); }

hipError_t launch_aggregate(int C, const AArgs& a, unsigned frames, hipStream_t s) {
    if (frames == 0) return hipSuccess;
    dim3 grid((a.W + kAggTileX - 1) / kAggTileX, (a.H + kAggTileY - 1) / kAggTileY, frames);
    if (a.N < 1 || a.N > kMaxN) return hipErrorInvalidValue;
    const size_t lds = (size_t)2 * kAggThreads * a.N * sizeof(uint32_t);  // the chunk's ordered list
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
