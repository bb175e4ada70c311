// kernels_bm.hip -- CDNA4 (gfx950) predictive block matching of the VBM3D
// collaborative filter: one wavefront per reference patch; lanes evaluate
// the candidates of a search window in parallel, selection is a wave-wide
// arg-min with a tie check (exact heap-select fallback in one lane when
// distances tie).
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
            // one scalar base per channel, kept opaque so that the copies
            // take the scalar-base form (no 64-bit vector address per copy)
            const char* cb[C];
#pragma unroll
            for (int c = 0; c < C; c++) {
                unsigned long long b = (unsigned long long)(const char*)(vid + fb + c * wh);
                unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
                asm volatile("" : "+s"(lo), "+s"(hi));
                cb[c] = (const char*)(((unsigned long long)hi << 32) | lo);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k * 64 >= cs) break;
                const int e = k * 64 + lane;
                const int er = small_div(e, L.tp0), ec = e - er * L.tp0;
                if (er < ah && ec < aw) {
                    const unsigned go = ((unsigned)er * W + (unsigned)ec) * 4u;
#pragma unroll
                    for (int c = 0; c < C; c++) lds_dma(cb[c] + go, tl0 + (unsigned)((c * cs + k * 64) * 4));
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
        // Npr x Npr box in the next frame.  Lane wn < 2 Nb keeps window wn
        // in registers (x0, y0, width, height, frame), which the copies and
        // the slot list read with readlane: no LDS round trip.
        int wx0 = 0, wy0 = 0, wwx = 0, wwy = 0, wfr = 0;
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
                    const float fx = fl ? fl[(size_t)ft * 2 * wh + (py + hk) * W + px + hk] : 0.f;
                    px = (unsigned)min(max((int)((float)px + roundf(fx)), 0), (int)(W - K));
                    const float fy = fl ? fl[(size_t)ft * 2 * wh + wh + (py + hk) * W + px + hk] : 0.f;
                    py = (unsigned)min(max((int)((float)py + roundf(fy)), 0), (int)(H - K));
                    const int x0 = max(0, (int)px - hs), y0 = max(0, (int)py - hs);
                    wx0 = x0;
                    wy0 = y0;
                    wwx = min((int)W - K, (int)px + hs) - x0 + 1;
                    wwy = min((int)H - K, (int)py + hs) - y0 + 1;
                    wfr = (int)(pt + dir);
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
                        const int aw = __builtin_amdgcn_readlane(wwx, wn) + K - 1;
                        const int ah = __builtin_amdgcn_readlane(wwy, wn) + K - 1;
                        const unsigned fb = (unsigned)__builtin_amdgcn_readlane(wfr, wn) * whc +
                                            (unsigned)__builtin_amdgcn_readlane(wy0, wn) * W +
                                            (unsigned)__builtin_amdgcn_readlane(wx0, wn);
                        // one scalar base per channel, kept opaque so that
                        // the copies take the scalar-base form (no 64-bit
                        // vector address arithmetic per copy)
                        const char* cb[C];
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            unsigned long long b = (unsigned long long)(const char*)(vid + fb + c * wh);
                            unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
                            asm volatile("" : "+s"(lo), "+s"(hi));
                            cb[c] = (const char*)(((unsigned long long)hi << 32) | lo);
                        }
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            if (k * 64 >= ne) break;
                            if (er[k] < ah && ec[k] < aw) {
                                // a 32-bit lane offset at the copy (not a
                                // hoisted 64-bit one) keeps the scalar-base form
                                unsigned go = goff[k];
                                asm volatile("" : "+v"(go));
#pragma unroll
                                for (int c = 0; c < C; c++)
                                    lds_dma(cb[c] + go, tl + (unsigned)((wn * L.ws + c * cs1 + k * 64) * 4));
                            }
                        }
                    }
                // no wait here: the copies land while the slot list below,
                // which does not read the tile, is formed
            }
            // slot lane of chain h; a valid slot's list entry holds its slot
            // (bits 16..22), the reference's own position (bit 15) and its
            // tile offset (bits 0..14; at most 4 windows of 12 x 13 x 3
            // words)
            unsigned long long bal[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                // the chain's windows 0 and 1 (window 0 again when Nb = 1),
                // wave-uniform; lanes of window 1 select its values, and a
                // candidate inside window 0 was seen
                const int ia = h * Nb, ib = h * Nb + (Nb > 1 ? 1 : 0);
                const int ax0 = __builtin_amdgcn_readlane(wx0, ia), ay0 = __builtin_amdgcn_readlane(wy0, ia);
                const int awx = __builtin_amdgcn_readlane(wwx, ia), awy = __builtin_amdgcn_readlane(wwy, ia);
                const int bx0 = __builtin_amdgcn_readlane(wx0, ib), by0 = __builtin_amdgcn_readlane(wy0, ib);
                const int bwx = __builtin_amdgcn_readlane(wwx, ib), bwy = __builtin_amdgcn_readlane(wwy, ib);
                const bool m1 = sm == 1;
                const int x0 = m1 ? bx0 : ax0, y0 = m1 ? by0 : ay0;
                const int wx = m1 ? bwx : awx, wy = m1 ? bwy : awy;
                const int qx = x0 + slx, qy = y0 + sly;
                const bool seen = m1 & (qx >= ax0) & (qx < ax0 + awx) & (qy >= ay0) & (qy < ay0 + awy);
                const bool valid = act[h] & (sm < Nb) & (sl < P2) & (slx < wx) & (sly < wy) & !seen;
                unsigned ent = 0;
                if (valid) {
                    const int wn = h * Nb + sm;
                    const unsigned toff = (unsigned)(wn * L.ws + sly * L.tp1 + slx);
                    const bool self = (unsigned)qx == rx && (unsigned)qy == ry;
                    ent = ((unsigned)(h * 64 + lane) << 16) | (self ? 0x8000u : 0u) | toff;
                    const int fr = m1 ? __builtin_amdgcn_readlane(wfr, ib) : __builtin_amdgcn_readlane(wfr, ia);
                    s_i[h * 64 + lane] = (unsigned)fr * whc + qy * W + qx;
                }
                bal[h] = __ballot(valid);
                if (valid) lst[(h ? __popcll(bal[0]) : 0) + lanes_below(bal[h], lane)] = ent;
            }
            __builtin_amdgcn_s_waitcnt(0);  // the window copies
            wave_sync();
            const int nv = __popcll(bal[0]) + __popcll(bal[1]);
            for (int e0 = 0; e0 < nv; e0 += 64) {
                if (e0 + lane < nv) {
                    const unsigned e = lst[e0 + lane];
                    const float sv = ssd(tile + (e & 0x7FFFu), L.tp1, cs1);
                    s_d[e >> 16] = sv / norm - ((e & 0x8000u) ? a.d : 0.f);
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
                const float fx = fl ? fl[(size_t)ft * 2 * wh + (py + hk) * W + px + hk] : 0.f;
                px = (unsigned)min(max((int)((float)px + roundf(fx)), 0), (int)(W - K));
                const float fy = fl ? fl[(size_t)ft * 2 * wh + wh + (py + hk) * W + px + hk] : 0.f;
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
                load_tile(tile + wn * L.ws, L.tp1, cs1, win[wn * 5], win[wn * 5 + 1], win[wn * 5 + 2],
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
                            toff[h] = wn * L.ws + ly * L.tp1 + lx;
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
    // the reference walks down from nsim - 1 while the distance exceeds tau
    // (stopping at 0): the highest pick that does not, found with one
    // ballot instead of a chain of dependent reads
    const unsigned long long keep = __ballot(lane < nsim && !(o_d[lane < nsim ? lane : 0] > a.tau));
    const int it = keep ? 63 - __clzll((long long)keep) : 0;
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
    unsigned lb = blockIdx.x;
    if (a.band) lb = band_order(lb, gridDim.x, max(1u, a.nrf / (unsigned)(8 * nw)));
    const long r = a.r_begin + (long)lb * nw + w;
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

int bm_lds_words(int K, int C, int Ns, int Npr, int Nb, int Nf) { return bm_layout(K, C, Ns, Npr, Nb, Nf).total; }

}  // namespace vb
