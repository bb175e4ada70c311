// kernels_gen.hip -- the general path of the VBM3D collaborative filter on
// CDNA4 (gfx950): every parameter set the reference accepts within the
// bounds of GenLimits (kernels.h), in particular the ones the fast kernels
// (kernels_bm/filter/agg.hip) leave out:
//   * temporal patches, kt > 1, with the temporal DCT along the slices
//     (src/vbm3d.cpp:1866-1956) and motion-compensated trajectories (mc,
//     patch_trajectory :378-425 with FLOAT_TRAJECTORIES);
//   * patch sizes up to 16 (the 12-point Kaiser window included);
//   * group sizes N that are not powers of two, and up to 32;
//   * search windows up to 15 x 15 and up to 16 matches per frame,
//     including predictive windows smaller than Nb (frames that yield
//     fewer than Nb picks keep older matches in the chain, :1140-1145).
// The arithmetic follows the oracle's operation order exactly, so results
// are bit-identical to it, as the fast kernels' are; the work layout is
// simpler (one wavefront per reference patch, pixels read through the
// caches, lane-0 heap selection for lists longer than a wavefront) and
// slower.
#include "kernels_common.h"

#pragma clang fp contract(off)

namespace vb {
namespace {

// Top-left corners of the kt slices of the patch at (px, py, pt)
// (patch_trajectory, src/vbm3d.cpp:378-425, FLOAT_TRAJECTORIES): with mc,
// the running float sum of the forward flow read at the current corner
// plus k/2, rounded half away from zero, clamped to the frame.
__device__ __forceinline__ void traj_step(const GenArgs& a, unsigned& px, unsigned& py, float& fx, float& fy,
                                          unsigned t) {
    const unsigned h = a.K / 2;
    const size_t fo = (size_t)t * 2 * a.wh + (size_t)(py + h) * a.W + px + h;
    fx = fx + (a.ff ? a.ff[fo] : 0.f);
    fy = fy + (a.ff ? a.ff[fo + a.wh] : 0.f);
    px = (unsigned)min((int)a.W - a.K, max(0, (int)roundf(fx)));
    py = (unsigned)min((int)a.H - a.K, max(0, (int)roundf(fy)));
}

// Exact partial selection (std::partial_sort order, ties included) of the
// `mid` smallest of list (ld, li)[0..m): the wave-parallel routine for
// lists that fit one wavefront, otherwise the heap select in lane 0.
__device__ void gen_select(const float* ld, const uint32_t* li, int m, int mid, float* od, uint32_t* oi,
                           VbEnt* scr, int lane) {
    if (m <= 64) {
        wave_select(ld, li, m, mid, od, oi, scr, lane);
        return;
    }
    if (lane == 0 && mid > 0) {
        for (int q = 0; q < m; q++) {
            scr[q].d = ld[q];
            scr[q].i = li[q];
        }
        vb_partial_sort(scr, m, mid);
        for (int q = 0; q < mid; q++) {
            od[q] = scr[q].d;
            oi[q] = scr[q].i;
        }
    }
    wave_sync();
}

// ---------------------------------------------------------------------
// Predictive block matching, general form (computeSimilarPatches /
// localSearch / patchDistance, src/vbm3d.cpp:930-1197).
// ---------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(128) void k_bm_gen(GenArgs a) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long r = a.r_begin + (long)blockIdx.x * (blockDim.x >> 6) + w;
    if (r >= a.r_end) return;
    const GenBMLay L = gen_bm_layout(a.Nb, a.Nf, a.Ns, a.Npr);
    float* base = smem + (size_t)w * L.total;
    int* win = (int*)(base + L.win);  // per window: x0, y0, wx, wy, frame
    float* wd = base + L.wd;
    uint32_t* wi = (uint32_t*)(base + L.wi);
    float* fd = base + L.fd;
    uint32_t* fi = (uint32_t*)(base + L.fi);
    float* bd = base + L.bd;
    uint32_t* bi = (uint32_t*)(base + L.bi);
    float* od = base + L.od;
    uint32_t* oi = (uint32_t*)(base + L.oi);
    uint32_t* chain = (uint32_t*)(base + L.chain);  // [0, 16) forward, [16, 32) backward
    int* rtr = (int*)(base + L.rtr);                 // reference trajectory, 2 per slice
    VbEnt* scr = (VbEnt*)(base + L.scr);

    const unsigned W = a.W, H = a.H, wh = a.wh, whc = a.whc;
    const int K = a.K, kt = a.kt, Nb = a.Nb;
    const float* __restrict__ vid = a.vid;
    unsigned rx, ry, rt, slot;
    if (a.ref_list) {
        const unsigned ridx = a.ref_list[r], rem = ridx % whc;
        rt = ridx / whc;
        ry = rem / W;
        rx = rem - ry * W;
        slot = (unsigned)(r - a.r_begin);
    } else {
        const unsigned t_ = (unsigned)(r / a.nrf), i_ = (unsigned)(r - (long)t_ * a.nrf);
        const unsigned xy = a.ref_xy[i_];
        rx = xy & 0xFFFFu;
        ry = xy >> 16;
        rt = t_;
        slot = (t_ % a.ring) * a.nrf + i_;
    }
    const float norm = (float)(unsigned)(K * K * kt * C);
    if (lane == 0) {
        unsigned px = rx, py = ry;
        float fx = (float)px, fy = (float)py;
        for (int ht = 0; ht < kt; ht++) {
            rtr[2 * ht] = (int)px;
            rtr[2 * ht + 1] = (int)py;
            if (a.mc && ht + 1 < kt) traj_step(a, px, py, fx, fy, rt + ht);
        }
    }
    wave_sync();

    int nwin = 0;
    long long n_eval = 0;
    // One localSearch (src/vbm3d.cpp:1013-1080): the window around (cx, cy)
    // in frame ct, its unseen candidates evaluated in raster order, then its
    // min(Nb, count) smallest appended to (dd, di) at *dn.
    auto local = [&](int cx, int cy, unsigned ct, int s, float* dd, uint32_t* di, int* dn) {
        const int h = (s - 1) / 2;
        const int x0 = max(0, cx - h), y0 = max(0, cy - h);
        const int wx = min((int)W - K, cx + h) - x0 + 1, wy = min((int)H - K, cy + h) - y0 + 1;
        const int me = nwin;
        if (lane == 0) {
            win[me * 5 + 0] = x0;
            win[me * 5 + 1] = y0;
            win[me * 5 + 2] = wx;
            win[me * 5 + 3] = wy;
            win[me * 5 + 4] = (int)ct;
        }
        nwin++;
        wave_sync();
        int m = 0;
        for (int b0 = 0; b0 < wx * wy; b0 += 64) {
            const int l = b0 + lane;
            bool valid = false;
            int qx = 0, qy = 0;
            if (l < wx * wy) {
                qy = y0 + l / wx;
                qx = x0 + l - (l / wx) * wx;
                valid = true;
                // already seen: inside an earlier window of the same frame
                for (int v = 0; v < me; v++) {
                    const int* e = win + v * 5;
                    if (e[4] == (int)ct && qx >= e[0] && qx < e[0] + e[2] && qy >= e[1] && qy < e[1] + e[3]) {
                        valid = false;
                        break;
                    }
                }
            }
            float dist = 0.f;
            if (valid) {
                unsigned px = (unsigned)qx, py = (unsigned)qy;
                float fx = (float)px, fy = (float)py;
                float acc = 0.f;
                for (int ht = 0; ht < kt; ht++) {
                    const float* rp = vid + (size_t)(rt + ht) * whc + (size_t)rtr[2 * ht + 1] * W + rtr[2 * ht];
                    const float* cp = vid + (size_t)(ct + ht) * whc + (size_t)py * W + px;
                    for (int c = 0; c < C; c++)
                        for (int y = 0; y < K; y++)
                            for (int x = 0; x < K; x++) {
                                const float dif = rp[c * wh + y * W + x] - cp[c * wh + y * W + x];
                                acc = acc + dif * dif;
                            }
                    if (a.mc && ht + 1 < kt) traj_step(a, px, py, fx, fy, ct + ht);
                }
                dist = acc / norm - (((unsigned)qx == rx && (unsigned)qy == ry) ? a.d : 0.f);
            }
            const unsigned long long bal = __ballot(valid);
            if (valid) {
                const int p = m + (int)lanes_below(bal, lane);
                wd[p] = dist;
                wi[p] = ct * whc + (unsigned)qy * W + (unsigned)qx;
            }
            m += __popcll(bal);
        }
        wave_sync();
        n_eval += m;
        const int nb = min(Nb, m);
        gen_select(wd, wi, m, nb, od, oi, scr, lane);
        if (lane < nb) {
            dd[*dn + lane] = od[lane];
            di[*dn + lane] = oi[lane];
        }
        *dn += nb;
        wave_sync();
    };

    // -- the reference frame (src/vbm3d.cpp:1114-1119)
    int nbest = 0;
    local((int)rx, (int)ry, rt, a.Ns, bd, bi, &nbest);
    if (lane < Nb) {
        chain[lane] = bi[lane];
        chain[16 + lane] = bi[lane];
    }
    wave_sync();
    // -- predictive frames, forward then backward (:1121-1167)
    const int nfw = (int)(min(rt + (unsigned)a.Nf, a.F - (unsigned)kt) - rt);
    const int nbw = (int)rt - max((int)rt - a.Nf, 0);
    for (int dir = 0; dir < 2; dir++) {
        const int nst = dir == 0 ? nfw : nbw;
        uint32_t* ch = chain + 16 * dir;
        const float* fl = dir == 0 ? a.ff : a.bf;
        for (int f = 0; f < nst; f++) {
            int nf = 0;
            for (int mm = 0; mm < Nb; mm++) {
                const unsigned mi = ch[mm];
                const unsigned pt = mi / whc;
                unsigned py = (mi - pt * whc) / W, px = mi - pt * whc - py * W;
                const unsigned ft = dir == 0 ? pt : pt - 1;
                const unsigned hk = K / 2;
                const float fx = fl ? fl[(size_t)ft * 2 * wh + (py + hk) * W + px + hk] : 0.f;
                px = (unsigned)min(max((int)((float)px + roundf(fx)), 0), (int)(W - K));
                const float fy = fl ? fl[(size_t)ft * 2 * wh + wh + (py + hk) * W + px + hk] : 0.f;
                py = (unsigned)min(max((int)((float)py + roundf(fy)), 0), (int)(H - K));
                local((int)px, (int)py, dir == 0 ? pt + 1 : pt - 1, a.Npr, fd, fi, &nf);
            }
            const int nb = min(Nb, nf);
            gen_select(fd, fi, nf, nb, od, oi, scr, lane);
            if (lane < nb) {
                ch[lane] = oi[lane];
                bd[nbest + lane] = od[lane];
                bi[nbest + lane] = oi[lane];
            }
            nbest += nb;
            wave_sync();
        }
    }
    // -- N best overall, tau cut, power-of-two group size (:1169-1184)
    const int nsim = min(a.N, nbest);
    gen_select(bd, bi, nbest, nsim, od, oi, scr, lane);
    int it = nsim - 1;
    while (od[it] > a.tau && it > 0) it--;
    unsigned nsx = 1;
    while (nsx * 2 <= (unsigned)(it + 1)) nsx *= 2;
    for (int n = lane; n < a.N; n += 64) {
        a.idx_out[(size_t)slot * a.N + n] = n < nsim ? oi[n] : 0xFFFFFFFFu;
        if (a.dist_out) a.dist_out[(size_t)slot * a.N + n] = n < nsim ? od[n] : __builtin_nanf("");
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

// ---------------------------------------------------------------------
// Group filtering, general form: 2D transforms per slice, temporal DCT,
// transform along the group, shrinkage, inverses (src/vbm3d.cpp:447-654,
// 676-928 with :1218-1386, 1405-1731, 1866-1956), in the oracle's
// operation order.  One wavefront per reference patch; lane = output
// coefficient (several per lane for large patches); the group lives in
// LDS as [k^2 kt][nSx] per channel.
// ---------------------------------------------------------------------
// 2D forward transform of the K x K patch in P (row-major) into out
// (coefficient layout of the oracle's bior_fwd / dct_fwd), T scratch.
__device__ void gen_fwd2d(const float* P, float* T, float* out, int K, bool bior, const TabsG& tb, int lane) {
    const int K2 = K * K, h = K / 2;
    if (bior) {
        for (int e = lane; e < 64; e += 64) {
            const int i = e >> 3, j = e & 7;
            float m = 0.f;
            for (int s = 0; s < 8; s++) m = m + P[i * 8 + s] * tb.biorF[j * 8 + s];
            T[e] = m;
        }
        wave_sync();
        for (int e = lane; e < 64; e += 64) {
            const int i = e >> 3, j = e & 7;
            float m = 0.f;
            for (int s = 0; s < 8; s++) m = m + T[s * 8 + i] * tb.biorF[j * 8 + s];
            out[e] = m;
        }
        wave_sync();
        return;
    }
    // rows: T[p][v] = REDFT10 of row p (parity-folded, q order)
    for (int e = lane; e < K2; e += 64) {
        const int p = e / K, v = e - p * K;
        const float* x = P + p * K;
        float m;
        if (h == 0) {
            m = x[0] * tb.dctF[v * K];  // K = 1: the centre sample alone
        } else if ((v & 1) == 0) {
            m = (x[0] + x[K - 1]) * tb.dctF[v * K];
            for (int q = 1; q < h; q++) m = m + (x[q] + x[K - 1 - q]) * tb.dctF[v * K + q];
            if (K & 1) m = m + x[h] * tb.dctF[v * K + h];
        } else {
            m = (x[0] - x[K - 1]) * tb.dctF[v * K];
            for (int q = 1; q < h; q++) m = m + (x[q] - x[K - 1 - q]) * tb.dctF[v * K + q];
        }
        T[e] = m;
    }
    wave_sync();
    // columns: out[u][v] = REDFT10 of column v, times coef_norm
    for (int e = lane; e < K2; e += 64) {
        const int u = e / K, v = e - u * K;
        float m;
        if (h == 0) {
            m = T[v] * tb.dctF[u * K];  // K = 1
        } else if ((u & 1) == 0) {
            m = (T[v] + T[(K - 1) * K + v]) * tb.dctF[u * K];
            for (int q = 1; q < h; q++) m = m + (T[q * K + v] + T[(K - 1 - q) * K + v]) * tb.dctF[u * K + q];
            if (K & 1) m = m + T[h * K + v] * tb.dctF[u * K + h];
        } else {
            m = (T[v] - T[(K - 1) * K + v]) * tb.dctF[u * K];
            for (int q = 1; q < h; q++) m = m + (T[q * K + v] - T[(K - 1 - q) * K + v]) * tb.dctF[u * K + q];
        }
        out[e] = m * tb.nrm[e];
    }
    wave_sync();
}

// One REDFT01 output o of the K values z[i * sz] (the oracle's E / O
// sums: rows o and K-1-o share them).
__device__ __forceinline__ float gen_redft01(const float* z, int sz, int o, int K, const TabsG& tb) {
    const int h = K / 2;
    const int r = o < K - 1 - o ? o : K - 1 - o;
    float e = z[0] * tb.dctI[r * K];
    for (int i = 2; i < K; i += 2) e = e + z[i * sz] * tb.dctI[r * K + i];
    if ((K & 1) && o == h) return e;
    float od = z[sz] * tb.dctI[r * K + 1];
    for (int i = 3; i < K; i += 2) od = od + z[i * sz] * tb.dctI[r * K + i];
    return o == r ? e + od : e - od;
}

// 2D inverse transform of coefficients S into the pixel patch out
// (row-major), P and T scratch (the oracle's bior_inv / dct_inv).
__device__ void gen_inv2d(const float* S, float* P, float* T, float* out, int K, bool bior, const TabsG& tb,
                          int lane) {
    const int K2 = K * K;
    if (bior) {
        for (int e = lane; e < 64; e += 64) {  // a[i][j] = sig[i + 8 j]; b = a B^T
            const int i = e >> 3, j = e & 7;
            float m = 0.f;
            for (int s = 0; s < 8; s++) m = m + S[i + s * 8] * tb.biorI[j * 8 + s];
            T[e] = m;
        }
        wave_sync();
        for (int e = lane; e < 64; e += 64) {  // a2[i][j] = sum_s b[s][i] B[j][s]; sig[8 j + i] = a2[i][j]
            const int i = e >> 3, j = e & 7;
            float m = 0.f;
            for (int s = 0; s < 8; s++) m = m + T[s * 8 + i] * tb.biorI[j * 8 + s];
            out[j * 8 + i] = m;
        }
        wave_sync();
        return;
    }
    for (int e = lane; e < K2; e += 64) P[e] = S[e] * tb.inv[e];
    wave_sync();
    for (int e = lane; e < K2; e += 64) {  // rows: T[u][o]
        const int u = e / K, o = e - u * K;
        T[e] = gen_redft01(P + u * K, 1, o, K, tb);
    }
    wave_sync();
    for (int e = lane; e < K2; e += 64) {  // columns: out[p][q] = dct_out * REDFT01(T[.][q])[p]
        const int p = e / K, q = e - p * K;
        out[e] = tb.dct_out * gen_redft01(T + q, K, p, K, tb);
    }
    wave_sync();
}

template <int NS>
__device__ void gen_shrink(float* zg, float* eg, int nkk, int step, bool had, float lambda, float sig, int lane,
                           int* cnt_out) {
    const float coef = 1.0f / (float)NS;
    int cnt = 0;
    for (int kk = lane; kk < nkk; kk += 64) {
        float z[NS], e[NS];
#pragma unroll
        for (int n = 0; n < NS; n++) {
            z[n] = zg[kk * NS + n];
            if (step == 2) e[n] = eg[kk * NS + n];
        }
        if (step == 1) {
            t3_fwd<NS>(z, had);
            const float thr = had ? lambda * sig * sqrtf((float)NS) : lambda * sig;
#pragma unroll
            for (int n = 0; n < NS; n++) {
                if ((n == 0 && kk == 0) || fabsf(z[n]) > thr) cnt++;
                else z[n] = 0.f;
            }
            t3_inv<NS>(z, had);
            if (had) {
#pragma unroll
                for (int n = 0; n < NS; n++) z[n] *= coef;
            }
#pragma unroll
            for (int n = 0; n < NS; n++) zg[kk * NS + n] = z[n];
        } else {
            t3_fwd<NS>(z, had);
            t3_fwd<NS>(e, had);
            const float s2 = sig * sig;
#pragma unroll
            for (int n = 0; n < NS; n++) {
                float value = e[n] * e[n];
                if (had) value = value * coef;
                value /= (value + s2);
                e[n] = had ? z[n] * value * coef : z[n] * value;
                z[n] = value * value;  // kept for the weight's sequential sum
            }
            t3_inv<NS>(e, had);
#pragma unroll
            for (int n = 0; n < NS; n++) {
                zg[kk * NS + n] = z[n];
                eg[kk * NS + n] = e[n];
            }
        }
    }
    *cnt_out = wave_sum_i32(cnt);
}

template <int C, int STEP>
__global__ __launch_bounds__(64) void k_filter_gen(GenArgs a) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x;
    const long r = a.r_begin + (long)blockIdx.x;
    if (r >= a.r_end) return;
    const int K = a.K, K2 = K * K, kt = a.kt, K2t = K2 * kt, N = a.N;
    const TabsG& tb = *a.tabs;
    const bool bior = a.T2D == 1, had = a.T3D == 2, temporal = kt > 1 && !bior;
    const unsigned t_ = (unsigned)(r / a.nrf), i_ = (unsigned)(r - (long)t_ * a.nrf);
    const unsigned slot = (t_ % a.ring) * a.nrf + i_;
    const int nsx = (int)a.nsx_out[slot];
    const uint32_t* mem = a.idx_out + (size_t)slot * N;
    const unsigned W = a.W, wh = a.wh, whc = a.whc;
    // LDS: zg, eg [K2t][nsx]; P, T, S scratch [K2]; member trajectories
    float* zg = smem;
    float* eg = zg + K2t * nsx;
    float* P = eg + (STEP == 2 ? K2t * nsx : 0);
    float* T = P + K2;
    float* S = T + K2;
    int* trj = (int*)(S + K2);  // [n][kt] x | y << 16
    if (lane < nsx) {
        const unsigned mi = mem[lane];
        const unsigned pt = mi / whc;
        unsigned py = (mi - pt * whc) / W, px = mi - pt * whc - py * W;
        float fx = (float)px, fy = (float)py;
        for (int ht = 0; ht < kt; ht++) {
            trj[lane * kt + ht] = (int)(px | (py << 16));
            if (a.mc && ht + 1 < kt) traj_step(a, px, py, fx, fy, pt + ht);
        }
    }
    wave_sync();
    if (a.mpos)
        for (int e = lane; e < nsx * kt; e += 64) a.mpos[((size_t)slot * N + e / kt) * kt + e % kt] = (uint32_t)trj[e];
    float wsel = 0.f;
    for (int c = 0; c < C; c++) {
        // forward 2D transforms of every slice, into group layout [kk][n]
        for (int g = 0; g < (STEP == 2 ? 2 : 1); g++) {
            const float* src = g == 0 ? a.noisy : a.basic;
            float* grp = g == 0 ? zg : eg;
            for (int n = 0; n < nsx; n++) {
                const unsigned pt = mem[n] / whc;
                for (int ht = 0; ht < kt; ht++) {
                    const unsigned pos = (unsigned)trj[n * kt + ht];
                    const float* pp = src + (size_t)(pt + ht) * whc + (size_t)c * wh + (size_t)(pos >> 16) * W + (pos & 0xFFFFu);
                    for (int e = lane; e < K2; e += 64) P[e] = pp[(e / K) * W + e % K];
                    wave_sync();
                    gen_fwd2d(P, T, S, K, bior, tb, lane);
                    for (int e = lane; e < K2; e += 64) grp[(ht * K2 + e) * nsx + n] = S[e];
                    wave_sync();
                }
            }
            if (temporal) {  // along each coefficient's slices (src/vbm3d.cpp:573-575, 816-822)
                for (int e = lane; e < K2 * nsx; e += 64) {
                    float* x = grp + e;  // coefficient e / nsx of member e % nsx
                    float y[8];
                    for (int u = 0; u < kt; u++) {
                        float m = 0.f;
                        for (int tt = 0; tt < kt; tt++) m = m + x[tt * K2 * nsx] * tb.tF[u * kt + tt];
                        y[u] = m;
                    }
                    for (int u = 0; u < kt; u++)
                        x[u * K2 * nsx] = (float)((double)(y[u] * tb.tnorm) * (u == 0 ? 0.7071067811865475 : 1.0));
                }
                wave_sync();
            }
        }
        // along the group, shrinkage, inverse (:1405-1666)
        const float sg = c == 0 ? a.sig[0] : (c == 1 ? a.sig[1] : a.sig[2]);
        int cnt = 0;
        switch (nsx) {
            case 1: gen_shrink<1>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
            case 2: gen_shrink<2>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
            case 4: gen_shrink<4>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
            case 8: gen_shrink<8>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
            case 16: gen_shrink<16>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
            default: gen_shrink<32>(zg, eg, K2t, STEP, had, a.lambda, sg, lane, &cnt); break;
        }
        wave_sync();
        float wc;
        if (STEP == 1) {
            wc = 1.f / (sg * sg * (float)cnt);
        } else {
            // the weight: one sequential sum of value^2 in the flat order
            float acc = 0.f;
            if (lane == 0)
                for (int f = 0; f < K2t * nsx; f++) acc = acc + zg[f];
            acc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), 0));
            const float s2 = sg * sg;
            wc = had ? (acc > 0.0f ? 1.0f / (s2 * acc) : 1.0f) : 1.f / (s2 * acc);
        }
        if (lane == c) wsel = wc;
        float* est = STEP == 2 ? eg : zg;
        if (temporal) {  // :589, :829-831
            for (int e = lane; e < K2 * nsx; e += 64) {
                float* x = est + e;
                float z[8];
                for (int u = 0; u < kt; u++)
                    z[u] = (float)((double)x[u * K2 * nsx] * (u == 0 ? 1.414213562373095 : 1.0));
                for (int tt = 0; tt < kt; tt++) {
                    float m = 0.f;
                    for (int u = 0; u < kt; u++) m = m + z[u] * tb.tI[tt * kt + u];
                    x[tt * K2 * nsx] = m * tb.tnorm;
                }
            }
            wave_sync();
        }
        // inverse 2D per slice, est[slot][n][c][slice][pixel]
        for (int n = 0; n < nsx; n++)
            for (int ht = 0; ht < kt; ht++) {
                for (int e = lane; e < K2; e += 64) S[e] = est[(ht * K2 + e) * nsx + n];
                wave_sync();
                float* dst = a.est + (((size_t)slot * N + n) * C + c) * K2t + ht * K2;
                gen_inv2d(S, P, T, dst, K, bior, tb, lane);
            }
    }
    if (lane < C) a.wgt[(size_t)slot * C + lane] = wsel;
}

// ---------------------------------------------------------------------
// Aggregation, general form: as the fast kernel (ordered gather per output
// tile, no atomics), with members that span kt frames along their
// trajectories (src/vbm3d.cpp:612-650, 860-924).  For output frame T a
// member of a reference patch in frame t' contributes its slice T - pt;
// candidates are scanned in (crop, frame, row, column, n) order.
// ---------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(kAggThreads) void k_aggregate_gen(AArgs a, const uint32_t* __restrict__ mpos,
                                                               const TabsG* __restrict__ tg) {
    constexpr int TX = kGenAggTileX, TY = kGenAggTileY, NT = kAggThreads, NS = kGenAggScan;
    constexpr int QS = 8, NWV = NT / 64, WX = TX / QS;
    static_assert(TX * TY == NT && (TX / QS) * (TY / QS) == NWV, "one thread per pixel, 8x8 per wavefront");
    // the ordered list of one chunk of NS scanned reference patches, at
    // most N entries each: slot, member | slice << 8, position
    extern __shared__ uint32_t agg_lists[];
    uint32_t* L_sl = agg_lists;
    uint32_t* L_nh = agg_lists + NS * a.N;
    uint32_t* L_pos = agg_lists + 2 * NS * a.N;
    __shared__ float s_kai[256];
    __shared__ int s_w[NWV];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const unsigned to = a.t_begin + blockIdx.z;
    const int qx0 = x0 + (wv % WX) * QS, qy0 = y0 + (wv / WX) * QS;
    const int px = qx0 + (lane & (QS - 1)), py = qy0 + lane / QS;
    const bool inside = px < (int)a.W && py < (int)a.H;
    const int K = a.K, K2 = K * K, N = a.N, kt = a.kt, K2t = K2 * kt;
    const unsigned W = a.W, wh = a.wh, whc = a.whc;
    for (int i = tid; i < K2; i += NT) s_kai[i] = tg->kaiser[i];
    float tn[C], td[C];
#pragma unroll
    for (int c = 0; c < C; c++) tn[c] = td[c] = 0.f;
    const int trlo = max(0, (int)to - a.Nf - (kt - 1)), trhi = min((int)a.Fr - 1, (int)to + a.Nf);
    const int rlo = y0 - a.R - (K - 1), rhi = y0 + TY - 1 + a.R;
    const int clo = x0 - a.R - (K - 1), chi = x0 + TX - 1 + a.R;
    // slice `ht` of member mi of the patch in `slot` lands at `pos`
    auto member_pos = [&](uint32_t slot, int n, uint32_t mi, int ht) -> uint32_t {
        if (mpos) return mpos[((size_t)slot * N + n) * kt + ht];
        const uint32_t rem = mi % wh;
        return (rem % W) | ((rem / W) << 16);
    };
    auto hits_tile = [&](uint32_t pos) {
        const int qx = (int)(pos & 0xFFFFu), qy = (int)(pos >> 16);
        return qx <= x0 + TX - 1 && qx + K - 1 >= x0 && qy <= y0 + TY - 1 && qy + K - 1 >= y0;
    };
    for (int ci = 0; ci < a.ncrop; ci++) {
        const int* meta = a.crop_meta + ci * 5;
        const int off = meta[0], nrows = meta[1], ncols = meta[2];
        const uint16_t* rows = a.grid_pos + meta[3];
        const uint16_t* cols = a.grid_pos + meta[4];
        auto lb = [](const uint16_t* v, int n, int x) {
            int lo = 0, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((int)v[mid] < x) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        const int ia = lb(rows, nrows, rlo), ib = lb(rows, nrows, rhi + 1);
        const int ja = lb(cols, ncols, clo), jb = lb(cols, ncols, chi + 1);
        const int nr = ib - ia, nc = jb - ja, ntr = trhi - trlo + 1;
        const long total = (nr > 0 && nc > 0 && ntr > 0) ? (long)ntr * nr * nc : 0;
        float pn[C], pd[C];
#pragma unroll
        for (int c = 0; c < C; c++) pn[c] = pd[c] = 0.f;
        for (long cb = 0; cb < total; cb += NS) {
            const long e = cb + tid;
            int cnt = 0, ns = 0;
            uint32_t slot = 0;
            if (tid < NS && e < total) {
                const long rest = e / nc, jj = e - rest * nc;
                const long tq = rest / nr, ii = rest - tq * nr;
                const unsigned tr = trlo + (unsigned)tq;
                slot = (tr % a.ring) * a.nrf + off + (ia + (unsigned)ii) * ncols + (ja + (unsigned)jj);
                ns = (int)a.nsx[slot];
                for (int n = 0; n < ns; n++) {
                    const uint32_t mi = a.idx[(size_t)slot * N + n];
                    const int ht = (int)to - (int)(mi / whc);
                    if (ht >= 0 && ht < kt && hits_tile(member_pos(slot, n, mi, ht))) cnt++;
                }
            }
            // ordered compaction: exclusive block scan of the counts
            int x = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) s_w[wv] = x;
            __syncthreads();
            int base = 0, tot = 0;
            for (int q = 0; q < NWV; q++) {
                base += q < wv ? s_w[q] : 0;
                tot += s_w[q];
            }
            int o = base + x - cnt;
            for (int n = 0; n < ns && cnt > 0; n++) {
                const uint32_t mi = a.idx[(size_t)slot * N + n];
                const int ht = (int)to - (int)(mi / whc);
                if (ht < 0 || ht >= kt) continue;
                const uint32_t pos = member_pos(slot, n, mi, ht);
                if (!hits_tile(pos)) continue;
                L_sl[o] = slot;
                L_nh[o] = (uint32_t)n | ((uint32_t)ht << 8);
                L_pos[o] = pos;
                o++;
            }
            __syncthreads();
            // every pixel takes the list entries that cover it, in order
            for (int j = 0; j < tot; j++) {
                const uint32_t pos = L_pos[j];
                const int dx = px - (int)(pos & 0xFFFFu), dy = py - (int)(pos >> 16);
                if (!inside || (unsigned)dx >= (unsigned)K || (unsigned)dy >= (unsigned)K) continue;
                const uint32_t sl = L_sl[j], nh = L_nh[j];
                const uint32_t n = nh & 0xFFu, ht = nh >> 8;
                const float kai = s_kai[dy * K + dx];
                const float* eb = a.est + ((size_t)sl * N + n) * C * K2t + ht * K2 + dy * K + dx;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const float ev = eb[(size_t)c * K2t];
                    if (a.raw) {
                        pn[c] = pn[c] + ev;
                    } else {
                        const float kw = kai * a.wgt[(size_t)sl * C + c];
                        pn[c] = pn[c] + kw * ev;
                        pd[c] = pd[c] + kw;
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

}  // namespace

// ---------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------
hipError_t launch_bm_gen(const GenArgs& a, hipStream_t s) {
    const long n = a.r_end - a.r_begin;
    if (n <= 0) return hipSuccess;
    const size_t per_wave = (size_t)gen_bm_layout(a.Nb, a.Nf, a.Ns, a.Npr).total * 4;
    const int waves = per_wave * 2 <= 64 * 1024 ? 2 : 1;
    if (per_wave > 64 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + waves - 1) / waves)), block(64 * waves);
    if (a.C == 1) hipLaunchKernelGGL((k_bm_gen<1>), grid, block, per_wave * waves, s, a);
    else hipLaunchKernelGGL((k_bm_gen<3>), grid, block, per_wave * waves, s, a);
    return hipGetLastError();
}

hipError_t launch_filter_gen(int step, const GenArgs& a, hipStream_t s) {
    const long n = a.r_end - a.r_begin;
    if (n <= 0) return hipSuccess;
    const size_t lds = (size_t)gen_filter_words(a.K, a.kt, a.N, step) * 4;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const dim3 grid((unsigned)n), block(64);
    if (a.C == 1) {
        if (step == 1) hipLaunchKernelGGL((k_filter_gen<1, 1>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_filter_gen<1, 2>), grid, block, lds, s, a);
    } else {
        if (step == 1) hipLaunchKernelGGL((k_filter_gen<3, 1>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_filter_gen<3, 2>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_aggregate_gen(int C, const AArgs& a, const uint32_t* mpos, const TabsG* tabs, unsigned frames,
                                hipStream_t s) {
    if (frames == 0) return hipSuccess;
    if (a.N < 1 || a.N > kGenMaxN) return hipErrorInvalidValue;
    dim3 grid((a.W + kGenAggTileX - 1) / kGenAggTileX, (a.H + kGenAggTileY - 1) / kGenAggTileY, frames);
    const size_t lds = (size_t)3 * kGenAggScan * a.N * sizeof(uint32_t);
    if (C == 1) hipLaunchKernelGGL((k_aggregate_gen<1>), grid, dim3(kAggThreads), lds, s, a, mpos, tabs);
    else hipLaunchKernelGGL((k_aggregate_gen<3>), grid, dim3(kAggThreads), lds, s, a, mpos, tabs);
    return hipGetLastError();
}

}  // namespace vb
