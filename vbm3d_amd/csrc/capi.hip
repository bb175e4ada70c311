// capi.hip -- host orchestration and C ABI of libvbm3d_hip.so.
//
// The orchestration replaces run_vbm3d's thread tiling and per-row loops
// (src/vbm3d.cpp:87-376, 527-653, 759-927) with a frame-chunked pipeline
// on one HIP stream:
//     for each chunk of reference frames:
//         k_bm        -> group indices and sizes (ring slot per reference)
//         k_filter    -> filtered group estimates and weights
//         k_aggregate -> output frames whose contributors are all done
// Per-reference results live in a ring of S + 2 Nf frames, so device
// memory does not grow with the clip length.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vbm3d_hip.h"
#include <rocblas/rocblas.h>

#include "kernels.h"
#include "select.h"

using namespace vb;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(VBM3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

// Fixed device memory that ensure() carves buffers from instead of
// allocating, while vbm3d_run_device_noalloc runs on this thread.
struct Arena {
    char* base = nullptr;  // device address of the scratch array
    size_t size = 0, used = 0, fixed = 0;  // fixed: carved once per call (kept across steps)
    bool overflow = false;
};
thread_local Arena* g_arena = nullptr;

hipError_t ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return hipSuccess;
    if (g_arena) {
        const size_t at = (g_arena->used + 255) & ~(size_t)255;
        if (at + bytes > g_arena->size) {
            g_arena->overflow = true;
            return hipErrorOutOfMemory;
        }
        b.p = g_arena->base + at;
        b.cap = bytes;
        g_arena->used = at + bytes;
        return hipSuccess;
    }
    if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess) return e;
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.cap = bytes;
    return e;
}

// ---------------------------------------------------------------------
// Reference-patch grid (src/Utilities/Utilities.cpp:152-173, 400-415;
// src/Utilities/LibVideoT.hpp:903-961; src/vbm3d.cpp:471-489, 700-719).
// ---------------------------------------------------------------------
struct GridDef {
    std::vector<std::vector<unsigned>> rows, cols;  // per crop, absolute
};

bool axis(std::vector<unsigned>& out, long b, long e, unsigned step) {
    out.clear();
    if (b < 0 || e < b || step == 0) return false;
    long v = b;
    for (; v <= e; v += step) out.push_back((unsigned)v);
    if ((long)out.back() < e) out.push_back((unsigned)e);
    return true;
}

bool build_grid(GridDef& g, unsigned W, unsigned H, unsigned k, unsigned p, int n, int nb) {
    g.rows.clear();
    g.cols.clear();
    if (W < k || H < k || p == 0) return false;
    if (nb == 1) {
        std::vector<unsigned> r, c;
        if (!axis(r, 0, (long)H - k, p) || !axis(c, 0, (long)W - k, p)) return false;
        g.rows.push_back(r);
        g.cols.push_back(c);
        return true;
    }
    if (nb < 2 || nb > 32 || (nb & (nb - 1))) return false;
    // tiles: the smallest divisor >= 2 and its cofactor, wider than tall
    unsigned b = 2;
    while ((unsigned)nb % b) b++;
    unsigned a = nb / b;
    if (b > a) std::swap(a, b);
    const unsigned nW = a, nH = b;
    const int wT = (int)std::ceil(float(W) / float(nW));
    const int hT = (int)std::ceil(float(H) / float(nH));
    const int border = 2 * n;
    for (int ty = 0; ty < (int)nH; ty++)
        for (int tx = 0; tx < (int)nW; tx++) {
            const int ox = std::max(0, tx * wT - border), oy = std::max(0, ty * hT - border);
            const int ex = std::min((int)W, (tx + 1) * wT + border);
            const int ey = std::min((int)H, (ty + 1) * hT + border);
            std::vector<unsigned> r, c;
            const long rb = oy == 0 ? 0 : n, cb = ox == 0 ? 0 : n;
            const long re = ey == (int)H ? (long)ey - oy - k : (long)ey - oy - k - n;
            const long ce = ex == (int)W ? (long)ex - ox - k : (long)ex - ox - k - n;
            if (!axis(r, rb, re, p) || !axis(c, cb, ce, p)) return false;
            for (auto& v : r) v += oy;
            for (auto& v : c) v += ox;
            g.rows.push_back(r);
            g.cols.push_back(c);
        }
    return true;
}

// Per-channel noise (estimate_sigma, src/Utilities/Utilities.cpp:282-333)
// as called by the reference with a bool colour flag (src/vbm3d.cpp:467):
// true -> YCbCr row, false -> YUV row.
void sigma_rows(float sigma, unsigned C, bool cs, float* out) {
    if (C == 1) {
        out[0] = sigma;
        return;
    }
    out[0] = sqrtf(0.299f * 0.299f + 0.587f * 0.587f + 0.114f * 0.114f) * sigma;
    if (cs) {
        out[1] = sqrtf(0.169f * 0.169f + 0.331f * 0.331f + 0.500f * 0.500f) * sigma;
        out[2] = sqrtf(0.500f * 0.500f + 0.419f * 0.419f + 0.081f * 0.081f) * sigma;
    } else {
        out[1] = sqrtf(0.14713f * 0.14713f + 0.28886f * 0.28886f + 0.436f * 0.436f) * sigma;
        out[2] = sqrtf(0.615f * 0.615f + 0.51498f * 0.51498f + 0.10001f * 0.10001f) * sigma;
    }
}

// Bior1.5 MATLAB basis (src/lib_transforms.cpp:141-148, 281-288), the
// same decimal values rounded to float.
const double kBF[8][8] = {
    {0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274},
    {0.219417649252501, 0.449283757993216, 0.449283757993216, 0.219417649252501, -0.219417649252501, -0.449283757993216, -0.449283757993216, -0.219417649252501},
    {0.569359398342846, 0.402347308162278, -0.402347308162278, -0.569359398342846, -0.083506045090284, 0.083506045090284, -0.083506045090284, 0.083506045090284},
    {-0.083506045090284, 0.083506045090284, -0.083506045090284, 0.083506045090284, 0.569359398342846, 0.402347308162278, -0.402347308162278, -0.569359398342846},
    {0.707106781186547, -0.707106781186547, 0, 0, 0, 0, 0, 0},
    {0, 0, 0.707106781186547, -0.707106781186547, 0, 0, 0, 0},
    {0, 0, 0, 0, 0.707106781186547, -0.707106781186547, 0, 0},
    {0, 0, 0, 0, 0, 0, 0.707106781186547, -0.707106781186547}};
const double kBI[8][8] = {
    {0.353553390593274, 0.353553390593274, 0.5, 0.0, 0.707106781186547, -0.121533978016438, 0.0, 0.121533978016438},
    {0.353553390593274, 0.353553390593274, 0.5, -0.0, -0.707106781186547, -0.121533978016438, 0.0, 0.121533978016438},
    {0.353553390593274, 0.353553390593274, -0.5, -0.0, 0.121533978016438, 0.707106781186548, -0.121533978016438, 0.0},
    {0.353553390593274, 0.353553390593274, -0.5, 0.0, 0.121533978016438, -0.707106781186548, -0.121533978016438, 0.0},
    {0.353553390593274, -0.353553390593274, 0.0, 0.5, 0.0, 0.121533978016438, 0.707106781186547, -0.121533978016438},
    {0.353553390593274, -0.353553390593274, 0.0, 0.5, 0.0, 0.121533978016438, -0.707106781186547, -0.121533978016438},
    {0.353553390593274, -0.353553390593274, 0.0, -0.5, -0.121533978016438, 0.0, 0.121533978016438, 0.707106781186547},
    {0.353553390593274, -0.353553390593274, 0.0, -0.5, -0.121533978016438, 0.0, 0.121533978016438, -0.707106781186547}};

// Kaiser(beta = 2) quarter tables (src/vbm3d.cpp:1755-1841).
float kaiser_at(unsigned k, unsigned a, unsigned b) {
    static const float q4[2][2] = {{0.1924f, 0.4055f}, {0.4055f, 0.8544f}};
    static const float q6[3][3] = {{0.1924f, 0.3368f, 0.4265f}, {0.3368f, 0.5893f, 0.7464f}, {0.4265f, 0.7464f, 0.9454f}};
    static const float q7[4][4] = {{0.1924f, 0.3151f, 0.4055f, 0.4387f}, {0.3151f, 0.5161f, 0.6640f, 0.7184f},
                                   {0.4055f, 0.6640f, 0.8544f, 0.9243f}, {0.4387f, 0.7184f, 0.9243f, 1.0000f}};
    static const float q8[4][4] = {{0.1924f, 0.2989f, 0.3846f, 0.4325f}, {0.2989f, 0.4642f, 0.5974f, 0.6717f},
                                   {0.3846f, 0.5974f, 0.7688f, 0.8644f}, {0.4325f, 0.6717f, 0.8644f, 0.9718f}};
    static const float q12[6][6] = {{0.1924f, 0.2615f, 0.3251f, 0.3782f, 0.4163f, 0.4362f},
                                    {0.2615f, 0.3554f, 0.4419f, 0.5139f, 0.5657f, 0.5927f},
                                    {0.3251f, 0.4419f, 0.5494f, 0.6390f, 0.7033f, 0.7369f},
                                    {0.3782f, 0.5139f, 0.6390f, 0.7433f, 0.8181f, 0.8572f},
                                    {0.4163f, 0.5657f, 0.7033f, 0.8181f, 0.9005f, 0.9435f},
                                    {0.4362f, 0.5927f, 0.7369f, 0.8572f, 0.9435f, 0.9885f}};
    if (k == 4) return q4[a][b];
    if (k == 6) return q6[a][b];
    if (k == 7) return q7[a][b];
    if (k == 8) return q8[a][b];
    if (k == 12) return q12[a][b];
    return 1.0f;
}

void fill_tabs(Tabs& t, unsigned k) {
    std::memset(&t, 0, sizeof t);
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 8; b++) {
            t.biorF[a * 8 + b] = (float)kBF[a][b];
            t.biorI[a * 8 + b] = (float)kBI[a][b];
        }
    // REDFT10 / REDFT01 closed forms, double then one rounding to float
    for (unsigned v = 0; v < k; v++)
        for (unsigned q = 0; q < k; q++) {
            t.dctF[v * k + q] = (float)(2.0 * std::cos(M_PI * (2.0 * q + 1.0) * v / (2.0 * k)));
            t.dctI[q * k + v] = v == 0 ? 1.0f : (float)(2.0 * std::cos(M_PI * v * (2.0 * q + 1.0) / (2.0 * k)));
        }
    // coef_norm / coef_norm_inv (src/vbm3d.cpp:1843-1863)
    const float coef = 0.5f / ((float)k);
    for (unsigned i = 0; i < k; i++)
        for (unsigned j = 0; j < k; j++) {
            if (i == 0 && j == 0) {
                t.nrm[i * k + j] = 0.5f * coef;
                t.inv[i * k + j] = 2.0f;
            } else if (i * j == 0) {
                t.nrm[i * k + j] = (float)(0.7071067811865475 * coef);
                t.inv[i * k + j] = (float)1.414213562373095;
            } else {
                t.nrm[i * k + j] = 1.0f * coef;
                t.inv[i * k + j] = 1.0f;
            }
        }
    for (unsigned p = 0; p < k; p++)
        for (unsigned q = 0; q < k; q++)
            t.kaiser[p * k + q] = kaiser_at(k, std::min(p, k - 1 - p), std::min(q, k - 1 - q));
    t.dct_out = 1.0f / (float)(k * 2);
}

// Tables of the general path: k up to 16, and the temporal DCT of length
// kt (src/vbm3d.cpp:1866-1956: REDFT10 / REDFT01 in closed form, double
// then one rounding to float; 1/sqrt(2 kt) rounded to float).
void fill_tabs_gen(TabsG& t, unsigned k, unsigned kt) {
    std::memset(&t, 0, sizeof t);
    Tabs b;
    fill_tabs(b, std::min(k, 8u));
    std::memcpy(t.biorF, b.biorF, sizeof t.biorF);
    std::memcpy(t.biorI, b.biorI, sizeof t.biorI);
    for (unsigned v = 0; v < k; v++)
        for (unsigned q = 0; q < k; q++) {
            t.dctF[v * k + q] = (float)(2.0 * std::cos(M_PI * (2.0 * q + 1.0) * v / (2.0 * k)));
            t.dctI[q * k + v] = v == 0 ? 1.0f : (float)(2.0 * std::cos(M_PI * v * (2.0 * q + 1.0) / (2.0 * k)));
        }
    const float coef = 0.5f / ((float)k);
    for (unsigned i = 0; i < k; i++)
        for (unsigned j = 0; j < k; j++) {
            if (i == 0 && j == 0) {
                t.nrm[i * k + j] = 0.5f * coef;
                t.inv[i * k + j] = 2.0f;
            } else if (i * j == 0) {
                t.nrm[i * k + j] = (float)(0.7071067811865475 * coef);
                t.inv[i * k + j] = (float)1.414213562373095;
            } else {
                t.nrm[i * k + j] = 1.0f * coef;
                t.inv[i * k + j] = 1.0f;
            }
        }
    for (unsigned p = 0; p < k; p++)
        for (unsigned q = 0; q < k; q++)
            t.kaiser[p * k + q] = kaiser_at(k, std::min(p, k - 1 - p), std::min(q, k - 1 - q));
    t.dct_out = 1.0f / (float)(k * 2);
    for (unsigned u = 0; u < kt; u++)
        for (unsigned tt = 0; tt < kt; tt++) {
            t.tF[u * kt + tt] = (float)(2.0 * std::cos(M_PI * (2.0 * tt + 1.0) * u / (2.0 * kt)));
            t.tI[tt * kt + u] = u == 0 ? 1.0f : (float)(2.0 * std::cos(M_PI * u * (2.0 * tt + 1.0) / (2.0 * kt)));
        }
    t.tnorm = (float)(1. / std::sqrt(2.0 * kt));
}

// Whether the fast kernels (kernels_bm/filter/agg.hip) implement these
// parameters; the general kernels (kernels_gen.hip) take the rest.
bool fast_ok(const vbm3d_params& p, unsigned W, unsigned H) {
    if (p.kt != 1 || p.k < 4 || p.k > 8) return false;
    if (p.N == 0 || p.N > (unsigned)kMaxN || (p.N & (p.N - 1))) return false;
    if (p.Nb == 0 || p.Nb > 8 || p.Nb * p.Nb > 64) return false;
    if (p.Ns == 0 || p.Ns > (unsigned)kMaxWin || p.Npr == 0 || p.Npr > (unsigned)kMaxWin) return false;
    if (2 * p.Nb * (unsigned)pow2ceil((int)(p.Npr * p.Npr)) > (unsigned)kMaxSlots) return false;
    if (p.Nb * (2 * p.Nf + 1) > (unsigned)kMaxBest) return false;
    // every predictive window must yield Nb candidates: the fast matching
    // keeps exactly Nb picks per frame
    const unsigned h = (p.Npr - 1) / 2 + 1;
    return std::min(W - p.k + 1, h) * std::min(H - p.k + 1, h) >= p.Nb;
}

// What the kernels implement (fast or general); everything else is
// rejected with a message.
int check(const vbm3d_params& p, unsigned W, unsigned H, unsigned F, unsigned C, int step) {
    char b[160];
    auto no = [&](int code, const char* m) {
        std::snprintf(b, sizeof b, "step %d: %s", step, m);
        return fail(code, b);
    };
    if (C != 1 && C != 3) return no(VBM3D_ERR_ARG, "only 1 or 3 channels are supported");
    if (W == 0 || H == 0 || F == 0) return no(VBM3D_ERR_ARG, "empty video");
    if (W > 65535 || H > 65535) return no(VBM3D_ERR_ARG, "frame dimensions above 65535");
    if ((unsigned long long)W * H * C * F >= (1ull << 32))
        return no(VBM3D_ERR_ARG, "video too large for 32-bit patch indices");
    if (p.k < 1 || p.k > (unsigned)kGenMaxK) return no(VBM3D_ERR_UNSUPPORTED, "patch size k must be in [1, 16]");
    if (W < p.k || H < p.k) return no(VBM3D_ERR_ARG, "frame smaller than a patch");
    if (p.kt < 1 || p.kt > (unsigned)kGenMaxKt) return no(VBM3D_ERR_UNSUPPORTED, "kt must be in [1, 8]");
    if (p.kt > F) return no(VBM3D_ERR_ARG, "fewer frames than kt");
    if (p.k * p.k * p.kt > (unsigned)kGenMaxK2t) return no(VBM3D_ERR_UNSUPPORTED, "k^2 kt > 256");
    if (p.N == 0 || p.N > (unsigned)kGenMaxN) return no(VBM3D_ERR_UNSUPPORTED, "N must be in [1, 32]");
    if (p.Nb == 0 || p.Nb > (unsigned)kGenMaxNb) return no(VBM3D_ERR_UNSUPPORTED, "Nb must be in [1, 16]");
    if (p.Ns == 0 || p.Ns > (unsigned)kGenMaxWin || p.Npr == 0 || p.Npr > (unsigned)kGenMaxWin)
        return no(VBM3D_ERR_UNSUPPORTED, "Ns and Npr must be in [1, 15]");
    if (p.Nf > (unsigned)kGenMaxNf) return no(VBM3D_ERR_UNSUPPORTED, "Nf must be <= 8");
    if (p.p == 0) return no(VBM3D_ERR_ARG, "step p must be > 0");
    if (p.T_2D != VBM3D_DCT && p.T_2D != VBM3D_BIOR) return no(VBM3D_ERR_ARG, "T_2D must be DCT (0) or BIOR (1)");
    if (p.T_2D == VBM3D_BIOR && p.k != 8) return no(VBM3D_ERR_ARG, "BIOR needs k == 8");
    // the reference-frame window must yield Nb candidates (src/vbm3d.cpp:1115-1119)
    const unsigned h = (p.Ns - 1) / 2 + 1;
    if (std::min(W - p.k + 1, h) * std::min(H - p.k + 1, h) < p.Nb)
        return no(VBM3D_ERR_UNSUPPORTED, "search window smaller than Nb");
    if (!fast_ok(p, W, H) && (size_t)gen_bm_layout(p.Nb, p.Nf, p.Ns, p.Npr).total * 4 > 64 * 1024)
        return no(VBM3D_ERR_UNSUPPORTED, "search shape too large for the general matching kernel");
    if (!fast_ok(p, W, H) && (size_t)gen_filter_words((int)p.k, (int)p.kt, (int)p.N, 2) * 4 > 64 * 1024)
        return no(VBM3D_ERR_UNSUPPORTED, "k^2 kt N too large for the general filtering kernel (2 k^2 kt N + 3 k^2 + N kt words <= 64 KiB)");
    return VBM3D_OK;
}

// Largest displacement between a group member and its reference patch:
// the reference-frame window's half width, plus per predictive frame the
// window's half width and the largest rounded flow.  Evaluated in 64 bits
// and capped at the frame size (a projection is clamped to the frame, so
// no member lies farther), so that a huge or non-finite flow value cannot
// wrap it.
int member_reach(const vbm3d_params& p, unsigned maxflow, unsigned W, unsigned H) {
    unsigned long long r = (p.Ns - 1) / 2 + (unsigned long long)p.Nf * ((p.Npr - 1) / 2 + (unsigned long long)maxflow);
    // a motion-compensated trajectory drifts by at most the rounded sum of
    // kt - 1 flow steps
    if (p.mc && p.kt > 1) r += (unsigned long long)(p.kt - 1) * ((unsigned long long)maxflow + 1);
    return (int)std::min<unsigned long long>(r, std::max(W, H));
}

// The same reach per frame distance: a member k frames from its reference
// patch lies within R0 + k Rs of it (the reference-frame window, then k
// predictive windows and flow steps), never beyond member_reach.  Options
// whose trajectories drift (mc with kt > 1) keep the single bound.
void member_reach_steps(const vbm3d_params& p, unsigned maxflow, unsigned W, unsigned H, int* R0, int* Rs) {
    const int R = member_reach(p, maxflow, W, H);
    if (p.mc && p.kt > 1) {
        *R0 = R;
        *Rs = 0;
        return;
    }
    *R0 = std::min<int>((int)(p.Ns - 1) / 2, R);
    *Rs = (int)std::min<unsigned long long>((p.Npr - 1) / 2 + (unsigned long long)maxflow, (unsigned long long)R);
}

// The aggregation's grid boxes (k_aggregate's layout of AArgs::boxtab):
// per crop, for each tile row and frame distance k the first grid row at
// or below y0 - R_k - (K - 1) and the first beyond y0 + TY - 1 + R_k, then
// the same for the tile columns with the division by their count;
// R_k = min(R, R0 + k Rs).
void box_table(const GridDef& g, unsigned W, unsigned H, int K, unsigned Nf, int R, int R0, int Rs,
               std::vector<int4>& box) {
    const int nty = (int)((H + kAggTileY - 1) / kAggTileY), ntx = (int)((W + kAggTileX - 1) / kAggTileX);
    auto lb = [](const std::vector<unsigned>& v, long x) {
        size_t lo = 0, hi = v.size();
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if ((long)v[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        return (int)lo;
    };
    box.clear();
    box.reserve(g.rows.size() * (size_t)(nty + ntx) * (Nf + 1));
    for (size_t c = 0; c < g.rows.size(); c++) {
        for (int ty = 0; ty < nty; ty++)
            for (unsigned k = 0; k <= Nf; k++) {
                const long Rk = std::min<long>(R, (long)R0 + (long)k * Rs), y0 = (long)ty * kAggTileY;
                box.push_back(make_int4(lb(g.rows[c], y0 - Rk - (K - 1)), lb(g.rows[c], y0 + kAggTileY + Rk), 0, 0));
            }
        for (int tx = 0; tx < ntx; tx++)
            for (unsigned k = 0; k <= Nf; k++) {
                const long Rk = std::min<long>(R, (long)R0 + (long)k * Rs), x0 = (long)tx * kAggTileX;
                const int ja = lb(g.cols[c], x0 - Rk - (K - 1)), jb = lb(g.cols[c], x0 + kAggTileX + Rk);
                const Div32 d = make_div32(jb > ja ? (unsigned)(jb - ja) : 1u);
                box.push_back(make_int4(ja, jb, (int)d.m, (int)(d.s1 | (d.s2 << 8))));
            }
    }
}

}  // namespace

struct ProfEvent {
    int kernel;
    int step;  // 1 or 2; 0 for the colour transforms
    hipEvent_t a, b;
};

struct vbm3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf noisy, basic, den, ff, bf, num_raw, den_raw;  // host-buffer entry points
    // vbm3d_run: colour round trips of the finished noisy and basic videos,
    // copied to the host on the side stream while the later steps compute
    DevBuf noisy_rt, basic_rt;
    hipStream_t side = nullptr;
    hipEvent_t ready[3] = {nullptr, nullptr, nullptr};  // noisy final, basic final, upload chunk
    DevBuf idx, nsx, nsim, dist, est, wgt, tabs, refxy, meta, gpos, rlist, scal, stats, counts;
    DevBuf boxtab;       // the aggregation's per-tile grid boxes
    DevBuf tabsg, mpos;  // general path: tables, members' slice corners
    DevBuf tv;           // TV-L1 flow workspace (flow.hip)
    DevBuf pyr;          // pyramid workspace (pyramid.hip)
    rocblas_handle blas = nullptr;  // the pyramid's DCT products
    size_t est_budget = 0;  // bytes for the per-patch rings; 0 = from free device memory
    unsigned chunk_frames = 0;  // per-clip path: frames per chunk; 0 = from the budget
    // kernel profiling (vbm3d_profile / vbm3d_profile_read)
    bool prof = false;    // bracket launches with events
    bool count = false;   // also accumulate work counters (atomics; not for timed runs)
    std::vector<ProfEvent> events;
    double contributions[3] = {0, 0, 0};  // output samples, per step
    double pix_per_cand[3] = {0, 0, 0};   // k*k*C of each step's matching
    // batched form: rings shared by the clips of a batch, the plan (job
    // arrays and queue prefixes), queue counters, and the captured graph
    DevBuf b_idx, b_nsx, b_est, b_wgt, b_refxy, b_meta, b_gpos, b_tabs, b_plan, b_cnt, b_save;
    hipGraph_t b_graph = nullptr;
    hipGraphExec_t b_exec = nullptr;
    std::string b_key;
    bool b_tuned = false;       // b_tune holds vbm3d_autotune_batch's choice
    vbm3d_batch_config b_tune{};
};

namespace {

struct StepIO {
    int step;
    float sigma;
    const float* noisy;    // colour-transformed
    const float* basic;    // step 2
    const float* ff;
    const float* bf;
    unsigned W, H, F, C;
    const vbm3d_params* p;
    bool cs;
    int nb;
    float* out;            // may be null
    const float* fallback;
    float* num_raw;        // may be null
    float* den_raw;
    unsigned maxflow;
    bool adjoint;          // step-2 adjoint: num_raw receives J^T gy
    const float* gy;
    const float* den;      // the forward pass's denominator
    // Optional: called before each chunk's launches with the number of
    // leading frames the chunk reads (the host-buffer entry uploads the
    // video while step 1 runs); max_chunk caps the chunk (0 = no cap).
    std::function<int(unsigned)> need_frames;
    unsigned max_chunk = 0;
};

// Brackets one launch with events when profiling is on.
template <class F>
hipError_t timed(vbm3d_ctx* ctx, int kernel, int step, hipStream_t s, F&& launch) {
    if (!ctx->prof) return launch();
    ProfEvent ev{kernel, step, nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev.a);
    if (e == hipSuccess) e = hipEventCreate(&ev.b);
    if (e == hipSuccess) e = hipEventRecord(ev.a, s);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(ev.b, s);
    ctx->events.push_back(ev);
    return e;
}

// Bytes the per-patch rings may take: VBM3D_EST_BUDGET_MB when set, else
// 60 % of the device memory that is free or already held by the rings.
size_t ring_budget(const vbm3d_ctx* ctx, size_t held) {
    if (g_arena) return g_arena->size - std::min(g_arena->size, g_arena->used + 4096);
    if (ctx->est_budget) return ctx->est_budget;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)32 << 30;
    return (size_t)((double)(fr + held) * 0.6);
}

int run_step(vbm3d_ctx* ctx, const StepIO& io, hipStream_t s) {
    const vbm3d_params& p = *io.p;
    int rc = check(p, io.W, io.H, io.F, io.C, io.step);
    if (rc) return rc;
    GridDef g;
    if (!build_grid(g, io.W, io.H, p.k, p.p, p.n, io.nb))
        return fail(VBM3D_ERR_ARG, "invalid reference grid for these sizes (ref_grid_nb, p, n)");
    // frame-local ref list in (crop, row, col) order, crop metadata
    std::vector<uint32_t> refxy;
    std::vector<int> meta;
    std::vector<uint16_t> gpos;
    for (size_t c = 0; c < g.rows.size(); c++) {
        meta.push_back((int)refxy.size());
        meta.push_back((int)g.rows[c].size());
        meta.push_back((int)g.cols[c].size());
        meta.push_back((int)gpos.size());
        for (unsigned v : g.rows[c]) gpos.push_back((uint16_t)v);
        meta.push_back((int)gpos.size());
        for (unsigned v : g.cols[c]) gpos.push_back((uint16_t)v);
        for (unsigned y : g.rows[c])
            for (unsigned x : g.cols[c]) refxy.push_back(x | (y << 16));
    }
    const unsigned nrf = (unsigned)refxy.size();
    const unsigned K = p.k, K2 = K * K, C = io.C, N = p.N;
    const unsigned Fr = io.F - p.kt + 1;
    const bool gen = !fast_ok(p, io.W, io.H);  // the general kernels
    if (gen && io.adjoint)
        return fail(VBM3D_ERR_UNSUPPORTED, "the step-2 adjoint runs the fast kernels' parameter sets only");
    const unsigned kt = p.kt, K2t = K2 * kt;
    Tabs tabs;
    fill_tabs(tabs, std::min(K, 8u));
    TabsG tabsg;
    if (gen) fill_tabs_gen(tabsg, K, kt);
    if (g_arena) {
        // the scratch array: this step's buffers start where the call's
        // fixed ones end (the previous step's are no longer read)
        g_arena->used = g_arena->fixed;
        for (DevBuf* b : {&ctx->idx, &ctx->nsx, &ctx->est, &ctx->wgt, &ctx->tabs, &ctx->refxy, &ctx->meta, &ctx->gpos,
                          &ctx->tabsg, &ctx->mpos, &ctx->boxtab})
            *b = DevBuf{};
    }
    HIP_TRY(ensure(ctx->tabs, sizeof(Tabs)));
    if (gen) HIP_TRY(ensure(ctx->tabsg, sizeof(TabsG)));
    HIP_TRY(ensure(ctx->refxy, refxy.size() * 4));
    HIP_TRY(ensure(ctx->meta, meta.size() * 4));
    HIP_TRY(ensure(ctx->gpos, gpos.size() * 2));
    // chunk size from the estimate budget; ring holds S + 2 Nf frames
    // the general path's members span kt frames, so the ring keeps kt - 1
    // more frames of reference patches for the aggregation
    const size_t mpos_words = gen && kt > 1 ? (size_t)N * kt : 0;
    const size_t per_frame = g_arena ? (size_t)nrf * (C * N * K2t + N + 1 + C + mpos_words) * sizeof(float)
                                     : (size_t)nrf * C * N * K2t * sizeof(float);
    const size_t budget = ring_budget(ctx, ctx->est.cap + ctx->idx.cap + ctx->nsx.cap + ctx->wgt.cap);
    size_t cap_frames = std::max<size_t>(1, budget / std::max<size_t>(per_frame, 1));
    unsigned S = (unsigned)std::max<long>(1, (long)cap_frames - 2 * (long)p.Nf - (long)(kt - 1));
    if (ctx->chunk_frames) S = std::min(S, ctx->chunk_frames);
    if (io.max_chunk) S = std::min(S, io.max_chunk);
    unsigned ring = std::min<unsigned>(Fr, S + 2 * p.Nf + (kt - 1));
    if (S >= Fr) { S = Fr; ring = Fr; }
    const size_t slots = (size_t)ring * nrf;
    // slot * kMaxN + n and the `est` plane index slot * C * N + n are 32-bit
    if ((slots * kMaxN * 3) >> 32) return fail(VBM3D_ERR_UNSUPPORTED, "too many reference patches in flight");

    HIP_TRY(ensure(ctx->idx, slots * N * 4));
    HIP_TRY(ensure(ctx->nsx, slots * 4));
    HIP_TRY(ensure(ctx->est, slots * C * N * K2t * 4));
    HIP_TRY(ensure(ctx->wgt, slots * C * 4));
    if (mpos_words) HIP_TRY(ensure(ctx->mpos, slots * mpos_words * 4));
    HIP_TRY(hipMemcpyAsync(ctx->tabs.p, &tabs, sizeof tabs, hipMemcpyHostToDevice, s));
    if (gen) HIP_TRY(hipMemcpyAsync(ctx->tabsg.p, &tabsg, sizeof tabsg, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->refxy.p, refxy.data(), refxy.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->meta.p, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->gpos.p, gpos.data(), gpos.size() * 2, hipMemcpyHostToDevice, s));

    float sig[3] = {0, 0, 0};
    sigma_rows(io.sigma, C, io.cs, sig);

    BMArgs bm{};
    bm.vid = io.step == 1 ? io.noisy : io.basic;
    bm.ff = io.ff;
    bm.bf = io.bf;
    bm.W = io.W; bm.H = io.H; bm.F = io.F; bm.C = C;
    bm.wh = io.W * io.H; bm.whc = bm.wh * C;
    bm.nrf = nrf;
    bm.ref_xy = (const uint32_t*)ctx->refxy.p;
    bm.ref_list = nullptr;
    bm.ring = ring;
    bm.Ns = p.Ns; bm.Npr = p.Npr; bm.Nb = p.Nb; bm.Nf = p.Nf; bm.N = p.N; bm.kt = p.kt;
    bm.d = p.d; bm.tau = p.tau;
    bm.dW = make_div32(io.W);
    bm.dWHC = make_div32(bm.whc);
    bm.idx_out = (uint32_t*)ctx->idx.p;
    bm.dist_out = nullptr;
    bm.nsim_out = nullptr;
    bm.nsx_out = (uint32_t*)ctx->nsx.p;
    bm.band = 1;
    bm.count_out = nullptr;
    if (ctx->count) {
        HIP_TRY(ensure(ctx->counts, (size_t)S * nrf * 2 * 4));
        bm.count_out = (uint32_t*)ctx->counts.p;
        ctx->pix_per_cand[io.step] = (double)K2 * C;
    }

    FArgs fa{};
    fa.noisy = io.noisy;
    fa.basic = io.basic;
    fa.W = io.W; fa.H = io.H; fa.F = io.F; fa.C = C; fa.wh = bm.wh; fa.whc = bm.whc;
    fa.nrf = nrf; fa.ring = ring;
    fa.idx = bm.idx_out; fa.nsx = bm.nsx_out;
    fa.N = p.N; fa.T2D = p.T_2D; fa.T3D = p.T_3D; fa.lambda = p.lambda3D;
    for (int c = 0; c < 3; c++) fa.sig[c] = sig[c];
    fa.tabs = (const Tabs*)ctx->tabs.p;
    fa.est = (float*)ctx->est.p;
    fa.wgt = (float*)ctx->wgt.p;
    fa.gy = io.gy;
    fa.den = io.den;
    fa.dW = bm.dW;
    fa.dWHC = bm.dWHC;

    AArgs ag{};
    ag.W = io.W; ag.H = io.H; ag.F = io.F; ag.C = C; ag.wh = bm.wh; ag.whc = bm.whc;
    ag.K = K; ag.N = N; ag.Nf = p.Nf; ag.kt = p.kt;
    ag.nrf = nrf; ag.ring = ring; ag.Fr = Fr;
    ag.ncrop = (int)g.rows.size();
    ag.crop_meta = (const int*)ctx->meta.p;
    ag.grid_pos = (const uint16_t*)ctx->gpos.p;
    ag.idx = bm.idx_out; ag.nsx = bm.nsx_out;
    ag.est = fa.est; ag.wgt = fa.wgt; ag.tabs = fa.tabs;
    ag.R = member_reach(p, io.maxflow, io.W, io.H);
    member_reach_steps(p, io.maxflow, io.W, io.H, &ag.R0, &ag.Rs);
    if (!gen) {
        // each tile's grid box per frame distance, as the kernel would
        // search it (one load per bound instead of a chain of dependent ones)
        std::vector<int4> box;
        box_table(g, io.W, io.H, (int)K, p.Nf, ag.R, ag.R0, ag.Rs, box);
        HIP_TRY(ensure(ctx->boxtab, box.size() * sizeof(int4)));
        HIP_TRY(hipMemcpyAsync(ctx->boxtab.p, box.data(), box.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        ag.boxtab = (const int4*)ctx->boxtab.p;
        ag.nty = (int)((io.H + kAggTileY - 1) / kAggTileY);
        ag.ntx = (int)((io.W + kAggTileX - 1) / kAggTileX);
    }
    ag.divW = make_div32(io.W);
    ag.fallback = io.fallback;
    ag.out = io.out;
    ag.num_out = io.num_raw;
    ag.den_out = io.den_raw;
    ag.raw = io.adjoint ? 1 : 0;

    GenArgs ga{};
    if (gen) {
        ga.vid = bm.vid;
        ga.noisy = io.noisy;
        ga.basic = io.basic;
        ga.ff = io.ff;
        ga.bf = io.bf;
        ga.W = io.W; ga.H = io.H; ga.F = io.F; ga.C = C; ga.wh = bm.wh; ga.whc = bm.whc;
        ga.nrf = nrf; ga.ring = ring;
        ga.ref_xy = bm.ref_xy;
        ga.K = (int)K; ga.kt = (int)kt; ga.mc = p.mc ? 1 : 0;
        ga.Ns = p.Ns; ga.Npr = p.Npr; ga.Nb = p.Nb; ga.Nf = p.Nf; ga.N = p.N;
        ga.d = p.d; ga.tau = p.tau;
        ga.T2D = p.T_2D; ga.T3D = p.T_3D; ga.lambda = p.lambda3D;
        for (int c = 0; c < 3; c++) ga.sig[c] = sig[c];
        ga.idx_out = bm.idx_out; ga.nsx_out = bm.nsx_out; ga.count_out = bm.count_out;
        ga.tabs = (const TabsG*)ctx->tabsg.p;
        ga.est = fa.est; ga.wgt = fa.wgt;
        ga.mpos = mpos_words ? (uint32_t*)ctx->mpos.p : nullptr;
    }

    unsigned t_done = 0;
    for (unsigned a = 0; a < Fr; a += S) {
        const unsigned e = std::min(Fr, a + S);
        bm.r_begin = fa.r_begin = (long)a * nrf;
        bm.r_end = fa.r_end = (long)e * nrf;
        // matching and filtering read frames up to e - 1 + Nf
        if (io.need_frames) {
            rc = io.need_frames(std::min(io.F, e + p.Nf + p.kt - 1));
            if (rc) return rc;
        }
        ga.r_begin = bm.r_begin;
        ga.r_end = bm.r_end;
        HIP_TRY(timed(ctx, 0, io.step, s, [&] { return gen ? launch_bm_gen(ga, s) : launch_bm((int)K, bm, s); }));
        if (ctx->count)
            HIP_TRY(launch_sum_counts(bm.count_out, bm.r_end - bm.r_begin,
                                      (unsigned long long*)ctx->stats.p + 4 * (io.step - 1), s));
        if (io.adjoint)
            HIP_TRY(launch_filter_adjoint((int)K, fa, s));
        else
            HIP_TRY(timed(ctx, 1, io.step, s, [&] {
                return gen ? launch_filter_gen(io.step, ga, s) : launch_filter((int)K, io.step, fa, s);
            }));
        const unsigned ready = e >= Fr ? io.F : (e > p.Nf ? e - p.Nf : 0);
        if (ready > t_done) {
            ag.t_begin = t_done;
            HIP_TRY(timed(ctx, 2, io.step, s, [&] {
                return gen ? launch_aggregate_gen((int)C, ag, ga.mpos, ga.tabs, ready - t_done, s)
                           : launch_aggregate((int)C, ag, ready - t_done, s);
            }));
            if (ctx->count) ctx->contributions[io.step] += (double)(ready - t_done) * io.W * io.H * C;
            t_done = ready;
        }
    }
    return VBM3D_OK;
}

int flow_absmax(vbm3d_ctx* ctx, const float* ff, const float* bf, unsigned W, unsigned H,
                unsigned F, hipStream_t s, unsigned* out) {
    *out = 0;
    if (!ff && !bf) return VBM3D_OK;
    HIP_TRY(ensure(ctx->scal, 4));
    HIP_TRY(hipMemsetAsync(ctx->scal.p, 0, 4, s));
    const size_t n = (size_t)(F > 1 ? F - 1 : 0) * W * H * 2;
    if (ff) HIP_TRY(launch_flow_absmax(ff, n, (unsigned*)ctx->scal.p, s));
    if (bf) HIP_TRY(launch_flow_absmax(bf, n, (unsigned*)ctx->scal.p, s));
    HIP_TRY(hipMemcpyAsync(out, ctx->scal.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return VBM3D_OK;
}

// run_vbm3d on device buffers (src/vbm3d.cpp:87-376).
// Early outputs of run_device for the host-buffer entry.  The noisy video
// is final once the forward colour transform has run (no step writes it),
// the basic estimate once step 1 has run; each is then copied, given the
// inverse colour transform the end of the run would apply, and `ready`
// recorded, so that its copy to the host can overlap the rest of the run.
// Without the colour round trip nothing is copied: the event marks the
// source buffer itself as final.
struct EarlyOut {
    float* noisy_rt;  // null without the round trip
    float* basic_rt;
    hipEvent_t noisy_ready, basic_ready;
};

// Upload of the noisy video by run_device itself for the host-buffer
// entry: `chunk` frames per copy on the side stream, each followed by its
// forward colour transform on the main stream, issued as step 1 (run in
// chunks of `step1_chunk` frames) reaches frames it has not read yet.
struct HostUpload {
    const float* src;  // the caller's video [F][C][H][W]
    unsigned chunk, step1_chunk;
    hipStream_t side;
    hipEvent_t ev;
};

int run_device(vbm3d_ctx* ctx, float sigma, float* noisy, const float* ff, const float* bf,
               unsigned W, unsigned H, unsigned F, unsigned C, const vbm3d_params* p1,
               const vbm3d_params* p2, int cs, int nb, float* basic, float* den, hipStream_t s,
               const EarlyOut* eo = nullptr, const HostUpload* hu = nullptr) {
    if (!ctx || !noisy || !basic || !den || !p1 || !p2) return fail(VBM3D_ERR_ARG, "null argument");
    if (C != 1 && C != 3) return fail(VBM3D_ERR_ARG, "only 1 or 3 channels are supported");
    if (p1->k > 0) { int rc = check(*p1, W, H, F, C, 1); if (rc) return rc; }
    if (p2->k > 0) { int rc = check(*p2, W, H, F, C, 2); if (rc) return rc; }
    unsigned maxflow = 0;
    int rc = flow_absmax(ctx, ff, bf, W, H, F, s, &maxflow);
    if (rc) return rc;
    if (hu && p1->k == 0) return fail(VBM3D_ERR_ARG, "internal: chunked upload needs step 1");
    if (cs && C == 3 && !hu)
        HIP_TRY(timed(ctx, 3, 0, s, [&] { return launch_color(p1->k > 0 ? noisy : basic, W, H, F, true, s); }));
    // chunked upload: frames [0, up) are on the device and colour-transformed
    unsigned up = 0;
    auto need = [&](unsigned fr) -> int {
        while (up < fr) {
            const unsigned n1 = std::min(F - up, hu->chunk);
            const size_t off = (size_t)up * W * H * C, cnt = (size_t)n1 * W * H * C;
            HIP_TRY(hipMemcpyAsync(noisy + off, hu->src + off, cnt * 4, hipMemcpyHostToDevice, hu->side));
            HIP_TRY(hipEventRecord(hu->ev, hu->side));
            HIP_TRY(hipStreamWaitEvent(s, hu->ev, 0));
            if (cs && C == 3)
                HIP_TRY(timed(ctx, 3, 0, s, [&] { return launch_color(noisy + off, W, H, n1, true, s); }));
            up += n1;
        }
        return VBM3D_OK;
    };
    const bool rt = cs && C == 3;
    auto early = [&](const float* src, float* dst, hipEvent_t ev) -> hipError_t {
        hipError_t e = hipSuccess;
        if (rt) {
            e = hipMemcpyAsync(dst, src, (size_t)W * H * F * C * 4, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = timed(ctx, 3, 0, s, [&] { return launch_color(dst, W, H, F, false, s); });
        }
        return e == hipSuccess ? hipEventRecord(ev, s) : e;
    };
    if (eo && !hu) HIP_TRY(early(noisy, eo->noisy_rt, eo->noisy_ready));
    StepIO io{};
    io.sigma = sigma; io.noisy = noisy; io.ff = ff; io.bf = bf;
    io.W = W; io.H = H; io.F = F; io.C = C; io.cs = cs != 0; io.nb = nb; io.maxflow = maxflow;
    if (p1->k > 0) {
        io.step = 1; io.p = p1; io.basic = nullptr;
        io.out = basic; io.fallback = noisy;  // den == 0 -> noisy (:169, :277)
        if (hu) {
            io.need_frames = need;
            io.max_chunk = hu->step1_chunk;
        }
        rc = run_step(ctx, io, s);
        io.need_frames = nullptr;
        io.max_chunk = 0;
        if (rc) return rc;
    }
    if (hu) {
        rc = need(F);  // every frame, whatever step 1 read
        if (rc) return rc;
        if (eo) HIP_TRY(early(noisy, eo->noisy_rt, eo->noisy_ready));
    }
    if (eo) HIP_TRY(early(basic, eo->basic_rt, eo->basic_ready));
    if (p2->k > 0) {
        io.step = 2; io.p = p2; io.basic = basic;
        io.out = den; io.fallback = nb == 1 ? basic : noisy;  // (:200, :343)
        rc = run_step(ctx, io, s);
        if (rc) return rc;
    } else {
        HIP_TRY(hipMemcpyAsync(den, basic, (size_t)W * H * F * C * 4, hipMemcpyDeviceToDevice, s));
    }
    if (cs && C == 3) {
        HIP_TRY(timed(ctx, 3, 0, s, [&] { return launch_color(den, W, H, F, false, s); }));
        if (!eo) {
            HIP_TRY(timed(ctx, 3, 0, s, [&] { return launch_color(noisy, W, H, F, false, s); }));
            HIP_TRY(timed(ctx, 3, 0, s, [&] { return launch_color(basic, W, H, F, false, s); }));
        }
    }
    if (!eo) HIP_TRY(hipStreamSynchronize(s));  // else the caller synchronises
    return VBM3D_OK;
}

hipError_t upload(DevBuf& b, const float* h, size_t n, hipStream_t s) {
    hipError_t e = ensure(b, n * 4);
    if (e != hipSuccess || !h) return e;
    return hipMemcpyAsync(b.p, h, n * 4, hipMemcpyHostToDevice, s);
}

}  // namespace

// ---------------------------------------------------------------------
// Batched form.  The clips of a batch share the rings; per step, per chunk
// of frames and per kernel class (patch size, channels) there is one
// persistent launch whose workgroups take items -- reference patches for
// matching and filtering, output tiles for aggregation -- from a global
// work queue until it is empty.  Items of one launch are ordered by clip,
// so each clip's per-patch results and each output sample are computed
// exactly as in the per-clip path.
// ---------------------------------------------------------------------
namespace {

struct BatchOp {
    int kind;  // 0 match, 1 filter, 2 aggregate, 3 colour, 4 copy
    int K = 0, C = 0, step = 0;
    size_t jobs = 0, prefix = 0;  // byte offsets in the plan buffer
    int njobs = 0, counter = 0, words = 0;
    long total = 0;
    float* v = nullptr;
    const float* src = nullptr;
    unsigned W = 0, H = 0, F = 0;
    int fwd = 0;
    size_t bytes = 0;
    int shape = 0;  // match: bm_shape shared by every job, else 0; aggregate: largest N
};

struct BatchPlan {
    std::vector<char> blob;  // job arrays and queue prefixes, uploaded to b_plan
    std::vector<BatchOp> ops;
    int counters = 0;
    size_t save_bytes = 0;
};

size_t blob_put(std::vector<char>& blob, const void* p, size_t n) {
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + n);
    std::memcpy(blob.data() + off, p, n);
    return off;
}

// One clip in one step.
struct Member {
    int clip;
    const vbm3d_params* p;
    unsigned K, C, N, nrf, Fr, ring = 0, maxflow = 0;
    size_t per_frame = 0;  // ring bytes per frame of reference patches
    size_t refxy = 0, meta = 0, gpos = 0, tab = 0;
    int ncrop = 0;
    size_t slot_base = 0, idx_base = 0, est_base = 0, wgt_base = 0;
    unsigned t_done = 0;
};

int n_cus(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    return n;
}

// Builds the whole batch as a list of launches.  Allocates the shared
// rings (sized by the estimate budget) and uploads the grids and tables;
// the job arrays go to plan.blob with device pointers already resolved
// against b_plan-independent buffers.
int plan_batch(vbm3d_ctx* ctx, const vbm3d_clip* clips, int n, int nb, const vbm3d_batch_config& cfg,
               BatchPlan& plan, hipStream_t s) {
    plan = BatchPlan();
    // planning may move the shared buffers: a captured graph is stale from here on
    HIP_TRY(hipStreamSynchronize(s));
    if (ctx->b_exec) { (void)hipGraphExecDestroy(ctx->b_exec); ctx->b_exec = nullptr; }
    if (ctx->b_graph) { (void)hipGraphDestroy(ctx->b_graph); ctx->b_graph = nullptr; }
    ctx->b_key.clear();
    std::vector<unsigned> maxflow(n, 0);
    for (int i = 0; i < n; i++) {
        const vbm3d_clip& c = clips[i];
        if (!c.noisy || !c.basic || !c.denoised) return fail(VBM3D_ERR_ARG, "clip " + std::to_string(i) + ": null buffer");
        if (c.C != 1 && c.C != 3) return fail(VBM3D_ERR_ARG, "clip " + std::to_string(i) + ": only 1 or 3 channels");
        for (int st = 1; st <= 2; st++) {
            const vbm3d_params& p = st == 1 ? c.p1 : c.p2;
            if (p.k == 0) continue;
            int rc = check(p, c.W, c.H, c.F, c.C, st);
            if (rc) return fail(rc, "clip " + std::to_string(i) + ": " + g_err);
            if (!fast_ok(p, c.W, c.H))
                return fail(VBM3D_ERR_UNSUPPORTED, "clip " + std::to_string(i) + ", step " + std::to_string(st) +
                                                       ": the batched form runs the fast kernels' parameter sets "
                                                       "(kt 1, k 4..8, N a power of two <= 16, Nb, Ns, Npr <= 8); "
                                                       "run this clip with vbm3d_run_device");
        }
        int rc = flow_absmax(ctx, c.fflow, c.bflow, c.W, c.H, c.F, s, &maxflow[i]);
        if (rc) return rc;
        if (c.C == 3 && c.color_space) plan.save_bytes += (size_t)c.W * c.H * c.F * 3 * 4 * (c.p1.k == 0 ? 2 : 1);
    }
    // host copies of grids and tables for both steps
    std::vector<uint32_t> h_refxy;
    std::vector<int> h_meta;
    std::vector<uint16_t> h_gpos;
    std::vector<Tabs> h_tabs;
    std::vector<std::vector<Member>> steps(3);
    for (int st = 1; st <= 2; st++)
        for (int i = 0; i < n; i++) {
            const vbm3d_clip& c = clips[i];
            const vbm3d_params& p = st == 1 ? c.p1 : c.p2;
            if (p.k == 0) continue;
            GridDef g;
            if (!build_grid(g, c.W, c.H, p.k, p.p, p.n, nb))
                return fail(VBM3D_ERR_ARG, "clip " + std::to_string(i) + ": invalid reference grid");
            Member m{};
            m.clip = i;
            m.p = &p;
            m.K = p.k; m.C = c.C; m.N = p.N;
            m.Fr = c.F - p.kt + 1;
            m.maxflow = maxflow[i];
            m.refxy = h_refxy.size();
            m.meta = h_meta.size();
            m.gpos = h_gpos.size();
            m.ncrop = (int)g.rows.size();
            for (size_t q = 0; q < g.rows.size(); q++) {
                h_meta.push_back((int)(h_refxy.size() - m.refxy));
                h_meta.push_back((int)g.rows[q].size());
                h_meta.push_back((int)g.cols[q].size());
                h_meta.push_back((int)(h_gpos.size() - m.gpos));
                for (unsigned v : g.rows[q]) h_gpos.push_back((uint16_t)v);
                h_meta.push_back((int)(h_gpos.size() - m.gpos));
                for (unsigned v : g.cols[q]) h_gpos.push_back((uint16_t)v);
                for (unsigned y : g.rows[q])
                    for (unsigned x : g.cols[q]) h_refxy.push_back(x | (y << 16));
            }
            m.nrf = (unsigned)(h_refxy.size() - m.refxy);
            m.per_frame = (size_t)m.nrf * (m.C * m.N * m.K * m.K * 4 + m.N * 4 + 4 + m.C * 4);
            Tabs t;
            fill_tabs(t, m.K);
            m.tab = h_tabs.size();
            h_tabs.push_back(t);
            steps[st].push_back(m);
        }
    // shared rings: sub-batches of members whose smallest rings fit the budget,
    // then the largest chunk that fits
    struct Sub { size_t b, e; unsigned S; };
    const size_t budget = ring_budget(ctx, ctx->b_est.cap + ctx->b_idx.cap + ctx->b_nsx.cap + ctx->b_wgt.cap);
    std::vector<std::vector<Sub>> subs(3);
    size_t need_slots = 0, need_idx = 0, need_est = 0, need_wgt = 0;
    for (int st = 1; st <= 2; st++) {
        std::vector<Member>& ms = steps[st];
        size_t i = 0;
        while (i < ms.size()) {
            size_t j = i, used = 0;
            while (j < ms.size()) {
                const vbm3d_params& p = *ms[j].p;
                const size_t r1 = std::min<size_t>(ms[j].Fr, 1 + 2 * p.Nf) * ms[j].per_frame;
                if (j > i && used + r1 > budget) break;
                used += r1;
                j++;
            }
            unsigned maxFr = 0;
            for (size_t q = i; q < j; q++) maxFr = std::max(maxFr, ms[q].Fr);
            auto ring_of = [&](const Member& m, unsigned S) -> unsigned {
                return S >= m.Fr ? m.Fr : std::min<unsigned>(m.Fr, S + 2 * m.p->Nf);
            };
            auto bytes_at = [&](unsigned S) {
                size_t b = 0;
                for (size_t q = i; q < j; q++) b += (size_t)ring_of(ms[q], S) * ms[q].per_frame;
                return b;
            };
            unsigned S = cfg.chunk_frames;
            if (S == 0) {
                S = maxFr;
                while (S > 1 && bytes_at(S) > budget) S--;
            }
            S = std::max(1u, std::min(S, maxFr));
            size_t sb = 0, ib = 0, eb = 0, wb = 0;
            for (size_t q = i; q < j; q++) {
                Member& m = ms[q];
                m.ring = ring_of(m, S);
                m.slot_base = sb;
                m.idx_base = ib;
                m.est_base = eb;
                m.wgt_base = wb;
                const size_t slots = (size_t)m.ring * m.nrf;
                sb += slots;
                ib += slots * m.N;
                eb += slots * m.C * m.N * m.K * m.K;
                wb += slots * m.C;
            }
            if (((size_t)sb * kMaxN * 3) >> 32) return fail(VBM3D_ERR_UNSUPPORTED, "too many reference patches in flight");
            need_slots = std::max(need_slots, sb);
            need_idx = std::max(need_idx, ib);
            need_est = std::max(need_est, eb);
            need_wgt = std::max(need_wgt, wb);
            subs[st].push_back(Sub{i, j, S});
            i = j;
        }
    }
    HIP_TRY(ensure(ctx->b_idx, std::max<size_t>(need_idx, 1) * 4));
    HIP_TRY(ensure(ctx->b_nsx, std::max<size_t>(need_slots, 1) * 4));
    HIP_TRY(ensure(ctx->b_est, std::max<size_t>(need_est, 1) * 4));
    HIP_TRY(ensure(ctx->b_wgt, std::max<size_t>(need_wgt, 1) * 4));
    HIP_TRY(ensure(ctx->b_refxy, std::max<size_t>(h_refxy.size(), 1) * 4));
    HIP_TRY(ensure(ctx->b_meta, std::max<size_t>(h_meta.size(), 1) * 4));
    HIP_TRY(ensure(ctx->b_gpos, std::max<size_t>(h_gpos.size(), 1) * 2));
    HIP_TRY(ensure(ctx->b_tabs, std::max<size_t>(h_tabs.size(), 1) * sizeof(Tabs)));
    if (!h_refxy.empty()) HIP_TRY(hipMemcpyAsync(ctx->b_refxy.p, h_refxy.data(), h_refxy.size() * 4, hipMemcpyHostToDevice, s));
    if (!h_meta.empty()) HIP_TRY(hipMemcpyAsync(ctx->b_meta.p, h_meta.data(), h_meta.size() * 4, hipMemcpyHostToDevice, s));
    if (!h_gpos.empty()) HIP_TRY(hipMemcpyAsync(ctx->b_gpos.p, h_gpos.data(), h_gpos.size() * 2, hipMemcpyHostToDevice, s));
    if (!h_tabs.empty()) HIP_TRY(hipMemcpyAsync(ctx->b_tabs.p, h_tabs.data(), h_tabs.size() * sizeof(Tabs), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host vectors go out of scope

    uint32_t* d_idx = (uint32_t*)ctx->b_idx.p;
    uint32_t* d_nsx = (uint32_t*)ctx->b_nsx.p;
    float* d_est = (float*)ctx->b_est.p;
    float* d_wgt = (float*)ctx->b_wgt.p;
    const uint32_t* d_refxy = (const uint32_t*)ctx->b_refxy.p;
    const int* d_meta = (const int*)ctx->b_meta.p;
    const uint16_t* d_gpos = (const uint16_t*)ctx->b_gpos.p;
    const Tabs* d_tabs = (const Tabs*)ctx->b_tabs.p;

    auto colour = [&](float* v, const vbm3d_clip& c, bool fwd) {
        BatchOp op;
        op.kind = 3; op.v = v; op.W = c.W; op.H = c.H; op.F = c.F; op.fwd = fwd ? 1 : 0;
        plan.ops.push_back(op);
    };
    for (int i = 0; i < n; i++)
        if (clips[i].C == 3 && clips[i].color_space) colour(clips[i].p1.k > 0 ? clips[i].noisy : clips[i].basic, clips[i], true);

    for (int st = 1; st <= 2; st++) {
        std::vector<Member>& ms = steps[st];
        for (const Sub& sub : subs[st]) {
            unsigned maxFr = 0;
            for (size_t q = sub.b; q < sub.e; q++) maxFr = std::max(maxFr, ms[q].Fr);
            for (unsigned a = 0; a < maxFr; a += sub.S) {
                // per class (K, C): matching and filtering jobs of this chunk
                struct Cls { int K, C; std::vector<BMArgs> bm; std::vector<FArgs> fa; std::vector<long> pre; int wb = 0, wf = 0; };
                std::vector<Cls> cls;
                std::vector<AArgs> ag1, ag3;
                std::vector<long> pre1{0}, pre3{0};
                for (size_t q = sub.b; q < sub.e; q++) {
                    Member& m = ms[q];
                    const vbm3d_clip& c = clips[m.clip];
                    const vbm3d_params& p = *m.p;
                    float sig[3] = {0, 0, 0};
                    sigma_rows(c.sigma, c.C, c.color_space != 0, sig);
                    const unsigned wh = c.W * c.H, whc = wh * c.C;
                    if (a < m.Fr) {
                        const unsigned e = std::min(m.Fr, a + sub.S);
                        BMArgs bm{};
                        bm.vid = st == 1 ? c.noisy : c.basic;
                        bm.ff = c.fflow; bm.bf = c.bflow;
                        bm.W = c.W; bm.H = c.H; bm.F = c.F; bm.C = c.C; bm.wh = wh; bm.whc = whc;
                        bm.nrf = m.nrf;
                        bm.ref_xy = d_refxy + m.refxy;
                        bm.ref_list = nullptr;
                        bm.r_begin = (long)a * m.nrf;
                        bm.r_end = (long)e * m.nrf;
                        bm.ring = m.ring;
                        bm.Ns = p.Ns; bm.Npr = p.Npr; bm.Nb = p.Nb; bm.Nf = p.Nf; bm.N = p.N; bm.kt = p.kt;
                        bm.d = p.d; bm.tau = p.tau;
                        bm.dW = make_div32(c.W);
                        bm.dWHC = make_div32(whc);
                        bm.idx_out = d_idx + m.idx_base;
                        bm.dist_out = nullptr;
                        bm.nsim_out = nullptr;
                        bm.nsx_out = d_nsx + m.slot_base;
                        bm.count_out = nullptr;
                        FArgs fa{};
                        fa.noisy = c.noisy;
                        fa.basic = c.basic;
                        fa.W = c.W; fa.H = c.H; fa.F = c.F; fa.C = c.C; fa.wh = wh; fa.whc = whc;
                        fa.nrf = m.nrf; fa.ring = m.ring;
                        fa.r_begin = bm.r_begin; fa.r_end = bm.r_end;
                        fa.idx = bm.idx_out; fa.nsx = bm.nsx_out;
                        fa.N = p.N; fa.T2D = p.T_2D; fa.T3D = p.T_3D; fa.lambda = p.lambda3D;
                        for (int k = 0; k < 3; k++) fa.sig[k] = sig[k];
                        fa.tabs = d_tabs + m.tab;
                        fa.est = d_est + m.est_base;
                        fa.wgt = d_wgt + m.wgt_base;
                        fa.gy = nullptr; fa.den = nullptr;
                        fa.dW = bm.dW; fa.dWHC = bm.dWHC;
                        Cls* cl = nullptr;
                        for (Cls& x : cls)
                            if (x.K == (int)m.K && x.C == (int)m.C) cl = &x;
                        if (!cl) {
                            cls.push_back(Cls{(int)m.K, (int)m.C, {}, {}, {0}});
                            cl = &cls.back();
                        }
                        cl->bm.push_back(bm);
                        cl->fa.push_back(fa);
                        cl->pre.push_back(cl->pre.back() + (bm.r_end - bm.r_begin));
                        cl->wb = std::max(cl->wb, bm_lds_words((int)m.K, (int)m.C, p.Ns, p.Npr, p.Nb, p.Nf));
                        cl->wf = std::max(cl->wf, filter_words((int)m.K, (int)m.C, p.N, st));
                    }
                    const unsigned e = std::min(m.Fr, a + sub.S);
                    const unsigned ready = e >= m.Fr ? c.F : (e > p.Nf ? e - p.Nf : 0);
                    if (ready > m.t_done) {
                        AArgs ag{};
                        ag.W = c.W; ag.H = c.H; ag.F = c.F; ag.C = c.C; ag.wh = wh; ag.whc = whc;
                        ag.K = m.K; ag.N = p.N; ag.Nf = p.Nf; ag.kt = p.kt;
                        ag.nrf = m.nrf; ag.ring = m.ring; ag.Fr = m.Fr;
                        ag.ncrop = m.ncrop;
                        ag.crop_meta = d_meta + m.meta;
                        ag.grid_pos = d_gpos + m.gpos;
                        ag.idx = d_idx + m.idx_base;
                        ag.nsx = d_nsx + m.slot_base;
                        ag.est = d_est + m.est_base;
                        ag.wgt = d_wgt + m.wgt_base;
                        ag.tabs = d_tabs + m.tab;
                        ag.R = member_reach(p, m.maxflow, c.W, c.H);
                        member_reach_steps(p, m.maxflow, c.W, c.H, &ag.R0, &ag.Rs);
                        ag.divW = make_div32(c.W);
                        ag.t_begin = m.t_done;
                        ag.fallback = st == 1 ? c.noisy : (nb == 1 ? c.basic : c.noisy);
                        ag.out = st == 1 ? c.basic : c.denoised;
                        ag.num_out = nullptr;
                        ag.den_out = nullptr;
                        ag.raw = 0;
                        const long tiles = (long)((c.W + kAggTileX - 1) / kAggTileX) * ((c.H + kAggTileY - 1) / kAggTileY) *
                                           (ready - m.t_done);
                        if (c.C == 1) { ag1.push_back(ag); pre1.push_back(pre1.back() + tiles); }
                        else { ag3.push_back(ag); pre3.push_back(pre3.back() + tiles); }
                        m.t_done = ready;
                    }
                }
                for (Cls& cl : cls) {
                    BatchOp op;
                    op.kind = 0; op.K = cl.K; op.C = cl.C; op.step = st;
                    op.jobs = blob_put(plan.blob, cl.bm.data(), cl.bm.size() * sizeof(BMArgs));
                    op.prefix = blob_put(plan.blob, cl.pre.data(), cl.pre.size() * sizeof(long));
                    op.njobs = (int)cl.bm.size(); op.total = cl.pre.back(); op.words = cl.wb;
                    op.shape = 1;
                    for (const BMArgs& b : cl.bm)
                        if (bm_shape(b) != 1) op.shape = 0;
                    op.counter = plan.counters++;
                    plan.ops.push_back(op);
                }
                for (Cls& cl : cls) {
                    BatchOp op;
                    op.kind = 1; op.K = cl.K; op.C = cl.C; op.step = st;
                    op.jobs = blob_put(plan.blob, cl.fa.data(), cl.fa.size() * sizeof(FArgs));
                    op.prefix = blob_put(plan.blob, cl.pre.data(), cl.pre.size() * sizeof(long));
                    op.njobs = (int)cl.fa.size(); op.total = cl.pre.back(); op.words = cl.wf;
                    op.counter = plan.counters++;
                    plan.ops.push_back(op);
                }
                for (int cc : {1, 3}) {
                    std::vector<AArgs>& v = cc == 1 ? ag1 : ag3;
                    std::vector<long>& pre = cc == 1 ? pre1 : pre3;
                    if (v.empty()) continue;
                    BatchOp op;
                    op.kind = 2; op.C = cc; op.step = st;
                    op.shape = 1;  // largest group size of the jobs
                    for (const AArgs& g : v) op.shape = std::max(op.shape, g.N);
                    op.jobs = blob_put(plan.blob, v.data(), v.size() * sizeof(AArgs));
                    op.prefix = blob_put(plan.blob, pre.data(), pre.size() * sizeof(long));
                    op.njobs = (int)v.size(); op.total = pre.back();
                    op.counter = plan.counters++;
                    plan.ops.push_back(op);
                }
            }
        }
        if (st == 2)
            for (int i = 0; i < n; i++)
                if (clips[i].p2.k == 0) {
                    BatchOp op;
                    op.kind = 4; op.v = clips[i].denoised; op.src = clips[i].basic;
                    op.bytes = (size_t)clips[i].W * clips[i].H * clips[i].F * clips[i].C * 4;
                    plan.ops.push_back(op);
                }
    }
    for (int i = 0; i < n; i++)
        if (clips[i].C == 3 && clips[i].color_space) {
            colour(clips[i].denoised, clips[i], false);
            colour(clips[i].noisy, clips[i], false);
            colour(clips[i].basic, clips[i], false);
        }
    HIP_TRY(ensure(ctx->b_plan, std::max<size_t>(plan.blob.size(), 1)));
    if (!plan.blob.empty()) HIP_TRY(hipMemcpy(ctx->b_plan.p, plan.blob.data(), plan.blob.size(), hipMemcpyHostToDevice));
    HIP_TRY(ensure(ctx->b_cnt, std::max<size_t>(plan.counters, 1) * sizeof(unsigned long long)));
    return VBM3D_OK;
}

// Enqueues the plan on s (also what a graph capture records).
int exec_batch(vbm3d_ctx* ctx, const BatchPlan& plan, const vbm3d_batch_config& cfg, hipStream_t s) {
    char* blob = (char*)ctx->b_plan.p;
    unsigned long long* cnt = (unsigned long long*)ctx->b_cnt.p;
    const int cus = n_cus(ctx->device);
    HIP_TRY(launch_zero_u64(cnt, plan.counters, s));
    for (const BatchOp& op : plan.ops) {
        if (op.kind <= 2) {
            WorkQueue q;
            q.counter = cnt + op.counter;
            q.item_begin = (const long*)(blob + op.prefix);
            q.njobs = op.njobs;
            q.total = op.total;
            q.grab = op.kind == 2 ? 1 : std::max(1, cfg.grab);
            if (op.kind == 2) {
                if (cfg.persistent) {
                    const long blocks = std::min<long>(op.total, (long)cfg.groups_per_cu * cus);
                    HIP_TRY(launch_aggregate_queue(op.C, (const AArgs*)(blob + op.jobs), q, (int)blocks, op.shape, s));
                } else {
                    HIP_TRY(launch_aggregate_multi(op.C, (const AArgs*)(blob + op.jobs), q, op.shape, s));
                }
                continue;
            }
            int waves = cfg.waves_per_group > 0 ? std::min(cfg.waves_per_group, 4) : 4;
            while (waves > 1 && (size_t)op.words * 4 * waves > 64 * 1024) waves--;
            if ((size_t)op.words * 4 > 64 * 1024) return fail(VBM3D_ERR_UNSUPPORTED, "patch working set above 64 KB of LDS");
            const long need = (op.total + (long)q.grab * waves - 1) / ((long)q.grab * waves);
            const long blocks = std::min<long>(need, (long)cfg.groups_per_cu * cus);
            if (!cfg.persistent) {
                if (op.kind == 0)
                    HIP_TRY(launch_bm_multi(op.K, op.C, (const BMArgs*)(blob + op.jobs), q, op.words, waves, op.shape, s));
                else
                    HIP_TRY(launch_filter_multi(op.K, op.C, op.step, (const FArgs*)(blob + op.jobs), q, op.words,
                                                waves, s));
            } else if (op.kind == 0) {
                HIP_TRY(launch_bm_queue(op.K, op.C, (const BMArgs*)(blob + op.jobs), q, op.words, waves, (int)blocks,
                                            op.shape, s));
            } else {
                HIP_TRY(launch_filter_queue(op.K, op.C, op.step, (const FArgs*)(blob + op.jobs), q, op.words, waves,
                                            (int)blocks, s));
            }
        } else if (op.kind == 3) {
            HIP_TRY(launch_color(op.v, op.W, op.H, op.F, op.fwd != 0, s));
        } else {
            HIP_TRY(hipMemcpyAsync(op.v, op.src, op.bytes, hipMemcpyDeviceToDevice, s));
        }
    }
    return VBM3D_OK;
}

vbm3d_batch_config default_batch_config() {
    vbm3d_batch_config c;
    c.chunk_frames = 0;
    c.waves_per_group = 0;
    c.grab = 1;
    c.groups_per_cu = 8;
    c.use_graph = 1;
    c.persistent = 0;
    return c;
}

int check_batch_config(const vbm3d_batch_config& c) {
    if (c.waves_per_group < 0 || c.waves_per_group > 4) return fail(VBM3D_ERR_ARG, "waves_per_group must be in [0, 4]");
    if (c.grab < 1 || c.grab > 64) return fail(VBM3D_ERR_ARG, "grab must be in [1, 64]");
    if (c.groups_per_cu < 1 || c.groups_per_cu > 64) return fail(VBM3D_ERR_ARG, "groups_per_cu must be in [1, 64]");
    if (c.persistent != 0 && c.persistent != 1) return fail(VBM3D_ERR_ARG, "persistent must be 0 or 1");
    return VBM3D_OK;
}

std::string batch_key(const vbm3d_clip* clips, int n, int nb, const vbm3d_batch_config& cfg) {
    std::string k((const char*)clips, (size_t)n * sizeof(vbm3d_clip));
    k.append((const char*)&nb, sizeof nb);
    k.append((const char*)&cfg, sizeof cfg);
    return k;
}

}  // namespace

// ---------------------------------------------------------------------
// vbm3d_run_device_noalloc: the device-buffer entry without allocation.
// Its per-patch rings, tables and flow maximum live in `scratch`, a
// file-scope device array that every GPU holds from the library's first
// use on it until the process ends.  Host code never passes `scratch`
// itself (its name is a host-side stand-in, not a device address): it
// takes the array's address per device with hipGetSymbolAddress once and
// uses that.
// ---------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)256 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;  // device address of `scratch`, per device

// Address of this device's `scratch` (dev must be current), looked up on
// the first call on the device.
int scratch_address(int dev, char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0) return fail(VBM3D_ERR_ARG, "no current device");
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VBM3D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (char*)p;
    }
    *out = g_scratch_ptr[dev];
    return VBM3D_OK;
}

// Scratch one step needs at its smallest chunking (rings of min(Fr,
// 2 Nf + 1) frames), as run_step carves it; 0 when the grid is invalid
// (run_step then reports the error).
size_t step_scratch_bytes(const vbm3d_params& p, unsigned W, unsigned H, unsigned F, unsigned C, int nb) {
    GridDef g;
    if (p.k == 0 || F < p.kt || !build_grid(g, W, H, p.k, p.p, p.n, nb)) return 0;
    size_t nrf = 0, npos = 0;
    for (size_t c = 0; c < g.rows.size(); c++) {
        nrf += g.rows[c].size() * g.cols[c].size();
        npos += g.rows[c].size() + g.cols[c].size();
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const bool gen = !fast_ok(p, W, H);
    const size_t Fr = F - p.kt + 1, ring = std::min<size_t>(Fr, 1 + 2 * (size_t)p.Nf + (p.kt - 1)), slots = ring * nrf;
    const size_t mpos_words = gen && p.kt > 1 ? (size_t)p.N * p.kt : 0;
    const size_t nbox = gen ? 0
                            : g.rows.size() * (size_t)((H + kAggTileY - 1) / kAggTileY + (W + kAggTileX - 1) / kAggTileX) *
                                  (p.Nf + 1);
    return al(sizeof(Tabs)) + (gen ? al(sizeof(TabsG)) : 0) + al(nrf * 4) + al(g.rows.size() * 5 * 4) + al(npos * 2) +
           al(nbox * sizeof(int4)) + slots * ((size_t)C * p.N * p.k * p.k * p.kt + p.N + 1 + C + mpos_words) * 4 + 4096;
}

}  // namespace

extern "C" {

const char* vbm3d_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// Context access for the flow entry points (flow.hip).
namespace vb {
hipStream_t ctx_stream(vbm3d_ctx* c) { return c->stream; }
int ctx_device(vbm3d_ctx* c) { return c->device; }
hipError_t ctx_flow_workspace(vbm3d_ctx* c, size_t bytes, void** p) {
    const hipError_t e = ensure(c->tv, bytes);
    *p = c->tv.p;
    return e;
}
int ctx_fail(int code, const std::string& msg) { return fail(code, msg); }
hipError_t ctx_pyr_workspace(vbm3d_ctx* c, size_t bytes, void** p) {
    const hipError_t e = ensure(c->pyr, bytes);
    *p = c->pyr.p;
    return e;
}
rocblas_handle ctx_blas(vbm3d_ctx* c) {
    if (!c->blas) {
        if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
            c->blas = nullptr;
            return nullptr;
        }
    }
    if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) return nullptr;
    return c->blas;
}
}  // namespace vb

extern "C" {

vbm3d_ctx* vbm3d_create(int device) {
    if (hipSetDevice(device) != hipSuccess) {
        g_err = "hipSetDevice failed";
        return nullptr;
    }
    vbm3d_ctx* c = new vbm3d_ctx();
    c->device = device;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        g_err = "hipStreamCreate failed";
        delete c;
        return nullptr;
    }
    if (const char* e = std::getenv("VBM3D_EST_BUDGET_MB")) c->est_budget = (size_t)std::atol(e) << 20;
    return c;
}

void vbm3d_destroy(vbm3d_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (DevBuf* b : {&c->noisy, &c->basic, &c->den, &c->ff, &c->bf, &c->num_raw, &c->den_raw, &c->noisy_rt, &c->basic_rt,
                      &c->idx, &c->nsx, &c->nsim, &c->dist, &c->est, &c->wgt, &c->tabs,
                      &c->refxy, &c->meta, &c->gpos, &c->rlist, &c->scal, &c->stats, &c->counts,
                      &c->b_idx, &c->b_nsx, &c->b_est, &c->b_wgt, &c->b_refxy, &c->b_meta, &c->b_gpos,
                      &c->b_tabs, &c->b_plan, &c->b_cnt, &c->b_save, &c->tabsg, &c->mpos, &c->tv, &c->pyr,
                      &c->boxtab})
        if (b->p) (void)hipFree(b->p);
    if (c->b_exec) (void)hipGraphExecDestroy(c->b_exec);
    if (c->b_graph) (void)hipGraphDestroy(c->b_graph);
    for (hipEvent_t e : c->ready)
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->blas) (void)rocblas_destroy_handle(c->blas);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void vbm3d_default_params_1(vbm3d_params* p, int k, int kt, int Nf, int Ns, int Npr, int Nb,
                            int pstep, int N, int d, float tau, float lambda3D, unsigned T_2D,
                            unsigned T_3D, float sigma) {
    // initializeParameters_1 (src/main.cpp:35-125)
    p->k = k < 0 ? 8 : k;
    p->kt = kt < 0 ? 1 : kt;
    p->Nf = Nf < 0 ? 4 : Nf;
    p->Ns = Ns < 0 ? 7 : Ns;
    p->Npr = Npr < 0 ? 5 : Npr;
    p->Nb = Nb < 0 ? 2 : Nb;
    p->d = (float)(d < 0 ? (7. * 7. * 255.) / (p->k * p->k * p->kt) : (d * d * 255.) / (p->k * p->k * p->kt));
    p->p = pstep < 0 ? p->k / 4 * 3 : pstep;
    p->N = N < 0 ? 8 : N;
    p->tau = tau < 0 ? ((sigma > 30) ? 4500.f : 3000.f) : tau;
    p->lambda3D = lambda3D < 0 ? 2.7f : lambda3D;
    p->T_2D = T_2D == 0xFFFFFFFFu ? (p->k == 8 ? VBM3D_BIOR : VBM3D_DCT) : (p->k == 8 ? T_2D : VBM3D_DCT);
    p->T_3D = T_3D == 0xFFFFFFFFu ? VBM3D_HAAR : T_3D;
    p->mc = 0;
    p->n = 16;
}

void vbm3d_default_params_2(vbm3d_params* p, int k, int kt, int Nf, int Ns, int Npr, int Nb,
                            int pstep, int N, int d, float tau, unsigned T_2D, unsigned T_3D,
                            float sigma) {
    // initializeParameters_2 (src/main.cpp:127-207)
    p->k = k < 0 ? ((sigma > 30) ? 8 : 7) : k;
    p->kt = kt < 0 ? 1 : kt;
    p->Nf = Nf < 0 ? 4 : Nf;
    p->Ns = Ns < 0 ? 7 : Ns;
    p->Npr = Npr < 0 ? 5 : Npr;
    p->Nb = Nb < 0 ? 2 : Nb;
    p->d = (float)(d < 0 ? (3. * 3. * 255.) / (p->k * p->k * p->kt) : (d * d * 255.) / (p->k * p->k * p->kt));
    p->p = pstep < 0 ? p->k / 2 : pstep;
    p->N = N < 0 ? 8 : N;
    p->tau = tau < 0 ? ((sigma > 30) ? 3000.f : 1500.f) : tau;
    p->lambda3D = 0.f;  // src/parameters.h:51
    p->T_2D = T_2D == 0xFFFFFFFFu ? VBM3D_DCT : (p->k == 8 ? T_2D : VBM3D_DCT);
    p->T_3D = T_3D == 0xFFFFFFFFu ? VBM3D_HAAR : T_3D;
    p->mc = 0;
    p->n = 16;
}

long vbm3d_grid(unsigned W, unsigned H, unsigned k, unsigned p, int n, int nb, unsigned* xs,
                unsigned* ys, unsigned* crop) {
    GridDef g;
    if (!build_grid(g, W, H, k, p, n, nb)) return -1;
    long cnt = 0;
    for (size_t c = 0; c < g.rows.size(); c++)
        for (unsigned y : g.rows[c])
            for (unsigned x : g.cols[c]) {
                if (xs) xs[cnt] = x;
                if (ys) ys[cnt] = y;
                if (crop) crop[cnt] = (unsigned)c;
                cnt++;
            }
    return cnt;
}

int vbm3d_run_device_noalloc(float sigma, float* d_noisy, const float* d_fflow, const float* d_bflow,
                             unsigned W, unsigned H, unsigned F, unsigned C, const vbm3d_params* p1,
                             const vbm3d_params* p2, int color_space, int ref_grid_nb, float* d_basic,
                             float* d_denoised, void* stream) {
    if (!d_noisy || !d_basic || !d_denoised || !p1 || !p2) return fail(VBM3D_ERR_ARG, "null argument");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    char* base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc) return rc;
    size_t need = 256;  // the flow maximum
    for (int st = 1; st <= 2; st++) {
        const vbm3d_params& p = st == 1 ? *p1 : *p2;
        if (p.k > 0) need = std::max(need, 256 + step_scratch_bytes(p, W, H, F, C, ref_grid_nb));
    }
    if (need > kScratchBytes) {
        char b[200];
        std::snprintf(b, sizeof b, "this clip needs %.1f MiB of scratch; vbm3d_run_device_noalloc has %zu MiB",
                      (double)need / (1 << 20), kScratchBytes >> 20);
        return fail(VBM3D_ERR_UNSUPPORTED, b);
    }
    vbm3d_ctx ctx;  // no stream and no buffers of its own
    ctx.device = dev;
    Arena arena;
    arena.base = base;
    arena.size = kScratchBytes;
    g_arena = &arena;
    hipError_t e = ensure(ctx.scal, 4);  // kept across both steps
    arena.fixed = arena.used;
    rc = e == hipSuccess ? run_device(&ctx, sigma, d_noisy, d_fflow, d_bflow, W, H, F, C, p1, p2, color_space,
                                      ref_grid_nb, d_basic, d_denoised, (hipStream_t)stream)
                         : fail(VBM3D_ERR_HIP, "scratch");
    g_arena = nullptr;
    if (rc && arena.overflow) return fail(VBM3D_ERR_UNSUPPORTED, "the clip does not fit the no-allocation scratch");
    return rc;
}

int vbm3d_run_device(vbm3d_ctx* ctx, float sigma, float* d_noisy, const float* d_fflow,
                     const float* d_bflow, unsigned W, unsigned H, unsigned F, unsigned C,
                     const vbm3d_params* p1, const vbm3d_params* p2, int color_space,
                     int ref_grid_nb, float* d_basic, float* d_denoised, void* stream) {
    if (!ctx) return fail(VBM3D_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return run_device(ctx, sigma, d_noisy, d_fflow, d_bflow, W, H, F, C, p1, p2, color_space,
                      ref_grid_nb, d_basic, d_denoised, s);
}

int vbm3d_run(vbm3d_ctx* ctx, float sigma, float* noisy, const float* fflow, const float* bflow,
              unsigned W, unsigned H, unsigned F, unsigned C, const vbm3d_params* p1,
              const vbm3d_params* p2, int color_space, int ref_grid_nb, float* basic,
              float* denoised) {
    if (!ctx || !noisy || !basic || !denoised || !p1 || !p2) return fail(VBM3D_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // a one-frame clip has no predictive search: its flows are never read
    if (F <= 1) fflow = bflow = nullptr;
    const size_t n = (size_t)W * H * F * C, nf = (size_t)(F > 1 ? F - 1 : 0) * W * H * 2;
    // With step 1 on, the noisy video is uploaded in chunks while step 1
    // runs on the frames already there (run_device's HostUpload).
    const bool chunked = p1->k > 0;
    if (chunked) HIP_TRY(ensure(ctx->noisy, n * 4));
    else HIP_TRY(upload(ctx->noisy, noisy, n, s));
    HIP_TRY(upload(ctx->basic, p1->k > 0 ? nullptr : basic, n, s));
    HIP_TRY(ensure(ctx->den, n * 4));
    if (fflow) HIP_TRY(upload(ctx->ff, fflow, nf, s));
    if (bflow) HIP_TRY(upload(ctx->bf, bflow, nf, s));
    // The noisy video changes only through the colour round trip; the
    // noisy and basic videos go to the host on the side stream as soon as
    // they are final, while the later steps compute.
    const bool rt = color_space && C == 3;
    if (rt) {
        HIP_TRY(ensure(ctx->noisy_rt, n * 4));
        HIP_TRY(ensure(ctx->basic_rt, n * 4));
    }
    if (!ctx->side) HIP_TRY(hipStreamCreate(&ctx->side));
    for (hipEvent_t& e : ctx->ready)
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const EarlyOut eo{rt ? (float*)ctx->noisy_rt.p : nullptr, rt ? (float*)ctx->basic_rt.p : nullptr,
                      ctx->ready[0], ctx->ready[1]};
    // 4 frames per copy; step 1 in chunks of 8 reference frames, so that
    // its first launches wait for 8 + Nf frames only
    const HostUpload hu{noisy, 4, 8, ctx->side, ctx->ready[2]};
    int rc = run_device(ctx, sigma, (float*)ctx->noisy.p, fflow ? (float*)ctx->ff.p : nullptr,
                        bflow ? (float*)ctx->bf.p : nullptr, W, H, F, C, p1, p2, color_space,
                        ref_grid_nb, (float*)ctx->basic.p, (float*)ctx->den.p, s, &eo, chunked ? &hu : nullptr);
    if (rc) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamSynchronize(s);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (rt) {
        e = hipStreamWaitEvent(ctx->side, eo.noisy_ready, 0);
        if (e == hipSuccess) e = hipMemcpyAsync(noisy, eo.noisy_rt, n * 4, hipMemcpyDeviceToHost, ctx->side);
    }
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->side, eo.basic_ready, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(basic, rt ? eo.basic_rt : ctx->basic.p, n * 4, hipMemcpyDeviceToHost, ctx->side);
    if (e == hipSuccess) e = hipMemcpyAsync(denoised, ctx->den.p, n * 4, hipMemcpyDeviceToHost, s);
    const hipError_t e1 = hipStreamSynchronize(ctx->side), e2 = hipStreamSynchronize(s);
    HIP_TRY(e);
    HIP_TRY(e1);
    HIP_TRY(e2);
    return VBM3D_OK;
}

int vbm3d_step(vbm3d_ctx* ctx, int step, float sigma, const float* noisy, const float* basic,
               const float* fflow, const float* bflow, unsigned W, unsigned H, unsigned F,
               unsigned C, const vbm3d_params* prm, int color_space, int ref_grid_nb, float* num,
               float* den) {
    if (!ctx || !noisy || !prm || !num || !den || (step == 2 && !basic) || (step != 1 && step != 2))
        return fail(VBM3D_ERR_ARG, "invalid argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // a one-frame clip has no predictive search: its flows are never read
    if (F <= 1) fflow = bflow = nullptr;
    const size_t n = (size_t)W * H * F * C, nf = (size_t)(F > 1 ? F - 1 : 0) * W * H * 2;
    int rc = check(*prm, W, H, F, C, step);
    if (rc) return rc;
    HIP_TRY(upload(ctx->noisy, noisy, n, s));
    if (step == 2) HIP_TRY(upload(ctx->basic, basic, n, s));
    if (fflow) HIP_TRY(upload(ctx->ff, fflow, nf, s));
    if (bflow) HIP_TRY(upload(ctx->bf, bflow, nf, s));
    HIP_TRY(ensure(ctx->num_raw, n * 4));
    HIP_TRY(ensure(ctx->den_raw, n * 4));
    const float* dff = fflow ? (const float*)ctx->ff.p : nullptr;
    const float* dbf = bflow ? (const float*)ctx->bf.p : nullptr;
    unsigned maxflow = 0;
    rc = flow_absmax(ctx, dff, dbf, W, H, F, s, &maxflow);
    if (rc) return rc;
    StepIO io{};
    io.step = step; io.sigma = sigma;
    io.noisy = (const float*)ctx->noisy.p;
    io.basic = step == 2 ? (const float*)ctx->basic.p : nullptr;
    io.ff = dff; io.bf = dbf;
    io.W = W; io.H = H; io.F = F; io.C = C; io.p = prm; io.cs = color_space != 0; io.nb = ref_grid_nb;
    io.out = nullptr; io.fallback = nullptr;
    io.num_raw = (float*)ctx->num_raw.p;
    io.den_raw = (float*)ctx->den_raw.p;
    io.maxflow = maxflow;
    rc = run_step(ctx, io, s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(num, ctx->num_raw.p, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(den, ctx->den_raw.p, n * 4, hipMemcpyDeviceToHost));
    return VBM3D_OK;
}

int vbm3d_step2_adjoint(vbm3d_ctx* ctx, float sigma, const float* d_noisy, const float* d_basic,
                        const float* d_fflow, const float* d_bflow, unsigned W, unsigned H,
                        unsigned F, unsigned C, const vbm3d_params* p2, int color_space,
                        int ref_grid_nb, const float* d_grad_out, float* d_grad_noisy,
                        float* d_final, void* stream) {
    if (!ctx || !d_noisy || !d_basic || !p2 || !d_grad_out || !d_grad_noisy)
        return fail(VBM3D_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    int rc = check(*p2, W, H, F, C, 2);
    if (rc) return rc;
    const size_t n = (size_t)W * H * F * C;
    HIP_TRY(ensure(ctx->num_raw, n * 4));
    HIP_TRY(ensure(ctx->den_raw, n * 4));
    unsigned maxflow = 0;
    rc = flow_absmax(ctx, d_fflow, d_bflow, W, H, F, s, &maxflow);
    if (rc) return rc;
    StepIO io{};
    io.step = 2; io.sigma = sigma;
    io.noisy = d_noisy; io.basic = d_basic; io.ff = d_fflow; io.bf = d_bflow;
    io.W = W; io.H = H; io.F = F; io.C = C; io.p = p2; io.cs = color_space != 0; io.nb = ref_grid_nb;
    io.maxflow = maxflow;
    // forward pass: the denominator (and, if asked, the output)
    io.out = d_final;
    io.fallback = ref_grid_nb == 1 ? d_basic : d_noisy;
    io.num_raw = (float*)ctx->num_raw.p;
    io.den_raw = (float*)ctx->den_raw.p;
    rc = run_step(ctx, io, s);
    if (rc) return rc;
    // transposed pass: the same groups (recomputed), gathered gradient
    io.adjoint = true;
    io.gy = d_grad_out;
    io.den = (const float*)ctx->den_raw.p;
    io.out = nullptr;
    io.fallback = nullptr;
    io.num_raw = d_grad_noisy;
    io.den_raw = nullptr;
    rc = run_step(ctx, io, s);
    if (rc) return rc;
    // den == 0 keeps the noisy value when the grid is tiled (src/vbm3d.cpp:343)
    if (ref_grid_nb > 1) HIP_TRY(launch_adjoint_fallback(d_grad_noisy, d_grad_out, io.den, n, s));
    HIP_TRY(hipStreamSynchronize(s));
    return VBM3D_OK;
}

int vbm3d_block_match(vbm3d_ctx* ctx, const float* vid, unsigned W, unsigned H, unsigned F,
                      unsigned C, const float* fflow, const float* bflow,
                      const vbm3d_params* prm, long nrefs, const uint32_t* ref_idx,
                      uint32_t* idx_out, float* dist_out, uint32_t* nsim_out, uint32_t* nsx_out) {
    if (!ctx || !vid || !prm || !ref_idx || !idx_out || !dist_out || !nsim_out || !nsx_out || nrefs < 0)
        return fail(VBM3D_ERR_ARG, "invalid argument");
    int rc = check(*prm, W, H, F, C, 0);
    if (rc) return rc;
    const unsigned whc = W * H * C;
    for (long r = 0; r < nrefs; r++) {
        const unsigned t = ref_idx[r] / whc, rem = ref_idx[r] % whc;
        if (t + prm->kt > F || rem >= W * H || rem % W > W - prm->k || rem / W > H - prm->k)
            return fail(VBM3D_ERR_ARG, "reference index outside the valid patch range");
    }
    if (nrefs == 0) return VBM3D_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // a one-frame clip has no predictive search: its flows are never read
    if (F <= 1) fflow = bflow = nullptr;
    const size_t n = (size_t)W * H * F * C, nf = (size_t)(F > 1 ? F - 1 : 0) * W * H * 2;
    HIP_TRY(upload(ctx->noisy, vid, n, s));
    if (fflow) HIP_TRY(upload(ctx->ff, fflow, nf, s));
    if (bflow) HIP_TRY(upload(ctx->bf, bflow, nf, s));
    HIP_TRY(ensure(ctx->rlist, nrefs * 4));
    HIP_TRY(ensure(ctx->idx, nrefs * prm->N * 4));
    HIP_TRY(ensure(ctx->dist, nrefs * prm->N * 4));
    HIP_TRY(ensure(ctx->nsim, nrefs * 4));
    HIP_TRY(ensure(ctx->nsx, nrefs * 4));
    HIP_TRY(hipMemcpyAsync(ctx->rlist.p, ref_idx, nrefs * 4, hipMemcpyHostToDevice, s));
    BMArgs bm{};
    bm.vid = (const float*)ctx->noisy.p;
    bm.ff = fflow ? (const float*)ctx->ff.p : nullptr;
    bm.bf = bflow ? (const float*)ctx->bf.p : nullptr;
    bm.W = W; bm.H = H; bm.F = F; bm.C = C; bm.wh = W * H; bm.whc = whc;
    bm.ref_list = (const uint32_t*)ctx->rlist.p;
    bm.r_begin = 0; bm.r_end = nrefs;
    bm.nrf = 1; bm.ring = 1;
    bm.Ns = prm->Ns; bm.Npr = prm->Npr; bm.Nb = prm->Nb; bm.Nf = prm->Nf; bm.N = prm->N; bm.kt = prm->kt;
    bm.d = prm->d; bm.tau = prm->tau;
    bm.dW = make_div32(W);
    bm.dWHC = make_div32(W * H * C);
    bm.idx_out = (uint32_t*)ctx->idx.p;
    bm.dist_out = (float*)ctx->dist.p;
    bm.nsim_out = (uint32_t*)ctx->nsim.p;
    bm.nsx_out = (uint32_t*)ctx->nsx.p;
    if (fast_ok(*prm, W, H)) {
        HIP_TRY(launch_bm((int)prm->k, bm, s));
    } else {
        GenArgs ga{};
        ga.vid = bm.vid; ga.ff = bm.ff; ga.bf = bm.bf;
        ga.W = W; ga.H = H; ga.F = F; ga.C = C; ga.wh = bm.wh; ga.whc = whc;
        ga.nrf = 1; ga.ring = 1;
        ga.ref_list = bm.ref_list;
        ga.r_begin = 0; ga.r_end = nrefs;
        ga.K = (int)prm->k; ga.kt = (int)prm->kt; ga.mc = prm->mc ? 1 : 0;
        ga.Ns = prm->Ns; ga.Npr = prm->Npr; ga.Nb = prm->Nb; ga.Nf = prm->Nf; ga.N = prm->N;
        ga.d = prm->d; ga.tau = prm->tau;
        ga.idx_out = bm.idx_out; ga.dist_out = bm.dist_out; ga.nsim_out = bm.nsim_out; ga.nsx_out = bm.nsx_out;
        HIP_TRY(launch_bm_gen(ga, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(idx_out, ctx->idx.p, nrefs * prm->N * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(dist_out, ctx->dist.p, nrefs * prm->N * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nsim_out, ctx->nsim.p, nrefs * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(nsx_out, ctx->nsx.p, nrefs * 4, hipMemcpyDeviceToHost));
    return VBM3D_OK;
}

int vbm3d_set_chunk_frames(vbm3d_ctx* ctx, unsigned frames) {
    if (!ctx) return fail(VBM3D_ERR_ARG, "invalid argument");
    ctx->chunk_frames = frames;
    return VBM3D_OK;
}

int vbm3d_bench_transforms(int K, int mode, long npatch, int reps, double* ms, double* max_err) {
    if ((K != 7 && K != 8) || (mode != 0 && mode != 1) || npatch <= 0 || npatch > (1L << 26) || reps <= 0 || !ms ||
        !max_err)
        return fail(VBM3D_ERR_ARG, "k must be 7 or 8, mode 0 (VALU) or 1 (MFMA), 0 < npatch <= 2^26, reps > 0");
    Tabs tb;
    fill_tabs(tb, (unsigned)K);
    XTabs t{};
    std::memcpy(t.F, tb.dctF, sizeof t.F);
    std::memcpy(t.I, tb.dctI, sizeof t.I);
    std::memcpy(t.nrm, tb.nrm, sizeof t.nrm);
    std::memcpy(t.inv, tb.inv, sizeof t.inv);
    t.dct_out = tb.dct_out;
    for (int u = 0; u < K; u++)
        for (int q = 0; q < K; q++)
            t.M[u][q] = (float)((u == 0 ? std::sqrt(1.0 / K) : std::sqrt(2.0 / K)) *
                                std::cos(M_PI * u * (2.0 * q + 1.0) / (2.0 * K)));
    const size_t n = (size_t)npatch * K * K;
    std::vector<float> h(n), o(n);
    uint32_t st = 12345u;
    for (size_t i = 0; i < n; i++) {
        st = st * 1664525u + 1013904223u;
        h[i] = (float)(st >> 8) * (255.0f / 16777216.0f);
    }
    DevBuf din, dout, dt;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int rc) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        for (DevBuf* b : {&din, &dout, &dt})
            if (b->p) (void)hipFree(b->p);
        return rc;
    };
    if (ensure(din, n * 4) != hipSuccess || ensure(dout, n * 4) != hipSuccess || ensure(dt, sizeof(XTabs)) != hipSuccess)
        return done(fail(VBM3D_ERR_NOMEM, "device allocation failed"));
    const float* in = (const float*)din.p;
    float* out = (float*)dout.p;
    const XTabs* tp = (const XTabs*)dt.p;
    if (hipMemcpy(din.p, h.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt.p, &t, sizeof t, hipMemcpyHostToDevice) != hipSuccess ||
        launch_xform_bench(K, mode, in, out, npatch, 1, tp, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(o.data(), dout.p, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(VBM3D_ERR_HIP, "transform benchmark: launch or copy failed"));
    double err = 0;
    for (size_t i = 0; i < n; i++) err = std::max(err, (double)std::fabs(o[i] - h[i]));
    *max_err = err;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return done(fail(VBM3D_ERR_HIP, "hipEventCreate failed"));
    std::vector<double> times;
    for (int k = 0; k < 4; k++) {
        float f = 0;
        if (hipEventRecord(e0, nullptr) != hipSuccess || launch_xform_bench(K, mode, in, out, npatch, reps, tp, nullptr) != hipSuccess ||
            hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&f, e0, e1) != hipSuccess)
            return done(fail(VBM3D_ERR_HIP, "transform benchmark: timed launch failed"));
        if (k > 0) times.push_back(f);
    }
    std::sort(times.begin(), times.end());
    *ms = times[1];
    return done(VBM3D_OK);
}

int vbm3d_profile(vbm3d_ctx* ctx, int enable) {
    if (!ctx) return fail(VBM3D_ERR_ARG, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (ProfEvent& e : ctx->events) {
        if (e.a) (void)hipEventDestroy(e.a);
        if (e.b) (void)hipEventDestroy(e.b);
    }
    ctx->events.clear();
    for (int k = 0; k < 3; k++) ctx->contributions[k] = ctx->pix_per_cand[k] = 0;
    ctx->prof = enable != 0;
    ctx->count = enable == 2;
    if (ctx->count) {
        HIP_TRY(ensure(ctx->stats, 8 * sizeof(unsigned long long)));
        HIP_TRY(hipMemset(ctx->stats.p, 0, 8 * sizeof(unsigned long long)));
    }
    return VBM3D_OK;
}

int vbm3d_profile_read(vbm3d_ctx* ctx, int kernel, double* ms, long* launches, double* work) {
    if (!ctx || !ms || !launches || !work) return fail(VBM3D_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    const int base = kernel & 7, step = kernel >> 3;
    if (base > 4 || step > 2) return fail(VBM3D_ERR_ARG, "unknown kernel code");
    for (ProfEvent& e : ctx->events) HIP_TRY(hipEventSynchronize(e.b));
    double tot = 0;
    long n = 0;
    for (ProfEvent& e : ctx->events)
        if (e.kernel == base && (step == 0 || e.step == step)) {
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, e.a, e.b));
            tot += t;
            n++;
        }
    unsigned long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ctx->stats.p && ctx->count)
        HIP_TRY(hipMemcpy(st, ctx->stats.p, sizeof st, hipMemcpyDeviceToHost));
    double w = 0;
    for (int k = 1; k <= 2; k++) {
        if (step != 0 && step != k) continue;
        const unsigned long long* sk = st + 4 * (k - 1);
        w += base == 0 ? (double)sk[0] : base == 1 ? (double)sk[1]
           : base == 2 ? ctx->contributions[k] : base == 4 ? (double)sk[0] * ctx->pix_per_cand[k] : 0.0;
    }
    *ms = tot;
    *launches = n;
    *work = w;
    return VBM3D_OK;
}

void vbm3d_selftest_partial_sort(float* dist, uint32_t* idx, unsigned len, unsigned mid) {
    std::vector<VbEnt> v(len);
    for (unsigned i = 0; i < len; i++) v[i] = VbEnt{dist[i], idx[i]};
    vb_partial_sort(v.data(), (int)len, (int)std::min(mid, len));
    for (unsigned i = 0; i < len; i++) {
        dist[i] = v[i].d;
        idx[i] = v[i].i;
    }
}

int vbm3d_run_batch(vbm3d_ctx* ctx, const vbm3d_clip* clips, int nclips, int ref_grid_nb,
                    const vbm3d_batch_config* cfgp, void* stream) {
    if (!ctx || (!clips && nclips > 0) || nclips < 0) return fail(VBM3D_ERR_ARG, "invalid argument");
    if (nclips == 0) return VBM3D_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    const vbm3d_batch_config cfg = cfgp ? *cfgp : (ctx->b_tuned ? ctx->b_tune : default_batch_config());
    int rc = check_batch_config(cfg);
    if (rc) return rc;
    // the batch runs on the context's stream, after the caller's work
    if (stream) HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    hipStream_t s = ctx->stream;
    bool flows = false;
    for (int i = 0; i < nclips; i++) flows = flows || clips[i].fflow || clips[i].bflow;
    const std::string key = batch_key(clips, nclips, ref_grid_nb, cfg);
    if (cfg.use_graph && !flows && ctx->b_exec && key == ctx->b_key) {
        HIP_TRY(hipGraphLaunch(ctx->b_exec, s));
        HIP_TRY(hipStreamSynchronize(s));
        return VBM3D_OK;
    }
    BatchPlan plan;
    rc = plan_batch(ctx, clips, nclips, ref_grid_nb, cfg, plan, s);
    if (rc) return rc;
    if (!cfg.use_graph || flows) {
        rc = exec_batch(ctx, plan, cfg, s);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        return VBM3D_OK;
    }
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = exec_batch(ctx, plan, cfg, s);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ce != hipSuccess) return fail(VBM3D_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(VBM3D_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    ctx->b_graph = g;
    ctx->b_exec = ex;
    ctx->b_key = key;
    HIP_TRY(hipGraphLaunch(ex, s));
    HIP_TRY(hipStreamSynchronize(s));
    return VBM3D_OK;
}

int vbm3d_autotune_batch(vbm3d_ctx* ctx, const vbm3d_clip* clips, int nclips, int ref_grid_nb,
                         vbm3d_batch_config* best, void* stream) {
    if (!ctx || !clips || nclips <= 0 || !best) return fail(VBM3D_ERR_ARG, "invalid argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (stream) HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    hipStream_t s = ctx->stream;
    vbm3d_batch_config cur = default_batch_config();
    cur.use_graph = 0;
    BatchPlan plan;
    int rc = plan_batch(ctx, clips, nclips, ref_grid_nb, cur, plan, s);
    if (rc) return rc;
    // inputs the trials mutate (colour round trips) are saved and restored
    HIP_TRY(ensure(ctx->b_save, std::max<size_t>(plan.save_bytes, 1)));
    std::vector<std::pair<float*, size_t>> saved;
    size_t off = 0;
    for (int i = 0; i < nclips; i++) {
        const vbm3d_clip& c = clips[i];
        if (!(c.C == 3 && c.color_space)) continue;
        const size_t bytes = (size_t)c.W * c.H * c.F * 3 * 4;
        for (float* v : {c.noisy, c.basic}) {
            if (v == c.basic && c.p1.k != 0) continue;
            HIP_TRY(hipMemcpyAsync((char*)ctx->b_save.p + off, v, bytes, hipMemcpyDeviceToDevice, s));
            saved.push_back({v, off});
            off += bytes;
        }
    }
    auto restore = [&]() -> int {
        for (int i = 0, k = 0; i < nclips; i++) {
            const vbm3d_clip& c = clips[i];
            if (!(c.C == 3 && c.color_space)) continue;
            const size_t bytes = (size_t)c.W * c.H * c.F * 3 * 4;
            for (float* v : {c.noisy, c.basic}) {
                if (v == c.basic && c.p1.k != 0) continue;
                HIP_TRY(hipMemcpyAsync(v, (char*)ctx->b_save.p + saved[k].second, bytes, hipMemcpyDeviceToDevice, s));
                k++;
            }
        }
        return VBM3D_OK;
    };
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    auto trial = [&](const vbm3d_batch_config& c, float* ms) -> int {
        int r = restore();
        if (r) return r;
        HIP_TRY(hipEventRecord(e0, s));
        r = exec_batch(ctx, plan, c, s);
        if (r) return r;
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(ms, e0, e1));
        return VBM3D_OK;
    };
    float best_ms = 0;
    rc = trial(cur, &best_ms);  // also warms up
    if (!rc) rc = trial(cur, &best_ms);
    // coordinate search: dispatch mode, waves per group, then (persistent
    // only) the take size and the resident groups per CU
    const int mode_c[] = {0, 1};
    const int waves_c[] = {1, 2, 4};
    const int grab_c[] = {1, 4, 16};
    const int gpc_c[] = {2, 4, 8, 16};
    for (int dim = 0; dim < 4 && !rc; dim++) {
        if (dim >= 2 && !cur.persistent) break;
        const int* vals = dim == 0 ? mode_c : dim == 1 ? waves_c : dim == 2 ? grab_c : gpc_c;
        const int nv = dim == 0 ? 2 : dim == 1 ? 3 : dim == 2 ? 3 : 4;
        vbm3d_batch_config keep = cur;
        for (int v = 0; v < nv && !rc; v++) {
            vbm3d_batch_config c = keep;
            if (dim == 0) c.persistent = vals[v];
            else if (dim == 1) c.waves_per_group = vals[v];
            else if (dim == 2) c.grab = vals[v];
            else c.groups_per_cu = vals[v];
            float ms = 0;
            rc = trial(c, &ms);
            if (!rc && ms < best_ms) { best_ms = ms; cur = c; }
        }
    }
    if (!rc) rc = restore();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    cur.use_graph = 1;
    *best = cur;
    ctx->b_tuned = true;
    ctx->b_tune = cur;
    return VBM3D_OK;
}

}  // extern "C"
