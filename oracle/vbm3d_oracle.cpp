/*
 * vbm3d_oracle.cpp -- CPU oracle for the VBM3D collaborative filter.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of
 * the product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * What it restates: tehret/vbm3d run_vbm3d (src/vbm3d.cpp:87-376) in FP32
 * with the reference's operation order.  Compiled with -ffp-contract=off
 * and without -ffast-math on x86-64 (SSE2 scalar float, no FMA), as the
 * reference's Makefile build (src/Makefile:25) compiles it.  Citations are
 * relative to the reference repository root.
 *
 * Pinning (the reference itself is unbuildable here: every translation
 * unit includes <fftw3.h> through src/Utilities/LibImages.h:17 and FFTW3 is
 * absent, and the reference ships no tests, fixtures or golden vectors):
 *   - transforms: closed forms (DCT vs scipy, Bior perfect reconstruction,
 *     Haar orthonormality), tests/test_oracle_transforms.py;
 *   - Kaiser tables: Kaiser(beta=2) rounded to 4 decimals;
 *   - grid: closed-form counts, tests/test_oracle_grid.py;
 *   - block matching: independent FP64 brute force on tie-free tiny clips,
 *     tests/test_oracle_bm.py;
 *   - shrinkage / aggregation / end to end: independent FP64 numpy
 *     restatement on tiny clips plus invariants, tests/test_oracle_e2e.py.
 * Parity unpinned: the 2D DCT's exact FP32 rounding (the reference calls
 * FFTW REDFT10/REDFT01, src/vbm3d.cpp:1290,1703, whose internal operation
 * order is not available); the oracle uses separable float matrix passes,
 * pinned only to the closed form within FP32 rounding.
 */
#include "vbm3d_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
long long g_stats[3][7];  // last orc_step's counters, per step (see StepIn::stats)

const unsigned kDCT = 0, kBIOR = 1, kHADAMARD = 2;

// ---------------------------------------------------------------------
// Video geometry: VideoSize (src/Utilities/LibVideoT.hpp:40-121).  The
// reference indexes with 32-bit unsigned arithmetic; so do we.
// ---------------------------------------------------------------------
struct Geo {
    unsigned W = 0, H = 0, F = 0, C = 0, wh = 0, whc = 0, whcf = 0;
    Geo() {}
    Geo(unsigned w, unsigned h, unsigned f, unsigned c)
        : W(w), H(h), F(f), C(c), wh(w * h), whc(w * h * c), whcf(w * h * c * f) {}
    unsigned index(unsigned x, unsigned y, unsigned t, unsigned c) const {
        return t * whc + c * wh + y * W + x;
    }
    void coords(unsigned idx, unsigned& x, unsigned& y, unsigned& t) const {
        t = idx / whc;
        y = (idx % wh) / W;
        x = idx % W;
    }
};

// Flow access fflow(x, y, t, c) on a [Ff][2][H][W] video
// (src/main.cpp:355; src/vbm3d.cpp:1129-1130).
inline float flow_at(const float* fl, const Geo& g, unsigned x, unsigned y,
                     unsigned t, unsigned c) {
    if (!fl) return 0.f;
    return fl[(size_t)t * g.wh * 2 + (size_t)c * g.wh + (size_t)y * g.W + x];
}

// patch_trajectory (src/vbm3d.cpp:378-425, FLOAT_TRAJECTORIES): the
// top-left corner of each of the kt slices of the patch at (px, py, pt).
// With mc, the next corner is the running float sum of the forward flow
// read at the current corner plus k/2 (both components before either
// moves), rounded half away from zero and clamped to the frame.
inline void trajectory(unsigned px, unsigned py, unsigned pt, unsigned k, unsigned kt, bool mc,
                       const float* ff, const Geo& g, unsigned* tx, unsigned* ty) {
    float fpx = (float)px, fpy = (float)py;
    for (unsigned ht = 0; ht < kt; ht++) {
        tx[ht] = px;
        ty[ht] = py;
        if (mc && ht + 1 < kt) {
            fpx += flow_at(ff, g, px + k / 2, py + k / 2, pt + ht, 0);
            fpy += flow_at(ff, g, px + k / 2, py + k / 2, pt + ht, 1);
            px = (unsigned)std::min((int)g.W - (int)k, std::max(0, (int)std::round(fpx)));
            py = (unsigned)std::min((int)g.H - (int)k, std::max(0, (int)std::round(fpy)));
        }
    }
}

// ---------------------------------------------------------------------
// Small helpers (src/Utilities/Utilities.cpp:152-173, 235-265, 400-415).
// ---------------------------------------------------------------------
unsigned pow2_floor(unsigned n) {  // closest_power_of_2
    unsigned r = 1;
    while (r * 2 <= n) r *= 2;
    return r;
}
bool is_pow2(unsigned n) { return n != 0 && (n & (n - 1)) == 0; }

void factor_tiles(unsigned n, unsigned& a, unsigned& b) {  // determineFactor
    if (n == 1) { a = b = 1; return; }
    b = 2;
    while (n % b > 0) b++;
    a = n / b;
    if (b > a) { a = b; b = n / a; }
}

// ind_initialize: begin, begin+step, ... <= end, then end if missed.
bool grid_1d(std::vector<unsigned>& out, long begin, long end, unsigned step) {
    out.clear();
    if (begin < 0 || end < begin || step == 0) return false;
    for (long v = begin; v <= end; v += step) out.push_back((unsigned)v);
    if ((long)out.back() < end) out.push_back((unsigned)end);
    return true;
}

struct Crop { int ox, oy, ex, ey; };

// subDivideTight (src/Utilities/LibVideoT.hpp:903-961).
std::vector<Crop> make_crops(unsigned W, unsigned H, int border, int nb) {
    unsigned nW, nH;
    factor_tiles((unsigned)nb, nW, nH);
    const int wT = (int)std::ceil(float(W) / float(nW));
    const int hT = (int)std::ceil(float(H) / float(nH));
    std::vector<Crop> out;
    for (int p = 0; p < (int)nH; p++)
        for (int q = 0; q < (int)nW; q++) {
            Crop c;
            c.ox = std::max(0, q * wT - border);
            c.oy = std::max(0, p * hT - border);
            c.ex = std::min((int)W, (q + 1) * wT + border);
            c.ey = std::min((int)H, (p + 1) * hT + border);
            out.push_back(c);
        }
    return out;
}

struct Grid {
    // one entry per crop: absolute row and column positions
    std::vector<std::vector<unsigned>> rows, cols;
};

// Reference grid per crop (src/vbm3d.cpp:471-489, 700-719).
bool make_grid(Grid& g, unsigned W, unsigned H, unsigned k, unsigned p,
               int n, int nb) {
    g.rows.clear();
    g.cols.clear();
    if (W < k || H < k || p == 0) return false;
    if (nb == 1) {
        std::vector<unsigned> r, c;
        if (!grid_1d(r, 0, (long)H - k, p) || !grid_1d(c, 0, (long)W - k, p))
            return false;
        g.rows.push_back(r);
        g.cols.push_back(c);
        return true;
    }
    if (nb < 1 || nb > 32 || !is_pow2((unsigned)nb)) return false;
    for (const Crop& cr : make_crops(W, H, 2 * n, nb)) {
        std::vector<unsigned> r, c;
        long rb = cr.oy == 0 ? 0 : n;
        long re = cr.ey == (int)H ? (long)cr.ey - cr.oy - k : (long)cr.ey - cr.oy - k - n;
        long cb = cr.ox == 0 ? 0 : n;
        long ce = cr.ex == (int)W ? (long)cr.ex - cr.ox - k : (long)cr.ex - cr.ox - k - n;
        if (!grid_1d(r, rb, re, p) || !grid_1d(c, cb, ce, p)) return false;
        for (auto& v : r) v += cr.oy;
        for (auto& v : c) v += cr.ox;
        g.rows.push_back(r);
        g.cols.push_back(c);
    }
    return true;
}

// ---------------------------------------------------------------------
// Noise level per channel: estimate_sigma
// (src/Utilities/Utilities.cpp:282-333).  The reference passes a bool as
// the colour-space code (src/vbm3d.cpp:467, 697): true selects the YCbCr
// row although the transform applied is YUV.
// ---------------------------------------------------------------------
bool sigma_table(float sigma, unsigned C, unsigned cs, float* out) {
    if (C == 1) { out[0] = sigma; return true; }
    if (C != 3) return false;
    if (cs == 0) {  // YUV
        out[0] = sqrtf(0.299f * 0.299f + 0.587f * 0.587f + 0.114f * 0.114f) * sigma;
        out[1] = sqrtf(0.14713f * 0.14713f + 0.28886f * 0.28886f + 0.436f * 0.436f) * sigma;
        out[2] = sqrtf(0.615f * 0.615f + 0.51498f * 0.51498f + 0.10001f * 0.10001f) * sigma;
    } else if (cs == 1) {  // YCbCr
        out[0] = sqrtf(0.299f * 0.299f + 0.587f * 0.587f + 0.114f * 0.114f) * sigma;
        out[1] = sqrtf(0.169f * 0.169f + 0.331f * 0.331f + 0.500f * 0.500f) * sigma;
        out[2] = sqrtf(0.500f * 0.500f + 0.419f * 0.419f + 0.081f * 0.081f) * sigma;
    } else if (cs == 2) {  // OPP
        out[0] = sqrtf(0.333f * 0.333f + 0.333f * 0.333f + 0.333f * 0.333f) * sigma;
        out[1] = sqrtf(0.5f * 0.5f + 0.0f * 0.0f + 0.5f * 0.5f) * sigma;
        out[2] = sqrtf(0.25f * 0.25f + 0.5f * 0.5f + 0.25f * 0.25f) * sigma;
    } else if (cs == 3) {
        out[0] = out[1] = out[2] = sigma;
    } else {
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------
// Colour transform, 3 channels (src/Utilities/LibVideoT.hpp:763-888).
// ---------------------------------------------------------------------
void color_transform(float* v, const Geo& g, bool forward) {
    if (g.C != 3) return;
    for (unsigned f = 0; f < g.F; f++) {
        float* r = v + (size_t)f * g.whc;
        float* gr = r + g.wh;
        float* b = gr + g.wh;
        for (unsigned k = 0; k < g.wh; k++) {
            const float R = r[k], G = gr[k], B = b[k];
            float o0, o1, o2;
            if (forward) {
                o0 = 0.299f * R + 0.587f * G + 0.114f * B;
                o1 = -0.14713f * R - 0.28886f * G + 0.436f * B;
                o2 = 0.615f * R - 0.51498f * G - 0.10001f * B;
            } else {
                // R, G, B here hold Y, U, V
                o0 = R + 1.13983f * B;
                o1 = R - 0.39465f * G - 0.5806f * B;
                o2 = R + 2.03211f * G;
            }
            r[k] = o0;
            gr[k] = o1;
            b[k] = o2;
        }
    }
}

// ---------------------------------------------------------------------
// Transform tables.
// ---------------------------------------------------------------------
// Bior1.5 analysis rows, MATLAB basis (src/lib_transforms.cpp:141-148):
// kBiorF[j][s] = basis function j at sample s.
const double kBiorF[8][8] = {
    {0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274,
     0.353553390593274, 0.353553390593274, 0.353553390593274, 0.353553390593274},
    {0.219417649252501, 0.449283757993216, 0.449283757993216, 0.219417649252501,
     -0.219417649252501, -0.449283757993216, -0.449283757993216, -0.219417649252501},
    {0.569359398342846, 0.402347308162278, -0.402347308162278, -0.569359398342846,
     -0.083506045090284, 0.083506045090284, -0.083506045090284, 0.083506045090284},
    {-0.083506045090284, 0.083506045090284, -0.083506045090284, 0.083506045090284,
     0.569359398342846, 0.402347308162278, -0.402347308162278, -0.569359398342846},
    {0.707106781186547, -0.707106781186547, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.707106781186547, -0.707106781186547, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.707106781186547, -0.707106781186547, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.707106781186547, -0.707106781186547}};
// Bior1.5 synthesis (src/lib_transforms.cpp:281-288): kBiorI[s][j] =
// weight of coefficient j in sample s.  Signed zeros kept.
const double kBiorI[8][8] = {
    {0.353553390593274, 0.353553390593274, 0.5, 0.0, 0.707106781186547, -0.121533978016438, 0.0, 0.121533978016438},
    {0.353553390593274, 0.353553390593274, 0.5, -0.0, -0.707106781186547, -0.121533978016438, 0.0, 0.121533978016438},
    {0.353553390593274, 0.353553390593274, -0.5, -0.0, 0.121533978016438, 0.707106781186548, -0.121533978016438, 0.0},
    {0.353553390593274, 0.353553390593274, -0.5, 0.0, 0.121533978016438, -0.707106781186548, -0.121533978016438, 0.0},
    {0.353553390593274, -0.353553390593274, 0.0, 0.5, 0.0, 0.121533978016438, 0.707106781186547, -0.121533978016438},
    {0.353553390593274, -0.353553390593274, 0.0, 0.5, 0.0, 0.121533978016438, -0.707106781186547, -0.121533978016438},
    {0.353553390593274, -0.353553390593274, 0.0, -0.5, -0.121533978016438, 0.0, 0.121533978016438, 0.707106781186547},
    {0.353553390593274, -0.353553390593274, 0.0, -0.5, -0.121533978016438, 0.0, 0.121533978016438, -0.707106781186547}};

const double kSqrt2 = 1.414213562373095;      // src/vbm3d.cpp:32
const double kSqrt2Inv = 0.7071067811865475;  // src/vbm3d.cpp:33

// Kaiser quarter tables (src/vbm3d.cpp:1755-1841); value at (p, q) is
// Q[min(p, k-1-p)][min(q, k-1-q)].
void kaiser_table(unsigned k, float* w) {
    static const float q4[2][2] = {{0.1924f, 0.4055f}, {0.4055f, 0.8544f}};
    static const float q6[3][3] = {{0.1924f, 0.3368f, 0.4265f},
                                   {0.3368f, 0.5893f, 0.7464f},
                                   {0.4265f, 0.7464f, 0.9454f}};
    static const float q7[4][4] = {{0.1924f, 0.3151f, 0.4055f, 0.4387f},
                                   {0.3151f, 0.5161f, 0.6640f, 0.7184f},
                                   {0.4055f, 0.6640f, 0.8544f, 0.9243f},
                                   {0.4387f, 0.7184f, 0.9243f, 1.0000f}};
    static const float q8[4][4] = {{0.1924f, 0.2989f, 0.3846f, 0.4325f},
                                   {0.2989f, 0.4642f, 0.5974f, 0.6717f},
                                   {0.3846f, 0.5974f, 0.7688f, 0.8644f},
                                   {0.4325f, 0.6717f, 0.8644f, 0.9718f}};
    static const float q12[6][6] = {
        {0.1924f, 0.2615f, 0.3251f, 0.3782f, 0.4163f, 0.4362f},
        {0.2615f, 0.3554f, 0.4419f, 0.5139f, 0.5657f, 0.5927f},
        {0.3251f, 0.4419f, 0.5494f, 0.6390f, 0.7033f, 0.7369f},
        {0.3782f, 0.5139f, 0.6390f, 0.7433f, 0.8181f, 0.8572f},
        {0.4163f, 0.5657f, 0.7033f, 0.8181f, 0.9005f, 0.9435f},
        {0.4362f, 0.5927f, 0.7369f, 0.8572f, 0.9435f, 0.9885f}};
    for (unsigned p = 0; p < k; p++)
        for (unsigned q = 0; q < k; q++) {
            const unsigned a = std::min(p, k - 1 - p), b = std::min(q, k - 1 - q);
            float v = 1.0f;
            if (k == 4) v = q4[a][b];
            else if (k == 6) v = q6[a][b];
            else if (k == 7) v = q7[a][b];
            else if (k == 8) v = q8[a][b];
            else if (k == 12) v = q12[a][b];
            w[p * k + q] = v;
        }
}

// DCT normalisation (src/vbm3d.cpp:1843-1863).
void dct_norm(unsigned k, float* nrm, float* inv) {
    const float coef = 0.5f / ((float)k);
    for (unsigned i = 0; i < k; i++)
        for (unsigned j = 0; j < k; j++) {
            if (i == 0 && j == 0) {
                nrm[i * k + j] = 0.5f * coef;
                inv[i * k + j] = 2.0f;
            } else if (i * j == 0) {
                nrm[i * k + j] = (float)(kSqrt2Inv * coef);
                inv[i * k + j] = (float)kSqrt2;
            } else {
                nrm[i * k + j] = 1.0f * coef;
                inv[i * k + j] = 1.0f;
            }
        }
}

struct Tables {
    unsigned k, kt;
    float biorF[64], biorI[64];          // float copies, row-major [j][s] / [s][j]
    float dctF[16 * 16], dctI[16 * 16];  // [v][q] REDFT10 and [q][v] REDFT01
    float nrm[256], inv[256], kaiser[256];
    float tF[64], tI[64];  // temporal REDFT10 [u][t] and REDFT01 [t][u], length kt
    float tnorm;           // 1/sqrt(2 kt) rounded to float (src/vbm3d.cpp:1876)
    explicit Tables(unsigned kk, unsigned ktt = 1) : k(kk), kt(ktt) {
        for (unsigned u = 0; u < kt; u++)
            for (unsigned t = 0; t < kt; t++) {
                tF[u * kt + t] = (float)(2.0 * std::cos(M_PI * (2.0 * t + 1.0) * u / (2.0 * kt)));
                tI[t * kt + u] = u == 0 ? 1.0f : (float)(2.0 * std::cos(M_PI * u * (2.0 * t + 1.0) / (2.0 * kt)));
            }
        tnorm = (float)(1. / std::sqrt(2.0 * kt));
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++) {
                biorF[a * 8 + b] = (float)kBiorF[a][b];
                biorI[a * 8 + b] = (float)kBiorI[a][b];
            }
        // Closed forms of FFTW's REDFT10 / REDFT01 (the FFTW manual's
        // definitions), evaluated in double and rounded once to float.
        for (unsigned v = 0; v < k; v++)
            for (unsigned q = 0; q < k; q++) {
                dctF[v * k + q] = (float)(2.0 * std::cos(M_PI * (2.0 * q + 1.0) * v / (2.0 * k)));
                dctI[q * k + v] = v == 0 ? 1.0f
                    : (float)(2.0 * std::cos(M_PI * v * (2.0 * q + 1.0) / (2.0 * k)));
            }
        dct_norm(k, nrm, inv);
        kaiser_table(k, kaiser);
    }
};


// Scalar-typed views of the tables: float copies for the FP32 oracle,
// exact double values for the FP64 re-evaluation.
template <class T>
struct TabT {
    T biorF[64], biorI[64], dctF[256], dctI[256], nrm[256], inv[256];
    T kaiser[256];
    T dct_out;  // 1/(2k), src/vbm3d.cpp:1707
    T tF[64], tI[64], tnorm;  // temporal DCT (kt > 1)
    unsigned k, kt;
};

TabT<float> make_tab_f(const Tables& t) {
    TabT<float> o;
    o.k = t.k;
    std::memcpy(o.biorF, t.biorF, sizeof o.biorF);
    std::memcpy(o.biorI, t.biorI, sizeof o.biorI);
    std::memcpy(o.dctF, t.dctF, sizeof o.dctF);
    std::memcpy(o.dctI, t.dctI, sizeof o.dctI);
    std::memcpy(o.nrm, t.nrm, sizeof o.nrm);
    std::memcpy(o.inv, t.inv, sizeof o.inv);
    std::memcpy(o.kaiser, t.kaiser, sizeof o.kaiser);
    o.dct_out = 1.0f / (float)(t.k * 2);
    o.kt = t.kt;
    std::memcpy(o.tF, t.tF, sizeof o.tF);
    std::memcpy(o.tI, t.tI, sizeof o.tI);
    o.tnorm = t.tnorm;
    return o;
}

TabT<double> make_tab_d(const Tables& t) {
    TabT<double> o;
    const unsigned k = t.k;
    o.k = k;
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 8; b++) {
            o.biorF[a * 8 + b] = kBiorF[a][b];
            o.biorI[a * 8 + b] = kBiorI[a][b];
        }
    for (unsigned v = 0; v < k; v++)
        for (unsigned q = 0; q < k; q++) {
            o.dctF[v * k + q] = 2.0 * std::cos(M_PI * (2.0 * q + 1.0) * v / (2.0 * k));
            o.dctI[q * k + v] = v == 0 ? 1.0 : 2.0 * std::cos(M_PI * v * (2.0 * q + 1.0) / (2.0 * k));
        }
    for (unsigned i = 0; i < k; i++)
        for (unsigned j = 0; j < k; j++) {
            double s = 0.5 / k;
            if (i == 0 && j == 0) { o.nrm[i * k + j] = 0.5 * s; o.inv[i * k + j] = 2.0; }
            else if (i * j == 0) { o.nrm[i * k + j] = std::sqrt(0.5) * s; o.inv[i * k + j] = std::sqrt(2.0); }
            else { o.nrm[i * k + j] = s; o.inv[i * k + j] = 1.0; }
        }
    for (unsigned a = 0; a < k * k; a++) o.kaiser[a] = t.kaiser[a];
    o.dct_out = 1.0 / (2.0 * k);
    const unsigned kt = t.kt;
    o.kt = kt;
    for (unsigned u = 0; u < kt; u++)
        for (unsigned tt = 0; tt < kt; tt++) {
            o.tF[u * kt + tt] = 2.0 * std::cos(M_PI * (2.0 * tt + 1.0) * u / (2.0 * kt));
            o.tI[tt * kt + u] = u == 0 ? 1.0 : 2.0 * std::cos(M_PI * u * (2.0 * tt + 1.0) / (2.0 * kt));
        }
    o.tnorm = 1.0 / std::sqrt(2.0 * kt);
    return o;
}

// temporal_transform (src/vbm3d.cpp:1866-1910) on one coefficient's kt
// values x[t] (stride sx): REDFT10 as sequential sums over t with the
// closed-form matrix, times 1/sqrt(2 kt) (a float product), then the DC
// times SQRT2_INV in double, rounded once to float.
template <class T>
void temporal_fwd(T* x, unsigned sx, const TabT<T>& tb) {
    T y[8];
    const unsigned kt = tb.kt;
    for (unsigned u = 0; u < kt; u++) {
        T m = 0;
        for (unsigned t = 0; t < kt; t++) m += x[t * sx] * tb.tF[u * kt + t];
        y[u] = m;
    }
    for (unsigned u = 0; u < kt; u++)
        x[u * sx] = (T)((double)(y[u] * tb.tnorm) * (u == 0 ? kSqrt2Inv : 1.0));
}

// temporal_inv_transform (src/vbm3d.cpp:1912-1956): the DC times SQRT2 in
// double (rounded to float), REDFT01 as sequential sums over u, times
// 1/sqrt(2 kt).
template <class T>
void temporal_inv(T* x, unsigned sx, const TabT<T>& tb) {
    T z[8];
    const unsigned kt = tb.kt;
    for (unsigned u = 0; u < kt; u++) z[u] = (T)((double)x[u * sx] * (u == 0 ? kSqrt2 : 1.0));
    for (unsigned t = 0; t < kt; t++) {
        T m = 0;
        for (unsigned u = 0; u < kt; u++) m += z[u] * tb.tI[t * kt + u];
        x[t * sx] = m * tb.tnorm;
    }
}

// ---------------------------------------------------------------------
// 2D transforms.
// ---------------------------------------------------------------------
// bior_2d_forward, MATLAB branch (src/lib_transforms.cpp:125-181).  x is
// the patch row-major (row i, column j); out[i*8+j] has i = horizontal
// frequency and j = vertical frequency (the second pass transposes).
template <class T>
void bior_fwd(const T* x, T* out, const TabT<T>& tb) {
    T t2[64];
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) {
            T m = 0;
            for (unsigned s = 0; s < 8; s++) m += x[i * 8 + s] * tb.biorF[j * 8 + s];
            t2[i * 8 + j] = m;
        }
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) {
            T m = 0;
            for (unsigned s = 0; s < 8; s++) m += t2[s * 8 + i] * tb.biorF[j * 8 + s];
            out[i * 8 + j] = m;
        }
}

// bior_2d_inverse, MATLAB branch (src/lib_transforms.cpp:270-321), in
// place on one 64-value chunk.
template <class T>
void bior_inv(T* sig, const TabT<T>& tb) {
    T a[64], b[64];
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) a[i * 8 + j] = sig[i + j * 8];
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) {
            T m = 0;
            for (unsigned s = 0; s < 8; s++) m += a[i * 8 + s] * tb.biorI[j * 8 + s];
            b[i * 8 + j] = m;
        }
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) {
            T m = 0;
            for (unsigned s = 0; s < 8; s++) m += b[s * 8 + i] * tb.biorI[j * 8 + s];
            a[i * 8 + j] = m;
        }
    for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) sig[j * 8 + i] = a[i * 8 + j];
}

// One 1D pass of REDFT10 on k values at stride sx, written at stride sy.
// Row v of the matrix is even (v even) or odd (v odd) about the centre, so
// each output takes the folded inputs s[q] = x[q] + x[k-1-q] or
// d[q] = x[q] - x[k-1-q], q < k/2, plus the centre sample for odd k and
// even v (the odd rows vanish there); the sum runs in q order.  This is
// FFTW's REDFT10 (the reference's transform) up to rounding, with about
// half the multiplications of the plain matrix product.
template <class T>
void dct_pass_fwd(const T* x, unsigned sx, T* y, unsigned sy, const TabT<T>& tb) {
    const unsigned k = tb.k, h = k / 2;
    T s[8], d[8];
    for (unsigned q = 0; q < h; q++) {
        s[q] = x[q * sx] + x[(k - 1 - q) * sx];
        d[q] = x[q * sx] - x[(k - 1 - q) * sx];
    }
    for (unsigned v = 0; v < k; v++) {
        T m;
        if (v % 2 == 0) {
            if (h == 0) {
                m = x[0] * tb.dctF[v * k];  // k = 1: the centre sample alone
            } else {
                m = s[0] * tb.dctF[v * k];
                for (unsigned q = 1; q < h; q++) m += s[q] * tb.dctF[v * k + q];
                if (k & 1) m += x[h * sx] * tb.dctF[v * k + h];
            }
        } else {
            m = d[0] * tb.dctF[v * k];
            for (unsigned q = 1; q < h; q++) m += d[q] * tb.dctF[v * k + q];
        }
        y[v * sy] = m;
    }
}

// One 1D pass of REDFT01 (matrix I[o][i] = dctI[o*k + i]).  Rows o and
// k-1-o differ only in the sign of the odd-i terms: E = sum over even i,
// O = sum over odd i (each in i order), y[o] = E + O, y[k-1-o] = E - O;
// the centre row of odd k is E alone (its odd terms vanish).
template <class T>
void dct_pass_inv(const T* z, unsigned sz, T* y, unsigned sy, const TabT<T>& tb) {
    const unsigned k = tb.k, h = k / 2;
    for (unsigned o = 0; o < h; o++) {
        T e = z[0] * tb.dctI[o * k], od = z[sz] * tb.dctI[o * k + 1];
        for (unsigned i = 2; i < k; i += 2) e += z[i * sz] * tb.dctI[o * k + i];
        for (unsigned i = 3; i < k; i += 2) od += z[i * sz] * tb.dctI[o * k + i];
        y[o * sy] = e + od;
        y[(k - 1 - o) * sy] = e - od;
    }
    if (k & 1) {
        T e = z[0] * tb.dctI[h * k];
        for (unsigned i = 2; i < k; i += 2) e += z[i * sz] * tb.dctI[h * k + i];
        y[h * sy] = e;
    }
}

// Orthonormal 2D DCT-II as REDFT10 on both axes followed by coef_norm
// (src/vbm3d.cpp:1245-1302).  Row pass (along columns q) first, then the
// column pass; out[u*k+v], u = vertical frequency.
template <class T>
void dct_fwd(const T* x, T* out, const TabT<T>& tb) {
    const unsigned k = tb.k;
    T t[256], c[16];
    for (unsigned p = 0; p < k; p++) dct_pass_fwd(x + p * k, 1, t + p * k, 1, tb);
    for (unsigned v = 0; v < k; v++) {
        dct_pass_fwd(t + v, k, c, 1, tb);
        for (unsigned u = 0; u < k; u++) out[u * k + v] = c[u] * tb.nrm[u * k + v];
    }
}

// dct_2d_inv (src/vbm3d.cpp:1678-1713): coef_norm_inv, REDFT01 on both
// axes (row pass first), then the 1/(2k) factor; in place.
template <class T>
void dct_inv(T* sig, const TabT<T>& tb) {
    const unsigned k = tb.k;
    T z[256], t[256], c[16];
    for (unsigned a = 0; a < k * k; a++) z[a] = sig[a] * tb.inv[a];
    for (unsigned u = 0; u < k; u++) dct_pass_inv(z + u * k, 1, t + u * k, 1, tb);
    for (unsigned q = 0; q < k; q++) {
        dct_pass_inv(t + q, k, c, 1, tb);
        for (unsigned p = 0; p < k; p++) sig[p * k + q] = tb.dct_out * c[p];
    }
}

// ---------------------------------------------------------------------
// 1D transforms along the group.
// ---------------------------------------------------------------------
// (a op b) * SQRT2_INV: float operands, double constant, float result
// (src/lib_transforms.cpp:468-469, 503-504).
template <class T>
inline T haar_scale(T s) { return (T)((double)s * kSqrt2Inv); }

// haar_forward (src/lib_transforms.cpp:453-476).
template <class T>
void haar_fwd(T* v, unsigned n) {
    T tmp[64];
    for (unsigned i = 0; i < 64; i++) tmp[i] = 0;
    int k = (int)n;
    while (1 < k) {
        k = k / 2;
        for (int i = 0; i < k; i++) {
            tmp[i] = haar_scale<T>(v[2 * i] + v[2 * i + 1]);
            tmp[i + k] = haar_scale<T>(v[2 * i] - v[2 * i + 1]);
        }
        for (int i = 0; i < k * 2; i++) v[i] = tmp[i];
    }
}

// haar_inverse (src/lib_transforms.cpp:489-512).
template <class T>
void haar_inv(T* v, unsigned n) {
    T tmp[64];
    for (unsigned i = 0; i < 64; i++) tmp[i] = 0;
    int k = 1;
    while (k * 2 <= (int)n) {
        for (int i = 0; i < k; ++i) {
            tmp[2 * i] = haar_scale<T>(v[i] + v[i + k]);
            tmp[2 * i + 1] = haar_scale<T>(v[i] - v[i + k]);
        }
        for (int i = 0; i < k * 2; i++) v[i] = tmp[i];
        k = k * 2;
    }
}

// hadamard_transform, unnormalised, recursive (src/lib_transforms.cpp:410-440).
template <class T>
void hadamard(T* v, unsigned n) {
    if (n <= 1) return;
    if (n == 2) {
        const T a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
        return;
    }
    const unsigned h = n / 2;
    T tmp[64];
    for (unsigned k = 0; k < h; k++) {
        const T a = v[2 * k], b = v[2 * k + 1];
        v[k] = a + b;
        tmp[k] = a - b;
    }
    for (unsigned k = 0; k < h; k++) v[h + k] = tmp[k];
    hadamard(v, h);
    hadamard(v + h, h);
}

// Relative distance below which two ranked values, a distance and tau, or
// a coefficient and the hard threshold count as a near tie: a few hundred
// FP32 ulps, the size of the rounding differences another operation order
// (FFTW's DCT, a -ffast-math build) could introduce.
constexpr double kNearRel = 1e-5;
inline bool near_rel(double a, double b) {
    return std::fabs(a - b) <= kNearRel * std::max(1.0, std::max(std::fabs(a), std::fabs(b)));
}

// A ranking decision near a tie: among the `keep` smallest of `v` (sorted
// copy), or between the last kept and the first dropped value.
// ---------------------------------------------------------------------
// Shrinkage along the group.  Group layout g[n + kk*nsx + c*k2*nsx]
// (src/vbm3d.cpp:569, 810-813).  `keep` records (FP32) or replays (FP64)
// the hard-threshold decisions, one byte per coefficient.
// ---------------------------------------------------------------------
template <class T>
void ht_filter(T* g, unsigned nsx, unsigned k2, unsigned C, const float* sig,
               float lambda, unsigned T3D, T* w, uint8_t* keep, bool replay,
               long long* near_thr = nullptr) {
    const unsigned per_c = k2 * nsx;
    const bool had = T3D == kHADAMARD;
    for (unsigned a = 0; a < k2 * C; a++) {
        if (had) hadamard(g + a * nsx, nsx);
        else haar_fwd(g + a * nsx, nsx);
    }
    T cnt[3];
    for (unsigned c = 0; c < C; c++) {
        // ht_filtering_hadamard scales the threshold by sqrt(nSx)
        // (src/vbm3d.cpp:1421, 1432); ht_filtering_haar does not (:1500).
        const float thr = had ? lambda * sig[c] * sqrtf((float)nsx) : lambda * sig[c];
        T n_keep = 0;
        for (unsigned k = 0; k < per_c; k++) {
            T& v = g[c * per_c + k];
            bool kp;
            if (replay) kp = keep[c * per_c + k] != 0;
            else {
                kp = (k == 0 || std::fabs((float)v) > thr);
                keep[c * per_c + k] = kp;
                if (near_thr && k != 0 && near_rel(std::fabs((double)v), (double)thr)) (*near_thr)++;
            }
            if (kp) n_keep++;
            else v = 0;
        }
        cnt[c] = n_keep;
    }
    for (unsigned a = 0; a < k2 * C; a++) {
        if (had) hadamard(g + a * nsx, nsx);
        else haar_inv(g + a * nsx, nsx);
    }
    if (had) {
        const float coef = 1.0f / (float)nsx;
        for (unsigned a = 0; a < per_c * C; a++) g[a] *= (T)coef;
    }
    for (unsigned c = 0; c < C; c++)
        w[c] = (T)1 / ((T)sig[c] * (T)sig[c] * cnt[c]);
}

// wiener_filtering_haar / _hadamard (src/vbm3d.cpp:1613-1666, 1539-1594),
// DCWIENER on: every coefficient, DC included, is Wiener-filtered.
template <class T>
void wiener_filter(T* z, T* e, unsigned nsx, unsigned k2, unsigned C,
                   const float* sig, unsigned T3D, T* w) {
    const unsigned per_c = k2 * nsx;
    const bool had = T3D == kHADAMARD;
    for (unsigned a = 0; a < k2 * C; a++) {
        if (had) { hadamard(z + a * nsx, nsx); hadamard(e + a * nsx, nsx); }
        else { haar_fwd(z + a * nsx, nsx); haar_fwd(e + a * nsx, nsx); }
    }
    const T coef = (T)(1.0f / (float)nsx);
    T acc[3];
    for (unsigned c = 0; c < C; c++) {
        T W = 0;
        const T s2 = (T)sig[c] * (T)sig[c];
        for (unsigned k = 0; k < per_c; k++) {
            const unsigned i = c * per_c + k;
            T value = e[i] * e[i];
            if (had) value = value * coef;
            value /= (value + s2);
            e[i] = had ? z[i] * value * coef : z[i] * value;
            W += value * value;
        }
        acc[c] = W;
    }
    for (unsigned a = 0; a < k2 * C; a++) {
        if (had) hadamard(e + a * nsx, nsx);
        else haar_inv(e + a * nsx, nsx);
    }
    for (unsigned c = 0; c < C; c++) {
        const T s2 = (T)sig[c] * (T)sig[c];
        if (had) w[c] = acc[c] > 0 ? (T)1 / (s2 * acc[c]) : (T)1;
        else w[c] = (T)1 / (s2 * acc[c]);
    }
}

// ---------------------------------------------------------------------
// Predictive block matching (src/vbm3d.cpp:930-1197).
// ---------------------------------------------------------------------
typedef std::pair<float, unsigned> Cand;
bool cmp_first(const Cand& a, const Cand& b) { return a.first < b.first; }  // Utilities.cpp:40-45

// Set of already-evaluated patch indices (the reference's
// unordered_map<unsigned,int> alreadySeen, src/vbm3d.cpp:1065).
struct Seen {
    std::vector<uint32_t> key, stamp;
    uint32_t gen = 0, mask = 0;
    explicit Seen(unsigned cap) {
        unsigned c = 64;
        while (c < 2 * cap) c *= 2;
        key.assign(c, 0);
        stamp.assign(c, 0);
        mask = c - 1;
    }
    void reset() { gen++; }
    bool insert(uint32_t k) {  // true if k was not seen before
        uint32_t h = (k * 2654435761u) & mask;
        while (stamp[h] == gen) {
            if (key[h] == k) return false;
            h = (h + 1) & mask;
        }
        stamp[h] = gen;
        key[h] = k;
        return true;
    }
};

inline bool ranking_near_tie(std::vector<float> v, size_t keep) {
    std::sort(v.begin(), v.end());
    const size_t upto = std::min(keep + 1, v.size());
    for (size_t j = 1; j < upto; j++)
        if (near_rel(v[j - 1], v[j])) return true;
    return false;
}

struct BM {
    const float* vid;
    Geo g;
    const float *ff, *bf;
    orc_params pr;
    Seen seen;
    long long n_eval = 0;
    long long near_tie = 0;  // references with a ranking decision near a tie
    long long near_tau = 0;  // references with a group candidate near tau
    bool tie_now = false;
    BM(const float* v, const Geo& gg, const float* f, const float* b, const orc_params& p)
        : vid(v), g(gg), ff(f), bf(b), pr(p),
          seen(p.Ns * p.Ns + 2 * p.Nf * p.Nb * p.Npr * p.Npr + 1) {}

    // patchDistance (src/vbm3d.cpp:930-1011): both patches follow their
    // trajectories over the kt slices (identity unless mc); one sequential
    // float sum in (slice, channel, row, column) order.
    float dist(unsigned r, unsigned q) const {
        unsigned rx, ry, rt, qx, qy, qt;
        g.coords(r, rx, ry, rt);
        g.coords(q, qx, qy, qt);
        const unsigned k = pr.k, kt = pr.kt;
        unsigned ax[16], ay[16], bx[16], by[16];
        trajectory(rx, ry, rt, k, kt, pr.mc != 0, ff, g, ax, ay);
        trajectory(qx, qy, qt, k, kt, pr.mc != 0, ff, g, bx, by);
        float d = 0.f, dif;
        for (unsigned ht = 0; ht < kt; ht++)
            for (unsigned c = 0; c < g.C; ++c)
                for (unsigned hy = 0; hy < k; hy++)
                    for (unsigned hx = 0; hx < k; hx++) {
                        dif = vid[g.index(ax[ht] + hx, ay[ht] + hy, rt + ht, c)] -
                              vid[g.index(bx[ht] + hx, by[ht] + hy, qt + ht, c)];
                        d += dif * dif;
                    }
        return d / (float)(k * k * kt * g.C);
    }

    // localSearch (src/vbm3d.cpp:1013-1080).
    void local(unsigned center, unsigned ref, unsigned s, std::vector<Cand>& out) {
        unsigned rx, ry, rt, px, py, pt;
        g.coords(ref, rx, ry, rt);
        g.coords(center, px, py, pt);
        const int h = ((int)s - 1) / 2;
        const unsigned x0 = std::max(0, (int)px - h), y0 = std::max(0, (int)py - h);
        const unsigned x1 = std::min((int)g.W - (int)pr.k, (int)px + h);
        const unsigned y1 = std::min((int)g.H - (int)pr.k, (int)py + h);
        std::vector<Cand> lst;
        lst.reserve((size_t)s * s);
        for (unsigned qy = y0; qy <= y1; qy++)
            for (unsigned qx = x0; qx <= x1; qx++) {
                const unsigned cur = g.index(qx, qy, pt, 0);
                if (!seen.insert(cur)) continue;
                n_eval++;
                const float d = dist(ref, cur) - ((qx == rx && qy == ry) ? pr.d : 0.f);
                lst.push_back(Cand(d, cur));
            }
        const int nb = std::min((int)pr.Nb, (int)lst.size());
        tie_now = tie_now || ranking_near_tie(dists(lst), nb);
        std::partial_sort(lst.begin(), lst.begin() + nb, lst.end(), cmp_first);
        for (int i = 0; i < nb; i++) out.push_back(lst[i]);
    }

    static std::vector<float> dists(const std::vector<Cand>& v) {
        std::vector<float> d(v.size());
        for (size_t i = 0; i < v.size(); i++) d[i] = v[i].first;
        return d;
    }

    // Flow projection of a match into the next (dir=+1) or previous frame
    // (src/vbm3d.cpp:1128-1131, 1152-1155); y uses the updated x.
    unsigned project(unsigned m, int dir) const {
        unsigned px, py, pt;
        g.coords(m, px, py, pt);
        const float* fl = dir > 0 ? ff : bf;
        const unsigned ft = dir > 0 ? pt : pt - 1;
        const unsigned h = pr.k / 2;
        int nx = (int)((float)px + std::round(flow_at(fl, g, px + h, py + h, ft, 0)));
        px = (unsigned)std::min(std::max(nx, 0), (int)(g.W - pr.k));
        int ny = (int)((float)py + std::round(flow_at(fl, g, px + h, py + h, ft, 1)));
        py = (unsigned)std::min(std::max(ny, 0), (int)(g.H - pr.k));
        return g.index(px, py, pt + dir, 0);
    }

    // computeSimilarPatches (src/vbm3d.cpp:1082-1197).  Returns nSx.
    unsigned run(unsigned pidx, float* od, unsigned* oi, unsigned* nsim_out) {
        seen.reset();
        tie_now = false;
        std::vector<Cand> best;
        best.reserve(pr.Nb * (2 * pr.Nf + 1));
        local(pidx, pidx, pr.Ns, best);
        std::vector<unsigned> pre(pr.Nb), post(pr.Nb);
        for (unsigned i = 0; i < pr.Nb; i++) pre[i] = post[i] = best[i].second;
        unsigned rx, ry, rt;
        g.coords(pidx, rx, ry, rt);
        std::vector<Cand> fb;
        const int nfw = (int)(std::min(rt + pr.Nf, g.F - pr.kt) - rt);
        for (int f = 0; f < nfw; f++) {
            fb.clear();
            for (unsigned m = 0; m < pr.Nb; m++) local(project(post[m], +1), pidx, pr.Npr, fb);
            const int nb = std::min((int)pr.Nb, (int)fb.size());
            tie_now = tie_now || ranking_near_tie(dists(fb), nb);
            std::partial_sort(fb.begin(), fb.begin() + nb, fb.end(), cmp_first);
            for (int i = 0; i < nb; i++) { post[i] = fb[i].second; best.push_back(fb[i]); }
        }
        const int nbw = (int)rt - std::max((int)rt - (int)pr.Nf, 0);
        for (int f = 0; f < nbw; f++) {
            fb.clear();
            for (unsigned m = 0; m < pr.Nb; m++) local(project(pre[m], -1), pidx, pr.Npr, fb);
            const int nb = std::min((int)pr.Nb, (int)fb.size());
            tie_now = tie_now || ranking_near_tie(dists(fb), nb);
            std::partial_sort(fb.begin(), fb.begin() + nb, fb.end(), cmp_first);
            for (int i = 0; i < nb; i++) { pre[i] = fb[i].second; best.push_back(fb[i]); }
        }
        const unsigned nsim = std::min(pr.N, (unsigned)best.size());
        tie_now = tie_now || ranking_near_tie(dists(best), nsim);
        near_tie += tie_now;
        std::partial_sort(best.begin(), best.begin() + nsim, best.end(), cmp_first);
        for (unsigned n = 0; n < nsim; n++) { od[n] = best[n].first; oi[n] = best[n].second; }
        bool nt = false;
        for (unsigned n = 1; n < nsim; n++) nt = nt || near_rel(od[n], pr.tau);
        near_tau += nt;
        unsigned it = nsim - 1;
        while (od[it] > pr.tau && it > 0) it--;
        if (nsim_out) *nsim_out = nsim;
        return pow2_floor(it + 1);
    }
};

// ---------------------------------------------------------------------
// Validation shared by every entry point.
// ---------------------------------------------------------------------
bool check_params(const orc_params& p, const Geo& g, int step) {
    char buf[256];
    auto fail = [&](const char* m) { std::snprintf(buf, sizeof buf, "step %d: %s", step, m); g_err = buf; return false; };
    if (p.k == 0) return fail("k must be > 0");
    if (p.kt == 0 || p.kt > 8) return fail("kt must be in [1, 8]");
    if (p.kt > g.F) return fail("fewer frames than kt");
    if (p.k > 16) return fail("k > 16 is not supported");
    if (p.N == 0 || p.N > 32) return fail("N must be in [1, 32]");
    if (p.Nb == 0 || p.Nb > 16) return fail("Nb must be in [1, 16]");
    if (p.Ns == 0 || p.Npr == 0 || p.Ns > 31 || p.Npr > 31) return fail("Ns and Npr must be in [1, 31]");
    if (p.p == 0) return fail("p must be > 0");
    if (p.T_2D != kDCT && p.T_2D != kBIOR) return fail("T_2D must be DCT (0) or BIOR (1)");
    if (p.T_2D == kBIOR && p.k != 8) return fail("BIOR needs k == 8");
    if (g.W < p.k || g.H < p.k) return fail("frame smaller than a patch");
    if (g.F < 1) return fail("no frames");
    if (g.C != 1 && g.C != 3) return fail("only 1 or 3 channels are supported");
    if ((uint64_t)g.W * g.H * g.C * g.F >= (1ull << 32)) return fail("video too large for 32-bit indices");
    // the first window must hold Nb candidates (src/vbm3d.cpp:1115-1119)
    const unsigned h = (p.Ns + 1) / 2;
    if (std::min(g.W - p.k + 1, h) * std::min(g.H - p.k + 1, h) < p.Nb)
        return fail("search window smaller than Nb");
    return true;
}

// Largest distance of a group member from its reference patch (sizes the
// crop's aggregation box), in 64 bits and capped at the frame size: a
// projection is clamped to the frame, so no member lies farther.
unsigned reach(const orc_params& p, const float* ff, const float* bf, const Geo& g) {
    float mf = 0.f;
    const size_t n = (size_t)(g.F > 1 ? g.F - 1 : 0) * g.wh * 2;
    for (const float* fl : {ff, bf})
        if (fl)
            for (size_t i = 0; i < n; i++) mf = std::max(mf, std::fmin(std::fabs(std::round(fl[i])), 65535.f));
    unsigned long long r = (p.Ns - 1) / 2 + (unsigned long long)p.Nf * ((p.Npr - 1) / 2 + (unsigned long long)mf);
    // a motion-compensated trajectory drifts by at most the rounded sum of
    // kt - 1 flow steps
    if (p.mc && p.kt > 1) r += (unsigned long long)(p.kt - 1) * ((unsigned long long)mf + 1);
    return (unsigned)std::min<unsigned long long>(r, std::max(g.W, g.H));
}

// ---------------------------------------------------------------------
// One step over one crop (vbm3d_1st_step src/vbm3d.cpp:447-654,
// vbm3d_2nd_step :676-928), aggregating into a box-limited buffer.
// ---------------------------------------------------------------------
struct Box { unsigned x0, y0, w, h; };

template <class T>
struct CropOut {
    Box box;
    std::vector<T> num, den;  // [F][C][box.h][box.w]
};

struct StepIn {
    int step;
    const float *noisy, *basic, *ff, *bf;
    Geo g;
    orc_params pr;
    float sig[3];
    // references, candidates evaluated, group members, references with a
    // ranking decision near a tie, references with a group candidate near
    // tau, hard-threshold coefficients near the threshold, coefficients
    // tested against it (the 3D DC is exempt)
    long long stats[7];
};

template <class T>
void run_crop(StepIn& in, const std::vector<unsigned>& rows,
              const std::vector<unsigned>& cols, unsigned R, CropOut<T>& out,
              bool replay, std::vector<uint8_t>* keep_log, size_t* keep_pos,
              std::vector<uint32_t>* grp_log, size_t* grp_pos) {
    const Geo& g = in.g;
    const orc_params& pr = in.pr;
    const unsigned k = pr.k, k2 = k * k, kt = pr.kt, k2t = k2 * kt, C = g.C, N = pr.N;
    const bool temporal = kt > 1 && pr.T_2D != kBIOR;  // src/vbm3d.cpp:573, 818
    Tables tables(k, kt);
    TabT<float> tf = make_tab_f(tables);
    TabT<double> td = make_tab_d(tables);
    const TabT<T>& tb = *reinterpret_cast<const TabT<T>*>(
        std::is_same<T, float>::value ? (const void*)&tf : (const void*)&td);

    Box& b = out.box;
    b.x0 = cols.front() > R ? cols.front() - R : 0;
    b.y0 = rows.front() > R ? rows.front() - R : 0;
    const unsigned x1 = std::min(g.W, cols.back() + k + R), y1 = std::min(g.H, rows.back() + k + R);
    b.w = x1 - b.x0;
    b.h = y1 - b.y0;
    out.num.assign((size_t)g.F * C * b.w * b.h, (T)0);
    out.den.assign(out.num.size(), (T)0);

    const float* mvid = in.step == 1 ? in.noisy : in.basic;
    BM bm(mvid, g, in.ff, in.bf, pr);
    std::vector<float> dist(N);
    std::vector<unsigned> idx(N);
    std::vector<T> tz((size_t)C * N * k2t), te((size_t)C * N * k2t), gz, ge, patch(k2);
    std::vector<uint8_t> keep((size_t)C * N * k2t);
    T w[3];
    unsigned trx[16], tr_y[16];

    for (unsigned t = 0; t + pr.kt <= g.F; ++t)
        for (unsigned yr : rows)
            for (unsigned xr : cols) {
                const unsigned pidx = g.index(xr, yr, t, 0);
                unsigned nsx;
                if (replay) {
                    nsx = (*grp_log)[(*grp_pos)++];
                    for (unsigned n = 0; n < nsx; n++) idx[n] = (*grp_log)[(*grp_pos)++];
                } else {
                    nsx = bm.run(pidx, dist.data(), idx.data(), nullptr);
                    if (grp_log) {
                        grp_log->push_back(nsx);
                        for (unsigned n = 0; n < nsx; n++) grp_log->push_back(idx[n]);
                    }
                }
                in.stats[0]++;
                in.stats[2] += nsx;
                // 2D transforms of every slice of the group members along
                // their trajectories (dct_2d_process :1218-1304,
                // bior_2d_process :1325-1386), table[c][n][slice][kk]
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++) {
                        unsigned px, py, pt;
                        g.coords(idx[n], px, py, pt);
                        trajectory(px, py, pt, k, kt, pr.mc != 0, in.ff, g, trx, tr_y);
                        for (int which = 0; which < (in.step == 2 ? 2 : 1); which++) {
                            const float* src = which == 0 ? in.noisy : in.basic;
                            for (unsigned ht = 0; ht < kt; ht++) {
                                for (unsigned p = 0; p < k; p++)
                                    for (unsigned q = 0; q < k; q++)
                                        patch[p * k + q] = (T)src[g.index(trx[ht] + q, tr_y[ht] + p, pt + ht, c)];
                                T* dst = (which == 0 ? tz.data() : te.data()) + (size_t)(c * N + n) * k2t + ht * k2;
                                if (pr.T_2D == kBIOR) bior_fwd(patch.data(), dst, tb);
                                else dct_fwd(patch.data(), dst, tb);
                            }
                        }
                    }
                // group layout [c][kk][n], kk = slice * k2 + pixel
                // (:565-570, :803-815)
                gz.assign((size_t)C * k2t * nsx, (T)0);
                ge.assign(gz.size(), (T)0);
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++)
                        for (unsigned kk = 0; kk < k2t; kk++) {
                            gz[n + kk * nsx + c * k2t * nsx] = tz[(size_t)(c * N + n) * k2t + kk];
                            if (in.step == 2) ge[n + kk * nsx + c * k2t * nsx] = te[(size_t)(c * N + n) * k2t + kk];
                        }
                // transform along each coefficient's slices (:573-575, :816-822)
                if (temporal)
                    for (T* grp : {gz.data(), ge.data()}) {
                        if (grp == ge.data() && in.step == 1) break;
                        for (unsigned c = 0; c < C; c++)
                            for (unsigned n = 0; n < nsx; n++)
                                for (unsigned kk = 0; kk < k2; kk++)
                                    temporal_fwd(grp + n + kk * nsx + c * k2t * nsx, k2 * nsx, tb);
                    }
                T* est;
                if (in.step == 1) {
                    uint8_t* kp = keep.data();
                    if (replay) { kp = keep_log->data() + *keep_pos; *keep_pos += (size_t)C * k2t * nsx; }
                    ht_filter(gz.data(), nsx, k2t, C, in.sig, pr.lambda3D, pr.T_3D, w, kp, replay, &in.stats[5]);
                    in.stats[6] += (long long)C * k2t * nsx - C;
                    if (!replay && keep_log) keep_log->insert(keep_log->end(), kp, kp + (size_t)C * k2t * nsx);
                    est = gz.data();
                } else {
                    wiener_filter(gz.data(), ge.data(), nsx, k2t, C, in.sig, pr.T_3D, w);
                    est = ge.data();
                }
                if (temporal)  // :589, :829-831
                    for (unsigned c = 0; c < C; c++)
                        for (unsigned n = 0; n < nsx; n++)
                            for (unsigned kk = 0; kk < k2; kk++)
                                temporal_inv(est + n + kk * nsx + c * k2t * nsx, k2 * nsx, tb);
                // the group chunks in [c][n][kk] order (:592-595, :841-844),
                // inverse 2D transform per slice (:603-609, :852-858)
                std::vector<T> chunk((size_t)C * nsx * k2t);
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++)
                        for (unsigned kk = 0; kk < k2t; kk++)
                            chunk[(size_t)(c * nsx + n) * k2t + kk] = est[n + kk * nsx + c * k2t * nsx];
                for (unsigned a = 0; a < C * nsx * kt; a++) {
                    if (pr.T_2D == kBIOR) bior_inv(chunk.data() + (size_t)a * k2, tb);
                    else dct_inv(chunk.data() + (size_t)a * k2, tb);
                }
                // aggregation along the members' trajectories (:612-650,
                // :860-924)
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++) {
                        unsigned px, py, pt;
                        g.coords(idx[n], px, py, pt);
                        trajectory(px, py, pt, k, kt, pr.mc != 0, in.ff, g, trx, tr_y);
                        for (unsigned ht = 0; ht < kt; ht++) {
                            const unsigned sx = trx[ht], sy = tr_y[ht];
                            if (sx < b.x0 || sy < b.y0 || sx + k > b.x0 + b.w || sy + k > b.y0 + b.h)
                                throw std::runtime_error("aggregation outside the crop box");
                            const T* e = chunk.data() + (size_t)(c * nsx + n) * k2t + ht * k2;
                            for (unsigned p = 0; p < k; p++)
                                for (unsigned q = 0; q < k; q++) {
                                    const size_t o = (((size_t)(pt + ht) * C + c) * b.h + (sy + p - b.y0)) * b.w +
                                                     (sx + q - b.x0);
                                    out.num[o] += tb.kaiser[p * k + q] * w[c] * e[p * k + q];
                                    out.den[o] += tb.kaiser[p * k + q] * w[c];
                                }
                        }
                    }
            }
    in.stats[1] += bm.n_eval;
    in.stats[3] += bm.near_tie;
    in.stats[4] += bm.near_tau;
}

// Full step: grid, crops, per-crop work (OpenMP over crops like
// src/vbm3d.cpp:254-265), then the ordered merge (:267-280).
template <class T>
int run_step(StepIn& in, int nb, int nthreads, T* num, T* den, bool replay,
             std::vector<std::vector<uint8_t>>* keeps,
             std::vector<std::vector<uint32_t>>* grps) {
    const Geo& g = in.g;
    Grid grid;
    if (!make_grid(grid, g.W, g.H, in.pr.k, in.pr.p, in.pr.n, nb)) {
        g_err = "invalid reference grid for these sizes";
        return -1;
    }
    const unsigned R = reach(in.pr, in.ff, in.bf, g);
    const int ncrop = (int)grid.rows.size();
    std::vector<CropOut<T>> outs(ncrop);
    std::vector<StepIn> ins(ncrop, in);
    std::string err;
    if (keeps && !replay) { keeps->assign(ncrop, {}); grps->assign(ncrop, {}); }
    const int nt = nthreads > 0 ? nthreads : ncrop;
    (void)nt;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int c = 0; c < ncrop; c++) {
        try {
            size_t kp = 0, gp = 0;
            run_crop<T>(ins[c], grid.rows[c], grid.cols[c], R, outs[c], replay,
                        keeps ? &(*keeps)[c] : nullptr, &kp, grps ? &(*grps)[c] : nullptr, &gp);
        } catch (const std::exception& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) { g_err = err; return -1; }
    for (int c = 0; c < ncrop; c++)
        for (int s = 0; s < 7; s++) in.stats[s] += ins[c].stats[s];
    const unsigned C = g.C;
    for (unsigned t = 0; t < g.F; t++)
        for (unsigned ch = 0; ch < C; ch++)
            for (unsigned y = 0; y < g.H; y++)
                for (unsigned x = 0; x < g.W; x++) {
                    const size_t o = g.index(x, y, t, ch);
                    T sn = 0, sd = 0;
                    for (int c = 0; c < ncrop; c++) {
                        const Box& b = outs[c].box;
                        T vn = 0, vd = 0;
                        if (x >= b.x0 && x < b.x0 + b.w && y >= b.y0 && y < b.y0 + b.h) {
                            const size_t i = (((size_t)t * C + ch) * b.h + (y - b.y0)) * b.w + (x - b.x0);
                            vn = outs[c].num[i];
                            vd = outs[c].den[i];
                        }
                        if (ncrop == 1) { sn = vn; sd = vd; }  // nb == 1: no merge loop
                        else { sn += vn; sd += vd; }
                    }
                    num[o] = sn;
                    den[o] = sd;
                }
    return 0;
}

// ---------------------------------------------------------------------
// Adjoint of step 2 with the groups, the Wiener gains and the weights
// frozen (FP64).  With the basic estimate fixed, step 2 is linear in the
// noisy video:  final = D^-1 sum_g Agg_g (kaiser w_g) G T3^-1 V_g T3 F Gather_g z,
// D = den.  Its transpose gathers u = grad/den with kaiser w_g, applies
// G^T, T3^T V_g T3^-T (= T3 V_g T3^-1 for the orthonormal Haar and the
// symmetric Hadamard) and F^T, and scatters back to the member patches.
// ---------------------------------------------------------------------
// Transposes of the 2D transforms as explicit double sums.  Patches are
// row-major x[p][q]; coefficient layouts are the forward transforms'.
void dct_fwd_T(const double* c, double* y, const TabT<double>& tb) {
    // forward: out = nrm o (F x F^T)  =>  x' = F^T (nrm o c) F
    const unsigned k = tb.k;
    for (unsigned p = 0; p < k; p++)
        for (unsigned q = 0; q < k; q++) {
            double m = 0;
            for (unsigned u = 0; u < k; u++)
                for (unsigned v = 0; v < k; v++)
                    m += tb.dctF[u * k + p] * tb.nrm[u * k + v] * c[u * k + v] * tb.dctF[v * k + q];
            y[p * k + q] = m;
        }
}

void dct_inv_T(const double* r, double* y, const TabT<double>& tb) {
    // inverse: pix = dct_out (I (c o inv) I^T)  =>  c' = inv o dct_out (I^T r I)
    const unsigned k = tb.k;
    for (unsigned u = 0; u < k; u++)
        for (unsigned v = 0; v < k; v++) {
            double m = 0;
            for (unsigned p = 0; p < k; p++)
                for (unsigned q = 0; q < k; q++)
                    m += tb.dctI[p * k + u] * r[p * k + q] * tb.dctI[q * k + v];
            y[u * k + v] = tb.inv[u * k + v] * tb.dct_out * m;
        }
}

void bior_fwd_T(const double* c, double* y, const TabT<double>& tb) {
    // forward: out = B x^T B^T (B = analysis)  =>  x'[s][q] = sum c[i][j] B[j][s] B[i][q]
    for (unsigned s = 0; s < 8; s++)
        for (unsigned q = 0; q < 8; q++) {
            double m = 0;
            for (unsigned i = 0; i < 8; i++)
                for (unsigned j = 0; j < 8; j++) m += c[i * 8 + j] * tb.biorF[j * 8 + s] * tb.biorF[i * 8 + q];
            y[s * 8 + q] = m;
        }
}

void bior_inv_T(const double* r, double* y, const TabT<double>& tb) {
    // inverse: pix = S c^T S^T (S = synthesis)  =>  c'[t][s] = sum r[j][i] S[j][s] S[i][t]
    for (unsigned t = 0; t < 8; t++)
        for (unsigned s = 0; s < 8; s++) {
            double m = 0;
            for (unsigned j = 0; j < 8; j++)
                for (unsigned i = 0; i < 8; i++) m += r[j * 8 + i] * tb.biorI[j * 8 + s] * tb.biorI[i * 8 + t];
            y[t * 8 + s] = m;
        }
}

// Adds the transpose of one crop's step-2 contribution, applied to u, to gz.
void adjoint_crop(StepIn& in, const std::vector<unsigned>& rows, const std::vector<unsigned>& cols,
                  const double* u, double* gz) {
    const Geo& g = in.g;
    const orc_params& pr = in.pr;
    const unsigned k = pr.k, k2 = k * k, C = g.C, N = pr.N;
    const bool had = pr.T_3D == kHADAMARD, bior = pr.T_2D == kBIOR;
    Tables tables(k);
    const TabT<double> tb = make_tab_d(tables);
    BM bm(in.basic, g, in.ff, in.bf, pr);
    std::vector<float> dist(N);
    std::vector<unsigned> idx(N);
    std::vector<double> patch(k2), coef(k2), e((size_t)C * k2 * N), gain((size_t)C * k2 * N),
        r((size_t)C * k2 * N), tmp(k2);
    for (unsigned t = 0; t + pr.kt <= g.F; ++t)
        for (unsigned yr : rows)
            for (unsigned xr : cols) {
                const unsigned nsx = bm.run(g.index(xr, yr, t, 0), dist.data(), idx.data(), nullptr);
                // Wiener gains and weights from the basic estimate (frozen)
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++) {
                        unsigned px, py, pt;
                        g.coords(idx[n], px, py, pt);
                        for (unsigned p = 0; p < k; p++)
                            for (unsigned q = 0; q < k; q++)
                                patch[p * k + q] = in.basic[g.index(px + q, py + p, pt, c)];
                        if (bior) bior_fwd(patch.data(), coef.data(), tb);
                        else dct_fwd(patch.data(), coef.data(), tb);
                        for (unsigned kk = 0; kk < k2; kk++) e[n + kk * nsx + c * k2 * nsx] = coef[kk];
                    }
                double w[3];
                for (unsigned c = 0; c < C; c++) {
                    const double s2 = (double)in.sig[c] * (double)in.sig[c];
                    const double cf = 1.0 / (double)nsx;
                    double W2 = 0;
                    for (unsigned kk = 0; kk < k2; kk++) {
                        double* ev = e.data() + kk * nsx + c * k2 * nsx;
                        if (had) hadamard(ev, nsx);
                        else haar_fwd(ev, nsx);
                        for (unsigned n = 0; n < nsx; n++) {
                            double v = ev[n] * ev[n];
                            if (had) v *= cf;
                            v /= (v + s2);
                            W2 += v * v;
                            gain[n + kk * nsx + c * k2 * nsx] = had ? v * cf : v;
                        }
                    }
                    w[c] = had ? (W2 > 0 ? 1.0 / (s2 * W2) : 1.0) : 1.0 / (s2 * W2);
                }
                // Agg^T then G^T, into group layout [c][kk][n]
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++) {
                        unsigned px, py, pt;
                        g.coords(idx[n], px, py, pt);
                        for (unsigned p = 0; p < k; p++)
                            for (unsigned q = 0; q < k; q++)
                                patch[p * k + q] = tb.kaiser[p * k + q] * w[c] * u[g.index(px + q, py + p, pt, c)];
                        if (bior) bior_inv_T(patch.data(), coef.data(), tb);
                        else dct_inv_T(patch.data(), coef.data(), tb);
                        for (unsigned kk = 0; kk < k2; kk++) r[n + kk * nsx + c * k2 * nsx] = coef[kk];
                    }
                // (T3^-1 V T3)^T = T3^-T V T3^T: Haar forward, gains, Haar inverse
                // (orthonormal); Hadamard is symmetric
                for (unsigned a = 0; a < k2 * C; a++) {
                    double* rv = r.data() + a * nsx;
                    if (had) hadamard(rv, nsx);
                    else haar_fwd(rv, nsx);
                    for (unsigned n = 0; n < nsx; n++) rv[n] *= gain[a * nsx + n];
                    if (had) hadamard(rv, nsx);
                    else haar_inv(rv, nsx);
                }
                // F^T and the scatter back to the member patches
                for (unsigned c = 0; c < C; c++)
                    for (unsigned n = 0; n < nsx; n++) {
                        unsigned px, py, pt;
                        g.coords(idx[n], px, py, pt);
                        for (unsigned kk = 0; kk < k2; kk++) coef[kk] = r[n + kk * nsx + c * k2 * nsx];
                        if (bior) bior_fwd_T(coef.data(), tmp.data(), tb);
                        else dct_fwd_T(coef.data(), tmp.data(), tb);
                        for (unsigned p = 0; p < k; p++)
                            for (unsigned q = 0; q < k; q++) gz[g.index(px + q, py + p, pt, c)] += tmp[p * k + q];
                    }
            }
}

bool prepare(StepIn& in, int step, float sigma, const float* noisy,
             const float* basic, const float* ff, const float* bf, unsigned W,
             unsigned H, unsigned F, unsigned C, const orc_params* prm,
             int color_space) {
    in.step = step;
    in.noisy = noisy;
    in.basic = basic;
    in.ff = ff;
    in.bf = bf;
    in.g = Geo(W, H, F, C);
    in.pr = *prm;
    for (auto& s : in.stats) s = 0;
    if (!check_params(in.pr, in.g, step)) return false;
    if (!sigma_table(sigma, C, color_space ? 1u : 0u, in.sig)) {
        g_err = "sigma table";
        return false;
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_init_params_1(orc_params* p, int k, int kt, int Nf, int Ns, int Npr,
                       int Nb, int pstep, int N, int d, float tau,
                       float lambda3D, unsigned T_2D, unsigned T_3D, float sigma) {
    // initializeParameters_1 (src/main.cpp:35-125)
    p->k = k < 0 ? 8 : k;
    p->kt = kt < 0 ? 1 : kt;
    p->Nf = Nf < 0 ? 4 : Nf;
    p->Ns = Ns < 0 ? 7 : Ns;
    p->Npr = Npr < 0 ? 5 : Npr;
    p->Nb = Nb < 0 ? 2 : Nb;
    p->d = d < 0 ? (float)((7. * 7. * 255.) / (p->k * p->k * p->kt))
                 : (float)((d * d * 255.) / (p->k * p->k * p->kt));
    p->p = pstep < 0 ? p->k / 4 * 3 : pstep;
    p->N = N < 0 ? 8 : N;
    p->tau = tau < 0 ? ((sigma > 30) ? 4500 : 3000) : tau;
    p->lambda3D = lambda3D < 0 ? 2.7f : lambda3D;
    if (T_2D == (unsigned)-1) p->T_2D = p->k == 8 ? kBIOR : kDCT;
    else p->T_2D = p->k == 8 ? T_2D : kDCT;
    p->T_3D = T_3D == (unsigned)-1 ? 3 : T_3D;
    p->mc = 0;
    p->n = 16;
}

void orc_init_params_2(orc_params* p, int k, int kt, int Nf, int Ns, int Npr,
                       int Nb, int pstep, int N, int d, float tau,
                       unsigned T_2D, unsigned T_3D, float sigma) {
    // initializeParameters_2 (src/main.cpp:127-207)
    p->k = k < 0 ? ((sigma > 30) ? 8 : 7) : k;
    p->kt = kt < 0 ? 1 : kt;
    p->Nf = Nf < 0 ? 4 : Nf;
    p->Ns = Ns < 0 ? 7 : Ns;
    p->Npr = Npr < 0 ? 5 : Npr;
    p->Nb = Nb < 0 ? 2 : Nb;
    p->d = d < 0 ? (float)((3. * 3. * 255.) / (p->k * p->k * p->kt))
                 : (float)((d * d * 255.) / (p->k * p->k * p->kt));
    p->p = pstep < 0 ? p->k / 2 : pstep;
    p->N = N < 0 ? 8 : N;
    p->tau = tau < 0 ? ((sigma > 30) ? 3000 : 1500) : tau;
    p->lambda3D = 0.f;  // default member initialiser, src/parameters.h:51
    if (T_2D == (unsigned)-1) p->T_2D = kDCT;
    else p->T_2D = p->k == 8 ? T_2D : kDCT;
    p->T_3D = T_3D == (unsigned)-1 ? 3 : T_3D;
    p->mc = 0;
    p->n = 16;
}

long orc_grid(unsigned W, unsigned H, unsigned k, unsigned p, int n, int nb,
              unsigned* xs, unsigned* ys, unsigned* crop) {
    Grid g;
    if (!make_grid(g, W, H, k, p, n, nb)) return -1;
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

int orc_block_match(const float* vid, unsigned W, unsigned H, unsigned F,
                    unsigned C, const float* fflow, const float* bflow,
                    const orc_params* prm, long nrefs, const uint32_t* ref_idx,
                    uint32_t* idx_out, float* dist_out, uint32_t* nsim_out,
                    uint32_t* nsx_out) {
    Geo g(W, H, F, C);
    if (!check_params(*prm, g, 0)) return -1;
    BM bm(vid, g, fflow, bflow, *prm);
    const unsigned N = prm->N;
    std::vector<float> d(N);
    std::vector<unsigned> ix(N);
    for (long r = 0; r < nrefs; r++) {
        unsigned nsim = 0;
        const unsigned nsx = bm.run(ref_idx[r], d.data(), ix.data(), &nsim);
        for (unsigned n = 0; n < N; n++) {
            idx_out[r * N + n] = n < nsim ? ix[n] : 0xFFFFFFFFu;
            dist_out[r * N + n] = n < nsim ? d[n] : NAN;
        }
        nsim_out[r] = nsim;
        nsx_out[r] = nsx;
    }
    return 0;
}

int orc_step(int step, float sigma, const float* noisy, const float* basic,
             const float* fflow, const float* bflow, unsigned W, unsigned H,
             unsigned F, unsigned C, const orc_params* prm, int color_space,
             int nb, int nthreads, float* num, float* den) {
    StepIn in;
    if (!prepare(in, step, sigma, noisy, basic, fflow, bflow, W, H, F, C, prm, color_space)) return -1;
    const int rc = run_step<float>(in, nb, nthreads, num, den, false, nullptr, nullptr);
    if (rc == 0 && (step == 1 || step == 2))
        for (int s = 0; s < 7; s++) g_stats[step][s] = in.stats[s];
    return rc;
}

void orc_last_stats(int step, long long* out) {
    for (int s = 0; s < 7; s++) out[s] = (step == 1 || step == 2) ? g_stats[step][s] : 0;
}

int orc_step_fp64(int step, float sigma, const float* noisy, const float* basic,
                  const float* fflow, const float* bflow, unsigned W,
                  unsigned H, unsigned F, unsigned C, const orc_params* prm,
                  int color_space, double* num, double* den) {
    StepIn in;
    if (!prepare(in, step, sigma, noisy, basic, fflow, bflow, W, H, F, C, prm, color_space)) return -1;
    std::vector<std::vector<uint8_t>> keeps;
    std::vector<std::vector<uint32_t>> grps;
    const size_t n = (size_t)W * H * F * C;
    std::vector<float> fn(n), fd(n);
    StepIn in32 = in;
    if (run_step<float>(in32, 1, 1, fn.data(), fd.data(), false, &keeps, &grps)) return -1;
    return run_step<double>(in, 1, 1, num, den, true, &keeps, &grps);
}

int orc_step2_adjoint_fp64(float sigma, const float* noisy, const float* basic,
                           const float* fflow, const float* bflow, unsigned W,
                           unsigned H, unsigned F, unsigned C, const orc_params* prm,
                           int color_space, int nb, const double* grad_out,
                           double* grad_noisy, double* final_out) {
    StepIn in;
    if (!basic) { g_err = "step 2 needs the basic estimate"; return -1; }
    if (prm->kt != 1) { g_err = "the adjoint is restated for kt == 1"; return -1; }
    if (!prepare(in, 2, sigma, noisy, basic, fflow, bflow, W, H, F, C, prm, color_space)) return -1;
    Grid grid;
    if (!make_grid(grid, W, H, prm->k, prm->p, prm->n, nb)) { g_err = "invalid reference grid"; return -1; }
    // den (and the forward value) in FP64 with the FP32 run's groups
    std::vector<std::vector<uint8_t>> keeps;
    std::vector<std::vector<uint32_t>> grps;
    const size_t n = (size_t)W * H * F * C;
    std::vector<float> fn(n), fd(n);
    std::vector<double> num(n), den(n), u(n);
    StepIn in32 = in;
    if (run_step<float>(in32, nb, 1, fn.data(), fd.data(), false, &keeps, &grps)) return -1;
    StepIn in64 = in;
    if (run_step<double>(in64, nb, 1, num.data(), den.data(), true, &keeps, &grps)) return -1;
    for (size_t i = 0; i < n; i++) {
        u[i] = den[i] != 0 ? grad_out[i] / den[i] : 0.0;
        grad_noisy[i] = 0.0;
        if (final_out)
            final_out[i] = den[i] != 0 ? num[i] / den[i] : (nb == 1 ? (double)basic[i] : (double)noisy[i]);
    }
    try {
        for (size_t c = 0; c < grid.rows.size(); c++) adjoint_crop(in, grid.rows[c], grid.cols[c], u.data(), grad_noisy);
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
    // den == 0 keeps the noisy value when nb > 1 (src/vbm3d.cpp:343)
    if (nb > 1)
        for (size_t i = 0; i < n; i++)
            if (den[i] == 0) grad_noisy[i] += grad_out[i];
    return 0;
}

int orc_run_vbm3d(float sigma, float* noisy, const float* fflow,
                  const float* bflow, unsigned W, unsigned H, unsigned F,
                  unsigned C, const orc_params* p1, const orc_params* p2,
                  int color_space, int nb, int nthreads, float* basic,
                  float* denoised) {
    // run_vbm3d (src/vbm3d.cpp:87-376)
    Geo g(W, H, F, C);
    const size_t n = (size_t)W * H * F * C;
    if (C != 1 && C != 3) { g_err = "only 1 or 3 channels are supported"; return -1; }
    if (color_space) color_transform(p1->k > 0 ? noisy : basic, g, true);
    std::vector<float> num(n), den(n);
    if (p1->k > 0) {
        if (orc_step(1, sigma, noisy, nullptr, fflow, bflow, W, H, F, C, p1, color_space, nb, nthreads, num.data(), den.data()))
            return -1;
        for (size_t i = 0; i < n; i++) basic[i] = den[i] == 0 ? noisy[i] : num[i] / den[i];
    }
    if (p2->k > 0) {
        if (orc_step(2, sigma, noisy, basic, fflow, bflow, W, H, F, C, p2, color_space, nb, nthreads, num.data(), den.data()))
            return -1;
        // den == 0 falls back to basic single-threaded (:200) and to noisy
        // with the tiled path (:342-343)
        const float* fb = nb == 1 ? basic : noisy;
        for (size_t i = 0; i < n; i++) denoised[i] = den[i] == 0 ? fb[i] : num[i] / den[i];
    } else {
        for (size_t i = 0; i < n; i++) denoised[i] = basic[i];
    }
    if (color_space) {
        color_transform(denoised, g, false);
        color_transform(noisy, g, false);
        color_transform(basic, g, false);
    }
    return 0;
}

void orc_bior_fwd(const float* patch64, float* out64) {
    Tables t(8);
    TabT<float> tb = make_tab_f(t);
    bior_fwd(patch64, out64, tb);
}
void orc_bior_inv(float* sig64) {
    Tables t(8);
    TabT<float> tb = make_tab_f(t);
    bior_inv(sig64, tb);
}
void orc_dct_fwd(unsigned k, const float* patch, float* out) {
    Tables t(k);
    TabT<float> tb = make_tab_f(t);
    dct_fwd(patch, out, tb);
}
void orc_dct_inv(unsigned k, float* sig) {
    Tables t(k);
    TabT<float> tb = make_tab_f(t);
    dct_inv(sig, tb);
}
void orc_temporal_fwd(unsigned kt, float* v) {
    Tables t(8, kt);
    TabT<float> tb = make_tab_f(t);
    temporal_fwd(v, 1, tb);
}
void orc_temporal_inv(unsigned kt, float* v) {
    Tables t(8, kt);
    TabT<float> tb = make_tab_f(t);
    temporal_inv(v, 1, tb);
}
void orc_haar_fwd(float* v, unsigned n) { haar_fwd(v, n); }
void orc_haar_inv(float* v, unsigned n) { haar_inv(v, n); }
void orc_hadamard(float* v, unsigned n) { hadamard(v, n); }
void orc_kaiser(unsigned k, float* out) { kaiser_table(k, out); }
void orc_dct_norm(unsigned k, float* nrm, float* inv) { dct_norm(k, nrm, inv); }
int orc_sigma_table(float sigma, unsigned C, unsigned cs, float* out) {
    return sigma_table(sigma, C, cs, out) ? 0 : -1;
}
void orc_color(float* vid, unsigned W, unsigned H, unsigned F, unsigned C, int forward) {
    color_transform(vid, Geo(W, H, F, C), forward != 0);
}
void orc_partial_sort(float* dist, uint32_t* idx, unsigned len, unsigned mid) {
    std::vector<Cand> v(len);
    for (unsigned i = 0; i < len; i++) v[i] = Cand(dist[i], idx[i]);
    std::partial_sort(v.begin(), v.begin() + mid, v.end(), cmp_first);
    for (unsigned i = 0; i < len; i++) { dist[i] = v[i].first; idx[i] = v[i].second; }
}
void orc_psnr(const float* a, const float* b, size_t n, double* psnr, double* rmse) {
    // VideoUtils::computePSNR (src/Utilities/LibVideoT.hpp:577-597)
    auto clip = [](float v) { return v < 0.f ? 0.f : (v > 255.f ? 255.f : v); };
    double sum = 0.f;
    for (size_t k = 0; k < n; k++) {
        const float d = clip(a[k]) - clip(b[k]);
        sum += d * d;
    }
    *rmse = sqrtf(sum / (float)n);
    *psnr = 20.f * log10f(255.f / *rmse);
}

}  // extern "C"
