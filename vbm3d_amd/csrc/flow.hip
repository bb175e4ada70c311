// flow.hip -- CDNA4 (gfx950) TV-L1 optical flow, the reference's flow
// companion (tvl1flow/tvl1flow_lib.c; Zach, Pock and Bischof 2007 with
// Chambolle's dual projection), and its C-ABI entry points.
//
// The pyramid, warps and the primal-dual iterations run as element-wise
// and stencil kernels over [ny][nx] float planes in HBM, one thread per
// pixel.  Every kernel evaluates the reference's expressions in its order
// and precision -- float where it stores float, double where its
// expressions promote (Gaussian taps and sums, bicubic cells, the
// normalisation, 1.0 + taut * |grad|) -- with no FMA contraction and
// IEEE-correct division and square root, so each pixel's value is the
// reference's sequential value.  The one reduction, the mean squared
// change that stops the iterations, is summed in double in a fixed order;
// the reference sums it in float, which can only move the stopping
// iteration when the two sums straddle epsilon^2.
//
// The iterations never return to the host one by one: a control word in
// device memory (active, iteration count) is updated by a one-workgroup
// kernel after each iteration, every iteration kernel exits at once when
// the word says the warp has converged, and the host launches iterations
// in batches and reads the word between batches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "vbm3d_hip.h"

#pragma clang fp contract(off)

struct vbm3d_ctx;

namespace vb {
// context access, defined in capi.hip
hipStream_t ctx_stream(vbm3d_ctx* c);
int ctx_device(vbm3d_ctx* c);
hipError_t ctx_flow_workspace(vbm3d_ctx* c, size_t bytes, void** p);
int ctx_fail(int code, const std::string& msg);
}  // namespace vb

namespace vb {
namespace {

constexpr int kMaxIter = 300;     // MAX_ITERATIONS (tvl1flow_lib.c:22)
constexpr int kMaxTaps = 64;      // Gaussian half-window bound of this implementation
constexpr int BX = 64, BY = 4;    // pixel blocks: one wavefront per row segment

struct Taps {
    double b[kMaxTaps];
    int size;
};

struct FlowCtl {
    int active;   // 1 while the current warp iterates
    int iter;     // iterations done in the current warp
    float error;  // last mean squared change
    unsigned done;  // blocks of k_iter_dual_ctl finished in this iteration
};

__device__ __forceinline__ int clampi(int x, int n, bool& out) {
    if (x < 0) {
        out = true;
        return 0;
    }
    if (x >= n) {
        out = true;
        return n - 1;
    }
    return x;
}

// Catmull-Rom cell (tvl1flow/bicubic_interpolation.c:120-131), double.
__device__ __forceinline__ double cell(double v0, double v1, double v2, double v3, double x) {
    const double c3 = 3.0 * (v1 - v2) + v3 - v0;
    const double c2 = 2.0 * v0 - 5.0 * v1 + 4.0 * v2 - v3 + x * c3;
    return v1 + 0.5 * x * (v2 - v0 + x * c2);
}

// The 4x4 neighbourhood of (uu, vv) with Neumann clamping; the row below
// the sample is clamped with the x direction's sign, as the reference
// does (bicubic_interpolation.c:156-232).  Returns false when a clamp
// happened (the sample lies outside).
struct Nbhd {
    int cx[4], cy[4];
    double fx, fy;
};
__device__ __forceinline__ bool nbhd(float uu, float vv, int nx, int ny, Nbhd& n) {
    const int sx = uu < 0 ? -1 : 1, sy = vv < 0 ? -1 : 1;
    bool out = false;
    const int iu = (int)uu, iv = (int)vv;
    const int x = clampi(iu, nx, out), y = clampi(iv, ny, out);
    n.cx[0] = clampi(iu - sx, nx, out);
    n.cy[0] = clampi(iv - sx, ny, out);
    n.cx[1] = x;
    n.cy[1] = y;
    n.cx[2] = clampi(iu + sx, nx, out);
    n.cy[2] = clampi(iv + sy, ny, out);
    n.cx[3] = clampi(iu + 2 * sx, nx, out);
    n.cy[3] = clampi(iv + 2 * sy, ny, out);
    n.fx = (double)(uu - (float)x);
    n.fy = (double)(vv - (float)y);
    return !out;
}

__device__ __forceinline__ float interp(const float* __restrict__ in, int nx, const Nbhd& n) {
    double col[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
        col[c] = cell(in[n.cx[c] + nx * n.cy[0]], in[n.cx[c] + nx * n.cy[1]], in[n.cx[c] + nx * n.cy[2]],
                      in[n.cx[c] + nx * n.cy[3]], n.fy);
    return (float)cell(col[0], col[1], col[2], col[3], n.fx);
}

// ---- normalisation to [0, 255] by the joint range (tvl1flow_lib.c:341-376)
__global__ void k_minmax(const float* __restrict__ a, const float* __restrict__ b, long n, float* part) {
    float mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        mn0 = fminf(mn0, a[i]);
        mx0 = fmaxf(mx0, a[i]);
        mn1 = fminf(mn1, b[i]);
        mx1 = fmaxf(mx1, b[i]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        mn0 = fminf(mn0, __shfl_xor(mn0, o, 64));
        mx0 = fmaxf(mx0, __shfl_xor(mx0, o, 64));
        mn1 = fminf(mn1, __shfl_xor(mn1, o, 64));
        mx1 = fmaxf(mx1, __shfl_xor(mx1, o, 64));
    }
    __shared__ float s[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s[w][0] = mn0;
        s[w][1] = mx0;
        s[w][2] = mn1;
        s[w][3] = mx1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            s[0][0] = fminf(s[0][0], s[q][0]);
            s[0][1] = fmaxf(s[0][1], s[q][1]);
            s[0][2] = fminf(s[0][2], s[q][2]);
            s[0][3] = fmaxf(s[0][3], s[q][3]);
        }
        for (int k = 0; k < 4; k++) part[blockIdx.x * 4 + k] = s[0][k];
    }
}

__global__ void k_normalize(const float* __restrict__ a, const float* __restrict__ b, float* oa, float* ob, long n,
                            const float* __restrict__ part, int nparts) {
    float mn = INFINITY, mx = -INFINITY;
    for (int q = 0; q < nparts; q++) {
        mn = fminf(mn, fminf(part[4 * q], part[4 * q + 2]));
        mx = fmaxf(mx, fmaxf(part[4 * q + 1], part[4 * q + 3]));
    }
    const float den = mx - mn;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (den > 0) {
            oa[i] = (float)(255.0 * (a[i] - mn) / den);
            ob[i] = (float)(255.0 * (b[i] - mn) / den);
        } else {
            oa[i] = a[i];
            ob[i] = b[i];
        }
    }
}

// ---- separable Gaussian, reflecting edges, double taps and sums
// (tvl1flow/mask.c:234-337): a line of n samples extended by `size` on
// each side, the left mirror excluding the edge sample, the right one
// including it.
__device__ __forceinline__ int reflect(int e, int n) { return e < 0 ? -e : (e >= n ? 2 * n - e - 1 : e); }

__global__ void k_gauss_rows(const float* __restrict__ in, float* out, int nx, int ny, Taps t) {
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j >= nx || i >= ny) return;
    const float* row = in + (size_t)i * nx;
    double sum = t.b[0] * (double)row[j];
    for (int k = 1; k < t.size; k++) sum += t.b[k] * ((double)row[reflect(j - k, nx)] + (double)row[reflect(j + k, nx)]);
    out[(size_t)i * nx + j] = (float)sum;
}

__global__ void k_gauss_cols(const float* __restrict__ in, float* out, int nx, int ny, Taps t) {
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j >= nx || i >= ny) return;
    double sum = t.b[0] * (double)in[(size_t)i * nx + j];
    for (int k = 1; k < t.size; k++)
        sum += t.b[k] * ((double)in[(size_t)reflect(i - k, ny) * nx + j] + (double)in[(size_t)reflect(i + k, ny) * nx + j]);
    out[(size_t)i * nx + j] = (float)sum;
}

// ---- bicubic resampling of a level (zoom.c:44-117): output (j1, i1)
// samples the input at (j1 / fx, i1 / fy); flows are then multiplied by
// `scale` (the 1 / zfactor of tvl1flow_lib.c:453-457) when apply.
__global__ void k_resample(const float* __restrict__ in, float* out, int nx, int ny, int nxx, int nyy, float fx,
                           float fy, float scale, int apply) {
    const int j1 = blockIdx.x * BX + threadIdx.x, i1 = blockIdx.y * BY + threadIdx.y;
    if (j1 >= nxx || i1 >= nyy) return;
    const float i2 = (float)i1 / fy, j2 = (float)j1 / fx;
    Nbhd n;
    (void)nbhd(j2, i2, nx, ny, n);
    const float g = interp(in, nx, n);
    out[(size_t)i1 * nxx + j1] = apply ? g * scale : g;
}

// ---- centred gradient with one-sided edges (mask.c:144-214)
__global__ void k_cgrad(const float* __restrict__ in, float* dx, float* dy, int nx, int ny) {
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j >= nx || i >= ny) return;
    const size_t k = (size_t)i * nx + j;
    const size_t jl = j > 0 ? k - 1 : k, jr = j < nx - 1 ? k + 1 : k;
    const size_t iu = i > 0 ? k - nx : k, id = i < ny - 1 ? k + nx : k;
    dx[k] = (float)(0.5 * (double)(in[jr] - in[jl]));
    dy[k] = (float)(0.5 * (double)(in[id] - in[iu]));
}

// ---- one warp's set-up (tvl1flow_lib.c:150-166): the target image and
// its derivatives at the flow-displaced positions (zero outside), |grad|^2
// and the constant part of the linearised residual.
__global__ void k_warp(const float* __restrict__ I0, const float* __restrict__ I1, const float* __restrict__ I1x,
                       const float* __restrict__ I1y, const float* __restrict__ u1, const float* __restrict__ u2,
                       float* wx, float* wy, float* grad, float* rho_c, int nx, int ny) {
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j >= nx || i >= ny) return;
    const size_t p = (size_t)i * nx + j;
    const float uu = (float)j + u1[p], vv = (float)i + u2[p];
    Nbhd n;
    float w = 0.f, gx = 0.f, gy = 0.f;
    if (nbhd(uu, vv, nx, ny, n)) {
        w = interp(I1, nx, n);
        gx = interp(I1x, nx, n);
        gy = interp(I1y, nx, n);
    }
    const float Ix2 = gx * gx, Iy2 = gy * gy;
    wx[p] = gx;
    wy[p] = gy;
    grad[p] = Ix2 + Iy2;
    rho_c[p] = w - gx * u1[p] - gy * u2[p] - I0[p];
}

// Backward-difference divergence with the reference's edge cases
// (mask.c:39-90).
__device__ __forceinline__ float div_at(const float* __restrict__ v1, const float* __restrict__ v2, int i, int j,
                                        int nx, int ny) {
    const size_t p = (size_t)i * nx + j;
    const bool L = j == 0, R = j == nx - 1, T = i == 0, B = i == ny - 1;
    if (!L && !R && !T && !B) return (v1[p] - v1[p - 1]) + (v2[p] - v2[p - nx]);
    if (T && L) return v1[0] + v2[0];
    if (T && R) return -v1[nx - 2] + v2[nx - 1];
    if (B && L) return v1[p] - v2[p - nx];
    if (B && R) return -v1[p - 1] - v2[p - nx];
    if (T) return v1[p] - v1[p - 1] + v2[p];
    if (B) return v1[p] - v1[p - 1] - v2[p - nx];
    if (L) return v1[p] + v2[p] - v2[p - nx];
    return -v1[p - 1] + v2[p] - v2[p - nx];
}

// ---- iteration, part 1 (tvl1flow_lib.c:170-231) at pixel (i, j):
// thresholding towards the data term, divergence of the dual field, primal
// update in place; returns the squared change of the flow.
__device__ __forceinline__ double primal_px(float* u1, float* u2, const float* __restrict__ wx,
                                            const float* __restrict__ wy, const float* __restrict__ grad,
                                            const float* __restrict__ rho_c, const float* __restrict__ p11,
                                            const float* __restrict__ p12, const float* __restrict__ p21,
                                            const float* __restrict__ p22, int i, int j, int nx, int ny, float l_t,
                                            float theta) {
    const size_t p = (size_t)i * nx + j;
    const float a = u1[p], b = u2[p], gx = wx[p], gy = wy[p], g = grad[p];
    const float rho = rho_c[p] + (gx * a + gy * b);
    float d1, d2;
    if (rho < -l_t * g) {
        d1 = l_t * gx;
        d2 = l_t * gy;
    } else if (rho > l_t * g) {
        d1 = -l_t * gx;
        d2 = -l_t * gy;
    } else if ((double)g < 1E-10) {
        d1 = d2 = 0.f;
    } else {
        const float fi = -rho / g;
        d1 = fi * gx;
        d2 = fi * gy;
    }
    const float v1 = a + d1, v2 = b + d2;
    const float n1 = v1 + theta * div_at(p11, p12, i, j, nx, ny);
    const float n2 = v2 + theta * div_at(p21, p22, i, j, nx, ny);
    u1[p] = n1;
    u2[p] = n2;
    return (double)((n1 - a) * (n1 - a) + (n2 - b) * (n2 - b));
}

// ---- iteration, part 2 (tvl1flow_lib.c:233-252) at pixel (i, j):
// forward gradient of the new flow and the projected dual update.
__device__ __forceinline__ void dual_px(const float* __restrict__ u1, const float* __restrict__ u2, float* p11,
                                        float* p12, float* p21, float* p22, int i, int j, int nx, int ny,
                                        float taut) {
    const size_t p = (size_t)i * nx + j;
    const float u1x = j < nx - 1 ? u1[p + 1] - u1[p] : 0.f, u1y = i < ny - 1 ? u1[p + nx] - u1[p] : 0.f;
    const float u2x = j < nx - 1 ? u2[p + 1] - u2[p] : 0.f, u2y = i < ny - 1 ? u2[p + nx] - u2[p] : 0.f;
    // |grad| in double: the squares of floats are exact, one rounding of
    // their sum and one of the root; as a float this is the C library's
    // hypot of the two values
    const float g1 = (float)sqrt((double)u1x * (double)u1x + (double)u1y * (double)u1y);
    const float g2 = (float)sqrt((double)u2x * (double)u2x + (double)u2y * (double)u2y);
    const float ng1 = (float)(1.0 + (double)(taut * g1)), ng2 = (float)(1.0 + (double)(taut * g2));
    p11[p] = (p11[p] + taut * u1x) / ng1;
    p12[p] = (p12[p] + taut * u1y) / ng1;
    p21[p] = (p21[p] + taut * u2x) / ng2;
    p22[p] = (p22[p] + taut * u2y) / ng2;
}

// Part 1 over the level; each workgroup (64 x 4 pixels) leaves its sum of
// squared changes in a fixed order: every row of 64 by a wavefront's
// butterfly, then the four rows in order.
__global__ void k_iter_primal(const FlowCtl* __restrict__ ctl, float* u1, float* u2, const float* __restrict__ wx,
                              const float* __restrict__ wy, const float* __restrict__ grad,
                              const float* __restrict__ rho_c, const float* __restrict__ p11,
                              const float* __restrict__ p12, const float* __restrict__ p21,
                              const float* __restrict__ p22, double* part, int nx, int ny, float l_t, float theta) {
    if (!ctl->active) return;  // uniform
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    double e = 0.0;
    if (j < nx && i < ny) e = primal_px(u1, u2, wx, wy, grad, rho_c, p11, p12, p21, p22, i, j, nx, ny, l_t, theta);
    for (int o = 32; o >= 1; o >>= 1) e += __shfl_xor(e, o, 64);
    __shared__ double s[BX * BY / 64];
    const int t = threadIdx.y * BX + threadIdx.x;
    if ((t & 63) == 0) s[t >> 6] = e;
    __syncthreads();
    if (t == 0) {
        double acc = 0.0;
        for (int q = 0; q < BX * BY / 64; q++) acc += s[q];
        part[blockIdx.y * gridDim.x + blockIdx.x] = acc;
    }
}

// Part 2 over the level.
__global__ void k_iter_dual(const FlowCtl* __restrict__ ctl, const float* __restrict__ u1,
                            const float* __restrict__ u2, float* p11, float* p12, float* p21, float* p22, int nx,
                            int ny, float taut) {
    if (!ctl->active) return;  // uniform
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j >= nx || i >= ny) return;
    dual_px(u1, u2, p11, p12, p21, p22, i, j, nx, ny, taut);
}

// ---- iteration, part 3: the mean squared change and the stopping test
// (tvl1flow_lib.c:159, 228): one workgroup sums the partials in a fixed
// order.
__global__ void k_iter_ctl(FlowCtl* ctl, const double* __restrict__ part, int nparts, float size, float eps2) {
    if (!ctl->active) return;  // uniform
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += part[q];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double s[16];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += s[q];
        float err = (float)t;
        err /= size;
        ctl->error = err;
        ctl->iter += 1;
        ctl->active = (err > eps2 && ctl->iter < kMaxIter) ? 1 : 0;
    }
}

__global__ void k_ctl_start(FlowCtl* ctl) {
    ctl->active = 1;
    ctl->iter = 0;
    ctl->error = INFINITY;
    ctl->done = 0;
}

// Parts 2 and 3 in one launch, for levels of at most kFuseBlocks blocks:
// the dual update, then the workgroup that finishes last (one atomic per
// workgroup on the control word) sums the partials and runs the stopping
// test exactly as k_iter_ctl does (256 threads, the same order), so the
// flow and the iteration count are unchanged.  No workgroup waits for
// another.
constexpr int kFuseBlocks = 1024;
__global__ void k_iter_dual_ctl(FlowCtl* ctl, const float* __restrict__ u1, const float* __restrict__ u2,
                                float* p11, float* p12, float* p21, float* p22, int nx, int ny, float taut,
                                const double* __restrict__ part, int nparts, float size, float eps2) {
    if (!ctl->active) return;  // uniform
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (j < nx && i < ny) dual_px(u1, u2, p11, p12, p21, p22, i, j, nx, ny, taut);
    const int t = threadIdx.y * BX + threadIdx.x;
    __shared__ int s_last;
    __syncthreads();
    if (t == 0) s_last = atomicAdd(&ctl->done, 1u) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (!s_last) return;
    double acc = 0.0;
    for (int q = t; q < nparts; q += BX * BY) acc += part[q];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double s[BX * BY / 64];
    if ((t & 63) == 0) s[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
        double tt = 0.0;
        for (int q = 0; q < BX * BY / 64; q++) tt += s[q];
        float err = (float)tt;
        err /= size;
        ctl->error = err;
        ctl->iter += 1;
        ctl->active = (err > eps2 && ctl->iter < kMaxIter) ? 1 : 0;
        ctl->done = 0;
    }
}

// Every iteration of one warp on a level of at most kSmallPx pixels (and
// kSmallBlocks blocks of 64 x 4), in one workgroup of 1024 threads: the
// same per-pixel updates as k_iter_primal and k_iter_dual, and the stopping
// test's sums in their order (a row of 64 pixels by a wavefront's
// butterfly, the four rows of a block in order, then the blocks' sums as
// k_iter_ctl adds them), so the flow and the iteration count are those of
// the three-launch form.  It replaces up to 3 x 300 launches and the host's
// reads of the control word; the count goes to *count.
constexpr int kSmallPx = 16384, kSmallBlocks = 256;
__global__ __launch_bounds__(1024) void k_iter_small(FlowCtl* ctl, float* u1, float* u2, const float* __restrict__ wx,
                                                     const float* __restrict__ wy, const float* __restrict__ grad,
                                                     const float* __restrict__ rho_c, float* p11, float* p12,
                                                     float* p21, float* p22, int nx, int ny, float l_t, float theta,
                                                     float taut, float size, float eps2, int* count) {
    __shared__ double s_row[kSmallBlocks * BY];
    __shared__ double s_blk[kSmallBlocks];
    __shared__ double s_w[4];
    __shared__ int s_active, s_iter;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int gx = (nx + BX - 1) / BX, gy = (ny + BY - 1) / BY, nblk = gx * gy;
    if (t == 0) {
        s_active = ctl->active;
        s_iter = ctl->iter;
    }
    for (;;) {
        __syncthreads();  // the previous iteration's dual updates and control word
        if (!s_active) break;  // workgroup-uniform
        // primal, one row of 64 pixels (block b, row r) per wavefront step
        for (int vr = w; vr < nblk * BY; vr += 16) {
            const int b = vr / BY, r = vr - b * BY;
            const int j = (b % gx) * BX + lane, i = (b / gx) * BY + r;
            double e = 0.0;
            if (j < nx && i < ny) e = primal_px(u1, u2, wx, wy, grad, rho_c, p11, p12, p21, p22, i, j, nx, ny, l_t, theta);
            for (int o = 32; o >= 1; o >>= 1) e += __shfl_xor(e, o, 64);
            if (lane == 0) s_row[vr] = e;
        }
        __syncthreads();
        for (int b = t; b < nblk; b += 1024) {
            double acc = 0.0;
            for (int q = 0; q < BY; q++) acc += s_row[b * BY + q];
            s_blk[b] = acc;
        }
        // dual, every pixel
        for (int q = t; q < nx * ny; q += 1024) {
            const int i = q / nx, j = q - i * nx;
            dual_px(u1, u2, p11, p12, p21, p22, i, j, nx, ny, taut);
        }
        __syncthreads();
        if (t < 256) {
            double acc = 0.0;
            for (int q = t; q < nblk; q += 256) acc += s_blk[q];
            for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (lane == 0) s_w[w] = acc;
        }
        __syncthreads();
        if (t == 0) {
            double tt = 0.0;
            for (int q = 0; q < 4; q++) tt += s_w[q];
            float err = (float)tt;
            err /= size;
            s_iter += 1;
            s_active = (err > eps2 && s_iter < kMaxIter) ? 1 : 0;
            ctl->error = err;
        }
    }
    if (t == 0) {
        ctl->iter = s_iter;
        ctl->active = 0;
        *count = s_iter;
    }
}

__global__ void k_fill(float* p, long n, float v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Gaussian taps (mask.c:246-264), on the host with the C library's exp.
Taps make_taps(double sigma) {
    Taps t{};
    const double den = 2 * sigma * sigma;
    t.size = (int)(5 * sigma) + 1;
    if (t.size > kMaxTaps) return t;
    for (int i = 0; i < t.size; i++) t.b[i] = 1 / (sigma * std::sqrt(2.0 * 3.1415926)) * std::exp(-i * i / den);
    double norm = 0;
    for (int i = 0; i < t.size; i++) norm += t.b[i];
    norm *= 2;
    norm -= t.b[0];
    for (int i = 0; i < t.size; i++) t.b[i] /= norm;
    return t;
}

dim3 grid2(int nx, int ny) { return dim3((unsigned)((nx + BX - 1) / BX), (unsigned)((ny + BY - 1) / BY)); }

struct Level {
    int nx, ny;
    float *I0, *I1, *u1, *u2;
};

#define FLOW_TRY(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return ctx_fail(VBM3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int gauss(float* img, float* tmp, int nx, int ny, double sigma, hipStream_t s) {
    const Taps t = make_taps(sigma);
    if (t.size > kMaxTaps) return ctx_fail(VBM3D_ERR_UNSUPPORTED, "Gaussian window too large");
    if (t.size >= nx || t.size >= ny)
        return ctx_fail(VBM3D_ERR_ARG, "image too small for its Gaussian window (" + std::to_string(nx) + "x" +
                                           std::to_string(ny) + ", half-width " + std::to_string(t.size) + ")");
    hipLaunchKernelGGL(k_gauss_rows, grid2(nx, ny), dim3(BX, BY), 0, s, img, tmp, nx, ny, t);
    hipLaunchKernelGGL(k_gauss_cols, grid2(nx, ny), dim3(BX, BY), 0, s, tmp, img, nx, ny, t);
    FLOW_TRY(hipGetLastError());
    return VBM3D_OK;
}

void zoom_size(int nx, int ny, int* nxx, int* nyy, float factor) {
    *nxx = (int)((float)nx * factor + 0.5);
    *nyy = (int)((float)ny * factor + 0.5);
}

}  // namespace

// The whole multiscale flow on device buffers (tvl1flow_lib.c:384-485).
int tvl1_device(vbm3d_ctx* ctx, const float* I0, const float* I1, float* u, float* v, int nx, int ny,
                const vbm3d_tvl1_params* prm, int* iters) {
    const float tau = prm->tau, lambda = prm->lambda, theta = prm->theta, zf = prm->zfactor, eps = prm->epsilon;
    const int nscales = prm->nscales, fscale = prm->fscale, warps = prm->nwarps;
    if (nx < 2 || ny < 2) return ctx_fail(VBM3D_ERR_ARG, "images must be at least 2x2");
    if (nscales < 1 || fscale < 0 || fscale >= nscales || warps < 1 || !(zf > 0.f && zf < 1.f) || !(theta > 0.f) ||
        !(tau > 0.f))
        return ctx_fail(VBM3D_ERR_ARG, "invalid TV-L1 parameters");
    if ((long)nx * ny > (1L << 30)) return ctx_fail(VBM3D_ERR_ARG, "image too large");
    hipStream_t s = ctx_stream(ctx);
    // level sizes and workspace: levels 0..nscales-1 (images; flows of
    // levels >= 1), then the per-level work planes at the finest size
    std::vector<Level> L(nscales);
    L[0].nx = nx;
    L[0].ny = ny;
    for (int k = 1; k < nscales; k++) {
        zoom_size(L[k - 1].nx, L[k - 1].ny, &L[k].nx, &L[k].ny, zf);
        if (L[k].nx < 2 || L[k].ny < 2) return ctx_fail(VBM3D_ERR_ARG, "too many scales for the image size");
    }
    const size_t n0 = (size_t)nx * ny;
    size_t words = 0;
    for (int k = 0; k < nscales; k++) words += (size_t)L[k].nx * L[k].ny * (k == 0 ? 2 : 4);
    const int nparts = (int)(grid2(nx, ny).x * grid2(nx, ny).y);
    const size_t work_planes = 11;  // I1x I1y wx wy grad rho_c p11 p12 p21 p22 tmp
    // planes, then the iteration partials, the min/max partials and the
    // control word, each 256-byte aligned
    const size_t bytes = (words + work_planes * n0) * 4 + (size_t)nparts * 8 + 256 * 4 + 4096 +
                         (size_t)nscales * warps * 4;
    void* base = nullptr;
    if (ctx_flow_workspace(ctx, bytes, &base) != hipSuccess)
        return ctx_fail(VBM3D_ERR_NOMEM, "device allocation of the flow workspace failed");
    float* f = (float*)base;
    for (int k = 0; k < nscales; k++) {
        const size_t sz = (size_t)L[k].nx * L[k].ny;
        L[k].I0 = f;
        f += sz;
        L[k].I1 = f;
        f += sz;
        if (k == 0) {
            L[k].u1 = u;
            L[k].u2 = v;
        } else {
            L[k].u1 = f;
            f += sz;
            L[k].u2 = f;
            f += sz;
        }
    }
    float *I1x = f, *I1y = f + n0, *wx = f + 2 * n0, *wy = f + 3 * n0, *grad = f + 4 * n0, *rho = f + 5 * n0;
    float *p11 = f + 6 * n0, *p12 = f + 7 * n0, *p21 = f + 8 * n0, *p22 = f + 9 * n0, *tmp = f + 10 * n0;
    f += work_planes * n0;
    double* part = (double*)(((uintptr_t)f + 255) & ~(uintptr_t)255);
    float* mm = (float*)(part + nparts);  // 64 workgroups x 4 min/max partials
    FlowCtl* ctl = (FlowCtl*)(((uintptr_t)(mm + 256) + 255) & ~(uintptr_t)255);
    int* counts = (int*)(ctl + 1);  // iterations of the single-workgroup levels, [level][warp]

    // normalisation and 0.8-sigma pre-smoothing (tvl1flow_lib.c:416-421)
    hipLaunchKernelGGL(k_minmax, dim3(64), dim3(256), 0, s, I0, I1, (long)n0, mm);
    hipLaunchKernelGGL(k_normalize, dim3(1024), dim3(256), 0, s, I0, I1, L[0].I0, L[0].I1, (long)n0, mm, 64);
    FLOW_TRY(hipGetLastError());
    int rc;
    if ((rc = gauss(L[0].I0, tmp, nx, ny, 0.8, s)) != VBM3D_OK) return rc;
    if ((rc = gauss(L[0].I1, tmp, nx, ny, 0.8, s)) != VBM3D_OK) return rc;
    // the pyramid (zoom.c:44-83): smooth a copy, resample
    const double zsig = (float)(0.6 * std::sqrt(1.0 / (zf * zf) - 1.0));
    for (int k = 1; k < nscales; k++) {
        const size_t sz = (size_t)L[k - 1].nx * L[k - 1].ny;
        for (int img = 0; img < 2; img++) {
            float* src = img ? L[k - 1].I1 : L[k - 1].I0;
            float* dst = img ? L[k].I1 : L[k].I0;
            FLOW_TRY(hipMemcpyAsync(wx, src, sz * 4, hipMemcpyDeviceToDevice, s));
            if ((rc = gauss(wx, tmp, L[k - 1].nx, L[k - 1].ny, zsig, s)) != VBM3D_OK) return rc;
            hipLaunchKernelGGL(k_resample, grid2(L[k].nx, L[k].ny), dim3(BX, BY), 0, s, wx, dst, L[k - 1].nx,
                               L[k - 1].ny, L[k].nx, L[k].ny, zf, zf, 1.f, 0);
        }
    }
    const long ncoarse = (long)L[nscales - 1].nx * L[nscales - 1].ny;
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, s, L[nscales - 1].u1, ncoarse, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, s, L[nscales - 1].u2, ncoarse, 0.f);
    FLOW_TRY(hipGetLastError());

    const float inv = (float)1.0 / zf, l_t = lambda * theta, taut = tau / theta, eps2 = eps * eps;
    auto upsample = [&](int k) -> int {
        const float fx = (float)L[k - 1].nx / L[k].nx, fy = (float)L[k - 1].ny / L[k].ny;
        hipLaunchKernelGGL(k_resample, grid2(L[k - 1].nx, L[k - 1].ny), dim3(BX, BY), 0, s, L[k].u1, L[k - 1].u1,
                           L[k].nx, L[k].ny, L[k - 1].nx, L[k - 1].ny, fx, fy, inv, 1);
        hipLaunchKernelGGL(k_resample, grid2(L[k - 1].nx, L[k - 1].ny), dim3(BX, BY), 0, s, L[k].u2, L[k - 1].u2,
                           L[k].nx, L[k].ny, L[k - 1].nx, L[k - 1].ny, fx, fy, inv, 1);
        FLOW_TRY(hipGetLastError());
        return VBM3D_OK;
    };
    int level = 0;
    FlowCtl host_ctl{};
    std::vector<char> small_level(nscales, 0);
    for (int k = nscales - 1; k >= fscale; k--, level++) {
        const int lx = L[k].nx, ly = L[k].ny;
        const long sz = (long)lx * ly;
        const dim3 g = grid2(lx, ly), b(BX, BY);
        const int np = (int)(g.x * g.y);
        const bool small = sz <= kSmallPx && np <= kSmallBlocks;
        small_level[level] = small;
        // one level (tvl1flow_lib.c:96-305)
        hipLaunchKernelGGL(k_cgrad, g, b, 0, s, L[k].I1, I1x, I1y, lx, ly);
        for (float* q : {p11, p12, p21, p22}) hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, s, q, sz, 0.f);
        FLOW_TRY(hipGetLastError());
        for (int w = 0; w < warps; w++) {
            hipLaunchKernelGGL(k_warp, g, b, 0, s, L[k].I0, L[k].I1, I1x, I1y, L[k].u1, L[k].u2, wx, wy, grad, rho,
                               lx, ly);
            hipLaunchKernelGGL(k_ctl_start, dim3(1), dim3(1), 0, s, ctl);
            if (small) {
                hipLaunchKernelGGL(k_iter_small, dim3(1), dim3(1024), 0, s, ctl, L[k].u1, L[k].u2, wx, wy, grad, rho,
                                   p11, p12, p21, p22, lx, ly, l_t, theta, taut, (float)sz, eps2,
                                   counts + level * warps + w);
                FLOW_TRY(hipGetLastError());
                continue;
            }
            int launched = 0;
            while (launched < kMaxIter) {
                const int batch = std::min(kMaxIter - launched, launched == 0 ? 24 : 48);
                for (int it = 0; it < batch; it++) {
                    hipLaunchKernelGGL(k_iter_primal, g, b, 0, s, ctl, L[k].u1, L[k].u2, wx, wy, grad, rho, p11, p12,
                                       p21, p22, part, lx, ly, l_t, theta);
                    if (np <= kFuseBlocks) {
                        hipLaunchKernelGGL(k_iter_dual_ctl, g, b, 0, s, ctl, L[k].u1, L[k].u2, p11, p12, p21, p22,
                                           lx, ly, taut, part, np, (float)sz, eps2);
                        continue;
                    }
                    hipLaunchKernelGGL(k_iter_dual, g, b, 0, s, ctl, L[k].u1, L[k].u2, p11, p12, p21, p22, lx, ly,
                                       taut);
                    hipLaunchKernelGGL(k_iter_ctl, dim3(1), dim3(256), 0, s, ctl, part, np, (float)sz, eps2);
                }
                launched += batch;
                FLOW_TRY(hipGetLastError());
                FLOW_TRY(hipMemcpyAsync(&host_ctl, ctl, sizeof(FlowCtl), hipMemcpyDeviceToHost, s));
                FLOW_TRY(hipStreamSynchronize(s));
                if (!host_ctl.active) break;
            }
            if (iters) iters[level * warps + w] = host_ctl.iter;
        }
        if (k == 0) break;
        if ((rc = upsample(k)) != VBM3D_OK) return rc;
    }
    for (int k = fscale - 1; k >= 1; k--)
        if ((rc = upsample(k)) != VBM3D_OK) return rc;
    if (iters) {
        const int nlev = nscales - fscale;  // levels solved
        std::vector<int> dev_counts((size_t)nlev * warps);
        if (!dev_counts.empty())
            FLOW_TRY(hipMemcpyAsync(dev_counts.data(), counts, dev_counts.size() * 4, hipMemcpyDeviceToHost, s));
        FLOW_TRY(hipStreamSynchronize(s));
        for (int lv = 0; lv < nlev; lv++)
            if (small_level[lv])
                for (int w = 0; w < warps; w++) iters[lv * warps + w] = dev_counts[(size_t)lv * warps + w];
    }
    FLOW_TRY(hipStreamSynchronize(s));
    return VBM3D_OK;
}

}  // namespace vb

extern "C" {

void vbm3d_tvl1_default_params(vbm3d_tvl1_params* p) {
    // tvl1flow/main.c:20-30
    p->tau = 0.25f;
    p->lambda = 0.15f;
    p->theta = 0.3f;
    p->nscales = 100;
    p->fscale = 0;
    p->zfactor = 0.5f;
    p->nwarps = 5;
    p->epsilon = 0.01f;
}

int vbm3d_tvl1_nscales(int nx, int ny, float zfactor, int nscales) {
    // tvl1flow/main.c:139-146
    const float N = (float)(1 + std::log(std::hypot((double)nx, (double)ny) / 16.0) / std::log((double)(1 / zfactor)));
    if (N < nscales) nscales = (int)N;
    return nscales;
}

int vbm3d_tvl1_flow_device(vbm3d_ctx* ctx, const float* d_I0, const float* d_I1, float* d_u, float* d_v, int nx,
                           int ny, const vbm3d_tvl1_params* prm, int* iters) {
    if (!ctx || !d_I0 || !d_I1 || !d_u || !d_v || !prm) return vb::ctx_fail(VBM3D_ERR_ARG, "null argument");
    if (hipSetDevice(vb::ctx_device(ctx)) != hipSuccess) return vb::ctx_fail(VBM3D_ERR_HIP, "hipSetDevice failed");
    return vb::tvl1_device(ctx, d_I0, d_I1, d_u, d_v, nx, ny, prm, iters);
}

int vbm3d_tvl1_flow(vbm3d_ctx* ctx, const float* I0, const float* I1, float* u, float* v, int nx, int ny,
                    const vbm3d_tvl1_params* prm, int* iters) {
    if (!ctx || !I0 || !I1 || !u || !v || !prm) return vb::ctx_fail(VBM3D_ERR_ARG, "null argument");
    if (nx < 2 || ny < 2 || (long)nx * ny > (1L << 30)) return vb::ctx_fail(VBM3D_ERR_ARG, "invalid image size");
    if (hipSetDevice(vb::ctx_device(ctx)) != hipSuccess) return vb::ctx_fail(VBM3D_ERR_HIP, "hipSetDevice failed");
    const size_t n = (size_t)nx * ny;
    float* d = nullptr;
    if (hipMalloc(&d, n * 4 * 4) != hipSuccess) return vb::ctx_fail(VBM3D_ERR_NOMEM, "device allocation failed");
    hipStream_t s = vb::ctx_stream(ctx);
    int rc = VBM3D_OK;
    if (hipMemcpyAsync(d, I0, n * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d + n, I1, n * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = vb::ctx_fail(VBM3D_ERR_HIP, "upload failed");
    if (rc == VBM3D_OK) rc = vb::tvl1_device(ctx, d, d + n, d + 2 * n, d + 3 * n, nx, ny, prm, iters);
    if (rc == VBM3D_OK &&
        (hipMemcpyAsync(u, d + 2 * n, n * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipMemcpyAsync(v, d + 3 * n, n * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = vb::ctx_fail(VBM3D_ERR_HIP, "download failed");
    (void)hipStreamSynchronize(s);
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
