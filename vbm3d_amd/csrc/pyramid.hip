// pyramid.hip -- CDNA4 (gfx950) multiscale pyramid of the reference's
// multiscale denoising (multiscaler/multiscaler.cpp, decompose.cpp,
// recompose.cpp, merge_coarse.cpp; src/utils/VBM3Ddenoising_multiscale.sh)
// and its C-ABI entry points.
//
// Images are planar float [c][rows][cols] in HBM.  The DCT pyramid takes
// the full-frame 2D DCT as two matrix products per channel -- T = X A_C^T
// along the rows, Y = A_R T down the columns, A_N the REDFT10 (or
// REDFT01) matrix -- in FP64 through rocBLAS's strided-batched DGEMM on
// the matrix cores, then one rounding to float: the transform's exact
// value to within double rounding (the reference's FFTW float transform
// differs from it in the last bits).  The Gaussian and Lanczos pyramids
// are separable convolutions, one thread per output sample, in the
// reference's float order: bit for bit its results.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "vbm3d_hip.h"

#pragma clang fp contract(off)

struct vbm3d_ctx;

namespace vb {
hipStream_t ctx_stream(vbm3d_ctx* c);
int ctx_device(vbm3d_ctx* c);
hipError_t ctx_pyr_workspace(vbm3d_ctx* c, size_t bytes, void** p);
rocblas_handle ctx_blas(vbm3d_ctx* c);
int ctx_fail(int code, const std::string& msg);
}  // namespace vb

namespace vb {
namespace {

constexpr int BX = 64, BY = 4;
constexpr int kMaxPyrTaps = 32;

struct PTaps {
    float k[kMaxPyrTaps];
    int n;
    float offset;
};

dim3 g2(int cols, int rows, int ch) {
    return dim3((unsigned)((cols + BX - 1) / BX), (unsigned)((rows + BY - 1) / BY), (unsigned)ch);
}

// ---- element kernels
__global__ void k_to_double(const float* __restrict__ in, double* out, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// float(Y) / (4 rows cols) for the forward DCT (multiscaler.cpp:52-56),
// float(Y) for the inverse
__global__ void k_from_double(const double* __restrict__ in, float* out, long n, float div, int apply) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = (float)in[i];
        out[i] = apply ? v / div : v;
    }
}

// copy the top-left h x w block of a rows x cols plane into an h x w
// plane (decompose.cpp:44-52), or back (recompose.cpp:50-56, only j <
// lim_r, k < lim_c)
__global__ void k_block(const float* __restrict__ src, int sr, int sc, float* dst, int dr, int dc, int h, int w) {
    const int q = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y, c = blockIdx.z;
    if (q >= w || r >= h) return;
    dst[((size_t)c * dr + r) * dc + q] = src[((size_t)c * sr + r) * sc + q];
}

__global__ void k_fill_zero(float* p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.f;
}

__global__ void k_axpy(float* a, const float* __restrict__ b, long n, float sign) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = sign > 0 ? a[i] + b[i] : a[i] - b[i];
}

// ---- separable convolution with clamped, rounded tap positions
// (multiscaler.cpp:136-185): pass 0 along the rows index, pass 1 along
// the columns, float sums in tap order
__global__ void k_conv_clamped(const float* __restrict__ in, float* out, int rows, int cols, PTaps t, int pass) {
    const int q = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y, c = blockIdx.z;
    if (q >= cols || r >= rows) return;
    const float* p = in + (size_t)c * rows * cols;
    float val = 0.0f;
    for (int i = 0; i < t.n; i++) {
        if (pass == 0) {
            int j = (int)roundf((float)r - t.offset + (float)i);
            j = min(max(j, 0), rows - 1);
            val += p[(size_t)j * cols + q] * t.k[i];
        } else {
            int j = (int)roundf((float)q - t.offset + (float)i);
            j = min(max(j, 0), cols - 1);
            val += p[(size_t)r * cols + j] * t.k[i];
        }
    }
    out[((size_t)c * rows + r) * cols + q] = val;
}

// ---- one-direction convolution, symmetric boundary, truncated tap
// positions (multiscaler.cpp:190-247); dir 0 along the rows index
__global__ void k_conv_dir(const float* __restrict__ in, float* out, int rows, int cols, PTaps t, int dir) {
    const int q = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y, c = blockIdx.z;
    if (q >= cols || r >= rows) return;
    const float* p = in + (size_t)c * rows * cols;
    const int X = dir == 0 ? rows : cols, X2 = 2 * X, x = dir == 0 ? r : q;
    float val = 0.0f;
    for (int i = 0; i < t.n; i++) {
        int j = (int)((float)x - t.offset + (float)i);
        while (j < 0) j += X2;
        while (j >= X2) j -= X2;
        if (j >= X) j = X2 - 1 - j;
        val += p[dir == 0 ? (size_t)j * cols + q : (size_t)r * cols + j] * t.k[i];
    }
    out[((size_t)c * rows + r) * cols + q] = val;
}

// out[2r][2q] = in[r][q] of a (rows x cols) -> (rows/2 x cols/2) subsample
__global__ void k_subsample(const float* __restrict__ in, float* out, int rows, int cols) {
    const int r2 = rows / 2, c2 = cols / 2;
    const int q = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y, c = blockIdx.z;
    if (q >= c2 || r >= r2) return;
    out[((size_t)c * r2 + r) * c2 + q] = in[((size_t)c * rows + 2 * r) * cols + 2 * q];
}

// one phase of the 2x upsampling (multiscaler.cpp:337-398): out[2j + rp]
// [2k + cp] = t[j][k]; for an uneven target the last row / column copies
// the phase's last row / column where the reference does
__global__ void k_up_phase(const float* __restrict__ t, float* out, int rows, int cols, int sx, int sy, int rp,
                           int cp) {
    const int q = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y, c = blockIdx.z;
    if (q >= cols || r >= rows) return;
    const float v = t[((size_t)c * rows + r) * cols + q];
    float* o = out + (size_t)c * sx * sy;
    o[(size_t)(2 * r + rp) * sy + 2 * q + cp] = v;
    // (rp, cp) = (0, 1): last column, rows 2j; (1, 0): last row, columns
    // 2k; (1, 1): last row, columns 2k+1, last column, rows 2j+1, corner
    const bool oddx = sx % 2 == 1, oddy = sy % 2 == 1;
    if (rp == 0 && cp == 1 && oddy && q == cols - 1) o[(size_t)(2 * r) * sy + sy - 1] = v;
    if (rp == 1 && cp == 0 && oddx && r == rows - 1) o[(size_t)(sx - 1) * sy + 2 * q] = v;
    if (rp == 1 && cp == 1) {
        if (oddx && r == rows - 1) o[(size_t)(sx - 1) * sy + 2 * q + 1] = v;
        if (oddy && q == cols - 1) o[(size_t)(2 * r + 1) * sy + sy - 1] = v;
        if (oddx && oddy && r == rows - 1 && q == cols - 1) o[(size_t)(sx - 1) * sy + sy - 1] = v;
    }
}

// ---- host-side taps (multiscaler.cpp:261-276, 392-418, 279-304)
PTaps gaussian_taps(int n, float sigma, float mean) {
    PTaps t{};
    t.n = n;
    t.offset = mean;
    double sum = 0.;
    for (int i = 0; i < n; i++) {
        const int val = (int)((i - mean) / sigma);
        t.k[i] = (float)std::exp(-0.5 * val * val);
        sum += t.k[i];
    }
    if (sum > 0.)
        for (int i = 0; i < n; i++) t.k[i] = (float)(t.k[i] / sum);
    return t;
}

float lanczos3(float x) {
    if (std::abs(x) >= 3) return 0.f;
    if (x == 0.) return 1.f;
    return (float)(std::sin(M_PI * x) / (M_PI * x) * std::sin(M_PI * x / 3.) / (M_PI * x / 3.));
}

PTaps lanczos_taps(int n, float offset) {
    PTaps t{};
    t.n = n;
    t.offset = offset;
    double sum = 0.;
    for (int i = 0; i < n; i++) {
        t.k[i] = (float)(lanczos3((float)((i - offset) / 2.)) / 2.);
        sum += t.k[i];
    }
    if (sum > 0.)
        for (int i = 0; i < n; i++) t.k[i] = (float)(t.k[i] / sum);
    return t;
}

PTaps down_taps(bool lanczos) {
    if (lanczos) return lanczos_taps(12, 5.5f);
    const float sigma = 1.39f, prec = 3.0f;
    const float offset = (float)(std::ceil(sigma * std::sqrt(2.0 * prec * std::log(10.0))) + 0.5);
    return gaussian_taps((int)(1 + 2 * offset), sigma, offset);
}

#define PYR_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return ctx_fail(VBM3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BLAS_TRY(expr)                                                                       \
    do {                                                                                     \
        rocblas_status s_ = (expr);                                                          \
        if (s_ != rocblas_status_success)                                                    \
            return ctx_fail(VBM3D_ERR_HIP, std::string(#expr) + ": " + rocblas_status_to_string(s_)); \
    } while (0)

// The device work of the pyramid entries: a context's stream and BLAS
// handle, a scratch arena, and the DCT matrices of the sizes in use.
struct Pyr {
    vbm3d_ctx* ctx;
    hipStream_t s;
    char* base = nullptr;
    size_t cap = 0, used = 0;

    template <class T>
    T* take(size_t n) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at + n * sizeof(T) > cap) return nullptr;
        used = at + n * sizeof(T);
        return (T*)(base + at);
    }
};

// REDFT10 (inverse = false: 2 cos(pi k (2n+1) / 2N)) or REDFT01 (1 for
// n = 0, else 2 cos(pi n (2k+1) / 2N)) as a row-major N x N table [k][n]
std::vector<double> dct_table(int N, bool inverse) {
    std::vector<double> a((size_t)N * N);
    for (int k = 0; k < N; k++)
        for (int n = 0; n < N; n++)
            a[(size_t)k * N + n] = inverse ? (n == 0 ? 1.0 : 2.0 * std::cos(M_PI * n * (2.0 * k + 1.0) / (2.0 * N)))
                                           : 2.0 * std::cos(M_PI * k * (2.0 * n + 1.0) / (2.0 * N));
    return a;
}

// In-place 2D DCT (REDFT10, / 4 rows cols) or inverse (REDFT01) of `ch`
// planes (multiscaler.cpp:21-98) with two strided-batched DGEMMs.
// Column-major view: a row-major rows x cols plane is the cols x rows
// matrix P^T; T^T = A_C P^T, then Y^T = T^T A_R^T.
int dct2(Pyr& P, float* img, int rows, int cols, int ch, bool inverse) {
    const size_t n = (size_t)ch * rows * cols;
    const size_t mark = P.used;
    double* x = P.take<double>(n);
    double* t = P.take<double>(n);
    double* ar = P.take<double>((size_t)rows * rows);
    double* ac = P.take<double>((size_t)cols * cols);
    if (!x || !t || !ar || !ac) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
    const std::vector<double> hr = dct_table(rows, inverse), hc = dct_table(cols, inverse);
    PYR_TRY(hipMemcpyAsync(ar, hr.data(), hr.size() * 8, hipMemcpyHostToDevice, P.s));
    PYR_TRY(hipMemcpyAsync(ac, hc.data(), hc.size() * 8, hipMemcpyHostToDevice, P.s));
    hipLaunchKernelGGL(k_to_double, dim3(1024), dim3(256), 0, P.s, img, x, (long)n);
    PYR_TRY(hipGetLastError());
    rocblas_handle h = ctx_blas(P.ctx);
    const double one = 1.0, zero = 0.0;
    const rocblas_stride st = (rocblas_stride)rows * cols;
    // T^T (cols x rows) = A_C (row-major table = col-major A_C^T, transposed) * P^T
    BLAS_TRY(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, cols, rows, cols,
                                           &one, ac, cols, 0, x, cols, st, &zero, t, cols, st, ch));
    // Y^T (cols x rows) = T^T * A_R^T (the row-major table read column-major)
    BLAS_TRY(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, cols, rows, rows, &one,
                                           t, cols, st, ar, rows, 0, &zero, x, cols, st, ch));
    hipLaunchKernelGGL(k_from_double, dim3(1024), dim3(256), 0, P.s, x, img, (long)n, (float)(4 * rows * cols),
                       inverse ? 0 : 1);
    PYR_TRY(hipGetLastError());
    PYR_TRY(hipStreamSynchronize(P.s));  // the host tables above are freed on return
    P.used = mark;
    return VBM3D_OK;
}

int conv_clamped(Pyr& P, float* img, float* tmp, int rows, int cols, int ch, const PTaps& t) {
    hipLaunchKernelGGL(k_conv_clamped, g2(cols, rows, ch), dim3(BX, BY), 0, P.s, img, tmp, rows, cols, t, 0);
    hipLaunchKernelGGL(k_conv_clamped, g2(cols, rows, ch), dim3(BX, BY), 0, P.s, tmp, img, rows, cols, t, 1);
    PYR_TRY(hipGetLastError());
    return VBM3D_OK;
}

// gdown / ldown (multiscaler.cpp:281-331, 421-443): blur a copy, keep
// every other sample; out is (rows/2) x (cols/2)
int down(Pyr& P, const float* img, float* out, int rows, int cols, int ch, bool lanczos) {
    const size_t n = (size_t)ch * rows * cols, mark = P.used;
    float* b = P.take<float>(n);
    float* tmp = P.take<float>(n);
    if (!b || !tmp) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
    PYR_TRY(hipMemcpyAsync(b, img, n * 4, hipMemcpyDeviceToDevice, P.s));
    int rc = conv_clamped(P, b, tmp, rows, cols, ch, down_taps(lanczos));
    if (rc) return rc;
    hipLaunchKernelGGL(k_subsample, g2(cols / 2, rows / 2, ch), dim3(BX, BY), 0, P.s, b, out, rows, cols);
    PYR_TRY(hipGetLastError());
    P.used = mark;
    return VBM3D_OK;
}

// gup / lup (multiscaler.cpp:337-398, 448-509): rows x cols -> sx x sy
int up(Pyr& P, const float* img, float* out, int rows, int cols, int ch, int sx, int sy, bool lanczos) {
    const size_t n = (size_t)ch * rows * cols, mark = P.used;
    float* t1 = P.take<float>(n);
    float* t2 = P.take<float>(n);
    if (!t1 || !t2) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
    const PTaps k1 = lanczos ? lanczos_taps(6, 2.25f) : gaussian_taps(5, 0.5f, 2.f);
    const PTaps k2 = lanczos ? lanczos_taps(6, 2.75f) : gaussian_taps(5, 0.5f, 1.5f);
    const dim3 g = g2(cols, rows, ch), b(BX, BY);
    hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, P.s, out, (long)ch * sx * sy);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, img, t1, rows, cols, k1, 0);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, t1, t2, rows, cols, k1, 1);
    hipLaunchKernelGGL(k_up_phase, g, b, 0, P.s, t2, out, rows, cols, sx, sy, 0, 0);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, t1, t2, rows, cols, k2, 1);
    hipLaunchKernelGGL(k_up_phase, g, b, 0, P.s, t2, out, rows, cols, sx, sy, 0, 1);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, img, t1, rows, cols, k2, 0);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, t1, t2, rows, cols, k1, 1);
    hipLaunchKernelGGL(k_up_phase, g, b, 0, P.s, t2, out, rows, cols, sx, sy, 1, 0);
    hipLaunchKernelGGL(k_conv_dir, g, b, 0, P.s, t1, t2, rows, cols, k2, 1);
    hipLaunchKernelGGL(k_up_phase, g, b, 0, P.s, t2, out, rows, cols, sx, sy, 1, 1);
    PYR_TRY(hipGetLastError());
    P.used = mark;
    return VBM3D_OK;
}

// Level sizes: each level halves the previous one (integer division).
void level_sizes(int rows, int cols, int levels, std::vector<int>& lr, std::vector<int>& lc) {
    lr.resize(levels);
    lc.resize(levels);
    for (int i = 0; i < levels; i++) {
        lr[i] = rows;
        lc[i] = cols;
        rows /= 2;
        cols /= 2;
    }
}

size_t level_words(int rows, int cols, int ch, int levels) {
    std::vector<int> lr, lc;
    level_sizes(rows, cols, levels, lr, lc);
    size_t n = 0;
    for (int i = 0; i < levels; i++) n += (size_t)ch * lr[i] * lc[i];
    return n;
}

// Workspace bytes of one image's pyramid operations: the DCT's two
// double planes and matrices, or the convolutions' float planes, and a
// full-size float plane for recomposition.
size_t pyr_bytes(int rows, int cols, int ch) {
    const size_t n = (size_t)ch * rows * cols;
    return n * 8 * 2 + ((size_t)rows * rows + (size_t)cols * cols) * 8 + n * 4 * 6 + 16 * 256;
}

int check_image(int rows, int cols, int ch, int levels, int type) {
    if (rows < 1 || cols < 1 || (ch != 1 && ch != 3) || (long)rows * cols > (1L << 28))
        return ctx_fail(VBM3D_ERR_ARG, "image: 1 or 3 channels, at most 2^28 pixels");
    if (levels < 1 || (type != 0 && type != 1 && type != 2))
        return ctx_fail(VBM3D_ERR_ARG, "levels >= 1 and type 0 (DCT), 1 (Gaussian) or 2 (Lanczos)");
    std::vector<int> lr, lc;
    level_sizes(rows, cols, levels, lr, lc);
    if (lr[levels - 1] < 1 || lc[levels - 1] < 1) return ctx_fail(VBM3D_ERR_ARG, "too many levels for the image size");
    return VBM3D_OK;
}

// decompose.cpp:17-106 on device planes: `out` holds the levels one after
// the other.
int decompose_dev(Pyr& P, const float* img, int rows, int cols, int ch, int levels, int type, float* out) {
    std::vector<int> lr, lc;
    level_sizes(rows, cols, levels, lr, lc);
    const size_t n = (size_t)ch * rows * cols, mark = P.used;
    int rc;
    if (type == 0) {
        float* d = P.take<float>(n);
        if (!d) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
        PYR_TRY(hipMemcpyAsync(d, img, n * 4, hipMemcpyDeviceToDevice, P.s));
        if ((rc = dct2(P, d, rows, cols, ch, false))) return rc;
        float* o = out;
        for (int i = 0; i < levels; i++) {
            hipLaunchKernelGGL(k_block, g2(lc[i], lr[i], ch), dim3(BX, BY), 0, P.s, d, rows, cols, o, lr[i], lc[i],
                               lr[i], lc[i]);
            PYR_TRY(hipGetLastError());
            if ((rc = dct2(P, o, lr[i], lc[i], ch, true))) return rc;
            o += (size_t)ch * lr[i] * lc[i];
        }
        P.used = mark;
        return VBM3D_OK;
    }
    PYR_TRY(hipMemcpyAsync(out, img, n * 4, hipMemcpyDeviceToDevice, P.s));
    float* prev = out;
    for (int i = 1; i < levels; i++) {
        float* cur = prev + (size_t)ch * lr[i - 1] * lc[i - 1];
        if ((rc = down(P, prev, cur, lr[i - 1], lc[i - 1], ch, type == 2))) return rc;
        prev = cur;
    }
    return VBM3D_OK;
}

// recompose.cpp:17-125 on device planes.
int recompose_dev(Pyr& P, const float* lev, int rows, int cols, int ch, int levels, int type, float factor, float* out) {
    std::vector<int> lr, lc;
    level_sizes(rows, cols, levels, lr, lc);
    std::vector<size_t> off(levels);
    size_t o = 0;
    for (int i = 0; i < levels; i++) {
        off[i] = o;
        o += (size_t)ch * lr[i] * lc[i];
    }
    const size_t mark = P.used;
    int rc;
    if (type == 0) {
        PYR_TRY(hipMemcpyAsync(out, lev, (size_t)ch * rows * cols * 4, hipMemcpyDeviceToDevice, P.s));
        if ((rc = dct2(P, out, rows, cols, ch, false))) return rc;
        float* im = P.take<float>((size_t)ch * rows * cols);
        if (!im) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
        for (int i = 1; i < levels; i++) {
            const int h = lr[i], w = lc[i];
            PYR_TRY(hipMemcpyAsync(im, lev + off[i], (size_t)ch * h * w * 4, hipMemcpyDeviceToDevice, P.s));
            if ((rc = dct2(P, im, h, w, ch, false))) return rc;
            // the copied block: j < h * factor, k < w * factor (float bounds)
            int bh = 0, bw = 0;
            while (bh < h && (float)bh < h * factor) bh++;
            while (bw < w && (float)bw < w * factor) bw++;
            if (bh > 0 && bw > 0) {
                hipLaunchKernelGGL(k_block, g2(bw, bh, ch), dim3(BX, BY), 0, P.s, im, h, w, out, rows, cols, bh, bw);
                PYR_TRY(hipGetLastError());
            }
        }
        if ((rc = dct2(P, out, rows, cols, ch, true))) return rc;
        P.used = mark;
        return VBM3D_OK;
    }
    const bool lz = type == 2;
    float* cur = P.take<float>((size_t)ch * rows * cols);
    float* nxt = P.take<float>((size_t)ch * rows * cols);
    float* small = P.take<float>((size_t)ch * rows * cols);
    if (!cur || !nxt || !small) return ctx_fail(VBM3D_ERR_NOMEM, "pyramid workspace too small");
    PYR_TRY(hipMemcpyAsync(cur, lev + off[levels - 1], (size_t)ch * lr[levels - 1] * lc[levels - 1] * 4,
                           hipMemcpyDeviceToDevice, P.s));
    for (int i = levels - 2; i >= 0; --i) {
        const int h = lr[i], w = lc[i], h2 = lr[i + 1], w2 = lc[i + 1];
        const float* image = lev + off[i];
        if ((rc = down(P, image, small, h, w, ch, lz))) return rc;
        const long n2 = (long)ch * h2 * w2;
        hipLaunchKernelGGL(k_axpy, dim3(256), dim3(256), 0, P.s, cur, small, n2, -1.f);
        PYR_TRY(hipGetLastError());
        // gblur(current, factor) (:279-304): skipped below 0.41
        if (!(factor < 0.41)) {
            const float prec = 3.0f;
            const float offset = (float)((int)std::ceil(factor * std::sqrt(2.0 * prec * std::log(10.0)))) - 0.5f;
            const int nt = (int)(2 * offset);
            if (nt > kMaxPyrTaps) return ctx_fail(VBM3D_ERR_UNSUPPORTED, "recomposition blur too wide");
            if ((rc = conv_clamped(P, cur, small, h2, w2, ch, gaussian_taps(nt, factor, offset)))) return rc;
        }
        if ((rc = up(P, cur, nxt, h2, w2, ch, h, w, lz))) return rc;
        hipLaunchKernelGGL(k_axpy, dim3(256), dim3(256), 0, P.s, nxt, image, (long)ch * h * w, 1.f);
        PYR_TRY(hipGetLastError());
        std::swap(cur, nxt);
    }
    PYR_TRY(hipMemcpyAsync(out, cur, (size_t)ch * rows * cols * 4, hipMemcpyDeviceToDevice, P.s));
    P.used = mark;
    return VBM3D_OK;
}

int pyr_open(vbm3d_ctx* ctx, Pyr& P, size_t bytes) {
    if (!ctx) return ctx_fail(VBM3D_ERR_ARG, "null context");
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return ctx_fail(VBM3D_ERR_HIP, "hipSetDevice failed");
    P.ctx = ctx;
    P.s = ctx_stream(ctx);
    void* b = nullptr;
    if (ctx_pyr_workspace(ctx, bytes, &b) != hipSuccess)
        return ctx_fail(VBM3D_ERR_NOMEM, "device allocation of the pyramid workspace failed");
    P.base = (char*)b;
    P.cap = bytes;
    P.used = 0;
    if (!ctx_blas(ctx)) return ctx_fail(VBM3D_ERR_HIP, "rocblas_create_handle failed");
    return VBM3D_OK;
}

// Device copies of host inputs, released on every return.
struct Dev {
    std::vector<void*> p;
    ~Dev() {
        for (void* q : p) (void)hipFree(q);
    }
    float* get(size_t n) {
        void* q = nullptr;
        if (hipMalloc(&q, n * 4 > 0 ? n * 4 : 4) != hipSuccess) return nullptr;
        p.push_back(q);
        return (float*)q;
    }
};

}  // namespace
}  // namespace vb

using namespace vb;

extern "C" {

int vbm3d_pyr_levels(int rows, int cols, int levels, int* lrows, int* lcols) {
    if (rows < 1 || cols < 1 || levels < 1 || !lrows || !lcols) return ctx_fail(VBM3D_ERR_ARG, "invalid arguments");
    std::vector<int> lr, lc;
    level_sizes(rows, cols, levels, lr, lc);
    for (int i = 0; i < levels; i++) {
        lrows[i] = lr[i];
        lcols[i] = lc[i];
    }
    return VBM3D_OK;
}

int vbm3d_pyr_decompose(vbm3d_ctx* ctx, const float* img, int rows, int cols, int ch, int levels, int type, float* out) {
    int rc = check_image(rows, cols, ch, levels, type);
    if (rc) return rc;
    if (!img || !out) return ctx_fail(VBM3D_ERR_ARG, "null buffer");
    Pyr P;
    if ((rc = pyr_open(ctx, P, pyr_bytes(rows, cols, ch)))) return rc;
    const size_t n = (size_t)ch * rows * cols, nl = level_words(rows, cols, ch, levels);
    Dev D;
    float* di = D.get(n);
    float* dout = D.get(nl);
    if (!di || !dout) return ctx_fail(VBM3D_ERR_NOMEM, "device allocation failed");
    PYR_TRY(hipMemcpyAsync(di, img, n * 4, hipMemcpyHostToDevice, P.s));
    if ((rc = decompose_dev(P, di, rows, cols, ch, levels, type, dout))) return rc;
    PYR_TRY(hipMemcpyAsync(out, dout, nl * 4, hipMemcpyDeviceToHost, P.s));
    PYR_TRY(hipStreamSynchronize(P.s));
    return VBM3D_OK;
}

int vbm3d_pyr_recompose(vbm3d_ctx* ctx, const float* levels_in, int rows, int cols, int ch, int levels, int type,
                        float factor, float* out) {
    int rc = check_image(rows, cols, ch, levels, type);
    if (rc) return rc;
    if (!levels_in || !out) return ctx_fail(VBM3D_ERR_ARG, "null buffer");
    Pyr P;
    if ((rc = pyr_open(ctx, P, pyr_bytes(rows, cols, ch)))) return rc;
    const size_t n = (size_t)ch * rows * cols, nl = level_words(rows, cols, ch, levels);
    Dev D;
    float* dl = D.get(nl);
    float* dout = D.get(n);
    if (!dl || !dout) return ctx_fail(VBM3D_ERR_NOMEM, "device allocation failed");
    PYR_TRY(hipMemcpyAsync(dl, levels_in, nl * 4, hipMemcpyHostToDevice, P.s));
    if ((rc = recompose_dev(P, dl, rows, cols, ch, levels, type, factor, dout))) return rc;
    PYR_TRY(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, P.s));
    PYR_TRY(hipStreamSynchronize(P.s));
    return VBM3D_OK;
}

int vbm3d_pyr_merge_coarse(vbm3d_ctx* ctx, const float* fine, int rows, int cols, const float* coarse, int crows,
                           int ccols, int ch, float factor, float* out) {
    int rc = check_image(rows, cols, ch, 1, 0);
    if (rc) return rc;
    if ((rc = check_image(crows, ccols, ch, 1, 0))) return rc;
    if (!fine || !coarse || !out) return ctx_fail(VBM3D_ERR_ARG, "null buffer");
    // the copied block must lie inside both images (merge_coarse.cpp:39-45)
    int bh = 0, bw = 0;
    while ((float)bh < crows * factor) bh++;
    while ((float)bw < ccols * factor) bw++;
    if (bh > crows || bw > ccols || bh > rows || bw > cols)
        return ctx_fail(VBM3D_ERR_ARG, "the coarse image's block does not fit (factor too large)");
    Pyr P;
    if ((rc = pyr_open(ctx, P, pyr_bytes(rows, cols, ch) + pyr_bytes(crows, ccols, ch)))) return rc;
    const size_t n = (size_t)ch * rows * cols, nc = (size_t)ch * crows * ccols;
    Dev D;
    float* df = D.get(n);
    float* dc = D.get(nc);
    if (!df || !dc) return ctx_fail(VBM3D_ERR_NOMEM, "device allocation failed");
    PYR_TRY(hipMemcpyAsync(df, fine, n * 4, hipMemcpyHostToDevice, P.s));
    PYR_TRY(hipMemcpyAsync(dc, coarse, nc * 4, hipMemcpyHostToDevice, P.s));
    if ((rc = dct2(P, df, rows, cols, ch, false)) || (rc = dct2(P, dc, crows, ccols, ch, false))) return rc;
    if (bh > 0 && bw > 0) {
        hipLaunchKernelGGL(k_block, g2(bw, bh, ch), dim3(BX, BY), 0, P.s, dc, crows, ccols, df, rows, cols, bh, bw);
        PYR_TRY(hipGetLastError());
    }
    if ((rc = dct2(P, df, rows, cols, ch, true))) return rc;
    PYR_TRY(hipMemcpyAsync(out, df, n * 4, hipMemcpyDeviceToHost, P.s));
    PYR_TRY(hipStreamSynchronize(P.s));
    return VBM3D_OK;
}

int vbm3d_multiscale_denoise(vbm3d_ctx* ctx, float sigma, const float* noisy, unsigned W, unsigned H, unsigned F,
                             unsigned C, const vbm3d_params* p1, const vbm3d_params* p2, int levels, int type,
                             float factor, int color_space, int ref_grid_nb, float* denoised, float* noisy_levels,
                             float* denoised_levels) {
    int rc = check_image((int)H, (int)W, (int)C, levels, type);
    if (rc) return rc;
    if (!noisy || !denoised || F == 0) return ctx_fail(VBM3D_ERR_ARG, "null buffer or empty video");
    if (!(sigma > 0.f)) return ctx_fail(VBM3D_ERR_ARG, "sigma must be positive");
    Pyr P;
    if ((rc = pyr_open(ctx, P, pyr_bytes((int)H, (int)W, (int)C)))) return rc;
    std::vector<int> lr, lc;
    level_sizes((int)H, (int)W, levels, lr, lc);
    const size_t fw = (size_t)C * H * W, flev = level_words((int)H, (int)W, (int)C, levels);
    Dev D;
    float* dn = D.get(fw * F);
    float* dlev = D.get(flev * F);   // per level: [F][C][h][w], level after level
    float* dden = D.get(flev * F);
    float* dbas = D.get(fw * F);
    float* dfr = D.get(flev);
    if (!dn || !dlev || !dden || !dbas || !dfr) return ctx_fail(VBM3D_ERR_NOMEM, "device allocation failed");
    std::vector<size_t> loff(levels);
    size_t o = 0;
    for (int i = 0; i < levels; i++) {
        loff[i] = o;
        o += (size_t)C * lr[i] * lc[i] * F;
    }
    PYR_TRY(hipMemcpyAsync(dn, noisy, fw * F * 4, hipMemcpyHostToDevice, P.s));
    // 1. every frame into its levels (decompose per frame)
    for (unsigned t = 0; t < F; t++) {
        if ((rc = decompose_dev(P, dn + fw * t, (int)H, (int)W, (int)C, levels, type, dfr))) return rc;
        size_t fo = 0;
        for (int i = 0; i < levels; i++) {
            const size_t ls = (size_t)C * lr[i] * lc[i];
            PYR_TRY(hipMemcpyAsync(dlev + loff[i] + ls * t, dfr + fo, ls * 4, hipMemcpyDeviceToDevice, P.s));
            fo += ls;
        }
    }
    if (noisy_levels) {
        PYR_TRY(hipMemcpyAsync(noisy_levels, dlev, flev * F * 4, hipMemcpyDeviceToHost, P.s));
    }
    PYR_TRY(hipStreamSynchronize(P.s));
    // 2. each level denoised as a video with sigma / 2^level, truncated to
    //    two decimals as the multiscale shell script's `bc` computes it
    //    (src/utils/VBM3Ddenoising_multiscale.sh:52-58)
    for (int i = 0; i < levels; i++) {
        const double sl_exact = (double)sigma / std::ldexp(1.0, i);
        const float sl = (float)(std::floor(sl_exact * 100.0 + 1e-9) / 100.0);
        vbm3d_params q1, q2;
        if (p1) q1 = *p1;
        else vbm3d_default_params_1(&q1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1.f, -1.f, (unsigned)-1, (unsigned)-1, sl);
        if (p2) q2 = *p2;
        else vbm3d_default_params_2(&q2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1.f, (unsigned)-1, (unsigned)-1, sl);
        if ((rc = vbm3d_run_device(ctx, sl, dlev + loff[i], nullptr, nullptr, (unsigned)lc[i], (unsigned)lr[i], F, C, &q1,
                                   &q2, color_space, ref_grid_nb, dbas, dden + loff[i], nullptr)))
            return rc;
    }
    if (denoised_levels) {
        PYR_TRY(hipMemcpyAsync(denoised_levels, dden, flev * F * 4, hipMemcpyDeviceToHost, P.s));
    }
    // 3. every frame recomposed from its denoised levels
    for (unsigned t = 0; t < F; t++) {
        size_t fo = 0;
        for (int i = 0; i < levels; i++) {
            const size_t ls = (size_t)C * lr[i] * lc[i];
            PYR_TRY(hipMemcpyAsync(dfr + fo, dden + loff[i] + ls * t, ls * 4, hipMemcpyDeviceToDevice, P.s));
            fo += ls;
        }
        if ((rc = recompose_dev(P, dfr, (int)H, (int)W, (int)C, levels, type, factor, dn + fw * t))) return rc;
    }
    PYR_TRY(hipMemcpyAsync(denoised, dn, fw * F * 4, hipMemcpyDeviceToHost, P.s));
    PYR_TRY(hipStreamSynchronize(P.s));
    return VBM3D_OK;
}

}  // extern "C"
