// tvl1_oracle.cpp -- CPU restatement of the TV-L1 optical flow that the
// reference ships beside the denoiser (tvl1flow/tvl1flow_lib.c, after Zach,
// Pock and Bischof 2007 with Chambolle's dual projection).  TEST
// INFRASTRUCTURE ONLY: the GPU flow kernels are checked against it; the
// product never links it.
//
// Every function states the reference lines it follows.  Arithmetic is
// written out in the reference's evaluation order and precision (float
// where the reference stores float, double where its expressions promote),
// compiled without contraction, so the results are the reference's own
// sequential results bit for bit; tests pin that against the reference
// library built from its sources (oracle/Makefile, target ref).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using Img = std::vector<float>;

// Neumann clamp of an interpolation index; `out` records a clamp
// (tvl1flow/bicubic_interpolation.c:31-46).
inline int clamp_idx(int x, int n, bool& out) {
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

// Catmull-Rom cell in double, the reference's Horner order
// (bicubic_interpolation.c:120-131).
inline double cubic_cell(const double v[4], double x) {
    const double c3 = 3.0 * (v[1] - v[2]) + v[3] - v[0];
    const double c2 = 2.0 * v[0] - 5.0 * v[1] + 4.0 * v[2] - v[3] + x * c3;
    return v[1] + 0.5 * x * (v[2] - v[0] + x * c2);
}

// Bicubic sample of `in` at (uu, vv) (bicubic_interpolation.c:156-232).
// The row below the sample is clamped with the x direction's sign, as the
// reference does (line 174).  Outside samples give 0 when border_out.
float bicubic_at(const float* in, float uu, float vv, int nx, int ny, bool border_out) {
    const int sx = uu < 0 ? -1 : 1, sy = vv < 0 ? -1 : 1;
    bool out = false;
    const int iu = (int)uu, iv = (int)vv;
    const int x = clamp_idx(iu, nx, out), y = clamp_idx(iv, ny, out);
    const int mx = clamp_idx(iu - sx, nx, out), my = clamp_idx(iv - sx, ny, out);
    const int dx = clamp_idx(iu + sx, nx, out), dy = clamp_idx(iv + sy, ny, out);
    const int ddx = clamp_idx(iu + 2 * sx, nx, out), ddy = clamp_idx(iv + 2 * sy, ny, out);
    if (out && border_out) return 0.f;
    const int cx[4] = {mx, x, dx, ddx}, cy[4] = {my, y, dy, ddy};
    // column c of the 4x4 neighbourhood, interpolated along y first
    double col[4];
    const double fy = (double)(vv - (float)y), fx = (double)(uu - (float)x);
    for (int c = 0; c < 4; c++) {
        double v[4];
        for (int r = 0; r < 4; r++) v[r] = in[cx[c] + nx * cy[r]];
        col[c] = cubic_cell(v, fy);
    }
    return (float)cubic_cell(col, fx);
}

// Gaussian smoothing in place, separable, reflecting boundary, double
// accumulation (tvl1flow/mask.c:234-337).
void gaussian(float* I, int xdim, int ydim, double sigma) {
    const double den = 2 * sigma * sigma;
    const int size = (int)(5 * sigma) + 1;
    std::vector<double> B(size);
    for (int i = 0; i < size; i++) B[i] = 1 / (sigma * std::sqrt(2.0 * 3.1415926)) * std::exp(-i * i / den);
    double norm = 0;
    for (int i = 0; i < size; i++) norm += B[i];
    norm *= 2;
    norm -= B[0];
    for (int i = 0; i < size; i++) B[i] /= norm;
    // one line of n samples at stride st: left mirror excludes the edge
    // sample, right mirror includes it
    auto line = [&](float* base, int n, int st) {
        std::vector<double> R(n + 2 * size);
        for (int i = 0; i < n; i++) R[size + i] = base[(size_t)i * st];
        for (int i = 0; i < size; i++) {
            R[i] = base[(size_t)(size - i) * st];
            R[size + n + i] = base[(size_t)(n - i - 1) * st];
        }
        for (int i = size; i < size + n; i++) {
            double sum = B[0] * R[i];
            for (int j = 1; j < size; j++) sum += B[j] * (R[i - j] + R[i + j]);
            base[(size_t)(i - size) * st] = (float)sum;
        }
    };
    for (int k = 0; k < ydim; k++) line(I + (size_t)k * xdim, xdim, 1);
    for (int k = 0; k < xdim; k++) line(I + k, ydim, xdim);
}

// Pyramid level size (tvl1flow/zoom.c:24-36).
void zoom_size(int nx, int ny, int* nxx, int* nyy, float factor) {
    *nxx = (int)((float)nx * factor + 0.5);
    *nyy = (int)((float)ny * factor + 0.5);
}

// Downsampling: Gaussian pre-smoothing, then bicubic resampling
// (zoom.c:44-83).
void zoom_out(const float* I, float* Iout, int nx, int ny, float factor) {
    Img Is(I, I + (size_t)nx * ny);
    int nxx, nyy;
    zoom_size(nx, ny, &nxx, &nyy, factor);
    const float sigma = (float)(0.6 * std::sqrt(1.0 / (factor * factor) - 1.0));
    gaussian(Is.data(), nx, ny, sigma);
    for (int i1 = 0; i1 < nyy; i1++)
        for (int j1 = 0; j1 < nxx; j1++) {
            const float i2 = (float)i1 / factor, j2 = (float)j1 / factor;
            Iout[i1 * nxx + j1] = bicubic_at(Is.data(), j2, i2, nx, ny, false);
        }
}

// Upsampling by bicubic resampling (zoom.c:91-117).
void zoom_in(const float* I, float* Iout, int nx, int ny, int nxx, int nyy) {
    const float fx = (float)nxx / nx, fy = (float)nyy / ny;
    for (int i1 = 0; i1 < nyy; i1++)
        for (int j1 = 0; j1 < nxx; j1++) {
            const float i2 = (float)i1 / fy, j2 = (float)j1 / fx;
            Iout[i1 * nxx + j1] = bicubic_at(I, j2, i2, nx, ny, false);
        }
}

// Centred differences with one-sided edges, times 0.5 (mask.c:144-214).
void centered_gradient(const float* in, float* dx, float* dy, int nx, int ny) {
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const int k = i * nx + j;
            const int jl = j > 0 ? k - 1 : k, jr = j < nx - 1 ? k + 1 : k;
            const int iu = i > 0 ? k - nx : k, id = i < ny - 1 ? k + nx : k;
            dx[k] = (float)(0.5 * (in[jr] - in[jl]));
            dy[k] = (float)(0.5 * (in[id] - in[iu]));
        }
}

// Forward differences, zero on the last row / column (mask.c:97-137).
void forward_gradient(const float* f, float* fx, float* fy, int nx, int ny) {
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const int p = i * nx + j;
            fx[p] = j < nx - 1 ? f[p + 1] - f[p] : 0.f;
            fy[p] = i < ny - 1 ? f[p + nx] - f[p] : 0.f;
        }
}

// Backward-difference divergence with the reference's edge cases and
// their operand order (mask.c:39-90).
float div_at(const float* v1, const float* v2, int i, int j, int nx, int ny) {
    const int p = i * nx + j;
    const bool L = j == 0, R = j == nx - 1, T = i == 0, Bm = i == ny - 1;
    if (!L && !R && !T && !Bm) return (v1[p] - v1[p - 1]) + (v2[p] - v2[p - nx]);
    if (T && L) return v1[0] + v2[0];
    if (T && R) return -v1[nx - 2] + v2[nx - 1];
    if (Bm && L) return v1[p] - v2[p - nx];
    if (Bm && R) return -v1[p - 1] - v2[p - nx];
    if (T) return v1[p] - v1[p - 1] + v2[p];
    if (Bm) return v1[p] - v1[p - 1] - v2[p - nx];
    if (L) return v1[p] + v2[p] - v2[p - nx];
    return -v1[p - 1] + v2[p] - v2[p - nx];  // right column
}

}  // namespace

extern "C" {

// One pyramid level (tvl1flow_lib.c:96-305).  iters_out (optional) gets
// the iteration count of each warp.
void orc_tvl1_scale(const float* I0, const float* I1, float* u1, float* u2, int nx, int ny, float tau,
                    float lambda, float theta, int warps, float epsilon, int* iters_out) {
    const size_t n = (size_t)nx * ny;
    const float l_t = lambda * theta;
    Img I1x(n), I1y(n), I1w(n), I1wx(n), I1wy(n), rho_c(n), v1(n), v2(n), grad(n);
    Img p11(n, 0.f), p12(n, 0.f), p21(n, 0.f), p22(n, 0.f), dp1(n), dp2(n), u1x(n), u1y(n), u2x(n), u2y(n);
    centered_gradient(I1, I1x.data(), I1y.data(), nx, ny);
    for (int w = 0; w < warps; w++) {
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) {
                const size_t p = (size_t)i * nx + j;
                const float uu = (float)(j + u1[p]), vv = (float)(i + u2[p]);
                I1w[p] = bicubic_at(I1, uu, vv, nx, ny, true);
                I1wx[p] = bicubic_at(I1x.data(), uu, vv, nx, ny, true);
                I1wy[p] = bicubic_at(I1y.data(), uu, vv, nx, ny, true);
            }
        for (size_t i = 0; i < n; i++) {
            const float Ix2 = I1wx[i] * I1wx[i], Iy2 = I1wy[i] * I1wy[i];
            grad[i] = Ix2 + Iy2;
            rho_c[i] = I1w[i] - I1wx[i] * u1[i] - I1wy[i] * u2[i] - I0[i];
        }
        int it = 0;
        float error = INFINITY;
        while (error > epsilon * epsilon && it < 300) {
            it++;
            // thresholding step towards the data term (lines 170-205)
            for (size_t i = 0; i < n; i++) {
                const float rho = rho_c[i] + (I1wx[i] * u1[i] + I1wy[i] * u2[i]);
                float d1, d2;
                if (rho < -l_t * grad[i]) {
                    d1 = l_t * I1wx[i];
                    d2 = l_t * I1wy[i];
                } else if (rho > l_t * grad[i]) {
                    d1 = -l_t * I1wx[i];
                    d2 = -l_t * I1wy[i];
                } else if (grad[i] < 1E-10) {
                    d1 = d2 = 0;
                } else {
                    const float fi = -rho / grad[i];
                    d1 = fi * I1wx[i];
                    d2 = fi * I1wy[i];
                }
                v1[i] = u1[i] + d1;
                v2[i] = u2[i] + d2;
            }
            for (int i = 0; i < ny; i++)
                for (int j = 0; j < nx; j++) {
                    dp1[(size_t)i * nx + j] = div_at(p11.data(), p12.data(), i, j, nx, ny);
                    dp2[(size_t)i * nx + j] = div_at(p21.data(), p22.data(), i, j, nx, ny);
                }
            // primal update and the mean squared change (lines 211-230)
            error = 0.0;
            for (size_t i = 0; i < n; i++) {
                const float a = u1[i], b = u2[i];
                u1[i] = v1[i] + theta * dp1[i];
                u2[i] = v2[i] + theta * dp2[i];
                error += (u1[i] - a) * (u1[i] - a) + (u2[i] - b) * (u2[i] - b);
            }
            error /= n;
            // dual update (lines 233-252)
            forward_gradient(u1, u1x.data(), u1y.data(), nx, ny);
            forward_gradient(u2, u2x.data(), u2y.data(), nx, ny);
            const float taut = tau / theta;
            for (size_t i = 0; i < n; i++) {
                const float g1 = (float)std::hypot((double)u1x[i], (double)u1y[i]);
                const float g2 = (float)std::hypot((double)u2x[i], (double)u2y[i]);
                const float ng1 = (float)(1.0 + taut * g1), ng2 = (float)(1.0 + taut * g2);
                p11[i] = (p11[i] + taut * u1x[i]) / ng1;
                p12[i] = (p12[i] + taut * u1y[i]) / ng1;
                p21[i] = (p21[i] + taut * u2x[i]) / ng2;
                p22[i] = (p22[i] + taut * u2y[i]) / ng2;
            }
        }
        if (iters_out) iters_out[w] = it;
    }
}

// Rescaling of both images to [0, 255] by their joint range
// (tvl1flow_lib.c:341-376).
void orc_tvl1_normalize(const float* I0, const float* I1, float* I0n, float* I1n, int size) {
    float mn0 = I0[0], mx0 = I0[0], mn1 = I1[0], mx1 = I1[0];
    for (int i = 1; i < size; i++) {
        if (I0[i] < mn0) mn0 = I0[i];
        if (I0[i] > mx0) mx0 = I0[i];
        if (I1[i] < mn1) mn1 = I1[i];
        if (I1[i] > mx1) mx1 = I1[i];
    }
    const float mx = mx0 > mx1 ? mx0 : mx1, mn = mn0 < mn1 ? mn0 : mn1, den = mx - mn;
    for (int i = 0; i < size; i++) {
        if (den > 0) {
            I0n[i] = (float)(255.0 * (I0[i] - mn) / den);
            I1n[i] = (float)(255.0 * (I1[i] - mn) / den);
        } else {
            I0n[i] = I0[i];
            I1n[i] = I1[i];
        }
    }
}

// The pyramid: normalisation, 0.8-sigma pre-smoothing, zoom_out levels,
// coarse-to-fine solve with bicubic upsampling of the flow scaled by
// 1/zfactor; levels finer than fscale are only upsampled
// (tvl1flow_lib.c:384-485).  iters (optional): nscales * warps counts,
// level-major from the coarsest solved level.
void orc_tvl1_multiscale(const float* I0, const float* I1, float* u1, float* u2, int nxx, int nyy, float tau,
                         float lambda, float theta, int nscales, int fscale, float zfactor, int warps,
                         float epsilon, int* iters) {
    std::vector<Img> I0s(nscales), I1s(nscales), u1s(nscales), u2s(nscales);
    std::vector<int> nx(nscales), ny(nscales);
    const size_t size = (size_t)nxx * nyy;
    nx[0] = nxx;
    ny[0] = nyy;
    I0s[0].resize(size);
    I1s[0].resize(size);
    u1s[0].assign(size, 0.f);
    u2s[0].assign(size, 0.f);
    orc_tvl1_normalize(I0, I1, I0s[0].data(), I1s[0].data(), (int)size);
    gaussian(I0s[0].data(), nx[0], ny[0], 0.8);
    gaussian(I1s[0].data(), nx[0], ny[0], 0.8);
    for (int s = 1; s < nscales; s++) {
        zoom_size(nx[s - 1], ny[s - 1], &nx[s], &ny[s], zfactor);
        const size_t sz = (size_t)nx[s] * ny[s];
        I0s[s].resize(sz);
        I1s[s].resize(sz);
        u1s[s].assign(sz, 0.f);
        u2s[s].assign(sz, 0.f);
        zoom_out(I0s[s - 1].data(), I0s[s].data(), nx[s - 1], ny[s - 1], zfactor);
        zoom_out(I1s[s - 1].data(), I1s[s].data(), nx[s - 1], ny[s - 1], zfactor);
    }
    int k = 0;
    const float inv = (float)1.0 / zfactor;
    auto upsample = [&](int s) {
        zoom_in(u1s[s].data(), u1s[s - 1].data(), nx[s], ny[s], nx[s - 1], ny[s - 1]);
        zoom_in(u2s[s].data(), u2s[s - 1].data(), nx[s], ny[s], nx[s - 1], ny[s - 1]);
        for (size_t i = 0; i < (size_t)nx[s - 1] * ny[s - 1]; i++) {
            u1s[s - 1][i] *= inv;
            u2s[s - 1][i] *= inv;
        }
    };
    for (int s = nscales - 1; s >= fscale; s--) {
        orc_tvl1_scale(I0s[s].data(), I1s[s].data(), u1s[s].data(), u2s[s].data(), nx[s], ny[s], tau, lambda,
                       theta, warps, epsilon, iters ? iters + (size_t)k * warps : nullptr);
        k++;
        if (!s) break;
        upsample(s);
    }
    for (int s = fscale - 1; s >= 1; s--) upsample(s);
    std::memcpy(u1, u1s[0].data(), size * sizeof(float));
    std::memcpy(u2, u2s[0].data(), size * sizeof(float));
}

// The number of levels the reference's command line uses: the requested
// count, capped so that the coarsest level stays about 16 pixels across
// (tvl1flow/main.c:139-146).
int orc_tvl1_nscales(int nx, int ny, float zfactor, int nscales) {
    const float N = (float)(1 + std::log(std::hypot((double)nx, (double)ny) / 16.0) / std::log((double)(1 / zfactor)));
    if (N < nscales) nscales = (int)N;
    return nscales;
}

// Building blocks, exported for the pins.
void orc_tvl1_gaussian(float* I, int nx, int ny, double sigma) { gaussian(I, nx, ny, sigma); }
float orc_tvl1_bicubic(const float* in, float uu, float vv, int nx, int ny, int border_out) {
    return bicubic_at(in, uu, vv, nx, ny, border_out != 0);
}
void orc_tvl1_zoom_out(const float* I, float* Iout, int nx, int ny, float factor) { zoom_out(I, Iout, nx, ny, factor); }
void orc_tvl1_zoom_in(const float* I, float* Iout, int nx, int ny, int nxx, int nyy) {
    zoom_in(I, Iout, nx, ny, nxx, nyy);
}
void orc_tvl1_divergence(const float* v1, const float* v2, float* div, int nx, int ny) {
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) div[(size_t)i * nx + j] = div_at(v1, v2, i, j, nx, ny);
}
void orc_tvl1_centered_gradient(const float* in, float* dx, float* dy, int nx, int ny) {
    centered_gradient(in, dx, dy, nx, ny);
}
void orc_tvl1_forward_gradient(const float* f, float* fx, float* fy, int nx, int ny) {
    forward_gradient(f, fx, fy, nx, ny);
}

}  // extern "C"
