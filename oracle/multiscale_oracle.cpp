// multiscale_oracle.cpp -- CPU restatement of the multiscale pyramid the
// reference ships for multiscale denoising (multiscaler/multiscaler.cpp,
// decompose.cpp, recompose.cpp, merge_coarse.cpp).  TEST INFRASTRUCTURE
// ONLY: the GPU pyramid kernels are checked against it; the product never
// links it.
//
// Images are planar float [c][rows][cols] (the reference interleaves
// channels; every operation here acts on one channel at a time, so the
// layout does not change a value).  The Gaussian and Lanczos pyramids are
// written in the reference's float operation order, compiled without
// contraction: bit for bit the reference's results.  The DCT pyramid
// replaces FFTW's REDFT10/REDFT01 (which this image does not have) by the
// transforms' definitions summed in double and rounded once to float:
// the exact transform, where FFTW's float result differs from it in the
// last bits.
#include <cmath>
#include <cstring>
#include <vector>

namespace {

using Img = std::vector<float>;

// 2D REDFT10 of every channel, divided by 4 rows cols (multiscaler.cpp:
// 21-58, the non-isometric normalisation).
void dct2(float* img, int rows, int cols, int ch) {
    std::vector<double> cr((size_t)rows * rows), cc((size_t)cols * cols), t((size_t)rows * cols);
    for (int k = 0; k < rows; k++)
        for (int n = 0; n < rows; n++) cr[(size_t)k * rows + n] = 2.0 * std::cos(M_PI * k * (2.0 * n + 1.0) / (2.0 * rows));
    for (int k = 0; k < cols; k++)
        for (int n = 0; n < cols; n++) cc[(size_t)k * cols + n] = 2.0 * std::cos(M_PI * k * (2.0 * n + 1.0) / (2.0 * cols));
    for (int c = 0; c < ch; c++) {
        float* x = img + (size_t)c * rows * cols;
        for (int r = 0; r < rows; r++)
            for (int k = 0; k < cols; k++) {
                double s = 0;
                for (int n = 0; n < cols; n++) s += cc[(size_t)k * cols + n] * x[(size_t)r * cols + n];
                t[(size_t)r * cols + k] = s;
            }
        for (int k = 0; k < rows; k++)
            for (int q = 0; q < cols; q++) {
                double s = 0;
                for (int n = 0; n < rows; n++) s += cr[(size_t)k * rows + n] * t[(size_t)n * cols + q];
                x[(size_t)k * cols + q] = (float)s;
            }
        const float d = (float)(4 * rows * cols);
        for (size_t i = 0; i < (size_t)rows * cols; i++) x[i] /= d;
    }
}

// 2D REDFT01 of every channel, unnormalised (multiscaler.cpp:60-98).
void idct2(float* img, int rows, int cols, int ch) {
    auto table = [](int N) {
        std::vector<double> a((size_t)N * N);
        for (int k = 0; k < N; k++)
            for (int n = 0; n < N; n++) a[(size_t)k * N + n] = n == 0 ? 1.0 : 2.0 * std::cos(M_PI * n * (2.0 * k + 1.0) / (2.0 * N));
        return a;
    };
    const std::vector<double> cr = table(rows), cc = table(cols);
    std::vector<double> t((size_t)rows * cols);
    for (int c = 0; c < ch; c++) {
        float* x = img + (size_t)c * rows * cols;
        for (int r = 0; r < rows; r++)
            for (int k = 0; k < cols; k++) {
                double s = 0;
                for (int n = 0; n < cols; n++) s += cc[(size_t)k * cols + n] * x[(size_t)r * cols + n];
                t[(size_t)r * cols + k] = s;
            }
        for (int k = 0; k < rows; k++)
            for (int q = 0; q < cols; q++) {
                double s = 0;
                for (int n = 0; n < rows; n++) s += cr[(size_t)k * rows + n] * t[(size_t)n * cols + q];
                x[(size_t)k * cols + q] = (float)s;
            }
    }
}

// Gaussian taps with the reference's integer truncation of (i - mean) /
// sigma, normalised by their double sum (multiscaler.cpp:261-276).
std::vector<float> gaussian_taps(int n, float sigma, float mean) {
    std::vector<float> k(n);
    double sum = 0.;
    for (int i = 0; i < n; i++) {
        const int val = (int)((i - mean) / sigma);
        k[i] = (float)std::exp(-0.5 * val * val);
        sum += k[i];
    }
    if (sum > 0.)
        for (int i = 0; i < n; i++) k[i] = (float)(k[i] / sum);
    return k;
}

float lanczos3(float x) {
    if (std::abs(x) >= 3) return 0.f;
    if (x == 0.) return 1.f;
    return (float)(std::sin(M_PI * x) / (M_PI * x) * std::sin(M_PI * x / 3.) / (M_PI * x / 3.));
}

// Lanczos-3 taps at half rate (multiscaler.cpp:404-418).
std::vector<float> lanczos_taps(int n, float offset) {
    std::vector<float> k(n);
    double sum = 0.;
    for (int i = 0; i < n; i++) {
        k[i] = (float)(lanczos3((float)((i - offset) / 2.)) / 2.);
        sum += k[i];
    }
    if (sum > 0.)
        for (int i = 0; i < n; i++) k[i] = (float)(k[i] / sum);
    return k;
}

// Separable convolution in place, rows axis first, taps at
// round(x - offset + i) clamped to the image (multiscaler.cpp:136-185).
void conv_clamped(float* img, int rows, int cols, int ch, const std::vector<float>& k, float offset) {
    const int n = (int)k.size();
    std::vector<float> tmp((size_t)rows * cols);
    for (int c = 0; c < ch; c++) {
        float* x = img + (size_t)c * rows * cols;
        for (int r = 0; r < rows; r++)
            for (int q = 0; q < cols; q++) {
                float val = 0.0;
                for (int i = 0; i < n; i++) {
                    int j = (int)std::round(r - offset + i);
                    j = std::min(std::max(j, 0), rows - 1);
                    val += x[(size_t)j * cols + q] * k[i];
                }
                tmp[(size_t)r * cols + q] = val;
            }
        for (int r = 0; r < rows; r++)
            for (int q = 0; q < cols; q++) {
                float val = 0.0;
                for (int i = 0; i < n; i++) {
                    int j = (int)std::round(q - offset + i);
                    j = std::min(std::max(j, 0), cols - 1);
                    val += tmp[(size_t)r * cols + j] * k[i];
                }
                x[(size_t)r * cols + q] = val;
            }
    }
}

// One-direction convolution with symmetric (half-sample) boundaries and
// truncated tap positions (multiscaler.cpp:190-247); dir 0 = along rows.
Img conv_dir(const float* img, int rows, int cols, int ch, const std::vector<float>& k, float offset, int dir) {
    const int n = (int)k.size();
    Img out((size_t)ch * rows * cols);
    const int X = dir == 0 ? rows : cols, X2 = 2 * X;
    for (int c = 0; c < ch; c++)
        for (int r = 0; r < rows; r++)
            for (int q = 0; q < cols; q++) {
                const int x = dir == 0 ? r : q;
                float val = 0.0;
                for (int i = 0; i < n; i++) {
                    int j = (int)(x - offset + i);
                    while (j < 0) j += X2;
                    while (j >= X2) j -= X2;
                    if (j >= X) j = X2 - 1 - j;
                    const size_t at = dir == 0 ? (size_t)j * cols + q : (size_t)r * cols + j;
                    val += img[(size_t)c * rows * cols + at] * k[i];
                }
                out[(size_t)c * rows * cols + (size_t)r * cols + q] = val;
            }
    return out;
}

// Blur, then keep every other sample (gdown :281-331 with sigma 1.39,
// ldown :421-443 with the Lanczos taps).
Img down(const float* img, int rows, int cols, int ch, bool lanczos) {
    std::vector<float> k;
    float offset;
    if (lanczos) {
        offset = 5.5f;
        k = lanczos_taps((int)(1 + 2 * offset), offset);
    } else {
        const float sigma = 1.39f, prec = 3.0f;
        offset = (float)(std::ceil(sigma * std::sqrt(2.0 * prec * std::log(10.0))) + 0.5);
        k = gaussian_taps((int)(1 + 2 * offset), sigma, offset);
    }
    Img b(img, img + (size_t)ch * rows * cols);
    conv_clamped(b.data(), rows, cols, ch, k, offset);
    const int r2 = rows / 2, c2 = cols / 2;
    Img out((size_t)ch * r2 * c2);
    for (int c = 0; c < ch; c++)
        for (int r = 0; r < r2; r++)
            for (int q = 0; q < c2; q++)
                out[((size_t)c * r2 + r) * c2 + q] = b[((size_t)c * rows + 2 * r) * cols + 2 * q];
    return out;
}

// Double the size to sx x sy (gup :337-398, lup :448-509): four phases of
// separable interpolation, the odd row / column of an uneven target
// copied from the last computed one.
Img up(const float* img, int rows, int cols, int ch, int sx, int sy, bool lanczos) {
    std::vector<float> k1, k2;
    float o1, o2;
    if (lanczos) {
        o1 = 2.25f;
        o2 = 2.75f;
        k1 = lanczos_taps(6, o1);
        k2 = lanczos_taps(6, o2);
    } else {
        o1 = 2.f;
        o2 = 1.5f;
        k1 = gaussian_taps(5, 0.5f, o1);
        k2 = gaussian_taps(5, 0.5f, o2);
    }
    Img out((size_t)ch * sx * sy, 0.f);
    auto O = [&](int c, int row, int col) -> float& { return out[((size_t)c * sx + row) * sy + col]; };
    auto at = [&](const Img& t, int c, int row, int col) { return t[((size_t)c * rows + row) * cols + col]; };
    Img t1 = conv_dir(img, rows, cols, ch, k1, o1, 0);
    Img t2 = conv_dir(t1.data(), rows, cols, ch, k1, o1, 1);
    for (int c = 0; c < ch; c++)
        for (int j = 0; j < rows; j++)
            for (int q = 0; q < cols; q++) O(c, 2 * j, 2 * q) = at(t2, c, j, q);
    t2 = conv_dir(t1.data(), rows, cols, ch, k2, o2, 1);
    for (int c = 0; c < ch; c++)
        for (int j = 0; j < rows; j++)
            for (int q = 0; q < cols; q++) O(c, 2 * j, 2 * q + 1) = at(t2, c, j, q);
    if (sy % 2 == 1)
        for (int c = 0; c < ch; c++)
            for (int j = 0; j < rows; j++) O(c, 2 * j, sy - 1) = at(t2, c, j, cols - 1);
    t1 = conv_dir(img, rows, cols, ch, k2, o2, 0);
    t2 = conv_dir(t1.data(), rows, cols, ch, k1, o1, 1);
    for (int c = 0; c < ch; c++)
        for (int j = 0; j < rows; j++)
            for (int q = 0; q < cols; q++) O(c, 2 * j + 1, 2 * q) = at(t2, c, j, q);
    if (sx % 2 == 1)
        for (int c = 0; c < ch; c++)
            for (int q = 0; q < cols; q++) O(c, sx - 1, 2 * q) = at(t2, c, rows - 1, q);
    t2 = conv_dir(t1.data(), rows, cols, ch, k2, o2, 1);
    for (int c = 0; c < ch; c++)
        for (int j = 0; j < rows; j++)
            for (int q = 0; q < cols; q++) O(c, 2 * j + 1, 2 * q + 1) = at(t2, c, j, q);
    if (sx % 2 == 1)
        for (int c = 0; c < ch; c++)
            for (int q = 0; q < cols; q++) O(c, sx - 1, 2 * q + 1) = at(t2, c, rows - 1, q);
    if (sy % 2 == 1)
        for (int c = 0; c < ch; c++)
            for (int j = 0; j < rows; j++) O(c, 2 * j + 1, sy - 1) = at(t2, c, j, cols - 1);
    if (sx % 2 == 1 && sy % 2 == 1)
        for (int c = 0; c < ch; c++) O(c, sx - 1, sy - 1) = at(t2, c, rows - 1, cols - 1);
    return out;
}

// Gaussian blur of standard deviation sigma, skipped below 0.41
// (gblur :279-304).
void gblur(float* img, int rows, int cols, int ch, float sigma) {
    if (sigma < 0.41) return;
    const float prec = 3.0f;
    const float offset = (float)((int)std::ceil(sigma * std::sqrt(2.0 * prec * std::log(10.0)))) - 0.5f;
    const int n = (int)(2 * offset);
    conv_clamped(img, rows, cols, ch, gaussian_taps(n, sigma, offset), offset);
}

}  // namespace

extern "C" {

// Level sizes: each level halves the previous one, integer division
// (decompose.cpp:46-53 for the DCT pyramid, gdown / ldown otherwise).
void orc_pyr_sizes(int rows, int cols, int levels, int* lr, int* lc) {
    for (int i = 0; i < levels; i++) {
        lr[i] = rows;
        lc[i] = cols;
        rows /= 2;
        cols /= 2;
    }
}

// decompose (decompose.cpp:17-106): `out` holds the levels one after the
// other, planar, level i of size lr[i] x lc[i].  type 0 DCT, 1 Gaussian,
// 2 Lanczos.
void orc_pyr_decompose(const float* img, int rows, int cols, int ch, int levels, int type, float* out) {
    std::vector<int> lr(levels), lc(levels);
    orc_pyr_sizes(rows, cols, levels, lr.data(), lc.data());
    float* o = out;
    if (type == 0) {
        Img d(img, img + (size_t)ch * rows * cols);
        dct2(d.data(), rows, cols, ch);
        for (int i = 0; i < levels; i++) {
            const int h = lr[i], w = lc[i];
            for (int c = 0; c < ch; c++)
                for (int r = 0; r < h; r++)
                    for (int q = 0; q < w; q++)
                        o[((size_t)c * h + r) * w + q] = d[((size_t)c * rows + r) * cols + q] * 1.0;
            idct2(o, h, w, ch);
            o += (size_t)ch * h * w;
        }
        return;
    }
    Img cur(img, img + (size_t)ch * rows * cols);
    std::memcpy(o, cur.data(), cur.size() * sizeof(float));
    o += cur.size();
    for (int i = 1; i < levels; i++) {
        cur = down(cur.data(), lr[i - 1], lc[i - 1], ch, type == 2);
        std::memcpy(o, cur.data(), cur.size() * sizeof(float));
        o += cur.size();
    }
}

// recompose (recompose.cpp:17-125): levels as decompose lays them out.
void orc_pyr_recompose(const float* lev, int rows, int cols, int ch, int levels, int type, float factor, float* out) {
    std::vector<int> lr(levels), lc(levels);
    orc_pyr_sizes(rows, cols, levels, lr.data(), lc.data());
    std::vector<size_t> off(levels);
    size_t o = 0;
    for (int i = 0; i < levels; i++) {
        off[i] = o;
        o += (size_t)ch * lr[i] * lc[i];
    }
    if (type == 0) {
        Img res(lev, lev + (size_t)ch * rows * cols);
        dct2(res.data(), rows, cols, ch);
        for (int i = 1; i < levels; i++) {
            const int h = lr[i], w = lc[i];
            Img im(lev + off[i], lev + off[i] + (size_t)ch * h * w);
            dct2(im.data(), h, w, ch);
            for (int j = 0; j < h * factor; ++j)
                for (int k = 0; k < w * factor; ++k)
                    for (int c = 0; c < ch; c++)
                        res[((size_t)c * rows + j) * cols + k] = im[((size_t)c * h + j) * w + k] * 1.0;
        }
        idct2(res.data(), rows, cols, ch);
        std::memcpy(out, res.data(), res.size() * sizeof(float));
        return;
    }
    const bool lz = type == 2;
    Img cur(lev + off[levels - 1], lev + off[levels - 1] + (size_t)ch * lr[levels - 1] * lc[levels - 1]);
    for (int i = levels - 2; i >= 0; --i) {
        const int h = lr[i], w = lc[i], h2 = lr[i + 1], w2 = lc[i + 1];
        const float* image = lev + off[i];
        Img small = down(image, h, w, ch, lz);
        for (size_t q = 0; q < cur.size(); q++) cur[q] = cur[q] - small[q];
        gblur(cur.data(), h2, w2, ch, factor);
        cur = up(cur.data(), h2, w2, ch, h, w, lz);
        for (size_t q = 0; q < cur.size(); q++) cur[q] = cur[q] + image[q];
    }
    std::memcpy(out, cur.data(), cur.size() * sizeof(float));
}

// merge_coarse (merge_coarse.cpp:16-49): the coarse image's low DCT
// frequencies replace the fine image's.
void orc_pyr_merge_coarse(const float* fine, int rows, int cols, const float* coarse, int crows, int ccols, int ch,
                          float factor, float* out) {
    Img f(fine, fine + (size_t)ch * rows * cols), c2(coarse, coarse + (size_t)ch * crows * ccols);
    dct2(f.data(), rows, cols, ch);
    dct2(c2.data(), crows, ccols, ch);
    for (int j = 0; j < crows * factor; ++j)
        for (int k = 0; k < ccols * factor; ++k)
            for (int c = 0; c < ch; c++) f[((size_t)c * rows + j) * cols + k] = c2[((size_t)c * crows + j) * ccols + k];
    idct2(f.data(), rows, cols, ch);
    std::memcpy(out, f.data(), f.size() * sizeof(float));
}

// Building blocks, exported for the pins.
void orc_pyr_dct2(float* img, int rows, int cols, int ch) { dct2(img, rows, cols, ch); }
void orc_pyr_idct2(float* img, int rows, int cols, int ch) { idct2(img, rows, cols, ch); }
void orc_pyr_gblur(float* img, int rows, int cols, int ch, float sigma) { gblur(img, rows, cols, ch, sigma); }
void orc_pyr_down(const float* img, int rows, int cols, int ch, int lanczos, float* out) {
    Img o = down(img, rows, cols, ch, lanczos != 0);
    std::memcpy(out, o.data(), o.size() * sizeof(float));
}
void orc_pyr_up(const float* img, int rows, int cols, int ch, int sx, int sy, int lanczos, float* out) {
    Img o = up(img, rows, cols, ch, sx, sy, lanczos != 0);
    std::memcpy(out, o.data(), o.size() * sizeof(float));
}

}  // extern "C"
