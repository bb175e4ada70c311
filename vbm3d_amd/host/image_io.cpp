// image_io.cpp -- see image_io.h.
#include "image_io.h"

#include <dlfcn.h>
#include <setjmp.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>

#include <jpeglib.h>
#include <png.h>

namespace {

std::string ext_of(const std::string& p)
{
    const size_t d = p.find_last_of('.');
    std::string e = d == std::string::npos ? "" : p.substr(d + 1);
    for (char& ch : e) ch = (char)std::tolower((unsigned char)ch);
    return e;
}

bool read_all(const std::string& path, std::vector<uint8_t>& buf, std::string& err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = path + ": cannot open"; return false; }
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

// interleaved [y][x][c] -> planar [c][y][x], with the reference's channel rule
void to_planar(const std::vector<float>& il, unsigned w, unsigned h, unsigned c,
               std::vector<float>& out, unsigned& cout)
{
    std::vector<float> pl((size_t)w * h * c);
    for (unsigned ch = 0, i = 0; ch < c; ch++)
        for (unsigned y = 0; y < h; y++)
            for (unsigned x = 0; x < w; x++, i++) pl[i] = il[((size_t)y * w + x) * c + ch];
    cout = c;
    if (c > 2) {
        const size_t wh = (size_t)w * h;
        size_t k = 0;
        while (k < wh && pl[k] == pl[wh + k] && pl[k] == pl[2 * wh + k]) k++;
        cout = k == wh ? 1 : 3;
    }
    pl.resize((size_t)w * h * cout);
    out.swap(pl);
}

// ---- PFM
bool load_pfm(const std::vector<uint8_t>& b, std::vector<float>& il, unsigned& w, unsigned& h,
              unsigned& c, std::string& err)
{
    std::string s(b.begin(), b.begin() + std::min<size_t>(b.size(), 256));
    std::istringstream is(s);
    std::string tag;
    double scale;
    is >> tag >> w >> h >> scale;
    if (tag != "PF" && tag != "Pf") { err = "not a PFM file"; return false; }
    c = tag == "PF" ? 3 : 1;
    size_t off = (size_t)is.tellg() + 1;
    const size_t n = (size_t)w * h * c;
    if (b.size() < off + n * 4) { err = "truncated PFM"; return false; }
    il.resize(n);
    const bool little = scale < 0;
    for (unsigned y = 0; y < h; y++) {  // rows stored bottom to top
        const uint8_t* row = b.data() + off + (size_t)(h - 1 - y) * w * c * 4;
        for (size_t i = 0; i < (size_t)w * c; i++) {
            uint8_t q[4];
            std::memcpy(q, row + 4 * i, 4);
            if (!little) std::swap(q[0], q[3]), std::swap(q[1], q[2]);
            std::memcpy(&il[(size_t)y * w * c + i], q, 4);
        }
    }
    return true;
}

bool save_pfm(const std::string& path, const std::vector<float>& il, unsigned w, unsigned h,
              unsigned c, std::string& err)
{
    if (c != 1 && c != 3) { err = "PFM holds 1 or 3 channels"; return false; }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = path + ": cannot write"; return false; }
    std::fprintf(f, "%s\n%u %u\n-1\n", c == 3 ? "PF" : "Pf", w, h);
    for (unsigned y = 0; y < h; y++)
        std::fwrite(il.data() + (size_t)(h - 1 - y) * w * c, 4, (size_t)w * c, f);
    std::fclose(f);
    return true;
}

// ---- PGM / PPM (binary)
bool load_pnm(const std::vector<uint8_t>& b, std::vector<float>& il, unsigned& w, unsigned& h,
              unsigned& c, std::string& err)
{
    size_t pos = 0;
    auto token = [&]() {
        std::string t;
        while (pos < b.size()) {
            if (b[pos] == '#') { while (pos < b.size() && b[pos] != '\n') pos++; continue; }
            if (std::isspace(b[pos])) { if (!t.empty()) break; pos++; continue; }
            t += (char)b[pos++];
        }
        return t;
    };
    const std::string tag = token();
    if (tag != "P5" && tag != "P6") { err = "not a binary PGM/PPM file"; return false; }
    c = tag == "P6" ? 3 : 1;
    w = std::stoul(token());
    h = std::stoul(token());
    const unsigned maxv = std::stoul(token());
    pos++;
    const size_t n = (size_t)w * h * c;
    const unsigned bps = maxv > 255 ? 2 : 1;
    if (b.size() < pos + n * bps) { err = "truncated PNM"; return false; }
    il.resize(n);
    for (size_t i = 0; i < n; i++)
        il[i] = bps == 1 ? b[pos + i] : (float)((b[pos + 2 * i] << 8) | b[pos + 2 * i + 1]);
    return true;
}

bool save_pnm(const std::string& path, const std::vector<float>& il, unsigned w, unsigned h,
              unsigned c, std::string& err)
{
    if (c != 1 && c != 3) { err = "PGM/PPM hold 1 or 3 channels"; return false; }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = path + ": cannot write"; return false; }
    std::fprintf(f, "%s\n%u %u\n255\n", c == 3 ? "P6" : "P5", w, h);
    std::vector<uint8_t> row((size_t)w * c);
    for (unsigned y = 0; y < h; y++) {
        for (size_t i = 0; i < row.size(); i++) {
            float v = std::round(il[(size_t)y * w * c + i]);
            row[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
        }
        std::fwrite(row.data(), 1, row.size(), f);
    }
    std::fclose(f);
    return true;
}

// ---- TIFF (baseline, uncompressed, chunky)
struct Tiff {
    const std::vector<uint8_t>& b;
    bool le;
    uint32_t u16(size_t o) const { return le ? b[o] | b[o + 1] << 8 : b[o] << 8 | b[o + 1]; }
    uint32_t u32(size_t o) const
    {
        return le ? (uint32_t)b[o] | b[o + 1] << 8 | b[o + 2] << 16 | (uint32_t)b[o + 3] << 24
                  : (uint32_t)b[o] << 24 | b[o + 1] << 16 | b[o + 2] << 8 | b[o + 3];
    }
};

bool load_tiff(const std::vector<uint8_t>& b, std::vector<float>& il, unsigned& w, unsigned& h,
               unsigned& c, std::string& err)
{
    if (b.size() < 8 || !((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M'))) {
        err = "not a TIFF file";
        return false;
    }
    Tiff t{b, b[0] == 'I'};
    if (t.u16(2) != 42) { err = "not a classic TIFF file"; return false; }
    size_t ifd = t.u32(4);
    const unsigned n = t.u16(ifd);
    unsigned bps = 8, fmt = 1, comp = 1, planar = 1, spp = 1;
    std::vector<uint32_t> offs, counts;
    w = h = 0;
    for (unsigned e = 0; e < n; e++) {
        const size_t o = ifd + 2 + 12 * e;
        const unsigned tag = t.u16(o), type = t.u16(o + 2);
        const uint32_t cnt = t.u32(o + 4);
        auto val = [&](uint32_t i) -> uint32_t {
            const size_t sz = type == 3 ? 2 : 4;
            const size_t at = cnt * sz <= 4 ? o + 8 + i * sz : t.u32(o + 8) + i * sz;
            return type == 3 ? t.u16(at) : t.u32(at);
        };
        switch (tag) {
            case 256: w = val(0); break;
            case 257: h = val(0); break;
            case 258: bps = val(0); break;
            case 259: comp = val(0); break;
            case 273: for (uint32_t i = 0; i < cnt; i++) offs.push_back(val(i)); break;
            case 277: spp = val(0); break;
            case 279: for (uint32_t i = 0; i < cnt; i++) counts.push_back(val(i)); break;
            case 284: planar = val(0); break;
            case 339: fmt = val(0); break;
            default: break;
        }
    }
    if (comp != 1) { err = "compressed TIFF is not supported"; return false; }
    if (planar != 1 && spp > 1) { err = "planar TIFF is not supported"; return false; }
    if (!((bps == 8 || bps == 16) && fmt == 1) && !(bps == 32 && fmt == 3)) {
        err = "TIFF sample format must be 8/16-bit unsigned or 32-bit float";
        return false;
    }
    c = spp;
    const size_t n_s = (size_t)w * h * c, bytes = bps / 8;
    std::vector<uint8_t> raw;
    raw.reserve(n_s * bytes);
    for (size_t s = 0; s < offs.size(); s++) {
        const size_t len = s < counts.size() ? counts[s] : n_s * bytes - raw.size();
        if (offs[s] + len > b.size()) { err = "truncated TIFF"; return false; }
        raw.insert(raw.end(), b.begin() + offs[s], b.begin() + offs[s] + len);
    }
    if (raw.size() < n_s * bytes) { err = "truncated TIFF"; return false; }
    il.resize(n_s);
    for (size_t i = 0; i < n_s; i++) {
        if (bps == 8) il[i] = raw[i];
        else if (bps == 16) il[i] = (float)(t.le ? raw[2 * i] | raw[2 * i + 1] << 8 : raw[2 * i] << 8 | raw[2 * i + 1]);
        else {
            uint8_t q[4];
            std::memcpy(q, &raw[4 * i], 4);
            if (!t.le) std::swap(q[0], q[3]), std::swap(q[1], q[2]);
            std::memcpy(&il[i], q, 4);
        }
    }
    return true;
}

// Little-endian TIFF; like iio (src/Utilities/iio.c:3642-3652), data
// whose values are all integers in [0, 255] are stored as 8-bit.
bool save_tiff(const std::string& path, const std::vector<float>& il, unsigned w, unsigned h,
               unsigned c, std::string& err)
{
    bool as8 = true;
    for (float v : il)
        if (!(v >= 0.f && v <= 255.f && v == std::floor(v))) { as8 = false; break; }
    const unsigned bps = as8 ? 8 : 32;
    const size_t data_bytes = il.size() * bps / 8;
    std::vector<uint8_t> out;
    auto p16 = [&](uint32_t v) { out.push_back(v & 255); out.push_back(v >> 8 & 255); };
    auto p32 = [&](uint32_t v) { p16(v & 0xFFFF); p16(v >> 16); };
    out = {'I', 'I'};
    p16(42);
    p32(8);
    const unsigned nent = 10;
    const size_t ifd_size = 2 + nent * 12 + 4;
    const size_t bps_at = 8 + ifd_size;          // c shorts of BitsPerSample
    const size_t fmt_at = bps_at + 2 * c;        // c shorts of SampleFormat
    const size_t data_at = (fmt_at + 2 * c + 15) & ~(size_t)15;
    p16(nent);
    auto ent = [&](unsigned tag, unsigned type, uint32_t cnt, uint32_t v) { p16(tag); p16(type); p32(cnt); p32(v); };
    ent(256, 4, 1, w);
    ent(257, 4, 1, h);
    ent(258, 3, c, c == 1 ? bps : (uint32_t)bps_at);
    ent(259, 3, 1, 1);
    ent(262, 3, 1, c >= 3 ? 2 : 1);
    ent(273, 4, 1, (uint32_t)data_at);
    ent(277, 3, 1, c);
    ent(278, 4, 1, h);
    ent(279, 4, 1, (uint32_t)data_bytes);
    ent(339, 3, c, c == 1 ? (as8 ? 1 : 3) : (uint32_t)fmt_at);
    p32(0);
    for (unsigned i = 0; i < c; i++) p16(bps);
    for (unsigned i = 0; i < c; i++) p16(as8 ? 1 : 3);
    out.resize(data_at, 0);
    if (as8) for (float v : il) out.push_back((uint8_t)v);
    else {
        const size_t o = out.size();
        out.resize(o + data_bytes);
        std::memcpy(out.data() + o, il.data(), data_bytes);
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = path + ": cannot write"; return false; }
    std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    return true;
}

// ---- PNG (libpng): 8- or 16-bit grey, grey + alpha, RGB or RGBA, any
// interlacing; palettes and low bit depths expand to 8 bits.  Samples load
// as their integer values (0-255 or 0-65535), as iio does
// (src/Utilities/iio.c); saving writes 8 bits, each float v stored as
// v + 0.5 clamped to [0, 255] and truncated (iio's float-to-uint8 rule,
// src/Utilities/iio.c:667, 736).
struct PngMem {
    const std::vector<uint8_t>* b;
    size_t pos;
    jmp_buf jb;
    char msg[200];
};

void png_fail(png_structp png, png_const_charp m) {
    PngMem* io = (PngMem*)png_get_error_ptr(png);
    std::snprintf(io->msg, sizeof io->msg, "%s", m);
    longjmp(io->jb, 1);
}

void png_quiet(png_structp, png_const_charp) {}

void png_read_mem(png_structp png, png_bytep out, png_size_t n) {
    PngMem* io = (PngMem*)png_get_io_ptr(png);
    if (io->pos + n > io->b->size()) png_error(png, "truncated PNG");
    std::memcpy(out, io->b->data() + io->pos, n);
    io->pos += n;
}

bool load_png(const std::vector<uint8_t>& b, std::vector<float>& il, unsigned& w, unsigned& h,
              unsigned& c, std::string& err)
{
    if (b.size() < 8 || png_sig_cmp(b.data(), 0, 8)) { err = "not a PNG file"; return false; }
    // everything libpng may leave half-built when it long-jumps out lives
    // on the heap, reached through pointers that do not change after setjmp
    std::unique_ptr<PngMem> io(new PngMem{&b, 0, {}, {0}});
    std::unique_ptr<std::vector<uint8_t>> buf(new std::vector<uint8_t>);
    std::unique_ptr<std::vector<png_bytep>> rowv(new std::vector<png_bytep>);
    png_structp png = png_create_read_struct(PNG_LIBPNG_VER_STRING, io.get(), png_fail, png_quiet);
    png_infop info = png ? png_create_info_struct(png) : nullptr;
    if (!png || !info) {
        png_destroy_read_struct(&png, &info, nullptr);
        err = "libpng initialisation failed";
        return false;
    }
    std::vector<uint8_t>& data = *buf;
    std::vector<png_bytep>& rows = *rowv;
    if (setjmp(io->jb)) {
        png_destroy_read_struct(&png, &info, nullptr);
        err = std::string("PNG: ") + io->msg;
        return false;
    }
    png_set_read_fn(png, io.get(), png_read_mem);
    png_read_info(png, info);
    png_set_palette_to_rgb(png);
    png_set_expand_gray_1_2_4_to_8(png);
    png_set_interlace_handling(png);
    png_read_update_info(png, info);
    w = png_get_image_width(png, info);
    h = png_get_image_height(png, info);
    c = png_get_channels(png, info);
    const unsigned depth = png_get_bit_depth(png, info);
    const size_t rb = png_get_rowbytes(png, info);
    data.resize(rb * h);
    rows.resize(h);
    for (unsigned y = 0; y < h; y++) rows[y] = data.data() + rb * y;
    png_read_image(png, rows.data());
    png_read_end(png, nullptr);
    png_destroy_read_struct(&png, &info, nullptr);
    il.resize((size_t)w * h * c);
    for (size_t i = 0; i < il.size(); i++)
        il[i] = depth == 16 ? (float)((data[2 * i] << 8) | data[2 * i + 1]) : (float)data[i];  // PNG is big-endian
    return true;
}

void png_write_mem(png_structp png, png_bytep in, png_size_t n) {
    std::vector<uint8_t>* out = (std::vector<uint8_t>*)png_get_io_ptr(png);
    out->insert(out->end(), in, in + n);
}

void png_flush_mem(png_structp) {}

bool save_png(const std::string& path, const std::vector<float>& il, unsigned w, unsigned h, unsigned c,
              std::string& err)
{
    if (c != 1 && c != 3) { err = path + ": PNG output takes 1 or 3 channels"; return false; }
    std::vector<uint8_t> px(il.size());
    for (size_t i = 0; i < il.size(); i++) {
        const double v = 0.5 + (double)il[i];
        px[i] = (uint8_t)(v > 0 ? (v < 255 ? v : 255) : 0);
    }
    std::unique_ptr<PngMem> io(new PngMem{nullptr, 0, {}, {0}});
    std::unique_ptr<std::vector<uint8_t>> obuf(new std::vector<uint8_t>);
    std::unique_ptr<std::vector<png_bytep>> rowv(new std::vector<png_bytep>(h));
    png_structp png = png_create_write_struct(PNG_LIBPNG_VER_STRING, io.get(), png_fail, png_quiet);
    png_infop info = png ? png_create_info_struct(png) : nullptr;
    if (!png || !info) {
        png_destroy_write_struct(&png, &info);
        err = "libpng initialisation failed";
        return false;
    }
    std::vector<uint8_t>& out = *obuf;
    std::vector<png_bytep>& rows = *rowv;
    if (setjmp(io->jb)) {
        png_destroy_write_struct(&png, &info);
        err = path + ": PNG: " + io->msg;
        return false;
    }
    png_set_write_fn(png, &out, png_write_mem, png_flush_mem);
    png_set_IHDR(png, info, w, h, 8, c == 1 ? PNG_COLOR_TYPE_GRAY : PNG_COLOR_TYPE_RGB, PNG_INTERLACE_NONE,
                 PNG_COMPRESSION_TYPE_DEFAULT, PNG_FILTER_TYPE_DEFAULT);
    for (unsigned y = 0; y < h; y++) rows[y] = px.data() + (size_t)y * w * c;
    png_write_info(png, info);
    png_write_image(png, rows.data());
    png_write_end(png, nullptr);
    png_destroy_write_struct(&png, &info);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = path + ": cannot write"; return false; }
    std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    return true;
}

// ---- JPEG (read only, as the reference's iio build reads it): libjpeg 9,
// loaded when the first JPEG frame is read, so that nothing else depends
// on it.  8-bit grey or RGB samples load as their integer values.
struct JpegApi {
    struct jpeg_error_mgr* (*std_error)(struct jpeg_error_mgr*);
    void (*create)(j_decompress_ptr, int, size_t);
    void (*mem_src)(j_decompress_ptr, const unsigned char*, unsigned long);
    int (*read_header)(j_decompress_ptr, boolean);
    boolean (*start)(j_decompress_ptr);
    JDIMENSION (*read_lines)(j_decompress_ptr, JSAMPARRAY, JDIMENSION);
    boolean (*finish)(j_decompress_ptr);
    void (*destroy)(j_decompress_ptr);
};

const JpegApi* jpeg_api(std::string& err) {
    static JpegApi api;
    static int state = 0;  // 0 not tried, 1 loaded, -1 unavailable
    if (state == 0) {
        void* h = nullptr;
        for (const char* name : {"libjpeg.so.9", "/opt/conda/lib/libjpeg.so.9"})
            if ((h = dlopen(name, RTLD_NOW | RTLD_LOCAL))) break;
        state = -1;
        if (h) {
            api.std_error = (decltype(api.std_error))dlsym(h, "jpeg_std_error");
            api.create = (decltype(api.create))dlsym(h, "jpeg_CreateDecompress");
            api.mem_src = (decltype(api.mem_src))dlsym(h, "jpeg_mem_src");
            api.read_header = (decltype(api.read_header))dlsym(h, "jpeg_read_header");
            api.start = (decltype(api.start))dlsym(h, "jpeg_start_decompress");
            api.read_lines = (decltype(api.read_lines))dlsym(h, "jpeg_read_scanlines");
            api.finish = (decltype(api.finish))dlsym(h, "jpeg_finish_decompress");
            api.destroy = (decltype(api.destroy))dlsym(h, "jpeg_destroy_decompress");
            if (api.std_error && api.create && api.mem_src && api.read_header && api.start && api.read_lines &&
                api.finish && api.destroy)
                state = 1;
        }
    }
    if (state != 1) {
        err = "JPEG input needs libjpeg 9 (libjpeg.so.9)";
        return nullptr;
    }
    return &api;
}

struct JpegErr {
    struct jpeg_error_mgr mgr;
    jmp_buf jb;
};

void jpeg_fail(j_common_ptr cinfo) { longjmp(((JpegErr*)cinfo->err)->jb, 1); }

bool load_jpeg(const std::vector<uint8_t>& b, std::vector<float>& il, unsigned& w, unsigned& h,
               unsigned& c, std::string& err)
{
    const JpegApi* J = jpeg_api(err);
    if (!J) return false;
    // state libjpeg may long-jump out of lives on the heap (see load_png)
    std::unique_ptr<jpeg_decompress_struct> cip(new jpeg_decompress_struct);
    std::unique_ptr<JpegErr> je(new JpegErr);
    std::unique_ptr<std::vector<uint8_t>> buf(new std::vector<uint8_t>);
    jpeg_decompress_struct& ci = *cip;
    std::vector<uint8_t>& data = *buf;
    ci.err = J->std_error(&je->mgr);
    je->mgr.error_exit = jpeg_fail;
    if (setjmp(je->jb)) {
        J->destroy(&ci);
        err = "corrupt JPEG data";
        return false;
    }
    J->create(&ci, JPEG_LIB_VERSION, sizeof ci);
    J->mem_src(&ci, b.data(), (unsigned long)b.size());
    J->read_header(&ci, TRUE);
    J->start(&ci);
    data.resize((size_t)ci.output_width * ci.output_height * ci.output_components);
    while (ci.output_scanline < ci.output_height) {
        JSAMPROW row = data.data() + (size_t)ci.output_scanline * ci.output_width * ci.output_components;
        J->read_lines(&ci, &row, 1);
    }
    J->finish(&ci);
    w = ci.output_width;
    h = ci.output_height;
    c = (unsigned)ci.output_components;
    J->destroy(&ci);
    il.assign(data.begin(), data.end());
    return true;
}

}  // namespace

// Interleaved samples of any supported file; `bytes` is true for the
// formats whose samples are 8-bit integers (PNG, JPEG and PNM at 8 bits,
// TIFF whose samples are all integers in [0, 255]).
bool load_interleaved(const std::string& path, std::vector<float>& il, unsigned& w, unsigned& h, unsigned& cc,
                      bool& bytes, std::string& err)
{
    std::vector<uint8_t> b;
    if (!read_all(path, b, err)) return false;
    const std::string e = ext_of(path);
    bool ok;
    bytes = false;
    if (e == "tif" || e == "tiff") ok = load_tiff(b, il, w, h, cc, err);
    else if (e == "pfm") ok = load_pfm(b, il, w, h, cc, err);
    else if (e == "pgm" || e == "ppm" || e == "pnm") ok = load_pnm(b, il, w, h, cc, err);
    else if (e == "png") ok = load_png(b, il, w, h, cc, err);
    else if (e == "jpg" || e == "jpeg") ok = load_jpeg(b, il, w, h, cc, err);
    else { err = path + ": unsupported image format (use .png, .jpg, .tif, .pfm, .pgm or .ppm)"; return false; }
    if (!ok) { err = path + ": " + err; return false; }
    if (e != "pfm") {
        bytes = true;
        for (float v : il)
            if (!(v >= 0.f && v <= 255.f && v == std::floor(v))) { bytes = false; break; }
    }
    return true;
}

bool load_frame(const std::string& path, std::vector<float>& planar, unsigned& w, unsigned& h,
                unsigned& c, std::string& err)
{
    std::vector<float> il;
    unsigned cc = 0;
    bool bytes;
    if (!load_interleaved(path, il, w, h, cc, bytes, err)) return false;
    to_planar(il, w, h, cc, planar, c);
    return true;
}

bool load_planar(const std::string& path, std::vector<float>& planar, unsigned& w, unsigned& h, unsigned& c,
                 std::string& err)
{
    std::vector<float> il;
    bool bytes;
    if (!load_interleaved(path, il, w, h, c, bytes, err)) return false;
    planar.resize(il.size());
    for (unsigned ch = 0, i = 0; ch < c; ch++)
        for (unsigned y = 0; y < h; y++)
            for (unsigned x = 0; x < w; x++, i++) planar[i] = il[((size_t)y * w + x) * c + ch];
    return true;
}

bool load_grey(const std::string& path, std::vector<float>& img, unsigned& w, unsigned& h, std::string& err)
{
    std::vector<float> il;
    unsigned cc = 0;
    bool bytes;
    if (!load_interleaved(path, il, w, h, cc, bytes, err)) return false;
    const size_t n = (size_t)w * h;
    img.resize(n);
    if (cc == 1) {
        std::memcpy(img.data(), il.data(), n * sizeof(float));
        return true;
    }
    if (cc != 3 && cc != 4) { err = path + ": non-scalarizable image"; return false; }
    for (size_t i = 0; i < n; i++) {
        const float* q = &il[i * cc];
        const double g = .299 * q[0] + .587 * q[1] + .114 * q[2];
        img[i] = bytes ? (float)(uint8_t)g : (float)g;
    }
    return true;
}

bool save_frame(const std::string& path, const std::vector<float>& planar, unsigned w, unsigned h,
                unsigned c, std::string& err)
{
    std::vector<float> il((size_t)w * h * c);
    for (unsigned ch = 0, k = 0; ch < c; ch++)
        for (unsigned y = 0; y < h; y++)
            for (unsigned x = 0; x < w; x++, k++) il[((size_t)y * w + x) * c + ch] = planar[k];
    const std::string e = ext_of(path);
    if (e == "tif" || e == "tiff") return save_tiff(path, il, w, h, c, err);
    if (e == "pfm") return save_pfm(path, il, w, h, c, err);
    if (e == "pgm" || e == "ppm" || e == "pnm") return save_pnm(path, il, w, h, c, err);
    if (e == "png") return save_png(path, il, w, h, c, err);
    err = path + ": unsupported output format (use .png, .tif, .pfm, .pgm or .ppm)";
    return false;
}

bool load_flow(const std::string& path, std::vector<float>& planar, unsigned& w, unsigned& h,
               std::string& err)
{
    std::vector<uint8_t> b;
    if (!read_all(path, b, err)) return false;
    if (ext_of(path) != "flo") {
        std::vector<float> il;
        unsigned c = 0;
        const std::string e = ext_of(path);
        bool ok = e == "tif" || e == "tiff" ? load_tiff(b, il, w, h, c, err)
                : e == "pfm" ? load_pfm(b, il, w, h, c, err) : (err = "unsupported flow format", false);
        if (!ok) { err = path + ": " + err; return false; }
        if (c != 2) { err = path + ": file doesn't correspond to an optical flow"; return false; }
        planar.resize((size_t)w * h * 2);
        for (unsigned ch = 0, k = 0; ch < 2; ch++)
            for (size_t i = 0; i < (size_t)w * h; i++, k++) planar[k] = il[i * 2 + ch];
        return true;
    }
    if (b.size() < 12) { err = path + ": truncated flow"; return false; }
    float tag;
    std::memcpy(&tag, b.data(), 4);
    if (tag != 202021.25f) { err = path + ": not a .flo file"; return false; }
    int32_t ww, hh;
    std::memcpy(&ww, b.data() + 4, 4);
    std::memcpy(&hh, b.data() + 8, 4);
    w = ww;
    h = hh;
    const size_t n = (size_t)w * h;
    if (b.size() < 12 + n * 8) { err = path + ": truncated flow"; return false; }
    planar.resize(2 * n);
    for (size_t i = 0; i < n; i++) {
        std::memcpy(&planar[i], b.data() + 12 + 8 * i, 4);
        std::memcpy(&planar[n + i], b.data() + 16 + 8 * i, 4);
    }
    return true;
}

bool save_flow(const std::string& path, const std::vector<float>& planar, unsigned w, unsigned h,
               std::string& err)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = path + ": cannot write"; return false; }
    const float tag = 202021.25f;
    const int32_t ww = (int32_t)w, hh = (int32_t)h;
    std::fwrite(&tag, 4, 1, f);
    std::fwrite(&ww, 4, 1, f);
    std::fwrite(&hh, 4, 1, f);
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) {
        std::fwrite(&planar[i], 4, 1, f);
        std::fwrite(&planar[n + i], 4, 1, f);
    }
    std::fclose(f);
    return true;
}

// C entry points of the frame and flow files, for callers outside C++
// (the command line's tests among them).  0 on success; -1 with the
// message in vbm3d_io_last_error() otherwise.  A load with out == NULL
// only reports the sizes.
namespace {
thread_local std::string g_io_err;
}

extern "C" {

const char* vbm3d_io_last_error(void) { return g_io_err.c_str(); }

int vbm3d_io_load_frame(const char* path, float* out, size_t cap, unsigned* w, unsigned* h, unsigned* c)
{
    std::vector<float> v;
    if (!load_frame(path, v, *w, *h, *c, g_io_err)) return -1;
    if (out) {
        if (cap < v.size()) { g_io_err = "buffer too small"; return -1; }
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    }
    return 0;
}

int vbm3d_io_save_frame(const char* path, const float* planar, unsigned w, unsigned h, unsigned c)
{
    std::vector<float> v(planar, planar + (size_t)w * h * c);
    return save_frame(path, v, w, h, c, g_io_err) ? 0 : -1;
}

int vbm3d_io_load_flow(const char* path, float* out, size_t cap, unsigned* w, unsigned* h)
{
    std::vector<float> v;
    if (!load_flow(path, v, *w, *h, g_io_err)) return -1;
    if (out) {
        if (cap < v.size()) { g_io_err = "buffer too small"; return -1; }
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    }
    return 0;
}

}  // extern "C"
