// image_io.cpp -- see image_io.h.
#include "image_io.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

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

}  // namespace

bool load_frame(const std::string& path, std::vector<float>& planar, unsigned& w, unsigned& h,
                unsigned& c, std::string& err)
{
    std::vector<uint8_t> b;
    if (!read_all(path, b, err)) return false;
    std::vector<float> il;
    unsigned cc = 0;
    const std::string e = ext_of(path);
    bool ok;
    if (e == "tif" || e == "tiff") ok = load_tiff(b, il, w, h, cc, err);
    else if (e == "pfm") ok = load_pfm(b, il, w, h, cc, err);
    else if (e == "pgm" || e == "ppm" || e == "pnm") ok = load_pnm(b, il, w, h, cc, err);
    else { err = path + ": unsupported image format (use .tif, .pfm, .pgm or .ppm)"; return false; }
    if (!ok) { err = path + ": " + err; return false; }
    to_planar(il, w, h, cc, planar, c);
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
    err = path + ": unsupported image format (use .tif, .pfm, .pgm or .ppm)";
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
