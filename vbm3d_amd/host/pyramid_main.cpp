// pyramid_main.cpp -- the multiscale pyramid programs on the GPU, drop-ins
// for the reference's decompose, recompose and merge_coarse
// (multiscaler/decompose.cpp, recompose.cpp, merge_coarse.cpp): the same
// positional arguments, options (-t type, -c factor, -h) and usage
// messages.  Built three times, PYR_TOOL selecting the program.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "image_io.h"
#include "vbm3d_hip.h"

#ifndef PYR_TOOL
#define PYR_TOOL 0
#endif

namespace {

// Removes "-name value" (or the flag "-name" when def is null) from argv
// and returns the value, or def when absent (the reference's pick_option
// contract, multiscaler.cpp:100-114).
const char* take_option(int* argc, char** argv, const char* name, const char* def) {
    const int has_value = def ? 1 : 0;
    for (int i = 0; i < *argc - has_value; i++) {
        if (argv[i][0] == '-' && std::strcmp(argv[i] + 1, name) == 0) {
            const char* r = has_value ? argv[i + 1] : argv[i];
            for (int j = i; j < *argc - has_value - 1; j++) argv[j] = argv[j + has_value + 1];
            *argc -= has_value + 1;
            return r;
        }
    }
    return def;
}

struct Ctx {
    vbm3d_ctx* c = nullptr;
    ~Ctx() {
        if (c) vbm3d_destroy(c);
    }
};

int fail_msg(const std::string& m) {
    std::cerr << "ERROR: " << m << std::endl;
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
#if PYR_TOOL == 0
    const bool usage = take_option(&argc, argv, "h", nullptr) != nullptr;
    const int type = std::atoi(take_option(&argc, argv, "t", "0"));
    if (argc != 5 || usage) {
        std::cerr << "Usage: " << argv[0] << " input prefix levels suffix [-t type: 0=DCT, 1=Gaussian, 2=Lanczos]"
                  << std::endl;
        return EXIT_FAILURE;
    }
    const std::string input = argv[1], prefix = argv[2], suffix = argv[4];
    const int levels = std::atoi(argv[3]);
    std::vector<float> img;
    unsigned w, h, c;
    std::string err;
    if (!load_planar(input, img, w, h, c, err)) return fail_msg(err);
    if (levels < 1) return fail_msg("levels must be >= 1");
    std::vector<int> lr(levels), lc(levels);
    if (vbm3d_pyr_levels((int)h, (int)w, levels, lr.data(), lc.data())) return fail_msg(vbm3d_last_error());
    size_t total = 0;
    for (int i = 0; i < levels; i++) total += (size_t)c * lr[i] * lc[i];
    std::vector<float> out(total);
    Ctx X;
    if (!(X.c = vbm3d_create(0))) return fail_msg(vbm3d_last_error());
    if (vbm3d_pyr_decompose(X.c, img.data(), (int)h, (int)w, (int)c, levels, type, out.data()))
        return fail_msg(vbm3d_last_error());
    size_t o = 0;
    for (int i = 0; i < levels; i++) {
        const size_t n = (size_t)c * lr[i] * lc[i];
        std::vector<float> lev(out.begin() + o, out.begin() + o + n);
        if (!save_frame(prefix + std::to_string(i) + suffix, lev, (unsigned)lc[i], (unsigned)lr[i], c, err))
            return fail_msg(err);
        o += n;
    }
    return EXIT_SUCCESS;
#elif PYR_TOOL == 1
    const float factor = (float)std::atof(take_option(&argc, argv, "c", ".8"));
    const bool usage = take_option(&argc, argv, "h", nullptr) != nullptr;
    const int type = std::atoi(take_option(&argc, argv, "t", "0"));
    if (argc != 5 || usage) {
        std::cerr << "Usage: " << argv[0]
                  << " prefix levels suffix output [-c factor] [-t type: 0=DCT, 1=Gaussian, 2=Lanczos]" << std::endl;
        return EXIT_FAILURE;
    }
    const std::string prefix = argv[1], suffix = argv[3], output = argv[4];
    const int levels = std::atoi(argv[2]);
    if (levels < 1) return fail_msg("levels must be >= 1");
    std::vector<float> all;
    unsigned w0 = 0, h0 = 0, c0 = 0;
    std::string err;
    for (int i = 0; i < levels; i++) {
        std::vector<float> lev;
        unsigned w, h, c;
        if (!load_planar(prefix + std::to_string(i) + suffix, lev, w, h, c, err)) return fail_msg(err);
        if (i == 0) {
            w0 = w;
            h0 = h;
            c0 = c;
        } else if (c != c0 || h != (h0 >> i) || w != (w0 >> i)) {
            return fail_msg("level " + std::to_string(i) + " does not halve the previous level's size");
        }
        all.insert(all.end(), lev.begin(), lev.end());
    }
    std::vector<float> out((size_t)c0 * h0 * w0);
    Ctx X;
    if (!(X.c = vbm3d_create(0))) return fail_msg(vbm3d_last_error());
    if (vbm3d_pyr_recompose(X.c, all.data(), (int)h0, (int)w0, (int)c0, levels, type, factor, out.data()))
        return fail_msg(vbm3d_last_error());
    if (!save_frame(output, out, w0, h0, c0, err)) return fail_msg(err);
    return EXIT_SUCCESS;
#else
    const float factor = (float)std::atof(take_option(&argc, argv, "c", ".8"));
    const bool usage = take_option(&argc, argv, "h", nullptr) != nullptr;
    if (argc != 4 || usage) {
        std::cerr << "Usage: " << argv[0] << " image coarse result [-c factor]" << std::endl;
        return EXIT_FAILURE;
    }
    std::vector<float> fine, coarse;
    unsigned w, h, c, cw, ch, cc;
    std::string err;
    if (!load_planar(argv[1], fine, w, h, c, err) || !load_planar(argv[2], coarse, cw, ch, cc, err)) return fail_msg(err);
    if (cc != c) return fail_msg("the images have different channel counts");
    std::vector<float> out(fine.size());
    Ctx X;
    if (!(X.c = vbm3d_create(0))) return fail_msg(vbm3d_last_error());
    if (vbm3d_pyr_merge_coarse(X.c, fine.data(), (int)h, (int)w, coarse.data(), (int)ch, (int)cw, (int)c, factor,
                               out.data()))
        return fail_msg(vbm3d_last_error());
    if (!save_frame(argv[3], out, w, h, c, err)) return fail_msg(err);
    return EXIT_SUCCESS;
#endif
}
