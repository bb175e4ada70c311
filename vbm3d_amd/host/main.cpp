/*
 * VBM3Ddenoising -- command line drop-in for the reference's CLI
 * (src/main.cpp:243-439): the same option table, default rules
 * (initializeParameters_1/_2, src/main.cpp:35-207), validation and exit
 * codes, noise model, PSNR/RMSE measure file and output files.  The
 * denoising runs on the GPU through run_vbm3d / libvbm3d_hip.so.
 *
 * Option parsing follows the reference's clo_option overloads
 * (src/Utilities/cmd_option.h:151-228): an option's value is the word
 * after it; options whose default is an int literal are read with atoi
 * (so "-tauHard 2.5" gives 2), float ones accept "a/b", bool ones are
 * false for "false", "off" or "0" (any case).
 *
 * Image formats: .png (read and write), .jpg/.jpeg (read), .tif/.tiff,
 * .pfm, .pgm/.ppm; flows: .flo (or 2-channel .tif/.pfm).
 */
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vbm3d_hip.h"
#include "image_io.h"
#include "vbm3d.h"

#define DCT 0
#define BIOR 1
#define HADAMARD 2
#define HAAR 3
#define NONE -1

namespace {

bool ieq(const char* a, const char* b)
{
    for (; *a && *b; a++, b++)
        if (std::tolower((unsigned char)*a) != std::tolower((unsigned char)*b)) return false;
    return *a == *b;
}

struct Cli {
    int argc;
    char** argv;
    bool help = false;

    Cli(int c, char** v) : argc(c), argv(v)
    {
        help = raw("-h") || raw("-help") || raw("--help");
    }
    // The word after `name`; the name itself when it is the last word;
    // null when absent.
    const char* raw(const char* name) const
    {
        int k = 0;
        while (k < argc && std::strcmp(argv[k], name)) ++k;
        if (k == argc) return nullptr;
        return k + 1 == argc ? argv[k] : argv[k + 1];
    }
    void show(const char* name, const std::string& v, const char* usage) const
    {
        if (help) std::fprintf(stderr, "    %-8s = %-12s : %s\n", name, v.c_str(), usage);
    }
    std::string str(const char* name, const char* def, const char* usage) const
    {
        const char* s = raw(name);
        std::string r = s ? s : def;
        show(name, r, usage);
        return r;
    }
    int i(const char* name, int def, const char* usage) const
    {
        const char* s = raw(name);
        const int r = s ? std::atoi(s) : def;
        show(name, std::to_string(r), usage);
        return r;
    }
    float f(const char* name, float def, const char* usage) const
    {
        const char* s = raw(name);
        float r = def;
        if (s) {
            float x = 0, y = 1;
            std::sscanf(s, "%g/%g", &x, &y);
            r = x / y;
        }
        char b[64];
        std::snprintf(b, sizeof b, "%g", r);
        show(name, b, usage);
        return r;
    }
    bool b(const char* name, bool def, const char* usage) const
    {
        const char* s = raw(name);
        const bool r = s ? !(ieq(s, "false") || ieq(s, "off") || ieq(s, "0")) : def;
        show(name, r ? "true" : "false", usage);
        return r;
    }
};

Parameters from_c(const vbm3d_params& c)
{
    Parameters p;
    p.k = c.k; p.kt = c.kt; p.p = c.p; p.N = c.N; p.tau = c.tau; p.d = c.d;
    p.Nf = c.Nf; p.Ns = c.Ns; p.Npr = c.Npr; p.Nb = c.Nb; p.T_2D = c.T_2D; p.T_3D = c.T_3D;
    p.lambda3D = c.lambda3D; p.mc = c.mc != 0; p.n = c.n;
    return p;
}

// printParameters (src/main.cpp:209-233)
void print_parameters(const Parameters& prms, const char* s)
{
    if (!prms.k) {
        std::printf("%s step skipped\n", s);
        return;
    }
    std::printf("Parameters %s step:\n", s);
    std::printf("\tpatch size k x k x kt = %dx%dx%d\n", prms.k, prms.k, prms.kt);
    std::printf("\tprocessing step p  = %d\n", prms.p);
    std::printf("\tpred. search Nf    = %d\n", prms.Nf);
    std::printf("\tpred. search Ns    = %d\n", prms.Ns);
    std::printf("\tpred. search Npr   = %d\n", prms.Npr);
    std::printf("\tpred. search Nb    = %d\n", prms.Nb);
    std::printf("\tpred. search d     = %f\n", prms.d);
    std::printf("\tpred. search N     = %d\n", prms.N);
    std::printf("\tpred. search tau   = %f\n", prms.tau);
    std::printf("\ttransform T_2D     = %d\n", prms.T_2D);
    std::printf("\ttransform T_3D     = %d\n", prms.T_3D);
    std::printf("\tthreshold lambda3D = %f\n", prms.lambda3D);
    std::printf("\tmotion comp.       = %s\n", prms.mc ? "true" : "false");
}

std::string frame_name(const std::string& pattern, unsigned f)
{
    char b[1024];
    std::snprintf(b, sizeof b, pattern.c_str(), f);
    return b;
}

// Video<float>::loadVideo (src/Utilities/LibVideoT.cpp:28-87)
void load_video(Video<float>& v, const std::string& pattern, unsigned first, unsigned last, unsigned step)
{
    v.clear();
    const unsigned frames = (last - first + 1) / step;
    unsigned t = 0;
    for (unsigned f = first; f <= last && t < frames; f += step, t++) {
        std::vector<float> fr;
        unsigned w, h, c;
        std::string err;
        if (!load_frame(frame_name(pattern, f), fr, w, h, c, err))
            throw std::runtime_error("loadVideo: " + err);
        if (t == 0) v.resize(w, h, frames, c);
        else if (w != v.sz.width || h != v.sz.height || c != v.sz.channels)
            throw std::runtime_error("loadVideo: size of " + frame_name(pattern, f) + " does not match video size");
        std::copy(fr.begin(), fr.end(), v.data.begin() + (size_t)t * v.sz.whc);
    }
}

// Video<float>::loadFullFlow (src/Utilities/LibVideoT.cpp:90-149)
void load_flows(Video<float>& v, const std::string& pattern, unsigned first, unsigned last, unsigned step)
{
    v.clear();
    const unsigned frames = (last - first + 1) / step;
    unsigned t = 0;
    for (unsigned f = first; f <= last && t < frames; f += step, t++) {
        std::vector<float> fr;
        unsigned w, h;
        std::string err;
        if (!load_flow(frame_name(pattern, f), fr, w, h, err)) throw std::runtime_error("loadFullFlow: " + err);
        if (t == 0) v.resize(w, h, frames, 2);
        else if (w != v.sz.width || h != v.sz.height)
            throw std::runtime_error("loadFullFlow: size of " + frame_name(pattern, f) + " does not match");
        std::copy(fr.begin(), fr.end(), v.data.begin() + (size_t)t * v.sz.whc);
    }
}

// Video<float>::saveVideo (src/Utilities/LibVideoT.cpp:152-188)
void save_video(const Video<float>& v, const std::string& pattern, unsigned first, unsigned step)
{
    for (unsigned t = 0; t < v.sz.frames; t++) {
        std::vector<float> fr(v.data.begin() + (size_t)t * v.sz.whc, v.data.begin() + (size_t)(t + 1) * v.sz.whc);
        std::string err;
        if (!save_frame(frame_name(pattern, first + t * step), fr, v.sz.width, v.sz.height, v.sz.channels, err))
            throw std::runtime_error("saveVideo: " + err);
    }
}

// VideoUtils::addNoise (src/Utilities/LibVideoT.hpp:541-564): MT19937
// seeded with the time, Box-Muller on two 53-bit uniforms.
void add_noise(const Video<float>& in, Video<float>& out, float sigma, bool verbose)
{
    if (verbose) std::printf("Add noise with sigma = %g\n", sigma);
    out = in;
    std::mt19937 mt((unsigned long)std::time(nullptr));
    auto res53 = [&]() {
        const unsigned long a = mt() >> 5, b = mt() >> 6;
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
    };
    for (unsigned k = 0; k < in.sz.whcf; k++) {
        const double a = res53(), b = res53();
        out(k) += (float)(sigma * sqrtl(-2.0l * std::log(a)) * std::cos(2.0l * M_PI * b));
    }
}

float clip(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// VideoUtils::computePSNR (src/Utilities/LibVideoT.hpp:577-597)
void compute_psnr(const Video<float>& a, const Video<float>& b, double& psnr, double& rmse)
{
    if (a.sz != b.sz) throw std::runtime_error("computePSNR: videos have different sizes");
    double sum = 0.f;
    for (unsigned k = 0; k < a.sz.whcf; k++) {
        const float d = clip(a(k), 0, 255) - clip(b(k), 0, 255);
        sum += d * d;
    }
    rmse = sqrtf(sum / (float)a.sz.whcf);
    psnr = 20.f * log10f(255.f / rmse);
}

// VideoUtils::computeDiff (src/Utilities/LibVideoT.hpp:611-630)
void compute_diff(const Video<float>& a, const Video<float>& b, Video<float>& d)
{
    if (a.sz != b.sz) throw std::runtime_error("computeDiff: videos have different sizes");
    d.resize(a.sz);
    for (unsigned k = 0; k < a.sz.whcf; k++) d(k) = clip((a(k) - b(k)) + (255.f - 0.f) / 2.f, 0.f, 255.f);
}

int run(int argc, char** argv)
{
    const Cli cli(argc, argv);
    if (cli.help) std::fprintf(stderr, "\n %s : Video denoising with VBM3D on the GPU\n\n", argv[0]);
    const std::string input_path = cli.str("-i", "", "< Input sequence");
    const std::string inbsc_path = cli.str("-b", "", "< Input basic sequence (replacing first pass)");
    const std::string noisy_path = cli.str("-nisy", "", "> Noisy sequence (only when -add is true)");
    const std::string final_path = cli.str("-deno", "deno_%03d.tiff", "> Denoised sequence");
    const std::string basic_path = cli.str("-bsic", "bsic_%03d.tiff", "> Basic denoised sequence");
    const std::string diff_path = cli.str("-diff", "", "> Difference sequence between noisy and output (only when -add is true)");
    const std::string diffi_path = cli.str("-diffi", "", "> Difference sequence between input and output");
    const std::string meas_path = cli.str("-meas", "measure.txt", "> Text file with PSNR/RMSE (only when -add is true)");
    const std::string fflow_path = cli.str("-fflow", "", "< Forward optical flow ");
    const std::string bflow_path = cli.str("-bflow", "", "< Backward optical flow ");

    const unsigned firstFrame = cli.i("-f", 0, "< Index of the first frame");
    const unsigned lastFrame = cli.i("-l", 0, "< Index of the last frame");
    const unsigned frameStep = cli.i("-s", 1, "< Frame step");

    const float fSigma = cli.f("-sigma", 0.f, "< Noise standard deviation");
    const bool addnoise = cli.b("-add", true, "< Add noise");
    const bool verbose = cli.b("-verbose", true, "> Verbose output");

    const int kHard = cli.i("-kHard", NONE, "< Spatial size of the patch (first pass)");
    const int ktHard = cli.i("-ktHard", NONE, "< Temporal size of the patch (first pass)");
    const int NfHard = cli.i("-NfHard", NONE, "< Number frames used before and after the reference (first pass)");
    const int NsHard = cli.i("-NsHard", NONE, "< Size of the search region in the reference frame (first pass)");
    const int NprHard = cli.i("-NprHard", NONE, "< Size of the search region in the other frames (first pass)");
    const int NbHard = cli.i("-NbHard", NONE, "< Maximum number of neighbors per frame (first pass)");
    const int pHard = cli.i("-pHard", NONE, "< Step between each patch (first pass)");
    const int NHard = cli.i("-NHard", NONE, "< Maximum number of neighbors (first pass)");
    const int dHard = cli.i("-dHard", NONE, "< Bias toward center patches (first pass)");
    const float tauHard = (float)cli.i("-tauHard", NONE, "< Distance threshold on neighbors (first pass)");
    const float lambda3D = (float)cli.i("-lambda3d", NONE, "< Coefficient threhsold (first pass)");
    const unsigned T_2D_hard = (unsigned)cli.i("-T2dh", NONE, "< 2D transform (first pass), choice is 1 (dct) or 2 (bior)");
    const unsigned T_3D_hard = (unsigned)cli.i("-T3dh", NONE, "< 1D transform (first pass), choice is 3 (hadamard) or 4 (haar)");

    const int kWien = cli.i("-kWien", NONE, "< Spatial size of the patch (second pass)");
    const int ktWien = cli.i("-ktWien", NONE, "< Temporal size of the patch (second pass)");
    const int NfWien = cli.i("-NfWien", NONE, "< Number of frames used before and after the reference (second pass)");
    const int NsWien = cli.i("-NsWien", NONE, "< Size of the search region in the reference frame (second pass)");
    const int NprWien = cli.i("-NprWien", NONE, "< Size of the search region in the other frames (second pass)");
    const int NbWien = cli.i("-NbWien", NONE, "< Maximum number of neighbors per frame (second pass)");
    const int pWien = cli.i("-pWien", NONE, "< Step between each patch (second pass)");
    const int NWien = cli.i("-NWien", NONE, "< Maximum number of neighbors (second pass)");
    const int dWien = cli.i("-dWien", NONE, "< Bias toward center patches (second pass)");
    const float tauWien = (float)cli.i("-tauWien", NONE, "< Distance threshold on neighbors (second pass)");
    const unsigned T_2D_wien = (unsigned)cli.i("-T2dw", NONE, "< 2D transform (second pass), choice is 0 (dct) or 1 (bior)");
    const unsigned T_3D_wien = (unsigned)cli.i("-T3dw", NONE, "< 1D transform (first pass), choice is 2 (hadamard) or 3 (haar)");

    const bool color_space = cli.b("-color", true, "< Transform the color space during the processing");
    const bool mc = cli.b("-mc", false, "< Motion compensation for 3d patches");

    //! Check inputs (src/main.cpp:298-323)
    if (input_path == "")
        return std::fprintf(stderr, "%s: no input images.\nTry `%s --help' for more information.\n", argv[0], argv[0]),
               EXIT_FAILURE;
    if (T_2D_hard != (unsigned)NONE && T_2D_hard != DCT && T_2D_hard != BIOR)
        return std::fprintf(stderr, "%s: unknown T_2d_hard.Try `%s --help' for the available choices.\n", argv[0], argv[0]),
               EXIT_FAILURE;
    if (T_2D_wien != (unsigned)NONE && T_2D_wien != DCT && T_2D_wien != BIOR)
        return std::fprintf(stderr, "%s: unknown T_2d_wien.Try `%s --help' for the available choices.\n", argv[0], argv[0]),
               EXIT_FAILURE;
    if (T_3D_hard != (unsigned)NONE && T_3D_hard != HAAR && T_3D_hard != HADAMARD)
        return std::fprintf(stderr, "%s: unknown T_3d_hard.Try `%s --help' for the available choices.\n", argv[0], argv[0]),
               EXIT_FAILURE;
    if (T_3D_wien != (unsigned)NONE && T_3D_wien != HAAR && T_3D_wien != HADAMARD)
        return std::fprintf(stderr, "%s: unknown T_3d_wien.Try `%s --help' for the available choices.\n", argv[0], argv[0]),
               EXIT_FAILURE;

    //! Init parameters (src/main.cpp:325-335)
    vbm3d_params c1, c2;
    vbm3d_default_params_1(&c1, kHard, ktHard, NfHard, NsHard, NprHard, NbHard, pHard, NHard, dHard,
                           tauHard, lambda3D, T_2D_hard, T_3D_hard, fSigma);
    vbm3d_default_params_2(&c2, kWien, ktWien, NfWien, NsWien, NprWien, NbWien, pWien, NWien, dWien,
                           tauWien, T_2D_wien, T_3D_wien, fSigma);
    Parameters prms_1 = from_c(c1), prms_2 = from_c(c2);
    prms_1.mc = mc;
    prms_2.mc = mc;
    if (verbose) {
        print_parameters(prms_1, "hard thresholding");
        print_parameters(prms_2, "Wiener filtering");
    }

    Video<float> vid, vid_noisy, vid_basic, vid_denoised, vid_diff, fflow, bflow;
    load_video(vid, input_path, firstFrame, lastFrame, frameStep);
    if (kHard == 0) load_video(vid_basic, inbsc_path, firstFrame, lastFrame, frameStep);
    if (fflow_path == "") fflow.resize(vid.sz.width, vid.sz.height, vid.sz.frames, 2);
    else load_flows(fflow, fflow_path, firstFrame, lastFrame - 1, frameStep);
    if (bflow_path == "") bflow.resize(vid.sz.width, vid.sz.height, vid.sz.frames, 2);
    else load_flows(bflow, bflow_path, firstFrame + 1, lastFrame, frameStep);
    if (fflow.sz.width != bflow.sz.width || fflow.sz.width != vid.sz.width ||
        fflow.sz.height != bflow.sz.height || fflow.sz.height != vid.sz.height)
        return std::fprintf(stderr, "%s: incompatible flows and/or frame sizes.\nTry `%s --help' for more information.\n",
                            argv[0], argv[0]),
               EXIT_FAILURE;

    vid_noisy.resize(vid.sz);
    vid_diff.resize(vid.sz);
    if (addnoise) add_noise(vid, vid_noisy, fSigma, verbose);
    else {
        std::printf("Noisy input video with sigma = %f\n", fSigma);
        vid_noisy = vid;
    }

    if (run_vbm3d(fSigma, vid_noisy, fflow, bflow, vid_basic, vid_denoised, prms_1, prms_2, color_space) != EXIT_SUCCESS)
        return EXIT_FAILURE;

    if (addnoise) {
        double psnr, rmse, psnr_basic, rmse_basic;
        compute_psnr(vid, vid_basic, psnr_basic, rmse_basic);
        compute_psnr(vid, vid_denoised, psnr, rmse);
        if (verbose) {
            std::printf("Basic estimate: PSNR= %5.2f RMSE= %6.3f\n", psnr_basic, rmse_basic);
            std::printf("Final estimate: PSNR= %5.2f RMSE= %6.3f\n", psnr, rmse);
        }
        FILE* file = std::fopen(meas_path.c_str(), "w");
        if (file) {
            std::fprintf(file, "-sigma = %f\n-PSNR_basic = %f\n-RSME_basic = %f\n-PSNR = %f\n-RMSE = %f\n",
                         fSigma, psnr_basic, rmse_basic, psnr, rmse);
            std::fclose(file);
        } else {
            std::fprintf(stderr, "Can't open %s\n", meas_path.c_str());
        }
        if (diff_path != "") compute_diff(vid_noisy, vid_denoised, vid_diff);
        if (noisy_path != "") save_video(vid_noisy, noisy_path, firstFrame, frameStep);
        if (diff_path != "") save_video(vid_diff, diff_path, firstFrame, frameStep);
    }
    if (diffi_path != "") {
        compute_diff(vid, vid_denoised, vid_diff);
        save_video(vid_diff, diffi_path, firstFrame, frameStep);
    }
    if (verbose) std::printf("Save output videos...\n");
    if (basic_path != "") save_video(vid_basic, basic_path, firstFrame, frameStep);
    if (prms_2.k) save_video(vid_denoised, final_path, firstFrame, frameStep);
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        // the reference lets std::runtime_error from file I/O escape main
        std::fprintf(stderr, "terminate called after throwing an instance of 'std::runtime_error'\n  what():  %s\n",
                     e.what());
        return 134;
    }
}
