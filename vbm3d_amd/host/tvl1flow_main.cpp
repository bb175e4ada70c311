// tvl1flow_main.cpp -- the TV-L1 flow command line on the GPU, a drop-in
// for the reference's tvl1flow program (tvl1flow/main.c): the same
// positional arguments, defaults, parameter resets and level cap; the
// flow is written as .flo (Middlebury) or, for other extensions, as a
// two-channel image (u, v).  nproc is accepted for compatibility and
// unused: the flow runs on GPU 0.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "image_io.h"
#include "vbm3d_hip.h"

namespace {

std::string ext_lower(const std::string& p) {
    const size_t d = p.find_last_of('.');
    std::string e = d == std::string::npos ? "" : p.substr(d + 1);
    for (char& ch : e) ch = (char)std::tolower((unsigned char)ch);
    return e;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr,
                     "Usage: %s I0 I1 [out nproc tau lambda theta nscales fscale zfactor nwarps epsilon verbose]\n",
                     argv[0]);
        return EXIT_FAILURE;
    }
    // positional arguments and defaults (tvl1flow/main.c:20-30, 70-84)
    vbm3d_tvl1_params p;
    vbm3d_tvl1_default_params(&p);
    int i = 1;
    const std::string in0 = argv[i++], in1 = argv[i++];
    const std::string out = argc > i ? argv[i] : "flow.flo";
    i++;
    int nproc = argc > i ? std::atoi(argv[i]) : 4;
    i++;
    if (argc > i) p.tau = (float)std::atof(argv[i]);
    i++;
    if (argc > i) p.lambda = (float)std::atof(argv[i]);
    i++;
    if (argc > i) p.theta = (float)std::atof(argv[i]);
    i++;
    if (argc > i) p.nscales = std::atoi(argv[i]);
    i++;
    if (argc > i) p.fscale = std::atoi(argv[i]);
    i++;
    if (argc > i) p.zfactor = (float)std::atof(argv[i]);
    i++;
    if (argc > i) p.nwarps = std::atoi(argv[i]);
    i++;
    if (argc > i) p.epsilon = (float)std::atof(argv[i]);
    i++;
    const int verbose = argc > i ? std::atoi(argv[i]) : 0;
    // out-of-range values fall back to the defaults (main.c:86-124)
    vbm3d_tvl1_params d;
    vbm3d_tvl1_default_params(&d);
    if (nproc < 0) {
        nproc = 4;
        if (verbose) std::fprintf(stderr, "warning: nproc changed to %d\n", nproc);
    }
    if (p.tau <= 0 || p.tau > 0.25) {
        p.tau = d.tau;
        if (verbose) std::fprintf(stderr, "warning: tau changed to %g\n", p.tau);
    }
    if (p.lambda <= 0) {
        p.lambda = d.lambda;
        if (verbose) std::fprintf(stderr, "warning: lambda changed to %g\n", p.lambda);
    }
    if (p.theta <= 0) {
        p.theta = d.theta;
        if (verbose) std::fprintf(stderr, "warning: theta changed to %g\n", p.theta);
    }
    if (p.nscales <= 0) {
        p.nscales = d.nscales;
        if (verbose) std::fprintf(stderr, "warning: nscales changed to %d\n", p.nscales);
    }
    if (p.zfactor <= 0 || p.zfactor >= 1) {
        p.zfactor = d.zfactor;
        if (verbose) std::fprintf(stderr, "warning: zfactor changed to %g\n", p.zfactor);
    }
    if (p.nwarps <= 0) {
        p.nwarps = d.nwarps;
        if (verbose) std::fprintf(stderr, "warning: nwarps changed to %d\n", p.nwarps);
    }
    if (p.epsilon <= 0) {
        p.epsilon = d.epsilon;
        if (verbose) std::fprintf(stderr, "warning: epsilon changed to %f\n", p.epsilon);
    }
    std::vector<float> I0, I1;
    unsigned nx, ny, nx2, ny2;
    std::string err;
    if (!load_grey(in0, I0, nx, ny, err) || !load_grey(in1, I1, nx2, ny2, err)) {
        std::fprintf(stderr, "ERROR: could not read image: %s\n", err.c_str());
        return EXIT_FAILURE;
    }
    if (nx != nx2 || ny != ny2) {
        std::fprintf(stderr, "ERROR: input images size mismatch %ux%u != %ux%u\n", nx, ny, nx2, ny2);
        return EXIT_FAILURE;
    }
    // level cap (main.c:139-146)
    p.nscales = vbm3d_tvl1_nscales((int)nx, (int)ny, p.zfactor, p.nscales);
    if (p.nscales < p.fscale) p.fscale = p.nscales;
    if (verbose)
        std::fprintf(stderr, "nproc=%d tau=%f lambda=%f theta=%f nscales=%d zfactor=%f nwarps=%d epsilon=%g\n",
                     nproc, p.tau, p.lambda, p.theta, p.nscales, p.zfactor, p.nwarps, p.epsilon);
    vbm3d_ctx* ctx = vbm3d_create(0);
    if (!ctx) {
        std::fprintf(stderr, "ERROR: %s\n", vbm3d_last_error());
        return EXIT_FAILURE;
    }
    std::vector<float> flow((size_t)2 * nx * ny);
    std::vector<int> iters((size_t)std::max(p.nscales - p.fscale, 1) * p.nwarps);
    const int rc = vbm3d_tvl1_flow(ctx, I0.data(), I1.data(), flow.data(), flow.data() + (size_t)nx * ny, (int)nx,
                                   (int)ny, &p, iters.data());
    if (rc != VBM3D_OK) {
        std::fprintf(stderr, "ERROR: %s\n", vbm3d_last_error());
        vbm3d_destroy(ctx);
        return EXIT_FAILURE;
    }
    if (verbose)
        for (int l = 0; l < p.nscales - p.fscale; l++)
            for (int w = 0; w < p.nwarps; w++)
                std::fprintf(stderr, "Level %d Warping: %d, Iterations: %d\n", p.nscales - 1 - l, w,
                             iters[(size_t)l * p.nwarps + w]);
    vbm3d_destroy(ctx);
    const bool ok = ext_lower(out) == "flo" ? save_flow(out, flow, nx, ny, err) : save_frame(out, flow, nx, ny, 2, err);
    if (!ok) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
