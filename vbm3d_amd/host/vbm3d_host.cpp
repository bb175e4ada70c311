/*
 * vbm3d_host.cpp -- run_vbm3d (src/vbm3d.cpp:87-376) over the C ABI of
 * libvbm3d_hip.so.  Host buffers in, host buffers out; one device context
 * per calling thread, created on first use and kept for later calls.
 */
#include "vbm3d.h"

#include <cstdio>
#include <cstdlib>
#include <memory>

#include "../../include/vbm3d_hip.h"

namespace {

int g_nb = -1;
int g_device = 0;

struct CtxHolder {
    vbm3d_ctx* ctx = nullptr;
    int device = -1;
    ~CtxHolder() { if (ctx) vbm3d_destroy(ctx); }
};
thread_local CtxHolder t_ctx;

vbm3d_params to_c(const Parameters& p)
{
    vbm3d_params o;
    o.k = p.k; o.kt = p.kt; o.p = p.p; o.N = p.N; o.tau = p.tau; o.d = p.d;
    o.Nf = p.Nf; o.Ns = p.Ns; o.Npr = p.Npr; o.Nb = p.Nb; o.T_2D = p.T_2D; o.T_3D = p.T_3D;
    o.lambda3D = p.lambda3D; o.mc = p.mc ? 1 : 0; o.n = p.n;
    return o;
}

// A flow video that is identically zero is passed as NULL (the kernels
// then skip the flow reads); a non-zero one must cover F-1 frames.
const float* flow_ptr(const Video<float>& f, const VideoSize& sz, bool& bad)
{
    if (f.data.empty()) return nullptr;
    bool any = false;
    for (float v : f.data) if (v != 0.f) { any = true; break; }
    if (!any) return nullptr;
    if (f.sz.width != sz.width || f.sz.height != sz.height || f.sz.channels != 2 ||
        f.sz.frames + 1 < sz.frames) { bad = true; return nullptr; }
    return f.data.data();
}

}  // namespace

void vbm3d_set_ref_grid_nb(int nb) { g_nb = nb; }

int vbm3d_get_ref_grid_nb()
{
    if (g_nb < 0) {
        const char* e = std::getenv("VBM3D_REF_GRID_NB");
        g_nb = e ? std::atoi(e) : 1;
    }
    return g_nb;
}

void vbm3d_set_device(int device) { g_device = device; }

int run_vbm3d(
    const float sigma
,   Video<float> &vid_noisy
,   Video<float> &fflow
,   Video<float> &bflow
,   Video<float> &vid_basic
,   Video<float> &vid_denoised
,   const Parameters& prms_1
,   const Parameters& prms_2
,   const bool color_space
){
    //! outputs take the noisy video's size (src/vbm3d.cpp:99-102)
    if (vid_basic.sz != vid_noisy.sz) vid_basic.resize(vid_noisy.sz);
    if (vid_denoised.sz != vid_noisy.sz) vid_denoised.resize(vid_noisy.sz);

    if (!t_ctx.ctx || t_ctx.device != g_device) {
        if (t_ctx.ctx) vbm3d_destroy(t_ctx.ctx);
        t_ctx.ctx = vbm3d_create(g_device);
        t_ctx.device = g_device;
        if (!t_ctx.ctx) {
            std::fprintf(stderr, "run_vbm3d: %s\n", vbm3d_last_error());
            return EXIT_FAILURE;
        }
    }
    bool bad = false;
    const float* ff = flow_ptr(fflow, vid_noisy.sz, bad);
    const float* bf = flow_ptr(bflow, vid_noisy.sz, bad);
    if (bad) {
        std::fprintf(stderr, "run_vbm3d: flow videos must be W x H x (F-1) x 2\n");
        return EXIT_FAILURE;
    }
    const vbm3d_params p1 = to_c(prms_1), p2 = to_c(prms_2);
    const VideoSize& s = vid_noisy.sz;
    const int rc = vbm3d_run(t_ctx.ctx, sigma, vid_noisy.data.data(), ff, bf, s.width, s.height,
                             s.frames, s.channels, &p1, &p2, color_space ? 1 : 0,
                             vbm3d_get_ref_grid_nb(), vid_basic.data.data(), vid_denoised.data.data());
    if (rc != VBM3D_OK) {
        std::fprintf(stderr, "run_vbm3d: %s\n", vbm3d_last_error());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
