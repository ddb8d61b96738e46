// twdic_api.cpp -- the extern "C" boundary (include/twdic.h): context
// lifetime, image upload, precompute orchestration and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>

#include "twdic_internal.h"

using namespace twdic;

namespace {

twdic_status fail(twdic_ctx* c, twdic_status st, const char* what)
{
    if (c) c->err = what;
    return st;
}

twdic_status hip_fail(twdic_ctx* c, hipError_t e, const char* where)
{
    if (c) {
        c->err = where;
        c->err += ": ";
        c->err += hipGetErrorString(e);
    }
    return TWDIC_EHIP;
}

#define TW_HIP(ctx, expr)                                         \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr);  \
    } while (0)

// Bytes from p to the end of the allocation that holds it; false when the
// runtime knows no allocation at p (the failed query's error is cleared so a
// later launch check does not report it).
bool device_span(const void* p, size_t* bytes)
{
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return false;
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
    if (q < b || q - b > size) return false;
    *bytes = size - (q - b);
    return true;
}

struct Need {
    const void* p;        // null: not given (optional output), not checked
    size_t bytes;
    const char* name;
};

// Every device buffer a call names must lie in an allocation that holds the
// bytes the call touches; checked before anything is launched.
twdic_status check_spans(twdic_ctx* c, const char* call, std::initializer_list<Need> needs)
{
    for (const Need& k : needs) {
        if (!k.p || !k.bytes) continue;
        size_t have = 0;
        if (!device_span(k.p, &have)) {
            if (c) c->err = std::string(call) + ": " + k.name + " is not in a device allocation";
            return TWDIC_EINVAL;
        }
        if (have < k.bytes) {
            if (c)
                c->err = std::string(call) + ": " + k.name + " holds " + std::to_string(have) + " bytes, the call needs " +
                         std::to_string(k.bytes);
            return TWDIC_EBOUNDS;
        }
    }
    return TWDIC_OK;
}

#define TW_SPANS(ctx, call, ...)                                           \
    do {                                                                   \
        twdic_status s_ = check_spans((ctx), (call), {__VA_ARGS__});       \
        if (s_ != TWDIC_OK) return s_;                                     \
    } while (0)

// Fftcc2D::recomputeNumPOI (TW_paDIC_FFTCC2D.cpp:88-94): float division,
// floor, truncation (a negative count, as the reference computes it, marks
// an invalid geometry).  A grid spacing below 1, which the reference would
// divide by, and a count outside int give 0: no valid grid either.
void num_poi(const twdic_config& c, int* nx, int* ny)
{
    auto count = [](int roi, int s, int m, int g) {
        if (g < 1) return 0;
        const float q = std::floor((float)((int64_t)roi - 2 * (int64_t)s - 2 * (int64_t)m) / (float)g) + 1.f;
        return (q > -2147483648.f && q < 2147483647.f) ? (int)q : 0;
    };
    *nx = count(c.roi_w, c.sx, c.mx, c.gx);
    *ny = count(c.roi_h, c.sy, c.my, c.gy);
}

const char* validate(const twdic_config& c)
{
    if (c.img_w <= 0 || c.img_h <= 0) return "image size must be positive";
    // kernels address images with 32-bit byte offsets and 24-bit row products
    if (c.img_w > 65535 || c.img_h > 65535 || (int64_t)c.img_w * c.img_h >= (int64_t)1 << 31)
        return "images are limited to 65535 px per side and 2^31 bytes";
    if (c.roi_w < 2 || c.roi_h < 2) return "ROI must be at least 2x2";
    if (c.roi_x < 0 || c.roi_y < 0 || c.roi_x + c.roi_w > c.img_w || c.roi_y + c.roi_h > c.img_h)
        return "ROI must lie inside the image";
    if (c.sx < 1 || c.sy < 1 || c.sx > 60 || c.sy > 60) return "subset half size must be in [1, 60]";
    if (c.gx < 1 || c.gy < 1) return "grid spacing must be >= 1";
    if (c.mx < 0 || c.my < 0) return "margins must be >= 0";
    if (c.fft_n < 2 || c.fft_n > kMaxFftN || (c.fft_n & 1)) return "fft_n must be even and in [2, 64]";
    if (c.max_iter < 1) return "max_iter must be >= 1";
    if (!(c.delta_p >= 0.f)) return "delta_p must be >= 0";
    if (c.interp != TWDIC_INTERP_LUT && c.interp != TWDIC_INTERP_FUSED && c.interp != TWDIC_INTERP_BSPLINE)
        return "unknown interp";
    if (c.interp == TWDIC_INTERP_BSPLINE &&
        (c.roi_x < 1 || c.roi_y < 1 || c.roi_x + c.roi_w + 2 > c.img_w || c.roi_y + c.roi_h + 2 > c.img_h))
        return "the B-spline LUT needs a 1-pixel frame left/above and 2 pixels right/below the ROI";
    int nx, ny;
    num_poi(c, &nx, &ny);
    if (nx <= 0 || ny <= 0) return "Number of POIs is below 0!";
    return nullptr;
}

twdic_status upload(twdic_ctx* c, uint8_t* dst, const uint8_t* img, int stride, int on_device)
{
    const twdic_config& g = c->cfg;
    if (!img || stride < g.img_w) return fail(c, TWDIC_EINVAL, "image pointer/stride invalid");
    if (on_device) {
        TW_SPANS(c, "image upload", {img, (size_t)stride * (g.img_h - 1) + g.img_w, "device image"});
        TW_HIP(c, hipMemcpy2DAsync(dst, c->dstride, img, stride, g.img_w, g.img_h, hipMemcpyDeviceToDevice, c->stream));
        return TWDIC_OK;
    }
    // Host image: copy into pinned staging, then an async DMA.  The staging
    // buffer is reused only after the previous DMA out of it has finished.
    size_t bytes = (size_t)g.img_w * g.img_h;
    if (c->h_stage_bytes < bytes) {
        if (c->h_stage) {
            TW_HIP(c, hipStreamSynchronize(c->stream));
            (void)hipHostFree(c->h_stage);
            c->h_stage = nullptr;
        }
        TW_HIP(c, hipHostMalloc((void**)&c->h_stage, bytes, hipHostMallocDefault));
        c->h_stage_bytes = bytes;
    } else {
        TW_HIP(c, hipEventSynchronize(c->ev[15]));
    }
    for (int y = 0; y < g.img_h; ++y) std::memcpy(c->h_stage + (size_t)y * g.img_w, img + (size_t)y * stride, g.img_w);
    TW_HIP(c, hipMemcpy2DAsync(dst, c->dstride, c->h_stage, g.img_w, g.img_w, g.img_h, hipMemcpyHostToDevice, c->stream));
    TW_HIP(c, hipEventRecord(c->ev[15], c->stream));
    return TWDIC_OK;
}

// The gradients need a one-pixel frame around the ROI (TW_utils.cpp:246-257,
// with the reference's off-by-one corrected); the reference throws there.
bool gradient_frame_ok(const twdic_config& g)
{
    return g.roi_x >= 1 && g.roi_y >= 1 && g.roi_x + g.roi_w + 1 <= g.img_w && g.roi_y + g.roi_h + 1 <= g.img_h;
}

twdic_status ensure_ref_grads(twdic_ctx* c)
{
    if (!c->ref_loaded) return fail(c, TWDIC_EINVAL, "reference image not set");
    if (!gradient_frame_ok(c->cfg)) return fail(c, TWDIC_EINVAL, "Error: Not enough boundary pixels for gradients!");
    if (c->ref_grads_valid) return TWDIC_OK;
    const twdic_config& g = c->cfg;
    TW_HIP(c, launch_gradients(c->d_ref, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_rx, c->d_ry,
                               nullptr, c->d_g2, c->stream));
    c->ref_grads_valid = true;
    return TWDIC_OK;
}

IcgnParams icgn_params(twdic_ctx* c, const int32_t* poi, int n, float* u, float* v, float* p6,
                       int32_t* iters, uint8_t* flags, float* dpnorm)
{
    const twdic_config& g = c->cfg;
    IcgnParams p{};
    p.ref = c->d_ref; p.tar = c->d_tar; p.stride = c->dstride;
    p.sx0 = g.roi_x; p.sy0 = g.roi_y; p.rw = g.roi_w; p.rh = g.roi_h;
    p.Sx = g.sx; p.Sy = g.sy; p.Wsub = 2 * g.sx + 1; p.Hsub = 2 * g.sy + 1; p.s = p.Wsub * p.Hsub;
    p.max_iter = g.max_iter; p.delta_p = g.delta_p;
    p.g2 = c->d_g2; p.lut = c->d_lut;
    p.rp = g.interp == TWDIC_INTERP_FUSED ? c->d_rp : nullptr;    // null when the plane does not fit
    p.refstate = c->d_refstate;
    p.poi = poi; p.n = n;
    p.lanes = icgn_lanes_for(n);
    p.u = u; p.v = v; p.p6 = p6; p.iters = iters; p.flags = flags; p.dpnorm = dpnorm;
    return p;
}

}  // namespace

namespace {

// The per-target interpolation data of the context's mode, on the context
// stream: the LUT (LUT, B-spline modes) or the row polynomials (fused mode).
// Built once per target; every call rebuilds it once the target buffer has
// been handed out, since the caller may then write it directly.
hipError_t launch_target_interp(twdic_ctx* c, hipStream_t st)
{
    const twdic_config& g = c->cfg;
    if (g.interp == TWDIC_INTERP_LUT)
        // cuBicubicCoefficients, TW_paDIC_cuICGN2D.cu:844-851
        return launch_bicubic_lut(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_lut, st);
    if (g.interp == TWDIC_INTERP_BSPLINE)
        // BicubicSplineCoefficients_s, TW_paDIC_ICGN2D_CPU.cpp:333-340
        return launch_bspline_lut(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_lut, st);
    if (c->d_rp) return launch_row_poly(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_rp, st);
    return hipSuccess;
}

twdic_status target_interp_prepare(twdic_ctx* c)
{
    bool& valid = c->cfg.interp == TWDIC_INTERP_FUSED ? c->tar_rp_valid : c->tar_pre_valid;
    if (valid && !c->tar_exposed) return TWDIC_OK;
    TW_HIP(c, launch_target_interp(c, c->stream));
    valid = true;
    return TWDIC_OK;
}

// Everything IC-GN needs before its kernel: reference gradients, the target
// LUT in the LUT/B-spline modes or the row polynomials in the fused mode, and
// the per-POI reference state, all on the context stream.
twdic_status icgn_prepare(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_p6,
                          int32_t* d_iters, uint8_t* d_flags, float* d_dpnorm, IcgnParams& p)
{
    twdic_status st = ensure_ref_grads(c);
    if (st != TWDIC_OK) return st;
    st = target_interp_prepare(c);
    if (st != TWDIC_OK) return st;
    if ((size_t)n > c->refstate_cap) {
        if (c->d_refstate) {
            TW_HIP(c, hipStreamSynchronize(c->stream));
            TW_HIP(c, hipStreamSynchronize(c->stream2));
            TW_HIP(c, hipFree(c->d_refstate));
            c->d_refstate = nullptr;
        }
        TW_HIP(c, hipMalloc((void**)&c->d_refstate, sizeof(double) * kRefStateStride * (size_t)n));
        c->refstate_cap = (size_t)n;
        c->refstate_version = ~0ull;
    }
    p = icgn_params(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, d_dpnorm);
    if (c->refstate_version != c->ref_version || c->refstate_poi != d_poi || c->refstate_n != n) {
        // per-reference state (RefAllSubetsNorm_Kernel + InverseHessian_Kernel,
        // TW_paDIC_cuICGN2D.cu:883-903), rebuilt when the reference or the POI set changed
        TW_HIP(c, launch_ref_state(p, c->stream));
        c->refstate_version = c->ref_version;
        c->refstate_poi = d_poi;
        c->refstate_n = n;
    }
    return TWDIC_OK;
}

// The parameter blocks of POIs [base, base + cnt) of a set.
FftccParams fftcc_chunk(FftccParams p, int base, int cnt)
{
    p.poi += 2 * (size_t)base;
    p.n = cnt;
    p.u += base; p.v += base; p.zncc += base;
    if (p.status) p.status += base;
    return p;
}

IcgnParams icgn_chunk(IcgnParams p, int base, int cnt)
{
    p.poi += 2 * (size_t)base;
    p.n = cnt;
    p.refstate += (size_t)base * kRefStateStride;
    p.u += base; p.v += base;
    if (p.p6) p.p6 += 6 * (size_t)base;
    if (p.iters) p.iters += base;
    if (p.flags) p.flags += base;
    if (p.dpnorm) p.dpnorm += base;
    return p;
}

FftccParams fftcc_params(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_zncc,
                         uint8_t* d_status)
{
    FftccParams p{};
    p.ref = c->d_ref; p.tar = c->d_tar; p.stride = c->dstride;
    p.img_w = c->cfg.img_w; p.img_h = c->cfg.img_h;
    p.poi = d_poi; p.n = n; p.N = c->cfg.fft_n; p.tw = c->d_fft_tw;
    p.tol = 1e-4f;
    p.u = d_u; p.v = d_v; p.zncc = d_zncc; p.status = d_status;
    return p;
}

void drop_graph(twdic_ctx* c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

}  // namespace

extern "C" {

twdic_status twdic_device_span(const void* d_ptr, size_t* bytes)
{
    if (!d_ptr || !bytes) return TWDIC_EINVAL;
    return device_span(d_ptr, bytes) ? TWDIC_OK : TWDIC_EINVAL;
}

twdic_status twdic_num_poi(const twdic_config* cfg, int32_t* nx, int32_t* ny)
{
    if (!cfg || !nx || !ny) return TWDIC_EINVAL;
    int a, b;
    num_poi(*cfg, &a, &b);
    *nx = a; *ny = b;
    return (a > 0 && b > 0) ? TWDIC_OK : TWDIC_EINVAL;
}

twdic_status twdic_create(const twdic_config* cfg, int device, twdic_ctx** out)
{
    if (!cfg || !out) return TWDIC_EINVAL;
    *out = nullptr;
    if (validate(*cfg)) return TWDIC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) return TWDIC_ENODEV;
    twdic_ctx* c = new (std::nothrow) twdic_ctx();
    if (!c) return TWDIC_ENOMEM;
    c->cfg = *cfg;
    c->device = device;
    num_poi(*cfg, &c->nx, &c->ny);
    c->dstride = cfg->img_w;
    if (const char* v = std::getenv("TWDIC_FFTCC_VARIANT")) c->fftcc_variant = std::atoi(v);
    if (const char* v = std::getenv("TWDIC_ICGN_VARIANT")) c->icgn_variant = std::atoi(v);
    if (const char* v = std::getenv("TWDIC_FRAME_CHUNKS"))
        c->frame_chunks = std::max(1, std::min(twdic_ctx::kMaxFrameChunks, std::atoi(v)));
    twdic_status st = TWDIC_OK;
    auto bail = [&](hipError_t e, const char* w) {
        hip_fail(c, e, w);
        st = (e == hipErrorOutOfMemory) ? TWDIC_ENOMEM : TWDIC_EHIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) bail(e, "hipSetDevice");
    if (st == TWDIC_OK && (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) bail(e, "hipStreamCreate");
    if (st == TWDIC_OK && (e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess) bail(e, "hipStreamCreate");
    if (st == TWDIC_OK && (e = dft_tables_init(device)) != hipSuccess) bail(e, "dft_tables_init");
    const size_t img = (size_t)c->dstride * cfg->img_h;
    const size_t roi = (size_t)cfg->roi_w * cfg->roi_h;
    struct { void** p; size_t b; } allocs[] = {
        {(void**)&c->d_ref, img}, {(void**)&c->d_tar, img},
        {(void**)&c->d_rx, roi * 4}, {(void**)&c->d_ry, roi * 4},
        {(void**)&c->d_tx, roi * 4}, {(void**)&c->d_ty, roi * 4}, {(void**)&c->d_txy, roi * 4},
        {(void**)&c->d_lut, roi * 64}, {(void**)&c->d_fft_tw, sizeof(float2) * kMaxFftN},
        {(void**)&c->d_g2, roi * 4},
    };
    for (auto& a : allocs) {
        if (st != TWDIC_OK) break;
        if ((e = hipMalloc(a.p, a.b)) != hipSuccess) bail(e, "hipMalloc");
        else if ((e = hipMemset(*a.p, 0, a.b)) != hipSuccess) bail(e, "hipMemset");
    }
    // fused mode: the target's row polynomials (16 bytes per ROI pixel);
    // without them (ROIs past 2^28 pixels) IC-GN reads the image bytes
    if (st == TWDIC_OK && cfg->interp == TWDIC_INTERP_FUSED &&
        row_poly_fits(cfg->roi_x, cfg->roi_y, cfg->roi_w, cfg->roi_h)) {
        if ((e = hipMalloc((void**)&c->d_rp, row_poly_bytes(cfg->roi_w, cfg->roi_h))) != hipSuccess)
            bail(e, "hipMalloc");
    }
    for (int i = 0; i < 16 && st == TWDIC_OK; ++i)
        if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) bail(e, "hipEventCreate");
    for (int i = 0; i <= twdic_ctx::kMaxFrameChunks && st == TWDIC_OK; ++i)
        if ((e = hipEventCreateWithFlags(&c->pev[i], hipEventDisableTiming)) != hipSuccess) bail(e, "hipEventCreate");
    if (st == TWDIC_OK) {
        // twiddles exp(-2 pi i e / N), rounded from double
        float2 tw[kMaxFftN];
        const int N = cfg->fft_n;
        for (int k = 0; k < N; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)N;
            tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        if ((e = hipMemcpy(c->d_fft_tw, tw, sizeof(float2) * N, hipMemcpyHostToDevice)) != hipSuccess)
            bail(e, "hipMemcpy twiddles");
    }
    if (st != TWDIC_OK) {
        twdic_destroy(c);
        return st;
    }
    *out = c;
    return TWDIC_OK;
}

twdic_status twdic_destroy(twdic_ctx* c)
{
    if (!c) return TWDIC_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    drop_graph(c);
    void* bufs[] = {c->d_ref, c->d_tar, c->d_rx, c->d_ry, c->d_tx, c->d_ty, c->d_txy, c->d_lut,
                    c->d_fft_tw, c->d_refstate, c->d_g2, c->d_rp};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : c->pev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TWDIC_OK;
}

const char* twdic_last_error(const twdic_ctx* c)
{
    return c ? c->err.c_str() : "null context";
}

twdic_status twdic_sync(twdic_ctx* c)
{
    if (!c) return TWDIC_EINVAL;
    TW_HIP(c, hipStreamSynchronize(c->stream));
    return TWDIC_OK;
}

void* twdic_stream(twdic_ctx* c)
{
    return c ? (void*)c->stream : nullptr;
}

twdic_status twdic_set_reference(twdic_ctx* c, const uint8_t* img, int stride, int on_device)
{
    if (!c) return TWDIC_EINVAL;
    twdic_status st = upload(c, c->d_ref, img, stride, on_device);
    if (st != TWDIC_OK) return st;
    c->ref_loaded = true;
    c->ref_grads_valid = false;
    ++c->ref_version;
    return TWDIC_OK;
}

twdic_status twdic_set_target(twdic_ctx* c, const uint8_t* img, int stride, int on_device)
{
    if (!c) return TWDIC_EINVAL;
    twdic_status st = upload(c, c->d_tar, img, stride, on_device);
    if (st != TWDIC_OK) return st;
    c->tar_loaded = true;
    c->tar_pre_valid = false;
    c->tar_rp_valid = false;
    return TWDIC_OK;
}

twdic_status twdic_poi_grid(twdic_ctx* c, int32_t* d_poi)
{
    if (!c || !d_poi) return TWDIC_EINVAL;
    const twdic_config& g = c->cfg;
    TW_SPANS(c, "twdic_poi_grid", {d_poi, 8 * (size_t)c->nx * c->ny, "d_poi"});
    TW_HIP(c, launch_poi_grid(d_poi, g.roi_x, g.roi_y, c->nx, c->ny, g.mx, g.my, g.sx, g.sy, g.gx, g.gy, c->stream));
    return TWDIC_OK;
}

twdic_status twdic_fftcc(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_zncc,
                         uint8_t* d_status)
{
    if (!c || n < 0 || (n > 0 && (!d_poi || !d_u || !d_v || !d_zncc))) return fail(c, TWDIC_EINVAL, "twdic_fftcc: bad arguments");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_fftcc: images not set");
    const size_t m = (size_t)n;
    TW_SPANS(c, "twdic_fftcc", {d_poi, 8 * m, "d_poi"}, {d_u, 4 * m, "d_u"}, {d_v, 4 * m, "d_v"},
             {d_zncc, 4 * m, "d_zncc"}, {d_status, m, "d_status"});
    TW_HIP(c, launch_fftcc(fftcc_params(c, d_poi, n, d_u, d_v, d_zncc, d_status), c->fftcc_variant, c->stream));
    return TWDIC_OK;
}

twdic_status twdic_precompute_target(twdic_ctx* c)
{
    if (!c) return TWDIC_EINVAL;
    if (!c->tar_loaded) return fail(c, TWDIC_EINVAL, "target image not set");
    twdic_status st = ensure_ref_grads(c);
    if (st != TWDIC_OK) return st;
    const twdic_config& g = c->cfg;
    TW_HIP(c, launch_gradients(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_tx, c->d_ty,
                               c->d_txy, nullptr, c->stream));
    if (g.interp == TWDIC_INTERP_BSPLINE)
        TW_HIP(c, launch_bspline_lut(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_lut, c->stream));
    else
        TW_HIP(c, launch_bicubic_lut(c->d_tar, c->dstride, g.roi_x, g.roi_y, g.roi_w, g.roi_h, c->d_lut, c->stream));
    c->tar_pre_valid = true;
    return TWDIC_OK;
}

twdic_status twdic_icgn(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_p6,
                        int32_t* d_iters, uint8_t* d_flags, float* d_dpnorm)
{
    if (!c || n < 0 || (n > 0 && (!d_poi || !d_u || !d_v))) return fail(c, TWDIC_EINVAL, "twdic_icgn: bad arguments");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_icgn: images not set");
    if (n == 0) return TWDIC_OK;
    const size_t m = (size_t)n;
    TW_SPANS(c, "twdic_icgn", {d_poi, 8 * m, "d_poi"}, {d_u, 4 * m, "d_u"}, {d_v, 4 * m, "d_v"},
             {d_p6, 24 * m, "d_p6"}, {d_iters, 4 * m, "d_iters"}, {d_flags, m, "d_flags"},
             {d_dpnorm, 4 * m, "d_dpnorm"});
    IcgnParams p;
    twdic_status st = icgn_prepare(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, d_dpnorm, p);
    if (st != TWDIC_OK) return st;
    TW_HIP(c, launch_icgn(p, c->cfg.interp, c->icgn_variant, c->stream));
    return TWDIC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// IC-GN without allocation.  The per-POI reference state (Cholesky factor,
// steepest-descent sums: kRefStateStride doubles per POI), which twdic_icgn
// keeps in a context buffer it grows with hipMalloc, lives here in one
// file-scope device array per GPU.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, once per device, and only that address reaches a
// kernel.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxPois = 65536;       // 65536 x 384 B = 24 MiB per GPU
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxPois * kRefStateStride];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the context's device (looked up on the
// first call there, with that device current).
twdic_status scratch_for(twdic_ctx* c, double** out)
{
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, TWDIC_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        TW_HIP(c, hipSetDevice(c->device));
        void* p = nullptr;
        TW_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(c, TWDIC_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[c->device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[c->device];
    return TWDIC_OK;
}

}  // namespace

extern "C" {

twdic_status twdic_icgn_noalloc(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_p6,
                                int32_t* d_iters, uint8_t* d_flags)
{
    if (!c || n < 0 || (n > 0 && (!d_poi || !d_u || !d_v)))
        return fail(c, TWDIC_EINVAL, "twdic_icgn_noalloc: bad arguments");
    if (n > kNoAllocMaxPois)
        return fail(c, TWDIC_EINVAL, "twdic_icgn_noalloc: at most 65536 POIs per call (24 MiB of per-GPU scratch)");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_icgn_noalloc: images not set");
    if (n == 0) return TWDIC_OK;
    const size_t m = (size_t)n;
    TW_SPANS(c, "twdic_icgn_noalloc", {d_poi, 8 * m, "d_poi"}, {d_u, 4 * m, "d_u"}, {d_v, 4 * m, "d_v"},
             {d_p6, 24 * m, "d_p6"}, {d_iters, 4 * m, "d_iters"}, {d_flags, m, "d_flags"});
    double* state = nullptr;
    twdic_status st = scratch_for(c, &state);
    if (st != TWDIC_OK) return st;
    st = ensure_ref_grads(c);
    if (st != TWDIC_OK) return st;
    st = target_interp_prepare(c);
    if (st != TWDIC_OK) return st;
    // the scratch holds whatever the last call on this GPU left: the state is
    // rebuilt for every POI of this call (each POI's wavefront writes every
    // field IC-GN reads), then the iterations read it, in stream order
    IcgnParams p = icgn_params(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, nullptr);
    p.refstate = state;
    TW_HIP(c, launch_ref_state(p, c->stream));
    TW_HIP(c, launch_icgn(p, c->cfg.interp, c->icgn_variant, c->stream));
    return TWDIC_OK;
}

twdic_status twdic_reset_poi_state(twdic_ctx* c)
{
    if (!c) return TWDIC_EINVAL;
    c->refstate_version = ~0ull;
    return TWDIC_OK;
}

twdic_status twdic_frame(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_zncc,
                         float* d_p6, int32_t* d_iters, uint8_t* d_flags)
{
    if (!c || n < 0 || (n > 0 && (!d_poi || !d_u || !d_v || !d_zncc))) return fail(c, TWDIC_EINVAL, "twdic_frame: bad arguments");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_frame: images not set");
    if (n == 0) return TWDIC_OK;
    const size_t m = (size_t)n;
    TW_SPANS(c, "twdic_frame", {d_poi, 8 * m, "d_poi"}, {d_u, 4 * m, "d_u"}, {d_v, 4 * m, "d_v"},
             {d_zncc, 4 * m, "d_zncc"}, {d_p6, 24 * m, "d_p6"}, {d_iters, 4 * m, "d_iters"}, {d_flags, m, "d_flags"});
    const int chunks = n >= 4096 * c->frame_chunks ? c->frame_chunks : 1;
    if (chunks == 1) {
        twdic_status st = twdic_fftcc(c, d_poi, n, d_u, d_v, d_zncc, nullptr);
        if (st != TWDIC_OK) return st;
        return twdic_icgn(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, nullptr);
    }
    // the per-reference state of the whole set first (context stream), then
    // chunk i: FFT-CC on the context stream, IC-GN on the second stream once
    // that chunk's integer guesses exist; FFT-CC of chunk i + 1 overlaps it.
    // Results are those of the two calls back to back: the same kernels over
    // the same POIs, each POI's work independent of the others.
    IcgnParams ip;
    twdic_status st = icgn_prepare(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, nullptr, ip);
    if (st != TWDIC_OK) return st;
    const FftccParams fp = fftcc_params(c, d_poi, n, d_u, d_v, d_zncc, nullptr);
    for (int i = 0; i < chunks; ++i) {
        const int b0 = (int)((long long)n * i / chunks), b1 = (int)((long long)n * (i + 1) / chunks);
        TW_HIP(c, launch_fftcc(fftcc_chunk(fp, b0, b1 - b0), c->fftcc_variant, c->stream));
        TW_HIP(c, hipEventRecord(c->pev[i], c->stream));
        TW_HIP(c, hipStreamWaitEvent(c->stream2, c->pev[i], 0));
        TW_HIP(c, launch_icgn(icgn_chunk(ip, b0, b1 - b0), c->cfg.interp, c->icgn_variant, c->stream2));
    }
    // later work on the context stream (the next target, a sync) waits for IC-GN
    TW_HIP(c, hipEventRecord(c->pev[chunks], c->stream2));
    TW_HIP(c, hipStreamWaitEvent(c->stream, c->pev[chunks], 0));
    return TWDIC_OK;
}

twdic_status twdic_frame_capture(twdic_ctx* c, const int32_t* d_poi, int n, float* d_u, float* d_v, float* d_zncc,
                                 float* d_p6, int32_t* d_iters, uint8_t* d_flags)
{
    if (!c || n <= 0 || !d_poi || !d_u || !d_v || !d_zncc) return fail(c, TWDIC_EINVAL, "twdic_frame_capture: bad arguments");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_frame_capture: images not set");
    const size_t m = (size_t)n;
    TW_SPANS(c, "twdic_frame_capture", {d_poi, 8 * m, "d_poi"}, {d_u, 4 * m, "d_u"}, {d_v, 4 * m, "d_v"},
             {d_zncc, 4 * m, "d_zncc"}, {d_p6, 24 * m, "d_p6"}, {d_iters, 4 * m, "d_iters"}, {d_flags, m, "d_flags"});
    // reference gradients and per-POI state now, outside the graph
    IcgnParams ip;
    twdic_status st = icgn_prepare(c, d_poi, n, d_u, d_v, d_p6, d_iters, d_flags, nullptr, ip);
    if (st != TWDIC_OK) return st;
    const FftccParams fp = fftcc_params(c, d_poi, n, d_u, d_v, d_zncc, nullptr);
    drop_graph(c);
    TW_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    // the per-target interpolation data is part of every replay
    hipError_t e = launch_target_interp(c, c->stream);
    if (e == hipSuccess) e = launch_fftcc(fp, c->fftcc_variant, c->stream);
    if (e == hipSuccess) e = launch_icgn(ip, c->cfg.interp, c->icgn_variant, c->stream);
    hipGraph_t gr = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return hip_fail(c, e != hipSuccess ? e : e2, "twdic_frame_capture");
    }
    c->graph = gr;
    e = hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(c);
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    c->graph_ip = ip;
    return TWDIC_OK;
}

twdic_status twdic_frame_replay(twdic_ctx* c)
{
    if (!c) return TWDIC_EINVAL;
    if (!c->graph_exec) return fail(c, TWDIC_EINVAL, "twdic_frame_replay: no captured frame");
    if (!c->ref_loaded || !c->tar_loaded) return fail(c, TWDIC_EINVAL, "twdic_frame_replay: images not set");
    if (c->d_refstate != c->graph_ip.refstate)
        return fail(c, TWDIC_EINVAL, "twdic_frame_replay: the per-POI state was reallocated; capture again");
    const IcgnParams& ip = c->graph_ip;
    if (c->refstate_version != c->ref_version || c->refstate_poi != ip.poi || c->refstate_n != ip.n) {
        twdic_status st = ensure_ref_grads(c);
        if (st != TWDIC_OK) return st;
        TW_HIP(c, launch_ref_state(ip, c->stream));
        c->refstate_version = c->ref_version;
        c->refstate_poi = ip.poi;
        c->refstate_n = ip.n;
    }
    TW_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
    // the graph rebuilt the LUT or the row polynomials
    (c->cfg.interp == TWDIC_INTERP_FUSED ? c->tar_rp_valid : c->tar_pre_valid) = true;
    return TWDIC_OK;
}

twdic_status twdic_get_buffer(twdic_ctx* c, int which, void** d_ptr, size_t* bytes)
{
    if (!c || !d_ptr) return TWDIC_EINVAL;
    const size_t roi = (size_t)c->cfg.roi_w * c->cfg.roi_h;
    const size_t img = (size_t)c->dstride * c->cfg.img_h;
    void* p = nullptr;
    size_t b = roi * 4;
    switch (which) {
    case TWDIC_BUF_RX: p = c->d_rx; break;
    case TWDIC_BUF_RY: p = c->d_ry; break;
    case TWDIC_BUF_TX: p = c->d_tx; break;
    case TWDIC_BUF_TY: p = c->d_ty; break;
    case TWDIC_BUF_TXY: p = c->d_txy; break;
    case TWDIC_BUF_LUT: p = c->d_lut; b = roi * 64; break;
    case TWDIC_BUF_REF: p = c->d_ref; b = img; break;
    case TWDIC_BUF_TAR: p = c->d_tar; b = img; c->tar_exposed = true; break;
    default: return fail(c, TWDIC_EINVAL, "unknown buffer");
    }
    if (which == TWDIC_BUF_RX || which == TWDIC_BUF_RY) {
        twdic_status st = ensure_ref_grads(c);
        if (st != TWDIC_OK) return st;
    }
    *d_ptr = p;
    if (bytes) *bytes = b;
    return TWDIC_OK;
}

// ctx == NULL: default stream, synchronous (the reference's wrappers use the
// legacy default stream).
static twdic_status track_finish(twdic_ctx* c, hipError_t e)
{
    if (e == hipSuccess && !c) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        if (c) return hip_fail(c, e, "tracking kernel");
        return TWDIC_EHIP;
    }
    return TWDIC_OK;
}

twdic_status twdic_track_update_poi(twdic_ctx* c, int32_t* d_poi, int n, const float* d_u, const float* d_v)
{
    if (n < 0 || (n > 0 && (!d_poi || !d_u || !d_v))) return c ? fail(c, TWDIC_EINVAL, "track_update_poi: bad arguments") : TWDIC_EINVAL;
    TW_SPANS(c, "track_update_poi", {d_poi, 8 * (size_t)n, "d_poi"}, {d_u, 4 * (size_t)n, "d_u"},
             {d_v, 4 * (size_t)n, "d_v"});
    if (c && c->refstate_poi == d_poi) c->refstate_version = ~0ull;     // the POIs moved
    return track_finish(c, launch_track_update_poi(d_poi, n, d_u, d_v, c ? c->stream : nullptr));
}

twdic_status twdic_track_current_poi(twdic_ctx* c, const int32_t* d_poi, int n, const float* d_u, const float* d_v,
                                     int32_t* d_out)
{
    if (n < 0 || (n > 0 && (!d_poi || !d_u || !d_v || !d_out)))
        return c ? fail(c, TWDIC_EINVAL, "track_current_poi: bad arguments") : TWDIC_EINVAL;
    TW_SPANS(c, "track_current_poi", {d_poi, 8 * (size_t)n, "d_poi"}, {d_u, 4 * (size_t)n, "d_u"},
             {d_v, 4 * (size_t)n, "d_v"}, {d_out, 8 * (size_t)n, "d_out"});
    return track_finish(c, launch_track_current_poi(d_poi, n, d_u, d_v, d_out, c ? c->stream : nullptr));
}

twdic_status twdic_track_current_poi_acc(twdic_ctx* c, const int32_t* d_poi, int n, const float* d_u,
                                         const float* d_v, const float* d_acc_u, const float* d_acc_v, int32_t* d_out)
{
    if (n < 0 || (n > 0 && (!d_poi || !d_u || !d_v || !d_acc_u || !d_acc_v || !d_out)))
        return c ? fail(c, TWDIC_EINVAL, "track_current_poi_acc: bad arguments") : TWDIC_EINVAL;
    TW_SPANS(c, "track_current_poi_acc", {d_poi, 8 * (size_t)n, "d_poi"}, {d_u, 4 * (size_t)n, "d_u"},
             {d_v, 4 * (size_t)n, "d_v"}, {d_acc_u, 4 * (size_t)n, "d_acc_u"}, {d_acc_v, 4 * (size_t)n, "d_acc_v"},
             {d_out, 8 * (size_t)n, "d_out"});
    return track_finish(c, launch_track_current_poi_acc(d_poi, n, d_u, d_v, d_acc_u, d_acc_v, d_out,
                                                        c ? c->stream : nullptr));
}

twdic_status twdic_track_accumulate_uv(twdic_ctx* c, int n, const float* d_u, const float* d_v, float* d_acc_u,
                                       float* d_acc_v)
{
    if (n < 0 || (n > 0 && (!d_u || !d_v || !d_acc_u || !d_acc_v)))
        return c ? fail(c, TWDIC_EINVAL, "track_accumulate_uv: bad arguments") : TWDIC_EINVAL;
    TW_SPANS(c, "track_accumulate_uv", {d_u, 4 * (size_t)n, "d_u"}, {d_v, 4 * (size_t)n, "d_v"},
             {d_acc_u, 4 * (size_t)n, "d_acc_u"}, {d_acc_v, 4 * (size_t)n, "d_acc_v"});
    return track_finish(c, launch_track_accumulate_uv(n, d_u, d_v, d_acc_u, d_acc_v, c ? c->stream : nullptr));
}

twdic_status twdic_write_csv(const char* path, int n, const int32_t* poi, const float* iu, const float* iv,
                             const float* p6, const int32_t* iters, const float* zncc, const uint8_t* fst,
                             const uint8_t* flags)
{
    if (!path || n < 0 || (n > 0 && (!poi || !iu || !iv || !p6 || !iters))) return TWDIC_EINVAL;
    FILE* f = std::fopen(path, "w");
    if (!f) return TWDIC_EINVAL;
    for (int i = 0; i < n; ++i) {
        const float* p = p6 + 6 * (size_t)i;
        std::fprintf(f, "%d,%d,%d,%g,%g,%g,%d,%g,%g,%g,%d,%g,%d,%d\n", poi[2 * i + 1], poi[2 * i], (int)iu[i],
                     (double)p[0], (double)p[1], (double)p[2], (int)iv[i], (double)p[3], (double)p[4], (double)p[5],
                     iters[i], zncc ? (double)zncc[i] : 0.0, fst ? (int)fst[i] : 0, flags ? (int)flags[i] : 0);
    }
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? TWDIC_OK : TWDIC_EINVAL;
}

twdic_status twdic_strain_ls(twdic_ctx* c, const float* d_field, int nx, int ny, int grid, int M, double* d_dfdx,
                             double* d_dfdy)
{
    if (!c || !d_field || nx < 1 || ny < 1 || grid < 1 || M < 1 || 2 * M + 1 > nx || 2 * M + 1 > ny)
        return c ? fail(c, TWDIC_EINVAL, "strain_ls: bad arguments") : TWDIC_EINVAL;
    const size_t no = (size_t)(nx - 2 * M) * (ny - 2 * M);
    TW_SPANS(c, "strain_ls", {d_field, 4 * (size_t)nx * ny, "d_field"}, {d_dfdx, 8 * no, "d_dfdx"},
             {d_dfdy, 8 * no, "d_dfdy"});
    TW_HIP(c, launch_strain_ls(d_field, nx, ny, grid, M, d_dfdx, d_dfdy, c->stream));
    return TWDIC_OK;
}

twdic_status twdic_malloc(twdic_ctx* c, void** d_ptr, size_t bytes)
{
    if (!c || !d_ptr) return TWDIC_EINVAL;
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e == hipErrorOutOfMemory) return fail(c, TWDIC_ENOMEM, "hipMalloc: out of memory");
    TW_HIP(c, e);
    return TWDIC_OK;
}

twdic_status twdic_free(twdic_ctx* c, void* d_ptr)
{
    if (!c) return TWDIC_EINVAL;
    if (!d_ptr) return TWDIC_OK;
    TW_HIP(c, hipStreamSynchronize(c->stream));
    TW_HIP(c, hipFree(d_ptr));
    return TWDIC_OK;
}

twdic_status twdic_memcpy(twdic_ctx* c, void* dst, const void* src, size_t bytes, int kind)
{
    if (!c || (bytes && (!dst || !src))) return TWDIC_EINVAL;
    if (!bytes) return TWDIC_OK;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) return fail(c, TWDIC_EINVAL, "memcpy kind");
    TW_SPANS(c, "twdic_memcpy", {kind != 1 ? dst : nullptr, bytes, "dst"}, {kind != 0 ? src : nullptr, bytes, "src"});
    TW_HIP(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
    if (kind != 2) TW_HIP(c, hipStreamSynchronize(c->stream));
    return TWDIC_OK;
}

twdic_status twdic_host_alloc(void** h_ptr, size_t bytes)
{
    if (!h_ptr) return TWDIC_EINVAL;
    *h_ptr = nullptr;
    return hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? TWDIC_OK : TWDIC_ENOMEM;
}

twdic_status twdic_host_free(void* h_ptr)
{
    if (!h_ptr) return TWDIC_OK;
    return hipHostFree(h_ptr) == hipSuccess ? TWDIC_OK : TWDIC_EHIP;
}

twdic_status twdic_memset(twdic_ctx* c, void* d_ptr, int value, size_t bytes)
{
    if (!c || (bytes && !d_ptr)) return TWDIC_EINVAL;
    if (!bytes) return TWDIC_OK;
    TW_SPANS(c, "twdic_memset", {d_ptr, bytes, "d_ptr"});
    TW_HIP(c, hipMemsetAsync(d_ptr, value, bytes, c->stream));
    return TWDIC_OK;
}

twdic_status twdic_event_record(twdic_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= 15) return TWDIC_EINVAL;
    TW_HIP(c, hipEventRecord(c->ev[slot], c->stream));
    return TWDIC_OK;
}

twdic_status twdic_event_elapsed(twdic_ctx* c, int a, int b, float* ms)
{
    if (!c || !ms || a < 0 || b < 0 || a >= 15 || b >= 15) return TWDIC_EINVAL;
    TW_HIP(c, hipEventSynchronize(c->ev[b]));
    TW_HIP(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return TWDIC_OK;
}

}  // extern "C"
