// host_check.cpp -- drives the library's host code (C ABI, batch and serving
// layers) through its normal and error paths, for the build whose host code
// carries AddressSanitizer and UndefinedBehaviorSanitizer (tw_amd/build.py,
// tw_host_check_asan).  Without a GPU it runs part A only: configuration
// validation over seeded random geometries, the POI count, the CSV writer,
// synthetic pairs and the create calls that must fail cleanly.  With a GPU it
// also runs part B: a context through every entry point (host and device
// images, FFT-CC, IC-GN, the frame step and its captured graph, the
// no-allocation IC-GN, tracking, strain, buffer checks that must refuse), a
// batch in the flat and direct forms (checked bit for bit against the
// context), and a server under a burst of requests.  Exit status 0 means every
// check held; the sanitizers abort on their own findings.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "twdic.h"
#include "twdic_batch.h"
#include "twdic_serve.h"
#include "twdic_synth.h"

static int g_fail = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #c); \
            ++g_fail;                                                            \
        }                                                                        \
    } while (0)

static twdic_config config(int w, int h, int S, int grid, int N)
{
    twdic_config c{};
    c.img_w = w; c.img_h = h;
    c.roi_x = 1; c.roi_y = 1; c.roi_w = w - 2; c.roi_h = h - 2;
    c.sx = S; c.sy = S; c.gx = grid; c.gy = grid; c.mx = 2; c.my = 2;
    c.fft_n = N; c.max_iter = 20; c.delta_p = 0.001f; c.interp = TWDIC_INTERP_FUSED;
    return c;
}

static void synth(int w, int h, uint64_t seed, std::vector<uint8_t>& ref, std::vector<uint8_t>& tar)
{
    twdic_speckle sp{};
    sp.w = w; sp.h = h; sp.nspots = w * h / 60; sp.seed = seed;
    sp.sigma_min = 1.5; sp.sigma_max = 3.0; sp.amp_min = 60; sp.amp_max = 160;
    twdic_deform d{};
    d.kind = TWDIC_DEFORM_AFFINE; d.u0 = 1.3; d.v0 = -0.6; d.exx = 0.002; d.eyy = -0.001; d.exy = 0.0005;
    d.xc = w / 2.0; d.yc = h / 2.0;
    ref.assign((size_t)w * h, 0);
    tar.assign((size_t)w * h, 0);
    CHECK(twdic_synth_pair(&sp, &d, ref.data(), tar.data(), w) == 0);
}

// ---------------------------------------------------------------- part A
static void part_a()
{
    std::mt19937 rng(20261203);
    auto r = [&](int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); };
    int valid = 0;
    for (int t = 0; t < 4000; ++t) {
        twdic_config c{};
        c.img_w = r(-3, 5000); c.img_h = r(-3, 5000);
        c.roi_x = r(-2, 40); c.roi_y = r(-2, 40);
        c.roi_w = r(-2, c.img_w + 3); c.roi_h = r(-2, c.img_h + 3);
        c.sx = r(-1, 70); c.sy = r(-1, 70); c.gx = r(-1, 30); c.gy = r(-1, 30);
        c.mx = r(-1, 10); c.my = r(-1, 10); c.fft_n = r(-2, 70); c.max_iter = r(-1, 30);
        c.delta_p = (float)r(-1, 5) * 1e-3f; c.interp = r(-1, 3);
        int32_t nx = -7, ny = -7;
        const twdic_status s = twdic_num_poi(&c, &nx, &ny);
        CHECK(s == TWDIC_OK || s == TWDIC_EINVAL);
        twdic_ctx* ctx = nullptr;
        const twdic_status cs = twdic_create(&c, 0, &ctx);
        if (s == TWDIC_OK) ++valid;
        CHECK(cs == TWDIC_OK || cs == TWDIC_EINVAL || cs == TWDIC_ENODEV || cs == TWDIC_ENOMEM);
        if (cs == TWDIC_OK) twdic_destroy(ctx);
        else CHECK(ctx == nullptr);
    }
    CHECK(valid > 0);
    CHECK(twdic_num_poi(nullptr, nullptr, nullptr) == TWDIC_EINVAL);
    CHECK(twdic_create(nullptr, 0, nullptr) == TWDIC_EINVAL);
    CHECK(twdic_last_error(nullptr) != nullptr);
    // the result table writer
    const int n = 5;
    std::vector<int32_t> poi(2 * n);
    std::vector<float> a(n, 1.f), p6(6 * n, 0.5f), z(n, 0.9f);
    std::vector<int32_t> it(n, 3);
    std::vector<uint8_t> fl(n, 0);
    for (int i = 0; i < n; ++i) { poi[2 * i] = 10 + i; poi[2 * i + 1] = 20 + i; }
    const std::string path = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") + "/tw_host_check.csv";
    CHECK(twdic_write_csv(path.c_str(), n, poi.data(), a.data(), a.data(), p6.data(), it.data(), z.data(), fl.data(),
                          fl.data()) == TWDIC_OK);
    std::remove(path.c_str());
    // synthetic pairs of odd sizes
    std::vector<uint8_t> ref, tar;
    synth(97, 61, 5, ref, tar);
    // creation of the batch and the server: refused without a device or with
    // bad configurations
    twdic_server_config sc{};
    sc.device = 0; sc.slots = 0; sc.queue_capacity = 4; sc.max_width = 64; sc.max_height = 64; sc.max_pois = 10;
    twdic_server* srv = nullptr;
    CHECK(twdic_server_create(&sc, &srv) == TWDIC_EINVAL && srv == nullptr);
    sc.slots = 2; sc.max_width = 70000;
    CHECK(twdic_server_create(&sc, &srv) == TWDIC_EINVAL && srv == nullptr);
    CHECK(twdic_batch_create(0, nullptr) == TWDIC_EINVAL);
}

// ---------------------------------------------------------------- part B
// equal element for element; otherwise reports how many differ and by how much
static bool same(const std::vector<float>& a, const std::vector<float>& b, const char* what)
{
    if (a.size() != b.size()) {
        std::fprintf(stderr, "%s: sizes %zu and %zu\n", what, a.size(), b.size());
        return false;
    }
    size_t nd = 0;
    double mx = 0;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::memcmp(&a[i], &b[i], sizeof(float)) != 0) {
            ++nd;
            mx = std::fmax(mx, std::fabs((double)a[i] - (double)b[i]));
        }
    if (nd) std::fprintf(stderr, "%s: %zu of %zu differ, max |diff| %g\n", what, nd, a.size(), mx);
    return nd == 0;
}

template <class T>
static T* dalloc(twdic_ctx* c, size_t n)
{
    void* p = nullptr;
    CHECK(twdic_malloc(c, &p, n * sizeof(T)) == TWDIC_OK);
    return (T*)p;
}

template <class T>
static std::vector<T> down(twdic_ctx* c, const T* d, size_t n)
{
    std::vector<T> h(n);
    CHECK(twdic_memcpy(c, h.data(), d, n * sizeof(T), 1) == TWDIC_OK);
    return h;
}

static void part_b()
{
    const int W = 256, H = 224;
    std::vector<uint8_t> ref, tar;
    synth(W, H, 11, ref, tar);
    // N = 2S as the reference: every window lies in the image, so FFT-CC
    // writes every POI (a window that leaves the image leaves its outputs
    // untouched, and IC-GN would then start from whatever the buffer held)
    twdic_config cfg = config(W, H, 10, 6, 20);
    twdic_ctx* c = nullptr;
    CHECK(twdic_create(&cfg, 0, &c) == TWDIC_OK);
    if (!c) return;
    int32_t nx = 0, ny = 0;
    CHECK(twdic_num_poi(&cfg, &nx, &ny) == TWDIC_OK);
    const int n = nx * ny;
    CHECK(twdic_set_reference(c, ref.data(), W, 0) == TWDIC_OK);
    CHECK(twdic_set_target(c, tar.data(), W, 0) == TWDIC_OK);
    int32_t* poi = dalloc<int32_t>(c, 2 * (size_t)n);
    float *u = dalloc<float>(c, n), *v = dalloc<float>(c, n), *z = dalloc<float>(c, n), *p6 = dalloc<float>(c, 6 * (size_t)n);
    int32_t* it = dalloc<int32_t>(c, n);
    uint8_t *fl = dalloc<uint8_t>(c, n), *st = dalloc<uint8_t>(c, n);
    CHECK(twdic_poi_grid(c, poi) == TWDIC_OK);
    CHECK(twdic_fftcc(c, poi, n, u, v, z, st) == TWDIC_OK);
    CHECK(twdic_icgn(c, poi, n, u, v, p6, it, fl, nullptr) == TWDIC_OK);
    CHECK(twdic_sync(c) == TWDIC_OK);
    const std::vector<float> u1 = down(c, u, n), v1 = down(c, v, n);
    const std::vector<uint8_t> f1 = down(c, fl, n);
    int ok = 0;
    for (int i = 0; i < n; ++i) ok += f1[i] == 0 && std::fabs(u1[i] - 1.3f) < 0.5f;
    CHECK(ok > n * 9 / 10);
    // frame step, captured graph and replay
    CHECK(twdic_frame(c, poi, n, u, v, z, p6, it, fl) == TWDIC_OK);
    CHECK(twdic_frame_capture(c, poi, n, u, v, z, p6, it, fl) == TWDIC_OK);
    for (int k = 0; k < 2; ++k) CHECK(twdic_frame_replay(c) == TWDIC_OK);
    CHECK(twdic_sync(c) == TWDIC_OK);
    CHECK(same(down(c, u, n), u1, "frame u") && same(down(c, v, n), v1, "frame v"));
    // a second plain FFT-CC + IC-GN
    CHECK(twdic_fftcc(c, poi, n, u, v, z, st) == TWDIC_OK);
    CHECK(twdic_icgn(c, poi, n, u, v, p6, it, fl, nullptr) == TWDIC_OK);
    CHECK(twdic_sync(c) == TWDIC_OK);
    CHECK(same(down(c, u, n), u1, "repeated u"));
    // no-allocation IC-GN from the FFT-CC guess
    CHECK(twdic_fftcc(c, poi, n, u, v, z, st) == TWDIC_OK);
    CHECK(twdic_icgn_noalloc(c, poi, n, u, v, p6, it, fl) == TWDIC_OK);
    CHECK(twdic_sync(c) == TWDIC_OK);
    CHECK(same(down(c, u, n), u1, "noalloc u"));
    // tracking and strain
    float *au = dalloc<float>(c, n), *av = dalloc<float>(c, n);
    int32_t* cur = dalloc<int32_t>(c, 2 * (size_t)n);
    CHECK(twdic_memset(c, au, 0, 4 * (size_t)n) == TWDIC_OK && twdic_memset(c, av, 0, 4 * (size_t)n) == TWDIC_OK);
    CHECK(twdic_track_accumulate_uv(c, n, u, v, au, av) == TWDIC_OK);
    CHECK(twdic_track_current_poi(c, poi, n, u, v, cur) == TWDIC_OK);
    double *dx = dalloc<double>(c, n), *dy = dalloc<double>(c, n);
    CHECK(twdic_strain_ls(c, u, nx, ny, cfg.gx, 2, dx, dy) == TWDIC_OK);
    CHECK(twdic_sync(c) == TWDIC_OK);
    // refusals: short buffer, host pointer, bad arguments
    float* tiny = dalloc<float>(c, 4);
    CHECK(twdic_fftcc(c, poi, n, tiny, v, z, nullptr) == TWDIC_EBOUNDS);
    std::vector<float> hostu(n);
    CHECK(twdic_fftcc(c, poi, n, hostu.data(), v, z, nullptr) == TWDIC_EINVAL);
    CHECK(twdic_icgn(c, poi, -1, u, v, nullptr, nullptr, nullptr, nullptr) == TWDIC_EINVAL);
    CHECK(twdic_set_target(c, nullptr, W, 0) == TWDIC_EINVAL);
    CHECK(std::strlen(twdic_last_error(c)) > 0);

    // batch: three pairs (two copies of this pair and one of a different size
    // and window), flat form and direct form, against the context bit for bit
    struct PairBuf {
        int w, h, S, N;
        uint8_t *ref, *tar;
        int32_t* poi;
        int n;
        float *u, *v;
        int32_t* it;
        uint8_t* fl;
    };
    std::vector<PairBuf> pb;
    // the third pair: 65 x 65 subsets (beyond the step table) and N = 64
    const int dims[3][4] = {{W, H, 10, 20}, {W, H, 10, 20}, {256, 256, 32, 64}};
    for (int i = 0; i < 3; ++i) {
        PairBuf p{dims[i][0], dims[i][1], dims[i][2], dims[i][3]};
        std::vector<uint8_t> r2, t2;
        synth(p.w, p.h, 11 + (i == 2), r2, t2);
        twdic_config pc = config(p.w, p.h, p.S, 6, p.N);
        int32_t px = 0, py = 0;
        CHECK(twdic_num_poi(&pc, &px, &py) == TWDIC_OK);
        p.n = px * py;
        p.ref = dalloc<uint8_t>(c, r2.size());
        p.tar = dalloc<uint8_t>(c, t2.size());
        CHECK(twdic_memcpy(c, p.ref, r2.data(), r2.size(), 0) == TWDIC_OK);
        CHECK(twdic_memcpy(c, p.tar, t2.data(), t2.size(), 0) == TWDIC_OK);
        p.poi = dalloc<int32_t>(c, 2 * (size_t)p.n);
        std::vector<int32_t> hp(2 * (size_t)p.n);
        for (int y = 0, k = 0; y < py; ++y)
            for (int x = 0; x < px; ++x, ++k) {
                hp[2 * k] = 1 + pc.my + pc.sy + y * pc.gy;
                hp[2 * k + 1] = 1 + pc.mx + pc.sx + x * pc.gx;
            }
        CHECK(twdic_memcpy(c, p.poi, hp.data(), hp.size() * 4, 0) == TWDIC_OK);
        p.u = dalloc<float>(c, p.n); p.v = dalloc<float>(c, p.n);
        p.it = dalloc<int32_t>(c, p.n); p.fl = dalloc<uint8_t>(c, p.n);
        pb.push_back(p);
    }
    auto make_pairs = [&]() {
        std::vector<twdic_pair> v;
        for (auto& p : pb) {
            twdic_pair q{};
            q.d_ref = p.ref; q.d_tar = p.tar; q.stride = p.w; q.img_w = p.w; q.img_h = p.h;
            q.roi_x = 1; q.roi_y = 1; q.roi_w = p.w - 2; q.roi_h = p.h - 2;
            q.sx = p.S; q.sy = p.S; q.fft_n = p.N; q.max_iter = 20; q.delta_p = 0.001f;
            q.d_poi_yx = p.poi; q.n = p.n; q.d_u = p.u; q.d_v = p.v; q.d_iters = p.it; q.d_flags = p.fl;
            v.push_back(q);
        }
        return v;
    };
    for (int form = 0; form < 2; ++form) {
        if (form == 1) setenv("TWDIC_BATCH_FORM", "direct", 1);
        twdic_batch* b = nullptr;
        CHECK(twdic_batch_create(0, &b) == TWDIC_OK);
        unsetenv("TWDIC_BATCH_FORM");
        if (!b) continue;
        std::vector<twdic_pair> pairs = make_pairs();
        // a bad pair is refused and leaves the batch empty
        std::vector<twdic_pair> bad = pairs;
        bad[1].d_u = tiny;
        CHECK(twdic_batch_set_pairs(b, bad.data(), 3) == TWDIC_EBOUNDS);
        CHECK(twdic_batch_set_pairs(b, pairs.data(), 3) == TWDIC_OK);
        for (int rep = 0; rep < 2; ++rep) CHECK(twdic_batch_run(b) == TWDIC_OK);
        CHECK(twdic_batch_sync(b) == TWDIC_OK);
        twdic_batch_info info{};
        CHECK(twdic_batch_info_get(b, &info) == TWDIC_OK && info.npairs == 3);
        const uint8_t* tars[3] = {pb[0].tar, nullptr, pb[2].tar};
        CHECK(twdic_batch_set_targets(b, tars) == TWDIC_OK);
        CHECK(twdic_batch_run(b) == TWDIC_OK && twdic_batch_sync(b) == TWDIC_OK);
        CHECK(same(down(c, pb[0].u, pb[0].n), u1, "batch pair 0 u") && same(down(c, pb[0].v, pb[0].n), v1, "batch pair 0 v"));
        CHECK(same(down(c, pb[1].u, pb[1].n), u1, "batch pair 1 u"));
        CHECK(twdic_batch_destroy(b) == TWDIC_OK);
    }

    // server: a burst of requests against one reference, results equal the context's
    twdic_server_config sc{};
    sc.device = 0; sc.slots = 4; sc.queue_capacity = 8; sc.hbm_budget = (size_t)256 << 20;
    sc.max_width = 512; sc.max_height = 512; sc.max_pois = 4096;
    twdic_server* s = nullptr;
    CHECK(twdic_server_create(&sc, &s) == TWDIC_OK);
    if (s) {
        CHECK(twdic_server_set_reference(s, 3, ref.data(), W, H, W) == TWDIC_OK);
        std::vector<int32_t> hp = down(c, poi, 2 * (size_t)n);
        int ps = -1;
        CHECK(twdic_server_add_poiset(s, hp.data(), n, &ps) == TWDIC_OK);
        const int R = 40;
        std::vector<std::vector<float>> ru(R, std::vector<float>(n)), rv(R, std::vector<float>(n));
        std::vector<uint64_t> tick(R);
        for (int k = 0; k < R; ++k) {
            twdic_request q{};
            q.ref_id = 3; q.poiset = ps; q.target = tar.data(); q.stride = W; q.subset = 10; q.fft_n = 20;
            q.max_iter = 20; q.delta_p = 0.001f; q.u = ru[k].data(); q.v = rv[k].data(); q.memory = TWDIC_MEM_PAGEABLE;
            CHECK(twdic_server_submit(s, &q, 0, &tick[k]) == TWDIC_OK);
        }
        for (int k = 0; k < R; ++k) CHECK(twdic_server_wait(s, tick[k]) == TWDIC_OK);
        for (int k = 0; k < R; ++k) CHECK(same(ru[k], u1, "served u") && same(rv[k], v1, "served v"));
        twdic_server_stats ss{};
        CHECK(twdic_server_stats_get(s, &ss) == TWDIC_OK && ss.completed == (uint64_t)R);
        twdic_request bad{};
        bad.ref_id = 99; bad.poiset = ps; bad.target = tar.data(); bad.stride = W; bad.subset = 10; bad.fft_n = 20;
        bad.u = ru[0].data(); bad.v = rv[0].data();
        uint64_t tk = 0;
        CHECK(twdic_server_submit(s, &bad, 0, &tk) != TWDIC_OK);
        CHECK(twdic_server_destroy(s) == TWDIC_OK);
    }
    CHECK(twdic_destroy(c) == TWDIC_OK);
}

int main()
{
    part_a();
    twdic_config probe = config(64, 64, 5, 8, 16);
    twdic_ctx* c = nullptr;
    const twdic_status s = twdic_create(&probe, 0, &c);
    if (s == TWDIC_OK) {
        twdic_destroy(c);
        part_b();
        std::printf("host_check: parts A and B, %d failed checks\n", g_fail);
    } else {
        std::printf("host_check: part A only (no device: status %d), %d failed checks\n", (int)s, g_fail);
    }
    return g_fail ? 1 : 0;
}
