// batch_api.cpp -- the batched form (include/twdic_batch.h): descriptors and
// work lists for many image pairs; the flat form (one launch per stage and
// kernel form over all pairs, the default), the direct form (per pair, the
// single-context kernels on four streams), and the persistent-kernel stages
// with hipGraph replay and the launch-parameter autotuner (switched off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "twdic_batch.h"
#include "twdic_internal.h"

using namespace twdic;

namespace {
enum { CNT_STATE = 0, CNT_F64 = 1, CNT_F32 = 2, CNT_FDFT = 3, CNT_ICGN = 4, NCNT = 5 };
}

struct twdic_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<twdic_pair> pairs;
    std::vector<PairDesc> h_desc;
    PairDesc* d_desc = nullptr;
    int2* d_items = nullptr;          // every POI as (pair, poi), pair-major: IC-GN and reference state
    int2* d_fitems[3] = {};           // FFT-CC items per window class (N = 64, 32, other)
    int64_t nf[3] = {};
    int64_t total = 0;
    uint32_t* d_g2 = nullptr;         // reference gradient planes of all pairs
    uint4* d_rp = nullptr;            // target row polynomials of all pairs (rebuilt every step)
    int max_rh = 0;                   // largest ROI height among the pairs with a plane
    bool all_rp = false;              // every pair has a plane
    double* d_state = nullptr;        // per-POI reference state of all pairs
    float* d_zncc_scratch = nullptr;  // for pairs without a ZNCC output
    std::map<int, float2*> tw;        // FFT twiddle tables per window edge
    int* d_counters = nullptr;
    bool state_valid = false;
    twdic_batch_tuning tune{4, 3, 1, 5, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_dirty = true;
    hipEvent_t ev[16] = {};
    size_t bytes = 0;
    // direct form: pairs are issued round-robin on these streams, which fork
    // from and join back into `stream`
    static constexpr int kLanes = 4;
    hipStream_t lane[kLanes] = {};
    hipEvent_t fork = nullptr, join[kLanes] = {};
    int lanes = 4;                    // TWDIC_BATCH_STREAMS (1..4)
    // flat form: one launch per stage and kernel form over all pairs
    bool flat = true;                 // TWDIC_BATCH_FORM=direct selects the direct form
    int2* d_mitems[4] = {};           // FFT-CC items (pair, POI) per matrix-core kernel form (mma_form)
    int64_t nm[4] = {};
    // IC-GN workgroup items (pair, first POI), list t + 2 w: step table fits
    // (t = 0) or not (t = 1), half-wavefront (w = 0) or whole wavefront (w = 1)
    // per POI, by the pair's POI count (icgn_lanes_for)
    int2* d_wg[4] = {};
    int64_t nwg[4] = {};
    // persistent form: grid sizes from each kernel's own occupancy, computed
    // when the pairs are set (never inside a stream capture)
    int pblocks_state = 0, pblocks_icgn = 0, pblocks_fftcc[3][5] = {};
};

namespace {

twdic_status bfail(twdic_batch* b, twdic_status st, const std::string& what)
{
    if (b) b->err = what;
    return st;
}

#define TB_HIP(b, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return bfail((b), TWDIC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int fclass(int N) { return N == 64 ? 0 : N == 32 ? 1 : 2; }

// The persistent-kernel form (work queues, hipGraph replay, autotuner)
// stalled on its first run on an MI355X and the cause has not been found; it
// is compiled only with -DTWDIC_BATCH_EXEC.  Without it, a batch runs in the
// flat form (one launch per stage and kernel form over the items of all
// pairs) or, with TWDIC_BATCH_FORM=direct, in the direct form (per pair, the
// single-context kernels launch_fftcc and launch_icgn on four streams).  The
// per-POI reference state is built per pair (launch_ref_state) in both.
#ifdef TWDIC_BATCH_EXEC
constexpr bool kPersistent = true;
#else
constexpr bool kPersistent = false;
#endif

void release(twdic_batch* b)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    void* bufs[] = {b->d_desc, b->d_items, b->d_fitems[0], b->d_fitems[1], b->d_fitems[2], b->d_g2, b->d_state,
                    b->d_zncc_scratch, b->d_counters, b->d_mitems[0], b->d_mitems[1], b->d_mitems[2],
                    b->d_mitems[3], b->d_wg[0], b->d_wg[1], b->d_wg[2], b->d_wg[3], b->d_rp};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    for (int f = 0; f < 4; ++f) { b->d_mitems[f] = nullptr; b->nm[f] = 0; }
    for (int t = 0; t < 4; ++t) { b->d_wg[t] = nullptr; b->nwg[t] = 0; }
    for (int c = 0; c < 3; ++c) b->nf[c] = 0;
    for (auto& kv : b->tw) (void)hipFree(kv.second);
    b->tw.clear();
    b->d_desc = nullptr; b->d_items = nullptr;
    b->d_fitems[0] = b->d_fitems[1] = b->d_fitems[2] = nullptr;
    b->d_g2 = nullptr; b->d_state = nullptr; b->d_zncc_scratch = nullptr; b->d_counters = nullptr;
    b->d_rp = nullptr; b->max_rh = 0; b->all_rp = false;
    b->bytes = 0;
}

const char* check_pair(const twdic_pair& p)
{
    if (!p.d_ref || !p.d_tar || (p.n > 0 && (!p.d_poi_yx || !p.d_u || !p.d_v))) return "null pointer in pair";
    if (p.img_w <= 0 || p.img_h <= 0 || p.stride < p.img_w) return "bad image size/stride";
    if (p.stride > 65535 || p.img_h > 65535 || (int64_t)p.stride * p.img_h >= (int64_t)1 << 31)
        return "images are limited to 65535 px per side (stride included) and 2^31 bytes";
    if (p.roi_x < 1 || p.roi_y < 1 || p.roi_w < 2 || p.roi_h < 2 || p.roi_x + p.roi_w + 1 > p.img_w ||
        p.roi_y + p.roi_h + 1 > p.img_h)
        return "ROI needs a one-pixel frame inside the image";
    if (p.sx < 1 || p.sy < 1 || p.sx > 60 || p.sy > 60) return "subset half size must be in [1, 60]";
    if (p.fft_n < 2 || p.fft_n > kMaxFftN || (p.fft_n & 1)) return "fft_n must be even and in [2, 64]";
    if (p.max_iter < 1 || !(p.delta_p >= 0.f) || p.n < 0) return "bad iteration parameters or POI count";
    return nullptr;
}

// Every device buffer of a pair must lie in a HIP allocation that holds what
// the pair's kernels touch (twdic_device_span); checked before anything runs.
twdic_status check_pair_spans(twdic_batch* b, int i, const twdic_pair& p)
{
    const size_t n = (size_t)p.n, img = (size_t)p.stride * (p.img_h - 1) + p.img_w;
    const struct {
        const void* ptr;
        size_t bytes;
        const char* name;
    } need[] = {{p.d_ref, img, "d_ref"}, {p.d_tar, img, "d_tar"},      {p.d_poi_yx, 8 * n, "d_poi_yx"},
                {p.d_u, 4 * n, "d_u"},   {p.d_v, 4 * n, "d_v"},        {p.d_zncc, 4 * n, "d_zncc"},
                {p.d_p6, 24 * n, "d_p6"}, {p.d_iters, 4 * n, "d_iters"}, {p.d_flags, n, "d_flags"}};
    for (const auto& k : need) {
        if (!k.ptr || !k.bytes) continue;
        size_t have = 0;
        if (twdic_device_span(k.ptr, &have) != TWDIC_OK)
            return bfail(b, TWDIC_EINVAL, "pair " + std::to_string(i) + ": " + k.name + " is not in a device allocation");
        if (have < k.bytes)
            return bfail(b, TWDIC_EBOUNDS, "pair " + std::to_string(i) + ": " + k.name + " holds " +
                                               std::to_string(have) + " bytes, the pair needs " +
                                               std::to_string(k.bytes));
    }
    return TWDIC_OK;
}

// Forgets the pairs after a failed set_pairs, so no later call launches on
// item lists or descriptors that were only partly built.
void clear_pairs(twdic_batch* b)
{
    release(b);
    b->pairs.clear();
    b->h_desc.clear();
    b->total = 0;
    b->state_valid = false;
    b->graph_dirty = true;
}

twdic_status upload_desc(twdic_batch* b)
{
    TB_HIP(b, hipMemcpyAsync(b->d_desc, b->h_desc.data(), sizeof(PairDesc) * b->h_desc.size(), hipMemcpyHostToDevice,
                             b->stream));
    TB_HIP(b, hipStreamSynchronize(b->stream));   // host descriptors may change right after
    return TWDIC_OK;
}

twdic_status build_state(twdic_batch* b)
{
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        const IcgnParams& q = b->h_desc[i].icgn;
        TB_HIP(b, launch_gradients(q.ref, q.stride, q.sx0, q.sy0, q.rw, q.rh, nullptr, nullptr, nullptr,
                                   const_cast<uint32_t*>(q.g2), b->stream));
    }
    TB_HIP(b, launch_fill_zero(b->d_counters + CNT_STATE, 1, b->stream));
    TB_HIP(b, launch_batch_ref_state(b->d_desc, b->d_items, (int)b->total, b->d_counters + CNT_STATE,
                                     b->pblocks_state, b->stream));
    b->state_valid = true;
    return TWDIC_OK;
}

// The per-frame sequence: reset counters, FFT-CC per window class, IC-GN.
twdic_status enqueue_frame(twdic_batch* b, bool fftcc, bool icgn)
{
    TB_HIP(b, launch_fill_zero(b->d_counters + CNT_F64, 4, b->stream));
    if (fftcc) {
        const int Ns[3] = {64, 32, 0};
        for (int c = 0; c < 3; ++c) {
            if (!b->nf[c]) continue;
            const int N = c == 2 ? 0 : Ns[c];
            TB_HIP(b, launch_batch_fftcc(b->d_desc, b->d_fitems[c], (int)b->nf[c], b->d_counters + CNT_F64 + c, N,
                                         b->tune.fftcc_variant,
                                         b->pblocks_fftcc[c][b->tune.fftcc_variant], b->stream));
        }
    }
    if (icgn) TB_HIP(b, launch_batch_row_poly(b->d_desc, (int)b->pairs.size(), b->max_rh, b->stream));
    if (icgn)
        TB_HIP(b, launch_batch_icgn(b->d_desc, b->d_items, (int)b->total, b->d_counters + CNT_ICGN,
                                    std::max(1, b->tune.icgn_chunk), b->pblocks_icgn, b->stream));
    return TWDIC_OK;
}

// Direct form: per-pair launches of the single-context kernels.
twdic_status build_state_direct(twdic_batch* b)
{
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        const IcgnParams& q = b->h_desc[i].icgn;
        TB_HIP(b, launch_gradients(q.ref, q.stride, q.sx0, q.sy0, q.rw, q.rh, nullptr, nullptr, nullptr,
                                   const_cast<uint32_t*>(q.g2), b->stream));
        if (q.n > 0) TB_HIP(b, launch_ref_state(q, b->stream));
    }
    b->state_valid = true;
    return TWDIC_OK;
}

twdic_status run_direct(twdic_batch* b, bool fftcc, bool icgn)
{
    if (!b->state_valid) {
        twdic_status st = build_state_direct(b);
        if (st != TWDIC_OK) return st;
    }
    // every pair's row polynomials of this step's targets, before the fork
    if (icgn) TB_HIP(b, launch_batch_row_poly(b->d_desc, (int)b->pairs.size(), b->max_rh, b->stream));
    // fork: every lane stream waits for the work already on the batch stream
    const int L = std::max(1, std::min(b->lanes, twdic_batch::kLanes));
    TB_HIP(b, hipEventRecord(b->fork, b->stream));
    for (int k = 0; k < L; ++k) TB_HIP(b, hipStreamWaitEvent(b->lane[k], b->fork, 0));
    for (size_t i = 0; i < b->h_desc.size(); ++i) {
        const PairDesc& d = b->h_desc[i];
        hipStream_t s = b->lane[i % L];
        if (fftcc && d.fftcc.n > 0) TB_HIP(b, launch_fftcc(d.fftcc, kFftccDftMfma, s));
        if (icgn && d.icgn.n > 0) TB_HIP(b, launch_icgn(d.icgn, TWDIC_INTERP_FUSED, 5, s));
    }
    // join: later work on the batch stream (events, syncs, the next step)
    // waits for every lane
    for (int k = 0; k < L; ++k) {
        TB_HIP(b, hipEventRecord(b->join[k], b->lane[k]));
        TB_HIP(b, hipStreamWaitEvent(b->stream, b->join[k], 0));
    }
    return TWDIC_OK;
}

// Flat form: FFT-CC per matrix-core kernel form, then IC-GN per step-table
// fit, each one launch over the items of every pair on the batch stream.  The
// same per-POI device code as the single-context kernels.
twdic_status enqueue_flat(twdic_batch* b, bool fftcc, bool icgn)
{
    if (fftcc)
        for (int f = 0; f < 4; ++f)
            if (b->nm[f]) TB_HIP(b, launch_batch_fftcc_mma(b->d_desc, b->d_mitems[f], (int)b->nm[f], f, b->stream));
    if (icgn) TB_HIP(b, launch_batch_row_poly(b->d_desc, (int)b->pairs.size(), b->max_rh, b->stream));
    if (icgn)
        for (int t = 0; t < 4; ++t)
            if (b->nwg[t])
                TB_HIP(b, launch_batch_icgn_wg(b->d_desc, b->d_wg[t], (int)b->nwg[t], (t & 1) == 0, t < 2 ? 32 : 64,
                                               b->tune.icgn_waves, b->tune.icgn_xcd_ranges != 0, b->all_rp,
                                               b->stream));
    return TWDIC_OK;
}

// The full step (FFT-CC and IC-GN) is replayed from a hipGraph when the
// tuning asks for it.  Its kernels read the pair descriptors from device
// memory, so new targets and references need no new capture; new pairs (item
// lists) or tuning mark it stale.  The per-POI reference state is rebuilt
// eagerly, before the launch, when a reference changed.
twdic_status run_flat(twdic_batch* b, bool fftcc, bool icgn)
{
    if (!b->state_valid) {
        twdic_status st = build_state_direct(b);
        if (st != TWDIC_OK) return st;
    }
    if (!(fftcc && icgn && b->tune.use_graph)) return enqueue_flat(b, fftcc, icgn);
    if (b->graph_dirty) {
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
        TB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const twdic_status st = enqueue_flat(b, true, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st != TWDIC_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return st != TWDIC_OK ? st : bfail(b, TWDIC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        }
        b->graph = g;
        TB_HIP(b, hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        b->graph_dirty = false;
    }
    TB_HIP(b, hipGraphLaunch(b->graph_exec, b->stream));
    return TWDIC_OK;
}

std::string signature(twdic_batch* b)
{
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, b->device);
    std::map<std::pair<int, int>, int64_t> hist;   // (N, subset half size) -> POIs
    for (const auto& p : b->pairs) hist[{p.fft_n, p.sx}] += p.n;
    std::ostringstream os;
    os << "arch=" << prop.gcnArchName << ";cus=" << prop.multiProcessorCount << ";pairs=" << b->pairs.size();
    for (const auto& kv : hist) os << ";N" << kv.first.first << "S" << kv.first.second << "=" << kv.second;
    return os.str();
}

}  // namespace

namespace {

twdic_status set_pairs_impl(twdic_batch* b, const twdic_pair* pairs, int npairs)
{
    release(b);
    b->pairs.assign(pairs, pairs + npairs);
    b->h_desc.assign(npairs, PairDesc{});
    size_t g2_elems = 0, state_elems = 0, zn = 0, rp_elems = 0;
    int64_t total = 0;
    for (const auto& p : b->pairs) {
        g2_elems += (size_t)p.roi_w * p.roi_h;
        if (row_poly_fits(p.roi_x, p.roi_y, p.roi_w, p.roi_h)) rp_elems += row_poly_bytes(p.roi_w, p.roi_h) / 16;
        state_elems += (size_t)p.n * kRefStateStride;
        if (!p.d_zncc) zn += p.n;
        total += p.n;
    }
    b->total = total;
    auto alloc = [&](void** ptr, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
        if (e == hipSuccess) b->bytes += bytes;
        return e;
    };
    TB_HIP(b, alloc((void**)&b->d_g2, g2_elems * sizeof(uint32_t)));
    TB_HIP(b, alloc((void**)&b->d_rp, rp_elems * sizeof(uint4)));
    TB_HIP(b, alloc((void**)&b->d_state, state_elems * sizeof(double)));
    TB_HIP(b, alloc((void**)&b->d_zncc_scratch, zn * sizeof(float)));
    TB_HIP(b, alloc((void**)&b->d_desc, sizeof(PairDesc) * std::max(1, npairs)));
    TB_HIP(b, alloc((void**)&b->d_items, sizeof(int2) * std::max<int64_t>(1, total)));
    TB_HIP(b, alloc((void**)&b->d_counters, sizeof(int) * NCNT));
    std::vector<int2> items;
    std::vector<int2> fitems[3];
    items.reserve(total);
    size_t g2_off = 0, st_off = 0, zn_off = 0, rp_off = 0;
    b->all_rp = true;
    for (int i = 0; i < npairs; ++i) {
        const twdic_pair& p = b->pairs[i];
        if (!b->tw.count(p.fft_n)) {
            float2 t[kMaxFftN];
            for (int k = 0; k < p.fft_n; ++k) {
                double a = -2.0 * M_PI * (double)k / (double)p.fft_n;
                t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
            }
            float2* d = nullptr;
            TB_HIP(b, alloc((void**)&d, sizeof(float2) * kMaxFftN));
            b->tw[p.fft_n] = d;
            TB_HIP(b, hipMemcpy(d, t, sizeof(float2) * p.fft_n, hipMemcpyHostToDevice));
        }
        IcgnParams& q = b->h_desc[i].icgn;
        q.ref = p.d_ref; q.tar = p.d_tar; q.stride = p.stride;
        q.sx0 = p.roi_x; q.sy0 = p.roi_y; q.rw = p.roi_w; q.rh = p.roi_h;
        q.Sx = p.sx; q.Sy = p.sy; q.Wsub = 2 * p.sx + 1; q.Hsub = 2 * p.sy + 1; q.s = q.Wsub * q.Hsub;
        q.max_iter = p.max_iter; q.delta_p = p.delta_p;
        q.lanes = icgn_lanes_for(p.n);    // as a single context on this pair's POIs
        q.g2 = b->d_g2 + g2_off; q.lut = nullptr;
        // the fused interpolation's row polynomials, as the single context
        // builds them (ROIs past 2^28 pixels read the image bytes there too)
        q.rp = nullptr;
        if (row_poly_fits(p.roi_x, p.roi_y, p.roi_w, p.roi_h)) {
            q.rp = b->d_rp + rp_off;
            rp_off += row_poly_bytes(p.roi_w, p.roi_h) / 16;
            b->max_rh = std::max(b->max_rh, p.roi_h);
        } else {
            b->all_rp = false;
        }
        q.refstate = b->d_state + st_off;
        q.poi = p.d_poi_yx; q.n = p.n;
        q.u = p.d_u; q.v = p.d_v; q.p6 = p.d_p6; q.iters = p.d_iters; q.flags = p.d_flags; q.dpnorm = nullptr;
        FftccParams& f = b->h_desc[i].fftcc;
        f.ref = p.d_ref; f.tar = p.d_tar; f.stride = p.stride; f.img_w = p.img_w; f.img_h = p.img_h;
        f.poi = p.d_poi_yx; f.n = p.n; f.N = p.fft_n; f.tw = b->tw[p.fft_n]; f.tol = 1e-4f;
        f.u = p.d_u; f.v = p.d_v; f.zncc = p.d_zncc ? p.d_zncc : b->d_zncc_scratch + zn_off; f.status = nullptr;
        g2_off += (size_t)p.roi_w * p.roi_h;
        st_off += (size_t)p.n * kRefStateStride;
        if (!p.d_zncc) zn_off += p.n;
        for (int k = 0; k < p.n; ++k) {
            items.push_back(make_int2(i, k));
            fitems[fclass(p.fft_n)].push_back(make_int2(i, k));
        }
    }
    if (total) TB_HIP(b, hipMemcpy(b->d_items, items.data(), sizeof(int2) * total, hipMemcpyHostToDevice));
    // flat form: FFT-CC items per kernel form, IC-GN items of 256 / lanes POIs of one pair
    std::vector<int2> mitems[4], wg[4];
    for (int i = 0; i < npairs; ++i) {
        const twdic_pair& p = b->pairs[i];
        auto& mi = mitems[mma_form(p.fft_n)];
        for (int k = 0; k < p.n; ++k) mi.push_back(make_int2(i, k));
        const int lanes = icgn_lanes_for(p.n), per_wg = 256 / lanes;
        auto& w = wg[(step_table_fits(2 * p.sx + 1, 2 * p.sy + 1) ? 0 : 1) + (lanes == 64 ? 2 : 0)];
        for (int k = 0; k < p.n; k += per_wg) w.push_back(make_int2(i, k));
    }
    for (int f = 0; f < 4; ++f) {
        b->nm[f] = (int64_t)mitems[f].size();
        if (b->nm[f] > INT32_MAX) return bfail(b, TWDIC_EINVAL, "more than 2^31 - 1 POIs in one FFT-CC kernel form");
        if (!b->nm[f]) continue;
        TB_HIP(b, alloc((void**)&b->d_mitems[f], sizeof(int2) * mitems[f].size()));
        TB_HIP(b, hipMemcpy(b->d_mitems[f], mitems[f].data(), sizeof(int2) * mitems[f].size(), hipMemcpyHostToDevice));
    }
    for (int t = 0; t < 4; ++t) {
        b->nwg[t] = (int64_t)wg[t].size();
        if (!b->nwg[t]) continue;
        TB_HIP(b, alloc((void**)&b->d_wg[t], sizeof(int2) * wg[t].size()));
        TB_HIP(b, hipMemcpy(b->d_wg[t], wg[t].data(), sizeof(int2) * wg[t].size(), hipMemcpyHostToDevice));
    }
    for (int c = 0; c < 3; ++c) {
        b->nf[c] = (int64_t)fitems[c].size();
        if (!b->nf[c]) continue;
        TB_HIP(b, alloc((void**)&b->d_fitems[c], sizeof(int2) * fitems[c].size()));
        TB_HIP(b, hipMemcpy(b->d_fitems[c], fitems[c].data(), sizeof(int2) * fitems[c].size(), hipMemcpyHostToDevice));
    }
    // every work item names a listed pair and one of its POIs (the device
    // loops trust the lists: a stray item would index past a pair's buffers)
    auto bad_item = [&](const std::vector<int2>& v, int step) {
        for (const int2& it : v)
            if (it.x < 0 || it.x >= npairs || it.y < 0 || it.y >= b->pairs[it.x].n || it.y % step) return true;
        return false;
    };
    bool bad = bad_item(items, 1) || bad_item(wg[0], 4) || bad_item(wg[1], 4);
    for (int f = 0; f < 4 && !bad; ++f) bad = bad_item(mitems[f], 1);
    for (int c = 0; c < 3 && !bad; ++c) bad = bad_item(fitems[c], 1);
    if (bad) return bfail(b, TWDIC_EINVAL, "internal: a work item names no POI of its pair");
    if (kPersistent) {
        const int Ns[3] = {64, 32, 0};
        b->pblocks_state = persistent_blocks_state(b->device);
        b->pblocks_icgn = persistent_blocks_icgn(b->device);
        for (int c = 0; c < 3; ++c)
            for (int v = 2; v <= 4; ++v) b->pblocks_fftcc[c][v] = persistent_blocks_fftcc(b->device, Ns[c], v);
    }
    b->state_valid = false;
    b->graph_dirty = true;
    return upload_desc(b);
}

}  // namespace

extern "C" {

twdic_status twdic_batch_create(int device, twdic_batch** out)
{
    if (!out) return TWDIC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) return TWDIC_ENODEV;
    twdic_batch* b = new (std::nothrow) twdic_batch();
    if (!b) return TWDIC_ENOMEM;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return TWDIC_EHIP;
    }
    if (dft_tables_init(device) != hipSuccess) {
        twdic_batch_destroy(b);
        return TWDIC_EHIP;
    }
    for (auto& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            twdic_batch_destroy(b);
            return TWDIC_EHIP;
        }
    bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < twdic_batch::kLanes && ok; ++k)
        ok = hipStreamCreateWithFlags(&b->lane[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        twdic_batch_destroy(b);
        return TWDIC_EHIP;
    }
    if (const char* v = std::getenv("TWDIC_BATCH_STREAMS")) b->lanes = std::atoi(v);
    if (const char* v = std::getenv("TWDIC_BATCH_FORM")) b->flat = std::strcmp(v, "direct") != 0;
    *out = b;
    return TWDIC_OK;
}

twdic_status twdic_batch_destroy(twdic_batch* b)
{
    if (!b) return TWDIC_EINVAL;
    (void)hipSetDevice(b->device);
    for (auto& l : b->lane)
        if (l) (void)hipStreamSynchronize(l);
    release(b);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int k = 0; k < twdic_batch::kLanes; ++k) {
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
        if (b->lane[k]) (void)hipStreamDestroy(b->lane[k]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TWDIC_OK;
}

const char* twdic_batch_last_error(const twdic_batch* b) { return b ? b->err.c_str() : "null batch"; }
void* twdic_batch_stream(twdic_batch* b) { return b ? (void*)b->stream : nullptr; }

twdic_status twdic_batch_set_pairs(twdic_batch* b, const twdic_pair* pairs, int npairs)
{
    if (!b || npairs < 0 || (npairs > 0 && !pairs)) return TWDIC_EINVAL;
    for (int i = 0; i < npairs; ++i)
        if (const char* why = check_pair(pairs[i]))
            return bfail(b, TWDIC_EINVAL, "pair " + std::to_string(i) + ": " + why);
    (void)hipSetDevice(b->device);
    for (int i = 0; i < npairs; ++i) {
        twdic_status st = check_pair_spans(b, i, pairs[i]);
        if (st != TWDIC_OK) return st;
    }
    const twdic_status st = set_pairs_impl(b, pairs, npairs);
    if (st != TWDIC_OK) clear_pairs(b);
    return st;
}


twdic_status twdic_batch_set_targets(twdic_batch* b, const uint8_t* const* d_tars)
{
    if (!b || !d_tars) return TWDIC_EINVAL;
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        if (!d_tars[i]) continue;
        twdic_pair q = b->pairs[i];
        q.d_tar = d_tars[i];
        twdic_status st = check_pair_spans(b, (int)i, q);
        if (st != TWDIC_OK) return st;
    }
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        if (!d_tars[i]) continue;
        b->pairs[i].d_tar = d_tars[i];
        b->h_desc[i].icgn.tar = d_tars[i];
        b->h_desc[i].fftcc.tar = d_tars[i];
    }
    return upload_desc(b);
}

twdic_status twdic_batch_set_references(twdic_batch* b, const uint8_t* const* d_refs)
{
    if (!b || !d_refs) return TWDIC_EINVAL;
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        if (!d_refs[i]) continue;
        twdic_pair q = b->pairs[i];
        q.d_ref = d_refs[i];
        twdic_status st = check_pair_spans(b, (int)i, q);
        if (st != TWDIC_OK) return st;
    }
    for (size_t i = 0; i < b->pairs.size(); ++i) {
        if (!d_refs[i]) continue;
        b->pairs[i].d_ref = d_refs[i];
        b->h_desc[i].icgn.ref = d_refs[i];
        b->h_desc[i].fftcc.ref = d_refs[i];
    }
    b->state_valid = false;
    return upload_desc(b);
}

twdic_status twdic_batch_run(twdic_batch* b)
{
    if (!b) return TWDIC_EINVAL;
    if (!b->total) return TWDIC_OK;
    (void)hipSetDevice(b->device);
    if (!kPersistent) return b->flat ? run_flat(b, true, true) : run_direct(b, true, true);
    if (!b->state_valid) {
        twdic_status st = build_state(b);
        if (st != TWDIC_OK) return st;
    }
    if (!b->tune.use_graph) return enqueue_frame(b, true, true);
    if (b->graph_dirty) {
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
        TB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        twdic_status st = enqueue_frame(b, true, true);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st != TWDIC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return bfail(b, TWDIC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        TB_HIP(b, hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        b->graph_dirty = false;
    }
    TB_HIP(b, hipGraphLaunch(b->graph_exec, b->stream));
    return TWDIC_OK;
}

twdic_status twdic_batch_run_icgn(twdic_batch* b)
{
    if (!b) return TWDIC_EINVAL;
    if (!b->total) return TWDIC_OK;
    (void)hipSetDevice(b->device);
    if (!kPersistent) return b->flat ? run_flat(b, false, true) : run_direct(b, false, true);
    if (!b->state_valid) {
        twdic_status st = build_state(b);
        if (st != TWDIC_OK) return st;
    }
    return enqueue_frame(b, false, true);
}

twdic_status twdic_batch_sync(twdic_batch* b)
{
    if (!b) return TWDIC_EINVAL;
    TB_HIP(b, hipStreamSynchronize(b->stream));
    return TWDIC_OK;
}

twdic_status twdic_batch_get_tuning(twdic_batch* b, twdic_batch_tuning* t)
{
    if (!b || !t) return TWDIC_EINVAL;
    *t = b->tune;
    return TWDIC_OK;
}

twdic_status twdic_batch_set_tuning(twdic_batch* b, const twdic_batch_tuning* t)
{
    if (!b || !t) return TWDIC_EINVAL;
    if (t->icgn_chunk < 1 || t->icgn_chunk > 64 || t->fftcc_variant < 2 || t->fftcc_variant > 4 ||
        (t->icgn_waves != 5 && t->icgn_waves != 6) || (t->icgn_xcd_ranges != 0 && t->icgn_xcd_ranges != 1))
        return bfail(b, TWDIC_EINVAL, "tuning out of range");
    b->tune = *t;
    b->graph_dirty = true;
    return TWDIC_OK;
}

namespace {

twdic_status append_tuning(twdic_batch* b, const char* table_path, const twdic_batch_tuning& t, float ms_per_step)
{
    if (!table_path) return TWDIC_OK;
    std::ofstream f(table_path, std::ios::app);
    if (!f) return bfail(b, TWDIC_EINVAL, std::string("cannot append to ") + table_path);
    f << signature(b) << '\t' << t.icgn_chunk << ' ' << t.fftcc_variant << ' ' << t.use_graph << ' ' << t.icgn_waves
      << ' ' << t.icgn_xcd_ranges << ' ' << ms_per_step << '\n';
    return TWDIC_OK;
}

}  // namespace

twdic_status twdic_batch_autotune(twdic_batch* b, const char* table_path, int reps)
{
    if (!b || reps < 1) return TWDIC_EINVAL;
    if (!kPersistent && !b->flat) return bfail(b, TWDIC_EINVAL, "the direct batched form has nothing to tune");
    if (!b->total) return TWDIC_OK;
    auto time_cfg = [&](const twdic_batch_tuning& t, float* ms) -> twdic_status {
        twdic_status st = twdic_batch_set_tuning(b, &t);
        if (st == TWDIC_OK) st = twdic_batch_run(b);                 // warm-up (and graph build)
        if (st != TWDIC_OK) return st;
        TB_HIP(b, hipEventRecord(b->ev[14], b->stream));
        for (int r = 0; r < reps && st == TWDIC_OK; ++r) st = twdic_batch_run(b);
        if (st != TWDIC_OK) return st;
        TB_HIP(b, hipEventRecord(b->ev[15], b->stream));
        TB_HIP(b, hipEventSynchronize(b->ev[15]));
        TB_HIP(b, hipEventElapsedTime(ms, b->ev[14], b->ev[15]));
        return TWDIC_OK;
    };
    twdic_batch_tuning best = b->tune;
    float best_ms = 1e30f;
    if (!kPersistent) {
        // flat form: IC-GN occupancy target x workgroup-to-XCD order
        for (int waves : {5, 6})
            for (int xcd : {0, 1}) {
                twdic_batch_tuning t = best;
                t.icgn_waves = waves;
                t.icgn_xcd_ranges = xcd;
                float ms;
                twdic_status st = time_cfg(t, &ms);
                if (st != TWDIC_OK) return st;
                if (ms < best_ms) { best_ms = ms; best = t; }
            }
        twdic_status st = twdic_batch_set_tuning(b, &best);
        if (st != TWDIC_OK) return st;
        return append_tuning(b, table_path, best, best_ms / reps);
    }
    // stage 1: FFT-CC occupancy target (only matters with N = 64 items)
    const int variants[3] = {2, 3, 4};
    for (int v : variants) {
        if (!b->nf[0] && v != best.fftcc_variant) continue;
        twdic_batch_tuning t = best;
        t.fftcc_variant = v;
        float ms;
        twdic_status st = time_cfg(t, &ms);
        if (st != TWDIC_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    // stage 2: IC-GN work-queue grab size
    const int chunks[5] = {1, 2, 4, 8, 16};
    for (int c : chunks) {
        twdic_batch_tuning t = best;
        t.icgn_chunk = c;
        float ms;
        twdic_status st = time_cfg(t, &ms);
        if (st != TWDIC_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    twdic_status st = twdic_batch_set_tuning(b, &best);
    if (st != TWDIC_OK) return st;
    return append_tuning(b, table_path, best, best_ms / reps);
}

twdic_status twdic_batch_load_tuning(twdic_batch* b, const char* table_path, int* found)
{
    if (!b || !table_path) return TWDIC_EINVAL;
    if (found) *found = 0;
    std::ifstream f(table_path);
    if (!f) return TWDIC_OK;
    const std::string sig = signature(b);
    std::string line;
    twdic_batch_tuning t = b->tune;
    bool hit = false;
    while (std::getline(f, line)) {        // last matching entry wins
        auto tab = line.find('\t');
        if (tab == std::string::npos || line.compare(0, tab, sig) != 0 || tab != sig.size()) continue;
        std::istringstream is(line.substr(tab + 1));
        twdic_batch_tuning c = b->tune;
        if (is >> c.icgn_chunk >> c.fftcc_variant >> c.use_graph) {
            // the flat-form fields follow; lines without them keep the defaults
            int w = 0, x = 0;
            if (is >> w >> x) { c.icgn_waves = w; c.icgn_xcd_ranges = x; }
            t = c;
            hit = true;
        }
    }
    if (hit) {
        twdic_status st = twdic_batch_set_tuning(b, &t);
        if (st != TWDIC_OK) return st;
    }
    if (found) *found = hit ? 1 : 0;
    return TWDIC_OK;
}

twdic_status twdic_batch_info_get(twdic_batch* b, twdic_batch_info* info)
{
    if (!b || !info) return TWDIC_EINVAL;
    info->npairs = (int)b->pairs.size();
    info->total_pois = b->total;
    for (int c = 0; c < 3; ++c) info->fftcc_items[c] = b->nf[c];
    info->device_bytes = b->bytes;
    info->tuning = b->tune;
    return TWDIC_OK;
}

twdic_status twdic_batch_record(twdic_batch* b, int slot)
{
    if (!b || slot < 0 || slot >= 14) return TWDIC_EINVAL;
    TB_HIP(b, hipEventRecord(b->ev[slot], b->stream));
    return TWDIC_OK;
}

twdic_status twdic_batch_elapsed(twdic_batch* b, int a, int c, float* ms)
{
    if (!b || !ms || a < 0 || c < 0 || a >= 14 || c >= 14) return TWDIC_EINVAL;
    TB_HIP(b, hipEventSynchronize(b->ev[c]));
    TB_HIP(b, hipEventElapsedTime(ms, b->ev[a], b->ev[c]));
    return TWDIC_OK;
}

}  // extern "C"
