// serve.cpp -- the serving loop (include/twdic_serve.h): a bounded admission
// queue, device slots (one HIP stream each) filled as they and device memory
// free up, resident references / POI sets / per-POI reference state in one
// device arena, and a dispatcher thread that issues each admitted request's
// stage sequence and reaps completions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "twdic_internal.h"
#include "twdic_serve.h"

using namespace twdic;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation, free ranges coalesced.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;   // offset -> size

    void init(char* b, size_t c)
    {
        base = b;
        cap = c;
        free_.clear();
        free_[0] = c;
    }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(bytes ? bytes : 1);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct RefEntry {
    int id = 0;
    uint64_t gen = 0;
    int w = 0, h = 0;
    size_t off = 0, bytes = 0;       // image (w*h, aligned) then g2 (4 (w-2)(h-2))
    size_t g2_rel = 0;
    hipEvent_t ready = nullptr;
    int users = 0;                   // queued + in-flight requests
    bool retired = false;
};

struct PoiSet {
    int n = 0;
    size_t off = 0;
};

struct StateEntry {
    uint64_t gen = 0;
    int poiset = 0, S = 0;
    size_t off = 0, bytes = 0;
    hipEvent_t ready = nullptr;
    int inflight = 0;
    uint64_t last_use = 0;
};

struct Job {
    uint64_t ticket = 0;
    twdic_request r{};
    RefEntry* ref = nullptr;
    Clock::time_point t_submit;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    Job job;
    StateEntry* state = nullptr;
    size_t off = 0, bytes = 0;       // target + outputs in the arena
    int n = 0;
    uint8_t* h_tar = nullptr;        // pinned staging
    char* h_out = nullptr;
};

}  // namespace

struct twdic_server {
    twdic_server_config cfg{};
    std::string err;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space;
    std::thread worker;
    bool stop = false;
    hipStream_t ctrl = nullptr;
    char* d_arena = nullptr;
    Arena arena;
    std::map<int, RefEntry*> refs;                 // current generation per id
    std::vector<RefEntry*> retired;
    uint64_t next_gen = 1;
    std::vector<PoiSet> poisets;
    std::map<std::tuple<uint64_t, int, int>, StateEntry*> states;
    uint64_t use_clock = 0;
    std::map<int, float2*> twiddles;               // per FFT window edge
    std::deque<Job> queue;
    std::vector<Slot> slots;
    std::unordered_map<uint64_t, int> ticket_state;  // 0 pending, 1 done, -1 failed
    uint64_t next_ticket = 1;
    // statistics
    twdic_server_stats st{};
    std::vector<double> lat_ms;
    Clock::time_point t0;
    int busy_now = 0;
};

namespace {

size_t out_bytes(int n) { return (size_t)n * 41; }   // u v zncc p6 iters (4 B each x 11) + flags

void free_ref(twdic_server* s, RefEntry* r)
{
    s->arena.release(r->off, r->bytes);
    if (r->ready) (void)hipEventDestroy(r->ready);
    delete r;
}

// Drops retired references nobody uses and the states built on them.
void collect(twdic_server* s)
{
    for (auto it = s->states.begin(); it != s->states.end();) {
        StateEntry* e = it->second;
        bool live = false;
        for (auto& kv : s->refs)
            if (kv.second->gen == e->gen) live = true;
        if (!live && e->inflight == 0) {
            s->arena.release(e->off, e->bytes);
            if (e->ready) (void)hipEventDestroy(e->ready);
            delete e;
            it = s->states.erase(it);
        } else {
            ++it;
        }
    }
    for (auto it = s->retired.begin(); it != s->retired.end();) {
        RefEntry* r = *it;
        bool has_state = false;
        for (auto& kv : s->states)
            if (kv.second->gen == r->gen) has_state = true;
        if (r->users == 0 && !has_state) {
            free_ref(s, r);
            it = s->retired.erase(it);
        } else {
            ++it;
        }
    }
}

// Arena allocation, evicting resident per-POI state that no request uses
// (least recently used first) until the block fits.
bool alloc_evict(twdic_server* s, size_t bytes, size_t* off, const StateEntry* keep)
{
    while (!s->arena.alloc(bytes, off)) {
        StateEntry* victim = nullptr;
        auto vit = s->states.end();
        for (auto it = s->states.begin(); it != s->states.end(); ++it) {
            StateEntry* e = it->second;
            if (e == keep || e->inflight) continue;
            if (!victim || e->last_use < victim->last_use) {
                victim = e;
                vit = it;
            }
        }
        if (!victim) return false;
        // an unused entry may still be read by nothing: its builder and every
        // reader completed (inflight counts them until reaped)
        s->arena.release(victim->off, victim->bytes);
        if (victim->ready) (void)hipEventDestroy(victim->ready);
        delete victim;
        s->states.erase(vit);
    }
    return true;
}

const float2* twiddles(twdic_server* s, int N)
{
    auto it = s->twiddles.find(N);
    if (it != s->twiddles.end()) return it->second;
    float2 t[kMaxFftN];
    for (int k = 0; k < N; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)N;
        t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(float2) * kMaxFftN) != hipSuccess) return nullptr;
    if (hipMemcpy(d, t, sizeof(float2) * N, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    s->twiddles[N] = d;
    return d;
}

void finish(twdic_server* s, Slot& sl, bool ok)
{
    Job& j = sl.job;
    if (ok) {
        const int n = sl.n;
        const char* h = sl.h_out;
        const twdic_request& r = j.r;
        std::memcpy(r.u, h, 4 * (size_t)n);
        std::memcpy(r.v, h + 4 * (size_t)n, 4 * (size_t)n);
        if (r.zncc) std::memcpy(r.zncc, h + 8 * (size_t)n, 4 * (size_t)n);
        if (r.p6) std::memcpy(r.p6, h + 12 * (size_t)n, 24 * (size_t)n);
        if (r.iters) std::memcpy(r.iters, h + 36 * (size_t)n, 4 * (size_t)n);
        if (r.flags) std::memcpy(r.flags, h + 40 * (size_t)n, (size_t)n);
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - j.t_submit).count();
        s->lat_ms.push_back(ms);
        s->st.completed++;
        s->st.pois += n;
    } else {
        s->st.failed++;
    }
    s->ticket_state[j.ticket] = ok ? 1 : -1;
    s->cv_done.notify_all();
    s->arena.release(sl.off, sl.bytes);
    if (sl.state) sl.state->inflight--;
    j.ref->users--;
    sl.state = nullptr;
    sl.busy = false;
    s->busy_now--;
}

// Issues one request on a free slot.  Returns false (nothing changed) when
// device memory is short; the request then waits at the head of the queue.
bool admit(twdic_server* s, Slot& sl, Job& j)
{
    const twdic_request& r = j.r;
    RefEntry* ref = j.ref;
    const PoiSet& ps = s->poisets[r.poiset];
    const int n = ps.n, w = ref->w, h = ref->h;
    const auto key = std::make_tuple(ref->gen, r.poiset, r.subset);
    StateEntry* e = nullptr;
    auto sit = s->states.find(key);
    bool build = false;
    if (sit != s->states.end()) {
        e = sit->second;
        s->st.state_hits++;
    } else {
        size_t off;
        const size_t bytes = (size_t)n * kRefStateStride * sizeof(double);
        if (!alloc_evict(s, bytes, &off, nullptr)) return false;
        e = new StateEntry();
        e->gen = ref->gen;
        e->poiset = r.poiset;
        e->S = r.subset;
        e->off = off;
        e->bytes = bytes;
        if (hipEventCreateWithFlags(&e->ready, hipEventDisableTiming) != hipSuccess) {
            s->arena.release(off, bytes);
            delete e;
            return false;
        }
        s->states[key] = e;
        build = true;
        s->st.state_builds++;
    }
    // the frame, its row polynomials (the fused interpolation's, as the
    // single context builds them) and the outputs
    const size_t tar_b = align_up((size_t)w * h);
    const bool rp_fits = row_poly_fits(1, 1, w - 2, h - 2);
    const size_t rp_b = rp_fits ? align_up(row_poly_bytes(w - 2, h - 2)) : 0;
    size_t off;
    if (!alloc_evict(s, tar_b + rp_b + out_bytes(n), &off, e)) {
        if (build) {      // undo: nothing was issued for it yet
            s->states.erase(key);
            s->arena.release(e->off, e->bytes);
            (void)hipEventDestroy(e->ready);
            delete e;
        }
        return false;
    }
    e->inflight++;
    e->last_use = ++s->use_clock;
    sl.busy = true;
    sl.job = j;
    sl.state = e;
    sl.off = off;
    sl.bytes = tar_b + rp_b + out_bytes(n);
    sl.n = n;
    s->busy_now++;
    s->st.admitted++;
    s->st.slots_busy_peak = std::max(s->st.slots_busy_peak, s->busy_now);

    char* base = s->d_arena;
    uint8_t* d_tar = (uint8_t*)(base + off);
    uint4* d_rp = rp_fits ? (uint4*)(base + off + tar_b) : nullptr;
    char* d_out = base + off + tar_b + rp_b;
    float* d_u = (float*)d_out;
    float* d_v = d_u + n;
    float* d_z = d_v + n;
    float* d_p6 = d_z + n;
    int32_t* d_it = (int32_t*)(d_p6 + 6 * (size_t)n);
    uint8_t* d_fl = (uint8_t*)(d_it + n);
    const uint8_t* d_ref = (const uint8_t*)(base + ref->off);
    const uint32_t* d_g2 = (const uint32_t*)(base + ref->off + ref->g2_rel);
    const int32_t* d_poi = (const int32_t*)(base + ps.off);
    double* d_state = (double*)(base + e->off);

    for (int y = 0; y < h; ++y) std::memcpy(sl.h_tar + (size_t)y * w, r.target + (size_t)y * r.stride, w);
    hipError_t err = hipSuccess;
    auto chk = [&](hipError_t x) {
        if (err == hipSuccess) err = x;
    };
    chk(hipMemcpyAsync(d_tar, sl.h_tar, (size_t)w * h, hipMemcpyHostToDevice, sl.st));
    chk(hipMemsetAsync(d_out, 0, out_bytes(n), sl.st));
    chk(hipStreamWaitEvent(sl.st, ref->ready, 0));

    IcgnParams ip{};
    ip.ref = d_ref; ip.tar = d_tar; ip.stride = w;
    ip.sx0 = 1; ip.sy0 = 1; ip.rw = w - 2; ip.rh = h - 2;
    ip.Sx = ip.Sy = r.subset; ip.Wsub = ip.Hsub = 2 * r.subset + 1; ip.s = ip.Wsub * ip.Hsub;
    ip.max_iter = r.max_iter; ip.delta_p = r.delta_p;
    ip.g2 = d_g2; ip.lut = nullptr; ip.rp = d_rp; ip.refstate = d_state;
    ip.poi = d_poi; ip.n = n; ip.lanes = icgn_lanes_for(n);
    ip.u = d_u; ip.v = d_v; ip.p6 = d_p6; ip.iters = d_it; ip.flags = d_fl; ip.dpnorm = nullptr;
    if (build) {
        chk(launch_ref_state(ip, sl.st));
        chk(hipEventRecord(e->ready, sl.st));
    } else {
        chk(hipStreamWaitEvent(sl.st, e->ready, 0));
    }
    FftccParams fp{};
    fp.ref = d_ref; fp.tar = d_tar; fp.stride = w; fp.img_w = w; fp.img_h = h;
    fp.poi = d_poi; fp.n = n; fp.N = r.fft_n; fp.tw = twiddles(s, r.fft_n); fp.tol = 1e-4f;
    fp.u = d_u; fp.v = d_v; fp.zncc = d_z; fp.status = nullptr;
    if (!fp.tw) chk(hipErrorOutOfMemory);
    if (err == hipSuccess) chk(launch_fftcc(fp, kFftccDftMfma, sl.st));
    if (err == hipSuccess && d_rp) chk(launch_row_poly(d_tar, w, 1, 1, w - 2, h - 2, d_rp, sl.st));
    if (err == hipSuccess) chk(launch_icgn(ip, TWDIC_INTERP_FUSED, 5, sl.st));
    chk(hipMemcpyAsync(sl.h_out, d_out, out_bytes(n), hipMemcpyDeviceToHost, sl.st));
    chk(hipEventRecord(sl.done, sl.st));
    if (err != hipSuccess) {
        // the stream may still hold work on these buffers: wait before releasing them
        (void)hipStreamSynchronize(sl.st);
        s->err = std::string("request issue: ") + hipGetErrorString(err);
        finish(s, sl, false);
    }
    return true;
}

// One thread issues and reaps everything, holding the server mutex across
// every runtime call; reference uploads hold the same mutex, so no two
// threads are ever inside the HIP runtime at once.
void dispatcher(twdic_server* s)
{
    (void)hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // reap completions
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            finish(s, sl, q == hipSuccess);
        }
        collect(s);
        // admit in arrival order while slots and memory allow
        for (Slot& sl : s->slots) {
            if (s->queue.empty()) break;
            if (sl.busy) continue;
            if (!admit(s, sl, s->queue.front())) {
                if (s->busy_now == 0) {
                    // nothing in flight and every unused block evicted: the
                    // request can never fit the budget
                    Job& j = s->queue.front();
                    s->ticket_state[j.ticket] = -1;
                    j.ref->users--;
                    s->st.failed++;
                    s->err = "request exceeds the device budget";
                    s->queue.pop_front();
                    s->cv_space.notify_all();
                    s->cv_done.notify_all();
                    continue;
                }
                break;
            }
            s->queue.pop_front();
            s->cv_space.notify_all();
        }
        if (s->stop && s->busy_now == 0) break;
        if (s->busy_now == 0 && s->queue.empty()) {
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        } else {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

twdic_status sfail(twdic_server* s, twdic_status st, const std::string& what)
{
    if (s) s->err = what;
    return st;
}

}  // namespace

extern "C" {

twdic_status twdic_server_create(const twdic_server_config* cfg, twdic_server** out)
{
    if (!cfg || !out) return TWDIC_EINVAL;
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 64 || cfg->queue_capacity < 1 || cfg->max_width < 4 ||
        cfg->max_height < 4 || cfg->max_pois < 1 || cfg->max_width > 65535 || cfg->max_height > 65535 ||
        (int64_t)cfg->max_width * cfg->max_height >= (int64_t)1 << 31)
        return TWDIC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || cfg->device < 0 || cfg->device >= ndev)
        return TWDIC_ENODEV;
    twdic_server* s = new (std::nothrow) twdic_server();
    if (!s) return TWDIC_ENOMEM;
    s->cfg = *cfg;
    auto bad = [&](twdic_status st) {
        twdic_server_destroy(s);
        return st;
    };
    if (hipSetDevice(cfg->device) != hipSuccess || dft_tables_init(cfg->device) != hipSuccess) return bad(TWDIC_EHIP);
    size_t budget = cfg->hbm_budget;
    if (!budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return bad(TWDIC_EHIP);
        budget = fr / 4;
    }
    budget = align_up(budget);
    if (hipMalloc((void**)&s->d_arena, budget) != hipSuccess) return bad(TWDIC_ENOMEM);
    s->arena.init(s->d_arena, budget);
    s->st.hbm_budget = budget;
    if (hipStreamCreateWithFlags(&s->ctrl, hipStreamNonBlocking) != hipSuccess) return bad(TWDIC_EHIP);
    s->slots.resize(cfg->slots);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_tar, (size_t)cfg->max_width * cfg->max_height, hipHostMallocDefault) !=
                hipSuccess ||
            hipHostMalloc((void**)&sl.h_out, out_bytes(cfg->max_pois), hipHostMallocDefault) != hipSuccess)
            return bad(TWDIC_ENOMEM);
    }
    s->t0 = Clock::now();
    s->worker = std::thread(dispatcher, s);
    *out = s;
    return TWDIC_OK;
}

twdic_status twdic_server_destroy(twdic_server* s)
{
    if (!s) return TWDIC_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Job& j : s->queue) {
            s->ticket_state[j.ticket] = -1;
            j.ref->users--;
            s->st.dropped++;
        }
        s->queue.clear();
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_space.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->cfg.device);
    for (Slot& sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.h_tar) (void)hipHostFree(sl.h_tar);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
    }
    for (auto& kv : s->states) {
        if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
        delete kv.second;
    }
    for (auto& kv : s->refs) {
        if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
        delete kv.second;
    }
    for (RefEntry* r : s->retired) {
        if (r->ready) (void)hipEventDestroy(r->ready);
        delete r;
    }
    for (auto& kv : s->twiddles) (void)hipFree(kv.second);
    if (s->ctrl) (void)hipStreamDestroy(s->ctrl);
    if (s->d_arena) (void)hipFree(s->d_arena);
    s->cv_done.notify_all();
    delete s;
    return TWDIC_OK;
}

const char* twdic_server_last_error(const twdic_server* s) { return s ? s->err.c_str() : "null server"; }

twdic_status twdic_server_set_reference(twdic_server* s, int ref_id, const uint8_t* img, int w, int h, int stride)
{
    if (!s || !img || w < 4 || h < 4 || w > s->cfg.max_width || h > s->cfg.max_height || stride < w)
        return sfail(s, TWDIC_EINVAL, "set_reference: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    (void)hipSetDevice(s->cfg.device);
    RefEntry* r = new RefEntry();
    r->id = ref_id;
    r->gen = s->next_gen++;
    r->w = w;
    r->h = h;
    r->g2_rel = align_up((size_t)w * h);
    r->bytes = r->g2_rel + (size_t)4 * (w - 2) * (h - 2);
    if (!alloc_evict(s, r->bytes, &r->off, nullptr)) {
        delete r;
        return sfail(s, TWDIC_ENOMEM, "set_reference: device budget exhausted");
    }
    uint8_t* d_img = (uint8_t*)(s->d_arena + r->off);
    uint32_t* d_g2 = (uint32_t*)(s->d_arena + r->off + r->g2_rel);
    hipError_t e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemcpy2DAsync(d_img, w, img, stride, w, h, hipMemcpyHostToDevice, s->ctrl);
    if (e == hipSuccess) e = launch_gradients(d_img, w, 1, 1, w - 2, h - 2, nullptr, nullptr, nullptr, d_g2, s->ctrl);
    if (e == hipSuccess) e = hipEventRecord(r->ready, s->ctrl);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctrl);     // the caller's image may go away
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->ctrl);
        free_ref(s, r);
        return sfail(s, TWDIC_EHIP, std::string("set_reference: ") + hipGetErrorString(e));
    }
    auto it = s->refs.find(ref_id);
    if (it != s->refs.end()) {
        it->second->retired = true;
        s->retired.push_back(it->second);
    }
    s->refs[ref_id] = r;
    collect(s);
    return TWDIC_OK;
}

twdic_status twdic_server_add_poiset(twdic_server* s, const int32_t* poi, int n, int* id)
{
    if (!s || !poi || !id || n < 1 || n > s->cfg.max_pois) return sfail(s, TWDIC_EINVAL, "add_poiset: bad arguments");
    std::lock_guard<std::mutex> lk(s->mu);
    (void)hipSetDevice(s->cfg.device);
    PoiSet p;
    p.n = n;
    if (!alloc_evict(s, (size_t)8 * n, &p.off, nullptr)) return sfail(s, TWDIC_ENOMEM, "add_poiset: device budget exhausted");
    hipError_t e = hipMemcpyAsync(s->d_arena + p.off, poi, (size_t)8 * n, hipMemcpyHostToDevice, s->ctrl);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctrl);
    if (e != hipSuccess) {
        s->arena.release(p.off, (size_t)8 * n);
        return sfail(s, TWDIC_EHIP, std::string("add_poiset: ") + hipGetErrorString(e));
    }
    s->poisets.push_back(p);
    *id = (int)s->poisets.size() - 1;
    return TWDIC_OK;
}

twdic_status twdic_server_submit(twdic_server* s, const twdic_request* r, int drop_if_full, uint64_t* ticket)
{
    if (!s || !r || !ticket) return TWDIC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto rit = s->refs.find(r->ref_id);
    if (rit == s->refs.end()) return sfail(s, TWDIC_EINVAL, "submit: unknown reference");
    if (r->poiset < 0 || r->poiset >= (int)s->poisets.size()) return sfail(s, TWDIC_EINVAL, "submit: unknown POI set");
    if (!r->target || r->stride < rit->second->w || !r->u || !r->v || r->subset < 1 || r->subset > 60 ||
        r->fft_n < 2 || r->fft_n > kMaxFftN || (r->fft_n & 1) || r->max_iter < 1 || !(r->delta_p >= 0.f) ||
        r->memory != TWDIC_MEM_PAGEABLE)
        return sfail(s, TWDIC_EINVAL, "submit: bad request");
    if (s->stop) return sfail(s, TWDIC_EINVAL, "submit: server stopping");
    if ((int)s->queue.size() >= s->cfg.queue_capacity) {
        if (drop_if_full) {
            s->st.submitted++;
            s->st.dropped++;
            return TWDIC_EBUSY;
        }
        s->cv_space.wait(lk, [&] { return s->stop || (int)s->queue.size() < s->cfg.queue_capacity; });
        if (s->stop) return sfail(s, TWDIC_EINVAL, "submit: server stopping");
        rit = s->refs.find(r->ref_id);
        if (rit == s->refs.end()) return sfail(s, TWDIC_EINVAL, "submit: unknown reference");
    }
    Job j;
    j.ticket = s->next_ticket++;
    j.r = *r;
    j.ref = rit->second;
    j.ref->users++;
    j.t_submit = Clock::now();
    s->queue.push_back(j);
    s->ticket_state[j.ticket] = 0;
    s->st.submitted++;
    *ticket = j.ticket;
    lk.unlock();
    s->cv_work.notify_one();
    return TWDIC_OK;
}

twdic_status twdic_server_poll(twdic_server* s, uint64_t ticket, int* state)
{
    if (!s || !state) return TWDIC_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->ticket_state.find(ticket);
    if (it == s->ticket_state.end()) return sfail(s, TWDIC_EINVAL, "poll: unknown ticket");
    *state = it->second;
    if (it->second != 0) s->ticket_state.erase(it);
    return TWDIC_OK;
}

twdic_status twdic_server_wait(twdic_server* s, uint64_t ticket)
{
    if (!s) return TWDIC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->ticket_state.find(ticket);
    if (it == s->ticket_state.end()) return sfail(s, TWDIC_EINVAL, "wait: unknown ticket");
    s->cv_done.wait(lk, [&] {
        auto i = s->ticket_state.find(ticket);
        return i == s->ticket_state.end() || i->second != 0;
    });
    it = s->ticket_state.find(ticket);
    const int v = it == s->ticket_state.end() ? -1 : it->second;
    if (it != s->ticket_state.end()) s->ticket_state.erase(it);
    return v == 1 ? TWDIC_OK : TWDIC_EHIP;
}

twdic_status twdic_server_drain(twdic_server* s)
{
    if (!s) return TWDIC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->busy_now == 0; });
    return TWDIC_OK;
}

twdic_status twdic_server_stats_get(twdic_server* s, twdic_server_stats* out)
{
    if (!s || !out) return TWDIC_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    twdic_server_stats st = s->st;
    std::vector<double> l = s->lat_ms;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {
            const double k = p / 100.0 * (double)(l.size() - 1);
            const size_t a = (size_t)std::floor(k), b = std::min(l.size() - 1, a + 1);
            return l[a] + (l[b] - l[a]) * (k - (double)a);
        };
        st.lat_p50_ms = pct(50);
        st.lat_p99_ms = pct(99);
        st.lat_max_ms = l.back();
    }
    st.wall_s = std::chrono::duration<double>(Clock::now() - s->t0).count();
    st.hbm_in_use = s->arena.used;
    st.hbm_peak = s->arena.peak;
    *out = st;
    return TWDIC_OK;
}

twdic_status twdic_server_stats_reset(twdic_server* s)
{
    if (!s) return TWDIC_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    const size_t budget = s->st.hbm_budget;
    s->st = twdic_server_stats{};
    s->st.hbm_budget = budget;
    s->lat_ms.clear();
    s->arena.peak = s->arena.used;
    s->t0 = Clock::now();
    return TWDIC_OK;
}

}  // extern "C"
