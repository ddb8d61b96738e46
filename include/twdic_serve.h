/*
 * twdic_serve.h -- serving loop for the paDIC path: correlation requests
 * arrive as frames do, are admitted into concurrent device slots as slots and
 * device memory free up (continuous batching: work of different requests
 * overlaps on the device), and complete asynchronously.  References and
 * per-POI reference state (the FP64 Cholesky factors) stay resident in HBM
 * across frames; admission is by device-memory footprint.
 *
 * No counterpart in the reference, which runs one frame at a time per worker
 * thread (FFTCCTWorkerThread::processFrameFFTCC_ICGN,
 * TW/TW_Core/fftcctworkerthread.cpp:399-572) fed by bounded frame queues
 * (Concurrent_Buffer::EnQueue/DeQueue, TW/TW_Engine/TW_Concurrent_Buffer.cpp:
 * 22-73, drop-if-full by default).  The admission queue here keeps those
 * queue semantics.  Each request runs the same device code as the
 * single-context API (include/twdic.h), so its results are bit-identical to
 * it, alone or in any mix.
 */
#ifndef TWDIC_SERVE_H
#define TWDIC_SERVE_H

#include "twdic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct twdic_server twdic_server;

typedef struct {
    int device;
    int slots;              /* requests in flight at once, one HIP stream each (1..64) */
    int queue_capacity;     /* admission queue length (Concurrent_Buffer size) */
    size_t hbm_budget;      /* device bytes for references, resident state and in-flight
                               requests (a w x h request holds its frame, 16 (w - 2) (h - 1)
                               bytes of row polynomials and 41 bytes per POI);
                               0 = a quarter of the free device memory */
    int max_width;          /* largest frame accepted */
    int max_height;
    int max_pois;           /* largest POI set accepted */
} twdic_server_config;

/* Where a request's target and output arrays live.  The server takes host
 * memory only (copied through each slot's pinned staging); the other kinds
 * are rejected with TWDIC_EINVAL. */
enum {
    TWDIC_MEM_PAGEABLE = 0,
    TWDIC_MEM_PINNED = 1,
    TWDIC_MEM_DEVICE = 2
};

typedef struct {
    int ref_id;             /* reference set by twdic_server_set_reference */
    int poiset;             /* POI set from twdic_server_add_poiset */
    const uint8_t* target;  /* frame, same size as the reference (where: see memory) */
    int stride;             /* bytes per target row */
    int subset;             /* subset half size S (subset 2S+1 square) */
    int fft_n;              /* FFT-CC window edge (even, <= 64; 32/64 use the fast engine) */
    int max_iter;
    float delta_p;
    /* outputs, n entries each (p6: 6n); u, v required, others may be NULL */
    float* u;
    float* v;
    float* zncc;
    float* p6;
    int32_t* iters;
    uint8_t* flags;
    int memory;             /* TWDIC_MEM_PAGEABLE; the outputs must stay valid until the
                               request is done, the frame is copied at admission */
} twdic_request;

typedef struct {
    uint64_t submitted, admitted, completed, dropped, failed;
    double pois;                     /* POIs of completed requests */
    double lat_p50_ms, lat_p99_ms, lat_max_ms;   /* submit -> results in the caller's buffers */
    double wall_s;                   /* since creation or the last stats reset */
    size_t hbm_budget, hbm_in_use, hbm_peak;
    int slots_busy_peak;
    uint64_t state_builds, state_hits;
} twdic_server_stats;

twdic_status twdic_server_create(const twdic_server_config* cfg, twdic_server** out);
/* Completes every admitted request, drops the queued ones, frees everything. */
twdic_status twdic_server_destroy(twdic_server* s);
const char* twdic_server_last_error(const twdic_server* s);

/* Sets (or replaces) reference ref_id.  Requests already submitted keep the
 * reference they were submitted against; later ones use the new one.  The
 * image is w x h (w, h within the configured maximum); the ROI is the image
 * minus a one-pixel frame. */
twdic_status twdic_server_set_reference(twdic_server* s, int ref_id, const uint8_t* img, int w, int h, int stride);
/* Registers a POI set (host int32[n][2] = {y, x}); it stays resident. */
twdic_status twdic_server_add_poiset(twdic_server* s, const int32_t* poi_yx, int n, int* poiset_id);

/* Queues a request.  When the admission queue is full: drop_if_full != 0
 * drops it (TWDIC_EBUSY, counted as dropped), otherwise the call blocks until
 * the queue has room.  *ticket identifies the request in poll/wait. */
twdic_status twdic_server_submit(twdic_server* s, const twdic_request* r, int drop_if_full, uint64_t* ticket);
/* *state: 0 queued or running, 1 done (outputs written), -1 failed.  A ticket
 * that reached 1 or -1 is forgotten after it has been reported once. */
twdic_status twdic_server_poll(twdic_server* s, uint64_t ticket, int* state);
/* Blocks until the request is done (TWDIC_OK) or failed (its error). */
twdic_status twdic_server_wait(twdic_server* s, uint64_t ticket);
/* Blocks until every submitted request has finished. */
twdic_status twdic_server_drain(twdic_server* s);

twdic_status twdic_server_stats_get(twdic_server* s, twdic_server_stats* st);
/* Clears the latency record and counters (not the resident state). */
twdic_status twdic_server_stats_reset(twdic_server* s);

#ifdef __cplusplus
}
#endif
#endif
