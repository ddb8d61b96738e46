/*
 * twdic_batch.h -- batched form of the paDIC hot path: many image pairs of
 * uneven sizes, mixed subset radii and FFT-CC windows behind one handle, with
 * resident per-pair reference state.
 *
 * Executions.  The default is the flat form: one launch per stage and
 * matrix-core kernel form over the POIs of every pair, each block running the
 * single-context per-POI code for its (pair, POI) item.
 * TWDIC_BATCH_FORM=direct selects the direct form: per pair, the
 * single-context kernels, round-robin on four streams.  Both give results
 * bit-identical to the single-context API.  The persistent
 * form (kernels that drain device work queues, one launch per stage and window
 * class, hipGraph replay, launch parameters chosen by an autotuner) is compiled
 * only with -DTWDIC_BATCH_EXEC: its first run on an MI355X stalled and the
 * cause has not been found.  In the default build the flat form reads
 * use_graph (replay of its whole step, FFT-CC and IC-GN, from a hipGraph
 * captured on the first run after new pairs or new tuning), icgn_waves and
 * icgn_xcd_ranges; twdic_batch_autotune tunes those two for the flat form and
 * returns TWDIC_EINVAL for the direct form.
 *
 * No counterpart in the reference, which runs one image pair per
 * cuFFTCC2D/cuICGN2D object and one CUDA block per POI
 * (TW_paDIC_cuFFTCC2D.cu:575-638, TW_paDIC_cuICGN2D.cu:927-945).  Every
 * per-POI result is computed by the same device code as the single-context
 * API (include/twdic.h) and is bit-identical to it: POIs do not interact.
 */
#ifndef TWDIC_BATCH_H
#define TWDIC_BATCH_H

#include "twdic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct twdic_batch twdic_batch;

/* One image pair.  Images, POIs and outputs are device buffers owned by the
 * caller (borrowed for the lifetime of the batch).  Subsets are
 * (2 sy + 1) x (2 sx + 1); fft_n is any even value up to 64 (the flat and
 * direct forms run every window size on the int8 matrix cores, as the
 * single-context API does).  Any output may be NULL except
 * d_u/d_v.  IC-GN interpolation is the fused bicubic evaluation. */
typedef struct {
    const uint8_t* d_ref;
    const uint8_t* d_tar;
    int stride;
    int img_w, img_h;
    int roi_x, roi_y, roi_w, roi_h;
    int sx, sy;
    int fft_n;
    int max_iter;
    float delta_p;
    const int32_t* d_poi_yx;   /* int32[n][2] = {y, x} */
    int n;
    float* d_u;                /* out: FFT-CC integer guess, then IC-GN result */
    float* d_v;
    float* d_zncc;
    float* d_p6;
    int32_t* d_iters;
    uint8_t* d_flags;
} twdic_pair;

/* Launch parameters the autotuner searches. */
typedef struct {
    int icgn_chunk;     /* POIs per IC-GN work-queue grab (1..64) */
    int fftcc_variant;  /* occupancy target (waves/SIMD) of the N = 64 engine: 2, 3 or 4 */
    int use_graph;      /* 1: replay the per-frame stage sequence as a hipGraph (flat and persistent forms) */
    int icgn_waves;     /* flat form: occupancy target (waves/SIMD) of the IC-GN kernel: 5 or 6 */
    int icgn_xcd_ranges;/* flat form: 1 gives each XCD a contiguous range of IC-GN workgroup items,
                           0 deals them out in block order */
} twdic_batch_tuning;

typedef struct {
    int npairs;
    int64_t total_pois;
    int64_t fftcc_items[3];    /* POIs per window class: N = 64, N = 32, other */
    size_t device_bytes;       /* batch-owned device memory */
    twdic_batch_tuning tuning;
} twdic_batch_info;

twdic_status twdic_batch_create(int device, twdic_batch** out);
twdic_status twdic_batch_destroy(twdic_batch* b);
const char* twdic_batch_last_error(const twdic_batch* b);
void* twdic_batch_stream(twdic_batch* b);

/* Defines the pairs (replacing any previous set): validates every pair,
 * allocates the per-pair reference gradients and per-POI reference state,
 * builds the work lists, and marks the reference state stale. */
twdic_status twdic_batch_set_pairs(twdic_batch* b, const twdic_pair* pairs, int npairs);
/* New target frames for the same pairs (d_tars[i] replaces pair i's target;
 * NULL keeps it).  Reference state stays resident. */
twdic_status twdic_batch_set_targets(twdic_batch* b, const uint8_t* const* d_tars);
/* New reference images (d_refs[i], NULL keeps it): state rebuilt on next run. */
twdic_status twdic_batch_set_references(twdic_batch* b, const uint8_t* const* d_refs);

/* One batched frame step, asynchronous on the batch stream: reference
 * gradients and per-POI state when stale, then FFT-CC over every POI, then
 * IC-GN over every POI from the FFT-CC guesses (flat form: one launch per
 * stage and kernel form; direct form: one launch per pair and stage;
 * persistent form: one launch per window class and stage). */
twdic_status twdic_batch_run(twdic_batch* b);
/* IC-GN only, from the guesses already in d_u/d_v. */
twdic_status twdic_batch_run_icgn(twdic_batch* b);
twdic_status twdic_batch_sync(twdic_batch* b);

twdic_status twdic_batch_get_tuning(twdic_batch* b, twdic_batch_tuning* t);
twdic_status twdic_batch_set_tuning(twdic_batch* b, const twdic_batch_tuning* t);
/* Times every candidate launch configuration on the current pairs (reps runs
 * each, after one warm-up), keeps the fastest, and appends it to table_path
 * (one line per workload signature) when table_path is not NULL. */
twdic_status twdic_batch_autotune(twdic_batch* b, const char* table_path, int reps);
/* Applies the entry for the current workload signature from table_path, if
 * any; *found tells whether one matched. */
twdic_status twdic_batch_load_tuning(twdic_batch* b, const char* table_path, int* found);
twdic_status twdic_batch_info_get(twdic_batch* b, twdic_batch_info* info);

/* Events on the batch stream for device timing. */
twdic_status twdic_batch_record(twdic_batch* b, int slot);
twdic_status twdic_batch_elapsed(twdic_batch* b, int slot_a, int slot_b, float* ms);

#ifdef __cplusplus
}
#endif
#endif
