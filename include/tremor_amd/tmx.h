/*
 * tmx.h - C-ABI of the MI355X template-matching library (libtremor_amd.so).
 *
 * Exact z-normalised Euclidean distance / Pearson correlation of every template
 * against every subsequence (window) of one long float32 waveform, for kNN and
 * threshold search, on one AMD Instinct MI355X (gfx950).  The answers are the
 * reference sequential scan's answers: same positions, same FP32 distances.
 *
 * Conventions
 *   - Plain C types, caller-owned inputs (copied), caller-allocated outputs,
 *     except threshold hits (library-owned until tmx_hits_free).
 *   - Every call returns 0 on success, > 0 for a recoverable condition,
 *     < 0 for an error; tmx_last_error() then describes it.  No exceptions
 *     cross this boundary.
 *   - Synchronous calls on one context are serialised by the context; they
 *     run on the context's own HIP stream on one device.  The serving loop
 *     (below) adds one stream per lane.
 *   - Distances are squared z-normalised Euclidean distances d = 2m(1 - rho)
 *     (ZnormSearch.hpp:23-31).  The missing-sample sentinel is -100000.0f
 *     (ZnormSearch.hpp:21): a window whose first sample is missing is never
 *     reported; missing samples inside a window count as 0.
 *
 * Reference interfaces each entry point replaces (files of the reference
 * repository, tremor/...):
 *   tmx_load_waveform    SkipSequentialScan::build           algos/sss/SkipSequential.cpp:138-198
 *                        (window statistics)                 algos/tremor/indexing.cpp:128-153
 *   tmx_knn              SkipSequentialScan::search_knn      algos/sss/SkipSequential.cpp:296-393
 *                        Tremor::searchIndex (I, D outputs)  algos/htremor/Tremor.hpp:172
 *   tmx_threshold        SkipSequentialScan::search_threshold algos/sss/SkipSequential.cpp:223-294
 *   tmx_distance_batch   l2_dist_znorm_subsequence           distance_computers/ZnormSearch.hpp:103-165
 *   tmx_window_stats     two-pass window statistics          algos/sss/SkipSequential.cpp:157-179
 *   tmx_prepare_template znormalize_inplace + query_block_order ZnormSearch.hpp:33-90
 *   tmx_set_fft_mode,    Tremor::setRefinementMode          algos/htremor/Tremor.hpp:37-44,161
 *   tmx_set_scan_mode    (scan / FFT / adaptive refinement)
 *   tmx_set_index_mode   Tremor's index search and leaf size  tremor_main.cpp:296-302 (--method tremor),
 *                        (--refinement leaf / adaptive)      :312-314, :341-343 (--leaf-size)
 */
#ifndef TREMOR_AMD_TMX_H
#define TREMOR_AMD_TMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tmx_ctx tmx_ctx;

/* Arithmetic of the candidate-generating inner product.  Every form is
 * followed by the exact verification pass, so all return identical answers. */
enum tmx_precision {
    TMX_PREC_FP32 = 0,   /* plain FP32 sliding dot product on the vector ALUs */
    TMX_PREC_BF16X3 = 1, /* Hankel GEMM on MFMA, split-BF16 (hi*hi + hi*lo + lo*hi), FP32 accumulation */
    TMX_PREC_FP16 = 2    /* Hankel GEMM on MFMA, FP16 operands, FP32 accumulation */
};

/* Which templates take the FFT (MASS overlap-save) form of the candidate pass. */
enum tmx_fft_mode {
    TMX_FFT_AUTO = 0,    /* cost model per template (long templates), as the reference's adaptive refinement */
    TMX_FFT_OFF = 1,     /* every template on the inner-product form chosen by `precision` */
    TMX_FFT_ALWAYS = 2   /* every template on the FFT form */
};

/* Which templates take the early-abandoning scan (the reference's sequential
 * scan with its per-block abandon test), decided once the kNN bounds are
 * known. */
enum tmx_scan_mode {
    TMX_SCAN_AUTO = 0,   /* cost model on a sampled abandon test, as the reference's adaptive refinement */
    TMX_SCAN_OFF = 1,    /* never */
    TMX_SCAN_ALWAYS = 2  /* every template */
};

/* Whether templates take the iSAX index (TREMOR's index search: leaves in
 * key order whose region bound passes, then the windows whose own word bound
 * passes, all exactly verified), decided once the kNN bounds are known. */
enum tmx_index_mode {
    TMX_INDEX_OFF = 0,    /* never (default) */
    TMX_INDEX_AUTO = 1,   /* templates whose index leaves at most 1% of their windows to visit (the
                             reference's rule: leave the index when more of them pass the bound) */
    TMX_INDEX_ALWAYS = 2  /* every eligible template (length a multiple of 16; 16 PAA segments) */
};

/* Launch form of the matrix-core pass (no counterpart in the reference: its
 * CPU scan has no launches).  Answers do not depend on it. */
enum tmx_launch {
    TMX_LAUNCH_AUTO = 0,  /* timed with the group sizes (tmx_cfg.autotune = 1), else the grid */
    TMX_LAUNCH_GRID = 1,  /* one workgroup per (waveform block, template group) item */
    TMX_LAUNCH_QUEUE = 2  /* one persistent workgroup per CU claiming items from per-XCD work
                             queues, stealing from the other XCDs' queues once its own is empty */
};

typedef struct {
    int32_t device;          /* HIP device ordinal */
    int32_t precision;       /* enum tmx_precision */
    int64_t cand_capacity;   /* candidate slots per call; 0 = default (grown on demand) */
    int32_t seed_target;     /* kNN bound seeding: target seed entries per template; 0 = default */
    int32_t verbose;         /* 1 = print per-phase timings to stderr */
    int32_t fft_mode;        /* enum tmx_fft_mode; ignored for TMX_PREC_FP32 (plain scan) */
    int32_t scan_mode;       /* enum tmx_scan_mode */
    int32_t verify_seq;      /* 1 = the exact recomputation takes every window's statistics by the
                                sequential sums (tests, diagnostics); 0 = default (a rigorous
                                parallel shortcut, sequential only where it cannot decide) */
    int32_t autotune;        /* 1 = default: the templates per matrix-core workgroup (4, 8, 16) and
                                the launch form (mfma_launch) are timed on a sampled pass the first
                                time a batch shape is seen and the fastest kept for the context;
                                0 = fixed (8); 4, 8 or 16 = that size, fixed */
    int64_t max_candidates;  /* most candidate windows one candidate pass may hold; 0 = default
                                (2^31 - 1, less if device memory is short).  A call whose templates
                                together would exceed it is split into halves of its templates, run
                                one after the other (same answers); one template alone exceeding it
                                is an error */
    int32_t lower_bounds;    /* 1 = the skip-sequential method's iSAX lower bound on the early-abandoning
                                scan (SkipSequentialScan's lower_bounds, SkipSequential.cpp:200-211,
                                :256, :321): a window whose bound exceeds the template's bound is
                                skipped; the iSAX words of each template length are built on first
                                use and kept; templates whose length is not a multiple of
                                paa_segments scan without bounds.  0 = default (off) */
    int32_t paa_segments;    /* PAA segments of those bounds, 1..64 (SkipSequentialScan's
                                paa_segments); default 16 */
    int32_t mfma_launch;     /* enum tmx_launch; default TMX_LAUNCH_AUTO */
    int32_t index_mode;      /* enum tmx_index_mode; default TMX_INDEX_OFF.  The index of a template
                                length (16-segment iSAX words, their keys sorted on the device, leaves
                                of at most leaf_size windows) is built on first use and kept with the
                                words; requires paa_segments = 16 when lower_bounds is on */
    int32_t leaf_size;       /* windows per index leaf before it splits (the reference's --leaf-size);
                                default 2000 */
} tmx_cfg;

typedef struct {
    double seconds_total;    /* wall time of the call, device work included */
    double seconds_seed;     /* kNN bound seeding */
    double seconds_main;     /* inner-product pass with fused candidate filter */
    double seconds_verify;   /* exact reference-order recomputation of the candidates */
    double seconds_select;   /* selection / merge */
    int64_t comparisons;     /* sum over templates of (n - m + 1) */
    int64_t candidates;      /* windows the candidate pass emitted */
    int32_t retries;         /* candidate-buffer overflows resolved by a tighter bound or a bigger buffer */
    float max_abs_rho_err;   /* max |rho_fast - rho_exact| over the verified candidates */
    double kernel_ms_main;   /* device time of the main candidate pass (HIP events on the context stream) */
    double flops_main;       /* algorithmic flops of that pass: sum over templates of 2 m (n - m + 1) */
    int32_t fft_templates;   /* templates that took the FFT form */
    int32_t fft_block;       /* FFT length B of those templates (0: none) */
    int64_t verified;        /* candidates left for the exact recomputation after the kNN bound was
                                tightened from the candidates' own upper bounds */
    int32_t scan_templates;  /* templates that took the early-abandoning scan */
    double scan_blocks;      /* their mean 32-sample blocks visited per window (sampled) */
    int32_t tuned_group;     /* templates per matrix-core workgroup used by this call */
    double seconds_tune;     /* time spent timing group sizes in this call (first call of a shape) */
    int32_t fft_lmax;        /* longest template of the FFT form's piece shape (pieces of fft_block
                                samples every fft_block - fft_lmax + 1 windows; 0: none) */
    int32_t splits;          /* times the call's templates were split in two because their
                                candidates exceeded max_candidates (0: one pass) */
    int64_t lb_skipped;      /* windows the lower bound skipped in the early-abandoning scan */
    int32_t tuned_launch;    /* enum tmx_launch of the matrix-core pass of this call */
    int32_t index_templates; /* templates that took the iSAX index */
    int64_t index_windows;   /* windows in their leaves whose region bound passed (before the
                                windows' own bounds) */
} tmx_stats;

typedef struct {
    int64_t count;           /* rows */
    int32_t *tmpl;           /* template index per row */
    int64_t *counter;        /* running row number within its template (the CSV "Counter") */
    uint64_t *pos;           /* window start */
    float *d;                /* squared z-normalised distance (reference FP32) */
} tmx_hits;

int tmx_cfg_default(tmx_cfg *cfg);
int tmx_create(tmx_ctx **out, const tmx_cfg *cfg);
void tmx_destroy(tmx_ctx *ctx);
const char *tmx_last_error(const tmx_ctx *ctx);
/* The setters below change the context's settings for later synchronous
 * calls and for serving loops started later; a running serving loop keeps
 * the settings it started with. */
int tmx_set_precision(tmx_ctx *ctx, int32_t precision);
int tmx_set_fft_mode(tmx_ctx *ctx, int32_t mode);
int tmx_set_scan_mode(tmx_ctx *ctx, int32_t mode);
/* enum tmx_index_mode and the index's leaf size (>= 1) for later calls. */
int tmx_set_index_mode(tmx_ctx *ctx, int32_t mode, int32_t leaf_size);

/* Copies the waveform to the device and builds its FP64 prefix sums of x and
 * x^2 (missing samples as 0).  Replaces any waveform loaded before.  Every
 * sample must be finite or the missing-sample sentinel: a NaN or infinite
 * sample is refused with an error naming the first one (nothing is loaded). */
int tmx_load_waveform(tmx_ctx *ctx, const float *x, uint64_t n);
uint64_t tmx_waveform_length(const tmx_ctx *ctx);

/* kNN search.  templates: raw (not normalised) samples, template i at
 * templates[offsets[i] .. offsets[i] + lengths[i]).  Output k slots per
 * template, best first, ties by position; unfilled slots are (0, FLT_MAX).
 * merge_offset > 0 keeps answers pairwise >= merge_offset apart. */
int tmx_knn(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets, int32_t count,
            int32_t k, int32_t merge_offset, uint64_t *out_pos, float *out_d, tmx_stats *stats);

/* Threshold search: every window with d <= 2m(1 - rho), per template sorted
 * by position; with merge_offset > 0, a hit closer than merge_offset + 1 to
 * the last kept hit is dropped.  *out is library-owned: tmx_hits_free. */
int tmx_threshold(tmx_ctx *ctx, const float *templates, const int32_t *lengths, const uint64_t *offsets,
                  int32_t count, float rho, int32_t merge_offset, tmx_hits **out, tmx_stats *stats);
void tmx_hits_free(tmx_hits *hits);

/* Exact reference distances (no early abandoning) of one raw template of
 * length m against the windows starting at starts[0..count); a window whose
 * first sample is missing or whose std is <= 1e-12 gets FLT_MAX. */
int tmx_distance_batch(tmx_ctx *ctx, const float *tmpl, int32_t m, const uint64_t *starts, int32_t count,
                       float *out_d);

/* Allocation-free variant of tmx_distance_batch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_tmpl (m raw samples), d_starts (count window starts) and d_out (count
 * distances) are device pointers of the context's device; the work is
 * enqueued on `stream` (a hipStream_t; NULL = the context's own stream) and
 * the call returns without waiting: synchronise the stream before reading
 * d_out.  Results are those of tmx_distance_batch.  The prepared template
 * lives in a fixed per-device scratch array of the library (about 264 KB),
 * so m is at most 65536 (longer: an error) and nothing is allocated or
 * freed.  Every call on one device shares that scratch: calls whose work can
 * overlap on one device (other streams, no ordering between them) race.  Do
 * not load another waveform while such work is in flight. */
int tmx_distance_batch_noalloc(tmx_ctx *ctx, const float *d_tmpl, int32_t m, const uint64_t *d_starts,
                               int32_t count, float *d_out, void *stream);

/* Exact reference window statistics (two-pass FP64, stored as float) of the
 * windows [first, first + count) for length m. */
int tmx_window_stats(tmx_ctx *ctx, int32_t m, uint64_t first, uint64_t count, float *mean, float *std);

/* Z-normalised template and block order exactly as the reference prepares
 * them, computed on the device.  order needs m / 32 slots. */
int tmx_prepare_template(tmx_ctx *ctx, const float *tmpl, int32_t m, float *out_z, int32_t *out_order);

/* Diagnostics: runs one of the library's own FFT plans on host data, so the
 * transform error the FFT forms' candidate filter allows for can be measured
 * against an exact transform.  No counterpart in the reference (its FFTW calls,
 * fft_fallback.cpp:149-150,170, are trusted with an assumed error constant,
 * fft_fallback.hpp:84-85).  B: a power of two in [2^14, 2^18].
 *   TMX_FFT_PROBE_R2C      the forward real-to-complex plan of the piece and
 *                          template spectra: in = count x B floats, out =
 *                          count x (B/2 + 1) complex (interleaved floats)
 *   TMX_FFT_PROBE_INVERSE  the unnormalised complex inverse of length B/2 of
 *                          the hipFFT form (B >= 2^16 in use): in and out =
 *                          count x B/2 complex
 * count: 1..64 transforms. */
enum tmx_fft_probe_kind {
    TMX_FFT_PROBE_R2C = 0,
    TMX_FFT_PROBE_INVERSE = 1
};
int tmx_fft_probe(tmx_ctx *ctx, int32_t B, int32_t kind, const float *in, int32_t count, float *out);

/* Diagnostics for tests of the iSAX index (its leaves are internal in the
 * reference: iSAXIndex.cpp's tree): the index of template length m with
 * leaves of at most leaf_size windows, built on first use as a search builds
 * it (m a multiple of 16, m <= n).  Copies the W = n - m + 1 window ids in key
 * order into ids (ids_cap >= W), and each leaf's first position in that order,
 * window count and depth (key bits its windows share), in key order, into the
 * leaf arrays (leaf_cap entries; the leaves beyond are not copied); *nleaves
 * gets the leaf count.  Also copies the 16 symbols of every window into words
 * when words is not NULL (ids_cap x 16 bytes, ids_cap >= W, also when ids is
 * NULL). */
int tmx_index_probe(tmx_ctx *ctx, int32_t m, int32_t leaf_size, uint32_t *ids, uint64_t ids_cap,
                    uint32_t *leaf_start, uint32_t *leaf_count, int32_t *leaf_depth, int64_t leaf_cap,
                    int64_t *nleaves, uint8_t *words);

/* Waits for all device work of the context's synchronous calls. */
int tmx_sync(tmx_ctx *ctx);

/* ---- Serving loop ---------------------------------------------------------
 * No counterpart in the reference (its searches are batch jobs); the boundary
 * follows its search entry points (tmx_knn / tmx_threshold above) per request.
 *
 * The waveform and its statistics stay resident; requests (one template each,
 * kNN or threshold, any length) are submitted from any thread and queue until
 * a lane is free.  A free lane takes the waiting requests in arrival order, as
 * many as fit the batch size and the memory budget, and runs them as one
 * batch on its own HIP stream; batches of different lanes overlap on the
 * device (continuous batching).  A request's answer does not depend on the
 * batch it ran in: every answer is the exact reference answer.
 * While the loop runs, tmx_load_waveform is refused; the synchronous calls
 * remain available (they run on their own stream). */

enum tmx_request_kind {
    TMX_REQ_KNN = 0,
    TMX_REQ_THRESHOLD = 1
};

typedef struct {
    int32_t lanes;           /* batches in flight at once, one stream each (default 2, at most 8) */
    int32_t max_batch;       /* requests per batch (default 64) */
    int64_t mem_budget;      /* HBM bytes admitted requests may hold; 0 = half the free memory at start */
    int32_t max_template;    /* longest template accepted (default 65536) */
    int32_t reserved;
} tmx_server_cfg;

typedef struct {
    int32_t kind;            /* enum tmx_request_kind */
    int32_t m;               /* template length */
    const float *tmpl;       /* raw template samples, copied by tmx_submit */
    int32_t k;               /* kNN: answers wanted, 1..4096 */
    float rho;               /* threshold: correlation threshold */
    int32_t merge_offset;    /* as in tmx_knn / tmx_threshold */
    int32_t reserved;
} tmx_request;

typedef struct {
    int32_t status;          /* 0 answered; < 0 failed (error says why) */
    int32_t kind;
    int64_t count;           /* kNN: k slots (unfilled (0, FLT_MAX)); threshold: hit rows by position */
    uint64_t *pos;           /* window starts (library-owned until tmx_result_free) */
    float *d;                /* squared z-normalised distances (reference FP32) */
    double submit_s;         /* steady-clock seconds at submission */
    double start_s;          /* its batch started */
    double done_s;           /* its answer was ready */
    int64_t batch;           /* batch sequence number */
    int32_t batch_size;      /* requests in that batch */
    int32_t lane;            /* lane that ran it */
    char error[160];
} tmx_result;

int tmx_server_cfg_default(tmx_server_cfg *cfg);
/* Starts the lanes (worker threads).  The derived structures the requests
 * will need are built on first use. */
int tmx_server_start(tmx_ctx *ctx, const tmx_server_cfg *cfg);
/* Answers every request submitted so far, then stops the lanes.  Answers not
 * yet collected stay available to tmx_poll / tmx_wait. */
int tmx_server_stop(tmx_ctx *ctx);
/* Thread-safe.  Returns a ticket > 0, or < 0 if the request is invalid or
 * the loop is not running. */
int64_t tmx_submit(tmx_ctx *ctx, const tmx_request *req);
/* 0: answered, *out filled and the ticket retired; 1: pending; < 0: unknown
 * ticket.  Thread-safe. */
int tmx_poll(tmx_ctx *ctx, int64_t ticket, tmx_result *out);
/* As tmx_poll, waiting up to timeout_s seconds (< 0: no limit); 1 on timeout. */
int tmx_wait(tmx_ctx *ctx, int64_t ticket, double timeout_s, tmx_result *out);
void tmx_result_free(tmx_result *res);

#ifdef __cplusplus
}
#endif
#endif
