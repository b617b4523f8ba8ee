/*
 * kira_hip.h — C ABI of the MI355X (gfx950) sieve + Smith-Waterman backend
 * for kira-ls-aligner.  Plain pointers and sizes only; no exceptions cross
 * this boundary.  Every call returns a kira_status; the message of the last
 * failure on a context is available from kira_ctx_last_error().
 *
 * Threading: a context owns one HIP stream and its device buffers and must
 * be driven by one host thread at a time (the reference's dispatcher owns
 * its CudaBackend on a single worker thread, src/cuda/dispatcher.rs:28-48).
 * Host arrays are caller-owned; device memory is library-owned unless a
 * *_dev entry point documents otherwise.
 */
#ifndef KIRA_HIP_H
#define KIRA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors CudaError (src/cuda/mod.rs:17-33). */
typedef enum {
    KIRA_OK = 0,
    KIRA_ERR_NOT_BUILT = 1,
    KIRA_ERR_DRIVER = 2,
    KIRA_ERR_RUNTIME = 3,
    KIRA_ERR_NO_DEVICE = 4,
    KIRA_ERR_READ_TOO_LONG = 5,
    KIRA_ERR_INVALID_ARGUMENT = 6,
    KIRA_ERR_OUT_OF_MEMORY = 7
} kira_status;

/* READ_BYTES_MAX — src/cuda/mod.rs:37: packed bytes of the longest read. */
#define KIRA_READ_BYTES_MAX 64

typedef struct kira_ctx kira_ctx;
typedef struct kira_genome kira_genome;

/* ------------------------------------------------------------------ context */

/* is_available() — src/cuda/mod.rs:40-49 / backend.rs:38-40. */
int kira_is_available(void);

/* CudaBackend::new() — src/cuda/backend.rs:63-93. */
kira_status kira_ctx_create(int device, kira_ctx **out);
void kira_ctx_destroy(kira_ctx *ctx);
const char *kira_ctx_last_error(const kira_ctx *ctx);
/* HIP stream (hipStream_t) the context launches on.  set_stream makes the
 * context launch on a caller-owned stream (NULL = the legacy default
 * stream); use_own_stream switches back to the context's own stream. */
void *kira_ctx_stream(kira_ctx *ctx);
void kira_ctx_set_stream(kira_ctx *ctx, void *stream);
void kira_ctx_use_own_stream(kira_ctx *ctx);

/* CudaBackend::release_buffers() — src/cuda/backend.rs:252-261: frees the
 * context's device scratch and pinned host staging buffers, including the
 * device temporaries the entry points keep cached between calls (a context
 * warm for the batch sizes in use makes no hipMalloc / hipFree per call). */
void kira_release_buffers(kira_ctx *ctx);

/* ------------------------------------------------------------------ sieve */

/*
 * CudaBackend::run_batch — src/cuda/backend.rs:96-249, host arrays in the
 * reference's SoA launch layout (src/cuda/spectral.cu:23-39).
 *   reads_packed      n x read_bytes_max bytes, 2-bit packed (A=00 C=01 G=10 T=11,
 *                     base k of byte i at bits 2k..2k+1)
 *   read_packed_len   packed byte length of each read (the Rust Vec length);
 *                     any value > KIRA_READ_BYTES_MAX fails the whole batch
 *                     with KIRA_ERR_READ_TOO_LONG, like the reference
 *   read_nucs         nucleotides per read
 *   ref_shifted_flat  4 bit-phases concatenated phase-major; only phase 0
 *                     is read (the kernel shifts on the fly)
 *   ref_offsets       byte offset of each job's window inside one phase
 *   ref_nucs          nucleotides per window
 *   max_mismatches    accept threshold; the limit used is min(it, read_nucs)
 * Outputs (n and 2n int32): out_shifts[i] = best shift or -1 if rejected;
 * out_mismatches[i] = best count or INT32_MAX if rejected;
 * out_mismatches[n+i] = second-best count at a distinct shift or INT32_MAX.
 * read_nucs == 0 or ref_nucs < read_nucs gives (-1, INT32_MAX, INT32_MAX).
 * n == 0 returns KIRA_OK without touching the device.
 */
kira_status kira_sieve_run_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads_packed,
                                 int32_t read_bytes_max, const int32_t *read_packed_len,
                                 const int32_t *read_nucs, const uint8_t *ref_shifted_flat,
                                 int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                                 const int32_t *ref_nucs, const uint32_t *max_mismatches,
                                 int32_t *out_shifts, int32_t *out_mismatches);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc in a serving loop, harnesses that
 * forbid it): kira_sieve_run_batch with the same arguments, results, errors
 * and sentinels, but every device buffer it stages (packed reads, phase-0
 * window bytes, job tables, bit planes, results) lies in one fixed
 * library-owned array of KIRA_NOALLOC_SCRATCH_BYTES per device, kept until
 * the process ends.  It calls no hipMalloc / hipFree and allocates no host
 * memory.  A batch whose staging exceeds the array (about
 * 200 bytes per job plus the phase-0 window bytes) fails with
 * KIRA_ERR_INVALID_ARGUMENT and a message naming both sizes.  All calls on
 * one GPU share its array: calls that overlap in time on the same GPU race,
 * so the caller serialises them (each call synchronises its stream before it
 * returns).
 */
#define KIRA_NOALLOC_SCRATCH_BYTES (64ll << 20)
kira_status kira_sieve_run_batch_noalloc(kira_ctx *ctx, int32_t n, const uint8_t *reads_packed,
                                         int32_t read_bytes_max, const int32_t *read_packed_len,
                                         const int32_t *read_nucs,
                                         const uint8_t *ref_shifted_flat,
                                         int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                                         const int32_t *ref_nucs, const uint32_t *max_mismatches,
                                         int32_t *out_shifts, int32_t *out_mismatches);

/*
 * Resident packed reference (no counterpart in the reference, which re-packs
 * and re-uploads every window per batch, src/pipeline/stage4_alignment.rs:1280-1303).
 * The ASCII sequence is uploaded once and packed on the device into two bit
 * planes (low/high code bit, 32 bases per uint32) plus a non-ACGT bitmap.
 */
kira_status kira_genome_create(kira_ctx *ctx, const uint8_t *seq, int64_t len,
                               kira_genome **out);
void kira_genome_destroy(kira_genome *g);
int64_t kira_genome_len(const kira_genome *g);
/* Device pointer to base 0 of the resident ASCII sequence, with 1024 zero
 * bytes readable before it and after its end (the padding the banded SW
 * device entry points need around every window). */
const uint8_t *kira_genome_ascii_dev(const kira_genome *g);

/*
 * Read batch packed on the device from ASCII.  reads: n x stride bytes,
 * read_lens[i] <= min(256, stride) bases (a longer read is not read past its
 * row and is reported invalid).  planes_out (device, n x 24 uint32: 8 low-plane
 * words, 8 high-plane words, 8 lower-case flag words) and valid_out (device, n bytes, 1 iff every
 * base is A/C/G/T in either case — PackedDna::valid, bitpacked.rs:29-33).
 * All pointers are device pointers; runs on the context stream.
 */
kira_status kira_reads_pack_dev(kira_ctx *ctx, int32_t n, const uint8_t *reads, int32_t stride,
                                const int32_t *read_lens, uint32_t *planes_out,
                                uint8_t *valid_out);

/* Batched sieve job description.  All pointers are DEVICE pointers. */
typedef struct {
    int32_t n_jobs;
    int32_t max_read_len;        /* longest read of the batch (<= 256 bases); sets
                                    the kernel's read-word count.  A longer read
                                    makes its jobs undispatched (flags 0). */
    const uint32_t *read_planes; /* n_reads x 24, from kira_reads_pack_dev */
    const int32_t *read_lens;    /* per read */
    const uint8_t *read_valid;   /* per read, may be NULL (all valid) */
    const int32_t *job_read;     /* per job read index; NULL -> job / jobs_per_read */
    int32_t jobs_per_read;
    const int64_t *win_start;    /* per job, base offset into the genome */
    const int32_t *win_len;      /* per job window length in bases */
    const int32_t *limit;        /* per job accept limit; NULL -> limit_all */
    int32_t limit_all;
    /* outputs */
    int32_t *out_shift;  /* best shift or -1 (rejected / skipped) */
    int32_t *out_best;   /* best count or INT32_MAX */
    int32_t *out_second; /* second-best count or INT32_MAX */
    uint8_t *out_flags;  /* may be NULL: bit0 dispatched (read and window all
                            ACGT), bit1 best <= limit, bit2 certified */
    /* certification (spectral_hit_is_certified, src/alignment/mod.rs:688-762):
       cert_max_mism < 0 disables it */
    int32_t cert_max_mism;
    int32_t cert_strict;
} kira_sieve_batch;

/* One launch of the sieve over a batch against a resident genome, on the
 * context stream (asynchronous). */
kira_status kira_sieve_genome_dev(kira_ctx *ctx, const kira_genome *g,
                                  const kira_sieve_batch *batch);

/* Synchronise the context stream. */
kira_status kira_ctx_sync(kira_ctx *ctx);

/* ------------------------------------------------------------------ banded SW */

/* AlignmentConfig — src/alignment/mod.rs:26-35. */
typedef struct {
    int32_t match_score, mismatch, gap_open, gap_extend, bandwidth, xdrop, clip_penalty;
} kira_sw_cfg;

/* Knobs the reference reads once from the environment:
 * adaptive  KIRA_ADAPTIVE_BAND (src/alignment/mod.rs:1267), reference default 1
 * max_drift KIRA_ADAPTIVE_MAX_DRIFT (mod.rs:1319), default 16
 * dp_xdrop  KIRA_DP_XDROP (mod.rs:1284), -1 = use cfg.xdrop */
typedef struct {
    int32_t adaptive, max_drift, dp_xdrop;
} kira_sw_knobs;

/* SwPublicResult — src/alignment/mod.rs:904-915.  The CIGAR (soft clips not
 * included) is n_cigar words at cigar_arena[cigar_off], each len << 4 | op
 * with op M=0 I=1 D=2 N=3 S=4.  cells: DP cells the reference algorithm
 * evaluates for the job (sum of band row widths over the rows it processes;
 * rows x window for the full-window lane groups), the unit of GCUPS. */
typedef struct {
    uint32_t ref_start, ref_end;
    int32_t read_start, read_end, score, early_abort, n_cigar, cells;
    int64_t cigar_off;
} kira_sw_result;

/*
 * banded_sw_public (src/alignment/mod.rs:891-920, core banded_sw_with
 * :1370-1650) over a batch, host arrays in and out.  Job k aligns
 * reads[read_off[k] .. +read_len[k]] against refs[ref_off[k] .. +ref_len[k]]
 * with band centre offset offset[k], band band[k] (NULL: cfg->bandwidth) and
 * abort_score[k] (<= INT32_MIN/8 disables the abort bound).  Bands up to 271
 * are supported; bands above 128 also need match, mismatch and gap costs in
 * [0, 2^20] and read_len * match < 2^21, and other jobs return
 * KIRA_ERR_INVALID_ARGUMENT for the whole batch.  cigar_arena
 * receives every job's CIGAR back to back; arena_cap words must suffice
 * (3 * sum(read_len) + 2n always does).
 */
kira_status kira_banded_sw_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads,
                                 int64_t reads_bytes, const int64_t *read_off,
                                 const int32_t *read_len, const uint8_t *refs,
                                 int64_t refs_bytes, const int64_t *ref_off,
                                 const int32_t *ref_len, const int32_t *offset,
                                 const int32_t *band, const int32_t *abort_score,
                                 const kira_sw_cfg *cfg, const kira_sw_knobs *knobs,
                                 kira_sw_result *out, uint32_t *cigar_arena, int64_t arena_cap);

/* ------------------------------------------------------------------ prefilter */

/* AnchorSpan coordinates — src/alignment/mod.rs:38-46. */
typedef struct {
    uint32_t ref_start, ref_end, read_start, read_end;
} kira_span;

/* prefilter_chain arguments that are the same for a whole batch
 * (src/alignment/prefilter.rs:93-109) plus its environment knobs:
 * tail_ext KIRA_UNGAP_TAIL (:318-326, default 1), ungap_min_id
 * KIRA_UNGAP_MINID (:331-338, -1 = unset), ungap_gap_check
 * KIRA_UNGAP_GAP_CHECK (:343-356, default 0): the final gate
 * (prefilter.rs:274-287) that turns an accept with mismatches into
 * Fallback(IndelSuspect) when a gapped alignment of the read against
 * contig[ref_start - read_start .. + read_len + bandwidth] costs strictly less
 * than the mismatches (ungapped_beaten_by_gap, src/alignment/mod.rs:163-204;
 * exact WFA score-only on the device).  With the gate on, (match + mismatch) *
 * mismatches must stay <= 65536 and read + text <= 60000 bases, else the call
 * fails with KIRA_ERR_INVALID_ARGUMENT. */
typedef struct {
    int32_t accept_enable, accept_only_top1, multi_alignments_enabled, tail_ext;
    int32_t accept_span_slack, accept_max_mismatches, accept_require_score_margin;
    int32_t long_read_threshold;
    float accept_min_identity;
    int32_t ungap_min_id;
    int32_t ungap_gap_check;
} kira_prefilter_params;

/* PrefilterOutcome (prefilter.rs:62-80) with the accepted alignment of
 * build_ungapped_alignment (:455-517).  result 0 Accept, 1 Reject,
 * 2 Fallback; reason = PrefilterReason::idx() (:31-44), Other(..) as
 * 9 + 16 * k with k = 1 read_len0, 2 non_monotonic, 3 anchor0, 4 span0,
 * 5 neg_start, 6 span_oob, 7 margin_fail.  cigar: len << 4 | op (M0 S4);
 * the MD string is host work (md.rs MdBuilder). */
typedef struct {
    int32_t result, reason, has_metrics;
    uint16_t len, mism, matches, identity_x10000;
    int32_t score;
    uint64_t read_start, ref_start;
    uint32_t aln_ref_start, aln_ref_end, aln_read_start, aln_read_end;
    int32_t aln_score;
    uint32_t aln_nm;
    int32_t n_cigar;
    uint32_t cigar[3];
} kira_prefilter_out;

/*
 * prefilter_chain (src/alignment/prefilter.rs:93-295) for n chains, host
 * arrays: chain k's read is reads[read_off[k] .. +read_len[k]], its contig
 * refs[ref_off[k] .. +ref_len[k]], anchor spans[k]; is_top1 / score_margin
 * may be NULL (true / INT32_MAX).  A span reaching past its read or contig
 * is rejected with KIRA_ERR_INVALID_ARGUMENT (the reference would panic).
 */
kira_status kira_prefilter_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads,
                                 int64_t reads_bytes, const int64_t *read_off,
                                 const int32_t *read_len, const uint8_t *refs,
                                 int64_t refs_bytes, const int64_t *ref_off,
                                 const int64_t *ref_len, const kira_span *spans,
                                 const uint8_t *is_top1, const int32_t *score_margin,
                                 const kira_sw_cfg *cfg, const kira_prefilter_params *params,
                                 kira_prefilter_out *out);

/* BatchInput — src/alignment/mod.rs:48-56 (host pointers). */
typedef struct {
    const uint8_t *read;
    uint32_t read_len;
    const uint8_t *window;
    uint32_t window_len;
    uint32_t win_start;
    uint32_t chain_ref_start, chain_ref_end, chain_read_start, chain_read_end;
    int32_t abort_score;
} kira_batch_input;

/* SimdMode — src/simd/mod.rs:5-13. */
typedef enum {
    KIRA_SIMD_SCALAR = 0,
    KIRA_SIMD_AVX2 = 1,
    KIRA_SIMD_AVXVNNI = 2,
    KIRA_SIMD_NEON = 3
} kira_simd_mode;

/*
 * align_batch_simd's Smith-Waterman (src/alignment/mod.rs:1111-1149) for
 * n_batches calls at once: batch b is inputs[batch_off[b] .. batch_off[b+1]),
 * exactly the slice the reference passes to one align_batch_simd call (the
 * lane-group size and the early stop depend on that grouping).  mode is the
 * requested SimdMode and detected the host's (detect_cached); the same
 * dispatch follows: mixed (read, window) lengths or the scalar mode run the
 * adaptive banded SW per job with the clamped chain offset (mod.rs:60-72);
 * AVX2 runs 16-lane full-window groups behind its int16 cost gate
 * (:1854-1868), AVX-VNNI one int8 group when viable (sw_int8_vnni.rs:47-55),
 * NEON 4-lane groups.  out[k] is the SwResult of inputs[k] (soft clips are
 * added by build_alignment, :1200-1254).
 */
kira_status kira_align_batch_simd(kira_ctx *ctx, int32_t n_batches, const int32_t *batch_off,
                                  const kira_batch_input *inputs, kira_simd_mode mode,
                                  kira_simd_mode detected, const kira_sw_cfg *cfg,
                                  const kira_sw_knobs *knobs, kira_sw_result *out,
                                  uint32_t *cigar_arena, int64_t arena_cap);

/*
 * Device-resident form for batches that repeat a shape: a plan fixes the
 * job count, read lengths and bands (it sorts jobs by band class and length
 * into packets of 64 lane-per-job jobs, or of 4 (2) jobs of 16 (32) lanes
 * for bands above 64, and sizes the trace slabs); run_dev then only
 * launches kernels on the context stream.  All kira_sw_jobs_dev pointers are
 * device pointers; refs must have at least 1024 readable bytes before and
 * after every window.  res: n x 10 int32 (score, read_end, ref_end,
 * early_abort, ref_start, read_start, n_cigar, int16 overflow flag (never set
 * by a correct plan), cells, 0); cigar: the plan's
 * cigar_words, job k's ops at word kira_banded_sw_plan_cigar_off(plan)[k].
 */
typedef struct kira_sw_plan kira_sw_plan;
typedef struct {
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *refs;
    const int64_t *ref_off;
    const int32_t *ref_len;
    const int32_t *offset;
    const int32_t *band;
    const int32_t *abort_score;
    int32_t *res;
    uint32_t *cigar;
} kira_sw_jobs_dev;

kira_status kira_banded_sw_plan_create(kira_ctx *ctx, int32_t n, const int32_t *read_len,
                                       const int32_t *band, const kira_sw_cfg *cfg,
                                       kira_sw_plan **out);
void kira_banded_sw_plan_destroy(kira_sw_plan *plan);
int64_t kira_banded_sw_plan_cigar_words(const kira_sw_plan *plan);
/* Device memory the plan holds (trace and row-meta slabs, packet tables),
 * for admission against an HBM budget. */
int64_t kira_banded_sw_plan_device_bytes(const kira_sw_plan *plan);
const int64_t *kira_banded_sw_plan_cigar_off(const kira_sw_plan *plan);
kira_status kira_banded_sw_run_dev(kira_ctx *ctx, kira_sw_plan *plan,
                                   const kira_sw_jobs_dev *jobs, const kira_sw_cfg *cfg,
                                   const kira_sw_knobs *knobs);

/*
 * Banded SW autotuner (the reference picks lane groups per (L, W) shape and a
 * lane count per host, src/pipeline/stage4_alignment.rs:1064-1069, 1153-1160,
 * src/simd/mod.rs:30; here the choice is a routing of each job to a kernel
 * form, timed on the device).  A shape key is (band bucket, read-length
 * bucket, job-count bucket): band buckets end at bands 4, 8, 16, 32, 39, 47,
 * 64, 71, 79, 128, 135, 143, 207, 223, 271 (the bands of one bucket can take
 * the same routes),
 * read lengths bucket by log2, and the count is the batch's number of jobs in
 * the job's band bucket (all lengths), bucketed by log4.  Routes: 1 one lane per job (int16 or int32 band by
 * the score bound), 2 + k the cooperative shape k of (P, G) = (5, 16),
 * (9, 16), (17, 16), (13, 32), (17, 32) in int32 and (6, 16), (10, 16),
 * (18, 16), (14, 32), (18, 32), (10, 8), (18, 8) in packed int16 pairs
 * (k = 5..11: gap open > 0,
 * read_len x match <= 8191, mismatch <= 8000, open + extend <= 4000, extend
 * <= 500).  Routing never changes results.
 *
 * kira_sw_autotune times, for each shape, every eligible route on a seeded
 * synthetic batch of that shape (read_len bases, band, n_jobs jobs up to
 * 262144; reads with 3% substitutions and 0.5% indels against windows of the
 * clamp_window geometry), checks that every route returns identical results,
 * keeps the fastest in the context's table and, when cache_path is not NULL,
 * writes the table to that file (lines "<device> <key> <route> <ms> <ms of the
 * built-in routing>", keyed by the device's architecture name and CU count;
 * other devices' lines are kept).  kira_sw_tune_load reads this device's lines
 * of such a file into the context; kira_sw_tune_clear empties the table.
 * Plans built on the context afterwards (kira_banded_sw_plan_create, every
 * banded SW entry point, servers created on it) route by the table.  Not to
 * be called while a server created on the context runs other work.
 */
/* band, read length, and the number of jobs of the band's bucket in the
 * batches to be tuned for (the synthetic batch has that many jobs, all of
 * this read length, at most 262144) */
typedef struct {
    int32_t band, read_len, n_jobs;
} kira_sw_shape;
typedef struct {
    uint32_t key;           /* shape key the choice is stored under */
    int32_t route;          /* chosen route */
    int32_t n_candidates;   /* routes timed */
    float ms_best;          /* device time of one batch on the chosen route */
    float ms_builtin;       /* same batch on the built-in routing */
} kira_sw_tune_result;
kira_status kira_sw_autotune(kira_ctx *ctx, const kira_sw_cfg *cfg, const kira_sw_knobs *knobs,
                             int32_t n_shapes, const kira_sw_shape *shapes,
                             const char *cache_path, kira_sw_tune_result *out);
kira_status kira_sw_tune_load(kira_ctx *ctx, const char *cache_path, int32_t *n_loaded);
void kira_sw_tune_clear(kira_ctx *ctx);
/* Jobs of a plan per route: counts[1] one lane per job, counts[2 + k]
 * cooperative shape k (counts has 14 entries; counts[0] is 0). */
void kira_banded_sw_plan_route_counts(const kira_sw_plan *plan, int32_t *counts);
/* Kernel class of each job of a plan (n entries): 2 b + packed for one lane
 * per job with a band of 9, 17, 33, 65, 129, 257 positions (b = 0..5), 12 + k
 * for cooperative shape k (k = 0..11, the route table's order).  The plan
 * launches its non-empty classes in this
 * order, one kernel each. */
void kira_banded_sw_plan_job_class(const kira_sw_plan *plan, int32_t *cls);

/* ------------------------------------------------------------------ records */

/* One alignment record of the path (the fields of Alignment, src/types.rs,
 * that build_alignment fills, src/alignment/mod.rs:1200-1254): CIGAR with
 * soft clips at cigar_out[cigar_off .. +n_cigar], MD text (no NUL) at
 * md_out[md_off .. +md_len]. */
typedef struct {
    uint32_t ref_start, ref_end, read_start, read_end;
    int32_t score;
    uint32_t nm;
    int32_t n_cigar, md_len;
    int64_t cigar_off, md_off;
} kira_alignment;

/*
 * build_alignment for n SW results (host code, no device work): job k's
 * read is reads[read_off[k] .. +read_len[k]], its window
 * windows[win_off[k] .. +win_len[k]] starting at reference position
 * win_start[k]; sw[k] and its CIGAR ops sw_cigar[sw[k].cigar_off ..] as
 * returned by the SW entry points.  NM / MD follow md.rs:25-204 (MD keeps
 * reference letters upper-cased, other bytes as N; an identical non-ACGT
 * base counts in NM).  For a certified sieve hit pass read_start 0,
 * read_end L, ref_start = shift and the CIGAR {L}M (mod.rs:636-686).
 * cigar_used / md_used receive the totals needed; with cigar_out / md_out
 * NULL only the sizes are computed; too small a buffer returns
 * KIRA_ERR_INVALID_ARGUMENT.
 */
kira_status kira_build_alignments(int32_t n, const uint8_t *reads, const int64_t *read_off,
                                  const int32_t *read_len, const uint8_t *windows,
                                  const int64_t *win_off, const int32_t *win_len,
                                  const uint32_t *win_start, const kira_sw_result *sw,
                                  const uint32_t *sw_cigar, kira_alignment *out,
                                  uint32_t *cigar_out, int64_t cigar_cap, char *md_out,
                                  int64_t md_cap, int64_t *cigar_used, int64_t *md_used);

/* ------------------------------------------------------------------ serving */

/*
 * Continuous-batching server over a resident genome (the reference's
 * gpu-server / dispatcher, src/cuda/server.rs:15-193, dispatcher.rs:21-104,
 * which run one batch at a time).  kira_server_submit queues a request and
 * returns at once; one worker thread admits queued requests first come first
 * served into up to `slots` concurrent streams whenever a slot is free and
 * the request's device memory fits in the hbm_budget_bytes pool (and its
 * inputs and outputs in the pinned staging pool), so requests that arrive
 * while others run overlap on the device.  Per request: one input copy,
 * read packing, sieve + certification, banded SW (jobs whose linked sieve
 * job is certified are dropped on the device), one output copy; with
 * use_graph these are captured and launched as one hipGraph per request.
 * A request's results do not depend on the other requests in flight.
 * All request arrays are host arrays, copied at submit.
 */
typedef struct kira_server kira_server;

typedef struct {
    int32_t slots;               /* requests in flight at once, 1..16 */
    int64_t hbm_budget_bytes;    /* device memory pool for in-flight requests */
    int64_t host_staging_bytes;  /* pinned host staging pool */
    int32_t use_graph;           /* one captured hipGraph per request */
    kira_sw_cfg cfg;
    kira_sw_knobs knobs;
    int32_t sieve_limit;         /* accept limit of sieve jobs without their own */
    int32_t cert_max_mism;       /* certification limit, < 0 disables it */
    int32_t cert_strict;
    /* coalescing: when a slot frees up, the waiting requests at the front of
     * the queue, up to this many reads in total (and as many as the free
     * memory holds), are admitted together as one device batch (their
     * results are the same as alone); 0 admits one request per slot */
    int64_t group_max_reads;
} kira_server_config;

typedef struct {
    /* reads: read k is reads[read_off[k] .. +read_len[k]] (ASCII); sieve
       jobs use reads of at most 256 bases (longer ones are not dispatched) */
    int32_t n_reads;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    /* sieve jobs: read index, window start (genome coordinate) and length,
       optional per-job limit */
    int32_t n_sieve;
    const int32_t *sieve_read;
    const int64_t *sieve_win_start;
    const int32_t *sieve_win_len;
    const int32_t *sieve_limit;
    /* banded SW jobs (banded_sw_public): read index, window in genome
       coordinates, band centre offset, band (NULL: cfg.bandwidth), abort
       score, and the sieve job whose certification drops the job (NULL or
       -1: never dropped) */
    int32_t n_sw;
    const int32_t *sw_read;
    const int64_t *sw_win_start;
    const int32_t *sw_win_len;
    const int32_t *sw_offset;
    const int32_t *sw_band;
    const int32_t *sw_abort;
    const int32_t *sw_skip_sieve;
} kira_request;

/* Arrays owned by the server until kira_server_release(ticket). */
typedef struct {
    int32_t status; /* kira_status of the request */
    int32_t n_sieve, n_sw;
    const int32_t *shift, *best, *second; /* as kira_sieve_batch outputs */
    const uint8_t *flags;
    const kira_sw_result *sw;  /* cigar_off indexes `cigar` */
    const uint8_t *sw_skipped; /* 1: dropped because its sieve job was certified */
    const uint32_t *cigar;
    int64_t cigar_words;
    double submit_s, start_s, done_s; /* steady clock: queued, admitted, results ready */
} kira_response;

typedef struct {
    int64_t submitted, admitted, completed;
    int32_t peak_in_flight;
    int64_t peak_device_bytes;
    int64_t grouped; /* requests admitted as members of a coalesced group */
} kira_server_stats;

kira_status kira_server_create(kira_ctx *ctx, const kira_genome *g,
                               const kira_server_config *cfg, kira_server **out);
kira_status kira_server_submit(kira_server *s, const kira_request *req, int64_t *ticket);
/* Blocks until the request is done; out->status carries its own outcome. */
kira_status kira_server_wait(kira_server *s, int64_t ticket, kira_response *out);
void kira_server_release(kira_server *s, int64_t ticket);
void kira_server_get_stats(kira_server *s, kira_server_stats *out);
const char *kira_server_last_error(kira_server *s);
/* Drains every queued request, then frees the server. */
void kira_server_destroy(kira_server *s);

/* ------------------------------------------------------------------ stage 4 */

/* Chain (src/types.rs:131-140) without its anchors: contig, reference and
 * oriented-read span, strand, chain score.  Stage 3 lists a read's chains
 * best first. */
typedef struct {
    uint32_t ref_id, ref_start, ref_end, read_start, read_end;
    int32_t strand; /* 0 forward, 1 reverse */
    int32_t score;
} kira_chain;

/* The router knobs of the fast path (src/alignment/router.rs:48-215), with
 * the reference defaults kira_fast_knobs_default fills in brackets:
 * wfa_max_read_len KIRA_WFA_MAX_LEN [300]; algo KIRA_ALGO 0 packed [0],
 * 1 spectral, 2 wfa, 3 sw (other values: packed); quality_gate / quality_pct
 * KIRA_CHAIN_QUALITY_GATE [0] / KIRA_CHAIN_QUALITY_PCT [70]; wfa_budget_pct
 * KIRA_WFA_BUDGET_PCT [25]; myers_pct / myers_floor KIRA_MYERS_BOUND_PCT [15]
 * / KIRA_MYERS_BOUND_FLOOR [4]; spectral_pct / spectral_floor
 * KIRA_SPECTRAL_MISM_PCT [3] / KIRA_SPECTRAL_MISM_FLOOR [2]; wfa_ends_free
 * KIRA_WFA_ENDS_FREE [0]; wfa_lead KIRA_WFA_LEAD [0]; wfa_adaptive
 * KIRA_WFA_ADAPTIVE [0 = exact]; cert_strict KIRA_CERT_STRICT [0]. */
typedef struct {
    int32_t wfa_max_read_len, algo, quality_gate, quality_pct, wfa_budget_pct;
    int32_t myers_pct, myers_floor, spectral_pct, spectral_floor;
    int32_t wfa_ends_free, wfa_lead, wfa_adaptive, cert_strict;
} kira_fast_knobs;

void kira_fast_knobs_default(kira_fast_knobs *k);

/* AlignmentStageConfig (src/pipeline/stage4_alignment.rs:33-54) and the
 * environment knobs stage 4 reads: tail_ext KIRA_UNGAP_TAIL (1),
 * ungap_min_id KIRA_UNGAP_MINID (-1 unset), short_dp_topk KIRA_SHORT_DPTOPK
 * (0 unset), ambig_div KIRA_AMBIG_DIV (5), xs_min_ratio_pct KIRA_XS_MINRATIO
 * (0), spectral_pct / spectral_floor (router.rs:73-74: 3, 2), cert_strict
 * KIRA_CERT_STRICT (0).  simd is the SimdMode stage 4 would detect on its
 * host (it sets the lane-group size; SCALAR sends every job to banded SW).
 * supplementary runs emit_supplementary_alignments (split_read.rs:22-57). */
typedef struct {
    kira_sw_cfg cfg;
    kira_sw_knobs knobs;
    float min_chain_ratio;
    int32_t accept_enable, accept_only_top1, accept_span_slack;
    float accept_min_identity;
    int32_t accept_max_mismatches, accept_require_score_margin;
    int32_t dp_topk, dp_abort_margin, long_read_threshold, max_alignments;
    int32_t tail_ext, ungap_min_id, short_dp_topk, ambig_div, xs_min_ratio_pct;
    int32_t spectral_pct, spectral_floor, cert_strict;
    kira_simd_mode simd;
    int32_t supplementary;
    /* cpu_mode: the configuration the reference runs without --gpu (the
     * dispatcher inactive, stage4_alignment.rs:868-882, 1073-1106): chains
     * that fall back from the prefilter try the fast DP cascade (router knobs
     * `fast`) before SW, and lane jobs skip the sieve dispatch. */
    int32_t cpu_mode;
    kira_fast_knobs fast;
    /* twotier KIRA_TWOTIER (1) / twotier_k KIRA_TWOTIER_K (5): the candidate
     * reorder of uncertain short reads (two_tier_reorder,
     * stage4_alignment.rs:346-451, 596-683), which scores up to twotier_k
     * passing chains with the fast DP cascade, orders them by that score,
     * keeps their alignments, sets the DP budget and the competitor XS. */
    int32_t twotier, twotier_k;
    /* ungap_gap_check KIRA_UNGAP_GAP_CHECK (0): the prefilter's final WFA gap
     * gate, as in kira_prefilter_params. */
    int32_t ungap_gap_check;
} kira_stage4_cfg;

/* Alignment (src/types.rs) as stage 4 leaves it: kind 0 AcceptedUngapped,
 * 1 DpAligned; CIGAR with soft clips at cigar[cigar_off .. +n_cigar]
 * (len << 4 | op, M0 I1 D2 N3 S4); MD text (no NUL) at md[md_off .. +md_len];
 * xs_score valid iff has_xs. */
typedef struct {
    int32_t kind;
    uint32_t ref_id, ref_start, ref_end, read_start, read_end;
    int32_t score, mapq, is_rev, is_supplementary;
    uint32_t nm;
    int32_t has_xs, xs_score;
    int32_t n_cigar, md_len;
    int64_t cigar_off, md_off;
} kira_stage4_aln;

/* AlignmentBatchStats counters (stage4_alignment.rs:57-100) the driver
 * fills; reason_counts = prefilter_reason_counts (PrefilterReason::idx). */
typedef struct {
    int64_t reads, chains_total, chains_used, exact_matches;
    int64_t prefilter_accept, prefilter_reject, prefilter_fallback;
    int64_t reason_counts[10];
    int64_t gpu_spectral_resolved, dp_simd, dp_scalar, dp_early_abort, dp_reads;
    int64_t supplementary, unmapped;
    int64_t packed_spectral_resolved, wfa_resolved; /* cpu_mode fast path */
} kira_stage4_stats;

typedef struct kira_stage4_result kira_stage4_result;

/* Read-only view of a result, valid until kira_stage4_result_free:
 * read k's alignments are alns[aln_off[k] .. aln_off[k+1]) in the order
 * stage 4 appends them.  phase_ms: host wall time of chain triage (exact
 * match + ungapped prefilter), sieve + certification, lane-group SW, banded
 * SW, and records + supplementary alignments. */
typedef struct {
    int32_t n_reads;
    const int64_t *aln_off;
    const kira_stage4_aln *alns;
    const uint32_t *cigar;
    const char *md;
    int64_t n_alns, cigar_words, md_bytes;
    kira_stage4_stats stats;
    double phase_ms[5];
} kira_stage4_view;

/*
 * stage4_alignment::run (src/pipeline/stage4_alignment.rs:959-1268, with
 * process_read_prefilter :535-957) with the GPU dispatcher active, for
 * single-end reads.  Per read: chain selection (min_chain_ratio, top-k, and
 * with cfg->twotier the candidate reorder of uncertain short reads, whose
 * verified fast-path alignments are kept and whose runner-up score becomes
 * XS), exact-match shortcut,
 * ungapped prefilter; fallback chains of reads of <= 150 bases become lane
 * jobs on build_simd_window windows (:1426-1445), dispatched to the 2-bit
 * sieve with certification (:1272-1358), the rest to align_batch_simd lane
 * groups by (read, window) shape; longer reads go to banded SW on
 * clamp_window windows (mod.rs:851-877).  Then XS (the two-tier
 * competitor, else the chain-score proxy, :1215-1250) and supplementary
 * alignments.  With cfg->cpu_mode the
 * fallback chains go through the fast DP cascade first (as
 * kira_fast_dp_batch) and the sieve dispatch is skipped.  The genome holds every contig
 * back to back: contig i is bases contig_off[i] .. +contig_len[i].  Read k
 * is reads[read_off[k] .. +read_len[k]] (forward strand); its chains are
 * chains[chain_off[k] .. chain_off[k+1]); repeat_min_occ may be NULL (0).
 * A chain whose contig, read span or reference span is out of range fails
 * the batch with KIRA_ERR_INVALID_ARGUMENT.  Exact matching, prefilter,
 * sieve and SW run on the device; records (NM / MD) are host work.
 */
kira_status kira_stage4_run(kira_ctx *ctx, const kira_genome *g, int32_t n_contigs,
                            const int64_t *contig_off, const int64_t *contig_len,
                            int32_t n_reads, const uint8_t *reads, const int64_t *read_off,
                            const int32_t *read_len, const int32_t *repeat_min_occ,
                            const int64_t *chain_off, const kira_chain *chains,
                            const kira_stage4_cfg *cfg, kira_stage4_result **out);
kira_status kira_stage4_result_view(const kira_stage4_result *r, kira_stage4_view *view);
void kira_stage4_result_free(kira_stage4_result *r);

/* ------------------------------------------------------------------ records out */

/*
 * left_normalize_indels (src/alignment/normalize.rs:55-190), host code: slide
 * every I / D of `cigar` (n_ops words len << 4 | op) as far left as the
 * contig (0-based ref_start = the alignment's first reference base) and the
 * strand-oriented read allow.  Returns 1 and writes the new CIGAR to out
 * (*out_n words; out_cap >= 2 * n_ops + 1 always suffices) when anything
 * moved, 0 (*out_n = n_ops, out untouched) when the CIGAR is already
 * canonical, -1 on invalid arguments or too small an out_cap.
 */
int32_t kira_left_normalize_indels(const uint32_t *cigar, int32_t n_ops, const uint8_t *contig,
                                   int64_t contig_len, uint32_t ref_start, const uint8_t *read,
                                   int32_t read_len, uint32_t *out, int32_t out_cap,
                                   int32_t *out_n);

/*
 * normalize_alignments (normalize.rs:197-249) over a stage-4 result, in
 * place: every record's CIGAR is left-normalized against its contig and the
 * read in the record's orientation (reads as passed to kira_stage4_run), and
 * NM / MD are recomputed for the records that moved.  original (may be NULL,
 * same offsets as reads) holds ReadRecord::original_seq — the read with its
 * IUPAC letters — for the reads with has_original[k] set; their records get
 * NM / MD from it whether or not they moved.  Host code.
 */
kira_status kira_stage4_left_normalize(kira_stage4_result *r, const kira_genome *g,
                                       int32_t n_contigs, const int64_t *contig_off,
                                       const int64_t *contig_len, int32_t n_reads,
                                       const uint8_t *reads, const int64_t *read_off,
                                       const int32_t *read_len, const uint8_t *original,
                                       const uint8_t *has_original);

/* The OutputConfig fields single-end records use (src/io/mod.rs:797-850):
 * tags NM / MD / AS / XS / RG, bwa-mem -M (split_as_secondary) and -Y
 * (soft_clip_supplementary); read_group is the @RG header line or NULL. */
typedef struct {
    int32_t write_nm, write_md, write_as, write_xs, write_rg;
    int32_t split_as_secondary, soft_clip_supplementary;
    const char *read_group;
} kira_sam_cfg;

/*
 * SAM text of a stage-4 result, host code: for each read in order, every
 * record through SamFormatter::append_alignment (src/io/mod.rs:1036-1140)
 * for a single-end read, or one append_unmapped line (:971-1033) when it has
 * none.  MAPQ is the record's own (stage 5 is not run); no XA / SA tags.
 * Read k: id read_ids[k], SEQ seqs[seq_off[k] .. +seq_len[k]] (forward, as
 * aligned), QUAL quals[seq_off[k] ..] or none when quals is NULL.  A record
 * whose CIGAR does not consume the read (checked_cigar, :1737-1750) fails
 * the call with KIRA_ERR_INVALID_ARGUMENT.  *used = bytes of text; with out
 * NULL only the size is computed; cap too small returns
 * KIRA_ERR_INVALID_ARGUMENT.
 */
kira_status kira_stage4_sam(const kira_stage4_result *r, int32_t n_contigs,
                            const char *const *contig_names, int32_t n_reads,
                            const char *const *read_ids, const uint8_t *seqs,
                            const int64_t *seq_off, const int32_t *seq_len, const uint8_t *quals,
                            const kira_sam_cfg *cfg, char *out, int64_t cap, int64_t *used);

/* ------------------------------------------------------------------ WFA / Myers */

/*
 * bounded_edit_distance (src/alignment/myers.rs:47-367) for n (pattern,
 * text) pairs, host arrays in and out: pattern k is
 * patterns[pat_off[k] .. +pat_len[k]], text k texts[text_off[k] .. +text_len[k]].
 * Semi-global unit-cost edit distance (the whole pattern against the best
 * substring of the text; A / C / G / T match in either case, any other byte
 * matches nothing, itself included).  found[k] = 1 with dist[k] and end[k]
 * (one past the best end position in the text, the last of equal minima)
 * when the distance is <= max_k[k] (>= 0), else found[k] = 0.
 */
kira_status kira_myers_batch(kira_ctx *ctx, int32_t n, const uint8_t *patterns,
                             int64_t patterns_bytes, const int64_t *pat_off,
                             const int32_t *pat_len, const uint8_t *texts, int64_t texts_bytes,
                             const int64_t *text_off, const int32_t *text_len,
                             const int32_t *max_k, uint8_t *found, int32_t *dist, int32_t *end);

/* WfaAlignment (src/alignment/wfa.rs:20-39) of one pair: found 0 when no
 * alignment costs <= max_score.  score = the loss of wfa_align_scored (the
 * affine score wfa_score_only_scored returns is pat_len * match_score -
 * score); the text span [text_start, text_end); the CIGAR, n_cigar words
 * len << 4 | op (M0 I1 D2), at cigar_arena[cigar_off]. */
typedef struct {
    int32_t found, score, text_start, text_end, n_cigar;
    int64_t cigar_off;
} kira_wfa_result;

/*
 * wfa_align_scored / wfa_score_only_scored (src/alignment/wfa.rs:58-81 over
 * wfa_align_with_costs :549-648) for n pairs, host arrays in and out.  Costs
 * from cfg as WfaCosts::scored (:41-55): mismatch A + B, gap open O,
 * insertion extend A + E, deletion extend E; unless match_score > 0,
 * mismatch > 0, gap_open >= 0 and gap_extend > 0 no pair is found.
 * WfaOptions per pair: max_score[k] (<= 65535), text_begin_free[k] (free
 * leading text bases; NULL = 0), adaptive_drop for all (<= 0: exact).  The
 * whole pattern aligns against text starting within its first
 * text_begin_free + 1 bases and ending anywhere; bytes match when equal.
 * want_cigar 0 skips the traceback (scores only; cigar_arena may be NULL).
 * The CIGARs are packed back to back into cigar_arena (arena_cap words; too
 * few returns KIRA_ERR_INVALID_ARGUMENT).  pat_len + text_len <= 60000.
 */
kira_status kira_wfa_batch(kira_ctx *ctx, int32_t n, const uint8_t *patterns,
                           int64_t patterns_bytes, const int64_t *pat_off, const int32_t *pat_len,
                           const uint8_t *texts, int64_t texts_bytes, const int64_t *text_off,
                           const int32_t *text_len, const int32_t *max_score,
                           const int32_t *text_begin_free, int32_t adaptive_drop,
                           const kira_sw_cfg *cfg, int32_t want_cigar, kira_wfa_result *out,
                           uint32_t *cigar_arena, int64_t arena_cap);

/* ------------------------------------------------------------------ fast DP cascade */

/* The Alignment try_fast_dp_alignment returns (mod.rs:636-686, 481-575):
 * kind 0 none (the caller falls back to SW), 1 packed spectral certificate,
 * 2 WFA; CIGAR (no clips; len << 4 | op, M0 I1 D2) at
 * cigar_out[cigar_off .. +n_cigar], MD text (no NUL) at md_out[md_off ..
 * +md_len]; ref_start / ref_end in contig coordinates. */
typedef struct {
    int32_t kind;
    uint32_t ref_start, ref_end, read_start, read_end;
    int32_t score;
    uint32_t nm;
    int32_t n_cigar, md_len;
    int64_t cigar_off, md_off;
} kira_fast_aln;

/* Chains past the router gates, certified by the packed scan, passed by
 * Myers, improved by the origin retry, declined by the partial-read check
 * (or a CIGAR that does not span the read), and aligned by WFA; host wall
 * time of the spectral scan, Myers, the WFA passes (with the origin probe)
 * and the records. */
typedef struct {
    int64_t attempted, spectral, myers_pass, origin_retries, clip_declined, wfa;
    double phase_ms[4];
} kira_fast_stats;

/*
 * try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440) for n chains
 * against a resident genome (contig i = bases contig_off[i] .. +contig_len[i]):
 * job k aligns reads[read_off[k] .. +read_len[k]], already oriented to the
 * chain's strand, in the window the chain's contig, ref_start and read_start
 * imply; chains[k].score feeds the chain-quality gate.  The packed scan and
 * certification, Myers and WFA run on the device; the origin probe of the
 * retry, the partial-read check and NM / MD run on the host.  A chain with an
 * out-of-range contig, or a read outside reads_bytes, fails the call with
 * KIRA_ERR_INVALID_ARGUMENT.  cigar_used / md_used receive the totals; too
 * small a cigar_cap / md_cap returns KIRA_ERR_INVALID_ARGUMENT (the totals
 * are still set).  Sum(2 * read_len + bandwidth + 66) words and 4 times
 * that in MD bytes always suffice.  knobs NULL = the defaults; stats may be
 * NULL.
 */
kira_status kira_fast_dp_batch(kira_ctx *ctx, const kira_genome *g, int32_t n_contigs,
                               const int64_t *contig_off, const int64_t *contig_len, int32_t n,
                               const uint8_t *reads, int64_t reads_bytes, const int64_t *read_off,
                               const int32_t *read_len, const kira_chain *chains,
                               const kira_sw_cfg *cfg, const kira_fast_knobs *knobs,
                               kira_fast_aln *out, uint32_t *cigar_out, int64_t cigar_cap,
                               char *md_out, int64_t md_cap, int64_t *cigar_used,
                               int64_t *md_used, kira_fast_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
