/*
 * kira_oracle.h — CPU restatement of kira-ls-aligner's sieve + Smith-Waterman
 * hot path, used ONLY as a test oracle and as the CPU baseline timed beside
 * the GPU path.  Nothing in the product links or calls this library.
 *
 * Every function cites the reference file:line it restates.  All arithmetic
 * is integer and follows the reference literally (including its int16/int8
 * saturating SIMD forms), so results are meant to be bit-identical.
 */
#ifndef KIRA_ORACLE_H
#define KIRA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AlignmentConfig — src/alignment/mod.rs:26-35 */
typedef struct {
    int32_t match_score;
    int32_t mismatch;
    int32_t gap_open;
    int32_t gap_extend;
    int32_t bandwidth;
    int32_t xdrop;
    int32_t clip_penalty;
} ko_cfg;

/* Knobs read once from the environment in the reference; explicit here.
 * adaptive: KIRA_ADAPTIVE_BAND (mod.rs:1267), default 1
 * max_drift: KIRA_ADAPTIVE_MAX_DRIFT (mod.rs:1319), default 16
 * dp_xdrop: KIRA_DP_XDROP (mod.rs:1284), -1 = use cfg.xdrop
 * cert_strict: KIRA_CERT_STRICT (mod.rs:1300), default 0 */
typedef struct {
    int32_t adaptive;
    int32_t max_drift;
    int32_t dp_xdrop;
    int32_t cert_strict;
} ko_knobs;

/* CIGAR op encoding (BAM order): len << 4 | op, op M=0 I=1 D=2 N=3 S=4.
 * CigarOp Display — src/types.rs:184-197. */
enum { KO_M = 0, KO_I = 1, KO_D = 2, KO_N = 3, KO_S = 4 };

/* SwResult — src/alignment/mod.rs:1190-1198 (cigar excludes soft clips). */
typedef struct {
    uint32_t ref_start;
    uint32_t ref_end;
    int32_t read_start;
    int32_t read_end;
    int32_t score;
    int32_t early_abort;
    int32_t n_cigar;
    uint32_t *cigar; /* caller-owned, capacity cigar_cap */
    int32_t cigar_cap;
    int64_t cells; /* banded SW: DP cells evaluated (sum of row widths) */
} ko_sw_result;

/* ---------------- sieve ---------------- */
int ko_pack(const uint8_t *seq, size_t n, uint8_t *bits);
void ko_shift_bits_right(const uint8_t *bits, size_t n, unsigned bit_offset, uint8_t *out);
void ko_pre_shift(const uint8_t *bits, size_t n, uint8_t *out4);
uint32_t ko_mismatch_count(const uint8_t *read, const uint8_t *ref, size_t n_bytes, size_t n_nucs);
int ko_scan_best_with_second(const uint8_t *read_bits, size_t read_bytes, size_t read_len,
                             int read_valid, const uint8_t *phases, size_t phase_len,
                             size_t ref_len, uint64_t *shift, uint64_t *best, uint64_t *second);
void ko_backend_run_batch(int32_t n, const uint8_t *reads_packed, int32_t read_bytes_max,
                          const int32_t *read_lens, const uint8_t *ref_flat,
                          int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                          const int64_t *ref_phase_bytes, const int32_t *ref_lens,
                          const uint32_t *max_mismatches, int32_t *out_shifts,
                          int32_t *out_mism);
void ko_backend_run_batch_mt(int32_t n, const uint8_t *reads_packed, int32_t read_bytes_max,
                             const int32_t *read_lens, const uint8_t *ref_flat,
                             int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                             const int64_t *ref_phase_bytes, const int32_t *ref_lens,
                             const uint32_t *max_mismatches, int32_t *out_shifts,
                             int32_t *out_mism, int threads);
uint32_t ko_spectral_max_mismatches(uint32_t read_len, uint32_t pct, uint32_t floor_);
int ko_spectral_hit_is_certified(const uint8_t *read, const uint8_t *ref_aligned, size_t len,
                                 size_t ref_len, uint64_t hit_mism, int has_second,
                                 uint64_t second, uint64_t max_mism, int strict);
/* Returns the NM and writes MD (NUL-terminated) — md.rs:147-151. */
uint32_t ko_ungapped_tags(const uint8_t *read, const uint8_t *ref, size_t len, char *md,
                          size_t md_cap);
int ko_build_simd_window(uint64_t ref_len, uint32_t chain_ref_start, uint32_t chain_read_start,
                         uint32_t read_len, int32_t bandwidth, uint32_t *win_start,
                         uint32_t *win_len);

/* ---------------- ungapped prefilter ---------------- */
typedef struct {
    uint32_t ref_start, ref_end, read_start, read_end;
} ko_span;

typedef struct {
    int32_t result;   /* 0 Accept, 1 Reject, 2 Fallback */
    int32_t reason;   /* PrefilterReason::idx() ; Other(..) = 9 + sub code*16 */
    int32_t has_metrics;
    uint16_t len, mism, matches, identity_x10000;
    int32_t score;
    uint64_t read_start, ref_start;
    /* accepted alignment (build_ungapped_alignment) */
    uint32_t aln_ref_start, aln_ref_end, aln_read_start, aln_read_end;
    int32_t aln_score;
    uint32_t aln_nm;
    int32_t n_cigar;
    uint32_t cigar[3];
} ko_prefilter_out;

uint16_t ko_identity_x10000(size_t len, size_t mism);
size_t ko_min_len_required(size_t read_len, uint16_t id_x10000, size_t mism, size_t span_slack);
void ko_prefilter_chain(const uint8_t *read, size_t read_len, const uint8_t *ref, size_t ref_len,
                        const ko_span *chain, const ko_cfg *cfg, int is_top1, int accept_enable,
                        int accept_only_top1, size_t accept_span_slack,
                        float accept_min_identity, size_t accept_max_mismatches,
                        int32_t accept_require_score_margin, int32_t score_margin,
                        size_t long_read_threshold, int multi_alignments_enabled,
                        int tail_ext, int32_t ungap_min_id /* -1 = unset */,
                        ko_prefilter_out *out, char *md, size_t md_cap);

/* ---------------- Smith-Waterman ---------------- */
void ko_banded_sw(const uint8_t *read, size_t q_len, const uint8_t *ref, size_t r_len,
                  int32_t offset, const ko_cfg *cfg, int32_t abort_score, const ko_knobs *kn,
                  ko_sw_result *res);

/* One BatchInput — src/alignment/mod.rs:48-73. */
typedef struct {
    const uint8_t *read;
    uint32_t read_len;
    const uint8_t *window;
    uint32_t window_len;
    uint32_t win_start;
    uint32_t chain_ref_start, chain_ref_end, chain_read_start, chain_read_end;
    int32_t abort_score;
} ko_batch_input;

/* Full-window lane-group SW: sw_batch_avx2 (lane_bits=16), sw_batch_int8 (8),
 * sw_batch_neon (32).  n <= group lanes; all inputs share (L, W). */
void ko_sw_group(const ko_batch_input *in, int32_t n, const ko_cfg *cfg, int lane_bits,
                 const ko_knobs *kn, ko_sw_result *res);

/* align_batch_simd SW dispatch — src/alignment/mod.rs:1111-1149.
 * mode/detected: 0 Scalar, 1 Avx2, 2 AvxVnni, 3 Neon. */
void ko_align_batch_sw(const ko_batch_input *in, int32_t n, const ko_cfg *cfg, int mode,
                       int detected, const ko_knobs *kn, ko_sw_result *res);

/* build_alignment — src/alignment/mod.rs:1200-1254; NM/MD via md.rs:162-204. */
typedef struct {
    uint32_t ref_start, ref_end, read_start, read_end;
    int32_t score;
    uint32_t nm;
    int32_t n_cigar;
} ko_alignment;
void ko_build_alignment(const uint8_t *read, size_t read_len, const uint8_t *window,
                        size_t window_len, uint32_t win_start, const ko_sw_result *sw,
                        ko_alignment *out, uint32_t *cigar_out, int32_t cigar_cap, char *md,
                        size_t md_cap);
uint32_t ko_nm_md_from_cigar(const uint8_t *read, size_t read_len, const uint8_t *ref,
                             size_t ref_len, size_t read_start, size_t ref_start,
                             const uint32_t *cigar, int32_t n_cigar, char *md, size_t md_cap);
size_t ko_cigar_string(const uint32_t *cigar, int32_t n, char *out, size_t cap);

/* CPU baselines, parallel over jobs with pthreads.  bands (nullable) replaces
 * cfg->bandwidth per job; out_cells (nullable) receives each job's cell count. */
void ko_banded_sw_batch_mt(int32_t n, const uint8_t *reads, const int64_t *read_off,
                           const int32_t *read_len, const uint8_t *refs, const int64_t *ref_off,
                           const int32_t *ref_len, const int32_t *offsets,
                           const int32_t *abort_scores, const int32_t *bands, const ko_cfg *cfg,
                           const ko_knobs *kn, int32_t *out_score, int32_t *out_end,
                           int64_t *out_cells, int threads);

/* ---- WFA / Myers fast path (kira_wfa_oracle.c) ---- */

/* bounded_edit_distance — src/alignment/myers.rs:47-367: semi-global edit
 * distance of the pattern against the text (best end position, last on
 * ties); 0 when it exceeds max_k. */
int ko_myers_bounded(const uint8_t *pattern, size_t m, const uint8_t *text, size_t n,
                     size_t max_k, size_t *dist, size_t *pos);

/* WfaAlignment — wfa.rs:115-126 (cigar caller-owned, capacity cigar_cap) */
typedef struct {
    int32_t score;                 /* WFA loss of the match-rewarding objective */
    int32_t text_start, text_end;
    int32_t n_cigar;
    uint32_t *cigar;
    int32_t cigar_cap;
} ko_wfa_aln;

/* wfa_align_scored / wfa_score_only_scored — wfa.rs:58-81 with WfaOptions
 * {max_score, adaptive_drop (<= 0: none), text_begin_free}.  1 found, 0 none
 * within max_score (or unsupported costs), -1 CIGAR capacity too small. */
int ko_wfa_align_scored(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n, const ko_cfg *cfg,
                        int32_t max_score, int32_t adaptive_drop, int32_t text_begin_free,
                        ko_wfa_aln *out);
int ko_wfa_score_only_scored(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n,
                             const ko_cfg *cfg, int32_t max_score, int32_t adaptive_drop,
                             int32_t text_begin_free, int32_t *score);

/* ---- the fast DP cascade (kira_fast_oracle.c) ---- */
/* The router knobs (src/alignment/router.rs:48-215); algo 0 packed,
 * 1 spectral, 2 wfa, 3 sw. */
typedef struct {
    int32_t wfa_max_read_len, algo, quality_gate, quality_pct, wfa_budget_pct, myers_pct,
        myers_floor, spectral_pct, spectral_floor, wfa_ends_free, wfa_lead, wfa_adaptive,
        cert_strict;
} ko_router;
/* kind 0 none (the caller falls back to SW), 1 packed certificate, 2 WFA */
typedef struct {
    int32_t kind;
    uint32_t ref_start, ref_end, read_start, read_end;
    int32_t score;
    uint32_t nm;
    int32_t n_cigar;
} ko_fast_out;
/* try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440) for one chain;
 * returns out->kind != 0. */
int ko_fast_dp(const uint8_t *read, size_t read_len, const uint8_t *contig, size_t contig_len,
               uint32_t ref_start, uint32_t read_start, int32_t chain_score, const ko_cfg *cfg,
               const ko_router *router, ko_fast_out *out);
/* ko_fast_dp with the record: *cigar (out->n_cigar words) and *md
 * (NUL-terminated) point into the scratch and stay valid until its next use.
 * One scratch per thread. */
typedef struct ko_fast_scratch ko_fast_scratch;
ko_fast_scratch *ko_fast_scratch_new(void);
void ko_fast_scratch_free(ko_fast_scratch *s);
int ko_fast_dp_rec(const uint8_t *read, size_t read_len, const uint8_t *contig,
                   size_t contig_len, uint32_t ref_start, uint32_t read_start,
                   int32_t chain_score, const ko_cfg *cfg, const ko_router *router,
                   ko_fast_scratch *s, ko_fast_out *out, const uint32_t **cigar,
                   const char **md);
/* the same over n chains on `threads` pthreads; chain k reads reads +
 * read_off[k] and contig chain_contig[k] of the genome */
void ko_fast_dp_batch_mt(int32_t n, const uint8_t *reads, const int64_t *read_off,
                         const int32_t *read_len, const uint8_t *genome,
                         const int64_t *contig_off, const int64_t *contig_len,
                         const int32_t *chain_contig, const uint32_t *chain_ref_start,
                         const uint32_t *chain_read_start, const int32_t *chain_score,
                         const ko_cfg *cfg, const ko_router *router, ko_fast_out *out,
                         int threads);

#ifdef __cplusplus
}
#endif
#endif
