/*
 * kira_stage4_oracle.h — the stage-4 driver restated in C over the oracle
 * (kira_stage4_oracle.c): the CPU baseline timed beside the device driver,
 * labelled "restated reference CPU path".  Test infrastructure only.
 */
#ifndef KIRA_STAGE4_ORACLE_H
#define KIRA_STAGE4_ORACLE_H

#include <stdint.h>

#include "kira_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One chain of stage 3 (src/pipeline/stage3_chaining.rs): contig, reference
 * span, strand-oriented read span, strand (0 forward, 1 reverse), score. */
typedef struct {
    uint32_t ref_id, ref_start, ref_end, read_start, read_end;
    int32_t strand, score;
} ko_chain;

/* AlignmentStageConfig (stage4_alignment.rs:33-54) with the environment
 * knobs stage 4 reads, as oracle/stage4.py's default_cfg names them.
 * simd: 0 scalar, 1 AVX2 (16 lanes), 2 AVX-VNNI (32), 3 NEON (4). */
typedef struct {
    ko_cfg sw;
    int32_t adaptive, max_drift, dp_xdrop;
    float min_chain_ratio;
    int32_t accept_enable, accept_only_top1, accept_span_slack;
    float accept_min_identity;
    int32_t accept_max_mismatches, accept_require_score_margin;
    int32_t dp_topk, dp_abort_margin, long_read_threshold, max_alignments;
    int32_t tail_ext, ungap_min_id, short_dp_topk, ambig_div, xs_min_ratio_pct;
    int32_t simd, supplementary;
    int32_t twotier, twotier_k, ungap_gap_check;
    ko_router router;
} ko_s4_cfg;

/* A batch: contig i = genome[contig_off[i] .. +contig_len[i]); read k =
 * reads[read_off[k] .. +read_len[k]) (forward strand), its chains
 * chains[chain_off[k] .. chain_off[k + 1]); repeat_min_occ may be NULL. */
typedef struct {
    int32_t n_reads;
    const uint8_t *genome;
    const int64_t *contig_off, *contig_len;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const int32_t *repeat_min_occ;
    const int64_t *chain_off;
    const ko_chain *chains;
} ko_s4_in;

/* An Alignment record of read `read`: kind 0 AcceptedUngapped, 1 DpAligned;
 * CIGAR (soft clips included) at cigar[cigar_off .. +n_cigar]; MD text at
 * md[md_off .. +md_len]; xs_score valid iff has_xs. */
typedef struct {
    int32_t read, kind;
    uint32_t ref_id, ref_start, ref_end, read_start, read_end;
    int32_t score, mapq, is_rev, is_supplementary;
    uint32_t nm;
    int32_t has_xs, xs_score;
    int32_t n_cigar, md_len;
    int64_t cigar_off, md_off;
} ko_s4_rec;

/* AlignmentBatchStats counters (stage4_alignment.rs:57-100). */
typedef struct {
    int64_t reads, chains_total, chains_used, exact_matches;
    int64_t prefilter_accept, prefilter_reject, prefilter_fallback;
    int64_t reason_counts[10];
    int64_t gpu_spectral_resolved, dp_simd, dp_scalar, dp_early_abort, dp_reads;
    int64_t supplementary, unmapped;
    int64_t packed_spectral_resolved, wfa_resolved;
} ko_s4_stats;

typedef struct ko_s4_result ko_s4_result;

/* Read k's records are alns[aln_off[k] .. aln_off[k + 1]) in the order
 * stage 4 appends them. */
typedef struct {
    int32_t n_reads;
    const int64_t *aln_off;
    const ko_s4_rec *alns;
    const uint32_t *cigar;
    const char *md;
    int64_t n_alns, cigar_words, md_bytes;
    ko_s4_stats stats;
} ko_s4_view;

/* stage4_alignment::run without the GPU dispatcher (fallback chains try the
 * fast DP cascade before SW), single-end, on `threads` pthreads. */
ko_s4_result *ko_stage4_run_mt(const ko_s4_in *in, const ko_s4_cfg *cfg, int threads);
void ko_stage4_view(const ko_s4_result *r, ko_s4_view *view);
void ko_stage4_free(ko_s4_result *r);

#ifdef __cplusplus
}
#endif
#endif
