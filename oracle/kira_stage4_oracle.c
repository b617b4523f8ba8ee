/*
 * kira_stage4_oracle.c — the stage-4 alignment driver in C over the oracle's
 * pinned pieces, on pthreads: the CPU baseline timed beside the device
 * driver (kira_stage4_run), labelled "restated reference CPU path".  Test
 * infrastructure only: the product never links or calls it.
 *
 * It restates src/pipeline/stage4_alignment.rs run() (:959-1268) with
 * process_read_prefilter (:535-957) in the configuration without the GPU
 * dispatcher (a fallback chain tries the fast DP cascade, :868-882, before
 * SW; no sieve dispatch), single-end, with the two-tier candidate reorder
 * (:346-451) and the supplementary alignments of src/pipeline/
 * split_read.rs:22-151: the steps of oracle/stage4.py with cpu_mode, in the
 * same order, and a CPU test compares every record of the two.  Like the
 * reference's rayon pools it spreads the per-read step over reads, the lane
 * groups over groups and the banded SW and supplementary jobs over jobs,
 * each on `threads` workers taking tasks through an atomic counter; every
 * task keeps its records apart and the records are put in read order at the
 * end, so the result does not depend on the thread count.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "kira_oracle.h"
#include "kira_stage4_oracle.h"

#define I32_MIN_ (-2147483647 - 1)
#define I32_MAX_ 2147483647

/* split_read.rs:10-19 */
#define SUPP_MIN_SCORE_RATIO 0.5f
#define SUPP_MIN_READ_SPAN 50u
#define SUPP_MAX_OVERLAP_PCT 50u
#define MAX_SUPP_PER_READ 4

#define READ_CHUNK 256 /* reads per task of the per-read step */
#define JOB_CHUNK 64   /* banded SW jobs per task */
#define MAX_CAND 64    /* two-tier candidates per read (KIRA_TWOTIER_K is 5) */

static int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static int32_t sat_i32(int64_t v) {
    return v > I32_MAX_ ? I32_MAX_ : (v < I32_MIN_ ? I32_MIN_ : (int32_t)v);
}
static int32_t wrap_i32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
static uint32_t usub(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

/* ------------------------------------------------------------ buffers */
typedef struct {
    void *p;
    size_t n, cap, elem;
} vec;

static void vec_init(vec *v, size_t elem) {
    v->p = NULL;
    v->n = v->cap = 0;
    v->elem = elem;
}

static void *vec_push(vec *v, size_t count) {
    if (v->n + count > v->cap) {
        size_t c = v->cap ? v->cap * 2 : 64;
        while (c < v->n + count) c *= 2;
        v->p = realloc(v->p, c * v->elem);
        v->cap = c;
    }
    void *at = (char *)v->p + v->n * v->elem;
    v->n += count;
    return at;
}

static void vec_free(vec *v) {
    free(v->p);
    vec_init(v, v->elem);
}

/* The records of one task (a chunk of reads, a lane group, a chunk of
 * jobs): CIGAR and MD offsets count from the task's own arenas. */
typedef struct {
    vec recs; /* ko_s4_rec */
    vec cig;  /* uint32_t */
    vec md;   /* char */
} part;

static void part_init(part *p) {
    vec_init(&p->recs, sizeof(ko_s4_rec));
    vec_init(&p->cig, sizeof(uint32_t));
    vec_init(&p->md, 1);
}

static void part_free(part *p) {
    vec_free(&p->recs);
    vec_free(&p->cig);
    vec_free(&p->md);
}

static part *parts_new(int32_t n) {
    part *p = (part *)calloc((size_t)n + 1, sizeof(part));
    for (int32_t k = 0; k < n; k++) part_init(&p[k]);
    return p;
}

static ko_s4_rec *add_rec(part *P, int32_t read, int32_t kind, uint32_t ref_id, int is_rev) {
    ko_s4_rec *a = (ko_s4_rec *)vec_push(&P->recs, 1);
    memset(a, 0, sizeof *a);
    a->read = read;
    a->kind = kind;
    a->ref_id = ref_id;
    a->is_rev = is_rev;
    a->cigar_off = (int64_t)P->cig.n;
    a->md_off = (int64_t)P->md.n;
    return a;
}

static void add_cig(part *P, ko_s4_rec *a, const uint32_t *ops, int32_t n) {
    if (n > 0) memcpy(vec_push(&P->cig, (size_t)n), ops, (size_t)n * 4);
    a->n_cigar += n;
}

static void add_md(part *P, ko_s4_rec *a, const char *md, size_t n) {
    if (n) memcpy(vec_push(&P->md, n), md, n);
    a->md_len = (int32_t)n;
}

/* ------------------------------------------------- per-thread scratch */
typedef struct {
    uint8_t *rc; /* the reverse complement of the current read */
    size_t rc_cap;
    uint32_t *swcig; /* a banded SW result's CIGAR */
    size_t swcig_cap;
    uint32_t *cig; /* build_alignment's CIGAR */
    size_t cig_cap;
    char *md;
    size_t md_cap;
    ko_fast_scratch *fast;
} scratch;

static void grow(void **p, size_t *cap, size_t need) {
    if (*cap >= need) return;
    free(*p);
    *p = malloc(need);
    *cap = need;
}

static void scratch_init(scratch *s) {
    memset(s, 0, sizeof *s);
    s->fast = ko_fast_scratch_new();
}

static void scratch_free(scratch *s) {
    free(s->rc);
    free(s->swcig);
    free(s->cig);
    free(s->md);
    ko_fast_scratch_free(s->fast);
    memset(s, 0, sizeof *s);
}

/* --------------------------------------------------------- the batch */
typedef struct {
    int32_t read;
    uint32_t ref_id;
    int32_t is_rev;
    ko_span span;
    int32_t abort;
    uint32_t win_start, win_len; /* lane jobs: the build_simd_window window */
} job;

typedef struct {
    int64_t used, exact, accept, reject, fallback, packed, wfa, dp_reads, reason[10];
} counts;

typedef struct {
    const ko_s4_in *in;
    const ko_s4_cfg *c;
    ko_knobs kn;
    int lanes;
    /* per read: the chain scores of the XS proxy and the two-tier competitor */
    int32_t *chain_best, *chain_second, *has_comp, *comp;
    /* per chunk of reads */
    int32_t n_chunks;
    part *p1;
    vec *lane_jobs, *scalar_jobs;
    counts *cnt;
    int32_t next;
} driver;

static const uint8_t *contig_of(const ko_s4_in *in, uint32_t id) {
    return in->genome + in->contig_off[id];
}

/* src/seq.rs:1-17: A/C/G/T of either case to the upper-case complement,
 * anything else to N */
static uint8_t comp_base(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    default: return 'N';
    }
}

/* read r on the strand of a chain (a reverse complement goes to s->rc) */
static const uint8_t *oriented(const ko_s4_in *in, int32_t r, int rev, scratch *s) {
    const uint8_t *f = in->reads + in->read_off[r];
    if (!rev) return f;
    const uint32_t L = (uint32_t)in->read_len[r];
    grow((void **)&s->rc, &s->rc_cap, (size_t)L + 1);
    for (uint32_t k = 0; k < L; k++) s->rc[k] = comp_base(f[L - 1 - k]);
    return s->rc;
}

/* exact_match_exists (mod.rs:818-832) */
static int exact_match(const uint8_t *read, uint32_t L, const uint8_t *contig, int64_t clen,
                       const ko_span *s) {
    if (s->read_start != 0 || s->read_end != L || s->ref_end < s->ref_start ||
        s->ref_end - s->ref_start != L)
        return 0;
    if ((int64_t)s->ref_start + L > clen) return 0;
    return memcmp(contig + s->ref_start, read, L) == 0;
}

/* exact_match_alignment (mod.rs:778-815): {L}M, MD "L", NM counts the bases
 * that are not A/C/G/T/= (identical_base_is_edit, md.rs:134-137) */
static void exact_record(part *P, int32_t r, const uint8_t *read, uint32_t L, const ko_span *s,
                         uint32_t ref_id, int is_rev, const ko_s4_cfg *c, int supp) {
    ko_s4_rec *a = add_rec(P, r, 0, ref_id, is_rev);
    a->ref_start = s->ref_start;
    a->ref_end = s->ref_start + L;
    a->read_start = 0;
    a->read_end = L;
    a->score = wrap_i32((int64_t)c->sw.match_score * L);
    uint32_t nm = 0;
    for (uint32_t k = 0; k < L; k++) {
        const uint8_t u = (read[k] >= 'a' && read[k] <= 'z') ? (uint8_t)(read[k] - 32) : read[k];
        nm += !(u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == '=');
    }
    a->nm = nm;
    a->is_supplementary = supp;
    const uint32_t op = (L << 4) | KO_M;
    add_cig(P, a, &op, 1);
    char t[16], buf[16];
    int n = 0;
    uint32_t v = L;
    do {
        t[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    for (int k = 0; k < n; k++) buf[k] = t[n - 1 - k];
    add_md(P, a, buf, (size_t)n);
}

/* The fast DP cascade's record (accept_fast_result, :289-299): read span
 * 0 .. L, MAPQ 0.  Returns the cascade's kind (1 certificate, 2 WFA) or 0. */
static int fast_record(part *P, int32_t r, const uint8_t *read, uint32_t L, const ko_chain *ch,
                       const driver *D, scratch *s) {
    const ko_s4_in *in = D->in;
    ko_fast_out o;
    const uint32_t *cig;
    const char *md;
    if (!ko_fast_dp_rec(read, L, contig_of(in, ch->ref_id), (size_t)in->contig_len[ch->ref_id],
                        ch->ref_start, ch->read_start, ch->score, &D->c->sw, &D->c->router,
                        s->fast, &o, &cig, &md))
        return 0;
    ko_s4_rec *a = add_rec(P, r, 1, ch->ref_id, ch->strand == 1);
    a->ref_start = o.ref_start;
    a->ref_end = o.ref_end;
    a->read_start = 0;
    a->read_end = L;
    a->score = o.score;
    a->nm = o.nm;
    add_cig(P, a, cig, o.n_cigar);
    add_md(P, a, md, strlen(md));
    return o.kind;
}

/* build_alignment (mod.rs:1200-1254) of an SW result, as a DpAligned record */
static void sw_record(part *P, int32_t r, const uint8_t *read, uint32_t L, const uint8_t *win,
                      uint32_t wlen, uint32_t win_start, const ko_sw_result *sw, uint32_t ref_id,
                      int is_rev, int supp, scratch *s) {
    ko_alignment b;
    const size_t cap = (size_t)sw->n_cigar + 2, mdcap = 4 * ((size_t)L + wlen) + 32;
    grow((void **)&s->cig, &s->cig_cap, cap * 4);
    grow((void **)&s->md, &s->md_cap, mdcap);
    ko_build_alignment(read, L, win, wlen, win_start, sw, &b, s->cig, (int32_t)cap, s->md, mdcap);
    ko_s4_rec *a = add_rec(P, r, 1, ref_id, is_rev);
    a->ref_start = b.ref_start;
    a->ref_end = b.ref_end;
    a->read_start = b.read_start;
    a->read_end = b.read_end;
    a->score = b.score;
    a->nm = b.nm;
    a->is_supplementary = supp;
    add_cig(P, a, s->cig, b.n_cigar);
    add_md(P, a, s->md, strlen(s->md));
}

/* align_chain_with_meta (mod.rs:835-883): the exact shortcut, else banded SW
 * on the clamp_window window (mod.rs:1766-1771).  Returns early_abort. */
static int scalar_job(part *P, const driver *D, const job *j, const uint8_t *read, uint32_t L,
                      int supp, scratch *s) {
    const ko_s4_in *in = D->in;
    const uint8_t *contig = contig_of(in, j->ref_id);
    const int64_t clen = in->contig_len[j->ref_id];
    if (exact_match(read, L, contig, clen, &j->span)) {
        exact_record(P, j->read, read, L, &j->span, j->ref_id, j->is_rev, D->c, supp);
        return 0;
    }
    const uint32_t flank = (uint32_t)imax(imax(D->c->sw.bandwidth, 1) * 2, 50);
    const uint32_t ws = usub(j->span.ref_start, flank);
    int64_t we = (int64_t)j->span.ref_end + flank;
    if (we > clen) we = clen;
    if (we < (int64_t)ws + 1) we = (int64_t)ws + 1;
    const uint32_t wlen = (uint32_t)(we - ws);
    const int64_t off = (int64_t)j->span.ref_start - ws - j->span.read_start;
    const size_t cap = (size_t)L + wlen + 4;
    grow((void **)&s->swcig, &s->swcig_cap, cap * 4);
    ko_sw_result sw;
    memset(&sw, 0, sizeof sw);
    sw.cigar = s->swcig;
    sw.cigar_cap = (int32_t)cap;
    ko_banded_sw(read, L, contig + ws, wlen, (int32_t)off, &D->c->sw, j->abort, &D->kn, &sw);
    sw_record(P, j->read, read, L, contig + ws, wlen, ws, &sw, j->ref_id, j->is_rev, supp, s);
    return sw.early_abort != 0;
}

/* ungapped_beaten_by_gap (mod.rs:163-204), the prefilter's final gate with
 * KIRA_UNGAP_GAP_CHECK (prefilter.rs:274-287) */
static int beaten_by_gap(const uint8_t *read, uint32_t L, const uint8_t *contig, int64_t clen,
                         const ko_span *sp, uint32_t mism, const ko_cfg *c) {
    if (mism == 0 || L == 0) return 0;
    const int64_t ws = (int64_t)sp->ref_start - sp->read_start;
    if (ws < 0 || ws >= clen) return 0;
    int64_t tl = (int64_t)L + imax(c->bandwidth, 1);
    if (tl > clen - ws) tl = clen - ws;
    if (tl < (int64_t)L) return 0;
    const int64_t cost = (int64_t)mism * ((int64_t)c->match_score + c->mismatch);
    if (cost <= 0) return 0;
    int32_t score;
    return ko_wfa_score_only_scored(read, (int32_t)L, contig + ws, (int32_t)tl, c,
                                    (int32_t)(cost - 1), 0, 0, &score) == 1;
}

/* ------------------------------------------------- per read (phase 1) */
typedef struct {
    int32_t ci;    /* chain index */
    int32_t score; /* A * L for an exact copy, the cascade's score, or I32_MIN */
    int32_t has;   /* an exact copy or a cascade result */
    int32_t slot;  /* the cascade record in the read's cache, or -1 */
    int32_t kind;  /* its kind (1 certificate, 2 WFA) */
} cand;

/* by score, best first; ties by chain index */
static int cand_cmp(const void *x, const void *y) {
    const cand *a = (const cand *)x, *b = (const cand *)y;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    return a->ci < b->ci ? -1 : (a->ci > b->ci);
}

/* two_tier_reorder (:346-451) over the np passing candidates cd[].ci: exact
 * copies score A * L, the rest the cascade's alignment score (its record is
 * kept in `cache`); sorts cd, writes the chain order and returns the DP
 * budget */
static int32_t two_tier(const driver *D, int32_t r, const ko_chain *cl, int32_t nc, cand *cd,
                        int32_t np, part *cache, int32_t *order, scratch *s) {
    const ko_s4_in *in = D->in;
    const ko_s4_cfg *c = D->c;
    const uint32_t L = (uint32_t)in->read_len[r];
    int any_none = 0;
    for (int32_t i = 0; i < np; i++) {
        const ko_chain *ch = &cl[cd[i].ci];
        const uint8_t *rd = oriented(in, r, ch->strand == 1, s);
        const ko_span sp = {ch->ref_start, ch->ref_end, ch->read_start, ch->read_end};
        cd[i].has = 1;
        cd[i].slot = -1;
        cd[i].kind = 0;
        if (exact_match(rd, L, contig_of(in, ch->ref_id), in->contig_len[ch->ref_id], &sp)) {
            cd[i].score = wrap_i32((int64_t)c->sw.match_score * L);
            continue;
        }
        const int32_t at = (int32_t)cache->recs.n;
        const int kind = fast_record(cache, r, rd, L, ch, D, s);
        if (kind) {
            cd[i].slot = at;
            cd[i].kind = kind;
            cd[i].score = ((const ko_s4_rec *)cache->recs.p)[at].score;
        } else {
            cd[i].has = 0;
            cd[i].score = I32_MIN_;
            any_none = 1;
        }
    }
    qsort(cd, (size_t)np, sizeof(cand), cand_cmp);
    int32_t dp_limit;
    if (any_none || (np >= 3 && cd[2].score == cd[0].score)) {
        dp_limit = 2;
    } else {
        const int32_t mrg =
            imax(sat_i32(((int64_t)c->sw.match_score + c->sw.mismatch) * 2), 1);
        const int32_t top = sat_i32((int64_t)cd[0].score - mrg);
        int32_t n_close = 0;
        for (int32_t i = 0; i < np && cd[i].score >= top; i++) n_close++;
        dp_limit = imin(imax(n_close, 1), 3);
    }
    uint8_t *listed = (uint8_t *)calloc((size_t)nc, 1);
    int32_t o = 0;
    for (int32_t i = 0; i < np; i++) {
        order[o++] = cd[i].ci;
        listed[cd[i].ci] = 1;
    }
    for (int32_t k = 0; k < nc; k++)
        if (!listed[k]) order[o++] = k;
    free(listed);
    return dp_limit;
}

/* process_read_prefilter (:535-957) for read r: records into the chunk's
 * part, fallback jobs into the chunk's lists */
static void one_read(driver *D, int32_t r, int32_t chunk, scratch *s) {
    const ko_s4_in *in = D->in;
    const ko_s4_cfg *c = D->c;
    part *P = &D->p1[chunk];
    counts *C = &D->cnt[chunk];
    const int64_t c0 = in->chain_off[r];
    const int32_t nc = (int32_t)(in->chain_off[r + 1] - c0);
    const ko_chain *cl = in->chains + c0;
    const uint32_t L = (uint32_t)in->read_len[r];
    const int bucket = L <= 120 ? 0 : (L <= 150 ? 1 : 2);
    D->chain_best[r] = D->chain_second[r] = 0;
    D->has_comp[r] = 0;
    if (nc == 0) return;
    /* select_chains (:549-708), mate guidance off */
    const int32_t best = cl[0].score;
    int32_t min_score = best > 0 ? (int32_t)ceilf((float)best * c->min_chain_ratio) : best;
    const int32_t second = nc > 1 ? cl[1].score : 0;
    const int32_t margin = nc > 1 ? wrap_i32((int64_t)best - second) : I32_MAX_;
    D->chain_best[r] = best;
    D->chain_second[r] = second;
    const int short_read = L <= 300;
    const int32_t q = best / (c->ambig_div != 0 ? c->ambig_div : 5);
    const int rmo = in->repeat_min_occ ? in->repeat_min_occ[r] : 0;
    const int ambiguous = nc > 1 && (margin <= imax(q, 8) || rmo > 1);
    const int weak = nc > 1 && best < (int32_t)(2 * L) / 3;
    if (short_read && weak && c->twotier) min_score = imin(min_score, best / 5);
    const int override_ = c->short_dp_topk >= 1;
    int32_t topk = short_read ? (override_ ? c->short_dp_topk : (ambiguous ? 2 : 1))
                              : imax(c->dp_topk, 1);
    int32_t *order = (int32_t *)malloc((size_t)nc * 4);
    for (int32_t k = 0; k < nc; k++) order[k] = k;
    part cache;
    part_init(&cache);
    cand cd[MAX_CAND];
    int32_t np = 0;
    if (c->twotier && short_read && (ambiguous || weak)) {
        for (int32_t k = 0; k < nc && np < c->twotier_k && np < MAX_CAND; k++)
            if (cl[k].score >= min_score) cd[np++].ci = k;
        if (np < 2) np = 0;
    }
    if (np) {
        const int32_t dp_limit = two_tier(D, r, cl, nc, cd, np, &cache, order, s);
        if (!override_) topk = dp_limit;
        if (topk == 1) {
            /* the runner-up's score becomes XS (:673-683) */
            int have = 0;
            int32_t mx = 0;
            for (int32_t i = 0; i < np; i++) {
                if (cd[i].ci == order[0] || !cd[i].has) continue;
                mx = have ? imax(mx, cd[i].score) : cd[i].score;
                have = 1;
            }
            D->has_comp[r] = have;
            D->comp[r] = mx;
        }
    }
    /* the selected chains, in order (the candidates are order[0 .. np)) */
    int dp_used = 0;
    int32_t rank = 0;
    for (int32_t p = 0; p < nc; p++) {
        const int32_t ci = order[p];
        const ko_chain *ch = &cl[ci];
        if (ch->score < min_score) continue;
        if (rank >= topk) break;
        const int32_t this_rank = rank++;
        C->used++;
        const int is_rev = ch->strand == 1;
        const uint8_t *rd = oriented(in, r, is_rev, s);
        const uint8_t *contig = contig_of(in, ch->ref_id);
        const int64_t clen = in->contig_len[ch->ref_id];
        const ko_span sp = {ch->ref_start, ch->ref_end, ch->read_start, ch->read_end};
        if (exact_match(rd, L, contig, clen, &sp)) {
            C->exact++;
            exact_record(P, r, rd, L, &sp, ch->ref_id, is_rev, c, 0);
            continue;
        }
        if (p < np && cd[p].slot >= 0) {
            /* a complete alignment the candidate search verified (:736-747) */
            const ko_s4_rec *x = &((const ko_s4_rec *)cache.recs.p)[cd[p].slot];
            ko_s4_rec *a = add_rec(P, r, 1, ch->ref_id, is_rev);
            const int64_t co = a->cigar_off, mo = a->md_off;
            *a = *x;
            a->cigar_off = co;
            a->md_off = mo;
            a->n_cigar = 0;
            add_cig(P, a, (const uint32_t *)cache.cig.p + x->cigar_off, x->n_cigar);
            add_md(P, a, (const char *)cache.md.p + x->md_off, (size_t)x->md_len);
            if (cd[p].kind == 1)
                C->packed++;
            else
                C->wfa++;
            dp_used = 1;
            continue;
        }
        ko_prefilter_out o;
        const size_t mdcap = 4 * (size_t)L + 32;
        grow((void **)&s->md, &s->md_cap, mdcap);
        ko_prefilter_chain(rd, L, contig, (size_t)clen, &sp, &c->sw, this_rank == 0,
                           c->accept_enable && c->max_alignments <= 1, c->accept_only_top1,
                           (size_t)c->accept_span_slack, c->accept_min_identity,
                           (size_t)c->accept_max_mismatches, c->accept_require_score_margin,
                           margin, (size_t)c->long_read_threshold, c->max_alignments > 1,
                           c->tail_ext, c->ungap_min_id, &o, s->md, mdcap);
        if (c->ungap_gap_check && o.result == 0 && o.mism > 0 &&
            beaten_by_gap(rd, L, contig, clen, &sp, o.mism, &c->sw)) {
            o.result = 2; /* Fallback (IndelSuspect) */
            o.reason = 8;
        }
        C->reason[o.reason < 9 ? o.reason : 9]++;
        const int32_t abort =
            o.has_metrics ? sat_i32((int64_t)o.score - c->dp_abort_margin) : I32_MIN_ / 4;
        if (o.result == 0) {
            /* build_ungapped_alignment (prefilter.rs:455-517) */
            C->accept++;
            ko_s4_rec *a = add_rec(P, r, 0, ch->ref_id, is_rev);
            a->ref_start = o.aln_ref_start;
            a->ref_end = o.aln_ref_end;
            a->read_start = o.aln_read_start;
            a->read_end = o.aln_read_end;
            a->score = o.aln_score;
            a->mapq = 60;
            a->nm = o.aln_nm;
            add_cig(P, a, o.cigar, o.n_cigar);
            add_md(P, a, s->md, strlen(s->md));
            continue;
        }
        if (o.result == 1) {
            C->reject++;
            continue;
        }
        C->fallback++;
        dp_used = 1;
        /* without the GPU dispatcher the fast DP cascade comes first
         * (try_fast_dp_alignment, :868-882) */
        const int kind = fast_record(P, r, rd, L, ch, D, s);
        if (kind) {
            if (kind == 1)
                C->packed++;
            else
                C->wfa++;
            continue;
        }
        job j;
        memset(&j, 0, sizeof j);
        j.read = r;
        j.ref_id = ch->ref_id;
        j.is_rev = is_rev;
        j.span = sp;
        j.abort = abort;
        if (bucket < 2 && D->lanes > 1 &&
            ko_build_simd_window((uint64_t)clen, ch->ref_start, ch->read_start, L,
                                 c->sw.bandwidth, &j.win_start, &j.win_len)) {
            *(job *)vec_push(&D->lane_jobs[chunk], 1) = j;
            continue;
        }
        *(job *)vec_push(&D->scalar_jobs[chunk], 1) = j;
    }
    C->dp_reads += dp_used;
    part_free(&cache);
    free(order);
}

static void *read_worker(void *arg) {
    driver *D = (driver *)arg;
    scratch s;
    scratch_init(&s);
    for (;;) {
        const int32_t k = __atomic_fetch_add(&D->next, 1, __ATOMIC_RELAXED);
        if (k >= D->n_chunks) break;
        const int32_t r1 = imin((k + 1) * READ_CHUNK, D->in->n_reads);
        for (int32_t r = k * READ_CHUNK; r < r1; r++) one_read(D, r, k, &s);
    }
    scratch_free(&s);
    return NULL;
}

/* ------------------------------------------------ lane groups (phase 2) */
typedef struct {
    const driver *D;
    const job *jobs;
    const int32_t *order;     /* jobs by (read length, window length), stable */
    const int32_t *group_off; /* group g = order[group_off[g] .. group_off[g + 1]) */
    int32_t n_groups;
    part *parts; /* one per group */
    int64_t *aborts;
    int32_t next;
} lane_work;

/* align_batch_simd (mod.rs:1111-1149) on one group of <= lanes jobs */
static void *lane_worker(void *arg) {
    lane_work *W = (lane_work *)arg;
    const driver *D = W->D;
    const ko_s4_in *in = D->in;
    scratch s;
    scratch_init(&s);
    ko_batch_input bi[32];
    ko_sw_result res[32];
    uint8_t *reads = NULL;
    size_t reads_cap = 0;
    uint32_t *cig = NULL;
    size_t cig_cap = 0;
    for (;;) {
        const int32_t g = __atomic_fetch_add(&W->next, 1, __ATOMIC_RELAXED);
        if (g >= W->n_groups) break;
        const int32_t a0 = W->group_off[g], n = W->group_off[g + 1] - a0;
        size_t total = 0, words = 0;
        for (int32_t t = 0; t < n; t++) {
            const job *j = &W->jobs[W->order[a0 + t]];
            total += (size_t)in->read_len[j->read];
            words += (size_t)in->read_len[j->read] + j->win_len + 4;
        }
        grow((void **)&reads, &reads_cap, total + 1);
        grow((void **)&cig, &cig_cap, words * 4);
        size_t at = 0, co = 0;
        for (int32_t t = 0; t < n; t++) {
            const job *j = &W->jobs[W->order[a0 + t]];
            const uint32_t L = (uint32_t)in->read_len[j->read];
            memcpy(reads + at, oriented(in, j->read, j->is_rev, &s), L);
            bi[t].read = reads + at;
            bi[t].read_len = L;
            bi[t].window = contig_of(in, j->ref_id) + j->win_start;
            bi[t].window_len = j->win_len;
            bi[t].win_start = j->win_start;
            bi[t].chain_ref_start = j->span.ref_start;
            bi[t].chain_ref_end = j->span.ref_end;
            bi[t].chain_read_start = j->span.read_start;
            bi[t].chain_read_end = j->span.read_end;
            bi[t].abort_score = j->abort;
            at += L;
            memset(&res[t], 0, sizeof res[t]);
            res[t].cigar = cig + co;
            res[t].cigar_cap = (int32_t)(L + j->win_len + 4);
            co += (size_t)res[t].cigar_cap;
        }
        ko_align_batch_sw(bi, n, &D->c->sw, D->c->simd, D->c->simd, &D->kn, res);
        for (int32_t t = 0; t < n; t++) {
            const job *j = &W->jobs[W->order[a0 + t]];
            sw_record(&W->parts[g], j->read, bi[t].read, bi[t].read_len, bi[t].window,
                      bi[t].window_len, j->win_start, &res[t], j->ref_id, j->is_rev, 0, &s);
            W->aborts[g] += res[t].early_abort != 0;
        }
    }
    free(reads);
    free(cig);
    scratch_free(&s);
    return NULL;
}

/* ---------------------------------------- banded SW and supplementary */
typedef struct {
    const driver *D;
    const job *jobs;
    int32_t n, n_chunks, supp;
    part *parts; /* one per chunk of JOB_CHUNK jobs */
    int64_t *aborts;
    int32_t next;
} job_work;

static void *job_worker(void *arg) {
    job_work *W = (job_work *)arg;
    scratch s;
    scratch_init(&s);
    for (;;) {
        const int32_t k = __atomic_fetch_add(&W->next, 1, __ATOMIC_RELAXED);
        if (k >= W->n_chunks) break;
        const int32_t t1 = imin((k + 1) * JOB_CHUNK, W->n);
        for (int32_t t = k * JOB_CHUNK; t < t1; t++) {
            const job *j = &W->jobs[t];
            const uint8_t *rd = oriented(W->D->in, j->read, j->is_rev, &s);
            W->aborts[k] += scalar_job(&W->parts[k], W->D, j, rd,
                                       (uint32_t)W->D->in->read_len[j->read], W->supp, &s);
        }
    }
    scratch_free(&s);
    return NULL;
}

static void run_threads(void *(*fn)(void *), void *arg, int threads) {
    if (threads <= 1) {
        fn(arg);
        return;
    }
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int k = 0; k < threads; k++) pthread_create(&th[k], NULL, fn, arg);
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    free(th);
}

/* runs the jobs (supp: as supplementary records) and returns their parts */
static part *run_jobs(const driver *D, const job *jobs, int32_t n, int supp, int threads,
                      int32_t *n_parts, int64_t *aborts) {
    job_work W;
    memset(&W, 0, sizeof W);
    W.D = D;
    W.jobs = jobs;
    W.n = n;
    W.n_chunks = (n + JOB_CHUNK - 1) / JOB_CHUNK;
    W.supp = supp;
    W.parts = parts_new(W.n_chunks);
    W.aborts = (int64_t *)calloc((size_t)W.n_chunks + 1, 8);
    run_threads(job_worker, &W, threads);
    for (int32_t k = 0; k < W.n_chunks; k++) *aborts += W.aborts[k];
    free(W.aborts);
    *n_parts = W.n_chunks;
    return W.parts;
}

/* ---------------------------------------------------------- the result */
struct ko_s4_result {
    ko_s4_view v;
    int64_t *aln_off;
    ko_s4_rec *alns;
    uint32_t *cigar;
    char *md;
};

/* all parts in the order their records are appended */
typedef struct {
    part **p;
    int32_t n, cap;
} part_list;

static void pl_add(part_list *l, part *p) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 256;
        l->p = (part **)realloc(l->p, (size_t)l->cap * sizeof(part *));
    }
    l->p[l->n++] = p;
}

/* records per read over the parts, as offsets: off[r] .. off[r + 1] */
static void count_by_read(const part_list *l, int32_t n, int64_t *off) {
    memset(off, 0, ((size_t)n + 1) * 8);
    for (int32_t q = 0; q < l->n; q++) {
        const ko_s4_rec *x = (const ko_s4_rec *)l->p[q]->recs.p;
        for (size_t k = 0; k < l->p[q]->recs.n; k++) off[x[k].read + 1]++;
    }
    for (int32_t r = 0; r < n; r++) off[r + 1] += off[r];
}

/* split_read.rs:128-148 */
static void orig_interval(uint32_t s, uint32_t e, int rev, uint32_t L, uint32_t *a, uint32_t *b) {
    if (rev) {
        *a = usub(L, e);
        *b = usub(L, s);
    } else {
        *a = s;
        *b = e;
    }
}

static uint32_t overlap_pct(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    const uint32_t al = usub(a1, a0), bl = usub(b1, b0);
    if (al == 0 || bl == 0) return 0;
    const uint32_t s = a0 > b0 ? a0 : b0, e = a1 < b1 ? a1 : b1;
    if (e <= s) return 0;
    uint64_t ov = (uint64_t)(e - s) * 100u;
    if (ov > 0xFFFFFFFFu) ov = 0xFFFFFFFFu;
    return (uint32_t)ov / al;
}

/* pick_supplementary_chains (split_read.rs:60-125) for a read whose first
 * record is p with CIGAR cig; returns the number of picks */
static int32_t pick_supp(const ko_s4_rec *p, const uint32_t *cig, const ko_chain *cl, int32_t nc,
                         int32_t *picks) {
    const int32_t ps = imax(cl[0].score, 1);
    const int32_t min_score = (int32_t)ceilf((float)ps * SUPP_MIN_SCORE_RATIO);
    uint32_t L = 0;
    for (int32_t k = 0; k < p->n_cigar; k++) {
        const uint32_t op = cig[k] & 15u;
        if (op == KO_M || op == KO_I || op == KO_S) L += cig[k] >> 4;
    }
    uint32_t pa, pb;
    orig_interval(p->read_start, p->read_end, p->is_rev, L, &pa, &pb);
    int32_t np = 0;
    for (int32_t k = 1; k < nc; k++) {
        const ko_chain *ch = &cl[k];
        if (ch->score < min_score) break;
        if (usub(ch->read_end, ch->read_start) < SUPP_MIN_READ_SPAN) continue;
        uint32_t ca, cb;
        orig_interval(ch->read_start, ch->read_end, ch->strand == 1, L, &ca, &cb);
        if (overlap_pct(pa, pb, ca, cb) > SUPP_MAX_OVERLAP_PCT) continue;
        int dup = 0;
        for (int32_t q = 0; q < np && !dup; q++) {
            const ko_chain *pc = &cl[picks[q]];
            uint32_t qa, qb;
            orig_interval(pc->read_start, pc->read_end, pc->strand == 1, L, &qa, &qb);
            dup = overlap_pct(qa, qb, ca, cb) > SUPP_MAX_OVERLAP_PCT;
        }
        if (dup) continue;
        picks[np++] = k;
        if (np >= MAX_SUPP_PER_READ) break;
    }
    return np;
}

/* merge sort of job indices by key (stable) */
static void sort_by_key(int32_t *idx, const uint64_t *key, int32_t n) {
    int32_t *tmp = (int32_t *)malloc(((size_t)n + 1) * 4);
    for (int32_t w = 1; w < n; w *= 2) {
        for (int32_t lo = 0; lo < n; lo += 2 * w) {
            const int32_t mid = imin(lo + w, n), hi = imin(lo + 2 * w, n);
            int32_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) tmp[o++] = key[idx[b]] < key[idx[a]] ? idx[b++] : idx[a++];
            while (a < mid) tmp[o++] = idx[a++];
            while (b < hi) tmp[o++] = idx[b++];
        }
        memcpy(idx, tmp, (size_t)n * 4);
    }
    free(tmp);
}

ko_s4_result *ko_stage4_run_mt(const ko_s4_in *in, const ko_s4_cfg *c, int threads) {
    if (threads < 1) threads = 1;
    const int32_t n = in->n_reads;
    driver D;
    memset(&D, 0, sizeof D);
    D.in = in;
    D.c = c;
    D.kn.adaptive = c->adaptive;
    D.kn.max_drift = c->max_drift;
    D.kn.dp_xdrop = c->dp_xdrop;
    D.lanes = c->simd == 2 ? 32 : (c->simd == 1 ? 16 : (c->simd == 3 ? 4 : 1));
    D.chain_best = (int32_t *)calloc((size_t)n + 1, 4);
    D.chain_second = (int32_t *)calloc((size_t)n + 1, 4);
    D.has_comp = (int32_t *)calloc((size_t)n + 1, 4);
    D.comp = (int32_t *)calloc((size_t)n + 1, 4);
    D.n_chunks = (n + READ_CHUNK - 1) / READ_CHUNK;
    D.p1 = parts_new(D.n_chunks);
    D.lane_jobs = (vec *)calloc((size_t)D.n_chunks + 1, sizeof(vec));
    D.scalar_jobs = (vec *)calloc((size_t)D.n_chunks + 1, sizeof(vec));
    D.cnt = (counts *)calloc((size_t)D.n_chunks + 1, sizeof(counts));
    for (int32_t k = 0; k < D.n_chunks; k++) {
        vec_init(&D.lane_jobs[k], sizeof(job));
        vec_init(&D.scalar_jobs[k], sizeof(job));
    }
    ko_s4_result *R = (ko_s4_result *)calloc(1, sizeof(ko_s4_result));
    ko_s4_stats *st = &R->v.stats;
    st->reads = n;
    st->chains_total = in->chain_off[n] - in->chain_off[0];

    /* phase 1: every read (process_read_prefilter over the batch) */
    run_threads(read_worker, &D, threads);
    part_list order = {NULL, 0, 0};
    vec lanes, scal;
    vec_init(&lanes, sizeof(job));
    vec_init(&scal, sizeof(job));
    for (int32_t k = 0; k < D.n_chunks; k++) {
        pl_add(&order, &D.p1[k]);
        if (D.lane_jobs[k].n)
            memcpy(vec_push(&lanes, D.lane_jobs[k].n), D.lane_jobs[k].p,
                   D.lane_jobs[k].n * sizeof(job));
        if (D.scalar_jobs[k].n)
            memcpy(vec_push(&scal, D.scalar_jobs[k].n), D.scalar_jobs[k].p,
                   D.scalar_jobs[k].n * sizeof(job));
        vec_free(&D.lane_jobs[k]);
        vec_free(&D.scalar_jobs[k]);
        const counts *C = &D.cnt[k];
        st->chains_used += C->used;
        st->exact_matches += C->exact;
        st->prefilter_accept += C->accept;
        st->prefilter_reject += C->reject;
        st->prefilter_fallback += C->fallback;
        st->packed_spectral_resolved += C->packed;
        st->wfa_resolved += C->wfa;
        st->dp_reads += C->dp_reads;
        for (int q = 0; q < 10; q++) st->reason_counts[q] += C->reason[q];
    }

    /* phase 2: lane groups, a BTreeMap over (read length, window length) cut
     * into chunks of `lanes` jobs (:1153-1186) */
    const job *lj = (const job *)lanes.p;
    const int32_t nl = (int32_t)lanes.n;
    int32_t *lorder = (int32_t *)malloc(((size_t)nl + 1) * 4);
    uint64_t *key = (uint64_t *)malloc(((size_t)nl + 1) * 8);
    for (int32_t t = 0; t < nl; t++) {
        key[t] = ((uint64_t)(uint32_t)in->read_len[lj[t].read] << 32) | lj[t].win_len;
        lorder[t] = t;
    }
    sort_by_key(lorder, key, nl);
    int32_t *goff = (int32_t *)malloc(((size_t)nl + 2) * 4);
    int32_t ng = 0;
    goff[0] = 0;
    for (int32_t i = 0; i < nl;) {
        int32_t e = i;
        while (e < nl && key[lorder[e]] == key[lorder[i]]) e++;
        for (int32_t b = i; b < e; b += D.lanes) goff[++ng] = imin(e, b + D.lanes);
        i = e;
    }
    free(key);
    lane_work LW;
    memset(&LW, 0, sizeof LW);
    LW.D = &D;
    LW.jobs = lj;
    LW.order = lorder;
    LW.group_off = goff;
    LW.n_groups = ng;
    LW.parts = parts_new(ng);
    LW.aborts = (int64_t *)calloc((size_t)ng + 1, 8);
    run_threads(lane_worker, &LW, threads);
    for (int32_t g = 0; g < ng; g++) {
        pl_add(&order, &LW.parts[g]);
        st->dp_early_abort += LW.aborts[g];
    }
    st->dp_simd = nl;
    free(LW.aborts);
    free(goff);
    free(lorder);

    /* phase 3: banded SW of the other fallback jobs (:1196-1205) */
    int32_t n_sp = 0;
    part *sparts = run_jobs(&D, (const job *)scal.p, (int32_t)scal.n, 0, threads, &n_sp,
                            &st->dp_early_abort);
    for (int32_t k = 0; k < n_sp; k++) pl_add(&order, &sparts[k]);
    st->dp_scalar = (int64_t)scal.n;

    /* each read's first record and its CIGAR */
    int64_t *off = (int64_t *)malloc(((size_t)n + 1) * 8);
    count_by_read(&order, n, off);
    ko_s4_rec **first = (ko_s4_rec **)calloc((size_t)n + 1, sizeof(ko_s4_rec *));
    const uint32_t **first_cig = (const uint32_t **)calloc((size_t)n + 1, sizeof(uint32_t *));
    for (int32_t q = order.n - 1; q >= 0; q--) {
        ko_s4_rec *x = (ko_s4_rec *)order.p[q]->recs.p;
        for (size_t k = order.p[q]->recs.n; k-- > 0;) {
            first[x[k].read] = &x[k];
            first_cig[x[k].read] = (const uint32_t *)order.p[q]->cig.p + x[k].cigar_off;
        }
    }

    /* XS on reads with exactly one record: the two-tier competitor's score,
     * else the chain-score proxy (:1215-1250) */
    for (int32_t r = 0; r < n; r++) {
        if (off[r + 1] - off[r] != 1) continue;
        ko_s4_rec *a = first[r];
        if (D.has_comp[r]) {
            a->has_xs = 1;
            a->xs_score = imin(imax(D.comp[r], 0), imax(a->score, 0));
            continue;
        }
        const int32_t cb = D.chain_best[r], cs = D.chain_second[r];
        if (cb <= 0 || cs <= 0) continue;
        const int32_t sub = wrap_i32((int64_t)a->score * cs / cb);
        if (sub > 0 && (int64_t)cs * 100 >= (int64_t)cb * c->xs_min_ratio_pct) {
            a->has_xs = 1;
            a->xs_score = sub;
        }
    }

    /* emit_supplementary_alignments (split_read.rs:22-57) */
    part *supp = NULL;
    int32_t n_supp = 0;
    vec sj;
    vec_init(&sj, sizeof(job));
    if (c->supplementary) {
        for (int32_t r = 0; r < n; r++) {
            const int64_t c0 = in->chain_off[r];
            const int32_t nc = (int32_t)(in->chain_off[r + 1] - c0);
            if (off[r + 1] == off[r] || nc < 2) continue;
            int32_t picks[MAX_SUPP_PER_READ];
            const int32_t np = pick_supp(first[r], first_cig[r], in->chains + c0, nc, picks);
            for (int32_t q = 0; q < np; q++) {
                const ko_chain *ch = &in->chains[c0 + picks[q]];
                job *j = (job *)vec_push(&sj, 1);
                memset(j, 0, sizeof *j);
                j->read = r;
                j->ref_id = ch->ref_id;
                j->is_rev = ch->strand == 1;
                j->span.ref_start = ch->ref_start;
                j->span.ref_end = ch->ref_end;
                j->span.read_start = ch->read_start;
                j->span.read_end = ch->read_end;
                j->abort = I32_MIN_ / 8;
            }
        }
        int64_t ignored = 0;
        supp = run_jobs(&D, (const job *)sj.p, (int32_t)sj.n, 1, threads, &n_supp, &ignored);
        for (int32_t k = 0; k < n_supp; k++) pl_add(&order, &supp[k]);
        st->supplementary = (int64_t)sj.n;
        count_by_read(&order, n, off);
    }

    /* records grouped by read, in the order they were appended; CIGAR and MD
     * arenas joined */
    int64_t cw = 0, mb = 0;
    for (int32_t q = 0; q < order.n; q++) {
        cw += (int64_t)order.p[q]->cig.n;
        mb += (int64_t)order.p[q]->md.n;
    }
    R->aln_off = off;
    R->alns = (ko_s4_rec *)malloc(((size_t)off[n] + 1) * sizeof(ko_s4_rec));
    R->cigar = (uint32_t *)malloc(((size_t)cw + 1) * 4);
    R->md = (char *)malloc((size_t)mb + 1);
    int64_t *pos = (int64_t *)malloc(((size_t)n + 1) * 8);
    memcpy(pos, off, ((size_t)n + 1) * 8);
    cw = mb = 0;
    for (int32_t q = 0; q < order.n; q++) {
        const part *P = order.p[q];
        const ko_s4_rec *x = (const ko_s4_rec *)P->recs.p;
        for (size_t k = 0; k < P->recs.n; k++) {
            ko_s4_rec a = x[k];
            a.cigar_off += cw;
            a.md_off += mb;
            R->alns[pos[a.read]++] = a;
        }
        if (P->cig.n) memcpy(R->cigar + cw, P->cig.p, P->cig.n * 4);
        if (P->md.n) memcpy(R->md + mb, P->md.p, P->md.n);
        cw += (int64_t)P->cig.n;
        mb += (int64_t)P->md.n;
    }
    free(pos);
    for (int32_t r = 0; r < n; r++) st->unmapped += off[r + 1] == off[r];
    R->v.n_reads = n;
    R->v.aln_off = R->aln_off;
    R->v.alns = R->alns;
    R->v.cigar = R->cigar;
    R->v.md = R->md;
    R->v.n_alns = off[n];
    R->v.cigar_words = cw;
    R->v.md_bytes = mb;

    for (int32_t q = 0; q < order.n; q++) part_free(order.p[q]);
    free(order.p);
    free(D.p1);
    free(LW.parts);
    free(sparts);
    free(supp);
    free(D.lane_jobs);
    free(D.scalar_jobs);
    free(D.cnt);
    vec_free(&lanes);
    vec_free(&scal);
    vec_free(&sj);
    free(first);
    free(first_cig);
    free(D.chain_best);
    free(D.chain_second);
    free(D.has_comp);
    free(D.comp);
    return R;
}

void ko_stage4_view(const ko_s4_result *r, ko_s4_view *v) { *v = r->v; }

void ko_stage4_free(ko_s4_result *r) {
    if (!r) return;
    free(r->aln_off);
    free(r->alns);
    free(r->cigar);
    free(r->md);
    free(r);
}
