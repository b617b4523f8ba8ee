/*
 * kira_fast_oracle.c — the reference's fast DP cascade in C, for the CPU
 * baseline timed beside the GPU fast path (and as a second restatement of
 * oracle/fastpath.py, which the tests pin).  Test infrastructure only: the
 * product never links or calls it.
 *
 * try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440): packed
 * spectral certificate (try_packed_spectral_alignment, mod.rs:624-686) ->
 * Myers reject (router.rs:196-202) -> scored WFA with the suspect-origin
 * retry (mod.rs:350-420) -> partial-read check (clipping_can_improve,
 * mod.rs:442-479) -> wfa_result_to_alignment (mod.rs:481-575), composed from
 * the oracle's pinned primitives (ko_scan_best_with_second,
 * ko_spectral_hit_is_certified, ko_myers_bounded, ko_wfa_align_scored,
 * ko_nm_md_from_cigar).  ko_fast_dp_batch_mt runs it over many chains on
 * pthreads (chunks of 64 chains handed out through an atomic counter).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "kira_oracle.h"

static int32_t imin32(int32_t a, int32_t b) { return a < b ? a : b; }
static int32_t imax32(int32_t a, int32_t b) { return a > b ? a : b; }
static int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }

/* worker scratch, grown on demand */
typedef struct {
    uint8_t *rb, *tb, *ph;
    size_t rb_cap, tb_cap, ph_cap;
    uint32_t *cig, *cig2;
    size_t cig_cap, cig2_cap;
    char *md;
    size_t md_cap;
} scratch;

static void grow(void **p, size_t *cap, size_t need) {
    if (*cap >= need) return;
    free(*p);
    *p = malloc(need);
    *cap = need;
}

static void scratch_free(scratch *s) {
    free(s->rb);
    free(s->tb);
    free(s->ph);
    free(s->cig);
    free(s->cig2);
    free(s->md);
}

/* try_packed_spectral_alignment (mod.rs:624-686): both sequences pack
 * (A/C/G/T only), the certified best shift becomes {L}M. */
static int spectral(const uint8_t *read, size_t L, const uint8_t *text, size_t tl,
                    uint32_t win_start, const ko_cfg *c, const ko_router *r, scratch *s,
                    ko_fast_out *out) {
    const size_t rbytes = (L + 3) / 4, tbytes = (tl + 3) / 4;
    grow((void **)&s->rb, &s->rb_cap, rbytes + 8);
    grow((void **)&s->tb, &s->tb_cap, tbytes + 8);
    grow((void **)&s->ph, &s->ph_cap, 4 * tbytes + 8);
    memset(s->rb, 0, rbytes + 8);
    memset(s->tb, 0, tbytes + 8);
    if (!ko_pack(read, L, s->rb) || !ko_pack(text, tl, s->tb)) return 0;
    ko_pre_shift(s->tb, tbytes, s->ph);
    uint64_t shift, best, second;
    if (!ko_scan_best_with_second(s->rb, rbytes, L, 1, s->ph, tbytes, tl, &shift, &best, &second))
        return 0;
    if (shift + L > tl) return 0;
    const uint8_t *ref = text + shift;
    const uint32_t max_mism =
        ko_spectral_max_mismatches((uint32_t)L, (uint32_t)r->spectral_pct, (uint32_t)r->spectral_floor);
    const int has_second = second != UINT64_MAX;
    if (!ko_spectral_hit_is_certified(read, ref, L, L, best, has_second, has_second ? second : 0,
                                      max_mism, r->cert_strict))
        return 0;
    grow((void **)&s->md, &s->md_cap, 4 * L + 32);
    out->kind = 1;
    out->nm = ko_ungapped_tags(read, ref, L, s->md, s->md_cap);
    out->score = (int32_t)((L - best) * (uint64_t)c->match_score - best * (uint64_t)c->mismatch);
    out->ref_start = win_start + (uint32_t)shift;
    out->ref_end = win_start + (uint32_t)shift + (uint32_t)L;
    out->read_start = 0;
    out->read_end = (uint32_t)L;
    out->n_cigar = 1;
    return 1;
}

/* clipping_can_improve (mod.rs:442-479) */
static int clipping_can_improve(const uint8_t *read, size_t L, const uint8_t *text, size_t tl,
                                int32_t text_start, const uint32_t *cig, int32_t n,
                                const ko_cfg *c) {
    size_t qi = 0, ti = (size_t)text_start;
    int64_t score = 0, min_prefix = 0, best_prefix = 0;
    for (int32_t k = 0; k < n; k++) {
        const uint32_t len = cig[k] >> 4, op = cig[k] & 15u;
        if (op == KO_M) {
            for (uint32_t x = 0; x < len; x++) {
                const int same = qi < L && ti < tl && read[qi] == text[ti];
                score += same ? c->match_score : -(int64_t)c->mismatch;
                qi++;
                ti++;
                if (score < min_prefix) min_prefix = score;
                if (score > best_prefix) best_prefix = score;
            }
        } else if (op == KO_I || op == KO_D) {
            score -= (int64_t)c->gap_open + (int64_t)c->gap_extend * len;
            if (op == KO_I)
                qi += len;
            else
                ti += len;
            if (score < min_prefix) min_prefix = score;
        } else {
            return 1;
        }
    }
    return min_prefix < 0 || best_prefix > score + c->clip_penalty;
}

/* wfa_result_to_alignment (mod.rs:481-575) */
static int wfa_to_alignment(const uint8_t *read, size_t L, const uint8_t *text, size_t tl,
                            uint32_t win_start, const ko_wfa_aln *w, const ko_cfg *c, scratch *s,
                            ko_fast_out *out) {
    int64_t cref = 0, cqry = 0;
    for (int32_t k = 0; k < w->n_cigar; k++) {
        const uint32_t len = w->cigar[k] >> 4, op = w->cigar[k] & 15u;
        if (op == KO_M || op == KO_D || op == KO_N) cref += len;
        if (op == KO_M || op == KO_I || op == KO_S) cqry += len;
    }
    if (cqry != (int64_t)L || w->text_start + cref > (int64_t)tl || w->text_end > (int64_t)tl)
        return 0;
    int64_t score = 0;
    size_t qi = 0, ti = (size_t)w->text_start;
    for (int32_t k = 0; k < w->n_cigar; k++) {
        const uint32_t len = w->cigar[k] >> 4, op = w->cigar[k] & 15u;
        if (op == KO_M) {
            for (uint32_t x = 0; x < len; x++) {
                score += read[qi] == text[ti] ? c->match_score : -(int64_t)c->mismatch;
                qi++;
                ti++;
            }
        } else if (op == KO_I || op == KO_D) {
            score -= (int64_t)c->gap_open + (int64_t)c->gap_extend * len;
            if (op == KO_I)
                qi += len;
            else
                ti += len;
        }
    }
    grow((void **)&s->md, &s->md_cap, 4 * (L + tl) + 32);
    out->kind = 2;
    out->nm = ko_nm_md_from_cigar(read, L, text, tl, 0, (size_t)w->text_start, w->cigar,
                                  w->n_cigar, s->md, s->md_cap);
    out->score = (int32_t)score;
    out->ref_start = win_start + (uint32_t)w->text_start;
    out->ref_end = win_start + (uint32_t)w->text_end;
    out->read_start = 0;
    out->read_end = (uint32_t)L;
    out->n_cigar = w->n_cigar;
    return 1;
}

static size_t common_prefix_len(const uint8_t *a, size_t na, const uint8_t *b, size_t nb) {
    const size_t n = na < nb ? na : nb;
    size_t i = 0;
    while (i < n && a[i] == b[i]) i++;
    return i;
}

static int wfa_run(const uint8_t *read, size_t L, const uint8_t *t, size_t tn, const ko_cfg *c,
                   int32_t budget, int32_t drop, int32_t tbf, uint32_t **buf, size_t *cap,
                   ko_wfa_aln *w) {
    const size_t need = 2 * (L + tn) + 8;
    grow((void **)buf, cap, need * sizeof(uint32_t));
    w->cigar = *buf;
    w->cigar_cap = (int32_t)need;
    return ko_wfa_align_scored(read, (int32_t)L, t, (int32_t)tn, c, budget, drop, tbf, w) == 1;
}

static int fast_dp_one(const uint8_t *read, size_t L, const uint8_t *contig, size_t clen,
                       uint32_t ref_start, uint32_t read_start, int32_t chain_score,
                       const ko_cfg *c, const ko_router *r, scratch *s, ko_fast_out *out) {
    memset(out, 0, sizeof *out);
    /* parse_algo (router.rs:28-37): anything unrecognised is packed */
    const int algo = (r->algo >= 0 && r->algo <= 3) ? r->algo : 0;
    if (L == 0 || L > (size_t)r->wfa_max_read_len || algo == 3) return 0;
    if (r->quality_gate) { /* fast_path_worth_attempting (router.rs:169-189) */
        if (chain_score <= 0) return 0;
        uint64_t q = (uint64_t)chain_score * 100u;
        if (q > UINT32_MAX) q = UINT32_MAX;
        if (q / L < (uint64_t)r->quality_pct) return 0;
    }
    const int64_t win_start = (int64_t)ref_start - (int64_t)read_start;
    if (win_start < 0 || win_start >= (int64_t)clen) return 0;
    const int32_t band = imax32(c->bandwidth, 1);
    const int64_t text_len = imin64((int64_t)L + band, (int64_t)clen - win_start);
    if (text_len < (int64_t)L) return 0;
    const uint8_t *text = contig + win_start;
    if (algo == 0 || algo == 1) {
        if (spectral(read, L, text, (size_t)text_len, (uint32_t)win_start, c, r, s, out)) return 1;
    }
    size_t dist, pos;
    if (!ko_myers_bounded(read, L, text, (size_t)text_len,
                          L * (size_t)r->myers_pct / 100 + (size_t)r->myers_floor, &dist, &pos))
        return 0;
    /* wfa_score_budget (router.rs:160-166) with mismatch = A + B and gap
     * extend = A + E, as the cascade calls it */
    const int64_t mb = ((int64_t)L * r->wfa_budget_pct) / 100 * ((int64_t)c->match_score + c->mismatch);
    const int64_t ib = ((int64_t)c->gap_open + ((int64_t)c->match_score + c->gap_extend) * 4) * 3;
    const int32_t budget = (int32_t)imin64(imax64(mb, ib), INT32_MAX);
    const int64_t lead = imin64(imin64(imax32(r->wfa_lead, 0), win_start), band);
    const uint8_t *wtext;
    int64_t wstart, wlen;
    int32_t tbf;
    if (lead > 0) {
        wtext = contig + win_start - lead;
        wstart = win_start - lead;
        wlen = text_len + lead;
        tbf = (int32_t)lead;
    } else {
        wtext = text;
        wstart = win_start;
        wlen = text_len;
        tbf = r->wfa_ends_free;
    }
    const int32_t drop = r->wfa_adaptive > 0 ? r->wfa_adaptive : 0;
    ko_wfa_aln w;
    if (!wfa_run(read, L, wtext, (size_t)wlen, c, budget, drop, tbf, &s->cig, &s->cig_cap, &w))
        return 0;
    const uint8_t *ftext = wtext;
    int64_t fstart = wstart, flen = wlen;
    const int suspect = w.n_cigar > 0 &&
                        ((w.cigar[0] & 15u) != 0 || ((w.cigar[0] >> 4) < 19 && w.n_cigar > 1));
    if (suspect && lead == 0 && r->wfa_ends_free == 0) {
        const int64_t slack = imin64(imin64(band, 32), win_start);
        if (slack > 0) {
            int64_t origin = win_start;
            size_t prefix = 0;
            const int64_t hi = imin64(win_start + slack, (int64_t)clen - 1);
            for (int64_t st = win_start - slack; st <= hi; st++) {
                const size_t avail = (size_t)imin64((int64_t)L, (int64_t)clen - st);
                const size_t m = common_prefix_len(read, L, contig + st, avail);
                if (m > prefix) {
                    prefix = m;
                    origin = st;
                }
            }
            const uint8_t *exp;
            int64_t estart, elen;
            int32_t free_;
            if (prefix >= 4) {
                exp = contig + origin;
                estart = origin;
                elen = imin64(origin + (int64_t)L + band, (int64_t)clen) - origin;
                free_ = 0;
            } else {
                exp = contig + win_start - slack;
                estart = win_start - slack;
                elen = win_start + text_len - estart;
                free_ = (int32_t)(2 * slack);
            }
            ko_wfa_aln w2;
            if (wfa_run(read, L, exp, (size_t)elen, c, w.score, drop, free_, &s->cig2,
                        &s->cig2_cap, &w2) &&
                w2.score < w.score) {
                /* the retry's CIGAR becomes the result: swap the buffers */
                uint32_t *tmp = s->cig;
                size_t tc = s->cig_cap;
                s->cig = s->cig2;
                s->cig_cap = s->cig2_cap;
                s->cig2 = tmp;
                s->cig2_cap = tc;
                w = w2;
                ftext = exp;
                fstart = estart;
                flen = elen;
            }
        }
    }
    const int64_t perfect = (int64_t)c->match_score * (int64_t)L;
    if ((int64_t)w.score * 10 > perfect * 3 &&
        clipping_can_improve(read, L, ftext, (size_t)flen, w.text_start, w.cigar, w.n_cigar, c))
        return 0;
    return wfa_to_alignment(read, L, ftext, (size_t)flen, (uint32_t)fstart, &w, c, s, out);
}

int ko_fast_dp(const uint8_t *read, size_t read_len, const uint8_t *contig, size_t contig_len,
               uint32_t ref_start, uint32_t read_start, int32_t chain_score, const ko_cfg *cfg,
               const ko_router *router, ko_fast_out *out) {
    scratch s;
    memset(&s, 0, sizeof s);
    const int r = fast_dp_one(read, read_len, contig, contig_len, ref_start, read_start,
                              chain_score, cfg, router, &s, out);
    scratch_free(&s);
    return r;
}

struct ko_fast_scratch {
    scratch s;
};

ko_fast_scratch *ko_fast_scratch_new(void) {
    return (ko_fast_scratch *)calloc(1, sizeof(ko_fast_scratch));
}

void ko_fast_scratch_free(ko_fast_scratch *f) {
    if (!f) return;
    scratch_free(&f->s);
    free(f);
}

/* The WFA CIGAR is the retry's or the main pass's, both left in s.cig (the
 * retry swaps the buffers); a certificate's is {L}M. */
int ko_fast_dp_rec(const uint8_t *read, size_t read_len, const uint8_t *contig,
                   size_t contig_len, uint32_t ref_start, uint32_t read_start,
                   int32_t chain_score, const ko_cfg *cfg, const ko_router *router,
                   ko_fast_scratch *f, ko_fast_out *out, const uint32_t **cigar,
                   const char **md) {
    if (!fast_dp_one(read, read_len, contig, contig_len, ref_start, read_start, chain_score, cfg,
                     router, &f->s, out))
        return 0;
    if (out->kind == 1) {
        grow((void **)&f->s.cig, &f->s.cig_cap, 4 * sizeof(uint32_t));
        f->s.cig[0] = ((uint32_t)read_len << 4) | KO_M;
    }
    *cigar = f->s.cig;
    *md = f->s.md;
    return 1;
}

typedef struct {
    int32_t n;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *genome;
    const int64_t *contig_off, *contig_len;
    const int32_t *chain_contig;
    const uint32_t *chain_ref_start, *chain_read_start;
    const int32_t *chain_score;
    const ko_cfg *cfg;
    const ko_router *router;
    ko_fast_out *out;
    int32_t next; /* next chunk, taken atomically */
} fast_job;

static void *fast_worker(void *arg) {
    fast_job *J = (fast_job *)arg;
    scratch s;
    memset(&s, 0, sizeof s);
    for (;;) {
        const int32_t lo = __atomic_fetch_add(&J->next, 64, __ATOMIC_RELAXED);
        if (lo >= J->n) break;
        const int32_t hi = imin32(lo + 64, J->n);
        for (int32_t k = lo; k < hi; k++) {
            const int32_t ci = J->chain_contig[k];
            fast_dp_one(J->reads + J->read_off[k], (size_t)J->read_len[k],
                        J->genome + J->contig_off[ci], (size_t)J->contig_len[ci],
                        J->chain_ref_start[k], J->chain_read_start[k], J->chain_score[k],
                        J->cfg, J->router, &s, &J->out[k]);
        }
    }
    scratch_free(&s);
    return NULL;
}

void ko_fast_dp_batch_mt(int32_t n, const uint8_t *reads, const int64_t *read_off,
                         const int32_t *read_len, const uint8_t *genome,
                         const int64_t *contig_off, const int64_t *contig_len,
                         const int32_t *chain_contig, const uint32_t *chain_ref_start,
                         const uint32_t *chain_read_start, const int32_t *chain_score,
                         const ko_cfg *cfg, const ko_router *router, ko_fast_out *out,
                         int threads) {
    if (threads < 1) threads = 1;
    fast_job J = {n,          reads,           read_off,         read_len,    genome,
                  contig_off, contig_len,      chain_contig,     chain_ref_start,
                  chain_read_start, chain_score, cfg, router, out, 0};
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int k = 0; k < threads; k++) pthread_create(&th[k], NULL, fast_worker, &J);
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    free(th);
}
