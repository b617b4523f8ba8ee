/*
 * kira_wfa_oracle.c — CPU restatement of the WFA / Myers fast path
 * (src/alignment/wfa.rs, myers.rs, router.rs and try_fast_dp_alignment_inner,
 * mod.rs:282-580).  Test infrastructure only: the product never links it.
 *
 * Pinned (tests/test_oracle_wfa.py) by the reference's own test oracles:
 * Myers against the naive semi-global edit distance of
 * tests/myers_bounded_edit.rs:9-31 and its known answers, the scored WFA
 * against the independent affine DP and CIGAR replay of tests/scored_wfa.rs.
 */
#include <stdlib.h>
#include <string.h>

#include "kira_oracle.h"

/* ------------------------------------------------------------------ Myers */

/* base_to_idx — myers.rs:8-16 */
static int base_idx(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
    }
}

/* bounded_edit_distance — myers.rs:47-70, 206-367 (the fixed-array path for
 * patterns of <= 256 bases runs the same recurrence, :101-204) */
int ko_myers_bounded(const uint8_t *pattern, size_t m, const uint8_t *text, size_t n,
                     size_t max_k, size_t *dist, size_t *pos) {
    if (m == 0) {
        *dist = 0;
        *pos = 0;
        return 1;
    }
    if (n == 0) {
        if (m <= max_k) {
            *dist = m;
            *pos = 0;
            return 1;
        }
        return 0;
    }
    const size_t blocks = (m + 63) / 64;
    uint64_t *peq = calloc(4 * blocks, 8), *vp = calloc(blocks, 8), *vn = calloc(blocks, 8);
    uint64_t *zero = calloc(blocks, 8);
    for (size_t i = 0; i < m; i++) {
        int c = base_idx(pattern[i]);
        if (c >= 0) peq[c * blocks + i / 64] |= 1ull << (i % 64);
    }
    const size_t top_used = m - (blocks - 1) * 64;
    const uint64_t top_mask = top_used == 64 ? ~0ull : ((1ull << top_used) - 1);
    const uint64_t high_bit = 1ull << (top_used - 1);
    for (size_t b = 0; b < blocks; b++) vp[b] = ~0ull;
    vp[blocks - 1] &= top_mask;
    size_t err = m, best = m, best_pos = 0;
    int ok = -1;
    for (size_t j = 0; j < n && ok < 0; j++) {
        int c = base_idx(text[j]);
        const uint64_t *eqv = c >= 0 ? peq + c * blocks : zero;
        uint64_t hin_p = 0, hin_n = 0;
        unsigned carry = 0;
        for (size_t b = 0; b < blocks; b++) {
            const uint64_t eq = eqv[b], vpb = vp[b], vnb = vn[b];
            const uint64_t x = eq | vnb;
            const uint64_t av = eq & vpb;
            const uint64_t s1 = av + vpb;
            const unsigned c1 = s1 < av;
            const uint64_t sum = s1 + carry;
            const unsigned c2 = sum < s1;
            carry = c1 | c2;
            const uint64_t d0 = (sum ^ vpb) | x;
            const uint64_t hp = vnb | ~(d0 | vpb);
            const uint64_t hn = d0 & vpb;
            const uint64_t hps = (hp << 1) | hin_p, hns = (hn << 1) | hin_n;
            hin_p = hp >> 63;
            hin_n = hn >> 63;
            const uint64_t mask = b == blocks - 1 ? top_mask : ~0ull;
            vp[b] = (hns | ~(d0 | hps)) & mask;
            vn[b] = (d0 & hps) & mask;
            if (b == blocks - 1) {
                if (hp & high_bit) err++;
                if (hn & high_bit) err--;
            }
        }
        if (err <= best) {
            best = err;
            best_pos = j + 1;
        }
        /* exact early reject */
        size_t rest = n - 1 - j;
        if (best > max_k && (err > rest ? err - rest : 0) > max_k) ok = 0;
    }
    free(peq);
    free(vp);
    free(vn);
    free(zero);
    if (ok == 0 || best > max_k) return 0;
    *dist = best;
    *pos = best_pos;
    return 1;
}

/* ------------------------------------------------------------------ WFA */

#define NONE_OFF (INT32_MIN / 4)

typedef struct {
    int32_t lo, hi; /* empty when hi < lo */
    int32_t *off;
} wf_t;

typedef struct {
    int32_t mismatch, gap_open, ins_extend, del_extend;
} wfa_costs;

static int32_t wf_get(const wf_t *w, int32_t k) {
    if (w->hi < w->lo || k < w->lo || k > w->hi) return NONE_OFF;
    return w->off[k - w->lo];
}

static wf_t wf_new(int32_t lo, int32_t hi) {
    wf_t w = {lo, hi, NULL};
    if (hi >= lo) {
        w.off = malloc(sizeof(int32_t) * (size_t)(hi - lo + 1));
        for (int32_t k = 0; k <= hi - lo; k++) w.off[k] = NONE_OFF;
    }
    return w;
}

typedef struct {
    wf_t *m, *i, *d;
    int32_t len; /* layers 0 .. len - 1 exist */
} hist_t;

/* get_src — wfa.rs:360-372 */
static const wf_t *get_src(const wf_t *h, int32_t len, int32_t s) {
    if (s < 0 || s >= len) return NULL;
    if (h[s].hi < h[s].lo) return NULL;
    return &h[s];
}

/* extend_diagonal_scalar — wfa.rs:207-224 */
static int32_t extend_diag(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n, int32_t k,
                           int32_t off) {
    for (;;) {
        int32_t i = off, j = off + k;
        if (i >= m || j < 0 || j >= n) break;
        if (p[i] != t[j]) break;
        off++;
    }
    return off;
}

static void extend_wf(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n, wf_t *w) {
    for (int32_t k = w->lo; k <= w->hi; k++) {
        int32_t o = w->off[k - w->lo];
        if (o == NONE_OFF) continue;
        w->off[k - w->lo] = extend_diag(p, m, t, n, k, o);
    }
}

/* prune_wavefront — wfa.rs:293-346 */
static void prune_wf(wf_t *w, int32_t drop) {
    if (w->hi < w->lo) return;
    int64_t max_ad = INT64_MIN;
    for (int32_t k = w->lo; k <= w->hi; k++) {
        int32_t o = w->off[k - w->lo];
        if (o != NONE_OFF && 2 * (int64_t)o + k > max_ad) max_ad = 2 * (int64_t)o + k;
    }
    if (max_ad == INT64_MIN) return;
    const int64_t th = max_ad - drop;
    int32_t nlo = w->lo;
    while (nlo <= w->hi) {
        int32_t o = w->off[nlo - w->lo];
        if (o != NONE_OFF && 2 * (int64_t)o + nlo >= th) break;
        nlo++;
    }
    int32_t nhi = w->hi;
    while (nhi >= nlo) {
        int32_t o = w->off[nhi - w->lo];
        if (o != NONE_OFF && 2 * (int64_t)o + nhi >= th) break;
        nhi--;
    }
    if (nlo > nhi) {
        free(w->off);
        w->off = NULL;
        w->lo = 0;
        w->hi = -1;
        return;
    }
    for (int32_t k = nlo; k <= nhi; k++) {
        int32_t o = w->off[k - w->lo];
        if (o != NONE_OFF && 2 * (int64_t)o + k < th) w->off[k - w->lo] = NONE_OFF;
    }
    if (nlo != w->lo || nhi != w->hi) {
        memmove(w->off, w->off + (nlo - w->lo), sizeof(int32_t) * (size_t)(nhi - nlo + 1));
        w->lo = nlo;
        w->hi = nhi;
    }
}

/* check_done — wfa.rs:348-357: first diagonal (from lo) whose offset
 * consumed the pattern */
static int check_done(const wf_t *w, int32_t m, int32_t *k_end) {
    for (int32_t k = w->lo; k <= w->hi; k++)
        if (w->off[k - w->lo] >= m) {
            *k_end = k;
            return 1;
        }
    return 0;
}

/* step — wfa.rs:374-525 */
static void wfa_step(int32_t s, wfa_costs pen, const hist_t *H, int32_t m, int32_t n, wf_t *nm,
                     wf_t *ni, wf_t *nd) {
    const wf_t *mx = get_src(H->m, H->len, s - pen.mismatch);
    const wf_t *moi = get_src(H->m, H->len, s - pen.gap_open - pen.ins_extend);
    const wf_t *mod_ = get_src(H->m, H->len, s - pen.gap_open - pen.del_extend);
    const wf_t *ie = get_src(H->i, H->len, s - pen.ins_extend);
    const wf_t *de = get_src(H->d, H->len, s - pen.del_extend);
    int32_t ilo = INT32_MAX, ihi = INT32_MIN, dlo = INT32_MAX, dhi = INT32_MIN;
    int32_t mlo = INT32_MAX, mhi = INT32_MIN;
    if (moi) { if (moi->lo - 1 < ilo) ilo = moi->lo - 1; if (moi->hi - 1 > ihi) ihi = moi->hi - 1; }
    if (ie) { if (ie->lo - 1 < ilo) ilo = ie->lo - 1; if (ie->hi - 1 > ihi) ihi = ie->hi - 1; }
    if (mod_) { if (mod_->lo + 1 < dlo) dlo = mod_->lo + 1; if (mod_->hi + 1 > dhi) dhi = mod_->hi + 1; }
    if (de) { if (de->lo + 1 < dlo) dlo = de->lo + 1; if (de->hi + 1 > dhi) dhi = de->hi + 1; }
    if (mx) { if (mx->lo < mlo) mlo = mx->lo; if (mx->hi > mhi) mhi = mx->hi; }
    if (ilo <= ihi) { if (ilo < mlo) mlo = ilo; if (ihi > mhi) mhi = ihi; }
    if (dlo <= dhi) { if (dlo < mlo) mlo = dlo; if (dhi > mhi) mhi = dhi; }
    *ni = ilo <= ihi ? wf_new(ilo, ihi) : wf_new(0, -1);
    for (int32_t k = ni->lo; k <= ni->hi; k++) {
        int32_t best = NONE_OFF, v;
        if (moi && (v = wf_get(moi, k + 1)) != NONE_OFF && v < m && v + 1 > best) best = v + 1;
        if (ie && (v = wf_get(ie, k + 1)) != NONE_OFF && v < m && v + 1 > best) best = v + 1;
        ni->off[k - ni->lo] = best;
    }
    *nd = dlo <= dhi ? wf_new(dlo, dhi) : wf_new(0, -1);
    for (int32_t k = nd->lo; k <= nd->hi; k++) {
        int32_t best = NONE_OFF, v;
        if (mod_ && (v = wf_get(mod_, k - 1)) != NONE_OFF && v + k <= n && v > best) best = v;
        if (de && (v = wf_get(de, k - 1)) != NONE_OFF && v + k <= n && v > best) best = v;
        nd->off[k - nd->lo] = best;
    }
    *nm = mlo <= mhi ? wf_new(mlo, mhi) : wf_new(0, -1);
    for (int32_t k = nm->lo; k <= nm->hi; k++) {
        int32_t best = NONE_OFF, v;
        if (mx && (v = wf_get(mx, k)) != NONE_OFF && v < m && v + k < n && v + 1 > best)
            best = v + 1;
        if ((v = wf_get(ni, k)) != NONE_OFF && v > best) best = v;
        if ((v = wf_get(nd, k)) != NONE_OFF && v > best) best = v;
        nm->off[k - nm->lo] = best;
    }
}

static void push_op(uint32_t **ops, int32_t *n, int32_t *cap, uint32_t op, uint32_t len) {
    if (*n > 0 && ((*ops)[*n - 1] & 15u) == op) {
        (*ops)[*n - 1] += len << 4;
        return;
    }
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 16;
        *ops = realloc(*ops, sizeof(uint32_t) * (size_t)*cap);
    }
    (*ops)[(*n)++] = (len << 4) | op;
}

static int32_t layer_get(const hist_t *H, int which, int32_t s, int32_t k) {
    if (s < 0 || s >= H->len) return NONE_OFF;
    const wf_t *w = which == 0 ? &H->m[s] : (which == 1 ? &H->i[s] : &H->d[s]);
    return wf_get(w, k);
}

/* build_cigar — wfa.rs:762-952; returns the start diagonal */
static int32_t build_cigar(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n,
                           const hist_t *H, int32_t s, int32_t k, wfa_costs pen, uint32_t **ops,
                           int32_t *nops) {
    int32_t cap = 0, cnt = 0;
    uint32_t *o = NULL;
    int32_t off = m;
    int layer = 0; /* 0 M, 1 I, 2 D */
    while (!(s == 0 && layer == 0 && off == 0)) {
        if (layer == 0) {
            int32_t walk = off;
            for (;;) {
                if (walk == 0) break;
                if (layer_get(H, 1, s, k) == walk) break;
                if (layer_get(H, 2, s, k) == walk) break;
                if (s >= pen.mismatch) {
                    int32_t pv = layer_get(H, 0, s - pen.mismatch, k);
                    if (pv != NONE_OFF && pv + 1 == walk) break;
                }
                int32_t i = walk - 1, j = walk - 1 + k;
                if (i < 0 || j < 0) break;
                if (i >= m || j >= n) break;
                if (p[i] != t[j]) break;
                walk--;
            }
            if (off - walk > 0) push_op(&o, &cnt, &cap, KO_M, (uint32_t)(off - walk));
            off = walk;
            if (s == 0 && off == 0) break;
            int took = 0;
            if (s >= pen.mismatch) {
                int32_t pv = layer_get(H, 0, s - pen.mismatch, k);
                if (pv != NONE_OFF && pv + 1 == off) {
                    push_op(&o, &cnt, &cap, KO_M, 1);
                    off--;
                    s -= pen.mismatch;
                    took = 1;
                }
            }
            if (!took && layer_get(H, 1, s, k) == off) {
                layer = 1;
                took = 1;
            }
            if (!took && layer_get(H, 2, s, k) == off) {
                layer = 2;
                took = 1;
            }
            if (!took) break;
        } else if (layer == 1) {
            int took = 0;
            if (s >= pen.gap_open + pen.ins_extend) {
                int32_t pv = layer_get(H, 0, s - pen.gap_open - pen.ins_extend, k + 1);
                if (pv != NONE_OFF && pv + 1 == off) {
                    push_op(&o, &cnt, &cap, KO_I, 1);
                    off--;
                    k++;
                    s -= pen.gap_open + pen.ins_extend;
                    layer = 0;
                    took = 1;
                }
            }
            if (!took && s >= pen.ins_extend) {
                int32_t pv = layer_get(H, 1, s - pen.ins_extend, k + 1);
                if (pv != NONE_OFF && pv + 1 == off) {
                    push_op(&o, &cnt, &cap, KO_I, 1);
                    off--;
                    k++;
                    s -= pen.ins_extend;
                    took = 1;
                }
            }
            if (!took) break;
        } else {
            int took = 0;
            if (s >= pen.gap_open + pen.del_extend) {
                int32_t pv = layer_get(H, 0, s - pen.gap_open - pen.del_extend, k - 1);
                if (pv != NONE_OFF && pv == off) {
                    push_op(&o, &cnt, &cap, KO_D, 1);
                    k--;
                    s -= pen.gap_open + pen.del_extend;
                    layer = 0;
                    took = 1;
                }
            }
            if (!took && s >= pen.del_extend) {
                int32_t pv = layer_get(H, 2, s - pen.del_extend, k - 1);
                if (pv != NONE_OFF && pv == off) {
                    push_op(&o, &cnt, &cap, KO_D, 1);
                    k--;
                    s -= pen.del_extend;
                    took = 1;
                }
            }
            if (!took) break;
        }
    }
    for (int32_t a = 0, b = cnt - 1; a < b; a++, b--) {
        uint32_t x = o[a];
        o[a] = o[b];
        o[b] = x;
    }
    *ops = o;
    *nops = cnt;
    return k;
}

static void hist_free(hist_t *H) {
    for (int32_t s = 0; s < H->len; s++) {
        free(H->m[s].off);
        free(H->i[s].off);
        free(H->d[s].off);
    }
    free(H->m);
    free(H->i);
    free(H->d);
}

/* WfaCosts::scored — wfa.rs:41-55 */
static int scored_costs(const ko_cfg *c, wfa_costs *pen) {
    if (c->match_score <= 0 || c->mismatch <= 0 || c->gap_open < 0 || c->gap_extend <= 0) return 0;
    int64_t x = (int64_t)c->match_score + c->mismatch, ie = (int64_t)c->match_score + c->gap_extend;
    if (x > INT32_MAX || ie > INT32_MAX) return 0;
    pen->mismatch = (int32_t)x;
    pen->gap_open = c->gap_open;
    pen->ins_extend = (int32_t)ie;
    pen->del_extend = c->gap_extend;
    return 1;
}

/* wfa_align_with_costs / wfa_score_with_costs — wfa.rs:549-648, 678-760.
 * want_cigar = 0 skips the traceback (the score is the same). */
static int wfa_run(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n, wfa_costs pen,
                   int32_t max_score, int32_t drop, int32_t tbf_in, int want_cigar,
                   ko_wfa_aln *out) {
    out->n_cigar = 0;
    if (m == 0) {
        out->score = 0;
        out->text_start = 0;
        out->text_end = 0;
        return 1;
    }
    int32_t tbf = tbf_in > 0 ? tbf_in : 0;
    int32_t lim = n - 1 > 0 ? n - 1 : 0;
    if (tbf > lim) tbf = lim;
    hist_t H;
    int32_t cap = (max_score > 0 ? max_score : 0) + 1;
    H.m = calloc((size_t)cap, sizeof(wf_t));
    H.i = calloc((size_t)cap, sizeof(wf_t));
    H.d = calloc((size_t)cap, sizeof(wf_t));
    for (int32_t s = 0; s < cap; s++) {
        H.m[s].hi = H.i[s].hi = H.d[s].hi = -1;
    }
    H.len = 0;
    wf_t m0 = wf_new(0, tbf);
    for (int32_t k = 0; k <= tbf; k++) m0.off[k] = 0;
    extend_wf(p, m, t, n, &m0);
    if (drop > 0) prune_wf(&m0, drop);
    H.m[0] = m0;
    H.len = 1;
    int32_t kend, s_done = -1;
    if (check_done(&H.m[0], m, &kend)) s_done = 0;
    for (int32_t s = 1; s_done < 0 && s <= max_score; s++) {
        wf_t nm, ni, nd;
        wfa_step(s, pen, &H, m, n, &nm, &ni, &nd);
        extend_wf(p, m, t, n, &nm);
        if (drop > 0) prune_wf(&nm, drop);
        H.m[s] = nm;
        H.i[s] = ni;
        H.d[s] = nd;
        H.len = s + 1;
        if (check_done(&H.m[s], m, &kend)) s_done = s;
    }
    int ok = s_done >= 0;
    if (ok) {
        out->score = s_done;
        out->text_end = m + kend;
        if (want_cigar) {
            uint32_t *ops = NULL;
            int32_t nops = 0;
            int32_t k0 = build_cigar(p, m, t, n, &H, s_done, kend, pen, &ops, &nops);
            out->text_start = k0 > 0 ? k0 : 0;
            if (nops > out->cigar_cap) {
                free(ops);
                hist_free(&H);
                return -1;
            }
            if (nops) memcpy(out->cigar, ops, sizeof(uint32_t) * (size_t)nops);
            out->n_cigar = nops;
            free(ops);
        }
    }
    hist_free(&H);
    return ok;
}

/* wfa_align_scored — wfa.rs:58-67: the loss (score) of the match-rewarding
 * objective, CIGAR, text_start / text_end; -1 if the CIGAR buffer is short */
int ko_wfa_align_scored(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n, const ko_cfg *cfg,
                        int32_t max_score, int32_t adaptive_drop, int32_t text_begin_free,
                        ko_wfa_aln *out) {
    wfa_costs pen;
    if (!scored_costs(cfg, &pen)) return 0;
    return wfa_run(p, m, t, n, pen, max_score, adaptive_drop, text_begin_free, 1, out);
}

/* wfa_score_only_scored — wfa.rs:69-81: the affine score m * A - loss */
int ko_wfa_score_only_scored(const uint8_t *p, int32_t m, const uint8_t *t, int32_t n,
                             const ko_cfg *cfg, int32_t max_score, int32_t adaptive_drop,
                             int32_t text_begin_free, int32_t *score) {
    wfa_costs pen;
    if (!scored_costs(cfg, &pen)) return 0;
    ko_wfa_aln a;
    memset(&a, 0, sizeof a);
    if (wfa_run(p, m, t, n, pen, max_score, adaptive_drop, text_begin_free, 0, &a) != 1) return 0;
    int64_t v = (int64_t)m * cfg->match_score;
    if (v > INT32_MAX) return 0;
    v -= a.score;
    if (v < INT32_MIN) return 0;
    *score = (int32_t)v;
    return 1;
}
