/*
 * kira_oracle.c — TEST INFRASTRUCTURE.  A plain-C restatement of the
 * reference CPU path (kira-ls-aligner v0.8.3, Rust) for the sieve + SW hot
 * path.  Only the tests, the smoke check and bench.py's cpu_baseline leg use
 * it, always as the checker or the timed CPU baseline, never as the product.
 *
 * Parity pins (checked in tests/test_oracle_*.py):
 *   pack / pre-shift: reference known answers (tests/bitpacked_scan.rs:44-59,
 *          tests/unit/alignment_bitpacked.rs:4-9,53-57) + closed form.
 *   mismatch count / scan: brute-force Hamming over every shift (tests/bitpacked_scan.rs:13-33),
 *          global-best + tied-second known answers (tests/unit/alignment_bitpacked.rs:59-76).
 *   batch backend: the reference's GPU corpus (tests/cuda_spectral.rs:13-137) vs brute force.
 *   certification: closed-form rules + hand cases; MD/NM known answers (md.rs:214-285).
 *   prefilter: tests/prefilter_accept.rs (5 gate outcomes), tests/ungapped_stats.rs.
 *   lane-group SW: tests/alignment.rs known answers; full-matrix brute force with the
 *          same tie rules on tiny inputs; CIGAR replay = score.
 *   banded SW: src/alignment/mod.rs:2368-2635 known answers; replay invariant.
 */
#include "kira_oracle.h"

#include <limits.h>
#ifdef __AVX2__
#include <immintrin.h>
#endif
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define I32_MIN_4 (INT32_MIN / 4)
#define I32_MIN_8 (INT32_MIN / 8)

/* ======================================================================
 * encode_base / pack_into : src/alignment/bitpacked.rs:12-20, 57-91
 * ====================================================================== */
static inline uint8_t encode_base(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}
static inline int is_acgt(uint8_t b) {
    return b == 'A' || b == 'a' || b == 'C' || b == 'c' || b == 'G' || b == 'g' ||
           b == 'T' || b == 't';
}

int ko_pack(const uint8_t *seq, size_t n, uint8_t *bits) {
    size_t bytes = (n + 3) / 4;
    int valid = 1;
    memset(bits, 0, bytes);
    for (size_t i = 0; i < n; i++) {
        valid &= is_acgt(seq[i]);
        bits[i / 4] |= (uint8_t)(encode_base(seq[i]) << (2 * (i % 4)));
    }
    return valid;
}

/* shift_bits_right_into / pre_shift_into : bitpacked.rs:95-119 */
void ko_shift_bits_right(const uint8_t *bits, size_t n, unsigned bit_offset, uint8_t *out) {
    if (bit_offset == 0) {
        memcpy(out, bits, n);
        return;
    }
    for (size_t i = 0; i < n; i++) {
        uint8_t lo = (uint8_t)(bits[i] >> bit_offset);
        uint8_t hi = (i + 1 < n) ? (uint8_t)(bits[i + 1] << (8 - bit_offset)) : 0;
        out[i] = lo | hi;
    }
}

void ko_pre_shift(const uint8_t *bits, size_t n, uint8_t *out4) {
    for (unsigned phase = 0; phase < 4; phase++)
        ko_shift_bits_right(bits, n, phase * 2, out4 + phase * n);
}

/* mismatch_count_u64 : bitpacked.rs:238-270.  (The AVX2 and scalar
 * variants at :272-354 return the same count; only the <=64-byte u64 form is
 * ever reached by the reads this path handles.) */
static inline uint32_t popc64(uint64_t x) { return (uint32_t)__builtin_popcountll(x); }
static inline uint32_t popc8(uint8_t x) { return (uint32_t)__builtin_popcount(x); }

uint32_t ko_mismatch_count(const uint8_t *read, const uint8_t *ref, size_t n_bytes,
                           size_t n_nucs) {
    const uint64_t PAIR_LOW = 0x5555555555555555ULL;
    size_t full_bytes = n_nucs / 4;
    size_t tail_pairs = n_nucs % 4;
    uint32_t total = 0;
    size_t i = 0;
    while (i + 8 <= full_bytes) {
        uint64_t a, b;
        memcpy(&a, read + i, 8);
        memcpy(&b, ref + i, 8);
        uint64_t x = a ^ b;
        total += popc64((x | (x >> 1)) & PAIR_LOW);
        i += 8;
    }
    while (i < full_bytes) {
        uint8_t x = read[i] ^ ref[i];
        total += popc8((uint8_t)((x | (x >> 1)) & 0x55));
        i++;
    }
    if (tail_pairs != 0 && full_bytes < n_bytes) {
        uint8_t x = read[full_bytes] ^ ref[full_bytes];
        uint8_t pm = (uint8_t)((x | (x >> 1)) & 0x55);
        uint8_t keep = (uint8_t)(((1u << (tail_pairs * 2)) - 1) & 0x55);
        total += popc8(pm & keep);
    }
    return total;
}

/* scan_best_with_second_raw : bitpacked.rs:154-199 */
int ko_scan_best_with_second(const uint8_t *read_bits, size_t read_bytes, size_t read_len,
                             int read_valid, const uint8_t *phases, size_t phase_len,
                             size_t ref_len, uint64_t *shift, uint64_t *best,
                             uint64_t *second) {
    size_t r = read_len;
    if (r == 0 || ref_len < r || !read_valid) return 0;
    size_t n_shifts = ref_len - r + 1;
    uint64_t best_shift = 0, best_m = UINT64_MAX, second_m = UINT64_MAX;
    for (size_t t = 0; t < n_shifts; t++) {
        size_t phase = t % 4, byte_off = t / 4;
        if (byte_off + read_bytes > phase_len) continue;
        uint64_t m = ko_mismatch_count(read_bits, phases + phase * phase_len + byte_off,
                                       read_bytes, r);
        if (m < best_m) {
            second_m = best_m;
            best_m = m;
            best_shift = t;
        } else if (m < second_m) {
            second_m = m;
        }
    }
    if (best_m == UINT64_MAX) return 0;
    *shift = best_shift;
    *best = best_m;
    *second = second_m; /* UINT64_MAX == None */
    return 1;
}

/* CudaBackend::run_batch (src/cuda/backend.rs:96-249) with the kernel's
 * result conventions (src/cuda/spectral.cu:149-212): limit =
 * min(max_mismatches, read_nucs); read_len == 0 or ref_len < read_len gives
 * (-1, INT32_MAX, INT32_MAX); a rejected best keeps its second. */
static void backend_one(int32_t i, int32_t n, const uint8_t *reads_packed,
                        int32_t read_bytes_max, const int32_t *read_lens,
                        const uint8_t *ref_flat, int64_t ref_bytes_per_phase,
                        const int64_t *ref_offsets, const int64_t *ref_phase_bytes,
                        const int32_t *ref_lens, const uint32_t *max_mismatches,
                        int32_t *out_shifts, int32_t *out_mism, uint8_t *scratch) {
    int32_t rl = read_lens[i], wl = ref_lens[i];
    if (rl == 0 || wl < rl) {
        out_shifts[i] = -1;
        out_mism[i] = INT32_MAX;
        out_mism[n + i] = INT32_MAX;
        return;
    }
    uint32_t lim = max_mismatches[i];
    if (lim > (uint32_t)rl) lim = (uint32_t)rl;
    if (lim > (uint32_t)INT32_MAX) lim = INT32_MAX;
    size_t pb = (size_t)ref_phase_bytes[i];
    for (int p = 0; p < 4; p++)
        memcpy(scratch + p * pb, ref_flat + p * ref_bytes_per_phase + ref_offsets[i], pb);
    uint64_t sh, be, se;
    size_t rb = ((size_t)rl + 3) / 4;
    int ok = ko_scan_best_with_second(reads_packed + (size_t)i * read_bytes_max, rb, (size_t)rl,
                                      1, scratch, pb, (size_t)wl, &sh, &be, &se);
    if (!ok) {
        out_shifts[i] = -1;
        out_mism[i] = INT32_MAX;
        out_mism[n + i] = INT32_MAX;
        return;
    }
    int acc = be <= lim;
    out_shifts[i] = acc ? (int32_t)sh : -1;
    out_mism[i] = acc ? (int32_t)be : INT32_MAX;
    out_mism[n + i] = se == UINT64_MAX ? INT32_MAX : (int32_t)se;
}

void ko_backend_run_batch(int32_t n, const uint8_t *reads_packed, int32_t read_bytes_max,
                          const int32_t *read_lens, const uint8_t *ref_flat,
                          int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                          const int64_t *ref_phase_bytes, const int32_t *ref_lens,
                          const uint32_t *max_mismatches, int32_t *out_shifts,
                          int32_t *out_mism) {
    size_t mx = 0;
    for (int32_t i = 0; i < n; i++)
        if ((size_t)ref_phase_bytes[i] > mx) mx = (size_t)ref_phase_bytes[i];
    uint8_t *scratch = (uint8_t *)malloc(4 * mx + 16);
    for (int32_t i = 0; i < n; i++)
        backend_one(i, n, reads_packed, read_bytes_max, read_lens, ref_flat,
                    ref_bytes_per_phase, ref_offsets, ref_phase_bytes, ref_lens,
                    max_mismatches, out_shifts, out_mism, scratch);
    free(scratch);
}

typedef struct {
    int32_t lo, hi, n;
    const uint8_t *reads_packed;
    int32_t read_bytes_max;
    const int32_t *read_lens;
    const uint8_t *ref_flat;
    int64_t ref_bytes_per_phase;
    const int64_t *ref_offsets, *ref_phase_bytes;
    const int32_t *ref_lens;
    const uint32_t *max_mismatches;
    int32_t *out_shifts, *out_mism;
} backend_task;

static void *backend_worker(void *arg) {
    backend_task *t = (backend_task *)arg;
    size_t mx = 0;
    for (int32_t i = t->lo; i < t->hi; i++)
        if ((size_t)t->ref_phase_bytes[i] > mx) mx = (size_t)t->ref_phase_bytes[i];
    uint8_t *scratch = (uint8_t *)malloc(4 * mx + 16);
    for (int32_t i = t->lo; i < t->hi; i++)
        backend_one(i, t->n, t->reads_packed, t->read_bytes_max, t->read_lens, t->ref_flat,
                    t->ref_bytes_per_phase, t->ref_offsets, t->ref_phase_bytes, t->ref_lens,
                    t->max_mismatches, t->out_shifts, t->out_mism, scratch);
    free(scratch);
    return NULL;
}

void ko_backend_run_batch_mt(int32_t n, const uint8_t *reads_packed, int32_t read_bytes_max,
                             const int32_t *read_lens, const uint8_t *ref_flat,
                             int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                             const int64_t *ref_phase_bytes, const int32_t *ref_lens,
                             const uint32_t *max_mismatches, int32_t *out_shifts,
                             int32_t *out_mism, int threads) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    backend_task *tk = (backend_task *)calloc((size_t)threads, sizeof(backend_task));
    for (int k = 0; k < threads; k++) {
        backend_task *t = &tk[k];
        t->lo = (int32_t)((int64_t)n * k / threads);
        t->hi = (int32_t)((int64_t)n * (k + 1) / threads);
        t->n = n;
        t->reads_packed = reads_packed;
        t->read_bytes_max = read_bytes_max;
        t->read_lens = read_lens;
        t->ref_flat = ref_flat;
        t->ref_bytes_per_phase = ref_bytes_per_phase;
        t->ref_offsets = ref_offsets;
        t->ref_phase_bytes = ref_phase_bytes;
        t->ref_lens = ref_lens;
        t->max_mismatches = max_mismatches;
        t->out_shifts = out_shifts;
        t->out_mism = out_mism;
        pthread_create(&th[k], NULL, backend_worker, t);
    }
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    free(th);
    free(tk);
}

/* router::spectral_max_mismatches : src/alignment/router.rs:191-195 */
uint32_t ko_spectral_max_mismatches(uint32_t read_len, uint32_t pct, uint32_t floor_) {
    uint32_t v = (uint32_t)(((uint64_t)read_len * pct) / 100);
    return v > floor_ ? v : floor_;
}

/* spectral_hit_is_certified + bounded_hamming : src/alignment/mod.rs:688-776 */
static size_t bounded_hamming(const uint8_t *a, size_t alen, const uint8_t *b, size_t blen,
                              size_t limit) {
    if (alen != blen) return limit + 1;
    size_t m = 0;
    for (size_t k = 0; k < alen; k++) {
        m += a[k] != b[k];
        if (m > limit) break;
    }
    return m;
}

int ko_spectral_hit_is_certified(const uint8_t *read, const uint8_t *ref, size_t len,
                                 size_t ref_len, uint64_t hit_mism, int has_second,
                                 uint64_t second, uint64_t max_mism, int strict) {
    if (len != ref_len || hit_mism > max_mism ||
        (has_second && second <= (hit_mism == UINT64_MAX ? UINT64_MAX : hit_mism + 1)))
        return 0;
    if (hit_mism == 0) return 1;
    if (hit_mism > 2) return 0;
    size_t pos[2] = {SIZE_MAX, SIZE_MAX};
    size_t n = 0;
    for (size_t i = 0; i < len; i++) {
        if (read[i] != ref[i]) {
            if (n < 2) pos[n] = i;
            n++;
        }
    }
    if (n != hit_mism) return 0;
    size_t edge = len < 8 ? len : 8;
    for (size_t k = 0; k < n; k++)
        if (pos[k] < edge || pos[k] + edge >= len) return 0;
    if (n == 2 && (pos[1] > pos[0] ? pos[1] - pos[0] : 0) <= 4) return 0;
    size_t p = pos[0];
    size_t suffix_len = len > p + 1 ? len - (p + 1) : 0;
    size_t min_suffix = strict ? 4 : 8;
    if (suffix_len >= min_suffix) {
        int read_ins = bounded_hamming(read + p + 1, len - p - 1, ref + p, (ref_len - 1) - p, 1) <= 1;
        int ref_ins = bounded_hamming(read + p, (len - 1) - p, ref + p + 1, ref_len - p - 1, 1) <= 1;
        if (read_ins || ref_ins) return 0;
        if (strict && suffix_len >= 6) {
            int r2 = bounded_hamming(read + p + 2, len - p - 2, ref + p, (ref_len - 2) - p, 1) <= 1;
            int f2 = bounded_hamming(read + p, (len - 2) - p, ref + p + 2, ref_len - p - 2, 1) <= 1;
            if (r2 || f2) return 0;
        }
    }
    return 1;
}

/* ======================================================================
 * MdBuilder — src/alignment/md.rs:25-151
 * ====================================================================== */
typedef struct {
    char *text;
    size_t cap, len;
    uint32_t run, nm;
} md_builder;

static void md_push(md_builder *m, char c) {
    if (m->text && m->len + 1 < m->cap) m->text[m->len] = c;
    m->len++;
}
static void md_flush(md_builder *m) {
    char buf[16];
    int k = 0;
    uint32_t v = m->run;
    do {
        buf[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) md_push(m, buf[--k]);
    m->run = 0;
}
static char md_ref_letter(uint8_t b) {
    if ((b >= 'A' && b <= 'Z') || (b >= 'a' && b <= 'z'))
        return (char)(b >= 'a' ? b - 32 : b);
    return 'N';
}
static int identical_base_is_edit(uint8_t b) {
    uint8_t u = (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b;
    return !(u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == '=');
}
static int eq_ignore_case(uint8_t a, uint8_t b) {
    uint8_t ua = (a >= 'a' && a <= 'z') ? (uint8_t)(a - 32) : a;
    uint8_t ub = (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b;
    return ua == ub;
}
static int md_column(md_builder *m, uint8_t q, uint8_t r) {
    if (q == '=' || eq_ignore_case(q, r)) {
        m->run++;
        m->nm += (uint32_t)identical_base_is_edit(q);
    } else {
        md_flush(m);
        md_push(m, md_ref_letter(r));
        m->nm++;
    }
    return q == r;
}
static void md_finish(md_builder *m) {
    md_flush(m);
    if (m->text && m->cap) m->text[m->len < m->cap ? m->len : m->cap - 1] = 0;
}

uint32_t ko_ungapped_tags(const uint8_t *read, const uint8_t *ref, size_t len, char *md,
                          size_t md_cap) {
    md_builder m = {md, md_cap, 0, 0, 0};
    for (size_t k = 0; k < len; k++) md_column(&m, read[k], ref[k]);
    md_finish(&m);
    return m.nm;
}

/* nm_md_from_cigar — src/alignment/md.rs:162-204 */
uint32_t ko_nm_md_from_cigar(const uint8_t *read, size_t read_len, const uint8_t *ref,
                             size_t ref_len, size_t read_start, size_t ref_start,
                             const uint32_t *cigar, int32_t n_cigar, char *md, size_t md_cap) {
    md_builder m = {md, md_cap, 0, 0, 0};
    size_t q = read_start, r = ref_start;
    for (int32_t k = 0; k < n_cigar; k++) {
        size_t len = cigar[k] >> 4;
        switch (cigar[k] & 15) {
        case KO_M:
            for (size_t t = 0; t < len; t++) {
                uint8_t qb = q < read_len ? read[q] : 'N';
                uint8_t rb = r < ref_len ? ref[r] : 'N';
                md_column(&m, qb, rb);
                q++;
                r++;
            }
            break;
        case KO_I:
            m.nm += (uint32_t)len;
            q += len;
            break;
        case KO_D: {
            size_t end = r + len < ref_len ? r + len : ref_len;
            size_t start = r < end ? r : end;
            if (end > start) {
                md_flush(&m);
                md_push(&m, '^');
                for (size_t t = start; t < end; t++) md_push(&m, md_ref_letter(ref[t]));
                m.nm += (uint32_t)(end - start);
            }
            m.nm += (uint32_t)(len - (end - start));
            r += len;
            break;
        }
        case KO_S:
            q += len;
            break;
        case KO_N:
            r += len;
            break;
        }
    }
    md_finish(&m);
    return m.nm;
}

size_t ko_cigar_string(const uint32_t *cigar, int32_t n, char *out, size_t cap) {
    static const char ops[] = "MIDNS";
    size_t len = 0;
    for (int32_t k = 0; k < n; k++) {
        char buf[16];
        int b = 0;
        uint32_t v = cigar[k] >> 4;
        do {
            buf[b++] = (char)('0' + v % 10);
            v /= 10;
        } while (v);
        while (b) {
            if (len + 1 < cap) out[len] = buf[b - 1];
            len++;
            b--;
        }
        if (len + 1 < cap) out[len] = ops[cigar[k] & 15];
        len++;
    }
    if (cap) out[len < cap ? len : cap - 1] = 0;
    return len;
}

/* build_simd_window : src/pipeline/stage4_alignment.rs:1426-1445 */
int ko_build_simd_window(uint64_t ref_len, uint32_t chain_ref_start, uint32_t chain_read_start,
                         uint32_t read_len, int32_t bandwidth, uint32_t *win_start,
                         uint32_t *win_len) {
    uint64_t band = (uint64_t)(bandwidth > 1 ? bandwidth : 1);
    uint64_t flank = band * 2;
    uint64_t wl = read_len + flank * 2;
    int32_t desired = (int32_t)chain_ref_start - (int32_t)chain_read_start - (int32_t)flank;
    if (desired < 0) return 0;
    if ((uint64_t)desired + wl > ref_len) return 0;
    *win_start = (uint32_t)desired;
    *win_len = (uint32_t)wl;
    return 1;
}

/* ======================================================================
 * prefilter : src/alignment/prefilter.rs:82-453
 * count_mismatches (kira-simd 0.1.0 via src/simd/mod.rs:56-58) = number of
 * unequal bytes of two equal-length slices.
 * ====================================================================== */
static size_t count_mismatches(const uint8_t *a, const uint8_t *b, size_t n) {
    size_t m = 0;
    for (size_t k = 0; k < n; k++) m += a[k] != b[k];
    return m;
}

uint16_t ko_identity_x10000(size_t len, size_t mism) {
    if (len == 0) return 0;
    uint32_t matches = (uint32_t)(len > mism ? len - mism : 0);
    uint64_t prod = (uint64_t)matches * 10000u;
    uint32_t id = prod > UINT32_MAX ? UINT32_MAX : (uint32_t)prod;
    id /= (uint32_t)len;
    return (uint16_t)(id > 65535 ? 65535 : id);
}

size_t ko_min_len_required(size_t read_len, uint16_t id, size_t mism, size_t span_slack) {
    size_t slack = 0;
    if (id >= 9933) slack = 15;
    else if (id >= 9900 && mism <= 2) slack = 10;
    else if (id >= 9870 && mism <= 3) slack = 5;
    if (span_slack > 0 && span_slack < slack) slack = span_slack;
    size_t min_len = read_len > slack ? read_len - slack : 0;
    size_t floor_ = (read_len * 85) / 100;
    return floor_ > min_len ? floor_ : min_len;
}

static void ungapped_extend(const uint8_t *read, size_t read_len, const uint8_t *ref,
                            size_t ref_len, const ko_span *c, const ko_cfg *cfg, int tail,
                            size_t *left_out, size_t *right_out) {
    size_t left = 0, right = 0;
    size_t max_left = c->read_start < c->ref_start ? c->read_start : c->ref_start;
    uint32_t mr_a = (uint32_t)read_len - c->read_end;
    uint32_t mr_b = (uint32_t)ref_len - c->ref_end;
    size_t max_right = mr_a < mr_b ? mr_a : mr_b;
    const size_t block = 32;
    int32_t score = 0, best = 0;
    size_t i = 0;
    int broke = 0;
    while (i + block <= max_left) {
        size_t ro = c->read_start - i - block, fo = c->ref_start - i - block;
        int32_t mism = (int32_t)count_mismatches(read + ro, ref + fo, block);
        score += ((int32_t)block - mism) * cfg->match_score - mism * cfg->mismatch;
        if (score > best) {
            best = score;
            left = i + block;
        } else if (best - score > cfg->xdrop) {
            broke = 1;
            break;
        }
        i += block;
    }
    if (tail && !broke && i < max_left) {
        size_t rem = max_left - i;
        size_t ro = c->read_start - i - rem, fo = c->ref_start - i - rem;
        int32_t mism = (int32_t)count_mismatches(read + ro, ref + fo, rem);
        score += ((int32_t)rem - mism) * cfg->match_score - mism * cfg->mismatch;
        if (score > best) left = i + rem;
    }
    score = 0;
    best = 0;
    i = 0;
    broke = 0;
    while (i + block <= max_right) {
        size_t ro = c->read_end + i, fo = c->ref_end + i;
        int32_t mism = (int32_t)count_mismatches(read + ro, ref + fo, block);
        score += ((int32_t)block - mism) * cfg->match_score - mism * cfg->mismatch;
        if (score > best) {
            best = score;
            right = i + block;
        } else if (best - score > cfg->xdrop) {
            broke = 1;
            break;
        }
        i += block;
    }
    if (tail && !broke && i < max_right) {
        size_t rem = max_right - i;
        size_t ro = c->read_end + i, fo = c->ref_end + i;
        int32_t mism = (int32_t)count_mismatches(read + ro, ref + fo, rem);
        score += ((int32_t)rem - mism) * cfg->match_score - mism * cfg->mismatch;
        if (score > best) right = i + rem;
    }
    *left_out = left;
    *right_out = right;
}

enum { R_ACCEPT = 0, R_REJECT = 1, R_FALLBACK = 2 };
enum {
    PR_ACCEPTED = 0, PR_DISABLED, PR_NOT_SHORT, PR_NOT_TOP1, PR_MAX_ALN, PR_SPAN_SMALL,
    PR_MISM_HIGH, PR_ID_LOW, PR_INDEL, PR_OTHER
};
#define PR_OTHER_(k) (PR_OTHER + 16 * (k))

void ko_prefilter_chain(const uint8_t *read, size_t read_len, const uint8_t *ref, size_t ref_len,
                        const ko_span *chain, const ko_cfg *cfg, int is_top1, int accept_enable,
                        int accept_only_top1, size_t accept_span_slack,
                        float accept_min_identity, size_t accept_max_mismatches,
                        int32_t accept_require_score_margin, int32_t score_margin,
                        size_t long_read_threshold, int multi_alignments_enabled,
                        int tail_ext, int32_t ungap_min_id, ko_prefilter_out *o, char *md,
                        size_t md_cap) {
    memset(o, 0, sizeof(*o));
    if (md && md_cap) md[0] = 0;
    o->result = R_FALLBACK;
    if (read_len == 0) { o->reason = PR_OTHER_(1); return; }
    if (chain->read_start > chain->read_end || chain->ref_start > chain->ref_end) {
        o->reason = PR_OTHER_(2);
        return;
    }
    size_t anchor_len = chain->read_end - chain->read_start;
    if (anchor_len == 0) { o->reason = PR_OTHER_(3); return; }
    size_t le, re;
    ungapped_extend(read, read_len, ref, ref_len, chain, cfg, tail_ext, &le, &re);
    size_t total = anchor_len + le + re;
    if (total == 0) { o->reason = PR_OTHER_(4); return; }
    int64_t urs = (int64_t)(int32_t)chain->read_start - (int64_t)(int32_t)le;
    int64_t ufs = (int64_t)(int32_t)chain->ref_start - (int64_t)(int32_t)le;
    if (urs < 0 || ufs < 0) { o->reason = PR_OTHER_(5); return; }
    if ((size_t)urs + total > read_len || (size_t)ufs + total > ref_len) {
        o->reason = PR_OTHER_(6);
        return;
    }
    const uint8_t *ru = read + urs, *fu = ref + ufs;
    size_t hamming = count_mismatches(ru, fu, total);
    int32_t score = (int32_t)(total > hamming ? total - hamming : 0) * cfg->match_score -
                    (int32_t)hamming * cfg->mismatch;
    uint16_t len16 = total > 65535 ? 65535 : (uint16_t)total;
    size_t hm = hamming < total ? hamming : total;
    uint16_t mism16 = hm > 65535 ? 65535 : (uint16_t)hm;
    uint16_t matches16 = len16 > mism16 ? (uint16_t)(len16 - mism16) : 0;
    uint16_t id = ko_identity_x10000(total, hamming);
    o->has_metrics = 1;
    o->len = len16;
    o->mism = mism16;
    o->matches = matches16;
    o->identity_x10000 = id;
    o->score = score;
    o->read_start = (uint64_t)urs;
    o->ref_start = (uint64_t)ufs;
    if (total < read_len / 2) { o->result = R_REJECT; o->reason = PR_SPAN_SMALL; return; }
    if (!accept_enable) { o->reason = PR_DISABLED; return; }
    if (multi_alignments_enabled) { o->reason = PR_MAX_ALN; return; }
    if (accept_only_top1 && !is_top1) { o->reason = PR_NOT_TOP1; return; }
    if (read_len > long_read_threshold) { o->reason = PR_NOT_SHORT; return; }
    if (accept_require_score_margin > 0 && score_margin < accept_require_score_margin) {
        o->reason = PR_OTHER_(7);
        return;
    }
    uint32_t rspan = chain->ref_end >= chain->ref_start ? chain->ref_end - chain->ref_start : 0;
    uint32_t qspan = chain->read_end >= chain->read_start ? chain->read_end - chain->read_start : 0;
    if (rspan != qspan) { o->reason = PR_INDEL; return; }
    size_t hard_max = (read_len * 5 + 149) / 150;
    size_t max_mism = accept_max_mismatches < hard_max ? accept_max_mismatches : hard_max;
    if (hamming > max_mism) { o->reason = PR_MISM_HIGH; return; }
    float f = roundf(accept_min_identity * 100.0f);
    uint16_t cfg_floor = f != f ? 0 : (f <= 0 ? 0 : (f >= 65535.0f ? 65535 : (uint16_t)f));
    uint16_t eff_floor = ungap_min_id >= 0 ? (uint16_t)ungap_min_id : cfg_floor;
    if (id < eff_floor) { o->reason = PR_ID_LOW; return; }
    size_t min_len = ko_min_len_required(read_len, id, hamming, accept_span_slack);
    if (total < min_len) { o->reason = PR_SPAN_SMALL; return; }
    /* build_ungapped_alignment — prefilter.rs:455-517 */
    size_t span = len16, rs = (size_t)urs, fs = (size_t)ufs;
    int n = 0;
    if (rs > 0) o->cigar[n++] = (uint32_t)(rs << 4) | KO_S;
    o->cigar[n++] = (uint32_t)(span << 4) | KO_M;
    if (rs + span < read_len) o->cigar[n++] = (uint32_t)((read_len - rs - span) << 4) | KO_S;
    o->n_cigar = n;
    md_builder m = {md, md_cap, 0, 0, 0};
    int32_t mm = 0;
    for (size_t k = 0; k < span; k++) mm += !md_column(&m, read[rs + k], ref[fs + k]);
    md_finish(&m);
    o->aln_nm = m.nm;
    o->aln_score = ((int32_t)span - mm) * cfg->match_score - mm * cfg->mismatch;
    o->aln_ref_start = (uint32_t)fs;
    o->aln_ref_end = (uint32_t)(fs + span);
    o->aln_read_start = (uint32_t)rs;
    o->aln_read_end = (uint32_t)(rs + span);
    o->result = R_ACCEPT;
    o->reason = PR_ACCEPTED;
}

/* ======================================================================
 * Smith-Waterman
 * ====================================================================== */
#define TR_STOP 0
#define TR_DIAG 1
#define TR_E 2
#define TR_F 3
#define TR_DIR 3
#define TR_EEXT 4
#define TR_FEXT 8

static void cig_push(ko_sw_result *res, int op) {
    if (res->n_cigar > 0 && (int)(res->cigar[res->n_cigar - 1] & 15) == op) {
        res->cigar[res->n_cigar - 1] += 16;
        return;
    }
    if (res->n_cigar < res->cigar_cap) res->cigar[res->n_cigar] = (1u << 4) | (uint32_t)op;
    res->n_cigar++;
}
static void cig_reverse(ko_sw_result *res) {
    int32_t n = res->n_cigar < res->cigar_cap ? res->n_cigar : res->cigar_cap;
    for (int32_t a = 0, b = n - 1; a < b; a++, b--) {
        uint32_t t = res->cigar[a];
        res->cigar[a] = res->cigar[b];
        res->cigar[b] = t;
    }
}

static inline int32_t sat_add32(int32_t a, int32_t b) {
    int64_t s = (int64_t)a + b;
    return s > INT32_MAX ? INT32_MAX : (s < INT32_MIN ? INT32_MIN : (int32_t)s);
}
static inline int32_t clamp32(int32_t v, int32_t lo, int32_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

/* banded_sw_with : src/alignment/mod.rs:1370-1650 (prev_cell /
 * prev_diag :1723-1754, push_cigar :1756-1764). */
void ko_banded_sw(const uint8_t *read, size_t q_len, const uint8_t *ref, size_t r_len,
                  int32_t offset, const ko_cfg *cfg, int32_t abort_score, const ko_knobs *kn,
                  ko_sw_result *res) {
    int32_t band = cfg->bandwidth > 1 ? cfg->bandwidth : 1;
    int32_t dp_xdrop = kn->dp_xdrop >= 0 ? kn->dp_xdrop : cfg->xdrop;
    int adaptive = kn->adaptive;
    int32_t gof = sat_add32(cfg->gap_open, cfg->gap_extend);
    int32_t ge = cfg->gap_extend;

    int32_t *row_starts = (int32_t *)malloc((q_len + 1) * sizeof(int32_t));
    int32_t *row_len = (int32_t *)calloc(q_len + 1, sizeof(int32_t));
    uint8_t **trace = (uint8_t **)calloc(q_len + 1, sizeof(uint8_t *));
    size_t cap = (size_t)2 * (size_t)band + 2;
    if (cap > r_len + 1) cap = r_len + 1;
    /* one trace slab for the call (the reference recycles per-thread row
     * buffers, mod.rs:1387-1409); row i uses its first row_len[i] bytes */
    uint8_t *trace_slab = (uint8_t *)malloc((q_len + 1) * cap);
    int64_t cells = 0;
    int32_t *prev_h = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *prev_e = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *cur_h = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *cur_e = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *cur_f = (int32_t *)malloc(cap * sizeof(int32_t));
    size_t prev_len = 0;
    int32_t prev_start = 1;
    for (size_t i = 0; i <= q_len; i++) row_starts[i] = 1;

    int32_t best_score = 0, best_qlen_score = 0;
    size_t best_i = 0, best_j = 0, best_qlen_j = 0;
    int early = 0;
    int32_t max_drift = adaptive ? kn->max_drift : 0;
    int32_t adaptive_center = 1 + offset;

    for (size_t i = 1; i <= q_len; i++) {
        int32_t static_center = (int32_t)i + offset;
        int32_t center = adaptive ? clamp32(adaptive_center, static_center - max_drift,
                                            static_center + max_drift)
                                  : static_center;
        int32_t j_start = center - band > 1 ? center - band : 1;
        int32_t j_end = center + band < (int32_t)r_len ? center + band : (int32_t)r_len;
        if (j_start > j_end) {
            row_starts[i] = 1;
            row_len[i] = 0;
            prev_len = 0;
            prev_start = 1;
            continue;
        }
        size_t rl = (size_t)(j_end - j_start + 1);
        row_starts[i] = j_start;
        row_len[i] = (int32_t)rl;
        trace[i] = trace_slab + i * cap;
        cells += (int64_t)rl;
        int32_t row_best = 0, row_best_j = -1;
        for (int32_t j = j_start; j <= j_end; j++) {
            size_t idx = (size_t)(j - j_start);
            int32_t qb = read[i - 1], rb = ref[j - 1];
            int32_t s = qb == rb ? cfg->match_score : -cfg->mismatch;
            int32_t h_diag = 0;
            int64_t di = (int64_t)j - 1 - prev_start;
            if (di >= 0 && (size_t)di < prev_len) h_diag = prev_h[di];
            int32_t h_match = h_diag + s;
            int64_t ui = (int64_t)j - prev_start;
            int up = ui >= 0 && (size_t)ui < prev_len;
            int32_t e_from_h = up ? prev_h[ui] - gof : I32_MIN_4;
            int32_t e_from_e = up ? prev_e[ui] - ge : I32_MIN_4;
            int32_t e = e_from_h > e_from_e ? e_from_h : e_from_e;
            int32_t f_from_h = I32_MIN_4, f_from_f = I32_MIN_4;
            if (idx > 0) {
                f_from_h = cur_h[idx - 1] - gof;
                f_from_f = cur_f[idx - 1] - ge;
            }
            int32_t f = f_from_h > f_from_f ? f_from_h : f_from_f;
            cur_e[idx] = e;
            cur_f[idx] = f;
            int32_t h = 0;
            uint8_t tr = 0;
            if (h_match >= e && h_match >= f && h_match > 0) {
                h = h_match;
                tr = TR_DIAG;
            } else if (e >= f && e > 0) {
                h = e;
                tr = TR_E;
            } else if (f > 0) {
                h = f;
                tr = TR_F;
            }
            cur_h[idx] = h;
            if (e_from_e > e_from_h) tr |= TR_EEXT;
            if (f_from_f > f_from_h) tr |= TR_FEXT;
            trace[i][idx] = tr;
            if (h > row_best) {
                row_best = h;
                row_best_j = j;
            }
            if (h > best_score) {
                best_score = h;
                best_i = i;
                best_j = (size_t)j;
            }
            if (i == q_len && h > best_qlen_score) {
                best_qlen_score = h;
                best_qlen_j = (size_t)j;
            }
        }
        int32_t *t;
        t = prev_h; prev_h = cur_h; cur_h = t;
        t = prev_e; prev_e = cur_e; cur_e = t;
        prev_len = rl;
        prev_start = j_start;
        if (adaptive) {
            int32_t raw = row_best_j >= 0 ? row_best_j + 1 : adaptive_center + 1;
            int32_t static_next = (int32_t)i + 1 + offset;
            adaptive_center = clamp32(raw, static_next - max_drift, static_next + max_drift);
        }
        if (dp_xdrop > 0 && best_score - row_best > dp_xdrop) {
            early = 1;
            break;
        }
        if (abort_score > I32_MIN_8) {
            int32_t remaining = (int32_t)(q_len - i);
            int32_t max_possible = best_score + remaining * cfg->match_score;
            if (max_possible < abort_score) {
                early = 1;
                break;
            }
        }
    }
    if (best_qlen_score > 0 && sat_add32(best_qlen_score, cfg->clip_penalty) > best_score) {
        best_score = best_qlen_score;
        best_i = q_len;
        best_j = best_qlen_j;
    }
    res->n_cigar = 0;
    int32_t i = (int32_t)best_i, j = (int32_t)best_j;
    int32_t read_end = i;
    uint32_t ref_end = (uint32_t)j;
    int state = 0; /* 0 H, 1 E, 2 F */
    while (i > 0 && j > 0) {
        int64_t idx = (int64_t)j - row_starts[i];
        if (idx < 0 || idx >= row_len[i]) break;
        uint8_t tr = trace[i][idx];
        if (state == 0) {
            int d = tr & TR_DIR;
            if (d == TR_STOP) break;
            if (d == TR_DIAG) {
                cig_push(res, KO_M);
                i--;
                j--;
            } else if (d == TR_E) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            cig_push(res, KO_I);
            i--;
            state = (tr & TR_EEXT) ? 1 : 0;
        } else {
            cig_push(res, KO_D);
            j--;
            state = (tr & TR_FEXT) ? 2 : 0;
        }
    }
    cig_reverse(res);
    res->ref_start = (uint32_t)j;
    res->ref_end = ref_end;
    res->read_start = i;
    res->read_end = read_end;
    res->score = best_score;
    res->early_abort = early;
    res->cells = cells;
    free(trace_slab);
    free(trace);
    free(row_starts);
    free(row_len);
    free(prev_h);
    free(prev_e);
    free(cur_h);
    free(cur_e);
    free(cur_f);
}

/* BatchInput::scalar_sw — src/alignment/mod.rs:60-72 */
static void scalar_sw(const ko_batch_input *in, const ko_cfg *cfg, const ko_knobs *kn,
                      ko_sw_result *res) {
    int64_t off = (int64_t)in->chain_ref_start - (int64_t)in->win_start -
                  (int64_t)in->chain_read_start;
    int64_t lo = -(int64_t)in->read_len, hi = (int64_t)in->window_len;
    off = off < lo ? lo : (off > hi ? hi : off);
    ko_banded_sw(in->read, in->read_len, in->window, in->window_len, (int32_t)off, cfg,
                 in->abort_score, kn, res);
}

/* traceback_interleaved — src/alignment/mod.rs:1667-1716 */
static void traceback_lane(const uint8_t *trace, size_t width, int32_t lanes, int32_t lane,
                           size_t tsize, int32_t i, int32_t j, ko_sw_result *res,
                           int32_t *oi, int32_t *oj) {
    res->n_cigar = 0;
    int state = 0;
    while (i > 0 && j > 0) {
        size_t idx = ((size_t)i * width + (size_t)j) * (size_t)lanes + (size_t)lane;
        if (idx >= tsize) break;
        uint8_t tr = trace[idx];
        if (state == 0) {
            int d = tr & TR_DIR;
            if (d == TR_STOP) break;
            if (d == TR_DIAG) {
                cig_push(res, KO_M);
                i--;
                j--;
            } else if (d == TR_E) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            cig_push(res, KO_I);
            i--;
            state = (tr & TR_EEXT) ? 1 : 0;
        } else {
            cig_push(res, KO_D);
            j--;
            state = (tr & TR_FEXT) ? 2 : 0;
        }
    }
    cig_reverse(res);
    *oi = i;
    *oj = j;
}

/* Lane arithmetic of the three full-window kernels: saturating to the lane
 * width (i16 _mm256_adds_epi16, i8 _mm256_adds_epi8) or plain i32 (NEON). */
static inline int32_t lane_add(int32_t a, int32_t b, int bits) {
    int64_t s = (int64_t)a + b;
    if (bits == 16) return s > 32767 ? 32767 : (s < -32768 ? -32768 : (int32_t)s);
    if (bits == 8) return s > 127 ? 127 : (s < -128 ? -128 : (int32_t)s);
    return (int32_t)s;
}

/* sw_batch_avx2_inner (mod.rs:1914-2135), sw_batch_int8_inner
 * (sw_int8_vnni.rs:141-371), sw_batch_neon (mod.rs:2139-2324): full-window
 * local affine SW over a lane group with the group-wide early stop. */
void ko_sw_group(const ko_batch_input *in, int32_t n, const ko_cfg *cfg, int bits,
                 const ko_knobs *kn, ko_sw_result *res) {
    int32_t lanes = n;
    size_t q = in[0].read_len, r = in[0].window_len;
    /* Lane constants exactly as each kernel builds them:
     * i16 mod.rs:1931-1935, i8 sw_int8_vnni.rs:157-160, i32 mod.rs:2151-2154. */
    int32_t neg, go, gx, vm, vx;
    int32_t gof = sat_add32(cfg->gap_open, cfg->gap_extend);
    if (bits == 16) {
        neg = -16384;
        go = (int16_t)(-gof);
        gx = (int16_t)(-cfg->gap_extend);
        vm = (int16_t)cfg->match_score;
        vx = (int16_t)(-cfg->mismatch);
    } else if (bits == 8) {
        neg = -100;
        go = (int8_t)(-(int8_t)gof);
        gx = (int8_t)(-(int8_t)cfg->gap_extend);
        vm = (int8_t)cfg->match_score;
        vx = (int8_t)(-(int8_t)cfg->mismatch);
    } else {
        neg = I32_MIN_4;
        go = -gof;
        gx = -cfg->gap_extend;
        vm = cfg->match_score;
        vx = -cfg->mismatch;
    }
    size_t tw = r + 1;
    size_t tsize = (q + 1) * tw * (size_t)lanes;
    uint8_t *trace = (uint8_t *)calloc(tsize, 1);
    int32_t *prev_h = (int32_t *)malloc(tw * lanes * sizeof(int32_t));
    int32_t *prev_e = (int32_t *)malloc(tw * lanes * sizeof(int32_t));
    int32_t *cur_h = (int32_t *)malloc(tw * lanes * sizeof(int32_t));
    int32_t *cur_e = (int32_t *)malloc(tw * lanes * sizeof(int32_t));
    for (size_t k = 0; k < tw * lanes; k++) {
        prev_h[k] = 0;
        cur_h[k] = 0;
        prev_e[k] = neg;
        cur_e[k] = neg;
    }
    int32_t best[64], bi[64], bj[64], bq[64], bqj[64], done[64], sat[64];
    for (int k = 0; k < lanes; k++) {
        best[k] = bi[k] = bj[k] = bq[k] = bqj[k] = done[k] = sat[k] = 0;
    }
    size_t rows_done = q;
    for (size_t i = 1; i <= q; i++) {
        for (int k = 0; k < lanes; k++) {
            cur_h[k] = 0;
            cur_e[k] = neg;
        }
        int32_t cur_f[64];
        for (int k = 0; k < lanes; k++) cur_f[k] = neg;
        for (size_t j = 1; j <= r; j++) {
            for (int k = 0; k < lanes; k++) {
                int32_t s = in[k].read[i - 1] == in[k].window[j - 1] ? vm : vx;
                int32_t hd = prev_h[(j - 1) * lanes + k];
                int32_t hm = lane_add(hd, s, bits);
                int32_t efh = lane_add(prev_h[j * lanes + k], go, bits);
                int32_t efe = lane_add(prev_e[j * lanes + k], gx, bits);
                int32_t e = efh > efe ? efh : efe;
                int32_t ffh = lane_add(cur_h[(j - 1) * lanes + k], go, bits);
                int32_t fff = lane_add(cur_f[k], gx, bits);
                int32_t f = ffh > fff ? ffh : fff;
                int32_t h = hm > e ? hm : e;
                h = h > f ? h : f;
                h = h > 0 ? h : 0;
                cur_h[j * lanes + k] = h;
                cur_e[j * lanes + k] = e;
                cur_f[k] = f;
                uint8_t tr;
                if (h == 0) tr = TR_STOP;
                else if (h == hm) tr = TR_DIAG;
                else if (h == e) tr = TR_E;
                else tr = TR_F;
                if (efe > efh) tr |= TR_EEXT;
                if (fff > ffh) tr |= TR_FEXT;
                trace[(i * tw + j) * lanes + k] = tr;
                if (h > best[k]) {
                    best[k] = h;
                    bi[k] = (int32_t)i;
                    bj[k] = (int32_t)j;
                }
                if (i == q && h > bq[k]) {
                    bq[k] = h;
                    bqj[k] = (int32_t)j;
                }
                if (bits == 8 && h > 120) sat[k] = 1;
            }
        }
        int32_t remaining = (int32_t)(q - i);
        int all_done = 1;
        for (int k = 0; k < lanes; k++) {
            if (done[k]) continue;
            int32_t ab = in[k].abort_score;
            if (ab > I32_MIN_8) {
                int32_t mp = best[k] + remaining * cfg->match_score;
                if (mp < ab) done[k] = 1;
            }
            if (!done[k]) all_done = 0;
        }
        if (all_done) {
            rows_done = i;
            break;
        }
        int32_t *t;
        t = prev_h; prev_h = cur_h; cur_h = t;
        t = prev_e; prev_e = cur_e; cur_e = t;
    }
    for (int k = 0; k < lanes; k++) {
        if (sat[k]) {
            scalar_sw(&in[k], cfg, kn, &res[k]);
            continue;
        }
        int32_t si, sj, bs;
        if (bq[k] > 0 && sat_add32(bq[k], cfg->clip_penalty) > best[k]) {
            si = (int32_t)q;
            sj = bqj[k];
            bs = bq[k];
        } else {
            si = bi[k];
            sj = bj[k];
            bs = best[k];
        }
        int32_t oi, oj;
        traceback_lane(trace, tw, lanes, k, tsize, si, sj, &res[k], &oi, &oj);
        res[k].ref_start = (uint32_t)oj;
        res[k].ref_end = (uint32_t)sj;
        res[k].read_start = oi;
        res[k].read_end = si;
        res[k].score = bs;
        res[k].early_abort = done[k];
        res[k].cells = (int64_t)(rows_done * r);
    }
    free(trace);
    free(prev_h);
    free(prev_e);
    free(cur_h);
    free(cur_e);
}

/* sw_batch_avx2 cost-fit gate — src/alignment/mod.rs:1854-1868 */
static int avx2_costs_fit(const ko_cfg *c, size_t q, size_t r) {
    int32_t mx = 32767;
    int fit = c->match_score >= 0 && c->match_score <= mx && c->mismatch >= 0 &&
              c->mismatch <= mx && c->gap_open >= 0 && c->gap_extend >= 0 &&
              sat_add32(c->gap_open, c->gap_extend) <= mx;
    size_t mn = q < r ? q : r;
    if (!fit || q > (size_t)mx || r > (size_t)mx || (int64_t)mn * c->match_score > mx) return 0;
    return 1;
}

/* int8_path_viable — src/alignment/sw_int8_vnni.rs:47-55 */
static int int8_viable(size_t read_len, const ko_cfg *c) {
    int64_t ms = c->match_score > 1 ? c->match_score : 1;
    int64_t upper = (int64_t)read_len * ms;
    return upper < 120 && c->match_score >= 0 && c->match_score <= 32 && c->mismatch >= 0 &&
           c->mismatch <= 32 && c->gap_open >= 0 && c->gap_open <= 32 && c->gap_extend >= 0 &&
           c->gap_extend <= 32;
}

#ifdef __AVX2__
/* ko_sw_group's 16-bit form in AVX2 registers: sixteen i16 lanes with
 * saturating adds, as sw_batch_avx2_inner computes them, over the same cells,
 * trace bytes (16 per cell, lane k at byte k) and per-lane bookkeeping, so the
 * results are those of ko_sw_group(bits = 16); a CPU test compares the two
 * forms.  Lanes n .. 15 repeat lane 0's bytes and are not read back.  The
 * cost gate (avx2_costs_fit) keeps scores and coordinates inside i16. */
static void sw_group16_avx2(const ko_batch_input *in, int32_t n, const ko_cfg *cfg,
                            const ko_knobs *kn, ko_sw_result *res) {
    enum { LN = 16 };
    const size_t q = in[0].read_len, r = in[0].window_len, tw = r + 1;
    const int32_t gof = sat_add32(cfg->gap_open, cfg->gap_extend);
    const __m256i Z = _mm256_setzero_si256(), NEG = _mm256_set1_epi16(-16384);
    const __m256i GO = _mm256_set1_epi16((int16_t)(-gof));
    const __m256i GX = _mm256_set1_epi16((int16_t)(-cfg->gap_extend));
    const __m256i VM = _mm256_set1_epi16((int16_t)cfg->match_score);
    const __m256i VX = _mm256_set1_epi16((int16_t)(-cfg->mismatch));
    const __m256i T1 = _mm256_set1_epi16(TR_DIAG), T2 = _mm256_set1_epi16(TR_E);
    const __m256i T3 = _mm256_set1_epi16(TR_F), TEE = _mm256_set1_epi16(TR_EEXT);
    const __m256i TFE = _mm256_set1_epi16(TR_FEXT);
    const size_t tsize = (q + 1) * tw * LN;
    uint8_t *trace = (uint8_t *)malloc(tsize);
    memset(trace, 0, tw * LN); /* row 0 */
    uint8_t *wt = (uint8_t *)malloc(r * LN + 32);
    for (size_t j = 0; j < r; j++)
        for (int k = 0; k < LN; k++) wt[j * LN + k] = in[k < n ? k : 0].window[j];
    int16_t *buf = (int16_t *)aligned_alloc(32, 4 * tw * LN * sizeof(int16_t));
    int16_t *prev_h = buf, *prev_e = buf + tw * LN, *cur_h = buf + 2 * tw * LN,
            *cur_e = buf + 3 * tw * LN;
    for (size_t j = 0; j < tw; j++) {
        _mm256_store_si256((__m256i *)(prev_h + j * LN), Z);
        _mm256_store_si256((__m256i *)(prev_e + j * LN), NEG);
    }
    __m256i best = Z, bi = Z, bj = Z, bq = Z, bqj = Z;
    int32_t done[LN] = {0};
    size_t rows_done = q;
    for (size_t i = 1; i <= q; i++) {
        _mm256_store_si256((__m256i *)cur_h, Z);
        _mm256_store_si256((__m256i *)cur_e, NEG);
        memset(trace + i * tw * LN, 0, LN); /* column 0 */
        uint8_t rb[LN];
        for (int k = 0; k < LN; k++) rb[k] = in[k < n ? k : 0].read[i - 1];
        const __m128i rv = _mm_loadu_si128((const __m128i *)rb);
        const __m256i vi = _mm256_set1_epi16((int16_t)i);
        const int last = i == q;
        __m256i hleft = Z, f = NEG, hdiag = Z;
        uint8_t *trow = trace + i * tw * LN;
        for (size_t j = 1; j <= r; j++) {
            const __m128i eq8 = _mm_cmpeq_epi8(_mm_loadu_si128((const __m128i *)(wt + (j - 1) * LN)), rv);
            const __m256i s = _mm256_blendv_epi8(VX, VM, _mm256_cvtepi8_epi16(eq8));
            const __m256i hm = _mm256_adds_epi16(hdiag, s);
            const __m256i hup = _mm256_load_si256((const __m256i *)(prev_h + j * LN));
            const __m256i efh = _mm256_adds_epi16(hup, GO);
            const __m256i efe =
                _mm256_adds_epi16(_mm256_load_si256((const __m256i *)(prev_e + j * LN)), GX);
            const __m256i e = _mm256_max_epi16(efh, efe);
            const __m256i ffh = _mm256_adds_epi16(hleft, GO);
            const __m256i fff = _mm256_adds_epi16(f, GX);
            f = _mm256_max_epi16(ffh, fff);
            __m256i h = _mm256_max_epi16(_mm256_max_epi16(hm, e), f);
            h = _mm256_max_epi16(h, Z);
            _mm256_store_si256((__m256i *)(cur_h + j * LN), h);
            _mm256_store_si256((__m256i *)(cur_e + j * LN), e);
            /* direction: stop if 0, else diagonal, E, F in that order */
            __m256i tr = T3;
            tr = _mm256_blendv_epi8(tr, T2, _mm256_cmpeq_epi16(h, e));
            tr = _mm256_blendv_epi8(tr, T1, _mm256_cmpeq_epi16(h, hm));
            tr = _mm256_blendv_epi8(tr, Z, _mm256_cmpeq_epi16(h, Z));
            tr = _mm256_or_si256(tr, _mm256_and_si256(_mm256_cmpgt_epi16(efe, efh), TEE));
            tr = _mm256_or_si256(tr, _mm256_and_si256(_mm256_cmpgt_epi16(fff, ffh), TFE));
            _mm_storeu_si128((__m128i *)(trow + j * LN),
                             _mm_packus_epi16(_mm256_castsi256_si128(tr),
                                              _mm256_extracti128_si256(tr, 1)));
            const __m256i vj = _mm256_set1_epi16((int16_t)j);
            const __m256i gt = _mm256_cmpgt_epi16(h, best);
            best = _mm256_max_epi16(best, h);
            bi = _mm256_blendv_epi8(bi, vi, gt);
            bj = _mm256_blendv_epi8(bj, vj, gt);
            if (last) {
                const __m256i gq = _mm256_cmpgt_epi16(h, bq);
                bq = _mm256_max_epi16(bq, h);
                bqj = _mm256_blendv_epi8(bqj, vj, gq);
            }
            hleft = h;
            hdiag = hup;
        }
        int16_t bl[LN];
        _mm256_storeu_si256((__m256i *)bl, best);
        const int32_t remaining = (int32_t)(q - i);
        int all_done = 1;
        for (int k = 0; k < n; k++) {
            if (done[k]) continue;
            const int32_t ab = in[k].abort_score;
            if (ab > I32_MIN_8 && bl[k] + remaining * cfg->match_score < ab) done[k] = 1;
            if (!done[k]) all_done = 0;
        }
        if (all_done) {
            rows_done = i;
            break;
        }
        int16_t *t = prev_h;
        prev_h = cur_h;
        cur_h = t;
        t = prev_e;
        prev_e = cur_e;
        cur_e = t;
    }
    int16_t vbest[LN], vbi[LN], vbj[LN], vbq[LN], vbqj[LN];
    _mm256_storeu_si256((__m256i *)vbest, best);
    _mm256_storeu_si256((__m256i *)vbi, bi);
    _mm256_storeu_si256((__m256i *)vbj, bj);
    _mm256_storeu_si256((__m256i *)vbq, bq);
    _mm256_storeu_si256((__m256i *)vbqj, bqj);
    for (int k = 0; k < n; k++) {
        int32_t si, sj, bs;
        if (vbq[k] > 0 && sat_add32(vbq[k], cfg->clip_penalty) > vbest[k]) {
            si = (int32_t)q;
            sj = vbqj[k];
            bs = vbq[k];
        } else {
            si = vbi[k];
            sj = vbj[k];
            bs = vbest[k];
        }
        int32_t oi, oj;
        traceback_lane(trace, tw, LN, k, tsize, si, sj, &res[k], &oi, &oj);
        res[k].ref_start = (uint32_t)oj;
        res[k].ref_end = (uint32_t)sj;
        res[k].read_start = oi;
        res[k].read_end = si;
        res[k].score = bs;
        res[k].early_abort = done[k];
        res[k].cells = (int64_t)(rows_done * r);
    }
    (void)kn;
    free(trace);
    free(wt);
    free(buf);
}
#endif

static void avx2_chunk(const ko_batch_input *in, int32_t n, const ko_cfg *cfg,
                       const ko_knobs *kn, ko_sw_result *res) {
    if (!avx2_costs_fit(cfg, in[0].read_len, in[0].window_len)) {
        for (int32_t k = 0; k < n; k++) scalar_sw(&in[k], cfg, kn, &res[k]);
        return;
    }
#ifdef __AVX2__
    sw_group16_avx2(in, n, cfg, kn, res);
#else
    ko_sw_group(in, n, cfg, 16, kn, res);
#endif
}

/* align_batch_simd SW part — src/alignment/mod.rs:1111-1149 and
 * sw_dispatch_avx_vnni :1824-1841. */
void ko_align_batch_sw(const ko_batch_input *in, int32_t n, const ko_cfg *cfg, int mode,
                       int detected, const ko_knobs *kn, ko_sw_result *res) {
    if (n <= 0) return;
    uint32_t L = in[0].read_len, W = in[0].window_len;
    for (int32_t k = 0; k < n; k++) {
        if (in[k].read_len != L || in[k].window_len != W) {
            for (int32_t t = 0; t < n; t++) scalar_sw(&in[t], cfg, kn, &res[t]);
            return;
        }
    }
    if (mode == 2 && detected == 2) {
        if (int8_viable(L, cfg) && n <= 32) {
            ko_sw_group(in, n, cfg, 8, kn, res);
            return;
        }
        for (int32_t s = 0; s < n; s += 16)
            avx2_chunk(in + s, n - s < 16 ? n - s : 16, cfg, kn, res + s);
        return;
    }
    if (mode == 1 && (detected == 1 || detected == 2)) {
        for (int32_t s = 0; s < n; s += 16)
            avx2_chunk(in + s, n - s < 16 ? n - s : 16, cfg, kn, res + s);
        return;
    }
    if (mode == 3 && detected == 3) {
        for (int32_t s = 0; s < n; s += 4)
            ko_sw_group(in + s, n - s < 4 ? n - s : 4, cfg, 32, kn, res + s);
        return;
    }
    for (int32_t k = 0; k < n; k++) scalar_sw(&in[k], cfg, kn, &res[k]);
}

void ko_build_alignment(const uint8_t *read, size_t read_len, const uint8_t *window,
                        size_t window_len, uint32_t win_start, const ko_sw_result *sw,
                        ko_alignment *out, uint32_t *cig, int32_t cap, char *md,
                        size_t md_cap) {
    int32_t n = 0;
    if (sw->read_start > 0 && n < cap) cig[n++] = ((uint32_t)sw->read_start << 4) | KO_S;
    for (int32_t k = 0; k < sw->n_cigar && n < cap; k++) cig[n++] = sw->cigar[k];
    if (sw->read_end < (int32_t)read_len && n < cap)
        cig[n++] = ((uint32_t)((int32_t)read_len - sw->read_end) << 4) | KO_S;
    out->n_cigar = n;
    out->ref_start = win_start + sw->ref_start;
    out->ref_end = win_start + sw->ref_end;
    out->read_start = (uint32_t)sw->read_start;
    out->read_end = (uint32_t)sw->read_end;
    out->score = sw->score;
    out->nm = ko_nm_md_from_cigar(read, read_len, window, window_len, 0, sw->ref_start, cig, n,
                                  md, md_cap);
}

/* ---------------- multi-threaded scalar SW baseline ---------------- */
typedef struct {
    int32_t lo, hi;
    const uint8_t *reads, *refs;
    const int64_t *read_off, *ref_off;
    const int32_t *read_len, *ref_len, *offsets, *abort_scores, *bands;
    const ko_cfg *cfg;
    const ko_knobs *kn;
    int32_t *out_score, *out_end;
    int64_t *out_cells;
} sw_task;

static void *sw_worker(void *arg) {
    sw_task *t = (sw_task *)arg;
    uint32_t cig[4096];
    for (int32_t k = t->lo; k < t->hi; k++) {
        ko_sw_result r;
        memset(&r, 0, sizeof(r));
        r.cigar = cig;
        r.cigar_cap = 4096;
        ko_cfg c = *t->cfg;
        if (t->bands) c.bandwidth = t->bands[k];
        ko_banded_sw(t->reads + t->read_off[k], (size_t)t->read_len[k], t->refs + t->ref_off[k],
                     (size_t)t->ref_len[k], t->offsets[k], &c, t->abort_scores[k], t->kn, &r);
        t->out_score[k] = r.score;
        t->out_end[k] = (int32_t)r.ref_end;
        if (t->out_cells) t->out_cells[k] = r.cells;
    }
    return NULL;
}

void ko_banded_sw_batch_mt(int32_t n, const uint8_t *reads, const int64_t *read_off,
                           const int32_t *read_len, const uint8_t *refs, const int64_t *ref_off,
                           const int32_t *ref_len, const int32_t *offsets,
                           const int32_t *abort_scores, const int32_t *bands,
                           const ko_cfg *cfg, const ko_knobs *kn, int32_t *out_score,
                           int32_t *out_end, int64_t *out_cells, int threads) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    sw_task *tk = (sw_task *)calloc((size_t)threads, sizeof(sw_task));
    for (int k = 0; k < threads; k++) {
        sw_task *t = &tk[k];
        t->lo = (int32_t)((int64_t)n * k / threads);
        t->hi = (int32_t)((int64_t)n * (k + 1) / threads);
        t->reads = reads;
        t->refs = refs;
        t->read_off = read_off;
        t->ref_off = ref_off;
        t->read_len = read_len;
        t->ref_len = ref_len;
        t->offsets = offsets;
        t->abort_scores = abort_scores;
        t->bands = bands;
        t->out_cells = out_cells;
        t->cfg = cfg;
        t->kn = kn;
        t->out_score = out_score;
        t->out_end = out_end;
        pthread_create(&th[k], NULL, sw_worker, t);
    }
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    free(th);
    free(tk);
}
