// Alignment records of the path, host side: build_alignment
// (src/alignment/mod.rs:1200-1254) turns a SwResult into the record's CIGAR
// (soft clips added), reference span (window start added) and its NM / MD
// tags, replaying the CIGAR over the read and window with the MdBuilder
// rules of src/alignment/md.rs:25-204.  Host code only (no device work).
#include <emmintrin.h>
#include <string.h>

#include <string>

#include "internal.hpp"

namespace {

constexpr uint32_t kOpM = 0, kOpI = 1, kOpD = 2, kOpN = 3, kOpS = 4;

uint8_t upper(uint8_t b) { return (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b; }

// MD text and edit distance of one alignment (md.rs:25-132).
// decimal digits of v appended to s
inline void append_u32(std::string &s, uint32_t v) {
    char buf[10];
    int n = 0;
    do {
        buf[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) s.push_back(buf[--n]);
}

// MD text appended straight to the caller's string
struct MdText {
    std::string &text;
    uint32_t run = 0, nm = 0;
    explicit MdText(std::string &out) : text(out) {}
    void flush() {
        append_u32(text, run);
        run = 0;
    }
    // letters keep their (upper-cased) identity, other bytes become N
    static char letter(uint8_t b) {
        uint8_t u = upper(b);
        return (u >= 'A' && u <= 'Z') ? (char)u : 'N';
    }
    static bool edit_when_equal(uint8_t q) {
        // an identical ambiguity code still counts as an edit
        const uint8_t u = upper(q);
        return !(u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == '=');
    }
    void column(uint8_t q, uint8_t r) {
        if (q == r || q == '=' || upper(q) == upper(r)) {
            run++;
            nm += edit_when_equal(q) ? 1u : 0u;
        } else {
            flush();
            text += letter(r);
            nm++;
        }
    }
    // n aligned columns of bytes inside both sequences.  Eight equal bytes
    // that are all upper-case A / C / G / T extend the match run at once
    // (column() would count each as a match, never as an edit).
    void columns(const uint8_t *q, const uint8_t *r, size_t n) {
        size_t t = 0;
        // sixteen at a time: equal bytes that are each one of A, C, G, T
        const __m128i cA = _mm_set1_epi8('A'), cC = _mm_set1_epi8('C'),
                      cG = _mm_set1_epi8('G'), cT = _mm_set1_epi8('T');
        for (; t + 16 <= n; t += 16) {
            const __m128i a = _mm_loadu_si128((const __m128i *)(q + t));
            const __m128i b = _mm_loadu_si128((const __m128i *)(r + t));
            const __m128i acgt = _mm_or_si128(
                _mm_or_si128(_mm_cmpeq_epi8(a, cA), _mm_cmpeq_epi8(a, cC)),
                _mm_or_si128(_mm_cmpeq_epi8(a, cG), _mm_cmpeq_epi8(a, cT)));
            // bit u: byte u is a plain match; runs of them extend the match
            // run at once, the other bytes take column()
            const unsigned m =
                (unsigned)_mm_movemask_epi8(_mm_and_si128(_mm_cmpeq_epi8(a, b), acgt));
            if (m == 0xFFFFu) {
                run += 16;
                continue;
            }
            for (unsigned u = 0; u < 16;) {
                const unsigned rest = m >> u;
                if (rest & 1u) {
                    const unsigned k = (unsigned)__builtin_ctz(~rest);  // <= 16 - u
                    run += k;
                    u += k;
                } else {
                    column(q[t + u], r[t + u]);
                    u++;
                }
            }
        }
        for (; t + 8 <= n; t += 8) {
            uint64_t a, b;
            memcpy(&a, q + t, 8);
            memcpy(&b, r + t, 8);
            if (a == b && acgt_bytes(a) == 0x8080808080808080ull) {
                run += 8;
                continue;
            }
            for (size_t u = t; u < t + 8; u++) column(q[u], r[u]);
        }
        for (; t < n; t++) column(q[t], r[t]);
    }
    // 0x80 in each byte of x that is 'A', 'C', 'G' or 'T', 0 in the others
    static uint64_t acgt_bytes(uint64_t x) {
        auto eq = [](uint64_t v, uint64_t c) {
            const uint64_t y = v ^ (0x0101010101010101ull * c);
            const uint64_t t = ((y & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | y;
            return ~t & 0x8080808080808080ull;
        };
        return eq(x, 'A') | eq(x, 'C') | eq(x, 'G') | eq(x, 'T');
    }
    void deletion(const uint8_t *ref, size_t n) {
        if (n == 0) return;
        flush();
        text += '^';
        for (size_t k = 0; k < n; k++) text += letter(ref[k]);
        nm += (uint32_t)n;
    }
};

// nm_md_from_cigar (md.rs:162-204); positions past either sequence read as N.
void replay(const uint8_t *read, size_t read_len, const uint8_t *ref, size_t ref_len,
            size_t ref_pos, const uint32_t *cig, int32_t n, MdText &md) {
    size_t q = 0, r = ref_pos;
    for (int32_t k = 0; k < n; k++) {
        const size_t len = cig[k] >> 4;
        switch (cig[k] & 15u) {
        case kOpM:
            if (q + len <= read_len && r + len <= ref_len) {
                md.columns(read + q, ref + r, len);
                q += len;
                r += len;
                break;
            }
            for (size_t t = 0; t < len; t++, q++, r++)
                md.column(q < read_len ? read[q] : (uint8_t)'N', r < ref_len ? ref[r] : (uint8_t)'N');
            break;
        case kOpI:
            md.nm += (uint32_t)len;
            q += len;
            break;
        case kOpD: {
            const size_t end = std::min(r + len, ref_len), start = std::min(r, end);
            md.deletion(ref + start, end - start);
            md.nm += (uint32_t)(len - (end - start));  // bases past the window end
            r += len;
            break;
        }
        case kOpS: q += len; break;
        case kOpN: r += len; break;
        default: break;
        }
    }
    md.flush();
}

}  // namespace

namespace kira {

uint32_t build_record(const uint8_t *read, int32_t L, const uint8_t *win, int64_t W,
                      const kira_sw_result &s, const uint32_t *ops, std::vector<uint32_t> &cig,
                      std::string &md_out) {
    const size_t c0 = cig.size();
    if (s.read_start > 0) cig.push_back(((uint32_t)s.read_start << 4) | kOpS);
    cig.insert(cig.end(), ops, ops + s.n_cigar);
    if (s.read_end < L) cig.push_back(((uint32_t)(L - s.read_end) << 4) | kOpS);
    MdText md(md_out);
    // the CIGAR carries the leading clip, so the replay starts at read
    // position 0 (mod.rs:1231-1233)
    replay(read, (size_t)L, win, (size_t)W, s.ref_start, cig.data() + c0,
           (int32_t)(cig.size() - c0), md);
    return md.nm;
}

uint32_t nm_md(const uint8_t *read, int32_t L, const uint8_t *ref, int64_t ref_len, int64_t ref_pos,
               const uint32_t *ops, int32_t n, std::string &md_out) {
    MdText md(md_out);
    replay(read, (size_t)L, ref, (size_t)ref_len, (size_t)ref_pos, ops, n, md);
    return md.nm;
}

bool left_normalize(std::vector<uint32_t> &ops, const uint8_t *ref, int64_t ref_len,
                    uint32_t ref_start, const uint8_t *read, int64_t read_len) {
    auto op_of = [](uint32_t w) { return w & 15u; };
    auto len_of = [](uint32_t w) { return w >> 4; };
    auto make = [](uint32_t len, uint32_t op) { return (len << 4) | op; };
    bool any = false;
    for (uint32_t w : ops)
        if (op_of(w) == kOpI || op_of(w) == kOpD) any = true;
    if (!any) return false;
    // bytes past either sequence compare as absent (Option::None)
    auto rb = [&](int64_t i) -> int { return i >= 0 && i < ref_len ? ref[i] : -1; };
    auto qb = [&](int64_t i) -> int { return i >= 0 && i < read_len ? read[i] : -1; };
    auto acgt = [](int b) { return b == 'A' || b == 'C' || b == 'G' || b == 'T'; };
    const int64_t rs = ref_start;
    bool changed = false;
    size_t i = 0;
    while (i < ops.size()) {
        const uint32_t kind = op_of(ops[i]);
        if (kind != kOpI && kind != kOpD) {
            i++;
            continue;
        }
        const int64_t g = len_of(ops[i]);
        // read / reference position of this op's start
        int64_t bq = 0, br = 0;
        for (size_t k = 0; k < i; k++) {
            const uint32_t o = op_of(ops[k]), l = len_of(ops[k]);
            if (o == kOpM || o == kOpI || o == kOpS) bq += l;
            if (o == kOpM || o == kOpD || o == kOpN) br += l;
        }
        // left-shift budget: matching aligned bases right before the block
        const int64_t m_run = (i > 0 && op_of(ops[i - 1]) == kOpM) ? len_of(ops[i - 1]) : 0;
        int64_t shift_max = 0;
        while (shift_max < m_run && br - shift_max >= 1 && bq - shift_max >= 1) {
            const int r = rb(rs + br - shift_max - 1);
            if (!acgt(r < 0 ? 'N' : r) || r != qb(bq - shift_max - 1)) break;
            shift_max++;
        }
        // one base per step: the base entering the block on the left must
        // equal the one leaving it on the right
        int64_t shift = 0;
        while (shift < shift_max) {
            const int before = rb(rs + br - shift - 1);
            const int after = kind == kOpD ? rb(rs + br + g - shift - 1) : qb(bq + g - shift - 1);
            if (before != after) break;
            shift++;
        }
        if (shift == 0) {
            i++;
            continue;
        }
        changed = true;
        if (shift == m_run) {
            // the gap crosses the whole M block: swap, merge with a same-kind
            // gap now on its left, and look at the merged gap again
            std::swap(ops[i - 1], ops[i]);
            i--;
            if (i > 0 && op_of(ops[i - 1]) == kind) {
                ops[i - 1] = make(len_of(ops[i - 1]) + len_of(ops[i]), kind);
                ops.erase(ops.begin() + (ptrdiff_t)i);
                i--;
            }
            continue;
        }
        ops[i - 1] = make(len_of(ops[i - 1]) - (uint32_t)shift, kOpM);
        ops.insert(ops.begin() + (ptrdiff_t)i + 1, make((uint32_t)shift, kOpM));
        i++;
    }
    if (!changed) return false;
    std::vector<uint32_t> compact;
    for (uint32_t w : ops) {
        if (!compact.empty() && op_of(compact.back()) == op_of(w))
            compact.back() = make(len_of(compact.back()) + len_of(w), op_of(w));
        else
            compact.push_back(w);
    }
    ops.swap(compact);
    return true;
}

uint32_t ungapped_tags(const uint8_t *read, const uint8_t *ref, int64_t n, std::string &md_out) {
    MdText md(md_out);
    if (n > 0) md.columns(read, ref, (size_t)n);
    md.flush();
    return md.nm;
}

}  // namespace kira

extern "C" kira_status kira_build_alignments(
    int32_t n, const uint8_t *reads, const int64_t *read_off, const int32_t *read_len,
    const uint8_t *windows, const int64_t *win_off, const int32_t *win_len,
    const uint32_t *win_start, const kira_sw_result *sw, const uint32_t *sw_cigar,
    kira_alignment *out, uint32_t *cigar_out, int64_t cigar_cap, char *md_out, int64_t md_cap,
    int64_t *cigar_used, int64_t *md_used) {
    if (n < 0 || (n > 0 && (!reads || !read_off || !read_len || !windows || !win_off ||
                            !win_len || !win_start || !sw || !sw_cigar)))
        return KIRA_ERR_INVALID_ARGUMENT;
    int64_t cpos = 0, mpos = 0;
    bool fits = true;
    std::vector<uint32_t> cig;
    std::string mdt;
    for (int32_t k = 0; k < n; k++) {
        const kira_sw_result &s = sw[k];
        const int32_t L = read_len[k];
        if (s.n_cigar < 0 || s.read_start < 0 || s.read_end > L || s.read_start > s.read_end)
            return KIRA_ERR_INVALID_ARGUMENT;
        cig.clear();
        mdt.clear();
        const uint32_t nm = kira::build_record(reads + read_off[k], L, windows + win_off[k],
                                               win_len[k], s, sw_cigar + s.cigar_off, cig, mdt);
        const int64_t nc = (int64_t)cig.size(), nmd = (int64_t)mdt.size();
        if (out) {
            kira_alignment &a = out[k];
            a.ref_start = win_start[k] + s.ref_start;
            a.ref_end = win_start[k] + s.ref_end;
            a.read_start = (uint32_t)s.read_start;
            a.read_end = (uint32_t)s.read_end;
            a.score = s.score;
            a.nm = nm;
            a.n_cigar = (int32_t)nc;
            a.md_len = (int32_t)nmd;
            a.cigar_off = cpos;
            a.md_off = mpos;
        }
        if (cigar_out && cpos + nc <= cigar_cap)
            memcpy(cigar_out + cpos, cig.data(), (size_t)nc * 4);
        else if (cigar_out)
            fits = false;
        if (md_out && mpos + nmd <= md_cap)
            memcpy(md_out + mpos, mdt.data(), (size_t)nmd);
        else if (md_out)
            fits = false;
        cpos += nc;
        mpos += nmd;
    }
    if (cigar_used) *cigar_used = cpos;
    if (md_used) *md_used = mpos;
    return fits ? KIRA_OK : KIRA_ERR_INVALID_ARGUMENT;
}

extern "C" int32_t kira_left_normalize_indels(const uint32_t *cigar, int32_t n_ops,
                                              const uint8_t *contig, int64_t contig_len,
                                              uint32_t ref_start, const uint8_t *read,
                                              int32_t read_len, uint32_t *out, int32_t out_cap,
                                              int32_t *out_n) {
    if (n_ops < 0 || (n_ops > 0 && !cigar) || contig_len < 0 || (contig_len > 0 && !contig) ||
        read_len < 0 || (read_len > 0 && !read) || !out_n)
        return -1;
    std::vector<uint32_t> ops(cigar, cigar + n_ops);
    if (!kira::left_normalize(ops, contig, contig_len, ref_start, read, read_len)) {
        *out_n = n_ops;
        return 0;
    }
    *out_n = (int32_t)ops.size();
    if ((int32_t)ops.size() > out_cap || !out) return -1;
    memcpy(out, ops.data(), ops.size() * 4);
    return 1;
}
