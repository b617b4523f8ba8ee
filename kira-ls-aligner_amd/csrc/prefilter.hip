// Ungapped prefilter (prefilter_chain, src/alignment/prefilter.rs:93-295) on
// the GPU: x-drop extension of the chain anchor in 32-byte blocks, span
// Hamming, score / identity and the accept gates.  One lane evaluates one
// chain; byte comparisons use 4-byte words (count_mismatches counts unequal
// bytes, kira-simd 0.1.0 through src/simd/mod.rs:56-58).
#include "internal.hpp"

namespace kira {
namespace {

// unequal bytes among 4 (SWAR): a byte of x is nonzero iff its high bit ends
// up set in ((x & 0x7f..) + 0x7f..) | x.
__device__ __forceinline__ uint32_t ne_bytes(uint32_t x) {
    uint32_t y = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
    return __popc(y & 0x80808080u);
}

__device__ __forceinline__ uint32_t load4(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

__device__ int32_t count_mismatches(const uint8_t *a, const uint8_t *b, int64_t n) {
    int32_t m = 0;
    int64_t k = 0;
    for (; k + 4 <= n; k += 4) m += (int32_t)ne_bytes(load4(a + k) ^ load4(b + k));
    for (; k < n; k++) m += a[k] != b[k];
    return m;
}

__device__ __forceinline__ bool md_match(uint8_t q, uint8_t r) {
    uint8_t uq = (q >= 'a' && q <= 'z') ? (uint8_t)(q - 32) : q;
    uint8_t ur = (r >= 'a' && r <= 'z') ? (uint8_t)(r - 32) : r;
    return q == '=' || uq == ur;
}
__device__ __forceinline__ bool identical_is_edit(uint8_t b) {
    uint8_t u = (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b;
    return !(u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == '=');
}

__device__ uint16_t identity_x10000(int64_t len, int64_t mism) {
    if (len == 0) return 0;
    uint32_t matches = (uint32_t)(len > mism ? len - mism : 0);
    uint64_t prod = (uint64_t)matches * 10000u;
    uint32_t id = prod > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)prod;
    id /= (uint32_t)len;
    return (uint16_t)(id > 65535u ? 65535u : id);
}

__device__ int64_t min_len_required(int64_t read_len, uint16_t id, int64_t mism, int64_t slack0) {
    int64_t slack = 0;
    if (id >= 9933) slack = 15;
    else if (id >= 9900 && mism <= 2) slack = 10;
    else if (id >= 9870 && mism <= 3) slack = 5;
    if (slack0 > 0 && slack0 < slack) slack = slack0;
    int64_t min_len = read_len > slack ? read_len - slack : 0;
    int64_t fl = (read_len * 85) / 100;
    return fl > min_len ? fl : min_len;
}

// ungapped_extend — prefilter.rs:357-453
__device__ void ungapped_extend(const uint8_t *read, int64_t read_len, const uint8_t *ref,
                                int64_t ref_len, const kira_span &c, const PrefilterLaunch &a,
                                int64_t &left, int64_t &right) {
    left = 0;
    right = 0;
    const int64_t max_left = c.read_start < c.ref_start ? c.read_start : c.ref_start;
    const uint32_t ma = (uint32_t)read_len - c.read_end, mb = (uint32_t)ref_len - c.ref_end;
    const int64_t max_right = ma < mb ? ma : mb;
    const int64_t block = 32;
    int32_t score = 0, best = 0;
    int64_t i = 0;
    bool broke = false;
    while (i + block <= max_left) {
        int64_t ro = c.read_start - i - block, fo = c.ref_start - i - block;
        int32_t mism = count_mismatches(read + ro, ref + fo, block);
        score += ((int32_t)block - mism) * a.match_score - mism * a.mismatch;
        if (score > best) {
            best = score;
            left = i + block;
        } else if (best - score > a.xdrop) {
            broke = true;
            break;
        }
        i += block;
    }
    if (a.p.tail_ext && !broke && i < max_left) {
        int64_t rem = max_left - i;
        int64_t ro = c.read_start - i - rem, fo = c.ref_start - i - rem;
        int32_t mism = count_mismatches(read + ro, ref + fo, rem);
        score += ((int32_t)rem - mism) * a.match_score - mism * a.mismatch;
        if (score > best) left = i + rem;
    }
    score = 0;
    best = 0;
    i = 0;
    broke = false;
    while (i + block <= max_right) {
        int64_t ro = c.read_end + i, fo = c.ref_end + i;
        int32_t mism = count_mismatches(read + ro, ref + fo, block);
        score += ((int32_t)block - mism) * a.match_score - mism * a.mismatch;
        if (score > best) {
            best = score;
            right = i + block;
        } else if (best - score > a.xdrop) {
            broke = true;
            break;
        }
        i += block;
    }
    if (a.p.tail_ext && !broke && i < max_right) {
        int64_t rem = max_right - i;
        int64_t ro = c.read_end + i, fo = c.ref_end + i;
        int32_t mism = count_mismatches(read + ro, ref + fo, rem);
        score += ((int32_t)rem - mism) * a.match_score - mism * a.mismatch;
        if (score > best) right = i + rem;
    }
}

enum { R_ACCEPT = 0, R_REJECT = 1, R_FALLBACK = 2 };
enum { PR_ACCEPTED = 0, PR_DISABLED, PR_NOT_SHORT, PR_NOT_TOP1, PR_MAX_ALN, PR_SPAN_SMALL,
       PR_MISM_HIGH, PR_ID_LOW, PR_INDEL, PR_OTHER };
#define PR_OTHER_(k) (PR_OTHER + 16 * (k))

__global__ void prefilter_kernel(PrefilterLaunch a) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const int64_t read_len = a.read_len[t], ref_len = a.ref_len[t];
    const uint8_t *read = a.reads + a.read_off[t];
    const uint8_t *ref = a.refs + a.ref_off[t];
    const kira_span c = a.spans[t];
    kira_prefilter_out o{};
    o.result = R_FALLBACK;
    do {
        if (read_len == 0) { o.reason = PR_OTHER_(1); break; }
        if (c.read_start > c.read_end || c.ref_start > c.ref_end) { o.reason = PR_OTHER_(2); break; }
        int64_t anchor = (int64_t)c.read_end - c.read_start;
        if (anchor == 0) { o.reason = PR_OTHER_(3); break; }
        int64_t le, re;
        ungapped_extend(read, read_len, ref, ref_len, c, a, le, re);
        int64_t total = anchor + le + re;
        if (total == 0) { o.reason = PR_OTHER_(4); break; }
        int64_t urs = (int64_t)(int32_t)c.read_start - (int64_t)(int32_t)le;
        int64_t ufs = (int64_t)(int32_t)c.ref_start - (int64_t)(int32_t)le;
        if (urs < 0 || ufs < 0) { o.reason = PR_OTHER_(5); break; }
        if (urs + total > read_len || ufs + total > ref_len) { o.reason = PR_OTHER_(6); break; }
        int32_t ham = count_mismatches(read + urs, ref + ufs, total);
        int32_t score = (int32_t)(total > ham ? total - ham : 0) * a.match_score - ham * a.mismatch;
        uint16_t len16 = total > 65535 ? 65535 : (uint16_t)total;
        int64_t hm = ham < total ? ham : total;
        uint16_t mism16 = hm > 65535 ? 65535 : (uint16_t)hm;
        uint16_t id = identity_x10000(total, ham);
        o.has_metrics = 1;
        o.len = len16;
        o.mism = mism16;
        o.matches = len16 > mism16 ? (uint16_t)(len16 - mism16) : 0;
        o.identity_x10000 = id;
        o.score = score;
        o.read_start = (uint64_t)urs;
        o.ref_start = (uint64_t)ufs;
        if (total < read_len / 2) { o.result = R_REJECT; o.reason = PR_SPAN_SMALL; break; }
        if (!a.p.accept_enable) { o.reason = PR_DISABLED; break; }
        if (a.p.multi_alignments_enabled) { o.reason = PR_MAX_ALN; break; }
        if (a.p.accept_only_top1 && !(a.is_top1 ? a.is_top1[t] : 1)) { o.reason = PR_NOT_TOP1; break; }
        if (read_len > a.p.long_read_threshold) { o.reason = PR_NOT_SHORT; break; }
        int32_t margin = a.score_margin ? a.score_margin[t] : INT32_MAX;
        if (a.p.accept_require_score_margin > 0 && margin < a.p.accept_require_score_margin) {
            o.reason = PR_OTHER_(7);
            break;
        }
        uint32_t rspan = c.ref_end >= c.ref_start ? c.ref_end - c.ref_start : 0;
        uint32_t qspan = c.read_end >= c.read_start ? c.read_end - c.read_start : 0;
        if (rspan != qspan) { o.reason = PR_INDEL; break; }
        int64_t hard = (read_len * 5 + 149) / 150;
        int64_t maxm = a.p.accept_max_mismatches < hard ? a.p.accept_max_mismatches : hard;
        if (ham > maxm) { o.reason = PR_MISM_HIGH; break; }
        if (id < a.id_floor) { o.reason = PR_ID_LOW; break; }
        if (total < min_len_required(read_len, id, ham, a.p.accept_span_slack)) {
            o.reason = PR_SPAN_SMALL;
            break;
        }
        // build_ungapped_alignment — prefilter.rs:455-517
        int64_t span = len16;
        int n = 0;
        if (urs > 0) o.cigar[n++] = ((uint32_t)urs << 4) | 4u;
        o.cigar[n++] = ((uint32_t)span << 4) | 0u;
        if (urs + span < read_len) o.cigar[n++] = ((uint32_t)(read_len - urs - span) << 4) | 4u;
        o.n_cigar = n;
        int32_t mm = 0;
        uint32_t nm = 0;
        for (int64_t k = 0; k < span; k++) {
            uint8_t q = read[urs + k], r = ref[ufs + k];
            mm += q != r;
            if (md_match(q, r)) nm += identical_is_edit(q) ? 1u : 0u;
            else nm++;
        }
        o.aln_nm = nm;
        o.aln_score = ((int32_t)span - mm) * a.match_score - mm * a.mismatch;
        o.aln_ref_start = (uint32_t)ufs;
        o.aln_ref_end = (uint32_t)(ufs + span);
        o.aln_read_start = (uint32_t)urs;
        o.aln_read_end = (uint32_t)(urs + span);
        o.result = R_ACCEPT;
        o.reason = PR_ACCEPTED;
    } while (0);
    a.out[t] = o;
}

}  // namespace

hipError_t launch_prefilter(const PrefilterLaunch &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    prefilter_kernel<<<(unsigned)((a.n + 255) / 256), 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace kira
