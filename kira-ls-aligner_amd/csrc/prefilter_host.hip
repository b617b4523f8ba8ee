// Host entry of the GPU ungapped prefilter (prefilter_chain, src/alignment/prefilter.rs:93-295).
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace {
kira_status fail(kira_ctx *ctx, kira_status st, const char *what, hipError_t e = hipSuccess) {
    if (ctx) {
        char buf[512];
        if (e != hipSuccess)
            snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
        else
            snprintf(buf, sizeof buf, "%s", what);
        ctx->err = buf;
    }
    return st;
}
}  // namespace

namespace kira {

uint16_t prefilter_id_floor(const kira_prefilter_params &p) {
    // (accept_min_identity * 100.0).round() as u16 (prefilter.rs:255), then
    // KIRA_UNGAP_MINID overrides it (:256)
    float f = roundf(p.accept_min_identity * 100.0f);
    uint16_t v = f != f ? 0 : (f <= 0 ? 0 : (f >= 65535.0f ? 65535 : (uint16_t)f));
    if (p.ungap_min_id >= 0) v = (uint16_t)p.ungap_min_id;
    return v;
}

kira_status prefilter_gap_check(kira_ctx *ctx, int32_t n, const uint8_t *d_reads,
                                const int64_t *read_off, const int32_t *read_len,
                                const uint8_t *d_refs, const int64_t *ref_off,
                                const int64_t *ref_len, const kira_span *spans,
                                const kira_sw_cfg &cfg, kira_prefilter_out *out) {
    constexpr int32_t kAccept = 0, kFallback = 2, kIndelSuspect = 8;
    WfaPen pen;
    if (!wfa_pen(cfg, pen)) return KIRA_OK;  // no WFA alignment exists: nothing is beaten
    std::vector<int32_t> sel, plen, tlen, ms, tbf;
    std::vector<int64_t> poff, toff;
    const int64_t band = std::max(cfg.bandwidth, 1);
    try {
        for (int32_t k = 0; k < n; k++) {
            const kira_prefilter_out &o = out[k];
            if (o.result != kAccept || o.mism == 0) continue;
            const int64_t L = read_len[k];
            const int64_t ws = (int64_t)spans[k].ref_start - (int64_t)spans[k].read_start;
            if (L == 0 || ws < 0 || ws >= ref_len[k]) continue;
            const int64_t tl = std::min<int64_t>(L + band, ref_len[k] - ws);
            if (tl < L) continue;
            const int64_t cost = (int64_t)o.mism * ((int64_t)cfg.match_score + cfg.mismatch);
            if (cost <= 0) continue;
            if (cost - 1 > 65535)
                return fail(ctx, KIRA_ERR_INVALID_ARGUMENT,
                            "gap check budget above 65535 (mismatches x (match + mismatch))");
            if (L + tl > 60000)
                return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "gap check pair above 60000 bases");
            sel.push_back(k);
            poff.push_back(read_off[k]);
            plen.push_back((int32_t)L);
            toff.push_back(ref_off[k] + ws);
            tlen.push_back((int32_t)tl);
            ms.push_back((int32_t)(cost - 1));
            tbf.push_back(0);
        }
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    const int32_t m = (int32_t)sel.size();
    if (m == 0) return KIRA_OK;
    hipStream_t s = ctx->stream;
    DevMem dm(ctx);
    int64_t *d_poff = nullptr, *d_toff = nullptr;
    int32_t *d_plen = nullptr, *d_tlen = nullptr, *d_ms = nullptr, *d_tbf = nullptr;
    hipError_t e;
    if ((e = dm.alloc(&d_poff, m)) || (e = dm.alloc(&d_toff, m)) || (e = dm.alloc(&d_plen, m)) ||
        (e = dm.alloc(&d_tlen, m)) || (e = dm.alloc(&d_ms, m)) || (e = dm.alloc(&d_tbf, m)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "gap check allocation", e);
    if ((e = hipMemcpyAsync(d_poff, poff.data(), 8 * (size_t)m, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_toff, toff.data(), 8 * (size_t)m, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_plen, plen.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_tlen, tlen.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_ms, ms.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_tbf, tbf.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "gap check upload", e);
    WfaDevJobs j;
    j.n = m;
    j.pat = d_reads;
    j.txt = d_refs;
    j.pat_off = d_poff;
    j.txt_off = d_toff;
    j.pat_len = d_plen;
    j.txt_len = d_tlen;
    j.max_score = d_ms;
    j.tbf = d_tbf;
    j.h_pat_len = plen.data();
    j.h_txt_len = tlen.data();
    j.h_max_score = ms.data();
    j.h_tbf = tbf.data();
    std::vector<int32_t> res;
    std::vector<uint32_t> cig;
    std::vector<int64_t> coff;
    // exact (unpruned) score-only WFA, as the reference probes (mod.rs:196-202)
    kira_status st = wfa_dev(ctx, j, pen, 0, false, res, cig, coff);
    if (st != KIRA_OK) return st;
    for (int32_t q = 0; q < m; q++) {
        if (!res[(size_t)q * kWfaRes]) continue;  // no cheaper gapped alignment
        kira_prefilter_out &o = out[sel[q]];
        o.result = kFallback;
        o.reason = kIndelSuspect;
        o.aln_ref_start = o.aln_ref_end = o.aln_read_start = o.aln_read_end = 0;
        o.aln_score = 0;
        o.aln_nm = 0;
        o.n_cigar = 0;
        o.cigar[0] = o.cigar[1] = o.cigar[2] = 0;
    }
    return KIRA_OK;
}

}  // namespace kira

extern "C" kira_status kira_prefilter_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads,
                                            int64_t reads_bytes, const int64_t *read_off,
                                            const int32_t *read_len, const uint8_t *refs,
                                            int64_t refs_bytes, const int64_t *ref_off,
                                            const int64_t *ref_len, const kira_span *spans,
                                            const uint8_t *is_top1, const int32_t *score_margin,
                                            const kira_sw_cfg *cfg,
                                            const kira_prefilter_params *prm,
                                            kira_prefilter_out *out) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (n == 0) return KIRA_OK;
    if (n < 0 || !reads || !read_off || !read_len || !refs || !ref_off || !ref_len || !spans ||
        !cfg || !prm || !out)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    for (int32_t k = 0; k < n; k++) {
        const kira_span &c = spans[k];
        if (read_off[k] < 0 || read_len[k] < 0 || read_off[k] + read_len[k] > reads_bytes ||
            ref_off[k] < 0 || ref_len[k] < 0 || ref_off[k] + ref_len[k] > refs_bytes ||
            (int64_t)c.read_end > read_len[k] || (int64_t)c.ref_end > ref_len[k] ||
            (int64_t)c.read_start > read_len[k] || (int64_t)c.ref_start > ref_len[k])
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "chain span outside its read or contig");
    }
    const uint16_t floor_ = kira::prefilter_id_floor(*prm);
    uint8_t *d_reads = nullptr, *d_refs = nullptr, *d_top = nullptr;
    int64_t *d_ro = nullptr, *d_fo = nullptr, *d_fl = nullptr;
    int32_t *d_rl = nullptr, *d_margin = nullptr;
    kira_span *d_spans = nullptr;
    kira_prefilter_out *d_out = nullptr;
    std::vector<void *> owned;
    struct Free {
        std::vector<void *> &v;
        ~Free() {
            for (void *x : v) (void)hipFree(x);
        }
    } guard{owned};
    auto al = [&](auto **p, size_t bytes) {
        hipError_t e = hipMalloc((void **)p, std::max<size_t>(bytes, 1));
        if (e == hipSuccess) owned.push_back((void *)*p);
        return e;
    };
    hipError_t e;
    if ((e = al(&d_reads, reads_bytes)) || (e = al(&d_refs, refs_bytes)) ||
        (e = al(&d_ro, n * 8)) || (e = al(&d_fo, n * 8)) || (e = al(&d_fl, n * 8)) ||
        (e = al(&d_rl, n * 4)) || (e = al(&d_spans, n * sizeof(kira_span))) ||
        (e = al(&d_out, n * sizeof(kira_prefilter_out))) ||
        (is_top1 && (e = al(&d_top, n))) || (score_margin && (e = al(&d_margin, n * 4))))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "prefilter allocation", e);
    hipStream_t s = ctx->stream;
    auto up = [&](void *d, const void *h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    if ((e = up(d_reads, reads, reads_bytes)) || (e = up(d_refs, refs, refs_bytes)) ||
        (e = up(d_ro, read_off, n * 8)) || (e = up(d_fo, ref_off, n * 8)) ||
        (e = up(d_fl, ref_len, n * 8)) || (e = up(d_rl, read_len, n * 4)) ||
        (e = up(d_spans, spans, n * sizeof(kira_span))) || (d_top && (e = up(d_top, is_top1, n))) ||
        (d_margin && (e = up(d_margin, score_margin, n * 4))))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    kira::PrefilterLaunch a{};
    a.n = n;
    a.reads = d_reads;
    a.read_off = d_ro;
    a.read_len = d_rl;
    a.refs = d_refs;
    a.ref_off = d_fo;
    a.ref_len = d_fl;
    a.spans = d_spans;
    a.is_top1 = d_top;
    a.score_margin = d_margin;
    a.match_score = cfg->match_score;
    a.mismatch = cfg->mismatch;
    a.xdrop = cfg->xdrop;
    a.id_floor = floor_;
    a.p = *prm;
    a.out = d_out;
    if ((e = kira::launch_prefilter(a, s))) return fail(ctx, KIRA_ERR_RUNTIME, "prefilter launch", e);
    if ((e = hipMemcpyAsync(out, d_out, n * sizeof(kira_prefilter_out), hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
    if (prm->ungap_gap_check)
        return kira::prefilter_gap_check(ctx, n, d_reads, read_off, read_len, d_refs, ref_off,
                                         ref_len, spans, *cfg, out);
    return KIRA_OK;
}
