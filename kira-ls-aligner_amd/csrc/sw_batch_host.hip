// Host side of align_batch_simd (src/alignment/mod.rs:1111-1149): splits each
// batch exactly as the reference dispatch does into full-window lane groups
// and per-job banded SW, runs both on the device and returns SwResults.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
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

int32_t sat_add(int32_t a, int32_t b) {
    int64_t s = (int64_t)a + b;
    return s > INT32_MAX ? INT32_MAX : (s < INT32_MIN ? INT32_MIN : (int32_t)s);
}

// sw_batch_avx2's cost gate (mod.rs:1854-1868).
bool avx2_costs_fit(const kira_sw_cfg *c, int64_t q, int64_t r) {
    const int32_t mx = 32767;
    bool fit = c->match_score >= 0 && c->match_score <= mx && c->mismatch >= 0 &&
               c->mismatch <= mx && c->gap_open >= 0 && c->gap_extend >= 0 &&
               sat_add(c->gap_open, c->gap_extend) <= mx;
    return fit && q <= mx && r <= mx && std::min(q, r) * (int64_t)c->match_score <= mx;
}

// int8_path_viable (sw_int8_vnni.rs:47-55).
bool int8_viable(int64_t read_len, const kira_sw_cfg *c) {
    int64_t ms = std::max(c->match_score, 1);
    auto in = [](int32_t v) { return v >= 0 && v <= 32; };
    return read_len * ms < 120 && in(c->match_score) && in(c->mismatch) && in(c->gap_open) &&
           in(c->gap_extend);
}

struct DevBuf {
    std::vector<void *> p;
    ~DevBuf() {
        for (void *x : p) (void)hipFree(x);
    }
    template <class T>
    hipError_t alloc(T **out, size_t n) {
        hipError_t e = hipMalloc((void **)out, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(*out);
        return e;
    }
};

}  // namespace

namespace kira {

void simd_groups(int32_t n_batches, const int32_t *batch_off, const uint32_t *read_len,
                 const uint32_t *win_len, kira_simd_mode mode, kira_simd_mode detected,
                 const kira_sw_cfg &cfg, std::vector<int32_t> &group_of,
                 std::vector<int32_t> &group_neg) {
    const int32_t n = batch_off[n_batches];
    group_of.assign(n, -1);
    group_neg.clear();
    auto add_group = [&](int32_t s, int32_t e, int32_t neg) {
        int32_t g = (int32_t)group_neg.size();
        group_neg.push_back(neg);
        for (int32_t k = s; k < e; k++) group_of[k] = g;
    };
    const int32_t kNeg16 = -16384, kNeg8 = -100, kNeg32 = INT32_MIN / 4;
    for (int32_t b = 0; b < n_batches; b++) {
        int32_t s = batch_off[b], e = batch_off[b + 1];
        if (e == s) continue;
        uint32_t L = read_len[s], W = win_len[s];
        bool homo = true;
        for (int32_t k = s; k < e; k++)
            if (read_len[k] != L || win_len[k] != W) homo = false;
        if (!homo) continue;  // every job scalar
        auto avx2_chunks = [&]() {
            for (int32_t c = s; c < e; c += 16) {
                int32_t ce = std::min(e, c + 16);
                if (avx2_costs_fit(&cfg, L, W)) add_group(c, ce, kNeg16);
            }
        };
        if (mode == KIRA_SIMD_AVXVNNI && detected == KIRA_SIMD_AVXVNNI) {
            if (int8_viable(L, &cfg) && e - s <= 32)
                add_group(s, e, kNeg8);
            else
                avx2_chunks();
        } else if (mode == KIRA_SIMD_AVX2 &&
                   (detected == KIRA_SIMD_AVX2 || detected == KIRA_SIMD_AVXVNNI)) {
            avx2_chunks();
        } else if (mode == KIRA_SIMD_NEON && detected == KIRA_SIMD_NEON) {
            for (int32_t c = s; c < e; c += 4) add_group(c, std::min(e, c + 4), kNeg32);
        }
    }
}

kira_status group_sw_dev(kira_ctx *ctx, const SwJobsHost &j, const int32_t *group_of,
                         const std::vector<int32_t> &group_neg, const kira_sw_cfg &cfg,
                         kira_sw_result *out, std::vector<uint32_t> &cigar) {
    const int32_t m = j.n;
    if (m == 0) return KIRA_OK;
    const int32_t ng = (int32_t)group_neg.size();
    // packets of 64 jobs ordered by (window, read) length
    std::vector<int32_t> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (j.ref_len[x] != j.ref_len[y]) return j.ref_len[x] > j.ref_len[y];
        return j.read_len[x] > j.read_len[y];
    });
    std::vector<int32_t> pj, prows, pwords, slot(m);
    std::vector<int64_t> ptoff, proff;
    int64_t toff = 0, roff = 0;
    for (int32_t s = 0; s < m; s += 64) {
        int32_t e = std::min(m, s + 64);
        int32_t pk = (int32_t)prows.size();
        int32_t rows = 0, maxw = 0;
        for (int32_t l = 0; l < 64; l++) {
            int32_t t = s + l < e ? order[s + l] : -1;
            pj.push_back(t);
            if (t >= 0) {
                slot[t] = pk * 64 + l;
                rows = std::max<int32_t>(rows, j.read_len[t]);
                maxw = std::max<int32_t>(maxw, j.ref_len[t]);
            }
        }
        int32_t words = 2 * ((maxw + 15) / 16);
        prows.push_back(rows);
        pwords.push_back(words);
        ptoff.push_back(toff);
        proff.push_back(roff);
        toff += (int64_t)rows * words * 64;
        roff += (int64_t)rows * 64;
    }
    std::vector<int64_t> coff(m);
    std::vector<int32_t> ccap(m);
    int64_t cw = 0;
    for (int32_t t = 0; t < m; t++) {
        ccap[t] = 3 * j.read_len[t] + 2;
        coff[t] = cw;
        cw += ccap[t];
    }
    DevBuf d;
    int64_t *d_ro, *d_fo, *d_coff, *d_ptoff, *d_proff;
    int32_t *d_rl, *d_fl, *d_ab, *d_gof, *d_neg, *d_gstop, *d_pj, *d_prows, *d_pwords,
        *d_counter, *d_slot, *d_res, *d_ccap;
    uint32_t *d_trace, *d_cig;
    int2 *d_bnd, *d_rstat;
    hipError_t e;
    const int32_t P = (int32_t)prows.size();
    if ((e = d.alloc(&d_ro, m)) || (e = d.alloc(&d_fo, m)) || (e = d.alloc(&d_coff, m)) ||
        (e = d.alloc(&d_ptoff, P)) || (e = d.alloc(&d_proff, P)) || (e = d.alloc(&d_rl, m)) ||
        (e = d.alloc(&d_fl, m)) || (e = d.alloc(&d_ab, m)) || (e = d.alloc(&d_gof, m)) ||
        (e = d.alloc(&d_neg, ng)) || (e = d.alloc(&d_gstop, ng)) ||
        (e = d.alloc(&d_pj, (size_t)P * 64)) || (e = d.alloc(&d_prows, P)) ||
        (e = d.alloc(&d_pwords, P)) || (e = d.alloc(&d_counter, 1)) ||
        (e = d.alloc(&d_slot, m)) || (e = d.alloc(&d_res, (size_t)m * kSwResWords)) ||
        (e = d.alloc(&d_ccap, m)) || (e = d.alloc(&d_trace, (size_t)toff)) ||
        (e = d.alloc(&d_cig, (size_t)cw)) || (e = d.alloc(&d_bnd, (size_t)roff)) ||
        (e = d.alloc(&d_rstat, (size_t)roff)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "lane-group allocation", e);
    hipStream_t s = ctx->stream;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    if ((e = up(d_ro, j.read_off, m * 8)) || (e = up(d_fo, j.ref_off, m * 8)) ||
        (e = up(d_coff, coff.data(), m * 8)) || (e = up(d_ptoff, ptoff.data(), P * 8)) ||
        (e = up(d_proff, proff.data(), P * 8)) || (e = up(d_rl, j.read_len, m * 4)) ||
        (e = up(d_fl, j.ref_len, m * 4)) || (e = up(d_ab, j.abort_score, m * 4)) ||
        (e = up(d_gof, group_of, m * 4)) || (e = up(d_neg, group_neg.data(), ng * 4)) ||
        (e = up(d_pj, pj.data(), pj.size() * 4)) || (e = up(d_prows, prows.data(), P * 4)) ||
        (e = up(d_pwords, pwords.data(), P * 4)) || (e = up(d_slot, slot.data(), m * 4)) ||
        (e = up(d_ccap, ccap.data(), m * 4)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    GroupLaunch a{};
    a.match_score = cfg.match_score;
    a.mismatch = cfg.mismatch;
    a.go = sat_add(cfg.gap_open, cfg.gap_extend);
    a.gap_extend = cfg.gap_extend;
    a.clip_penalty = cfg.clip_penalty;
    a.reads = j.d_reads;
    a.read_off = d_ro;
    a.read_len = d_rl;
    a.refs = j.d_refs;
    a.ref_off = d_fo;
    a.ref_len = d_fl;
    a.abort_score = d_ab;
    a.group_of = d_gof;
    a.neg = d_neg;
    a.group_stop = d_gstop;
    a.packet_jobs = d_pj;
    a.packet_rows = d_prows;
    a.packet_words = d_pwords;
    a.packet_trace_off = d_ptoff;
    a.packet_row_off = d_proff;
    a.n_packets = P;
    a.counter = d_counter;
    a.trace = d_trace;
    a.bnd = d_bnd;
    a.rstat = d_rstat;
    a.job_slot = d_slot;
    a.res = d_res;
    a.cigar = d_cig;
    a.cigar_off = d_coff;
    a.cigar_cap = d_ccap;
    // the keyed maximum shifts scores left by 4: every H <= min(L, W) * match
    // and every candidate >= -max(mismatch, open + extend) must stay in 2^27
    {
        const int64_t kLim = int64_t(1) << 20;
        int64_t hmax = 0;
        for (int32_t t = 0; t < m; t++)
            hmax = std::max<int64_t>(hmax, (int64_t)std::min(j.read_len[t], j.ref_len[t]) *
                                               std::max(cfg.match_score, 0));
        a.keyed = cfg.match_score >= 0 && cfg.match_score <= kLim && cfg.mismatch >= 0 &&
                  cfg.mismatch <= kLim && cfg.gap_open >= 0 && cfg.gap_extend >= 0 &&
                  a.go <= kLim && hmax < (int64_t(1) << 27);
    }
    if ((e = launch_group_sw(a, m, ng, std::min(P, 256 * 16), s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "lane-group launch", e);
    return fetch_sw_results(ctx, m, d_res, d_cig, coff.data(), false, out, cigar);
}

}  // namespace kira

extern "C" kira_status kira_align_batch_simd(kira_ctx *ctx, int32_t n_batches,
                                             const int32_t *batch_off,
                                             const kira_batch_input *in, kira_simd_mode mode,
                                             kira_simd_mode detected, const kira_sw_cfg *cfg,
                                             const kira_sw_knobs *kn, kira_sw_result *out,
                                             uint32_t *cigar_arena, int64_t arena_cap) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (n_batches <= 0) return KIRA_OK;
    if (!batch_off || !cfg || !kn || batch_off[0] != 0)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    const int32_t n = batch_off[n_batches];
    if (n == 0) return KIRA_OK;
    if (n < 0 || !in || !out) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    for (int32_t b = 0; b < n_batches; b++)
        if (batch_off[b + 1] < batch_off[b])
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "batch offsets must not decrease");

    // Dispatch each batch (mod.rs:1118-1149).
    std::vector<uint32_t> rl(n), wl(n);
    for (int32_t k = 0; k < n; k++) {
        rl[k] = in[k].read_len;
        wl[k] = in[k].window_len;
    }
    std::vector<int32_t> group_of, group_neg;
    kira::simd_groups(n_batches, batch_off, rl.data(), wl.data(), mode, detected, *cfg, group_of,
                      group_neg);

    // every read and window once on the device (windows between kRefPad zero bytes)
    std::vector<int64_t> ro(n), fo(n);
    int64_t rb = 0, fb = kira::kRefPad;
    for (int32_t k = 0; k < n; k++) {
        ro[k] = rb;
        fo[k] = fb;
        rb += rl[k];
        fb += wl[k];
    }
    std::vector<uint8_t> reads((size_t)rb), refs((size_t)fb + kira::kRefPad, 0);
    for (int32_t k = 0; k < n; k++) {
        if (rl[k]) memcpy(reads.data() + ro[k], in[k].read, rl[k]);
        if (wl[k]) memcpy(refs.data() + fo[k], in[k].window, wl[k]);
    }
    DevBuf d;
    uint8_t *d_reads, *d_refs;
    hipError_t e;
    hipStream_t s = ctx->stream;
    if ((e = d.alloc(&d_reads, reads.size() + 8)) || (e = d.alloc(&d_refs, refs.size())))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "batch allocation", e);
    if ((!reads.empty() &&
         (e = hipMemcpyAsync(d_reads, reads.data(), reads.size(), hipMemcpyHostToDevice, s))) ||
        (e = hipMemcpyAsync(d_refs, refs.data(), refs.size(), hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);

    std::vector<uint32_t> cig;
    std::vector<kira_sw_result> res(n);
    for (int pass = 0; pass < 2; pass++) {
        // pass 0: banded SW jobs (BatchInput::scalar_sw, mod.rs:60-72); pass 1: lane groups
        std::vector<int32_t> sel;
        for (int32_t k = 0; k < n; k++)
            if ((group_of[k] >= 0) == (pass == 1)) sel.push_back(k);
        const int32_t m = (int32_t)sel.size();
        if (m == 0) continue;
        std::vector<int64_t> sro(m), sfo(m);
        std::vector<int32_t> srl(m), sfl(m), off(m), ab(m), gof(m);
        for (int32_t t = 0; t < m; t++) {
            const int32_t k = sel[t];
            const kira_batch_input &x = in[k];
            sro[t] = ro[k];
            sfo[t] = fo[k];
            srl[t] = (int32_t)rl[k];
            sfl[t] = (int32_t)wl[k];
            int64_t o = (int64_t)x.chain_ref_start - (int64_t)x.win_start -
                        (int64_t)x.chain_read_start;
            o = std::max<int64_t>(o, -(int64_t)x.read_len);
            o = std::min<int64_t>(o, (int64_t)x.window_len);
            off[t] = (int32_t)o;
            ab[t] = x.abort_score;
            gof[t] = group_of[k];
        }
        kira::SwJobsHost j{m,          d_reads, sro.data(), srl.data(), d_refs,
                           sfo.data(), sfl.data(), off.data(), nullptr, ab.data()};
        std::vector<kira_sw_result> r(m);
        kira_status st = pass == 0
                             ? kira::banded_sw_dev(ctx, j, *cfg, *kn, r.data(), cig)
                             : kira::group_sw_dev(ctx, j, gof.data(), group_neg, *cfg, r.data(), cig);
        if (st != KIRA_OK) return st;
        for (int32_t t = 0; t < m; t++) res[sel[t]] = r[t];
    }
    int64_t pos = 0;
    for (int32_t k = 0; k < n; k++) {
        out[k] = res[k];
        const int32_t nc = res[k].n_cigar;
        if (pos + nc > arena_cap || (nc > 0 && !cigar_arena))
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "cigar arena too small");
        out[k].cigar_off = pos;
        if (nc) memcpy(cigar_arena + pos, cig.data() + res[k].cigar_off, (size_t)nc * 4);
        pos += nc;
    }
    return KIRA_OK;
}
