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

kira_status GroupRun::launch(const SwJobsHost &j, const int32_t *group_of,
                            const std::vector<int32_t> &group_neg, const kira_sw_cfg &cfg,
                            int slot) {
    const int32_t m = n = j.n;
    if (m == 0) return KIRA_OK;
    const int32_t ng = (int32_t)group_neg.size();
    // packets of 64 jobs ordered by (window, read) length, both descending
    std::vector<int32_t> order;
    {
        bool fits = true;
        std::vector<uint32_t> key(m);
        for (int32_t t = 0; t < m; t++) {
            fits &= j.ref_len[t] >= 0 && j.ref_len[t] <= 65535 && j.read_len[t] >= 0 &&
                    j.read_len[t] <= 65535;
            key[t] = ((uint32_t)(65535 - j.ref_len[t]) << 16) | (uint32_t)(65535 - j.read_len[t]);
        }
        if (fits) {
            stable_order_by_key(key.data(), m, order);
        } else {
            order.resize(m);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
                if (j.ref_len[x] != j.ref_len[y]) return j.ref_len[x] > j.ref_len[y];
                return j.read_len[x] > j.read_len[y];
            });
        }
    }
    const int32_t P = (m + 63) / 64;
    // the systolic DP takes windows of up to 64 strips (1024 columns)
    int32_t max_w = 0;
    for (int32_t t = 0; t < m; t++) max_w = std::max<int32_t>(max_w, j.ref_len[t]);
    const bool sys = max_w <= 64 * 16;
    // pinned block: per-job arrays, then per-packet tables, then the tasks
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t m8 = al((size_t)m * 8), m4 = al((size_t)m * 4), p8 = al((size_t)P * 8),
                 p4 = al((size_t)P * 4);
    const size_t o_ro = 0, o_fo = o_ro + m8, o_coff = o_fo + m8, o_rl = o_coff + m8,
                 o_fl = o_rl + m4, o_ab = o_fl + m4, o_gof = o_ab + m4, o_slot = o_gof + m4,
                 o_ccap = o_slot + m4, o_neg = o_ccap + m4,
                 o_pj = o_neg + al((size_t)std::max(ng, 1) * 4), o_ptoff = o_pj + P * 256,
                 o_proff = o_ptoff + p8, o_prows = o_proff + p8, o_pwords = o_prows + p4,
                 o_tasks = o_pwords + p4, total = o_tasks + (sys ? (size_t)P * 64 * 8 : 0);
    void *hp = nullptr;
    hipError_t e;
    if ((e = pinned_buf(ctx, slot, total, &hp)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "lane-group staging", e);
    uint8_t *h = (uint8_t *)hp;
    int64_t *coff = (int64_t *)(h + o_coff), *ptoff = (int64_t *)(h + o_ptoff),
            *proff = (int64_t *)(h + o_proff);
    int32_t *slot_of = (int32_t *)(h + o_slot), *ccap = (int32_t *)(h + o_ccap),
            *pj = (int32_t *)(h + o_pj), *prows = (int32_t *)(h + o_prows),
            *pwords = (int32_t *)(h + o_pwords);
    int2 *tasks = (int2 *)(h + o_tasks);
    int32_t n_tasks = 0;
    int64_t toff = 0, roff = 0;
    for (int32_t pk = 0; pk < P; pk++) {
        const int32_t s0 = pk * 64, e0 = std::min(m, s0 + 64);
        int32_t rows = 0, maxw = 0;
        for (int32_t l = 0; l < 64; l++) {
            const int32_t t = s0 + l < e0 ? order[s0 + l] : -1;
            pj[(size_t)pk * 64 + l] = t;
            if (t >= 0) {
                slot_of[t] = pk * 64 + l;
                rows = std::max<int32_t>(rows, j.read_len[t]);
                maxw = std::max<int32_t>(maxw, j.ref_len[t]);
            }
        }
        const int32_t words = 2 * ((maxw + 15) / 16);
        prows[pk] = rows;
        pwords[pk] = words;
        ptoff[pk] = toff;
        proff[pk] = roff;
        if (sys) {
            // per slot (rows + G - 1) steps of `words` words; tasks of 64 / G jobs
            const int32_t G = std::max(words / 2, 1), per = 64 / G;
            toff += (int64_t)(rows + G - 1) * words * 64;
            for (int32_t l = 0; l < e0 - s0; l += per) tasks[n_tasks++] = make_int2(pk, l);
        } else {
            toff += (int64_t)rows * words * 64;
        }
        roff += (int64_t)rows * 64;
    }
    int64_t cw = 0;
    for (int32_t t = 0; t < m; t++) {
        ccap[t] = 3 * j.read_len[t] + 2;
        coff[t] = cw;
        cw += ccap[t];
    }
    cig_words = cw;
    memcpy(h + o_ro, j.read_off, (size_t)m * 8);
    memcpy(h + o_fo, j.ref_off, (size_t)m * 8);
    memcpy(h + o_rl, j.read_len, (size_t)m * 4);
    memcpy(h + o_fl, j.ref_len, (size_t)m * 4);
    memcpy(h + o_ab, j.abort_score, (size_t)m * 4);
    memcpy(h + o_gof, group_of, (size_t)m * 4);
    if (ng) memcpy(h + o_neg, group_neg.data(), (size_t)ng * 4);
    trace_mark(".grp_plan");
    uint8_t *d_blob = nullptr;
    int32_t *d_gstop, *d_counter;
    uint32_t *d_trace;
    int2 *d_bnd, *d_rstat;
    if ((e = dm.alloc(&d_blob, total)) || (e = dm.alloc(&d_gstop, std::max(ng, 1))) ||
        (e = dm.alloc(&d_counter, 1)) || (e = dm.alloc(&d_res, (size_t)m * kSwResWords)) ||
        (e = dm.alloc(&d_trace, (size_t)toff)) || (e = dm.alloc(&d_cig, (size_t)cw)) ||
        (e = dm.alloc(&d_bnd, sys ? 1 : (size_t)roff)) || (e = dm.alloc(&d_rstat, (size_t)roff)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "lane-group allocation", e);
    hipStream_t s = ctx->stream;
    if ((e = hipMemcpyAsync(d_blob, h, total, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    trace_mark(".grp_upload");
    d_cig_off = (const int64_t *)(d_blob + o_coff);
    GroupLaunch a{};
    a.match_score = cfg.match_score;
    a.mismatch = cfg.mismatch;
    a.go = sat_add(cfg.gap_open, cfg.gap_extend);
    a.gap_extend = cfg.gap_extend;
    a.clip_penalty = cfg.clip_penalty;
    a.reads = j.d_reads;
    a.read_off = (const int64_t *)(d_blob + o_ro);
    a.read_len = (const int32_t *)(d_blob + o_rl);
    a.refs = j.d_refs;
    a.ref_off = (const int64_t *)(d_blob + o_fo);
    a.ref_len = (const int32_t *)(d_blob + o_fl);
    a.abort_score = (const int32_t *)(d_blob + o_ab);
    a.group_of = (const int32_t *)(d_blob + o_gof);
    a.neg = (const int32_t *)(d_blob + o_neg);
    a.group_stop = d_gstop;
    a.packet_jobs = (const int32_t *)(d_blob + o_pj);
    a.packet_rows = (const int32_t *)(d_blob + o_prows);
    a.packet_words = (const int32_t *)(d_blob + o_pwords);
    a.packet_trace_off = (const int64_t *)(d_blob + o_ptoff);
    a.packet_row_off = (const int64_t *)(d_blob + o_proff);
    a.n_packets = P;
    a.counter = d_counter;
    a.trace = d_trace;
    a.bnd = d_bnd;
    a.rstat = d_rstat;
    a.job_slot = (const int32_t *)(d_blob + o_slot);
    a.res = d_res;
    a.cigar = d_cig;
    a.cigar_off = d_cig_off;
    a.cigar_cap = (const int32_t *)(d_blob + o_ccap);
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
    a.systolic = sys;
    a.tasks = (const int2 *)(d_blob + o_tasks);
    a.n_tasks = n_tasks;
    // persistent waves on the atomic task counter: the one-lane form at most
    // 16 a CU, the systolic form (≈90 VGPRs) at most 20 a CU
    const int n_waves = sys ? std::min(n_tasks, 256 * 20) : std::min(P, 256 * 16);
    if ((e = launch_group_sw(a, m, ng, n_waves, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "lane-group launch", e);
    trace_mark(".grp_launch");
    return KIRA_OK;
}

kira_status GroupRun::fetch(kira_sw_result *out, std::vector<uint32_t> &cigar) {
    const kira_status st = fetch_sw_results(ctx, ctx->stream, dm, n, d_res, d_cig, d_cig_off,
                                            cig_words, false, out, cigar);
    trace_mark(".grp_fetch");
    return st;
}

kira_status group_sw_dev(kira_ctx *ctx, const SwJobsHost &j, const int32_t *group_of,
                         const std::vector<int32_t> &group_neg, const kira_sw_cfg &cfg,
                         kira_sw_result *out, std::vector<uint32_t> &cigar) {
    GroupRun run(ctx);
    kira_status st = run.launch(j, group_of, group_neg, cfg, kPinGroup);
    if (st != KIRA_OK) return st;
    return run.fetch(out, cigar);
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
    kira::DevMem d(ctx);
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
