// Host side of the banded SW entry points: job packetisation (the device
// counterpart of the reference's per-job rayon loop, stage4_alignment.rs:1196-1205),
// trace slab sizing, and the host-array batch entry banded_sw_public
// (src/alignment/mod.rs:891-920).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "internal.hpp"

namespace {

// Job classes.  Lane-per-job: a job with band w runs in the smallest
// C >= 2w + 1 of kLaneC, with the band stored as packed int16 when its
// scores provably fit, else as int32 (class = band index * 2 + packed).
// Bands above 64 (C = 257 would spill the register file) run 16 lanes per
// job with P = kCoopP[k] positions a lane (16 P >= 2w + 1) when the costs
// keep its int32 scan and arg-max key exact; otherwise C = 257 lane-per-job.
constexpr int kLaneC[] = {9, 17, 33, 65, 129, 257};
constexpr int kNumLane = 6;
constexpr int kCoopP[] = {9, 17, 25, 33};
constexpr int kNumCoop = 4;
constexpr int kNumClasses = 2 * kNumLane + kNumCoop;
constexpr int kMaxBand = (16 * 33 - 1) / 2;  // 263

bool class_is_coop(int c) { return c >= 2 * kNumLane; }
int class_coop_p(int c) { return kCoopP[c - 2 * kNumLane]; }
int class_jobs_per_packet(int c) { return class_is_coop(c) ? 4 : 64; }
int class_rowdw(int c) {
    return class_is_coop(c) ? (class_coop_p(c) + 7) / 8 : (kLaneC[c / 2] + 7) / 8;
}

int lane_class_of(int32_t w) {
    int need = 2 * w + 1;
    for (int k = 0; k < kNumLane; k++)
        if (kLaneC[k] >= need) return k;
    return -1;
}

int coop_class_of(int32_t w) {
    for (int k = 0; k < kNumCoop; k++)
        if (16 * kCoopP[k] >= 2 * w + 1) return 2 * kNumLane + k;
    return -1;
}

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

template <class T>
hipError_t dev_alloc(T **p, size_t count) {
    return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}

int32_t sat_add(int32_t a, int32_t b) {
    int64_t s = (int64_t)a + b;
    return s > INT32_MAX ? INT32_MAX : (s < INT32_MIN ? INT32_MIN : (int32_t)s);
}

}  // namespace

struct kira_sw_plan {
    int32_t n = 0;
    int32_t n_packets = 0;
    int32_t class_first[kNumClasses] = {};
    int32_t class_count[kNumClasses] = {};
    int32_t *packet_jobs = nullptr, *packet_rows = nullptr, *packet_rowdw = nullptr;
    int64_t *packet_trace_off = nullptr, *packet_meta_off = nullptr;
    int32_t *job_slot = nullptr, *cigar_cap = nullptr, *counter = nullptr;
    int64_t *cigar_off = nullptr;
    uint32_t *trace = nullptr;
    int2 *meta = nullptr;
    int64_t cigar_words = 0;
    int64_t device_bytes = 0;
    std::vector<int64_t> host_cigar_off;
    void release() {
        void *ptrs[] = {packet_jobs, packet_rows, packet_rowdw, packet_trace_off,
                        packet_meta_off, job_slot, cigar_cap, counter, cigar_off, trace, meta};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

extern "C" {

kira_status kira_banded_sw_plan_create(kira_ctx *ctx, int32_t n, const int32_t *read_len,
                                       const int32_t *band, const kira_sw_cfg *cfg,
                                       kira_sw_plan **out) {
    if (!ctx || !out || !cfg || n < 0 || (n > 0 && !read_len))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    *out = nullptr;
    std::vector<int> cls(n);
    const int32_t go = sat_add(cfg->gap_open, cfg->gap_extend);
    const bool costs_small = cfg->match_score >= 0 && cfg->mismatch >= 0 && cfg->gap_open >= 0 &&
                             cfg->gap_extend >= 0 && go <= 16000;
    // 16-lane jobs: the F scan stays inside int32 and the arg-max key
    // (score << 10 | position) needs score < 2^21
    const int32_t kCoopCost = 1 << 20;
    const bool coop_costs = cfg->match_score >= 0 && cfg->mismatch >= 0 && cfg->gap_open >= 0 &&
                            cfg->gap_extend >= 0 && cfg->match_score <= kCoopCost &&
                            cfg->mismatch <= kCoopCost && go <= kCoopCost;
    for (int32_t k = 0; k < n; k++) {
        int32_t w = std::max(band ? band[k] : cfg->bandwidth, 1);
        if (read_len[k] < 0) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "negative read length");
        // every H is at most read_len * match: int16 storage cannot clamp
        bool pack = costs_small && (int64_t)read_len[k] * cfg->match_score < 32767;
        bool coop = coop_costs && (int64_t)read_len[k] * cfg->match_score < (1 << 21);
        if (w > 64 && coop) {
            cls[k] = coop_class_of(w);
            if (cls[k] < 0) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "band wider than 263");
            continue;
        }
        int b = lane_class_of(w);
        if (b < 0)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT,
                        w <= kMaxBand ? "band above 128 needs costs <= 2^20 and read_len * match < 2^21"
                                      : "band wider than 263");
        cls[k] = 2 * b + (pack ? 1 : 0);
    }
    // packets: jobs sorted by (band class, read length descending)
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (cls[x] != cls[y]) return cls[x] < cls[y];
        return read_len[x] > read_len[y];
    });
    kira_sw_plan *p = new (std::nothrow) kira_sw_plan();
    if (!p) return KIRA_ERR_OUT_OF_MEMORY;
    p->n = n;
    std::vector<int32_t> pj, prows, prowdw, slot(n);
    std::vector<int64_t> ptoff, pmoff;
    int64_t toff = 0, moff = 0;
    for (int c = 0; c < kNumClasses; c++) p->class_first[c] = -1;
    for (int32_t s = 0; s < n;) {
        int c = cls[order[s]];
        const int per = class_jobs_per_packet(c), stride = 64 / per;
        int32_t e = s;
        while (e < n && e - s < per && cls[order[e]] == c) e++;
        int32_t pk = (int32_t)prows.size();
        if (p->class_first[c] < 0) p->class_first[c] = pk;
        p->class_count[c]++;
        int32_t rows = read_len[order[s]];
        int32_t rowdw = class_rowdw(c);
        for (int32_t l = 0; l < 64; l++) {
            int32_t job = (l % stride == 0 && s + l / stride < e) ? order[s + l / stride] : -1;
            pj.push_back(job);
            if (job >= 0) slot[job] = pk * 64 + l;
        }
        prows.push_back(rows);
        prowdw.push_back(rowdw);
        ptoff.push_back(toff);
        pmoff.push_back(moff);
        toff += (int64_t)rows * rowdw * 64;
        moff += (int64_t)rows * 64;
        s = e;
    }
    p->n_packets = (int32_t)prows.size();
    std::vector<int64_t> coff(n);
    std::vector<int32_t> ccap(n);
    int64_t acc = 0;
    for (int32_t k = 0; k < n; k++) {
        ccap[k] = 3 * read_len[k] + 2;
        coff[k] = acc;
        acc += ccap[k];
    }
    p->cigar_words = acc;
    p->device_bytes = toff * 4 + moff * (int64_t)sizeof(int2) +
                      (int64_t)std::max(p->n_packets, 1) * (64 * 4 + 4 * 2 + 8 * 2) +
                      (int64_t)n * (4 + 4 + 8) + kNumClasses * 4;
    p->host_cigar_off = coff;
    hipError_t e = hipSetDevice(ctx->device);
    int32_t P = std::max(p->n_packets, 1);
    if (e || (e = dev_alloc(&p->packet_jobs, (size_t)P * 64)) ||
        (e = dev_alloc(&p->packet_rows, P)) || (e = dev_alloc(&p->packet_rowdw, P)) ||
        (e = dev_alloc(&p->packet_trace_off, P)) || (e = dev_alloc(&p->packet_meta_off, P)) ||
        (e = dev_alloc(&p->job_slot, n)) || (e = dev_alloc(&p->cigar_cap, n)) ||
        (e = dev_alloc(&p->cigar_off, n)) || (e = dev_alloc(&p->counter, kNumClasses)) ||
        (e = dev_alloc(&p->trace, (size_t)toff)) || (e = dev_alloc(&p->meta, (size_t)moff))) {
        p->release();
        delete p;
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "plan allocation", e);
    }
    hipStream_t s = ctx->stream;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    if ((e = up(p->packet_jobs, pj.data(), pj.size() * 4)) ||
        (e = up(p->packet_rows, prows.data(), prows.size() * 4)) ||
        (e = up(p->packet_rowdw, prowdw.data(), prowdw.size() * 4)) ||
        (e = up(p->packet_trace_off, ptoff.data(), ptoff.size() * 8)) ||
        (e = up(p->packet_meta_off, pmoff.data(), pmoff.size() * 8)) ||
        (e = up(p->job_slot, slot.data(), (size_t)n * 4)) ||
        (e = up(p->cigar_cap, ccap.data(), (size_t)n * 4)) ||
        (e = up(p->cigar_off, coff.data(), (size_t)n * 8)) || (e = hipStreamSynchronize(s))) {
        p->release();
        delete p;
        return fail(ctx, KIRA_ERR_RUNTIME, "plan upload", e);
    }
    *out = p;
    return KIRA_OK;
}

void kira_banded_sw_plan_destroy(kira_sw_plan *plan) {
    if (!plan) return;
    plan->release();
    delete plan;
}

int64_t kira_banded_sw_plan_cigar_words(const kira_sw_plan *plan) {
    return plan ? plan->cigar_words : 0;
}

int64_t kira_banded_sw_plan_device_bytes(const kira_sw_plan *plan) {
    return plan ? plan->device_bytes : 0;
}

const int64_t *kira_banded_sw_plan_cigar_off(const kira_sw_plan *plan) {
    return plan ? plan->host_cigar_off.data() : nullptr;
}

kira_status kira_banded_sw_run_dev(kira_ctx *ctx, kira_sw_plan *plan,
                                   const kira_sw_jobs_dev *j, const kira_sw_cfg *cfg,
                                   const kira_sw_knobs *kn) {
    if (!ctx || !plan || !j || !cfg || !kn)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (plan->n == 0) return KIRA_OK;
    int32_t go = sat_add(cfg->gap_open, cfg->gap_extend);
    // pads are INT32_MIN/4 + cost and must stay negative
    if (go >= (1 << 29) || cfg->gap_extend >= (1 << 29) || go < 0 || cfg->gap_extend < 0)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "gap costs outside the supported range");
    kira::SwLaunch a{};
    a.match_score = cfg->match_score;
    a.mismatch = cfg->mismatch;
    a.go = go;
    a.gap_extend = cfg->gap_extend;
    a.clip_penalty = cfg->clip_penalty;
    a.bandwidth = cfg->bandwidth;
    a.adaptive = kn->adaptive ? 1 : 0;
    a.max_drift = kn->max_drift;
    a.dp_xdrop = kn->dp_xdrop >= 0 ? kn->dp_xdrop : cfg->xdrop;
    a.reads = j->reads;
    a.read_off = j->read_off;
    a.read_len = j->read_len;
    a.refs = j->refs;
    a.ref_off = j->ref_off;
    a.ref_len = j->ref_len;
    a.offset = j->offset;
    a.band = j->band;
    a.abort_score = j->abort_score;
    a.trace = plan->trace;
    a.row_meta = plan->meta;
    a.res = j->res;
    a.job_slot = plan->job_slot;
    a.cigar = j->cigar;
    a.cigar_off = plan->cigar_off;
    a.cigar_cap = plan->cigar_cap;
    hipError_t e = hipSuccess;
    for (int c = 0; c < kNumClasses; c++) {
        if (plan->class_count[c] == 0) continue;
        kira::SwLaunch ac = a;
        int32_t f = plan->class_first[c];
        ac.packet_jobs = plan->packet_jobs + (int64_t)f * 64;
        ac.packet_rows = plan->packet_rows + f;
        ac.packet_rowdw = plan->packet_rowdw + f;
        ac.packet_trace_off = plan->packet_trace_off + f;
        ac.packet_meta_off = plan->packet_meta_off + f;
        ac.n_packets = plan->class_count[c];
        ac.counter = plan->counter + c;
        if (class_is_coop(c)) {
            int waves = std::min(plan->class_count[c], 256 * 20);
            e = kira::launch_banded_sw_coop(ac, class_coop_p(c), waves, ctx->stream);
        } else {
            int waves = std::min(plan->class_count[c], 256 * 16);
            e = kira::launch_banded_sw(ac, kLaneC[c / 2], (c & 1) != 0, waves, ctx->stream);
        }
        if (e) return fail(ctx, KIRA_ERR_RUNTIME, "banded SW launch", e);
    }
    return KIRA_OK;
}

kira_status kira_banded_sw_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads,
                                 int64_t reads_bytes, const int64_t *read_off,
                                 const int32_t *read_len, const uint8_t *refs,
                                 int64_t refs_bytes, const int64_t *ref_off,
                                 const int32_t *ref_len, const int32_t *offset,
                                 const int32_t *band, const int32_t *abort_score,
                                 const kira_sw_cfg *cfg, const kira_sw_knobs *kn,
                                 kira_sw_result *out, uint32_t *cigar_arena, int64_t arena_cap) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (n == 0) return KIRA_OK;
    if (n < 0 || !reads || !read_off || !read_len || !refs || !ref_off || !ref_len || !offset ||
        !abort_score || !cfg || !kn || !out || reads_bytes < 0 || refs_bytes < 0)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    for (int32_t k = 0; k < n; k++) {
        if (read_off[k] < 0 || read_len[k] < 0 || read_off[k] + read_len[k] > reads_bytes ||
            ref_off[k] < 0 || ref_len[k] < 0 || ref_off[k] + ref_len[k] > refs_bytes)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "job outside its buffers");
    }
    kira_sw_plan *plan = nullptr;
    kira_status st = kira_banded_sw_plan_create(ctx, n, read_len, band, cfg, &plan);
    if (st != KIRA_OK) return st;
    struct Guard {
        kira_sw_plan *p;
        std::vector<void *> d;
        ~Guard() {
            for (void *x : d) (void)hipFree(x);
            kira_banded_sw_plan_destroy(p);
        }
    } g{plan, {}};
    uint8_t *d_reads, *d_refs;
    int64_t *d_roff, *d_foff;
    int32_t *d_rlen, *d_flen, *d_off, *d_band = nullptr, *d_abort, *d_res;
    uint32_t *d_cig;
    hipError_t e;
    std::vector<int64_t> foff(n);
    for (int32_t k = 0; k < n; k++) foff[k] = ref_off[k] + kira::kRefPad;
    hipStream_t s = ctx->stream;
    if ((e = dev_alloc(&d_reads, (size_t)reads_bytes + 8)) ||
        (g.d.push_back(d_reads), e = dev_alloc(&d_refs, (size_t)refs_bytes + 2 * kira::kRefPad)) ||
        (g.d.push_back(d_refs), e = dev_alloc(&d_roff, n)) ||
        (g.d.push_back(d_roff), e = dev_alloc(&d_foff, n)) ||
        (g.d.push_back(d_foff), e = dev_alloc(&d_rlen, n)) ||
        (g.d.push_back(d_rlen), e = dev_alloc(&d_flen, n)) ||
        (g.d.push_back(d_flen), e = dev_alloc(&d_off, n)) ||
        (g.d.push_back(d_off), e = dev_alloc(&d_abort, n)) ||
        (g.d.push_back(d_abort), e = dev_alloc(&d_res, (size_t)n * kira::kSwResWords)) ||
        (g.d.push_back(d_res), e = dev_alloc(&d_cig, (size_t)plan->cigar_words)) ||
        (g.d.push_back(d_cig), band && (e = dev_alloc(&d_band, n))))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "batch allocation", e);
    if (d_band) g.d.push_back(d_band);
    if ((e = hipMemsetAsync(d_refs, 0, (size_t)refs_bytes + 2 * kira::kRefPad, s)) ||
        (reads_bytes && (e = hipMemcpyAsync(d_reads, reads, (size_t)reads_bytes,
                                            hipMemcpyHostToDevice, s))) ||
        (refs_bytes && (e = hipMemcpyAsync(d_refs + kira::kRefPad, refs, (size_t)refs_bytes,
                                           hipMemcpyHostToDevice, s))) ||
        (e = hipMemcpyAsync(d_roff, read_off, (size_t)n * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_foff, foff.data(), (size_t)n * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_rlen, read_len, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_flen, ref_len, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_off, offset, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_abort, abort_score, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (d_band && (e = hipMemcpyAsync(d_band, band, (size_t)n * 4, hipMemcpyHostToDevice, s))))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    kira_sw_jobs_dev jd{d_reads, d_roff, d_rlen, d_refs, d_foff, d_flen, d_off,
                        d_band, d_abort, d_res, d_cig};
    if ((st = kira_banded_sw_run_dev(ctx, plan, &jd, cfg, kn)) != KIRA_OK) return st;
    std::vector<int32_t> res((size_t)n * kira::kSwResWords);
    std::vector<uint32_t> cig((size_t)plan->cigar_words);
    if ((e = hipMemcpyAsync(res.data(), d_res, res.size() * 4, hipMemcpyDeviceToHost, s)) ||
        (e = hipMemcpyAsync(cig.data(), d_cig, cig.size() * 4, hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
    int64_t pos = 0;
    for (int32_t k = 0; k < n; k++) {
        const int32_t *r = &res[(size_t)k * kira::kSwResWords];
        kira_sw_result &o = out[k];
        o.score = r[0];
        o.read_end = r[1];
        o.ref_end = (uint32_t)r[2];
        o.early_abort = r[3];
        o.ref_start = (uint32_t)r[4];
        o.read_start = r[5];
        o.n_cigar = r[6];
        o.cells = r[8];
        if (r[6] < 0) return fail(ctx, KIRA_ERR_RUNTIME, "CIGAR capacity exceeded");
        if (r[7] != 0) return fail(ctx, KIRA_ERR_RUNTIME, "int16 band overflow");
        if (pos + r[6] > arena_cap || (r[6] > 0 && !cigar_arena))
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "cigar arena too small");
        o.cigar_off = pos;
        if (r[6] > 0)
            memcpy(cigar_arena + pos, &cig[(size_t)plan->host_cigar_off[k]], (size_t)r[6] * 4);
        pos += r[6];
    }
    return KIRA_OK;
}

}  // extern "C"
