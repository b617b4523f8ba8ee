// Serving loop: continuous batching of read batches against a resident
// genome.  The reference's GPU server (src/cuda/server.rs:15-193) feeds one
// CudaBackend through a single dispatcher thread (src/cuda/dispatcher.rs:21-104)
// and runs one batch at a time; here one worker thread admits queued
// requests into up to `slots` concurrent streams as soon as a slot and the
// request's device memory (carved from one pool of hbm_budget_bytes) are
// free, so batches that arrive while others run overlap on the device.
//
// A request is a batch of reads with sieve jobs (read, window) and banded SW
// jobs (read, window, offset, band, abort, optional sieve job whose
// certification drops it).  Per request, on its slot's stream: one
// host-to-device copy of every input from pinned staging, read packing,
// the sieve + certification, the banded SW kernels (certified jobs skipped
// on the device), one device-to-host copy of every output, an event.  With
// use_graph the whole sequence is captured once per request and launched as
// a hipGraph.  Results of a request do not depend on what else is in flight:
// each request owns its memory and every kernel reads only its own inputs
// and the read-only genome.
//
// Coalescing (group_max_reads > 0): when a slot frees up and several
// requests are waiting, the ones at the front of the queue (up to
// group_max_reads reads, and as many as the free memory holds) are admitted
// together as one device batch: their inputs are staged back to back with
// the read and job indices shifted, one plan covers all their SW jobs, and
// the outputs are split back per request.  Every job's result depends only
// on its own inputs, so a request's results are the same alone or grouped;
// what grouping changes is the size of each launch (a request's long SW
// classes alone leave most of the chip idle).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;
double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit range allocator over one preallocated block.
class RangePool {
    std::map<size_t, size_t> free_;  // offset -> size
    size_t cap_ = 0, used_ = 0;

   public:
    void init(size_t cap) {
        cap_ = cap;
        used_ = 0;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    bool alloc(size_t n, size_t &off) {
        n = align256(std::max<size_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align256(std::max<size_t>(n, 1));
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Byte layout of one request's staging (host, pinned) and device blocks.
struct Layout {
    // input section (same offsets on host and device)
    size_t reads, read_off, read_len, s_read, s_start, s_len, s_limit;
    size_t w_roff, w_rlen, w_foff, w_flen, w_off, w_band, w_abort, w_skip, tables, in_bytes;
    // output section (offsets relative to the output section); the packed
    // CIGARs (o_cigar) follow it in host memory only: the device writes them
    // there itself, so a copy moves only the ops the jobs produced
    size_t o_shift, o_best, o_second, o_flags, o_res, o_coff, out_bytes, o_cigar, host_out_bytes;
    // device-only scratch (relative to the scratch section): read planes,
    // the trace slabs and the plan's CIGAR arena (worst case per job)
    size_t planes, valid, slabs, arena, scratch_bytes;
};

struct Request {
    int64_t ticket = 0;
    double t_submit = 0, t_start = 0;
    int32_t nr = 0, ns = 0, nw = 0, max_sieve_len = 0;
    // copies of the caller's arrays (first touched by the parallel copies)
    kira::PodVec<uint8_t> reads;
    kira::PodVec<int64_t> read_off;
    kira::PodVec<int32_t> read_len;
    kira::PodVec<int32_t> s_read, s_len, s_limit;
    kira::PodVec<int64_t> s_start;
    kira::PodVec<int64_t> w_roff, w_foff;
    kira::PodVec<int32_t> w_rlen, w_flen, w_off, w_band, w_abort, w_skip;
    int64_t reads_bytes = 0;  // total bytes of the reads
    // a group: requests admitted together as one device batch.  Member k
    // keeps its own arrays and is staged at bases[k] (first read, read byte,
    // sieve job, SW job) of the group's; the group itself holds only what
    // its plan needs (w_rlen, w_band) and the totals.  The members are left
    // unchanged, so a group that cannot be placed hands them back as they came.
    std::vector<std::unique_ptr<Request>> members;
    std::vector<std::array<int64_t, 4>> bases;
    kira::SwPlanLayout plan;
    Layout lay{};
    bool laid_out = false;
    size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
    hipGraphExec_t exec = nullptr;
};

struct Result {
    int32_t status = KIRA_OK;
    std::string err;
    kira::PodVec<int32_t> shift, best, second;
    kira::PodVec<uint8_t> flags, skipped;
    kira::PodVec<kira_sw_result> sw;
    kira::PodVec<uint32_t> cigar;
    double t_submit = 0, t_start = 0, t_done = 0;
};

}  // namespace

struct kira_server {
    int device = 0;
    const kira_genome *genome = nullptr;
    kira_server_config cfg{};
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    // per slot: streams the SW classes of long jobs fan out to (not used
    // under graph capture)
    std::vector<kira::SwFanout> fans;
    std::vector<std::unique_ptr<Request>> slot_req;  // null = slot free
    uint8_t *dpool = nullptr, *hpool = nullptr;
    uint8_t *hpool_dev = nullptr;  // hpool as the device addresses it
    RangePool dalloc, halloc;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Request>> pending;
    std::unordered_map<int64_t, std::unique_ptr<Result>> results;
    int64_t next_ticket = 1;
    bool stop = false;
    std::string err;
    kira_server_stats stats{};
    // the creating context's banded SW routes, copied at creation
    kira::SwTuneTable tune;

    void run();
    bool layout(Request &r, std::string &why);
    bool admit(std::unique_ptr<Request> &r, int slot, std::string &why);
    std::unique_ptr<Request> make_group(std::vector<std::unique_ptr<Request>> &&members);
    void stage(const Request &m, const std::array<int64_t, 4> &base, uint8_t *hin,
               const Layout &L);
    void fail_request(Request &r, kira_status st, const std::string &why);
    hipError_t enqueue(Request &r, hipStream_t s, kira::SwFanout *fan);
    void finish(int slot, hipError_t e);
};

bool kira_server::layout(Request &r, std::string &why) {
    if (r.laid_out) return true;
    if (const char *w = kira::sw_plan_layout(r.nw, r.w_rlen.data(), r.w_band.data(), cfg.cfg,
                                             r.plan, &tune)) {
        why = w;
        return false;
    }
    Layout &L = r.lay;
    size_t o = 0;
    auto sec = [&](size_t bytes) {
        size_t at = o;
        o += align256(std::max<size_t>(bytes, 1));
        return at;
    };
    L.reads = sec((size_t)r.reads_bytes);
    L.read_off = sec((size_t)r.nr * 8);
    L.read_len = sec((size_t)r.nr * 4);
    L.s_read = sec((size_t)r.ns * 4);
    L.s_start = sec((size_t)r.ns * 8);
    L.s_len = sec((size_t)r.ns * 4);
    L.s_limit = sec((size_t)r.ns * 4);
    L.w_roff = sec((size_t)r.nw * 8);
    L.w_rlen = sec((size_t)r.nw * 4);
    L.w_foff = sec((size_t)r.nw * 8);
    L.w_flen = sec((size_t)r.nw * 4);
    L.w_off = sec((size_t)r.nw * 4);
    L.w_band = sec((size_t)r.nw * 4);
    L.w_abort = sec((size_t)r.nw * 4);
    L.w_skip = sec((size_t)r.nw * 4);
    L.tables = sec(kira::sw_plan_table_bytes(r.plan));
    L.in_bytes = o;
    o = 0;
    L.o_shift = sec((size_t)r.ns * 4);
    L.o_best = sec((size_t)r.ns * 4);
    L.o_second = sec((size_t)r.ns * 4);
    L.o_flags = sec((size_t)r.ns);
    L.o_res = sec((size_t)r.nw * kira::kSwResWords * 4);
    L.o_coff = sec((size_t)r.nw * 8);
    L.out_bytes = o;
    L.o_cigar = sec((size_t)r.plan.cigar_words * 4);
    L.host_out_bytes = o;
    o = 0;
    L.planes = sec((size_t)r.nr * kira::kReadWords * 4);
    L.valid = sec((size_t)r.nr);
    L.slabs = sec(kira::sw_plan_slab_bytes(r.plan));
    L.arena = sec((size_t)r.plan.cigar_words * 4);
    L.scratch_bytes = o;
    r.dev_bytes = L.in_bytes + L.out_bytes + L.scratch_bytes;
    r.host_bytes = L.in_bytes + L.host_out_bytes;
    r.laid_out = true;
    return true;
}

hipError_t kira_server::enqueue(Request &r, hipStream_t s, kira::SwFanout *fan) {
    const Layout &L = r.lay;
    uint8_t *din = dpool + r.dev_off, *dout = din + L.in_bytes, *dscr = dout + L.out_bytes;
    uint8_t *hin = hpool + r.host_off, *hout = hin + L.in_bytes;
    uint8_t *hout_dev = hpool_dev + r.host_off + L.in_bytes;
    hipError_t e = hipMemcpyAsync(din, hin, L.in_bytes, hipMemcpyHostToDevice, s);
    if (e) return e;
    uint32_t *planes = (uint32_t *)(dscr + L.planes);
    uint8_t *valid = dscr + L.valid;
    int32_t *o_shift = (int32_t *)(dout + L.o_shift), *o_best = (int32_t *)(dout + L.o_best);
    int32_t *o_second = (int32_t *)(dout + L.o_second);
    uint8_t *o_flags = dout + L.o_flags;
    if (r.ns > 0) {
        if ((e = kira::launch_pack_reads_ascii(r.nr, din + L.reads, 0,
                                               (const int64_t *)(din + L.read_off),
                                               (const int32_t *)(din + L.read_len), planes,
                                               valid, s)))
            return e;
        kira::SieveLaunch a{};
        a.n_jobs = r.ns;
        a.read_planes = planes;
        a.read_lens = (const int32_t *)(din + L.read_len);
        a.read_valid = valid;
        a.job_read = (const int32_t *)(din + L.s_read);
        a.jobs_per_read = 1;
        a.win_start = (const int64_t *)(din + L.s_start);
        a.win_len = (const int32_t *)(din + L.s_len);
        a.limit = (const int32_t *)(din + L.s_limit);
        a.limit_all = cfg.sieve_limit;
        a.out_shift = o_shift;
        a.out_best = o_best;
        a.out_second = o_second;
        a.out_flags = o_flags;
        a.cert_max_mism = cfg.cert_max_mism;
        a.cert_strict = cfg.cert_strict;
        a.g_lh = genome->planes.lh;
        a.g_il = genome->planes.il;
        a.g_nz = genome->planes.nz;
        a.g_bases = genome->planes.n_bases;
        int rw = std::max(1, std::min(8, (r.max_sieve_len + 31) / 32));
        if ((e = kira::launch_sieve(a, rw, s))) return e;
    }
    if (r.nw > 0) {
        kira::SwPlanDev d = kira::sw_plan_bind(r.plan, din + L.tables, dscr + L.slabs);
        kira_sw_jobs_dev j{};
        j.reads = din + L.reads;
        j.read_off = (const int64_t *)(din + L.w_roff);
        j.read_len = (const int32_t *)(din + L.w_rlen);
        j.refs = genome->ascii;
        j.ref_off = (const int64_t *)(din + L.w_foff);
        j.ref_len = (const int32_t *)(din + L.w_flen);
        j.offset = (const int32_t *)(din + L.w_off);
        j.band = (const int32_t *)(din + L.w_band);
        j.abort_score = (const int32_t *)(din + L.w_abort);
        j.res = (int32_t *)(dout + L.o_res);
        j.cigar = (uint32_t *)(dscr + L.arena);
        const char *why = kira::sw_plan_launch(r.plan, d, j, cfg.cfg, cfg.knobs,
                                               r.ns > 0 ? o_flags : nullptr,
                                               (const int32_t *)(din + L.w_skip), s, &e, fan);
        if (why) return e ? e : hipErrorInvalidValue;
        // the ops, packed, straight into the pinned host buffer
        if ((e = kira::launch_cigar_compact(r.nw, j.res, j.cigar, d.cigar_off,
                                            (int64_t *)(dout + L.o_coff),
                                            (uint32_t *)(hout_dev + L.o_cigar), s)))
            return e;
    }
    return hipMemcpyAsync(hout, dout, L.out_bytes, hipMemcpyDeviceToHost, s);
}

void kira_server::stage(const Request &m, const std::array<int64_t, 4> &base, uint8_t *hin,
                        const Layout &L) {
    auto put = [&](size_t off, const void *src, size_t bytes) {
        kira::parallel_memcpy(hin + off, src, bytes);
    };
    const int64_t rb = base[0], bb = base[1], sb = base[2], wb = base[3];
    put(L.reads + (size_t)bb, m.reads.data(), (size_t)m.reads_bytes);
    put(L.read_len + (size_t)rb * 4, m.read_len.data(), (size_t)m.nr * 4);
    put(L.s_start + (size_t)sb * 8, m.s_start.data(), (size_t)m.ns * 8);
    put(L.s_len + (size_t)sb * 4, m.s_len.data(), (size_t)m.ns * 4);
    put(L.s_limit + (size_t)sb * 4, m.s_limit.data(), (size_t)m.ns * 4);
    put(L.w_rlen + (size_t)wb * 4, m.w_rlen.data(), (size_t)m.nw * 4);
    put(L.w_foff + (size_t)wb * 8, m.w_foff.data(), (size_t)m.nw * 8);
    put(L.w_flen + (size_t)wb * 4, m.w_flen.data(), (size_t)m.nw * 4);
    put(L.w_off + (size_t)wb * 4, m.w_off.data(), (size_t)m.nw * 4);
    put(L.w_band + (size_t)wb * 4, m.w_band.data(), (size_t)m.nw * 4);
    put(L.w_abort + (size_t)wb * 4, m.w_abort.data(), (size_t)m.nw * 4);
    // indices and read offsets shifted by the member's bases
    int64_t *roff = (int64_t *)(hin + L.read_off) + rb;
    kira::parallel_chunks(m.nr, 16384, [&](int, int64_t b, int64_t e) {
        for (int64_t k = b; k < e; k++) roff[k] = m.read_off[k] + bb;
    });
    int32_t *sread = (int32_t *)(hin + L.s_read) + sb;
    kira::parallel_chunks(m.ns, 16384, [&](int, int64_t b, int64_t e) {
        for (int64_t k = b; k < e; k++) sread[k] = m.s_read[k] + (int32_t)rb;
    });
    int64_t *wroff = (int64_t *)(hin + L.w_roff) + wb;
    int32_t *wskip = (int32_t *)(hin + L.w_skip) + wb;
    kira::parallel_chunks(m.nw, 16384, [&](int, int64_t b, int64_t e) {
        for (int64_t k = b; k < e; k++) {
            wroff[k] = m.w_roff[k] + bb;
            wskip[k] = m.w_skip[k] >= 0 ? m.w_skip[k] + (int32_t)sb : -1;
        }
    });
}

std::unique_ptr<Request> kira_server::make_group(std::vector<std::unique_ptr<Request>> &&members) {
    auto g = std::make_unique<Request>();
    int64_t nw = 0;
    for (const auto &m : members) nw += m->nw;
    g->w_rlen.resize((size_t)nw);
    g->w_band.resize((size_t)nw);
    g->t_submit = members.front()->t_submit;
    for (auto &m : members) {
        g->bases.push_back(std::array<int64_t, 4>{g->nr, g->reads_bytes, g->ns, g->nw});
        if (m->nw) {
            memcpy(g->w_rlen.data() + g->nw, m->w_rlen.data(), (size_t)m->nw * 4);
            memcpy(g->w_band.data() + g->nw, m->w_band.data(), (size_t)m->nw * 4);
        }
        g->nr += m->nr;
        g->reads_bytes += m->reads_bytes;
        g->ns += m->ns;
        g->nw += m->nw;
        g->max_sieve_len = std::max(g->max_sieve_len, m->max_sieve_len);
        g->t_submit = std::min(g->t_submit, m->t_submit);
    }
    g->members = std::move(members);
    return g;
}

void kira_server::fail_request(Request &r, kira_status st, const std::string &why) {
    auto one = [&](const Request &m) {
        auto res = std::make_unique<Result>();
        res->status = st;
        res->err = why;
        res->t_submit = m.t_submit;
        res->t_start = r.t_start ? r.t_start : m.t_submit;
        res->t_done = now_s();
        results[m.ticket] = std::move(res);
    };
    if (r.members.empty())
        one(r);
    else
        for (const auto &m : r.members) one(*m);
    cv_done.notify_all();
}

bool kira_server::admit(std::unique_ptr<Request> &rp, int slot, std::string &why) {
    Request &r = *rp;
    const Layout &L = r.lay;
    // stage every input in pinned memory: one copy per request (per member
    // of a group, at its bases)
    uint8_t *hin = hpool + r.host_off;
    if (r.members.empty()) {
        stage(r, std::array<int64_t, 4>{}, hin, L);
    } else {
        for (size_t k = 0; k < r.members.size(); k++) stage(*r.members[k], r.bases[k], hin, L);
    }
    kira::sw_plan_pack_tables(r.plan, hin + L.tables);
    hipStream_t s = streams[slot];
    hipError_t e;
    if (cfg.use_graph) {
        hipGraph_t graph = nullptr;
        if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) {
            why = hipGetErrorString(e);
            return false;
        }
        hipError_t eq = enqueue(r, s, nullptr);
        e = hipStreamEndCapture(s, &graph);
        if (eq || e) {
            if (graph) (void)hipGraphDestroy(graph);
            why = hipGetErrorString(eq ? eq : e);
            return false;
        }
        e = hipGraphInstantiate(&r.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e || (e = hipGraphLaunch(r.exec, s))) {
            why = hipGetErrorString(e);
            return false;
        }
    } else if ((e = enqueue(r, s, fans[slot].ok ? &fans[slot] : nullptr))) {
        why = hipGetErrorString(e);
        return false;
    }
    if ((e = hipEventRecord(events[slot], s))) {
        why = hipGetErrorString(e);
        return false;
    }
    slot_req[slot] = std::move(rp);
    return true;
}

void kira_server::finish(int slot, hipError_t e) {
    std::unique_ptr<Request> rp = std::move(slot_req[slot]);
    Request &r = *rp;
    const Layout &L = r.lay;
    const uint8_t *hout = hpool + r.host_off + L.in_bytes;
    const int32_t *rr = (const int32_t *)(hout + L.o_res);
    const int64_t *coff = (const int64_t *)(hout + L.o_coff);
    const uint32_t *cg = (const uint32_t *)(hout + L.o_cigar);
    // the results of sieve jobs [s0, s0 + ns) and SW jobs [w0, w0 + nw) of
    // the launch (one request, or one member of a group)
    auto make = [&](const Request &m, int64_t s0, int64_t ns, int64_t w0, int64_t nw) {
        auto res = std::make_unique<Result>();
        res->t_submit = m.t_submit;
        res->t_start = r.t_start;
        if (e != hipSuccess) {
            res->status = KIRA_ERR_RUNTIME;
            res->err = hipGetErrorString(e);
            res->t_done = now_s();
            return res;
        }
        auto get32 = [&](kira::PodVec<int32_t> &v, size_t off) {
            v.resize((size_t)ns);
            kira::parallel_memcpy(v.data(), hout + off + (size_t)s0 * 4, (size_t)ns * 4);
        };
        get32(res->shift, L.o_shift);
        get32(res->best, L.o_best);
        get32(res->second, L.o_second);
        res->flags.resize((size_t)ns);
        kira::parallel_memcpy(res->flags.data(), hout + L.o_flags + s0, (size_t)ns);
        res->sw.resize((size_t)nw);
        res->skipped.resize((size_t)nw);
        // the device packed every job's ops in job order at coff[k]: one copy
        // of this range's packed words, then the per-job fields in parallel
        int64_t c0 = 0, words = 0;
        if (nw > 0) {
            const int32_t last = rr[(size_t)(w0 + nw - 1) * kira::kSwResWords + 6];
            c0 = coff[w0];
            words = coff[w0 + nw - 1] + (last > 0 ? last : 0) - c0;
        }
        res->cigar.resize((size_t)words);
        kira::parallel_memcpy(res->cigar.data(), cg + c0, (size_t)words * 4);
        std::vector<int32_t> bad(kira::host_threads(), -1);  // last failing job per thread
        kira::parallel_chunks(nw, 4096, [&](int t, int64_t b, int64_t e2) {
            for (int64_t k = b; k < e2; k++) {
                const int32_t *q = rr + (size_t)(w0 + k) * kira::kSwResWords;
                kira_sw_result &o = res->sw[k];
                o.score = q[0];
                o.read_end = q[1];
                o.ref_end = (uint32_t)q[2];
                o.early_abort = q[3];
                o.ref_start = (uint32_t)q[4];
                o.read_start = q[5];
                o.n_cigar = q[6];
                o.cells = q[8];
                o.cigar_off = coff[w0 + k] - c0;
                res->skipped[k] = (uint8_t)(q[9] == 1);
                if (q[6] < 0 || q[7] != 0) {
                    o.n_cigar = 0;
                    bad[t] = (int32_t)k;
                }
            }
        });
        int32_t worst = -1;
        for (int32_t k : bad) worst = std::max(worst, k);
        if (worst >= 0) {
            const int32_t *q = rr + (size_t)(w0 + worst) * kira::kSwResWords;
            res->status = KIRA_ERR_RUNTIME;
            res->err = q[6] < 0 ? "CIGAR capacity exceeded" : "int16 band overflow";
        }
        res->t_done = now_s();
        return res;
    };
    std::vector<std::pair<int64_t, std::unique_ptr<Result>>> out;
    if (r.members.empty()) {
        out.emplace_back(r.ticket, make(r, 0, r.ns, 0, r.nw));
    } else {
        for (size_t k = 0; k < r.members.size(); k++) {
            const Request &m = *r.members[k];
            out.emplace_back(m.ticket, make(m, r.bases[k][2], m.ns, r.bases[k][3], m.nw));
        }
    }
    if (r.exec) (void)hipGraphExecDestroy(r.exec);
    std::lock_guard<std::mutex> lk(mu);
    dalloc.release(r.dev_off, r.dev_bytes);
    halloc.release(r.host_off, r.host_bytes);
    for (auto &x : out) {
        stats.completed++;
        results[x.first] = std::move(x.second);
    }
    cv_done.notify_all();
}

void kira_server::run() {
    (void)hipSetDevice(device);
    const int S = (int)streams.size();
    for (;;) {
        bool progressed = false;
        // completions
        for (int k = 0; k < S; k++) {
            if (!slot_req[k]) continue;
            hipError_t q = hipEventQuery(events[k]);
            if (q == hipErrorNotReady) continue;
            finish(k, q);
            progressed = true;
        }
        // admissions, first come first served
        std::unique_lock<std::mutex> lk(mu);
        int in_flight = 0;
        for (int k = 0; k < S; k++) in_flight += slot_req[k] ? 1 : 0;
        bool no_group = false;  // a group that did not fit: its first member goes alone
        while (!pending.empty() && in_flight < S) {
            Request &r = *pending.front();
            std::string why;
            bool ok = layout(r, why);
            if (ok && r.dev_bytes > dalloc.capacity()) {
                ok = false;
                why = "request larger than the device memory budget";
            }
            if (ok && r.host_bytes > halloc.capacity()) {
                ok = false;
                why = "request larger than the host staging pool";
            }
            if (!ok) {
                std::unique_ptr<Request> bad = std::move(pending.front());
                pending.pop_front();
                fail_request(*bad,
                             why.find("budget") != std::string::npos ||
                                     why.find("staging") != std::string::npos
                                 ? KIRA_ERR_OUT_OF_MEMORY
                                 : KIRA_ERR_INVALID_ARGUMENT,
                             why);
                progressed = true;
                continue;
            }
            // coalescing: the requests behind it that fit with it (reads and
            // free memory; the merged layout is about the sum of the parts)
            size_t n_group = 1;
            if (cfg.group_max_reads > 0 && !no_group) {
                int64_t reads = r.nr;
                size_t dev = r.dev_bytes, host = r.host_bytes;
                const size_t dfree = dalloc.capacity() - dalloc.used(),
                             hfree = halloc.capacity() - halloc.used();
                for (size_t k = 1; k < pending.size(); k++) {
                    Request &q = *pending[k];
                    std::string w2;
                    if (!layout(q, w2) || reads + q.nr > cfg.group_max_reads) break;
                    const size_t d2 = dev + q.dev_bytes, h2 = host + q.host_bytes;
                    if (d2 + d2 / 16 + ((size_t)1 << 20) > dfree ||
                        h2 + h2 / 16 + ((size_t)1 << 20) > hfree)
                        break;
                    reads += q.nr;
                    dev = d2;
                    host = h2;
                    n_group = k + 1;
                }
            }
            std::unique_ptr<Request> rp;
            if (n_group > 1) {
                std::vector<std::unique_ptr<Request>> mem;
                for (size_t k = 0; k < n_group; k++) {
                    mem.push_back(std::move(pending.front()));
                    pending.pop_front();
                }
                lk.unlock();
                rp = make_group(std::move(mem));
                std::string w3;
                const bool lok = layout(*rp, w3);
                lk.lock();
                if (!lok) {
                    fail_request(*rp, KIRA_ERR_INVALID_ARGUMENT, w3);
                    progressed = true;
                    continue;
                }
            }
            Request &a = rp ? *rp : *pending.front();
            size_t doff, hoff;
            bool got = dalloc.alloc(a.dev_bytes, doff);
            if (got && !halloc.alloc(a.host_bytes, hoff)) {
                dalloc.release(doff, a.dev_bytes);
                got = false;
            }
            if (!got) {
                if (!rp) break;  // wait for memory to free up
                // the group does not fit after all: back to the queue, in
                // order, and the first goes alone
                for (auto it = rp->members.rbegin(); it != rp->members.rend(); ++it)
                    pending.push_front(std::move(*it));
                no_group = true;
                continue;
            }
            a.dev_off = doff;
            a.host_off = hoff;
            int slot = 0;
            while (slot_req[slot]) slot++;
            if (!rp) {
                rp = std::move(pending.front());
                pending.pop_front();
            }
            stats.admitted += rp->members.empty() ? 1 : (int64_t)rp->members.size();
            stats.grouped += (int64_t)rp->members.size();
            stats.peak_device_bytes = std::max<int64_t>(stats.peak_device_bytes, dalloc.used());
            lk.unlock();
            rp->t_start = now_s();
            std::string why2;
            Request *raw = rp.get();
            const bool launched = admit(rp, slot, why2);  // on success the slot owns it
            lk.lock();
            if (!launched) {
                dalloc.release(raw->dev_off, raw->dev_bytes);
                halloc.release(raw->host_off, raw->host_bytes);
                if (raw->exec) (void)hipGraphExecDestroy(raw->exec);
                fail_request(*raw, KIRA_ERR_RUNTIME, why2);
            }
            progressed = true;
            no_group = false;
            in_flight = 0;
            for (int k = 0; k < S; k++) in_flight += slot_req[k] ? 1 : 0;
            stats.peak_in_flight = std::max<int32_t>(stats.peak_in_flight, in_flight);
        }
        if (stop && pending.empty() && in_flight == 0) return;
        if (!progressed) cv_work.wait_for(lk, std::chrono::microseconds(in_flight ? 50 : 2000));
    }
}

extern "C" {

kira_status kira_server_create(kira_ctx *ctx, const kira_genome *g, const kira_server_config *c,
                               kira_server **out) {
    if (!ctx || !g || !c || !out || c->slots < 1 || c->slots > 16 || c->hbm_budget_bytes <= 0 ||
        c->host_staging_bytes <= 0)
        return KIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    kira_server *s = new (std::nothrow) kira_server();
    if (!s) return KIRA_ERR_OUT_OF_MEMORY;
    s->device = ctx->device;
    s->tune = ctx->tune;
    s->genome = g;
    s->cfg = *c;
    hipError_t e = hipSetDevice(ctx->device);
    if (!e) e = hipMalloc((void **)&s->dpool, (size_t)c->hbm_budget_bytes);
    if (!e) e = hipHostMalloc((void **)&s->hpool, (size_t)c->host_staging_bytes, 0);
    if (!e) e = hipHostGetDevicePointer((void **)&s->hpool_dev, s->hpool, 0);
    for (int k = 0; !e && k < c->slots; k++) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (!e) s->streams.push_back(st);
        if (!e) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (!e) s->events.push_back(ev);
    }
    // without its streams a slot keeps every SW class on its own stream
    s->fans.resize(c->slots);
    if (!e)
        for (auto &f : s->fans) (void)kira::sw_fanout_create(f);
    if (e) {
        for (auto st : s->streams) (void)hipStreamDestroy(st);
        for (auto ev : s->events) (void)hipEventDestroy(ev);
        if (s->dpool) (void)hipFree(s->dpool);
        if (s->hpool) (void)hipHostFree(s->hpool);
        if (ctx) ctx->err = std::string("server create: ") + hipGetErrorString(e);
        delete s;
        return KIRA_ERR_OUT_OF_MEMORY;
    }
    s->dalloc.init((size_t)c->hbm_budget_bytes);
    s->halloc.init((size_t)c->host_staging_bytes);
    s->slot_req.resize(c->slots);
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return KIRA_OK;
}

kira_status kira_server_submit(kira_server *s, const kira_request *q, int64_t *ticket) {
    if (!s || !q || !ticket || q->n_reads < 0 || q->n_sieve < 0 || q->n_sw < 0)
        return KIRA_ERR_INVALID_ARGUMENT;
    auto bad = [&](const char *why) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = why;
        return KIRA_ERR_INVALID_ARGUMENT;
    };
    if ((q->n_reads > 0 && (!q->reads || !q->read_off || !q->read_len)) ||
        (q->n_sieve > 0 && (!q->sieve_read || !q->sieve_win_start || !q->sieve_win_len)) ||
        (q->n_sw > 0 && (!q->sw_read || !q->sw_win_start || !q->sw_win_len || !q->sw_offset ||
                         !q->sw_abort)))
        return bad("missing request array");
    const int64_t G = s->genome->planes.n_bases;
    auto r = std::make_unique<Request>();
    r->nr = q->n_reads;
    r->ns = q->n_sieve;
    r->nw = q->n_sw;
    for (int32_t k = 0; k < r->nr; k++)
        if (q->read_off[k] < 0 || q->read_len[k] < 0) return bad("negative read offset / length");
    // reads are re-based to one contiguous copy
    r->read_off.resize(r->nr);
    r->read_len.resize(r->nr);
    memcpy(r->read_len.data(), q->read_len, (size_t)r->nr * 4);
    int64_t pos = 0;
    for (int32_t k = 0; k < r->nr; k++) {
        r->read_off[k] = pos;
        pos += q->read_len[k];
    }
    r->reads.resize((size_t)pos);
    r->reads_bytes = pos;
    kira::parallel_chunks(r->nr, 2048, [&](int, int64_t b, int64_t e) {
        for (int64_t k = b; k < e; k++)
            if (q->read_len[k])
                memcpy(r->reads.data() + r->read_off[k], q->reads + q->read_off[k],
                       (size_t)q->read_len[k]);
    });
    // the job arrays copied and checked in parallel over job ranges
    r->s_read.resize(r->ns);
    r->s_start.resize(r->ns);
    r->s_len.resize(r->ns);
    r->s_limit.resize(r->ns);
    {
        const int T = kira::host_threads();
        std::vector<int32_t> max_len((size_t)T, 0);
        std::vector<uint8_t> bad_read((size_t)T, 0), bad_win((size_t)T, 0);
        kira::parallel_chunks(r->ns, 16384, [&](int t, int64_t b, int64_t e) {
            int32_t mx = 0;
            for (int64_t k = b; k < e; k++) {
                const int32_t rd = q->sieve_read[k];
                const int64_t ws = q->sieve_win_start[k];
                const int32_t wl = q->sieve_win_len[k];
                r->s_read[k] = rd;
                r->s_start[k] = ws;
                r->s_len[k] = wl;
                r->s_limit[k] = q->sieve_limit ? q->sieve_limit[k] : s->cfg.sieve_limit;
                if (rd < 0 || rd >= r->nr) {
                    bad_read[(size_t)t] = 1;
                    continue;
                }
                if (ws < 0 || wl < 0 || ws + wl > G) bad_win[(size_t)t] = 1;
                mx = std::max(mx, r->read_len[rd]);
            }
            max_len[(size_t)t] = mx;
        });
        for (int t = 0; t < T; t++) {
            if (bad_read[(size_t)t]) return bad("sieve job read out of range");
            if (bad_win[(size_t)t]) return bad("sieve window outside the genome");
            r->max_sieve_len = std::max(r->max_sieve_len, max_len[(size_t)t]);
        }
    }
    r->max_sieve_len = std::min(r->max_sieve_len, kira::kMaxReadBases);
    r->w_roff.resize(r->nw);
    r->w_rlen.resize(r->nw);
    r->w_foff.resize(r->nw);
    r->w_flen.resize(r->nw);
    r->w_off.resize(r->nw);
    r->w_band.resize(r->nw);
    r->w_abort.resize(r->nw);
    r->w_skip.resize(r->nw);
    {
        // 1 read out of range, 2 window outside the genome, 3 skip index
        const int T = kira::host_threads();
        std::vector<uint8_t> why((size_t)T, 0);
        kira::parallel_chunks(r->nw, 16384, [&](int t, int64_t b, int64_t e) {
            for (int64_t k = b; k < e; k++) {
                const int32_t rd = q->sw_read[k];
                if (rd < 0 || rd >= r->nr) {
                    why[(size_t)t] = why[(size_t)t] ? why[(size_t)t] : 1;
                    continue;
                }
                if (q->sw_win_start[k] < 0 || q->sw_win_len[k] < 0 ||
                    q->sw_win_start[k] + q->sw_win_len[k] > G)
                    why[(size_t)t] = why[(size_t)t] ? why[(size_t)t] : 2;
                r->w_roff[k] = r->read_off[rd];
                r->w_rlen[k] = r->read_len[rd];
                r->w_foff[k] = q->sw_win_start[k];
                r->w_flen[k] = q->sw_win_len[k];
                r->w_off[k] = q->sw_offset[k];
                r->w_abort[k] = q->sw_abort[k];
                r->w_band[k] = q->sw_band ? q->sw_band[k] : s->cfg.cfg.bandwidth;
                r->w_skip[k] = q->sw_skip_sieve ? q->sw_skip_sieve[k] : -1;
                if (r->w_skip[k] >= r->ns) why[(size_t)t] = why[(size_t)t] ? why[(size_t)t] : 3;
            }
        });
        for (int t = 0; t < T; t++) {
            if (why[(size_t)t] == 1) return bad("SW job read out of range");
            if (why[(size_t)t] == 2) return bad("SW window outside the genome");
            if (why[(size_t)t] == 3) return bad("SW skip index out of range");
        }
    }
    r->t_submit = now_s();
    // the job plan and buffer layout on the submitting thread, so the worker
    // (which stages, launches and collects) does not serialise them; a
    // request that cannot be laid out is reported by the worker
    {
        std::string why;
        (void)s->layout(*r, why);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return KIRA_ERR_INVALID_ARGUMENT;
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->stats.submitted++;
    s->pending.push_back(std::move(r));
    s->cv_work.notify_one();
    return KIRA_OK;
}

kira_status kira_server_wait(kira_server *s, int64_t ticket, kira_response *out) {
    if (!s || !out) return KIRA_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) return KIRA_ERR_INVALID_ARGUMENT;
    s->cv_done.wait(lk, [&] { return s->results.count(ticket) > 0; });
    const Result &r = *s->results[ticket];
    memset(out, 0, sizeof(*out));
    out->status = r.status;
    out->n_sieve = (int32_t)r.shift.size();
    out->n_sw = (int32_t)r.sw.size();
    out->shift = r.shift.data();
    out->best = r.best.data();
    out->second = r.second.data();
    out->flags = r.flags.data();
    out->sw = r.sw.data();
    out->sw_skipped = r.skipped.data();
    out->cigar = r.cigar.data();
    out->cigar_words = (int64_t)r.cigar.size();
    out->submit_s = r.t_submit;
    out->start_s = r.t_start;
    out->done_s = r.t_done;
    if (r.status != KIRA_OK) s->err = r.err;
    return KIRA_OK;
}

void kira_server_release(kira_server *s, int64_t ticket) {
    if (!s) return;
    std::lock_guard<std::mutex> lk(s->mu);
    s->results.erase(ticket);
}

void kira_server_get_stats(kira_server *s, kira_server_stats *out) {
    if (!s || !out) return;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
}

const char *kira_server_last_error(kira_server *s) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->err.c_str();
}

void kira_server_destroy(kira_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (auto ev : s->events) (void)hipEventDestroy(ev);
    for (auto &f : s->fans) {
        for (hipStream_t st : f.aux)
            if (st) (void)hipStreamSynchronize(st);
        kira::sw_fanout_destroy(f);
    }
    if (s->dpool) (void)hipFree(s->dpool);
    if (s->hpool) (void)hipHostFree(s->hpool);
    delete s;
}

}  // extern "C"
