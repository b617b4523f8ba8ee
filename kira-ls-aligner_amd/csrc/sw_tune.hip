// Autotuner of the banded SW routing.  The reference picks its lane groups
// per (L, W) shape and a lane count per host (stage4_alignment.rs:1064-1069,
// 1153-1160; src/simd/mod.rs:30); here each job is routed to a kernel form
// (one lane per job, or one of the cooperative (P, G) shapes), and which form
// is fastest depends on the band, the read length and how many such jobs a
// batch holds (occupancy).  The tuner times every eligible route on a seeded
// synthetic batch per shape, checks that all routes give identical results,
// and keeps the fastest per shape key in the context (and in a cache file).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

kira_status fail(kira_ctx *ctx, kira_status st, const char *what, hipError_t e = hipSuccess) {
    char buf[512];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    ctx->err = buf;
    return st;
}

// Architecture name and CU count of the context's device: the cache key.
std::string device_key(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    std::string arch = p.gcnArchName;
    for (char &c : arch)
        if (c == ' ' || c == '\t') c = '_';
    return arch + "/" + std::to_string(p.multiProcessorCount) + "cu";
}

struct XorShift {
    uint64_t s;
    uint64_t next() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
};

// Synthetic jobs of one shape: windows are slices of a random text (clamp_window
// geometry: flank max(2w, 50) on each side, offset = flank), reads are the
// window's middle with 3% substitutions and 0.5% one-base indels.
struct Synth {
    std::vector<uint8_t> text;  // kRefPad zero bytes, the text, kRefPad zero bytes
    std::vector<uint8_t> reads;
    std::vector<int64_t> read_off, ref_off;
    std::vector<int32_t> read_len, ref_len, offset, band, abort_score;
};

void make_synth(int32_t w, int32_t L, int32_t n, uint64_t seed, Synth &S) {
    XorShift rng{seed * 0x9E3779B97F4A7C15ull + 0x1234567ull};
    const int32_t flank = std::max(2 * w, 50);
    const int32_t W = L + 2 * flank;
    const int64_t text_len = std::max<int64_t>(4 * (int64_t)W, 1 << 22);
    S.text.assign((size_t)(text_len + 2 * kira::kRefPad), 0);
    for (int64_t k = 0; k < text_len; k++)
        S.text[(size_t)(kira::kRefPad + k)] = "ACGT"[rng.next() & 3];
    S.reads.clear();
    S.reads.reserve((size_t)n * (L + 8));
    for (int32_t k = 0; k < n; k++) {
        const int64_t ws = (int64_t)(rng.next() % (uint64_t)(text_len - W));
        const uint8_t *src = &S.text[(size_t)(kira::kRefPad + ws + flank)];
        S.read_off.push_back((int64_t)S.reads.size());
        int32_t got = 0;
        for (int32_t p = 0; got < L && p < L + flank; p++) {
            const uint32_t u = (uint32_t)(rng.next() % 1000);
            if (u < 25) continue;  // deletion
            if (u < 50) {          // insertion before the base
                S.reads.push_back("ACGT"[rng.next() & 3]);
                if (++got == L) break;
            }
            uint8_t b = src[p];
            if (rng.next() % 100 < 3) b = "ACGT"[(rng.next() & 3)];
            S.reads.push_back(b);
            got++;
        }
        S.read_len.push_back(got);
        S.ref_off.push_back(kira::kRefPad + ws);
        S.ref_len.push_back(W);
        S.offset.push_back(flank);
        S.band.push_back(w);
        S.abort_score.push_back(INT32_MIN / 8);
    }
    S.reads.resize(S.reads.size() + 8, 0);  // readable past the last read
}

struct Run {
    float ms = 0;
    std::vector<int32_t> res;
    std::vector<uint32_t> cig;
};

// One routing of the synthetic batch: plan with the route forced, one
// warm-up launch and two timed ones (the faster counts), results kept.
kira_status run_route(kira_ctx *ctx, const Synth &S, const kira_sw_cfg &cfg,
                      const kira_sw_knobs &kn, int route, const kira_sw_jobs_dev &jd_in,
                      Run &out) {
    const int32_t n = (int32_t)S.read_len.size();
    kira::SwPlanLayout L;
    if (const char *why = kira::sw_plan_layout(n, S.read_len.data(), S.band.data(), cfg, L,
                                               nullptr, route))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, why);
    const size_t tb = kira::sw_plan_table_bytes(L), sb = kira::sw_plan_slab_bytes(L);
    std::vector<uint8_t> host(tb);
    kira::sw_plan_pack_tables(L, host.data());
    hipStream_t s = ctx->stream;
    kira::DevMem dm(ctx);
    uint8_t *mem = nullptr;
    uint32_t *d_cig = nullptr;
    hipError_t e;
    if ((e = dm.alloc(&mem, tb + sb)) || (e = dm.alloc(&d_cig, (size_t)L.cigar_words)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "tuning plan allocation", e);
    if ((e = hipMemcpyAsync(mem, host.data(), tb, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "tuning plan upload", e);
    const kira::SwPlanDev d = kira::sw_plan_bind(L, mem, mem + tb);
    kira_sw_jobs_dev jd = jd_in;
    jd.cigar = d_cig;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if ((e = hipEventCreate(&ev[0])) || (e = hipEventCreate(&ev[1]))) {
        if (ev[0]) (void)hipEventDestroy(ev[0]);
        return fail(ctx, KIRA_ERR_RUNTIME, "tuning events", e);
    }
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            (void)hipEventDestroy(e[0]);
            (void)hipEventDestroy(e[1]);
        }
    } eg{ev};
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        if ((e = hipEventRecord(ev[0], s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "tuning event", e);
        const char *why =
            kira::sw_plan_launch(L, d, jd, cfg, kn, nullptr, nullptr, s, &e, &ctx->fan);
        if (why) return fail(ctx, KIRA_ERR_RUNTIME, why, e);
        float ms = 0;
        if ((e = hipEventRecord(ev[1], s)) || (e = hipEventSynchronize(ev[1])) ||
            (e = hipEventElapsedTime(&ms, ev[0], ev[1])))
            return fail(ctx, KIRA_ERR_RUNTIME, "tuning timing", e);
        if (rep > 0) best = std::min(best, ms);
    }
    out.ms = best;
    out.res.resize((size_t)n * kira::kSwResWords);
    out.cig.resize((size_t)L.cigar_words);
    if ((e = hipMemcpyAsync(out.res.data(), jd.res, out.res.size() * 4, hipMemcpyDeviceToHost,
                            s)) ||
        (e = hipMemcpyAsync(out.cig.data(), d_cig, out.cig.size() * 4, hipMemcpyDeviceToHost,
                            s)) ||
        (e = hipStreamSynchronize(s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "tuning results", e);
    // only the written part of each job's CIGAR slot counts (cigar_off / the
    // op count do not depend on the routing)
    for (int32_t k = 0; k < n; k++) {
        const int32_t c = out.res[(size_t)k * kira::kSwResWords + 6];
        const int64_t o = L.cigar_off[k];
        for (int64_t x = o + std::max(c, 0); x < o + L.cigar_cap[k]; x++) out.cig[(size_t)x] = 0;
    }
    return KIRA_OK;
}

bool write_cache(const std::string &path, const std::string &dev,
                 const std::map<uint32_t, std::pair<int, std::pair<float, float>>> &entries) {
    std::vector<std::string> keep;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string d;
            uint32_t key = 0;
            if (!(ls >> d >> key) || line[0] == '#') continue;
            if (d == dev && entries.count(key)) continue;
            keep.push_back(line);
        }
    }
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out) return false;
        out << "# kira banded SW routes: device key route ms_route ms_builtin\n";
        for (const auto &l : keep) out << l << "\n";
        for (const auto &kv : entries)
            out << dev << " " << kv.first << " " << kv.second.first << " "
                << kv.second.second.first << " " << kv.second.second.second << "\n";
        if (!out) return false;
    }
    return rename(tmp.c_str(), path.c_str()) == 0;
}

}  // namespace

extern "C" {

kira_status kira_sw_autotune(kira_ctx *ctx, const kira_sw_cfg *cfg, const kira_sw_knobs *knobs,
                             int32_t n_shapes, const kira_sw_shape *shapes,
                             const char *cache_path, kira_sw_tune_result *out) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (!cfg || !knobs || n_shapes < 0 || (n_shapes > 0 && (!shapes || !out)))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    const std::string dev = device_key(ctx->device);
    std::map<uint32_t, std::pair<int, std::pair<float, float>>> entries;
    for (int32_t sh = 0; sh < n_shapes; sh++) {
        const int32_t w = std::max(shapes[sh].band, 1), L = std::max(shapes[sh].read_len, 1);
        if (shapes[sh].n_jobs <= 0 || w > 271 || L > 60000)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "tuning shape outside 1..271 bands");
        const uint32_t key = kira::sw_tune_key(kira::sw_band_bucket(w), kira::sw_len_bucket(L),
                                               kira::sw_count_bucket(shapes[sh].n_jobs));
        // the timed batch: the shape's job count, at most 262144 jobs and 64 MB
        // of reads
        const int32_t n = (int32_t)std::max<int64_t>(
            1, std::min<int64_t>({(int64_t)shapes[sh].n_jobs, 262144, (64ll << 20) / L}));
        Synth S;
        try {
            make_synth(w, L, n, key, S);
        } catch (const std::bad_alloc &) {
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "tuning batch");
        }
        // candidates: the built-in routing and every route the shape may take
        // (one lane per job is not a candidate where a 257-position band would
        // spill the register file and a cooperative shape holds the band)
        // and, of each cooperative form (int32, packed) and group width, the
        // smallest shape that holds the band, unless it holds more than twice
        // the form's smallest (a larger shape only adds idle positions)
        std::vector<int> cand{-1};
        const uint32_t ok = kira::sw_routes_for(w, L, *cfg);
        const bool coop_any = (ok >> kira::kSwRouteCoop) != 0;
        if (((ok >> kira::kSwRouteLane) & 1u) && !(w > 64 && coop_any))
            cand.push_back(kira::kSwRouteLane);
        for (int form = 0; form < 2; form++) {
            const int k0 = form ? kira::kSwCoop32 : 0;
            const int k1 = form ? kira::sw_num_coop_shapes() : kira::kSwCoop32;
            int min_cap = 1 << 30;
            for (int k = k0; k < k1; k++) {
                int P, G;
                kira::sw_coop_shape(k, &P, &G);
                if ((ok >> (kira::kSwRouteCoop + k)) & 1u) min_cap = std::min(min_cap, P * G);
            }
            for (int gw = 8; gw <= 32; gw *= 2) {
                int best_k = -1, best_cap = 1 << 30;
                for (int k = k0; k < k1; k++) {
                    int P, G;
                    kira::sw_coop_shape(k, &P, &G);
                    if (G == gw && ((ok >> (kira::kSwRouteCoop + k)) & 1u) && P * G < best_cap) {
                        best_k = k;
                        best_cap = P * G;
                    }
                }
                if (best_k >= 0 && best_cap <= 2 * min_cap)
                    cand.push_back(kira::kSwRouteCoop + best_k);
            }
        }
        kira::DevMem dm(ctx);
        uint8_t *d_reads = nullptr, *d_text = nullptr;
        int64_t *d_ro = nullptr, *d_fo = nullptr;
        int32_t *d_rl = nullptr, *d_fl = nullptr, *d_off = nullptr, *d_band = nullptr,
                *d_ab = nullptr, *d_res = nullptr;
        hipStream_t s = ctx->stream;
        if ((e = dm.alloc(&d_reads, S.reads.size())) || (e = dm.alloc(&d_text, S.text.size())) ||
            (e = dm.alloc(&d_ro, n)) || (e = dm.alloc(&d_fo, n)) || (e = dm.alloc(&d_rl, n)) ||
            (e = dm.alloc(&d_fl, n)) || (e = dm.alloc(&d_off, n)) || (e = dm.alloc(&d_band, n)) ||
            (e = dm.alloc(&d_ab, n)) || (e = dm.alloc(&d_res, (size_t)n * kira::kSwResWords)))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "tuning allocation", e);
        auto up = [&](void *dst, const void *src, size_t bytes) {
            return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        };
        if ((e = up(d_reads, S.reads.data(), S.reads.size())) ||
            (e = up(d_text, S.text.data(), S.text.size())) ||
            (e = up(d_ro, S.read_off.data(), (size_t)n * 8)) ||
            (e = up(d_fo, S.ref_off.data(), (size_t)n * 8)) ||
            (e = up(d_rl, S.read_len.data(), (size_t)n * 4)) ||
            (e = up(d_fl, S.ref_len.data(), (size_t)n * 4)) ||
            (e = up(d_off, S.offset.data(), (size_t)n * 4)) ||
            (e = up(d_band, S.band.data(), (size_t)n * 4)) ||
            (e = up(d_ab, S.abort_score.data(), (size_t)n * 4)))
            return fail(ctx, KIRA_ERR_RUNTIME, "tuning upload", e);
        const kira_sw_jobs_dev jd{d_reads, d_ro, d_rl, d_text, d_fo, d_fl, d_off,
                                  d_band,  d_ab, d_res, nullptr};
        Run base;
        int best_route = kira::kSwRouteDefault;
        float best_ms = 0;
        for (size_t c = 0; c < cand.size(); c++) {
            Run r;
            kira_status st = run_route(ctx, S, *cfg, *knobs, cand[c], jd, r);
            if (st != KIRA_OK) return st;
            if (c == 0) {
                base = std::move(r);
                best_ms = base.ms;
                continue;
            }
            if (r.res != base.res || r.cig != base.cig) {
                // name the first differing job and field
                size_t at = 0;
                while (at < r.res.size() && r.res[at] == base.res[at]) at++;
                char buf[256];
                if (at < r.res.size()) {
                    const size_t job = at / kira::kSwResWords, f = at % kira::kSwResWords;
                    snprintf(buf, sizeof buf,
                             "banded SW routes disagree on a tuning batch: band %d, read length "
                             "%d, route %d vs built-in, job %zu (read %d bases), result word "
                             "%zu: %d vs %d",
                             w, L, cand[c], job, S.read_len[job], f, r.res[at], base.res[at]);
                } else {
                    snprintf(buf, sizeof buf,
                             "banded SW routes disagree on a tuning batch: band %d, read length "
                             "%d, route %d vs built-in, CIGAR words",
                             w, L, cand[c]);
                }
                return fail(ctx, KIRA_ERR_RUNTIME, buf);
            }
            if (r.ms < best_ms) {
                best_ms = r.ms;
                best_route = cand[c];
            }
        }
        try {
            ctx->tune.route[key] = (uint8_t)best_route;
        } catch (const std::bad_alloc &) {
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "tuning table");
        }
        entries[key] = {best_route, {best_ms, base.ms}};
        out[sh] = kira_sw_tune_result{key, best_route, (int32_t)cand.size() - 1, best_ms, base.ms};
    }
    if (cache_path && !entries.empty() && !write_cache(cache_path, dev, entries))
        return fail(ctx, KIRA_ERR_RUNTIME, "cannot write the tuning cache");
    return KIRA_OK;
}

kira_status kira_sw_tune_load(kira_ctx *ctx, const char *cache_path, int32_t *n_loaded) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (!cache_path) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    std::ifstream in(cache_path);
    if (!in) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "cannot read the tuning cache");
    const std::string dev = device_key(ctx->device);
    int32_t n = 0;
    std::string line;
    try {
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ls(line);
            std::string d;
            uint32_t key = 0;
            int route = -1;
            if (!(ls >> d >> key >> route) || route < kira::kSwRouteDefault ||
                route >= kira::kSwRouteCoop + kira::sw_num_coop_shapes())
                return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "malformed tuning cache line");
            if (d != dev) continue;
            ctx->tune.route[key] = (uint8_t)route;
            n++;
        }
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "tuning table");
    }
    if (n_loaded) *n_loaded = n;
    return KIRA_OK;
}

void kira_sw_tune_clear(kira_ctx *ctx) {
    if (ctx) ctx->tune.route.clear();
}

}  // extern "C"
