// The reference's fast DP cascade, try_fast_dp_alignment_inner
// (src/alignment/mod.rs:285-440), over a batch of chains against a resident
// genome: packed spectral certificate, Myers reject, scored WFA with the
// suspect-origin retry, the partial-read check, and the record
// (wfa_result_to_alignment, mod.rs:481-575).
//
// Device work: the spectral scan + certification (one wave per chain, one
// lane per shift), Myers (one lane per chain) and WFA (one wave per chain),
// reading the windows straight from the genome's resident ASCII copy.  Host
// work: the router gates and window arithmetic, the origin probe of the
// retry (a common-prefix scan of at most 65 starts), the partial-read check
// and NM / MD, on the genome's host copy.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

using kira::host_threads;
using kira::parallel_chunks;

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

using kira::DevMem;

// bitpacked.rs:89-91: A / C / G / T in either case
__device__ __forceinline__ bool packable(uint32_t b) {
    const uint32_t u = b & 0xDFu;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

// bounded_hamming (src/alignment/mod.rs:764-776) on LDS bytes, lengths equal
__device__ int32_t bounded_hamming(const uint8_t *a, const uint8_t *b, int32_t len, int32_t limit) {
    int32_t m = 0;
    for (int32_t k = 0; k < len; k++) {
        m += a[k] != b[k];
        if (m > limit) break;
    }
    return m;
}

// spectral_hit_is_certified (src/alignment/mod.rs:688-762) on bytes
__device__ bool certify_bytes(const uint8_t *read, const uint8_t *ref, int32_t L, int32_t best,
                              int32_t second, int32_t max_mism, bool strict) {
    if (best > max_mism) return false;
    if (second >= 0 && second <= best + 1) return false;
    if (best == 0) return true;
    if (best > 2) return false;
    int32_t pos[2] = {-1, -1}, n = 0;
    for (int32_t i = 0; i < L; i++)
        if (read[i] != ref[i]) {
            if (n < 2) pos[n] = i;
            n++;
        }
    if (n != best) return false;
    const int32_t edge = L < 8 ? L : 8;
    for (int32_t k = 0; k < n; k++)
        if (pos[k] < edge || pos[k] + edge >= L) return false;
    if (n == 2 && pos[1] - pos[0] <= 4) return false;
    const int32_t p = pos[0];
    const int32_t suffix = L - (p + 1);
    if (suffix >= (strict ? 4 : 8)) {
        if (bounded_hamming(read + p + 1, ref + p, L - p - 1, 1) <= 1) return false;
        if (bounded_hamming(read + p, ref + p + 1, L - 1 - p, 1) <= 1) return false;
        if (strict && suffix >= 6) {
            if (bounded_hamming(read + p + 2, ref + p, L - p - 2, 1) <= 1) return false;
            if (bounded_hamming(read + p, ref + p + 2, L - 2 - p, 1) <= 1) return false;
        }
    }
    return true;
}

// try_packed_spectral_alignment (mod.rs:624-686) for one chain per wave:
// both sequences must pack (A/C/G/T either case), then every shift's 2-bit
// Hamming distance (case-insensitive) gives the first best shift and the
// second-best count (scan_best_with_second_raw, bitpacked.rs:153-199), and
// lane 0 certifies the hit on bytes.  out[3 i ..] = certified, shift,
// mismatches.
__global__ __launch_bounds__(64) void spectral_kernel(
    int32_t n, const uint8_t *__restrict__ reads, const int64_t *__restrict__ roff,
    const int32_t *__restrict__ rlen, const uint8_t *__restrict__ gen,
    const int64_t *__restrict__ toff, const int32_t *__restrict__ tlen, int32_t pct,
    int32_t floor_, int32_t strict, int32_t *__restrict__ out) {
    extern __shared__ uint8_t sbuf[];
    const int32_t i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const int32_t L = rlen[i], T = tlen[i];
    uint8_t *R = sbuf, *X = sbuf + ((L + 7) & ~3);
    const uint8_t *rs = reads + roff[i], *ts = gen + toff[i];
    bool bad = false;
    for (int32_t k = lane; k < L + T; k += 64) {
        const uint32_t b = k < L ? rs[k] : ts[k - L];
        if (k < L)
            R[k] = (uint8_t)b;
        else
            X[k - L] = (uint8_t)b;
        bad |= !packable(b);
    }
    // pad the tails so four-byte reads past the ends stay inside the buffer
    if (lane < 4) {
        R[L + lane] = 0;
        X[T + lane] = 0;
    }
    const bool any_bad = __ballot(bad) != 0;
    __syncthreads();
    if (any_bad || L == 0 || T < L) {
        if (lane == 0) out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0;
        return;
    }
    // per lane: best (count, first shift) and second-best count
    int32_t bm = INT32_MAX, bs = INT32_MAX, sm = INT32_MAX;
    for (int32_t t = lane; t <= T - L; t += 64) {
        int32_t c = 0;
        for (int32_t q = 0; q < L; q += 4) {
            uint32_t a = 0, b = 0;
#pragma unroll
            for (int z = 0; z < 4; z++) {
                a |= (uint32_t)R[q + z] << (8 * z);
                b |= (uint32_t)X[t + q + z] << (8 * z);
            }
            uint32_t x = (a ^ b) & 0xDFDFDFDFu;
            const int32_t rem = L - q;
            if (rem < 4) x &= (1u << (8 * rem)) - 1u;
            x |= x >> 4;
            x |= x >> 2;
            x |= x >> 1;
            c += __builtin_popcount(x & 0x01010101u);
        }
        if (c < bm) {
            sm = bm;
            bm = c;
            bs = t;
        } else if (c < sm) {
            sm = c;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t obm = __shfl_xor(bm, o, 64), obs = __shfl_xor(bs, o, 64),
                      osm = __shfl_xor(sm, o, 64);
        const bool mine = bm < obm || (bm == obm && bs < obs);
        const int32_t nsm = mine ? min(sm, obm) : min(osm, bm);
        if (!mine) {
            bm = obm;
            bs = obs;
        }
        sm = nsm;
    }
    if (lane == 0) {
        const int32_t max_mism = max(L * pct / 100, floor_);
        const bool ok = certify_bytes(R, X + bs, L, bm, sm == INT32_MAX ? -1 : sm, max_mism,
                                      strict != 0);
        out[3 * i] = ok ? 1 : 0;
        out[3 * i + 1] = bs;
        out[3 * i + 2] = bm;
    }
}

int64_t common_prefix_len(const uint8_t *a, int64_t an, const uint8_t *b, int64_t bn) {
    const int64_t n = std::min(an, bn);
    int64_t i = 0;
    while (i < n && a[i] == b[i]) i++;
    return i;
}

// clipping_can_improve (mod.rs:442-479)
bool clipping_can_improve(const uint8_t *read, int32_t L, const uint8_t *text, int64_t tn,
                          int32_t text_start, const uint32_t *ops, int32_t nops,
                          const kira_sw_cfg &c) {
    int64_t qi = 0, ti = text_start;
    int32_t score = 0, min_prefix = 0, best_prefix = 0;
    for (int32_t k = 0; k < nops; k++) {
        const uint32_t op = ops[k] & 15u, len = ops[k] >> 4;
        if (op == 0) {
            for (uint32_t z = 0; z < len; z++) {
                const bool same = qi < L && ti < tn && read[qi] == text[ti];
                score += same ? c.match_score : -c.mismatch;
                qi++;
                ti++;
                min_prefix = std::min(min_prefix, score);
                best_prefix = std::max(best_prefix, score);
            }
        } else if (op == 1 || op == 2) {
            score -= c.gap_open + c.gap_extend * (int32_t)len;
            if (op == 1)
                qi += len;
            else
                ti += len;
            min_prefix = std::min(min_prefix, score);
        } else {
            return true;
        }
    }
    const int64_t lim = (int64_t)score + c.clip_penalty;
    return min_prefix < 0 || best_prefix > std::min<int64_t>(lim, INT32_MAX);
}

struct Job {
    int32_t k;           // chain index
    int32_t contig;
    int64_t ws;          // window start in the contig
    int32_t L, tl;       // read and window lengths
    int32_t stage;       // 0 none, 1 certified, 2 WFA
    int32_t shift, mism; // certificate
    // final WFA
    int32_t loss, text_start, text_end, n_ops;
    const uint32_t *ops; // the CIGAR of the alignment kept
    int64_t ftext;       // final text: offset from genome base 0
    int32_t ftext_len;
    int32_t first_loss;
    int32_t retry;       // index into the retry list, or -1
};

// number of positions where a[k] == b[k], k < n (eight at a time: a zero
// byte of a ^ b is an equal position)
inline uint32_t count_equal_bytes(const uint8_t *a, const uint8_t *b, uint32_t n) {
    uint32_t eq = 0, t = 0;
    for (; t + 8 <= n; t += 8) {
        uint64_t x, y;
        memcpy(&x, a + t, 8);
        memcpy(&y, b + t, 8);
        const uint64_t z = x ^ y;
        const uint64_t u = ((z & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | z;
        eq += (uint32_t)__builtin_popcountll(~u & 0x8080808080808080ull);
    }
    for (; t < n; t++) eq += a[t] == b[t] ? 1u : 0u;
    return eq;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

namespace kira {

kira_status fast_dp_run(kira_ctx *ctx, const kira_genome *g, const int64_t *contig_off,
                        const int64_t *contig_len, int32_t n, const uint8_t *reads,
                        int64_t reads_bytes, const uint8_t *d_reads_in, const int64_t *read_off,
                        const int32_t *read_len, const kira_chain *chains,
                        const kira_sw_cfg *cfg, const kira_fast_knobs &kn_in,
                        PodVec<kira_fast_aln> &res_out, std::vector<uint32_t> &cig_out,
                        std::string &md_out, kira_fast_stats &st_local, bool want_md) {
    memset(&st_local, 0, sizeof st_local);
    try {
        res_out.resize((size_t)std::max(n, 0));
        cig_out.clear();
        md_out.clear();
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    parallel_chunks(n, 16384, [&](int, int64_t b, int64_t e) {
        memset((void *)(res_out.data() + b), 0, (size_t)(e - b) * sizeof(kira_fast_aln));
    });
    if (n == 0) return KIRA_OK;
    // parse_algo (router.rs:28-37): anything unrecognised is packed; the
    // percentages and floors are u32 knobs
    kira_fast_knobs kn = kn_in;
    if (kn.algo < 0 || kn.algo > 3) kn.algo = 0;
    kn.quality_pct = std::max(kn.quality_pct, 0);
    kn.wfa_budget_pct = std::max(kn.wfa_budget_pct, 0);
    kn.myers_pct = std::max(kn.myers_pct, 0);
    kn.myers_floor = std::max(kn.myers_floor, 0);
    kn.spectral_pct = std::max(kn.spectral_pct, 0);
    kn.spectral_floor = std::max(kn.spectral_floor, 0);
    kn.wfa_ends_free = std::max(kn.wfa_ends_free, 0);
    const kira_sw_cfg c = *cfg;
    const int32_t band = std::max(c.bandwidth, 1);
    kira::WfaPen pen;
    const bool pen_ok = kira::wfa_pen(c, pen);
    const uint8_t *gh = g->host.data();
    // ---- router gates and windows (host)
    PodVec<Job> jobs;
    PodVec<kira_fast_aln> &res = res_out;
    try {
        // the chains that pass, in chain order: each range gates its chains
        // and counts, then places them
        auto gate = [&](int32_t k, Job &j) -> bool {
            const int32_t L = read_len[k];
            if (L == 0 || L > kn.wfa_max_read_len || kn.algo == 3) return false;
            if (kn.quality_gate) {  // fast_path_worth_attempting (router.rs:169-189)
                const int32_t cs = chains[k].score;
                if (cs <= 0) return false;
                const uint64_t r = std::min<uint64_t>((uint64_t)cs * 100u, UINT32_MAX) / (uint64_t)L;
                if (r < (uint64_t)std::max(kn.quality_pct, 0)) return false;
            }
            const int64_t ws = (int64_t)chains[k].ref_start - (int64_t)chains[k].read_start;
            const int32_t cid = (int32_t)chains[k].ref_id;
            const int64_t clen = contig_len[cid];
            if (ws < 0 || ws >= clen) return false;
            const int64_t tl = std::min<int64_t>((int64_t)L + band, clen - ws);
            if (tl < L) return false;
            j = Job{};
            j.k = k;
            j.contig = cid;
            j.ws = ws;
            j.L = L;
            j.tl = (int32_t)tl;
            return true;
        };
        const int64_t kGrain = 8192;
        const int T = host_threads();
        const int Tn = (int)std::max<int64_t>(1, std::min<int64_t>(T, ((int64_t)n + kGrain - 1) / kGrain));
        std::vector<int64_t> cnt((size_t)T + 1, 0);
        parallel_chunks(n, kGrain, [&](int t, int64_t b, int64_t e) {
            Job j;
            int64_t c0 = 0;
            for (int64_t k = b; k < e; k++) c0 += gate((int32_t)k, j) ? 1 : 0;
            cnt[(size_t)t + 1] = c0;
        });
        for (int t = 0; t < Tn; t++) cnt[(size_t)t + 1] += cnt[(size_t)t];
        jobs.resize((size_t)cnt[(size_t)Tn]);
        parallel_chunks(n, kGrain, [&](int t, int64_t b, int64_t e) {
            int64_t o = cnt[(size_t)t];
            for (int64_t k = b; k < e; k++)
                if (gate((int32_t)k, jobs[(size_t)o])) o++;
        });
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    st_local.attempted = (int64_t)jobs.size();
    hipError_t e = hipSetDevice(ctx->device);
    if (e) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    hipStream_t st = ctx->stream;
    DevMem dm(ctx);
    const uint8_t *d_reads = d_reads_in;
    if (!d_reads) {
        uint8_t *d = nullptr;
        if ((e = dm.alloc(&d, reads_bytes)))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path allocation", e);
        if (reads_bytes && (e = hipMemcpyAsync(d, reads, reads_bytes, hipMemcpyHostToDevice, st)))
            return fail(ctx, KIRA_ERR_DRIVER, "fast path upload", e);
        d_reads = d;
    }
    // one launch's job lists (read offset, read len, text offset, text len,
    // two extras), written straight into pinned buffer 3 of the context side
    // by side and uploaded with one copy; every phase has synchronised before
    // the next one rewrites them
    struct Lists {
        int64_t *roff = nullptr, *toff = nullptr;
        int32_t *rlen = nullptr, *tlen = nullptr, *x = nullptr, *y = nullptr;
        uint8_t *base = nullptr;
        size_t n = 0, bytes = 0, a8 = 0, a4 = 0;
        hipError_t begin(kira_ctx *cx, size_t m) {
            a8 = (m * 8 + 255) & ~(size_t)255;
            a4 = (m * 4 + 255) & ~(size_t)255;
            bytes = 2 * a8 + 4 * a4;
            void *p = nullptr;
            hipError_t r = kira::pinned_buf(cx, kira::kPinFast, bytes, &p);
            if (r != hipSuccess) return r;
            base = (uint8_t *)p;
            roff = (int64_t *)base;
            toff = (int64_t *)(base + a8);
            rlen = (int32_t *)(base + 2 * a8);
            tlen = (int32_t *)(base + 2 * a8 + a4);
            x = (int32_t *)(base + 2 * a8 + 2 * a4);
            y = (int32_t *)(base + 2 * a8 + 3 * a4);
            n = 0;
            return hipSuccess;
        }
    } lists;
    struct DevLists {
        int64_t *roff = nullptr, *toff = nullptr;
        int32_t *rlen = nullptr, *tlen = nullptr, *x = nullptr, *y = nullptr;
    };
    auto upload = [&](DevLists &d) -> hipError_t {
        uint8_t *db = nullptr;
        hipError_t r;
        if ((r = dm.alloc(&db, lists.bytes))) return r;
        d.roff = (int64_t *)db;
        d.toff = (int64_t *)(db + lists.a8);
        d.rlen = (int32_t *)(db + 2 * lists.a8);
        d.tlen = (int32_t *)(db + 2 * lists.a8 + lists.a4);
        d.x = (int32_t *)(db + 2 * lists.a8 + 2 * lists.a4);
        d.y = (int32_t *)(db + 2 * lists.a8 + 3 * lists.a4);
        if (lists.n == 0) return hipSuccess;
        return hipMemcpyAsync(db, lists.base, lists.bytes, hipMemcpyHostToDevice, st);
    };
    auto text_of = [&](const Job &j) { return contig_off[j.contig] + j.ws; };
    try {
        // ---- packed spectral certificate
        auto t_sp = std::chrono::steady_clock::now();
        std::vector<int32_t> cand;
        if (kn.algo == 0 || kn.algo == 1) {
            if ((e = lists.begin(ctx, jobs.size())))
                return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path staging", e);
            int32_t max_lt = 0;
            const int64_t nq = (int64_t)jobs.size();
            std::vector<int32_t> part_max((size_t)host_threads(), 0);
            parallel_chunks(nq, 16384, [&](int t, int64_t b, int64_t e) {
                int32_t mx = 0;
                for (int64_t q = b; q < e; q++) {
                    const Job &j = jobs[q];
                    lists.roff[q] = read_off[j.k];
                    lists.rlen[q] = j.L;
                    lists.toff[q] = text_of(j);
                    lists.tlen[q] = j.tl;
                    lists.x[q] = 0;
                    lists.y[q] = 0;
                    mx = std::max(mx, j.L + j.tl);
                }
                part_max[(size_t)t] = mx;
            });
            lists.n = (size_t)nq;
            for (int32_t v : part_max) max_lt = std::max(max_lt, v);
            const int32_t m = (int32_t)lists.n;
            trace_mark(".fp_sp_lists");
            if (m) {
                const size_t lds = (size_t)max_lt + 16;
                if (lds > 65536)
                    return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "fast path window too long");
                DevLists d;
                int32_t *d_out = nullptr;
                if ((e = upload(d)) || (e = dm.alloc(&d_out, 3 * (size_t)m)))
                    return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path spectral", e);
                hipLaunchKernelGGL(spectral_kernel, dim3(m), dim3(64), lds, st, m, d_reads, d.roff,
                                   d.rlen, g->ascii, d.toff, d.tlen, kn.spectral_pct,
                                   kn.spectral_floor, kn.cert_strict, d_out);
                trace_mark(".fp_sp_launch");
                std::vector<int32_t> so(3 * (size_t)m);
                if ((e = hipGetLastError()) ||
                    (e = hipMemcpyAsync(so.data(), d_out, 4 * so.size(), hipMemcpyDeviceToHost, st)) ||
                    (e = hipStreamSynchronize(st)))
                    return fail(ctx, KIRA_ERR_RUNTIME, "fast path spectral", e);
                trace_mark(".fp_sp_sync");
                parallel_chunks(m, 16384, [&](int, int64_t b, int64_t e0) {
                    for (int64_t q = b; q < e0; q++)
                        if (so[3 * q]) {
                            jobs[q].stage = 1;
                            jobs[q].shift = so[3 * q + 1];
                            jobs[q].mism = so[3 * q + 2];
                        }
                });
            }
        }
        parallel_filter((int64_t)jobs.size(), 16384,
                        [&](int64_t q) { return jobs[q].stage == 0; }, cand);
        st_local.phase_ms[0] = ms_since(t_sp);
        trace_mark(".fp_spectral");
        auto t_my = std::chrono::steady_clock::now();
        // ---- Myers reject (router.rs:196-202)
        std::vector<int32_t> wfa_jobs;
        if (!cand.empty() && pen_ok) {
            if ((e = lists.begin(ctx, cand.size())))
                return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path staging", e);
            const int32_t m = (int32_t)cand.size();
            parallel_chunks(m, 16384, [&](int, int64_t b, int64_t e0) {
                for (int64_t i = b; i < e0; i++) {
                    const Job &j = jobs[cand[i]];
                    lists.roff[i] = read_off[j.k];
                    lists.rlen[i] = j.L;
                    lists.toff[i] = text_of(j);
                    lists.tlen[i] = j.tl;
                    lists.x[i] = (int32_t)((int64_t)j.L * kn.myers_pct / 100 + kn.myers_floor);
                    lists.y[i] = 0;
                }
            });
            lists.n = (size_t)m;
            DevLists d;
            uint8_t *d_found = nullptr;
            int32_t *d_dist = nullptr, *d_end = nullptr;
            if ((e = upload(d)) || (e = dm.alloc(&d_found, m)) || (e = dm.alloc(&d_dist, m)) ||
                (e = dm.alloc(&d_end, m)))
                return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path Myers", e);
            kira_status rs = kira::myers_dev(ctx, m, d_reads, d.roff, d.rlen, lists.rlen,
                                             g->ascii, d.toff, d.tlen, d.x, d_found, d_dist, d_end);
            if (rs != KIRA_OK) return rs;
            std::vector<uint8_t> found(m);
            if ((e = hipMemcpyAsync(found.data(), d_found, m, hipMemcpyDeviceToHost, st)) ||
                (e = hipStreamSynchronize(st)))
                return fail(ctx, KIRA_ERR_RUNTIME, "fast path Myers", e);
            parallel_filter(m, 16384, [&](int64_t i) { return found[i] != 0; }, wfa_jobs);
            parallel_chunks((int64_t)wfa_jobs.size(), 16384, [&](int, int64_t b, int64_t e0) {
                for (int64_t i = b; i < e0; i++) wfa_jobs[i] = cand[wfa_jobs[i]];
            });
            st_local.myers_pass = (int64_t)wfa_jobs.size();
        }
        st_local.phase_ms[1] = ms_since(t_my);
        trace_mark(".fp_myers");
        // ---- scored WFA (wfa_score_budget, router.rs:160-166).  The records
        // are built while the device runs: the certificates' during the main
        // pass, the WFA alignments that need no retry during the retry pass,
        // the retried ones last.
        auto t_wfa = std::chrono::steady_clock::now();
        const int32_t drop = kn.wfa_adaptive > 0 ? kn.wfa_adaptive : 0;
        auto launch_wfa = [&](const std::vector<int32_t> &list, const std::vector<int64_t> &tofs,
                              const std::vector<int32_t> &tlens,
                              const std::vector<int32_t> &budget, const std::vector<int32_t> &tbf,
                              kira::WfaRun &run) -> kira_status {
            if ((e = lists.begin(ctx, list.size())))
                return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path staging", e);
            parallel_chunks((int64_t)list.size(), 16384, [&](int, int64_t b, int64_t e0) {
                for (int64_t i = b; i < e0; i++) {
                    const Job &j = jobs[list[i]];
                    lists.roff[i] = read_off[j.k];
                    lists.rlen[i] = j.L;
                    lists.toff[i] = tofs[i];
                    lists.tlen[i] = tlens[i];
                    lists.x[i] = budget[i];
                    lists.y[i] = tbf[i];
                }
            });
            lists.n = list.size();
            trace_mark(".wfa_lists");
            DevLists d;
            if ((e = upload(d))) return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "fast path WFA", e);
            kira::WfaDevJobs wj;
            wj.n = (int32_t)list.size();
            wj.pat = d_reads;
            wj.txt = g->ascii;
            wj.pat_off = d.roff;
            wj.txt_off = d.toff;
            wj.pat_len = d.rlen;
            wj.txt_len = d.tlen;
            wj.max_score = d.x;
            wj.tbf = d.y;
            wj.h_pat_len = lists.rlen;
            wj.h_txt_len = lists.tlen;
            wj.h_max_score = lists.x;
            wj.h_tbf = lists.y;
            return run.launch(wj, pen, drop, true);
        };
        // ---- records (host): partial-read check, wfa_result_to_alignment;
        // per phase (0 certificates, 1 WFA without retry, 2 retried WFA) one
        // part per thread over the same ranges of jobs
        const int T = host_threads();
        struct Part {
            std::vector<uint32_t> cig;
            std::string md;
            int64_t spectral = 0, wfa = 0, declined = 0;
        };
        std::vector<Part> parts((size_t)3 * T);
        const int64_t nj = (int64_t)jobs.size();
        std::vector<int64_t> bounds(T + 1, nj);
        for (int t = 0; t <= T; t++) bounds[t] = nj * t / T;
        auto phase_of = [](const Job &j) -> int {
            return j.stage == 1 ? 0 : (j.stage == 2 ? (j.retry < 0 ? 1 : 2) : -1);
        };
        auto record = [&](int64_t q, Part &P) {
            const Job &j = jobs[q];
            kira_fast_aln &a = res[j.k];
            const uint8_t *rd = reads + read_off[j.k];
            if (j.stage == 1) {
                const uint8_t *ref = gh + contig_off[j.contig] + j.ws + j.shift;
                a.kind = 1;
                a.ref_start = (uint32_t)(j.ws + j.shift);
                a.ref_end = a.ref_start + (uint32_t)j.L;
                a.read_start = 0;
                a.read_end = (uint32_t)j.L;
                a.score = (j.L - j.mism) * c.match_score - j.mism * c.mismatch;
                a.cigar_off = (int64_t)P.cig.size();
                a.n_cigar = 1;
                P.cig.push_back((uint32_t)j.L << 4);
                a.md_off = (int64_t)P.md.size();
                if (want_md) a.nm = kira::ungapped_tags(rd, ref, j.L, P.md);
                a.md_len = (int32_t)((int64_t)P.md.size() - a.md_off);
                P.spectral++;
                return;
            }
            if (j.stage != 2) return;
            const uint8_t *text = gh + j.ftext;
            const uint32_t *ops = j.ops;
            const int64_t perfect = std::min<int64_t>(
                std::max<int64_t>((int64_t)c.match_score * j.L, INT32_MIN), INT32_MAX);
            if ((int64_t)j.loss * 10 > perfect * 3 &&
                clipping_can_improve(rd, j.L, text, j.ftext_len, j.text_start, ops,
                                     j.n_ops, c)) {
                P.declined++;
                return;
            }
            int64_t cref = 0, cq = 0;
            for (int32_t z = 0; z < j.n_ops; z++) {
                const uint32_t op = ops[z] & 15u, len = ops[z] >> 4;
                if (op == 0 || op == 2 || op == 3) cref += len;
                if (op == 0 || op == 1 || op == 4) cq += len;
            }
            if (cq != j.L || j.text_start + cref > j.ftext_len || j.text_end > j.ftext_len) {
                P.declined++;
                return;
            }
            int32_t score = 0;
            int64_t qi = 0, ti = j.text_start;
            for (int32_t z = 0; z < j.n_ops; z++) {
                const uint32_t op = ops[z] & 15u, len = ops[z] >> 4;
                if (op == 0) {
                    // the per-column sum (wrapping like the reference's i32)
                    const uint32_t eq = count_equal_bytes(rd + qi, text + ti, len);
                    score = (int32_t)((uint32_t)score + eq * (uint32_t)c.match_score -
                                      (len - eq) * (uint32_t)c.mismatch);
                    qi += len;
                    ti += len;
                } else if (op == 1 || op == 2) {
                    score -= c.gap_open + c.gap_extend * (int32_t)len;
                    if (op == 1)
                        qi += len;
                    else
                        ti += len;
                }
            }
            const int64_t fstart = j.ftext - contig_off[j.contig];
            a.kind = 2;
            a.ref_start = (uint32_t)(fstart + j.text_start);
            a.ref_end = (uint32_t)(fstart + j.text_end);
            a.read_start = 0;
            a.read_end = (uint32_t)j.L;
            a.score = score;
            a.cigar_off = (int64_t)P.cig.size();
            a.n_cigar = j.n_ops;
            P.cig.insert(P.cig.end(), ops, ops + j.n_ops);
            a.md_off = (int64_t)P.md.size();
            if (want_md)
                a.nm = kira::nm_md(rd, j.L, text, j.ftext_len, j.text_start, ops, j.n_ops,
                                   P.md);
            a.md_len = (int32_t)((int64_t)P.md.size() - a.md_off);
            P.wfa++;
        };
        double rec_ms = 0;
        auto build = [&](int ph) {
            const auto t0 = std::chrono::steady_clock::now();
            parallel_chunks(T, 1, [&](int, int64_t t0_, int64_t t1_) {
                for (int64_t t = t0_; t < t1_; t++)
                    for (int64_t q = bounds[t]; q < bounds[t + 1]; q++)
                        if (phase_of(jobs[q]) == ph) record(q, parts[(size_t)ph * T + t]);
            });
            rec_ms += ms_since(t0);
        };
        // the CIGARs of both passes stay alive until the records are built
        std::vector<int32_t> wres1, wres2;
        std::vector<uint32_t> wcig1, wcig2;
        std::vector<int64_t> woff1, woff2;
        std::vector<int32_t> retry_jobs, retry_len, retry_free;
        std::vector<int64_t> retry_text;
        kira::WfaRun main_run(ctx), retry_run(ctx);
        if (!wfa_jobs.empty()) {
            const size_t m = wfa_jobs.size();
            std::vector<int64_t> tofs(m);
            std::vector<int32_t> tlens(m), budget(m), tbf(m);
            const int64_t xb = (int64_t)c.match_score + c.mismatch;
            const int64_t ieb = (int64_t)c.match_score + c.gap_extend;
            parallel_chunks((int64_t)m, 16384, [&](int, int64_t b0, int64_t e0) {
            for (int64_t i = b0; i < e0; i++) {
                Job &j = jobs[wfa_jobs[i]];
                const int64_t mb = ((int64_t)j.L * kn.wfa_budget_pct) / 100 * xb;
                const int64_t ib = (c.gap_open + ieb * 4) * 3;
                budget[i] = (int32_t)std::min<int64_t>(std::max(mb, ib), INT32_MAX);
                const int64_t lead = std::min<int64_t>({(int64_t)std::max(kn.wfa_lead, 0), j.ws,
                                                        (int64_t)band});
                tofs[i] = text_of(j) - lead;
                tlens[i] = (int32_t)(j.tl + lead);
                tbf[i] = lead > 0 ? (int32_t)lead : kn.wfa_ends_free;
                j.ftext = tofs[i];
                j.ftext_len = tlens[i];
            }
            });
            kira_status rs = launch_wfa(wfa_jobs, tofs, tlens, budget, tbf, main_run);
            if (rs != KIRA_OK) return rs;
        }
        build(0);  // the certificates, while the main pass runs
        if (!wfa_jobs.empty()) {
            const size_t m = wfa_jobs.size();
            kira_status rs = main_run.finish(wres1, wcig1, woff1);
            if (rs != KIRA_OK) return rs;
            std::vector<uint8_t> is_suspect(m, 0);
            parallel_chunks((int64_t)m, 4096, [&](int, int64_t b0, int64_t e0) {
                for (int64_t i = b0; i < e0; i++) {
                    Job &j = jobs[wfa_jobs[i]];
                    const int32_t *r = &wres1[i * kira::kWfaRes];
                    j.retry = -1;
                    if (!r[0]) continue;
                    j.stage = 2;
                    j.loss = j.first_loss = r[1];
                    j.text_start = r[2];
                    j.text_end = r[3];
                    j.n_ops = r[4];
                    j.ops = wcig1.data() + woff1[i];
                    // suspect origin (mod.rs:373-383): a leading gap, or a
                    // short leading match followed by more ops
                    const uint32_t f = j.n_ops ? j.ops[0] : 0;
                    const bool suspect =
                        j.n_ops > 0 && ((f & 15u) != 0 || ((f >> 4) < 19 && j.n_ops > 1));
                    const int64_t lead = std::min<int64_t>(
                        {(int64_t)std::max(kn.wfa_lead, 0), j.ws, (int64_t)band});
                    if (!(suspect && lead == 0 && kn.wfa_ends_free == 0)) continue;
                    if (std::min<int64_t>({(int64_t)band, 32, j.ws}) <= 0) continue;
                    is_suspect[i] = 1;
                }
            });
            // the suspects in pass order; each job knows its retry slot
            std::vector<int32_t> suspects;
            parallel_filter((int64_t)m, 4096, [&](int64_t i) { return is_suspect[i] != 0; },
                            suspects);
            parallel_chunks((int64_t)suspects.size(), 4096, [&](int, int64_t b0, int64_t e0) {
                for (int64_t k = b0; k < e0; k++) {
                    suspects[k] = wfa_jobs[suspects[k]];
                    jobs[suspects[k]].retry = (int32_t)k;
                }
            });
            trace_mark(".wfa_suspects");
            // the origin probe (mod.rs:389-399): the longest exact read prefix
            // among the starts within the slack
            const size_t ns = suspects.size();
            retry_jobs = suspects;
            retry_text.resize(ns);
            retry_len.resize(ns);
            retry_free.resize(ns);
            parallel_chunks((int64_t)ns, 512, [&](int, int64_t b0, int64_t e0) {
                for (int64_t i = b0; i < e0; i++) {
                    const Job &j = jobs[suspects[i]];
                    const int64_t slack = std::min<int64_t>({(int64_t)band, 32, j.ws});
                    const uint8_t *contig = gh + contig_off[j.contig];
                    const int64_t clen = contig_len[j.contig];
                    const uint8_t *rd = reads + read_off[j.k];
                    int64_t origin = j.ws, prefix = 0;
                    const int64_t hi = std::min<int64_t>(j.ws + slack, clen - 1);
                    for (int64_t s0 = j.ws - slack; s0 <= hi; s0++) {
                        const int64_t mlen = common_prefix_len(rd, j.L, contig + s0, clen - s0);
                        if (mlen > prefix) {
                            prefix = mlen;
                            origin = s0;
                        }
                    }
                    if (prefix >= 4) {
                        const int64_t oend = std::min<int64_t>(origin + j.L + band, clen);
                        retry_text[i] = contig_off[j.contig] + origin;
                        retry_len[i] = (int32_t)(oend - origin);
                        retry_free[i] = 0;
                    } else {
                        retry_text[i] = contig_off[j.contig] + j.ws - slack;
                        retry_len[i] = (int32_t)(j.tl + slack);
                        retry_free[i] = (int32_t)(2 * slack);
                    }
                }
            });
            trace_mark(".wfa_probe");
        }
        // ---- the bounded ends-free retry (mod.rs:384-426)
        if (!retry_jobs.empty()) {
            const size_t m = retry_jobs.size();
            std::vector<int32_t> budget(m);
            for (size_t i = 0; i < m; i++) budget[i] = jobs[retry_jobs[i]].first_loss;
            kira_status rs = launch_wfa(retry_jobs, retry_text, retry_len, budget, retry_free,
                                        retry_run);
            if (rs != KIRA_OK) return rs;
        }
        build(1);  // the WFA alignments that need no retry, while the retry pass runs
        if (!retry_jobs.empty()) {
            const size_t m = retry_jobs.size();
            kira_status rs = retry_run.finish(wres2, wcig2, woff2);
            if (rs != KIRA_OK) return rs;
            for (size_t i = 0; i < m; i++) {
                Job &j = jobs[retry_jobs[i]];
                const int32_t *r = &wres2[i * kira::kWfaRes];
                if (!r[0] || r[1] >= j.loss) continue;
                st_local.origin_retries++;
                j.loss = r[1];
                j.text_start = r[2];
                j.text_end = r[3];
                j.n_ops = r[4];
                j.ops = wcig2.data() + woff2[i];
                j.ftext = retry_text[i];
                j.ftext_len = retry_len[i];
            }
        }
        trace_mark(".fp_wfa");
        build(2);
        st_local.phase_ms[2] = ms_since(t_wfa) - rec_ms;
        // place each part's CIGAR words and MD text after the previous parts'
        int64_t cig_total = 0, md_total = 0;
        std::vector<int64_t> cig_base((size_t)3 * T), md_base((size_t)3 * T);
        for (size_t x = 0; x < parts.size(); x++) {
            cig_base[x] = cig_total;
            md_base[x] = md_total;
            cig_total += (int64_t)parts[x].cig.size();
            md_total += (int64_t)parts[x].md.size();
            st_local.spectral += parts[x].spectral;
            st_local.wfa += parts[x].wfa;
            st_local.clip_declined += parts[x].declined;
        }
        cig_out.resize((size_t)cig_total);
        md_out.resize((size_t)md_total);
        parallel_chunks(T, 1, [&](int, int64_t t0_, int64_t t1_) {
            for (int64_t t = t0_; t < t1_; t++) {
                for (int64_t q = bounds[t]; q < bounds[t + 1]; q++) {
                    kira_fast_aln &a = res[jobs[q].k];
                    const int ph = phase_of(jobs[q]);
                    if (a.kind == 0 || ph < 0) continue;
                    a.cigar_off += cig_base[(size_t)ph * T + t];
                    a.md_off += md_base[(size_t)ph * T + t];
                }
                for (int ph = 0; ph < 3; ph++) {
                    const Part &P = parts[(size_t)ph * T + t];
                    if (!P.cig.empty())
                        memcpy(cig_out.data() + cig_base[(size_t)ph * T + t], P.cig.data(),
                               4 * P.cig.size());
                    if (!P.md.empty())
                        memcpy(&md_out[(size_t)md_base[(size_t)ph * T + t]], P.md.data(),
                               P.md.size());
                }
            }
        });
        // host time of building the records (the first two phases overlap
        // the WFA kernels)
        st_local.phase_ms[3] = rec_ms;
        trace_mark(".fp_records");
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    return KIRA_OK;
}

}  // namespace kira

extern "C" kira_status kira_fast_dp_batch(
    kira_ctx *ctx, const kira_genome *g, int32_t n_contigs, const int64_t *contig_off,
    const int64_t *contig_len, int32_t n, const uint8_t *reads, int64_t reads_bytes,
    const int64_t *read_off, const int32_t *read_len, const kira_chain *chains,
    const kira_sw_cfg *cfg, const kira_fast_knobs *knobs_in, kira_fast_aln *out,
    uint32_t *cigar_out, int64_t cigar_cap, char *md_out, int64_t md_cap, int64_t *cigar_used,
    int64_t *md_used, kira_fast_stats *stats) {
    if (!ctx || !g || !cfg || n < 0 || n_contigs < 0 || reads_bytes < 0)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    kira_fast_knobs kn;
    kira_fast_knobs_default(&kn);
    if (knobs_in) kn = *knobs_in;
    kira_fast_stats st;
    memset(&st, 0, sizeof st);
    if (cigar_used) *cigar_used = 0;
    if (md_used) *md_used = 0;
    if (n == 0) {
        if (stats) *stats = st;
        return KIRA_OK;
    }
    if (!read_off || !read_len || !chains || !out || (reads_bytes && !reads) ||
        (n_contigs && (!contig_off || !contig_len)))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    const int64_t glen = (int64_t)g->host.size();
    for (int32_t c = 0; c < n_contigs; c++)
        if (contig_off[c] < 0 || contig_len[c] < 0 || contig_off[c] > glen ||
            contig_len[c] > glen - contig_off[c])
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "contig outside the genome");
    for (int32_t k = 0; k < n; k++) {
        if (read_off[k] < 0 || read_len[k] < 0 || read_off[k] > reads_bytes ||
            read_len[k] > reads_bytes - read_off[k])
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "read outside its buffer");
        if (chains[k].ref_id >= (uint32_t)n_contigs)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "chain contig out of range");
    }
    kira::PodVec<kira_fast_aln> res;
    std::vector<uint32_t> cig;
    std::string md;
    kira_status rs = kira::fast_dp_run(ctx, g, contig_off, contig_len, n, reads, reads_bytes,
                                       nullptr, read_off, read_len, chains, cfg, kn, res, cig, md,
                                       st, true);
    if (rs != KIRA_OK) return rs;
    if (cigar_used) *cigar_used = (int64_t)cig.size();
    if (md_used) *md_used = (int64_t)md.size();
    if (stats) *stats = st;
    if ((int64_t)cig.size() > cigar_cap || (int64_t)md.size() > md_cap ||
        (!cig.empty() && !cigar_out) || (!md.empty() && !md_out))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "CIGAR or MD buffer too small");
    if (!cig.empty()) memcpy(cigar_out, cig.data(), 4 * cig.size());
    if (!md.empty()) memcpy(md_out, md.data(), md.size());
    memcpy(out, res.data(), sizeof(kira_fast_aln) * (size_t)n);
    return KIRA_OK;
}

extern "C" void kira_fast_knobs_default(kira_fast_knobs *k) {
    if (!k) return;
    *k = kira_fast_knobs{300, 0, 0, 70, 25, 15, 4, 3, 2, 0, 0, 0, 0};
}
