// Stage-4 alignment driver: src/pipeline/stage4_alignment.rs run() (:959-1268)
// with process_read_prefilter (:535-957), the GPU dispatcher active or not,
// single-end reads, the two-tier candidate reorder on or off.  Chain selection,
// job routing and the record text (NM / MD) are host work; exact matching,
// the ungapped prefilter, the 2-bit sieve with certification, the lane-group
// SW and the banded SW run on the device.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "internal.hpp"

struct kira_stage4_result {
    std::vector<int64_t> aln_off;
    kira::PodVec<kira_stage4_aln> alns;
    kira::PodVec<uint32_t> cigar;
    std::string md;
    kira_stage4_stats stats{};
    double phase_ms[5] = {};
};

namespace {

constexpr uint32_t kOpM = 0, kOpI = 1, kOpS = 4;

// split_read.rs:10-19
constexpr float kSuppMinScoreRatio = 0.5f;
constexpr uint32_t kSuppMinReadSpan = 50;
constexpr uint32_t kSuppMaxOverlapPct = 50;
constexpr size_t kMaxSuppPerRead = 4;

// exact_match_exists (mod.rs:818-832) on candidate chains whose spans
// already cover the whole read: one wave per chain compares the oriented
// read with the contig bytes, 64 at a time.
__global__ void exact_match_kernel(int32_t n, const uint8_t *__restrict__ reads,
                                   const int64_t *__restrict__ read_off,
                                   const int32_t *__restrict__ read_len,
                                   const uint8_t *__restrict__ genome,
                                   const int64_t *__restrict__ ref_pos,
                                   uint8_t *__restrict__ out) {
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= n) return;  // whole wave
    const uint8_t *q = reads + read_off[w];
    const uint8_t *r = genome + ref_pos[w];
    const int32_t L = read_len[w];
    bool equal = true;
    for (int32_t base = 0; base < L; base += 64) {  // wave-uniform trip count
        const int32_t i = base + lane;
        const bool d = i < L && q[i] != r[i];
        if (__any(d)) {
            equal = false;
            break;
        }
    }
    if (lane == 0) out[w] = equal ? 1 : 0;
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

kira_status fail(kira_ctx *ctx, kira_status st, const char *what, hipError_t e = hipSuccess) {
    if (ctx) {
        char buf[512];
        if (e != hipSuccess)
            snprintf(buf, sizeof buf, "stage4: %s: %s", what, hipGetErrorString(e));
        else
            snprintf(buf, sizeof buf, "stage4: %s", what);
        ctx->err = buf;
    }
    return st;
}

// device allocations released together
using kira::DevMem;

int32_t sat_sub(int32_t a, int32_t b) {
    int64_t v = (int64_t)a - b;
    return (int32_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(INT32_MIN, v));
}
int32_t wrap_i32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
int32_t sat_mul(int32_t a, int32_t b) {
    int64_t v = (int64_t)a * b;
    return (int32_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(INT32_MIN, v));
}
// `x as i32` of a float (saturating, NaN -> 0)
int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483647.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
uint32_t usub(uint32_t a, uint32_t b) { return a > b ? a - b : 0; }

// identical_base_is_edit (md.rs:134-137)
bool identical_is_edit(uint8_t b) {
    uint8_t u = (b >= 'a' && b <= 'z') ? (uint8_t)(b - 32) : b;
    return !(u == 'A' || u == 'C' || u == 'G' || u == 'T' || u == '=');
}

// src/seq.rs:1-17
uint8_t comp_base(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    default: return 'N';
    }
}

// original_interval / read_overlap_pct (split_read.rs:128-148)
std::pair<uint32_t, uint32_t> original_interval(uint32_t s, uint32_t e, bool rev, uint32_t L) {
    return rev ? std::make_pair(usub(L, e), usub(L, s)) : std::make_pair(s, e);
}
uint32_t read_overlap_pct(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    uint32_t al = usub(a1, a0), bl = usub(b1, b0);
    if (al == 0 || bl == 0) return 0;
    uint32_t s = std::max(a0, b0), e = std::min(a1, b1);
    if (e <= s) return 0;
    uint64_t ov = (uint64_t)(e - s) * 100u;
    return (uint32_t)std::min<uint64_t>(ov, 0xFFFFFFFFu) / al;
}

// The records one thread makes in one build() call, in item order: CIGAR /
// MD offsets are relative to the arena until the result is assembled.
struct Arena {
    std::vector<int32_t> read;  // read of each record
    std::vector<kira_stage4_aln> recs;
    std::vector<uint32_t> cig;
    std::string md;
    void add(int32_t r, const kira_stage4_aln &a) {
        read.push_back(r);
        recs.push_back(a);
    }
};

// A record's place: arena index << 32 | index in the arena.
inline int64_t rec_handle(size_t arena, size_t k) { return ((int64_t)arena << 32) | (int64_t)k; }

using kira::host_threads;
using kira::parallel_chunks;

// a chain sent to the prefilter (and the exact-match check)
struct Sel {
    int32_t read;
    int64_t chain;
    uint8_t top1;
    int32_t margin;
    int32_t cached;  // the two-tier search's verified alignment (index), or -1
};

// an SW job: oriented read, window [win, win + wlen) in contig coordinates
struct Job {
    int32_t read;
    int64_t chain;
    int64_t win;
    int32_t wlen;
    int32_t abort;
};
using JobVec = kira::PodVec<Job>;

struct Driver {
    kira_ctx *ctx;
    const kira_genome *g;
    const int64_t *contig_off, *contig_len;
    int32_t n;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const int64_t *chain_off;
    const kira_chain *chains;
    const kira_stage4_cfg &c;
    kira_stage4_result &R;
    hipStream_t s;

    // forward and reverse-complement copy of every read, back to back
    kira::Pinned<uint8_t> orient;
    std::vector<int64_t> ooff;  // read k: forward at ooff[k], reverse at ooff[k] + L
    uint8_t *d_orient = nullptr;
    DevMem mem;
    // every record so far, arena by arena in record order (never moved: the
    // result is assembled once, at the end)
    std::vector<std::unique_ptr<Arena>> arenas;
    const int32_t *rmo = nullptr;  // ReadRecord::repeat_min_occ per read, may be NULL

    int32_t rmo_of(int32_t r) const { return rmo ? rmo[r] : 0; }

    int64_t oriented_off(int32_t r, bool rev) const { return ooff[r] + (rev ? read_len[r] : 0); }
    const uint8_t *oriented(int32_t r, bool rev) const {
        return orient.data() + oriented_off(r, rev);
    }
    const uint8_t *contig(uint32_t id) const { return g->host.data() + contig_off[id]; }

    // Builds the records of items [0, n) in parallel (f(i, arena) adds zero
    // or more) after every record so far, in item order: thread t's arena
    // holds the t-th contiguous range of items.
    template <class F>
    void build(int64_t n, F &&f) {
        if (n <= 0) return;
        build_at(reserve_arenas(), n, f);
    }
    // Places for one build() call's records after every record so far, to
    // be filled later by build_at (so records can be made out of order).
    size_t reserve_arenas() {
        const size_t a0 = arenas.size();
        for (int t = 0; t < host_threads(); t++) arenas.emplace_back(new Arena());
        return a0;
    }
    template <class F>
    void build_at(size_t a0, int64_t n, F &&f) {
        if (n <= 0) return;
        parallel_chunks(n, 256, [&](int t, int64_t b, int64_t e) {
            Arena &A = *arenas[a0 + (size_t)t];
            // about one record of a few CIGAR ops and MD characters an item
            A.read.reserve((size_t)(e - b));
            A.recs.reserve((size_t)(e - b));
            A.cig.reserve((size_t)(e - b) * 4);
            A.md.reserve((size_t)(e - b) * 16);
            for (int64_t i = b; i < e; i++) f(i, A);
        });
    }

    kira_stage4_aln &rec(int64_t h) {
        return arenas[(size_t)(h >> 32)]->recs[(size_t)(h & 0xFFFFFFFF)];
    }
    const uint32_t *rec_cigar(int64_t h) {
        return arenas[(size_t)(h >> 32)]->cig.data() + rec(h).cigar_off;
    }

    kira_stage4_aln base_aln(int32_t kind, const kira_chain &ch) const {
        kira_stage4_aln a{};
        a.kind = kind;
        a.ref_id = ch.ref_id;
        a.is_rev = ch.strand == 1;
        return a;
    }

    // exact_match_alignment (mod.rs:778-815)
    kira_stage4_aln exact_aln(int32_t r, const kira_chain &ch, Arena &A) const {
        const int32_t L = read_len[r];
        const uint8_t *q = oriented(r, ch.strand == 1);
        kira_stage4_aln a = base_aln(0, ch);
        a.ref_start = ch.ref_start;
        a.ref_end = ch.ref_start + (uint32_t)L;
        a.read_start = 0;
        a.read_end = (uint32_t)L;
        a.score = wrap_i32((int64_t)c.cfg.match_score * L);
        a.mapq = 0;
        uint32_t nm = 0;
        for (int32_t k = 0; k < L; k++) nm += identical_is_edit(q[k]) ? 1 : 0;
        a.nm = nm;
        a.cigar_off = (int64_t)A.cig.size();
        a.n_cigar = 1;
        A.cig.push_back(((uint32_t)L << 4) | kOpM);
        a.md_off = (int64_t)A.md.size();
        A.md += std::to_string(L);
        a.md_len = (int32_t)(A.md.size() - a.md_off);
        return a;
    }

    // build_alignment (mod.rs:1200-1254) of an SW result over window
    // [win, win + wlen) of the chain's contig
    kira_stage4_aln sw_aln(int32_t r, const kira_chain &ch, int64_t win, int32_t wlen,
                           const kira_sw_result &sw, const uint32_t *ops, Arena &A) const {
        kira_stage4_aln a = base_aln(1, ch);
        const int32_t L = read_len[r];
        a.ref_start = (uint32_t)win + sw.ref_start;
        a.ref_end = (uint32_t)win + sw.ref_end;
        a.read_start = (uint32_t)sw.read_start;
        a.read_end = (uint32_t)sw.read_end;
        a.score = sw.score;
        a.mapq = 0;
        a.cigar_off = (int64_t)A.cig.size();
        a.md_off = (int64_t)A.md.size();
        a.nm = kira::build_record(oriented(r, ch.strand == 1), L, contig(ch.ref_id) + win, wlen,
                                  sw, ops, A.cig, A.md);
        a.n_cigar = (int32_t)(A.cig.size() - a.cigar_off);
        a.md_len = (int32_t)(A.md.size() - a.md_off);
        return a;
    }

    bool exact_candidate(int32_t r, const kira_chain &ch) const {
        const uint32_t L = (uint32_t)read_len[r];
        return ch.read_start == 0 && ch.read_end == L && ch.ref_end >= ch.ref_start &&
               ch.ref_end - ch.ref_start == L &&
               (int64_t)ch.ref_start + L <= contig_len[ch.ref_id];
    }

    // Banded SW over device-resident reads and windows (offset: each job's
    // band centre offset), launched on the stream; run.fetch() returns the
    // results and CIGARs in job order.
    kira_status banded_launch(kira::BandedRun &run, const JobVec &jobs,
                              const kira::PodVec<int32_t> &offset, int slot,
                              hipStream_t stream = nullptr) {
        const int32_t m = (int32_t)jobs.size();
        if (m == 0) return KIRA_OK;
        kira::PodVec<int64_t> ro(m), fo(m);
        kira::PodVec<int32_t> rl(m), fl(m), ab(m);
        parallel_chunks(m, 16384, [&](int, int64_t b0, int64_t e0) {
            for (int64_t t = b0; t < e0; t++) {
                const Job &j = jobs[t];
                const kira_chain &ch = chains[j.chain];
                ro[t] = oriented_off(j.read, ch.strand == 1);
                rl[t] = read_len[j.read];
                fo[t] = contig_off[ch.ref_id] + j.win;
                fl[t] = j.wlen;
                ab[t] = j.abort;
            }
        });
        kira::SwJobsHost sj{m,         d_orient,  ro.data(),     rl.data(), g->ascii,
                            fo.data(), fl.data(), offset.data(), nullptr,   ab.data()};
        return run.launch(sj, c.cfg, c.knobs, slot, stream);
    }

    // Groups the records by read without moving them: read r's records are
    // rec(perm[k]) for k in [first[r], first[r + 1]), in record order.
    void group_by_read(std::vector<int64_t> &first, std::vector<int64_t> &perm) const {
        first.assign((size_t)n + 1, 0);
        for (const auto &A : arenas)
            for (int32_t r : A->read) first[r + 1]++;
        for (int32_t r = 0; r < n; r++) first[r + 1] += first[r];
        std::vector<int64_t> pos(first.begin(), first.end() - 1);
        perm.resize((size_t)first[n]);
        for (size_t a = 0; a < arenas.size(); a++) {
            const std::vector<int32_t> &rd = arenas[a]->read;
            for (size_t k = 0; k < rd.size(); k++) perm[pos[rd[k]]++] = rec_handle(a, k);
        }
    }

    // chain.ref_start as i32 - win_start as i32 - chain.read_start as i32
    // (align_oriented_chain_with_meta, mod.rs:867-869)
    kira::PodVec<int32_t> chain_offsets(const JobVec &jobs) const {
        kira::PodVec<int32_t> off(jobs.size());
        parallel_chunks((int64_t)jobs.size(), 16384, [&](int, int64_t b0, int64_t e0) {
            for (int64_t t = b0; t < e0; t++) {
                const kira_chain &ch = chains[jobs[t].chain];
                off[t] = wrap_i32((int64_t)(int32_t)ch.ref_start - (int32_t)(uint32_t)jobs[t].win -
                                  (int32_t)ch.read_start);
            }
        });
        return off;
    }

    // clamp_window (mod.rs:1766-1771) job of a chain
    Job clamp_job(int32_t r, int64_t chain, int32_t abort) const {
        const kira_chain &ch = chains[chain];
        const uint32_t ref_len = (uint32_t)contig_len[ch.ref_id];
        const uint32_t flank = std::max<uint32_t>((uint32_t)std::max(c.cfg.bandwidth, 1) * 2u, 50u);
        const uint32_t start = usub(ch.ref_start, flank);
        uint32_t end = std::min<uint32_t>(ch.ref_end + flank, ref_len);
        end = std::max(end, start + 1);
        return Job{r, chain, (int64_t)start, (int32_t)(end - start), abort};
    }

    kira_status run();
};

kira_status Driver::run() {
    kira_stage4_stats &st = R.stats;
    st.reads = n;
    const int32_t lanes = c.simd == KIRA_SIMD_AVXVNNI ? 32
                          : c.simd == KIRA_SIMD_AVX2  ? 16
                          : c.simd == KIRA_SIMD_NEON  ? 4
                                                      : 1;
    const int32_t ambig_div = c.ambig_div >= 1 ? c.ambig_div : 5;
    auto t0 = Clock::now();
    kira::StepTrace trace("stage4");

    // oriented reads
    int64_t total = 0;
    ooff.resize(n);
    for (int32_t r = 0; r < n; r++) {
        ooff[r] = total;
        total += 2 * (int64_t)read_len[r];
    }
    hipError_t e;
    if ((e = orient.resize(ctx, kira::kPinOrient, (size_t)total)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "pinned read staging", e);
    uint8_t comp[256];
    for (int b = 0; b < 256; b++) comp[b] = comp_base((uint8_t)b);
    parallel_chunks(n, 512, [&](int, int64_t b, int64_t e) {
        for (int64_t r = b; r < e; r++) {
            const uint8_t *src = reads + read_off[r];
            uint8_t *f = orient.data() + ooff[r];
            const int32_t L = read_len[r];
            memcpy(f, src, (size_t)L);
            for (int32_t k = 0; k < L; k++) f[L + k] = comp[src[L - 1 - k]];
        }
    });

    kira::trace_mark("orient");
    // the oriented reads go to the device once; every phase reads them there
    if ((e = mem.alloc(&d_orient, (size_t)total + 8)) ||
        (total && (e = hipMemcpyAsync(d_orient, orient.data(), (size_t)total,
                                      hipMemcpyHostToDevice, s))))
        return fail(ctx, e == hipErrorOutOfMemory ? KIRA_ERR_OUT_OF_MEMORY : KIRA_ERR_RUNTIME,
                    "read upload", e);

    kira::trace_mark("upload");
    // chain selection (:549-708, mate guidance off), in two passes: the
    // two-tier candidate search of uncertain short reads (:596-627) runs its
    // fast cascades for the whole batch at once, then each read selects
    kira::PodVec<Sel> sels;
    std::vector<int32_t> chain_best(n, 0), chain_second(n, 0), min_score_of(n, 0),
        topk_of(n, 1), margin_of(n, INT32_MAX);
    std::vector<uint8_t> has_comp(n, 0);
    std::vector<int32_t> comp_of(n, 0);
    // two-tier candidates: (read, chain) in read order, passing order
    struct Cand {
        int32_t read;
        int64_t chain;
        int32_t score;   // INT32_MIN: the cascade found nothing
        int32_t fast;    // index into tt_res, or -1
    };
    std::vector<Cand> cand;
    // per read, in parallel: the selection bounds and the number of two-tier
    // candidates (cand_off[r + 1]); then, after the prefix sum, the candidates
    std::vector<int64_t> cand_off(n + 1, 0);
    const int32_t kmax = c.twotier_k >= 2 ? c.twotier_k : 5;
    parallel_chunks(n, 1024, [&](int, int64_t rb, int64_t re) {
        for (int32_t r = (int32_t)rb; r < (int32_t)re; r++) {
            const int64_t c0 = chain_off[r], nc = chain_off[r + 1] - c0;
            if (nc == 0) continue;
            const kira_chain *cl = chains + c0;
            const int32_t best = cl[0].score;
            int32_t min_score = best > 0 ? f2i(ceilf((float)best * c.min_chain_ratio)) : best;
            const int32_t second = nc > 1 ? cl[1].score : 0;
            const int32_t margin = nc > 1 ? wrap_i32((int64_t)best - second) : INT32_MAX;
            chain_best[r] = best;
            chain_second[r] = second;
            margin_of[r] = margin;
            const int32_t L = read_len[r];
            const bool short_read = L <= 300;
            const bool ambiguous =
                nc > 1 && (margin <= std::max(best / ambig_div, 8) || rmo_of(r) > 1);
            const bool weak = nc > 1 && best < wrap_i32((int64_t)L * 2) / 3;
            if (short_read && weak && c.twotier) min_score = std::min(min_score, best / 5);
            min_score_of[r] = min_score;
            topk_of[r] = short_read
                             ? (c.short_dp_topk >= 1 ? c.short_dp_topk : (ambiguous ? 2 : 1))
                             : std::max(c.dp_topk, 1);
            if (!(c.twotier && short_read && (ambiguous || weak))) continue;
            int32_t np = 0;
            for (int64_t k = 0; k < nc && np < kmax; k++)
                if (cl[k].score >= min_score) np++;
            if (np >= 2) cand_off[r + 1] = np;
        }
    });
    st.chains_total += chain_off[n] - chain_off[0];
    for (int32_t r = 0; r < n; r++) cand_off[r + 1] += cand_off[r];
    cand.resize((size_t)cand_off[n]);
    parallel_chunks(n, 1024, [&](int, int64_t rb, int64_t re) {
        for (int32_t r = (int32_t)rb; r < (int32_t)re; r++) {
            int64_t o = cand_off[r];
            if (o == cand_off[r + 1]) continue;
            const int64_t c0 = chain_off[r], nc = chain_off[r + 1] - c0;
            for (int64_t k = 0; k < nc && o < cand_off[r + 1]; k++)
                if (chains[c0 + k].score >= min_score_of[r])
                    cand[(size_t)o++] = Cand{r, c0 + k, INT32_MIN, -1};
        }
    });
    kira::trace_mark("select1");
    // exact copies score A * L (exact_match_exists, mod.rs:818-832); the rest
    // go through the fast cascade together
    kira::PodVec<kira_fast_aln> tt_res;
    std::vector<uint32_t> tt_cig;
    std::string tt_md;
    {
        kira::PodVec<int32_t> fk, frl;
        kira::PodVec<int64_t> fro;
        kira::PodVec<kira_chain> fch;
        // exact copies get their score; the others are listed for the
        // cascade in candidate order (count per range, then place)
        const int64_t nc = (int64_t)cand.size();
        const int64_t kGrain = 1024;
        const int T = host_threads();
        const int Tn = (int)std::max<int64_t>(1, std::min<int64_t>(T, (nc + kGrain - 1) / kGrain));
        std::vector<int64_t> cnt((size_t)T + 1, 0);
        std::vector<uint8_t> is_copy(cand.size(), 0);
        parallel_chunks(nc, kGrain, [&](int t, int64_t b, int64_t e) {
            int64_t others = 0;
            for (int64_t i = b; i < e; i++) {
                Cand &x = cand[i];
                const kira_chain &ch = chains[x.chain];
                const int32_t L = read_len[x.read];
                is_copy[i] = exact_candidate(x.read, ch) &&
                             memcmp(oriented(x.read, ch.strand == 1),
                                    contig(ch.ref_id) + ch.ref_start, (size_t)L) == 0;
                if (is_copy[i])
                    x.score = wrap_i32((int64_t)c.cfg.match_score * L);
                else
                    others++;
            }
            cnt[(size_t)t + 1] = others;
        });
        for (int t = 0; t < Tn; t++) cnt[(size_t)t + 1] += cnt[(size_t)t];
        const int64_t nf = cnt[(size_t)Tn];
        fk.resize((size_t)nf);
        fro.resize((size_t)nf);
        frl.resize((size_t)nf);
        fch.resize((size_t)nf);
        parallel_chunks(nc, kGrain, [&](int t, int64_t b, int64_t e) {
            int64_t o = cnt[(size_t)t];
            for (int64_t i = b; i < e; i++) {
                if (is_copy[i]) continue;
                const Cand &x = cand[i];
                const kira_chain &ch = chains[x.chain];
                fk[(size_t)o] = (int32_t)i;
                fro[(size_t)o] = oriented_off(x.read, ch.strand == 1);
                frl[(size_t)o] = read_len[x.read];
                fch[(size_t)o] = ch;
                o++;
            }
        });
        kira::trace_mark("tt_prep");
        if (!fk.empty()) {
            kira_fast_stats fst;
            kira_status fs = kira::fast_dp_run(ctx, g, contig_off, contig_len, (int32_t)fk.size(),
                                               orient.data(), (int64_t)orient.size(), d_orient,
                                               fro.data(), frl.data(), fch.data(), &c.cfg, c.fast,
                                               tt_res, tt_cig, tt_md, fst, false);
            if (fs != KIRA_OK) return fs;
            for (size_t i = 0; i < fk.size(); i++)
                if (tt_res[i].kind != 0) {
                    cand[fk[i]].score = tt_res[i].score;
                    cand[fk[i]].fast = (int32_t)i;
                }
        }
    }
    kira::trace_mark("tt_cascade");
    // each read selects; threads take read ranges and their selections are
    // joined in read order
    std::vector<std::vector<Sel>> part((size_t)host_threads());
    parallel_chunks(n, 512, [&](int t, int64_t rb, int64_t re) {
        std::vector<Sel> &out = part[(size_t)t];
        // per-read scratch, reused across reads
        std::vector<int64_t> order, idx;
        std::vector<int32_t> cached_of;  // per entry of `order`
        std::vector<uint8_t> in;
        for (int32_t r = (int32_t)rb; r < (int32_t)re; r++) {
            const int64_t c0 = chain_off[r], nc = chain_off[r + 1] - c0;
            if (nc == 0) continue;
            const kira_chain *cl = chains + c0;
            int32_t topk = topk_of[r];
            order.clear();
            cached_of.clear();
            const int64_t a0 = cand_off[r], a1 = cand_off[r + 1];
            if (a1 > a0) {
                // two_tier_reorder (:346-451): by score, ties by chain index (a
                // total order: the chains of a read are distinct)
                idx.clear();
                for (int64_t i = a0; i < a1; i++) idx.push_back(i);
                std::sort(idx.begin(), idx.end(), [&](int64_t u, int64_t v) {
                    if (cand[u].score != cand[v].score) return cand[u].score > cand[v].score;
                    return cand[u].chain < cand[v].chain;
                });
                const int32_t s0 = cand[idx[0]].score;
                const int32_t mrg = std::max(sat_mul(wrap_i32((int64_t)c.cfg.match_score +
                                                              c.cfg.mismatch), 2), 1);
                bool any_none = false;
                for (int64_t i : idx) any_none |= cand[i].score == INT32_MIN;
                int32_t dp_limit;
                if (any_none) {
                    dp_limit = 2;
                } else if (idx.size() >= 3 && cand[idx[2]].score == s0) {
                    dp_limit = 2;
                } else {
                    const int32_t floor_s = sat_sub(s0, mrg);
                    int32_t cnt = 0;
                    for (int64_t i : idx) {
                        if (cand[i].score < floor_s) break;
                        cnt++;
                    }
                    dp_limit = std::min(std::max(cnt, 1), 3);
                }
                in.assign(nc, 0);
                for (int64_t i : idx) {
                    order.push_back(cand[i].chain);
                    cached_of.push_back(cand[i].fast);
                    in[cand[i].chain - c0] = 1;
                }
                for (int64_t k = 0; k < nc; k++)
                    if (!in[k]) {
                        order.push_back(c0 + k);
                        cached_of.push_back(-1);
                    }
                if (c.short_dp_topk < 1) topk = dp_limit;
                if (topk == 1) {
                    // the runner-up's affine score becomes XS (:673-683)
                    bool have = false;
                    int32_t mx = INT32_MIN;
                    for (int64_t i : idx) {
                        if (cand[i].chain == order[0] || cand[i].score == INT32_MIN) continue;
                        mx = have ? std::max(mx, cand[i].score) : cand[i].score;
                        have = true;
                    }
                    has_comp[r] = have;
                    comp_of[r] = mx;
                }
            } else {
                for (int64_t k = 0; k < nc; k++) {
                    order.push_back(c0 + k);
                    cached_of.push_back(-1);
                }
            }
            int32_t selected = 0;
            for (size_t p = 0; p < order.size(); p++) {
                if (chains[order[p]].score < min_score_of[r]) continue;
                if (selected >= topk) break;
                out.push_back(
                    Sel{r, order[p], (uint8_t)(selected == 0), margin_of[r], cached_of[p]});
                selected++;
            }
            (void)cl;
        }
    });
    {
        std::vector<size_t> at(part.size() + 1, 0);
        for (size_t t = 0; t < part.size(); t++) at[t + 1] = at[t] + part[t].size();
        sels.resize(at[part.size()]);
        parallel_chunks((int64_t)part.size(), 1, [&](int, int64_t b, int64_t e) {
            for (int64_t t = b; t < e; t++)
                if (!part[t].empty())
                    memcpy(sels.data() + at[t], part[t].data(), part[t].size() * sizeof(Sel));
        });
    }
    st.chains_used = (int64_t)sels.size();

    kira::trace_mark("select2");
    // exact-match check and prefilter of every selected chain, on the device
    const int32_t ns = (int32_t)sels.size();
    std::vector<uint8_t> exact(ns, 0);
    kira::Pinned<kira_prefilter_out> pf;
    if (ns > 0 && (e = pf.resize(ctx, kira::kPinPfOut, (size_t)ns)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "pinned prefilter staging", e);
    if (ns > 0) {
        // the job arrays side by side in one pinned block, uploaded with one copy
        const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t o_fo = al((size_t)ns * 8), o_fl = o_fo + al((size_t)ns * 8),
                     o_rl = o_fl + al((size_t)ns * 8), o_mg = o_rl + al((size_t)ns * 4),
                     o_sp = o_mg + al((size_t)ns * 4),
                     o_top = o_sp + al((size_t)ns * sizeof(kira_span)),
                     o_end = o_top + al((size_t)ns);
        kira::Pinned<uint8_t> hb;
        if ((e = hb.resize(ctx, kira::kPinPfJobs, o_end)))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "pinned prefilter staging", e);
        int64_t *ro = (int64_t *)hb.data(), *fo = (int64_t *)(hb.data() + o_fo),
                *fl = (int64_t *)(hb.data() + o_fl);
        int32_t *rl = (int32_t *)(hb.data() + o_rl), *mg = (int32_t *)(hb.data() + o_mg);
        kira_span *sp = (kira_span *)(hb.data() + o_sp);
        uint8_t *top = hb.data() + o_top;
        std::vector<int64_t> ex_ro, ex_pos;
        std::vector<int32_t> ex_rl, ex_sel;
        std::vector<uint8_t> is_ex((size_t)ns);
        parallel_chunks(ns, 2048, [&](int, int64_t b0, int64_t e0) {
            for (int64_t k = b0; k < e0; k++) {
                const Sel &x = sels[k];
                const kira_chain &ch = chains[x.chain];
                ro[k] = oriented_off(x.read, ch.strand == 1);
                rl[k] = read_len[x.read];
                fo[k] = contig_off[ch.ref_id];
                fl[k] = contig_len[ch.ref_id];
                sp[k] = kira_span{ch.ref_start, ch.ref_end, ch.read_start, ch.read_end};
                top[k] = x.top1;
                mg[k] = x.margin;
                is_ex[k] = exact_candidate(x.read, ch) ? 1 : 0;
            }
        });
        for (int32_t k = 0; k < ns; k++) {
            if (!is_ex[k]) continue;
            ex_sel.push_back(k);
            ex_ro.push_back(ro[k]);
            ex_rl.push_back(rl[k]);
            ex_pos.push_back(fo[k] + chains[sels[k].chain].ref_start);
        }
        kira::trace_mark("pf_prep");
        const int32_t ne = (int32_t)ex_sel.size();
        DevMem dm(ctx);
        uint8_t *db = nullptr;
        int64_t *d_ero = nullptr, *d_epos = nullptr;
        int32_t *d_erl = nullptr;
        uint8_t *d_ex = nullptr;
        kira_prefilter_out *d_out;
        if ((e = dm.alloc(&db, o_end)) || (e = dm.alloc(&d_out, ns)) ||
            (ne && ((e = dm.alloc(&d_ero, ne)) || (e = dm.alloc(&d_epos, ne)) ||
                    (e = dm.alloc(&d_erl, ne)) || (e = dm.alloc(&d_ex, ne)))))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "triage allocation", e);
        const int64_t *d_ro = (const int64_t *)db, *d_fo = (const int64_t *)(db + o_fo),
                      *d_fl = (const int64_t *)(db + o_fl);
        const int32_t *d_rl = (const int32_t *)(db + o_rl), *d_mg = (const int32_t *)(db + o_mg);
        const kira_span *d_sp = (const kira_span *)(db + o_sp);
        const uint8_t *d_top = db + o_top;
        auto up = [&](void *d, const void *h, size_t bytes) {
            return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s);
        };
        if ((e = up(db, hb.data(), o_end)) ||
            (ne && ((e = up(d_ero, ex_ro.data(), ne * 8)) ||
                    (e = up(d_epos, ex_pos.data(), ne * 8)) ||
                    (e = up(d_erl, ex_rl.data(), ne * 4)))))
            return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
        if (ne) {
            exact_match_kernel<<<(unsigned)((ne + 3) / 4), 256, 0, s>>>(ne, d_orient, d_ero, d_erl,
                                                                       g->ascii, d_epos, d_ex);
            if ((e = hipGetLastError())) return fail(ctx, KIRA_ERR_RUNTIME, "exact-match launch", e);
        }
        kira::PrefilterLaunch a{};
        a.n = ns;
        a.reads = d_orient;
        a.read_off = d_ro;
        a.read_len = d_rl;
        a.refs = g->ascii;
        a.ref_off = d_fo;
        a.ref_len = d_fl;
        a.spans = d_sp;
        a.is_top1 = d_top;
        a.score_margin = d_mg;
        a.match_score = c.cfg.match_score;
        a.mismatch = c.cfg.mismatch;
        a.xdrop = c.cfg.xdrop;
        a.p.accept_enable = (c.accept_enable && c.max_alignments <= 1) ? 1 : 0;
        a.p.accept_only_top1 = c.accept_only_top1;
        a.p.multi_alignments_enabled = c.max_alignments > 1 ? 1 : 0;
        a.p.tail_ext = c.tail_ext;
        a.p.accept_span_slack = c.accept_span_slack;
        a.p.accept_max_mismatches = c.accept_max_mismatches;
        a.p.accept_require_score_margin = c.accept_require_score_margin;
        a.p.long_read_threshold = c.long_read_threshold;
        a.p.accept_min_identity = c.accept_min_identity;
        a.p.ungap_min_id = c.ungap_min_id;
        a.p.ungap_gap_check = c.ungap_gap_check;
        a.id_floor = kira::prefilter_id_floor(a.p);
        a.out = d_out;
        if ((e = kira::launch_prefilter(a, s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "prefilter launch", e);
        std::vector<uint8_t> ex(ne);
        if ((e = hipMemcpyAsync(pf.data(), d_out, ns * sizeof(kira_prefilter_out),
                                hipMemcpyDeviceToHost, s)) ||
            (ne && (e = hipMemcpyAsync(ex.data(), d_ex, ne, hipMemcpyDeviceToHost, s))) ||
            (e = hipStreamSynchronize(s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
        kira::trace_mark("pf_gpu");
        for (int32_t k = 0; k < ne; k++) exact[ex_sel[k]] = ex[k];
        if (c.ungap_gap_check) {
            kira_status gs = kira::prefilter_gap_check(ctx, ns, d_orient, ro, rl, g->ascii, fo,
                                                       fl, sp, c.cfg, pf.data());
            if (gs != KIRA_OK) return gs;
        }
    }

    kira::trace_mark("pf_done");
    // outcomes in chain order (:758-934): 0 nothing, 1 a record here, 2 fallback
    JobVec simd, scalar;
    std::vector<uint8_t> dp_used(n, 0), outcome(ns, 0);
    std::vector<int32_t> abort_of(ns, 0);
    std::vector<int32_t> accepted;  // selected chains that end in a record here
    const int32_t band = std::max(c.cfg.bandwidth, 1);
    {
        // in parallel over chain ranges; each range counts into its own stats
        struct Cnt {
            int64_t exact = 0, packed = 0, wfa = 0, accept = 0, reject = 0, fallback = 0;
            int64_t reason[10] = {};
        };
        std::vector<Cnt> cnt((size_t)host_threads());
        parallel_chunks(ns, 1024, [&](int t, int64_t b0, int64_t e0) {
            Cnt &C = cnt[(size_t)t];
            for (int64_t k = b0; k < e0; k++) {
                const Sel &x = sels[k];
                if (exact[k]) {
                    C.exact++;
                    outcome[k] = 1;
                    continue;
                }
                if (x.cached >= 0) {
                    // the candidate search's verified alignment (:736-747)
                    if (tt_res[x.cached].kind == 1)
                        C.packed++;
                    else
                        C.wfa++;
                    __atomic_store_n(&dp_used[x.read], (uint8_t)1, __ATOMIC_RELAXED);
                    outcome[k] = 1;
                    continue;
                }
                const kira_prefilter_out &o = pf[k];
                C.reason[o.reason < 9 ? o.reason : 9]++;
                abort_of[k] = o.has_metrics ? sat_sub(o.score, c.dp_abort_margin) : INT32_MIN / 4;
                if (o.result == 0) {
                    C.accept++;
                    outcome[k] = 1;
                    continue;
                }
                if (o.result == 1) {
                    C.reject++;
                    continue;
                }
                C.fallback++;
                __atomic_store_n(&dp_used[x.read], (uint8_t)1, __ATOMIC_RELAXED);
                outcome[k] = 2;
            }
        });
        for (const Cnt &C : cnt) {
            st.exact_matches += C.exact;
            st.packed_spectral_resolved += C.packed;
            st.wfa_resolved += C.wfa;
            st.prefilter_accept += C.accept;
            st.prefilter_reject += C.reject;
            st.prefilter_fallback += C.fallback;
            for (int r = 0; r < 10; r++) st.reason_counts[r] += C.reason[r];
        }
    }
    // without the GPU dispatcher, a fallback chain tries the fast DP cascade
    // first (try_fast_dp_alignment, :868-882; accept_fast_result, :289-299)
    std::vector<int32_t> fast_of(ns, -1);
    kira::PodVec<kira_fast_aln> fres;
    std::vector<uint32_t> fcig;
    std::string fmd;
    if (c.cpu_mode) {
        std::vector<int32_t> fk, frl;
        std::vector<int64_t> fro;
        std::vector<kira_chain> fch;
        for (int32_t k = 0; k < ns; k++) {
            if (outcome[k] != 2) continue;
            const Sel &x = sels[k];
            const kira_chain &ch = chains[x.chain];
            fk.push_back(k);
            fro.push_back(oriented_off(x.read, ch.strand == 1));
            frl.push_back(read_len[x.read]);
            fch.push_back(ch);
        }
        kira_fast_stats fst;
        kira_status fs = kira::fast_dp_run(ctx, g, contig_off, contig_len, (int32_t)fk.size(),
                                           orient.data(), (int64_t)orient.size(), d_orient,
                                           fro.data(), frl.data(), fch.data(), &c.cfg, c.fast,
                                           fres, fcig, fmd, fst, true);
        if (fs != KIRA_OK) return fs;
        for (size_t i = 0; i < fk.size(); i++) {
            if (fres[i].kind == 0) continue;
            fast_of[fk[i]] = (int32_t)i;
            if (fres[i].kind == 1)
                st.packed_spectral_resolved++;
            else
                st.wfa_resolved++;
        }
    }
    {
        // route every chain (0 none, 1 a record here, 2 lane job, 3 scalar
        // job) in parallel, then place each list's items in chain order
        std::vector<uint8_t> route(ns);
        JobVec job_of(ns);
        const int T = host_threads();
        std::vector<std::array<int64_t, 4>> cnt((size_t)T, std::array<int64_t, 4>{});
        const int64_t kGrain = 1024;
        const int Tn = (int)std::max<int64_t>(1, std::min<int64_t>(T, (ns + kGrain - 1) / kGrain));
        parallel_chunks(ns, kGrain, [&](int t, int64_t b0, int64_t e0) {
            for (int64_t k = b0; k < e0; k++) {
                uint8_t rt = 0;
                if (outcome[k] != 0) {
                    if (outcome[k] == 1 || fast_of[k] >= 0) {
                        rt = 1;
                    } else {
                        const Sel &x = sels[k];
                        const kira_chain &ch = chains[x.chain];
                        const int32_t L = read_len[x.read];
                        const int32_t abort = abort_of[k];
                        const int bucket = L <= 120 ? 0 : (L <= 150 ? 1 : 2);
                        rt = 3;
                        if (bucket < 2 && lanes > 1) {
                            // build_simd_window (:1426-1445)
                            const int32_t flank = band * 2;
                            const int64_t wlen = (int64_t)L + 2 * (int64_t)flank;
                            const int32_t desired = wrap_i32((int64_t)(int32_t)ch.ref_start -
                                                             (int32_t)ch.read_start - flank);
                            if (desired >= 0 && desired + wlen <= contig_len[ch.ref_id]) {
                                job_of[k] = Job{x.read, x.chain, desired, (int32_t)wlen, abort};
                                rt = 2;
                            }
                        }
                        if (rt == 3) job_of[k] = clamp_job(x.read, x.chain, abort);
                    }
                }
                route[k] = rt;
                cnt[(size_t)t][rt]++;
            }
        });
        // chunk t's first position in each list
        std::array<int64_t, 4> tot{};
        std::vector<std::array<int64_t, 4>> base((size_t)T);
        for (int t = 0; t < Tn; t++)
            for (int r = 0; r < 4; r++) {
                base[(size_t)t][r] = tot[r];
                tot[r] += cnt[(size_t)t][r];
            }
        accepted.resize((size_t)tot[1]);
        simd.resize((size_t)tot[2]);
        scalar.resize((size_t)tot[3]);
        parallel_chunks(ns, kGrain, [&](int t, int64_t b0, int64_t e0) {
            std::array<int64_t, 4> pos = base[(size_t)t];
            for (int64_t k = b0; k < e0; k++) {
                switch (route[k]) {
                case 1: accepted[(size_t)pos[1]++] = (int32_t)k; break;
                case 2: simd[(size_t)pos[2]++] = job_of[k]; break;
                case 3: scalar[(size_t)pos[3]++] = job_of[k]; break;
                default: break;
                }
            }
        });
    }
    kira::trace_mark("route");
    for (int32_t r = 0; r < n; r++) st.dp_reads += dp_used[r];
    R.phase_ms[0] = ms_since(t0);
    // the records of the accepted chains are built later, while the device
    // runs the SW kernels
    auto build_accepted = [&]() {
    build((int64_t)accepted.size(), [&](int64_t i, Arena &A) {
        const int32_t k = accepted[i];
        const Sel &x = sels[k];
        const kira_chain &ch = chains[x.chain];
        if (exact[k]) {
            A.add(x.read, exact_aln(x.read, ch, A));
            return;
        }
        if (fast_of[k] >= 0 || x.cached >= 0) {
            // the fast path's Alignment (mod.rs:636-686, 481-575), from the
            // candidate search or the fallback
            const bool tt = x.cached >= 0;
            const kira_fast_aln &f = tt ? tt_res[x.cached] : fres[fast_of[k]];
            const std::vector<uint32_t> &fc = tt ? tt_cig : fcig;
            const std::string &fm = tt ? tt_md : fmd;
            kira_stage4_aln a = base_aln(1, ch);
            a.ref_start = f.ref_start;
            a.ref_end = f.ref_end;
            a.read_start = f.read_start;
            a.read_end = f.read_end;
            a.score = f.score;
            a.mapq = 0;
            a.cigar_off = (int64_t)A.cig.size();
            a.n_cigar = f.n_cigar;
            A.cig.insert(A.cig.end(), fc.begin() + f.cigar_off,
                         fc.begin() + f.cigar_off + f.n_cigar);
            a.md_off = (int64_t)A.md.size();
            if (tt) {
                // the candidate search builds no MD: only the kept ones need it
                const uint8_t *q = oriented(x.read, ch.strand == 1);
                a.nm = f.kind == 1 ? kira::ungapped_tags(q, contig(ch.ref_id) + f.ref_start,
                                                         (int64_t)read_len[x.read], A.md)
                                   : kira::nm_md(q, read_len[x.read], contig(ch.ref_id),
                                                 contig_len[ch.ref_id], f.ref_start,
                                                 A.cig.data() + a.cigar_off, f.n_cigar, A.md);
                a.md_len = (int32_t)((int64_t)A.md.size() - a.md_off);
            } else {
                a.nm = f.nm;
                A.md.append(fm, (size_t)f.md_off, (size_t)f.md_len);
                a.md_len = f.md_len;
            }
            A.add(x.read, a);
            return;
        }
        // build_ungapped_alignment (prefilter.rs:455-517)
        const kira_prefilter_out &o = pf[k];
        kira_stage4_aln a = base_aln(0, ch);
        a.ref_start = o.aln_ref_start;
        a.ref_end = o.aln_ref_end;
        a.read_start = o.aln_read_start;
        a.read_end = o.aln_read_end;
        a.score = o.aln_score;
        a.mapq = 60;
        a.cigar_off = (int64_t)A.cig.size();
        a.n_cigar = o.n_cigar;
        A.cig.insert(A.cig.end(), o.cigar, o.cigar + o.n_cigar);
        a.md_off = (int64_t)A.md.size();
        a.nm = kira::ungapped_tags(oriented(x.read, ch.strand == 1) + o.aln_read_start,
                                   contig(ch.ref_id) + o.aln_ref_start,
                                   (int64_t)o.aln_read_end - o.aln_read_start, A.md);
        a.md_len = (int32_t)(A.md.size() - a.md_off);
        A.add(x.read, a);
    });
    };

    // gpu_dispatch_spectral (:1272-1358): sieve + certification
    t0 = Clock::now();
    std::vector<uint8_t> keep(simd.size(), 1);
    const int32_t nj = (int32_t)simd.size();
    std::vector<int32_t> cert, sv_shift, sv_best;
    if (nj > 0 && !c.cpu_mode) {
        // job arrays side by side in one pinned block, one copy
        const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t o_ws = al((size_t)nj * 8), o_rl = o_ws + al((size_t)nj * 8),
                     o_wl = o_rl + al((size_t)nj * 4), o_lim = o_wl + al((size_t)nj * 4),
                     o_end = o_lim + al((size_t)nj * 4);
        kira::Pinned<uint8_t> hb;
        if ((e = hb.resize(ctx, kira::kPinSieve, o_end)))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "pinned sieve staging", e);
        int64_t *ro = (int64_t *)hb.data(), *ws = (int64_t *)(hb.data() + o_ws);
        int32_t *rl = (int32_t *)(hb.data() + o_rl), *wl = (int32_t *)(hb.data() + o_wl),
                *lim = (int32_t *)(hb.data() + o_lim);
        std::vector<int32_t> part_max((size_t)host_threads(), 0);
        parallel_chunks(nj, 2048, [&](int tt, int64_t b0, int64_t e0) {
            int32_t mx = 0;
            for (int64_t t = b0; t < e0; t++) {
                const Job &j = simd[t];
                const kira_chain &ch = chains[j.chain];
                ro[t] = oriented_off(j.read, ch.strand == 1);
                rl[t] = read_len[j.read];
                ws[t] = contig_off[ch.ref_id] + j.win;
                wl[t] = j.wlen;
                // spectral_max_mismatches (router.rs:191-195)
                lim[t] = std::max<int32_t>((int32_t)((int64_t)rl[t] * c.spectral_pct / 100),
                                           c.spectral_floor);
                mx = std::max(mx, rl[t]);
            }
            part_max[(size_t)tt] = mx;
        });
        int32_t max_len = 0;
        for (int32_t v : part_max) max_len = std::max(max_len, v);
        DevMem dm(ctx);
        uint8_t *d_blob;
        int32_t *d_shift, *d_best, *d_second;
        uint32_t *d_planes;
        uint8_t *d_valid, *d_flags;
        if ((e = dm.alloc(&d_blob, o_end)) || (e = dm.alloc(&d_shift, nj)) ||
            (e = dm.alloc(&d_best, nj)) || (e = dm.alloc(&d_second, nj)) ||
            (e = dm.alloc(&d_planes, (size_t)nj * 24)) || (e = dm.alloc(&d_valid, nj)) ||
            (e = dm.alloc(&d_flags, nj)))
            return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "sieve allocation", e);
        const int64_t *d_ro = (const int64_t *)d_blob, *d_ws = (const int64_t *)(d_blob + o_ws);
        const int32_t *d_rl = (const int32_t *)(d_blob + o_rl),
                      *d_wl = (const int32_t *)(d_blob + o_wl),
                      *d_lim = (const int32_t *)(d_blob + o_lim);
        if ((e = hipMemcpyAsync(d_blob, hb.data(), o_end, hipMemcpyHostToDevice, s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
        if ((e = kira::launch_pack_reads_ascii(nj, d_orient, 0, d_ro, d_rl, d_planes, d_valid, s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "read packing", e);
        kira_sieve_batch b{};
        b.n_jobs = nj;
        b.max_read_len = max_len;
        b.read_planes = d_planes;
        b.read_lens = d_rl;
        b.read_valid = d_valid;
        b.job_read = nullptr;
        b.jobs_per_read = 1;
        b.win_start = d_ws;
        b.win_len = d_wl;
        b.limit = d_lim;
        b.out_shift = d_shift;
        b.out_best = d_best;
        b.out_second = d_second;
        b.out_flags = d_flags;
        // the certification limit is the job's own dispatch limit
        b.cert_max_mism = INT32_MAX;
        b.cert_strict = c.cert_strict;
        kira::trace_mark("sieve_prep");
        kira_status kst = kira_sieve_genome_dev(ctx, g, &b);
        if (kst != KIRA_OK) return kst;
        std::vector<int32_t> &shift = sv_shift, &best = sv_best;
        shift.resize(nj);
        best.resize(nj);
        std::vector<uint8_t> flags(nj);
        if ((e = hipMemcpyAsync(shift.data(), d_shift, nj * 4, hipMemcpyDeviceToHost, s)) ||
            (e = hipMemcpyAsync(best.data(), d_best, nj * 4, hipMemcpyDeviceToHost, s)) ||
            (e = hipMemcpyAsync(flags.data(), d_flags, nj, hipMemcpyDeviceToHost, s)) ||
            (e = hipStreamSynchronize(s)))
            return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
        kira::trace_mark("sieve_gpu");
        for (int32_t t = 0; t < nj; t++)
            if (flags[t] & 4) {
                cert.push_back(t);
                keep[t] = 0;
            }
        st.gpu_spectral_resolved += (int64_t)cert.size();
    }
    R.phase_ms[1] = ms_since(t0);
    auto build_certified = [&]() {
        build((int64_t)cert.size(), [&](int64_t i, Arena &A) {
            // build_certified_spectral_alignment (mod.rs:636-686)
            const int32_t t = cert[i];
            const Job &j = simd[t];
            const kira_chain &ch = chains[j.chain];
            const int32_t L = read_len[j.read], mm = sv_best[t];
            const int32_t *shift = sv_shift.data();
            kira_stage4_aln a = base_aln(1, ch);
            a.ref_start = (uint32_t)(j.win + shift[t]);
            a.ref_end = a.ref_start + (uint32_t)L;
            a.read_start = 0;
            a.read_end = (uint32_t)L;
            a.score = wrap_i32((int64_t)(L - mm) * c.cfg.match_score - (int64_t)mm * c.cfg.mismatch);
            a.mapq = 0;
            a.cigar_off = (int64_t)A.cig.size();
            a.n_cigar = 1;
            A.cig.push_back(((uint32_t)L << 4) | kOpM);
            a.md_off = (int64_t)A.md.size();
            a.nm = kira::ungapped_tags(oriented(j.read, ch.strand == 1),
                                       contig(ch.ref_id) + j.win + shift[t], L, A.md);
            a.md_len = (int32_t)(A.md.size() - a.md_off);
            A.add(j.read, a);
        });
    };

    // align_batch_simd lane groups: BTreeMap over (read, window) length,
    // chunks of `lanes` jobs (:1153-1186); the groups and the jobs outside
    // any group are launched, then the scalar jobs, and the host builds the
    // accepted and certified records while the device runs them
    t0 = Clock::now();
    // the side stream (scalar jobs, below) starts after what the context
    // stream holds now, not after the lane groups
    const bool use_side = ctx->side && !scalar.empty();
    if (use_side && (e = hipEventRecord(ctx->side_fork, s)) != hipSuccess)
        return fail(ctx, KIRA_ERR_RUNTIME, "side stream", e);
    std::vector<int32_t> lane_order;  // lane jobs (indices into simd) in group order
    kira::GroupRun lane_groups(ctx);
    kira::BandedRun lane_banded(ctx), scalar_run(ctx);
    std::vector<int32_t> lane_sel[2];  // positions in lane_order of each pass
    {
        std::vector<int32_t> kept;
        for (int32_t t = 0; t < nj; t++)
            if (keep[t]) kept.push_back(t);
        const int32_t m = (int32_t)kept.size();
        // stable order by (read length, window length)
        bool fits = true;
        std::vector<uint32_t> key(m);
        for (int32_t i = 0; i < m; i++) {
            const Job &j = simd[kept[i]];
            fits &= read_len[j.read] <= 65535 && j.wlen <= 65535;
            key[i] = ((uint32_t)read_len[j.read] << 16) | (uint32_t)(j.wlen & 0xFFFF);
        }
        std::vector<int32_t> ord;
        if (fits) {
            kira::stable_order_by_key(key.data(), m, ord);
        } else {
            ord.resize(m);
            for (int32_t i = 0; i < m; i++) ord[i] = i;
            std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
                const Job &a = simd[kept[x]], &b = simd[kept[y]];
                const int32_t lx = read_len[a.read], ly = read_len[b.read];
                return lx != ly ? lx < ly : a.wlen < b.wlen;
            });
        }
        lane_order.resize(m);
        for (int32_t i = 0; i < m; i++) lane_order[i] = kept[ord[i]];
        std::vector<int32_t> batch_off{0};
        for (int32_t i = 0; i < m;) {
            int32_t e = i;
            const Job &ji = simd[lane_order[i]];
            while (e < m && read_len[simd[lane_order[e]].read] == read_len[ji.read] &&
                   simd[lane_order[e]].wlen == ji.wlen)
                e++;
            for (int32_t b = i; b < e; b += lanes) batch_off.push_back(std::min(e, b + lanes));
            i = e;
        }
        kira::trace_mark("lane_order");
        std::vector<uint32_t> rl(m), wl(m);
        for (int32_t i = 0; i < m; i++) {
            rl[i] = (uint32_t)read_len[simd[lane_order[i]].read];
            wl[i] = (uint32_t)simd[lane_order[i]].wlen;
        }
        std::vector<int32_t> group_of, group_neg;
        kira::simd_groups((int32_t)batch_off.size() - 1, batch_off.data(), rl.data(), wl.data(),
                          c.simd, c.simd, c.cfg, group_of, group_neg);
        for (int pass = 0; pass < 2; pass++) {
            // pass 0: jobs outside any lane group run BatchInput::scalar_sw
            // (banded SW, offset clamped to [-L, W], mod.rs:60-72); pass 1: groups
            std::vector<int32_t> &sel = lane_sel[pass];
            for (int32_t i = 0; i < m; i++)
                if ((group_of[i] >= 0) == (pass == 1)) sel.push_back(i);
            const int32_t ms = (int32_t)sel.size();
            if (ms == 0) continue;
            kira::PodVec<int64_t> ro(ms), fo(ms);
            kira::PodVec<int32_t> srl(ms), sfl(ms), off(ms), ab(ms), gof(ms);
            parallel_chunks(ms, 16384, [&](int, int64_t b0, int64_t e0) {
                for (int64_t t = b0; t < e0; t++) {
                    const Job &j = simd[lane_order[sel[t]]];
                    const kira_chain &ch = chains[j.chain];
                    ro[t] = oriented_off(j.read, ch.strand == 1);
                    srl[t] = read_len[j.read];
                    fo[t] = contig_off[ch.ref_id] + j.win;
                    sfl[t] = j.wlen;
                    int64_t o = (int64_t)ch.ref_start - j.win - (int64_t)ch.read_start;
                    o = std::max<int64_t>(o, -(int64_t)srl[t]);
                    off[t] = (int32_t)std::min<int64_t>(o, (int64_t)j.wlen);
                    ab[t] = j.abort;
                    gof[t] = group_of[sel[t]];
                }
            });
            kira::SwJobsHost sj{ms,        d_orient,   ro.data(),  srl.data(), g->ascii,
                                fo.data(), sfl.data(), off.data(), nullptr,    ab.data()};
            kira_status kst =
                pass == 0 ? lane_banded.launch(sj, c.cfg, c.knobs, kira::kPinBanded)
                          : lane_groups.launch(sj, gof.data(), group_neg, c.cfg, kira::kPinGroup);
            if (kst != KIRA_OK) return kst;
        }
    }
    kira::trace_mark("lane_launch");
    double lane_ms = ms_since(t0);
    // scalar jobs: align_chain_with_meta (stage4:1196-1205), on the side
    // stream beside the lane groups (which leave most of the chip idle)
    t0 = Clock::now();
    {
        hipStream_t ss = s;
        if (use_side) {
            if ((e = hipStreamWaitEvent(ctx->side, ctx->side_fork, 0)) != hipSuccess)
                return fail(ctx, KIRA_ERR_RUNTIME, "side stream", e);
            ss = ctx->side;
        }
        kira_status kst =
            banded_launch(scalar_run, scalar, chain_offsets(scalar), kira::kPinBanded2, ss);
        if (kst != KIRA_OK) return kst;
    }
    kira::trace_mark("scalar_launch");
    double scalar_ms = ms_since(t0);

    t0 = Clock::now();
    build_accepted();
    kira::trace_mark("acc_records");
    R.phase_ms[0] += ms_since(t0);
    t0 = Clock::now();
    build_certified();
    kira::trace_mark("cert_records");
    R.phase_ms[1] += ms_since(t0);

    // places for the lane-group records, then the scalar ones; the scalar
    // jobs (on the side stream) usually finish first and are built first
    const size_t lane_arenas = reserve_arenas(), scalar_arenas = reserve_arenas();
    t0 = Clock::now();
    {
        kira::PodVec<kira_sw_result> res(scalar.size());
        std::vector<uint32_t> cig;
        kira_status kst = scalar_run.fetch(res.data(), cig);
        if (kst != KIRA_OK) return kst;
        kira::trace_mark("scalar_sw");
        build_at(scalar_arenas, (int64_t)scalar.size(), [&](int64_t t, Arena &A) {
            const Job &j = scalar[t];
            A.add(j.read, sw_aln(j.read, chains[j.chain], j.win, j.wlen, res[t],
                                 cig.data() + res[t].cigar_off, A));
        });
        for (const kira_sw_result &x : res) {
            st.dp_scalar++;
            if (x.early_abort) st.dp_early_abort++;
        }
    }
    kira::trace_mark("scalar_records");
    R.phase_ms[3] = scalar_ms + ms_since(t0);

    t0 = Clock::now();
    {
        const int32_t m = (int32_t)lane_order.size();
        kira::PodVec<kira_sw_result> out(m);
        std::vector<uint32_t> cig;
        for (int pass = 0; pass < 2; pass++) {
            const std::vector<int32_t> &sel = lane_sel[pass];
            const int32_t ms = (int32_t)sel.size();
            if (ms == 0) continue;
            kira::PodVec<kira_sw_result> r(ms);
            kira_status kst =
                pass == 0 ? lane_banded.fetch(r.data(), cig) : lane_groups.fetch(r.data(), cig);
            if (kst != KIRA_OK) return kst;
            for (int32_t t = 0; t < ms; t++) out[sel[t]] = r[t];
        }
        kira::trace_mark("lane_sw");
        build_at(lane_arenas, m, [&](int64_t i, Arena &A) {
            const Job &j = simd[lane_order[i]];
            A.add(j.read, sw_aln(j.read, chains[j.chain], j.win, j.wlen, out[i],
                                 cig.data() + out[i].cigar_off, A));
        });
        for (int32_t i = 0; i < m; i++) {
            st.dp_simd++;
            if (out[i].early_abort) st.dp_early_abort++;
        }
    }
    kira::trace_mark("lane_records");
    R.phase_ms[2] = lane_ms + ms_since(t0);

    t0 = Clock::now();
    std::vector<int64_t> first, perm;
    group_by_read(first, perm);

    // XS: the two-tier competitor's affine score, else the chain-score proxy
    // (:1215-1250), on the reads with exactly one record before the
    // supplementary ones are added
    auto assign_xs = [&]() {
        parallel_chunks(n, 2048, [&](int, int64_t rb, int64_t re) {
            for (int32_t r = (int32_t)rb; r < (int32_t)re; r++) {
                if (first[r + 1] - first[r] != 1) continue;
                if (has_comp[r]) {
                    kira_stage4_aln &a = rec(perm[first[r]]);
                    a.has_xs = 1;
                    a.xs_score = std::min(std::max(comp_of[r], 0), std::max(a.score, 0));
                    continue;
                }
                const int32_t cb = chain_best[r], cs = chain_second[r];
                if (cb <= 0 || cs <= 0) continue;
                kira_stage4_aln &a = rec(perm[first[r]]);
                if (a.has_xs) continue;
                const int32_t sub = wrap_i32((int64_t)a.score * cs / cb);
                if (sub > 0 && (int64_t)cs * 100 >= (int64_t)cb * c.xs_min_ratio_pct) {
                    a.has_xs = 1;
                    a.xs_score = sub;
                }
            }
        });
    };

    // emit_supplementary_alignments (split_read.rs:22-57)
    if (c.supplementary) {
        // picks in the order split_read.rs appends them: job -1 = exact match
        struct Pick {
            int32_t read;
            int64_t chain;
            int32_t job;
        };
        // threads take read ranges; their picks are joined in read order
        struct Part {
            std::vector<Pick> all;
            JobVec sj;
        };
        std::vector<Part> parts((size_t)host_threads());
        parallel_chunks(n, 512, [&](int t, int64_t rb, int64_t re) {
            Part &P = parts[(size_t)t];
            std::vector<int64_t> picks;
            for (int32_t r = (int32_t)rb; r < (int32_t)re; r++) {
                const int64_t c0 = chain_off[r], nc = chain_off[r + 1] - c0;
                if (first[r + 1] == first[r] || nc < 2) continue;
                const kira_stage4_aln &p = rec(perm[first[r]]);
                const uint32_t *pc = rec_cigar(perm[first[r]]);
                const kira_chain *cl = chains + c0;
                // pick_supplementary_chains (split_read.rs:60-125)
                const int32_t ps = std::max(cl[0].score, 1);
                const int32_t min_score = f2i(ceilf((float)ps * kSuppMinScoreRatio));
                uint32_t L = 0;
                for (int32_t k = 0; k < p.n_cigar; k++) {
                    uint32_t op = pc[k];
                    if ((op & 15u) == kOpM || (op & 15u) == kOpI || (op & 15u) == kOpS)
                        L += op >> 4;
                }
                auto pi = original_interval(p.read_start, p.read_end, p.is_rev != 0, L);
                picks.clear();
                for (int64_t k = 1; k < nc; k++) {
                    const kira_chain &ch = cl[k];
                    if (ch.score < min_score) break;
                    if (usub(ch.read_end, ch.read_start) < kSuppMinReadSpan) continue;
                    auto ci = original_interval(ch.read_start, ch.read_end, ch.strand == 1, L);
                    if (read_overlap_pct(pi.first, pi.second, ci.first, ci.second) >
                        kSuppMaxOverlapPct)
                        continue;
                    bool dup = false;
                    for (int64_t q : picks) {
                        const kira_chain &pc = chains[q];
                        auto qi =
                            original_interval(pc.read_start, pc.read_end, pc.strand == 1, L);
                        if (read_overlap_pct(qi.first, qi.second, ci.first, ci.second) >
                            kSuppMaxOverlapPct) {
                            dup = true;
                            break;
                        }
                    }
                    if (dup) continue;
                    picks.push_back(c0 + k);
                    if (picks.size() >= kMaxSuppPerRead) break;
                }
                for (int64_t q : picks) {
                    const kira_chain &ch = chains[q];
                    // align_oriented_chain_with_meta: exact shortcut, else banded SW
                    if (exact_candidate(r, ch) &&
                        memcmp(oriented(r, ch.strand == 1), contig(ch.ref_id) + ch.ref_start,
                               (size_t)read_len[r]) == 0) {
                        P.all.push_back(Pick{r, q, -1});
                        continue;
                    }
                    P.all.push_back(Pick{r, q, (int32_t)P.sj.size()});
                    P.sj.push_back(clamp_job(r, q, INT32_MIN / 8));
                }
            }
        });
        std::vector<Pick> all;
        JobVec sj;
        for (Part &P : parts) {
            const int32_t base = (int32_t)sj.size();
            for (Pick x : P.all) {
                if (x.job >= 0) x.job += base;
                all.push_back(x);
            }
            sj.insert(sj.end(), P.sj.begin(), P.sj.end());
        }
        kira::trace_mark("supp_picks");
        kira::PodVec<kira_sw_result> res(sj.size());
        std::vector<uint32_t> cig;
        kira::BandedRun supp_run(ctx);
        kira_status kst = banded_launch(supp_run, sj, chain_offsets(sj), kira::kPinBanded);
        if (kst != KIRA_OK) return kst;
        // XS while the device aligns the supplementary chains (it reads and
        // sets fields the picks above do not use)
        assign_xs();
        kira::trace_mark("xs");
        kst = supp_run.fetch(res.data(), cig);
        if (kst != KIRA_OK) return kst;
        kira::trace_mark("supp_sw");
        const size_t a_supp = arenas.size();
        build((int64_t)all.size(), [&](int64_t i, Arena &A) {
            const Pick &p = all[i];
            const kira_chain &ch = chains[p.chain];
            kira_stage4_aln a;
            if (p.job < 0) {
                a = exact_aln(p.read, ch, A);
            } else {
                const Job &j = sj[p.job];
                a = sw_aln(p.read, ch, j.win, j.wlen, res[p.job], cig.data() + res[p.job].cigar_off,
                           A);
            }
            a.is_supplementary = 1;
            A.add(p.read, a);
        });
        st.supplementary += (int64_t)all.size();
        // regroup: each read's supplementary records follow its others; the
        // picks are in read order, so the new records are too
        std::vector<int64_t> supp_h;
        supp_h.reserve(all.size());
        for (size_t ai = a_supp; ai < arenas.size(); ai++)
            for (size_t k = 0; k < arenas[ai]->recs.size(); k++) supp_h.push_back(rec_handle(ai, k));
        std::vector<int64_t> sfirst((size_t)n + 1, 0);
        for (const Pick &x : all) sfirst[(size_t)x.read + 1]++;
        for (int32_t r = 0; r < n; r++) sfirst[r + 1] += sfirst[r];
        std::vector<int64_t> first2((size_t)n + 1), perm2(perm.size() + supp_h.size());
        for (int32_t r = 0; r <= n; r++) first2[r] = first[r] + sfirst[r];
        parallel_chunks(n, 2048, [&](int, int64_t rb, int64_t re) {
            for (int64_t r = rb; r < re; r++) {
                int64_t o = first2[r];
                for (int64_t k = first[r]; k < first[r + 1]; k++) perm2[o++] = perm[k];
                for (int64_t k = sfirst[r]; k < sfirst[r + 1]; k++) perm2[o++] = supp_h[k];
            }
        });
        first.swap(first2);
        perm.swap(perm2);
    } else {
        assign_xs();
        kira::trace_mark("xs");
    }

    kira::trace_mark("supp_records");
    R.aln_off = first;
    // the result: CIGAR words and MD text of the arenas back to back, the
    // records in read order with their offsets moved accordingly
    {
        const size_t na = arenas.size();
        std::vector<int64_t> cig_base(na + 1, 0), md_base(na + 1, 0);
        for (size_t a = 0; a < na; a++) {
            cig_base[a + 1] = cig_base[a] + (int64_t)arenas[a]->cig.size();
            md_base[a + 1] = md_base[a] + (int64_t)arenas[a]->md.size();
        }
        R.cigar.resize((size_t)cig_base[na]);
        R.md.resize((size_t)md_base[na]);
        parallel_chunks((int64_t)na, 1, [&](int, int64_t b, int64_t e) {
            for (int64_t a = b; a < e; a++) {
                const Arena &A = *arenas[(size_t)a];
                if (!A.cig.empty())
                    memcpy(R.cigar.data() + cig_base[a], A.cig.data(), A.cig.size() * 4);
                if (!A.md.empty()) memcpy(&R.md[(size_t)md_base[a]], A.md.data(), A.md.size());
            }
        });
        R.alns.resize(perm.size());
        parallel_chunks((int64_t)perm.size(), 2048, [&](int, int64_t b, int64_t e) {
            for (int64_t k = b; k < e; k++) {
                const int64_t h = perm[k];
                kira_stage4_aln a = rec(h);
                a.cigar_off += cig_base[(size_t)(h >> 32)];
                a.md_off += md_base[(size_t)(h >> 32)];
                R.alns[k] = a;
            }
        });
    }
    for (int32_t r = 0; r < n; r++) st.unmapped += R.aln_off[r + 1] == R.aln_off[r] ? 1 : 0;
    kira::trace_mark("final");
    R.phase_ms[4] = ms_since(t0);
    return KIRA_OK;
}

}  // namespace

extern "C" {

kira_status kira_stage4_run(kira_ctx *ctx, const kira_genome *g, int32_t n_contigs,
                            const int64_t *contig_off, const int64_t *contig_len,
                            int32_t n_reads, const uint8_t *reads, const int64_t *read_off,
                            const int32_t *read_len, const int32_t *repeat_min_occ,
                            const int64_t *chain_off, const kira_chain *chains,
                            const kira_stage4_cfg *cfg, kira_stage4_result **out) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (!g || !cfg || !out || n_reads < 0 || n_contigs < 0 ||
        (n_contigs > 0 && (!contig_off || !contig_len)) ||
        (n_reads > 0 && (!reads || !read_off || !read_len || !chain_off)))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    *out = nullptr;
    kira::StepTrace trace("stage4_call");
    const int64_t glen = (int64_t)g->host.size();
    for (int32_t i = 0; i < n_contigs; i++)
        if (contig_off[i] < 0 || contig_len[i] < 0 || contig_off[i] + contig_len[i] > glen)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "contig outside the genome");
    if (n_reads > 0 && (chain_off[0] != 0 || (chain_off[n_reads] > 0 && !chains)))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid chain offsets");
    // 0 when read r and its chains are well formed, else the error
    auto check_read = [&](int32_t r) -> const char * {
        if (read_len[r] < 0 || read_off[r] < 0 || chain_off[r + 1] < chain_off[r])
            return "invalid read or chain offsets";
        for (int64_t k = chain_off[r]; k < chain_off[r + 1]; k++) {
            const kira_chain &ch = chains[k];
            if (ch.ref_id >= (uint32_t)n_contigs || contig_len[ch.ref_id] == 0 ||
                ch.read_start > (uint32_t)read_len[r] || ch.read_end > (uint32_t)read_len[r] ||
                ch.ref_start > contig_len[ch.ref_id] || ch.ref_end > contig_len[ch.ref_id] ||
                (ch.strand != 0 && ch.strand != 1))
                return "chain outside its read or contig";
        }
        return nullptr;
    };
    // in parallel over reads; the first bad read (by index) names the error
    std::atomic<int32_t> first_bad{INT32_MAX};
    try {
        kira::parallel_chunks(n_reads, 8192, [&](int, int64_t b, int64_t e) {
            for (int64_t r = b; r < e; r++)
                if (check_read((int32_t)r)) {
                    int32_t cur = first_bad.load();
                    while ((int32_t)r < cur && !first_bad.compare_exchange_weak(cur, (int32_t)r)) {
                    }
                    break;
                }
        });
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    if (first_bad.load() != INT32_MAX)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, check_read(first_bad.load()));
    kira::trace_mark("validate");
    kira_stage4_result *res = new (std::nothrow) kira_stage4_result();
    if (!res) return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "result allocation");
    kira_status st;
    try {
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) {
            delete res;
            return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
        }
        Driver d{ctx,      g,        contig_off, contig_len, n_reads, reads, read_off,
                 read_len, chain_off, chains,    *cfg,       *res,    ctx->stream};
        d.rmo = repeat_min_occ;
        d.mem.ctx = ctx;
        st = d.run();
        kira::trace_mark("run");
    } catch (const std::bad_alloc &) {
        st = fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    kira::trace_mark("teardown");
    if (st != KIRA_OK) {
        delete res;
        return st;
    }
    *out = res;
    return KIRA_OK;
}

kira_status kira_stage4_result_view(const kira_stage4_result *r, kira_stage4_view *v) {
    if (!r || !v) return KIRA_ERR_INVALID_ARGUMENT;
    v->n_reads = (int32_t)r->aln_off.size() - 1;
    v->aln_off = r->aln_off.data();
    v->alns = r->alns.data();
    v->cigar = r->cigar.data();
    v->md = r->md.data();
    v->n_alns = (int64_t)r->alns.size();
    v->cigar_words = (int64_t)r->cigar.size();
    v->md_bytes = (int64_t)r->md.size();
    v->stats = r->stats;
    for (int k = 0; k < 5; k++) v->phase_ms[k] = r->phase_ms[k];
    return KIRA_OK;
}

void kira_stage4_result_free(kira_stage4_result *r) { delete r; }

}  // extern "C"

namespace {

// append_reverse_complement (src/io/mod.rs:1509-1533): IUPAC-aware, case kept
uint8_t sam_comp(uint8_t b) {
    switch (b) {
    case 'A': return 'T';
    case 'a': return 't';
    case 'C': return 'G';
    case 'c': return 'g';
    case 'G': return 'C';
    case 'g': return 'c';
    case 'T': case 'U': return 'A';
    case 't': case 'u': return 'a';
    case 'R': return 'Y';
    case 'Y': return 'R';
    case 'S': return 'S';
    case 'W': return 'W';
    case 'K': return 'M';
    case 'M': return 'K';
    case 'B': return 'V';
    case 'D': return 'H';
    case 'H': return 'D';
    case 'V': return 'B';
    default: return 'N';
    }
}

void put_u(std::string &o, uint64_t v) { o += std::to_string(v); }
void put_i(std::string &o, int64_t v) { o += std::to_string(v); }

// extract_rg_id (src/io/mod.rs:1788-1799)
std::string rg_id(const char *rg) {
    std::string t(rg);
    if (t.compare(0, 4, "@RG\t") == 0) t = t.substr(4);
    size_t p = 0;
    while (p <= t.size()) {
        size_t e = t.find('\t', p);
        if (e == std::string::npos) e = t.size();
        if (t.compare(p, 3, "ID:") == 0 && e - p >= 3) return t.substr(p + 3, e - p - 3);
        p = e + 1;
    }
    return t;
}

const char kOpChar[] = "MIDNS";

}  // namespace

extern "C" {

kira_status kira_stage4_left_normalize(kira_stage4_result *r, const kira_genome *g,
                                       int32_t n_contigs, const int64_t *contig_off,
                                       const int64_t *contig_len, int32_t n_reads,
                                       const uint8_t *reads, const int64_t *read_off,
                                       const int32_t *read_len, const uint8_t *original,
                                       const uint8_t *has_original) {
    if (!r || !g || n_contigs < 0 || (n_contigs > 0 && (!contig_off || !contig_len)) ||
        n_reads != (int32_t)r->aln_off.size() - 1 ||
        (n_reads > 0 && (!reads || !read_off || !read_len)))
        return KIRA_ERR_INVALID_ARGUMENT;
    const int64_t glen = (int64_t)g->host.size();
    for (int32_t i = 0; i < n_contigs; i++)
        if (contig_off[i] < 0 || contig_len[i] < 0 || contig_off[i] + contig_len[i] > glen)
            return KIRA_ERR_INVALID_ARGUMENT;
    for (const kira_stage4_aln &a : r->alns)
        if (a.ref_id >= (uint32_t)n_contigs) return KIRA_ERR_INVALID_ARGUMENT;
    try {
        // normalize_alignments (normalize.rs:197-249): every record of every
        // read; NM / MD recomputed for the records that moved
        kira::PodVec<uint32_t> cig;
        std::string md;
        cig.reserve(r->cigar.size());
        md.reserve(r->md.size());
        std::vector<uint8_t> rc, orc;
        std::vector<uint32_t> ops;
        for (int32_t k = 0; k < n_reads; k++) {
            const uint8_t *fwd = reads + read_off[k];
            const int32_t L = read_len[k];
            // ReadRecord::original_seq: the read's IUPAC letters, which NM / MD
            // are recomputed from for every record (normalize.rs:213-244)
            const uint8_t *orig =
                (original && has_original && has_original[k]) ? original + read_off[k] : nullptr;
            bool have_rc = false, have_orc = false;
            for (int64_t q = r->aln_off[k]; q < r->aln_off[k + 1]; q++) {
                kira_stage4_aln &a = r->alns[q];
                ops.assign(r->cigar.begin() + a.cigar_off,
                           r->cigar.begin() + a.cigar_off + a.n_cigar);
                const uint8_t *read = fwd;
                if (a.is_rev) {
                    if (!have_rc) {
                        rc.resize((size_t)L);
                        for (int32_t t = 0; t < L; t++) rc[t] = comp_base(fwd[L - 1 - t]);
                        have_rc = true;
                    }
                    read = rc.data();
                }
                const uint8_t *contig = g->host.data() + contig_off[a.ref_id];
                const bool moved =
                    kira::left_normalize(ops, contig, contig_len[a.ref_id], a.ref_start, read, L);
                const int64_t c0 = (int64_t)cig.size(), m0 = (int64_t)md.size();
                cig.insert(cig.end(), ops.begin(), ops.end());
                const uint8_t *tags_read = read;
                if (orig) {
                    tags_read = orig;
                    if (a.is_rev) {
                        if (!have_orc) {
                            orc.resize((size_t)L);
                            for (int32_t t = 0; t < L; t++) orc[t] = sam_comp(orig[L - 1 - t]);
                            have_orc = true;
                        }
                        tags_read = orc.data();
                    }
                }
                if (moved || orig)
                    a.nm = kira::nm_md(tags_read, L, contig, contig_len[a.ref_id], a.ref_start,
                                       ops.data(), (int32_t)ops.size(), md);
                else
                    md.append(r->md, (size_t)a.md_off, (size_t)a.md_len);
                a.cigar_off = c0;
                a.n_cigar = (int32_t)ops.size();
                a.md_off = m0;
                a.md_len = (int32_t)(md.size() - m0);
            }
        }
        r->cigar.swap(cig);
        r->md.swap(md);
    } catch (const std::bad_alloc &) {
        return KIRA_ERR_OUT_OF_MEMORY;
    }
    return KIRA_OK;
}

kira_status kira_stage4_sam(const kira_stage4_result *r, int32_t n_contigs,
                            const char *const *contig_names, int32_t n_reads,
                            const char *const *read_ids, const uint8_t *seqs,
                            const int64_t *seq_off, const int32_t *seq_len, const uint8_t *quals,
                            const kira_sam_cfg *cfg, char *out, int64_t cap, int64_t *used) {
    if (!r || !cfg || !used || n_contigs < 0 || (n_contigs > 0 && !contig_names) ||
        n_reads != (int32_t)r->aln_off.size() - 1 ||
        (n_reads > 0 && (!read_ids || !seqs || !seq_off || !seq_len)))
        return KIRA_ERR_INVALID_ARGUMENT;
    for (const kira_stage4_aln &a : r->alns)
        if (a.ref_id >= (uint32_t)n_contigs) return KIRA_ERR_INVALID_ARGUMENT;
    const std::string rg = (cfg->write_rg && cfg->read_group) ? rg_id(cfg->read_group) : "";
    const bool have_rg = cfg->write_rg && cfg->read_group;
    std::vector<std::string> parts((size_t)host_threads());
    std::vector<char> bad(parts.size(), 0);
    try {
        parallel_chunks(n_reads, 2048, [&](int t, int64_t b, int64_t e) {
            std::string &o = parts[t];
            for (int64_t k = b; k < e; k++) {
                const uint8_t *seq = seqs + seq_off[k];
                const int64_t n = seq_len[k];
                const uint8_t *qual = quals ? quals + seq_off[k] : nullptr;
                if (r->aln_off[k] == r->aln_off[k + 1]) {
                    // append_unmapped (io/mod.rs:971-1033) without a mate
                    o += read_ids[k];
                    o += "\t4\t*\t0\t0\t*\t*\t0\t0\t";
                    if (n == 0) o += '*';
                    else o.append((const char *)seq, (size_t)n);
                    o += '\t';
                    if (qual && n > 0) o.append((const char *)qual, (size_t)n);
                    else o += '*';
                    o += '\n';
                    continue;
                }
                for (int64_t q = r->aln_off[k]; q < r->aln_off[k + 1]; q++) {
                    // append_alignment (io/mod.rs:1036-1140), single-end
                    const kira_stage4_aln &a = r->alns[q];
                    const uint32_t *ops = r->cigar.data() + a.cigar_off;
                    const int32_t nc = a.n_cigar;
                    uint64_t consumed = 0;
                    for (int32_t i = 0; i < nc; i++) {
                        const uint32_t op = ops[i] & 15u;
                        if (op == kOpM || op == kOpI || op == kOpS) consumed += ops[i] >> 4;
                    }
                    if (nc == 0 || consumed != (uint64_t)n) {  // checked_cigar (:1737-1750)
                        bad[t] = 1;
                        return;
                    }
                    uint16_t flag = 0;
                    if (a.is_rev) flag |= 0x10;
                    if (a.is_supplementary) flag |= cfg->split_as_secondary ? 0x100 : 0x800;
                    // bwa-mem hard-clips supplementary segments unless -Y
                    const bool hard = a.is_supplementary && !cfg->soft_clip_supplementary;
                    uint32_t lead = 0, trail = 0;
                    if (hard) {
                        if ((ops[0] & 15u) == kOpS) lead = ops[0] >> 4;
                        if (nc > 1 && (ops[nc - 1] & 15u) == kOpS) trail = ops[nc - 1] >> 4;
                    }
                    const int64_t ke = n > (int64_t)trail ? n - trail : 0;
                    const int64_t ks = std::min<int64_t>(lead, ke);
                    o += read_ids[k];
                    o += '\t';
                    put_u(o, flag);
                    o += '\t';
                    o += contig_names[a.ref_id];
                    o += '\t';
                    put_u(o, (uint64_t)a.ref_start + 1);
                    o += '\t';
                    put_u(o, (uint32_t)a.mapq);
                    o += '\t';
                    for (int32_t i = 0; i < nc; i++) {
                        const uint32_t op = ops[i] & 15u;
                        put_u(o, ops[i] >> 4);
                        o += (op == kOpS && hard && (i == 0 || i + 1 == nc)) ? 'H' : kOpChar[op];
                    }
                    o += "\t*\t0\t0\t";
                    if (ks == ke) {
                        o += '*';
                    } else if (a.is_rev) {
                        for (int64_t i = n - ks - 1; i >= n - ke; i--) o += (char)sam_comp(seq[i]);
                    } else {
                        o.append((const char *)seq + ks, (size_t)(ke - ks));
                    }
                    o += '\t';
                    if (!qual || ks == ke) {
                        o += '*';
                    } else if (a.is_rev) {
                        for (int64_t i = n - ks - 1; i >= n - ke; i--) o += (char)qual[i];
                    } else {
                        o.append((const char *)qual + ks, (size_t)(ke - ks));
                    }
                    if (cfg->write_nm) {
                        o += "\tNM:i:";
                        put_u(o, a.nm);
                    }
                    if (cfg->write_md) {
                        o += "\tMD:Z:";
                        o.append(r->md, (size_t)a.md_off, (size_t)a.md_len);
                    }
                    if (cfg->write_as) {
                        o += "\tAS:i:";
                        put_i(o, a.score);
                    }
                    if (cfg->write_xs && a.has_xs) {
                        o += "\tXS:i:";
                        put_i(o, a.xs_score);
                    }
                    if (have_rg) {
                        o += "\tRG:Z:";
                        o += rg;
                    }
                    o += '\n';
                }
            }
        });
    } catch (const std::bad_alloc &) {
        return KIRA_ERR_OUT_OF_MEMORY;
    }
    for (char x : bad)
        if (x) return KIRA_ERR_INVALID_ARGUMENT;
    int64_t total = 0;
    for (const std::string &x : parts) total += (int64_t)x.size();
    *used = total;
    if (!out) return KIRA_OK;
    if (total > cap) return KIRA_ERR_INVALID_ARGUMENT;
    int64_t pos = 0;
    for (const std::string &x : parts) {
        memcpy(out + pos, x.data(), x.size());
        pos += (int64_t)x.size();
    }
    return KIRA_OK;
}

}  // extern "C"
