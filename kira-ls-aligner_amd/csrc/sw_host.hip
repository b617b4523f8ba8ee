// Host side of the banded SW entry points: job packetisation (the device
// counterpart of the reference's per-job rayon loop, stage4_alignment.rs:1196-1205),
// trace slab sizing, and the host-array batch entry banded_sw_public
// (src/alignment/mod.rs:891-920).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace {

// Job classes.  Lane-per-job: a job with band w runs in the smallest
// C >= 2w + 1 of kLaneC, with the band stored as packed int16 when its
// scores provably fit, else as int32 (class = band index * 2 + packed).
// Cooperative: kCoop[k].g lanes per job with P = kCoop[k].p positions a lane
// (G P >= 2w + 1), in int32 when the costs keep its scan and arg-max key
// exact, or computed in packed int16 pairs when every value fits int16.
// Bands above 64 (C = 257 would spill the register file) run cooperatively
// where the costs allow, otherwise C = 257 lane per job.
constexpr int kLaneC[] = {9, 17, 33, 65, 129, 257};
constexpr int kNumLane = 6;
struct CoopShape {
    int p, g;
};
constexpr CoopShape kCoop[] = {{5, 16}, {9, 16}, {17, 16}, {13, 32}, {17, 32},
                               // packed int16 forms: p positions = 2 x the registers
                               {6, 16}, {10, 16}, {18, 16}, {14, 32}, {18, 32},
                               {10, 8}, {18, 8}};
constexpr int kNumCoop = 12;
constexpr int kNumCoop32 = kira::kSwCoop32;  // kCoop[0..4] int32, kCoop[5..11] packed int16
constexpr int kNumClasses = 2 * kNumLane + kNumCoop;
static_assert(kNumClasses == kira::kSwClasses, "class table");
constexpr int kMaxBand = (32 * 17 - 1) / 2;  // 271
constexpr int kMinLaneJobs = 65536;  // one 64-job wave on each of the 1024 SIMDs
// 8-lane packed jobs fill a wave with 8 jobs: 32768 of them are 4 waves a
// SIMD.  Fewer leave SIMDs idle and the 16-lane shape (twice the waves) wins:
// 5.8k long windows of bands 40-71 took 2.5 ms as (10, 16) jobs, 3.0 ms as
// (18, 8); 100k pairs of 100-300 bp took 2.7 ms as (18, 8) jobs at band 64
// (3.3 as (10, 16)) and 1.9 ms as (10, 8) at band 32 (2.1 lane per job).
constexpr int kNarrowMinJobs = 32768;
constexpr int64_t kLongJobCells = 1 << 16;

bool class_is_coop(int c) { return c >= 2 * kNumLane; }
bool class_is_coop16(int c) { return c >= 2 * kNumLane + kNumCoop32; }
CoopShape class_coop(int c) { return kCoop[c - 2 * kNumLane]; }
int class_jobs_per_packet(int c) { return class_is_coop(c) ? 64 / class_coop(c).g : 64; }
int class_rowdw(int c) {
    return class_is_coop(c) ? (class_coop(c).p + 7) / 8 : (kLaneC[c / 2] + 7) / 8;
}

int lane_class_of(int32_t w) {
    int need = 2 * w + 1;
    for (int k = 0; k < kNumLane; k++)
        if (kLaneC[k] >= need) return k;
    return -1;
}

// the class of cooperative shape (p, g)
int coop_class_by_shape(int p, int g) {
    for (int k = 0; k < kNumCoop; k++)
        if (kCoop[k].p == p && kCoop[k].g == g) return 2 * kNumLane + k;
    return -1;
}

// the first cooperative shape of the int32 or the packed int16 form (table
// order: 16-lane shapes by size, then 32-lane, then 8-lane) that holds the
// band
int coop_class_of(int32_t w, bool packed) {
    for (int k = packed ? kNumCoop32 : 0; k < (packed ? kNumCoop : kNumCoop32); k++)
        if (kCoop[k].g * kCoop[k].p >= 2 * w + 1) return 2 * kNumLane + k;
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

// Packs each job's CIGAR (left at src_off[k] in a slot sized for the worst
// case) to dst_off[k], so only the ops themselves cross PCIe.
__global__ void compact_cigar_kernel(int32_t n, const int32_t *__restrict__ res,
                                     const uint32_t *__restrict__ src,
                                     const int64_t *__restrict__ src_off,
                                     const int64_t *__restrict__ dst_off,
                                     uint32_t *__restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t c = res[k * kira::kSwResWords + 6];
    const uint32_t *a = src + src_off[k];
    uint32_t *b = dst + dst_off[k];
    for (int32_t i = 0; i < c; i++) b[i] = a[i];
}

// Exclusive prefix sum of the jobs' CIGAR op counts (negative = overflow,
// counted as 0) in one workgroup of 16 waves: wave w owns jobs
// [n w / 16, n (w + 1) / 16) and walks them 64 at a time (lane i reads job
// base + i, so a wave's loads are adjacent), first summing, then, after the
// waves' sums are scanned, scanning each group of 64 across the lanes with a
// running carry.
__global__ __launch_bounds__(1024) void cigar_offsets_kernel(int32_t n,
                                                             const int32_t *__restrict__ res,
                                                             int64_t *__restrict__ off) {
    __shared__ int64_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t lo = (int64_t)n * w / 16, hi = (int64_t)n * (w + 1) / 16;
    auto count = [&](int64_t k) -> int64_t {
        if (k >= hi) return 0;
        const int32_t c = res[k * kira::kSwResWords + 6];
        return c > 0 ? c : 0;
    };
    int64_t sum = 0;
    for (int64_t base = lo; base < hi; base += 64) sum += count(base + lane);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) wsum[w] = sum;
    __syncthreads();
    int64_t carry = 0;
    for (int i = 0; i < w; i++) carry += wsum[i];
    for (int64_t base = lo; base < hi; base += 64) {
        const int64_t k = base + lane;
        const int64_t c = count(k);
        int64_t x = c;  // inclusive scan of the group across the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (k < hi) off[k] = carry + x - c;
        carry += __shfl(x, 63, 64);
    }
}

// device allocations released together
using kira::DevMem;

}  // namespace

namespace kira {

hipError_t launch_cigar_compact(int32_t n, const int32_t *res, const uint32_t *arena,
                                const int64_t *arena_off, int64_t *dst_off, uint32_t *dst,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    cigar_offsets_kernel<<<1, 1024, 0, s>>>(n, res, dst_off);
    compact_cigar_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, res, arena, arena_off,
                                                                     dst_off, dst);
    return hipGetLastError();
}

}  // namespace kira

namespace kira {

// Band bucket edges: the largest band of each bucket.  Each edge is where the
// smallest lane class (2w + 1 <= 9, 17, 33, 65, 129, 257) or the set of
// cooperative shapes that hold 2w + 1 (80, 144, 272, 416, 544 positions in
// int32, 96, 160, 288, 448, 576 packed) changes, so every band of a bucket
// has the same eligible routes.
constexpr int32_t kBandEdges[] = {4,  8,   16,  32,  39,  47,  64, 71,
                                  79, 128, 135, 143, 207, 223, 271};
constexpr int kNumBandBuckets = sizeof(kBandEdges) / sizeof(kBandEdges[0]);

int sw_band_bucket(int32_t w) {
    for (int k = 0; k < kNumBandBuckets; k++)
        if (w <= kBandEdges[k]) return k;
    return kNumBandBuckets;
}

int sw_len_bucket(int32_t read_len) {
    int b = 0;
    for (uint32_t v = (uint32_t)std::max(read_len, 1); v > 1; v >>= 1) b++;
    return std::min(b, 31);
}

int sw_count_bucket(int64_t jobs) {
    int b = 0;
    for (int64_t v = std::max<int64_t>(jobs, 1); v >= 4; v >>= 2) b++;
    return std::min(b, 15);
}

uint32_t sw_tune_key(int band_bucket, int len_bucket, int count_bucket) {
    return ((uint32_t)band_bucket << 16) | ((uint32_t)len_bucket << 8) | (uint32_t)count_bucket;
}

int sw_num_coop_shapes() { return kNumCoop; }

void sw_coop_shape(int k, int *P, int *G) {
    *P = kCoop[k].p;
    *G = kCoop[k].g;
}

namespace {
// The plan's cost bounds (below): the 16-lane jobs keep their F scan and
// arg-max key exact only with costs <= 2^20 and read_len * match < 2^21.
bool coop_costs_ok(const kira_sw_cfg &cfg) {
    const int32_t go = sat_add(cfg.gap_open, cfg.gap_extend);
    const int32_t kCoopCost = 1 << 20;
    return cfg.match_score >= 0 && cfg.mismatch >= 0 && cfg.gap_open >= 0 &&
           cfg.gap_extend >= 0 && cfg.match_score <= kCoopCost && cfg.mismatch <= kCoopCost &&
           go <= kCoopCost;
}
}  // namespace

// The packed int16 cooperative form: its x4 keyed values (4 value + tag)
// stay inside int16 (scores <= 8191, mismatch <= 8000, open + extend <=
// 4000, extend <= 500), and gap open > 0 (its F-extend bit compares against
// the pre-F H, which equals the reference's comparison only then).
bool coop16_ok(int32_t read_len, const kira_sw_cfg &cfg) {
    const int32_t go = sat_add(cfg.gap_open, cfg.gap_extend);
    return cfg.match_score >= 0 && cfg.mismatch >= 0 && cfg.gap_open > 0 &&
           cfg.gap_extend >= 0 && cfg.mismatch <= 8000 && go <= 4000 &&
           cfg.gap_extend <= 500 && (int64_t)read_len * cfg.match_score <= 8191;
}

uint32_t sw_routes_for(int32_t w, int32_t read_len, const kira_sw_cfg &cfg) {
    w = std::max(w, 1);
    uint32_t r = 1u << kSwRouteDefault;
    if (lane_class_of(w) >= 0) r |= 1u << kSwRouteLane;
    if (coop_costs_ok(cfg) && (int64_t)read_len * cfg.match_score < (1 << 21))
        for (int k = 0; k < kNumCoop32; k++)
            if (kCoop[k].g * kCoop[k].p >= 2 * w + 1) r |= 1u << (kSwRouteCoop + k);
    if (coop16_ok(read_len, cfg))
        for (int k = kNumCoop32; k < kNumCoop; k++)
            if (kCoop[k].g * kCoop[k].p >= 2 * w + 1) r |= 1u << (kSwRouteCoop + k);
    return r;
}

const char *sw_plan_layout(int32_t n, const int32_t *read_len, const int32_t *band,
                           const kira_sw_cfg &cfg, SwPlanLayout &L, const SwTuneTable *tune,
                           int force_route) {
    L = SwPlanLayout();
    L.n = n;
    std::vector<int> cls(n);
    const int32_t go = sat_add(cfg.gap_open, cfg.gap_extend);
    // packed bands: also bounds the diagonal candidate (>= -mismatch) so the
    // kernel's keyed maximum (value << 2 | priority) stays inside int32
    const bool costs_small = cfg.match_score >= 0 && cfg.mismatch >= 0 && cfg.gap_open >= 0 &&
                             cfg.gap_extend >= 0 && go <= 16000 && cfg.mismatch <= 16000;
    // 16-lane jobs: the F scan stays inside int32 and the arg-max key
    // (score << 10 | position) needs score < 2^21
    const bool coop_costs = coop_costs_ok(cfg);
    // 0: lane per job only, 1: int32 cooperative shapes, 2: packed ones too
    std::vector<uint8_t> coop_ok(n);
    for (int32_t k = 0; k < n; k++) {
        int32_t w = std::max(band ? band[k] : cfg.bandwidth, 1);
        if (read_len[k] < 0) return "negative read length";
        // every H is at most read_len * match: int16 storage cannot clamp
        bool pack = costs_small && (int64_t)read_len[k] * cfg.match_score < 32767;
        bool coop = coop_costs && (int64_t)read_len[k] * cfg.match_score < (1 << 21);
        const bool c16 = coop && coop16_ok(read_len[k], cfg);
        coop_ok[k] = coop ? (c16 ? 2 : 1) : 0;
        // the widest band every form holds (the int32 (17, 32) shape); the
        // packed (18, 32) shape would hold 287 but the limit does not depend
        // on the costs
        if (w > kMaxBand) return "band wider than 271";
        // bands of 17 .. 71 run as packed 8- or 16-lane jobs where the costs
        // allow (decided below by how many there are), wider ones as packed
        // 16- or 32-lane jobs: at band 64, 100k pairs of 100-300 bp took
        // 3.7 ms lane per job (C = 129), 3.3 ms as (10, 16) and 2.7 ms as
        // (18, 8) int16 jobs
        if (c16 && w > 16 && w <= 71) {
            cls[k] = -1;
            continue;
        }
        if (w > 64 && coop) {
            cls[k] = coop_class_of(w, c16);
            continue;
        }
        // one long lane-per-job job serialises read_len x (2w + 1) cells on
        // one lane and becomes the straggler of its class: run it on 16
        // lanes instead
        if (coop && w > 16 && (int64_t)read_len[k] * (2 * w + 1) > kLongJobCells &&
            coop_class_of(w, c16) >= 0) {
            cls[k] = coop_class_of(w, c16);
            continue;
        }
        int b = lane_class_of(w);
        if (b < 0)
            return w <= kMaxBand ? "band above 128 needs costs <= 2^20 and read_len * match < 2^21"
                                 : "band wider than 271";
        cls[k] = 2 * b + (pack ? 1 : 0);
    }
    {
        int64_t narrow[2] = {0, 0};  // packed jobs of bands 17 .. 39, 40 .. 71
        for (int32_t k = 0; k < n; k++)
            if (cls[k] < 0) narrow[std::max(band ? band[k] : cfg.bandwidth, 1) > 39]++;
        const int c8[2] = {coop_class_by_shape(10, 8), coop_class_by_shape(18, 8)};
        for (int32_t k = 0; k < n; k++) {
            if (cls[k] >= 0) continue;
            const int32_t w = std::max(band ? band[k] : cfg.bandwidth, 1);
            cls[k] = narrow[w > 39] >= kNarrowMinJobs ? c8[w > 39] : coop_class_of(w, true);
        }
    }
    // A lane-per-job class runs one wave per 64 jobs, so a class of fewer
    // than 64 x 1024 jobs leaves SIMDs idle, and a lone wave on a SIMD waits
    // on each dependent instruction and LDS read of its row loop (37k jobs
    // of band 32 took 4.8 ms lane per job, 2.3 ms as (9, 16) jobs).  Such
    // jobs run as 16-lane jobs (16x the waves) in the smallest 16-lane
    // class that holds the band, packed int16 where the costs allow.
    {
        int count[kNumClasses] = {};
        for (int32_t k = 0; k < n; k++) count[cls[k]]++;
        const int c65 = 2 * 3, c65p = 2 * 3 + 1, c129 = 2 * 4, c129p = 2 * 4 + 1;
        for (int32_t k = 0; k < n; k++) {
            const int c = cls[k];
            if ((c == c65 || c == c65p || c == c129 || c == c129p) && count[c] < kMinLaneJobs &&
                coop_ok[k])
                cls[k] = coop_class_of(std::max(band ? band[k] : cfg.bandwidth, 1),
                                       coop_ok[k] == 2);
        }
    }
    // tuned (or forced) routes override the built-in choice for the jobs that
    // may take them; the shape key's count is the batch's number of jobs in
    // the job's band bucket (the jobs a kernel class shares the chip with)
    if (force_route >= 0 || (tune && !tune->route.empty())) {
        int64_t per_band[kNumBandBuckets + 1] = {};
        std::vector<uint32_t> gk(n);
        for (int32_t k = 0; k < n; k++) {
            const int32_t w = std::max(band ? band[k] : cfg.bandwidth, 1);
            const int bb = sw_band_bucket(w);
            gk[k] = sw_tune_key(bb, sw_len_bucket(read_len[k]), 0);
            per_band[bb]++;
        }
        for (int32_t k = 0; k < n; k++) {
            int route = force_route;
            if (route < 0) {
                const int bb = (int)(gk[k] >> 16);
                auto it = tune->route.find(gk[k] | (uint32_t)sw_count_bucket(per_band[bb]));
                if (it == tune->route.end()) continue;
                route = it->second;
            }
            const int32_t w = std::max(band ? band[k] : cfg.bandwidth, 1);
            if (route <= kSwRouteDefault || !((sw_routes_for(w, read_len[k], cfg) >> route) & 1u))
                continue;
            if (route == kSwRouteLane) {
                const bool pack =
                    costs_small && (int64_t)read_len[k] * cfg.match_score < 32767;
                cls[k] = 2 * lane_class_of(w) + (pack ? 1 : 0);
            } else {
                cls[k] = 2 * kNumLane + (route - kSwRouteCoop);
            }
        }
    }
    // packets: jobs sorted by (band class, read length descending), stable
    std::vector<int32_t> order;
    {
        constexpr int32_t kLenMax = (1 << 27) - 1;
        static_assert(kNumClasses <= 32, "class field of the sort key");
        bool fits = true;
        std::vector<uint32_t> key(n);
        for (int32_t k = 0; k < n; k++) {
            fits &= read_len[k] <= kLenMax;
            key[k] = ((uint32_t)cls[k] << 27) | (uint32_t)(kLenMax - std::min(read_len[k], kLenMax));
        }
        if (fits) {
            stable_order_by_key(key.data(), n, order);
        } else {
            order.resize(n);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
                if (cls[x] != cls[y]) return cls[x] < cls[y];
                return read_len[x] > read_len[y];
            });
        }
    }
    L.job_slot.assign(n, 0);
    for (int c = 0; c < kNumClasses; c++) L.class_first[c] = -1;
    {
        // packets: each class's jobs in order, class_jobs_per_packet(c) a packet
        int64_t per_class[kNumClasses] = {}, n_pk = 0;
        for (int32_t k = 0; k < n; k++) per_class[cls[k]]++;
        for (int c = 0; c < kNumClasses; c++)
            n_pk += (per_class[c] + class_jobs_per_packet(c) - 1) / class_jobs_per_packet(c);
        L.packet_jobs.resize((size_t)n_pk * 64);
        L.packet_rows.resize((size_t)n_pk);
        L.packet_rowdw.resize((size_t)n_pk);
        L.packet_trace_off.resize((size_t)n_pk);
        L.packet_meta_off.resize((size_t)n_pk);
    }
    int32_t pk = 0;
    for (int32_t s = 0; s < n; pk++) {
        int c = cls[order[s]];
        const int per = class_jobs_per_packet(c), stride = 64 / per;
        int32_t e = s;
        while (e < n && e - s < per && cls[order[e]] == c) e++;
        if (L.class_first[c] < 0) L.class_first[c] = pk;
        L.class_count[c]++;
        int32_t rows = read_len[order[s]];
        int32_t rowdw = class_rowdw(c);
        int32_t *pj = L.packet_jobs.data() + (size_t)pk * 64;
        for (int32_t l = 0; l < 64; l++) {
            int32_t job = (l % stride == 0 && s + l / stride < e) ? order[s + l / stride] : -1;
            pj[l] = job;
            if (job >= 0) L.job_slot[job] = pk * 64 + l;
        }
        L.packet_rows[pk] = rows;
        L.packet_rowdw[pk] = rowdw;
        L.packet_trace_off[pk] = L.trace_words;
        L.packet_meta_off[pk] = L.meta_rows;
        L.trace_words += (int64_t)rows * rowdw * 64;
        L.meta_rows += (int64_t)rows * 64;
        s = e;
    }
    L.n_packets = pk;
    L.cigar_cap.resize(n);
    L.cigar_off.resize(n);
    for (int32_t k = 0; k < n; k++) {
        L.cigar_cap[k] = 3 * read_len[k] + 2;
        L.cigar_off[k] = L.cigar_words;
        L.cigar_words += L.cigar_cap[k];
    }
    return nullptr;
}

namespace {
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
// table sections in order: packet_jobs, packet_rows, packet_rowdw,
// packet_trace_off, packet_meta_off, job_slot, cigar_cap, cigar_off
struct TableSections {
    size_t off[8], bytes[8], total;
};
TableSections table_sections(const SwPlanLayout &L) {
    TableSections t{};
    size_t b[8] = {L.packet_jobs.size() * 4,      L.packet_rows.size() * 4,
                   L.packet_rowdw.size() * 4,     L.packet_trace_off.size() * 8,
                   L.packet_meta_off.size() * 8,  L.job_slot.size() * 4,
                   L.cigar_cap.size() * 4,        L.cigar_off.size() * 8};
    size_t o = 0;
    for (int k = 0; k < 8; k++) {
        t.off[k] = o;
        t.bytes[k] = b[k];
        o += align256(std::max<size_t>(b[k], 1));
    }
    t.total = o;
    return t;
}
}  // namespace

size_t sw_plan_table_bytes(const SwPlanLayout &L) { return table_sections(L).total; }

void sw_plan_pack_tables(const SwPlanLayout &L, uint8_t *dst) {
    TableSections t = table_sections(L);
    const void *src[8] = {L.packet_jobs.data(),      L.packet_rows.data(),
                          L.packet_rowdw.data(),     L.packet_trace_off.data(),
                          L.packet_meta_off.data(),  L.job_slot.data(),
                          L.cigar_cap.data(),        L.cigar_off.data()};
    for (int k = 0; k < 8; k++)
        if (t.bytes[k]) memcpy(dst + t.off[k], src[k], t.bytes[k]);
}

size_t sw_plan_slab_bytes(const SwPlanLayout &L) {
    return align256((size_t)std::max<int64_t>(L.trace_words, 1) * 4) +
           align256((size_t)std::max<int64_t>(L.meta_rows, 1) * sizeof(int2)) +
           align256(kNumClasses * 4);
}

SwPlanDev sw_plan_bind(const SwPlanLayout &L, const uint8_t *tab, uint8_t *slabs) {
    TableSections t = table_sections(L);
    SwPlanDev d{};
    d.packet_jobs = (const int32_t *)(tab + t.off[0]);
    d.packet_rows = (const int32_t *)(tab + t.off[1]);
    d.packet_rowdw = (const int32_t *)(tab + t.off[2]);
    d.packet_trace_off = (const int64_t *)(tab + t.off[3]);
    d.packet_meta_off = (const int64_t *)(tab + t.off[4]);
    d.job_slot = (const int32_t *)(tab + t.off[5]);
    d.cigar_cap = (const int32_t *)(tab + t.off[6]);
    d.cigar_off = (const int64_t *)(tab + t.off[7]);
    size_t tb = align256((size_t)std::max<int64_t>(L.trace_words, 1) * 4);
    size_t mb = align256((size_t)std::max<int64_t>(L.meta_rows, 1) * sizeof(int2));
    d.trace = (uint32_t *)slabs;
    d.meta = (int2 *)(slabs + tb);
    d.counter = (int32_t *)(slabs + tb + mb);
    return d;
}

hipError_t sw_fanout_create(SwFanout &f) {
    hipError_t e = hipEventCreateWithFlags(&f.fork, hipEventDisableTiming);
    for (int k = 0; !e && k < kSwFan; k++) {
        e = hipStreamCreateWithFlags(&f.aux[k], hipStreamNonBlocking);
        if (!e) e = hipEventCreateWithFlags(&f.join[k], hipEventDisableTiming);
    }
    f.ok = !e;
    if (e) sw_fanout_destroy(f);
    return e;
}

void sw_fanout_destroy(SwFanout &f) {
    for (int k = 0; k < kSwFan; k++) {
        if (f.aux[k]) (void)hipStreamDestroy(f.aux[k]);
        if (f.join[k]) (void)hipEventDestroy(f.join[k]);
        f.aux[k] = nullptr;
        f.join[k] = nullptr;
    }
    if (f.fork) (void)hipEventDestroy(f.fork);
    f.fork = nullptr;
    f.ok = false;
}

const char *sw_plan_launch(const SwPlanLayout &L, const SwPlanDev &d, const kira_sw_jobs_dev &j,
                           const kira_sw_cfg &cfg, const kira_sw_knobs &kn,
                           const uint8_t *skip_flags, const int32_t *skip_src, hipStream_t s,
                           hipError_t *err, SwFanout *fan) {
    *err = hipSuccess;
    if (L.n == 0) return nullptr;
    int32_t go = sat_add(cfg.gap_open, cfg.gap_extend);
    // pads are INT32_MIN/4 + cost and must stay negative
    if (go >= (1 << 29) || cfg.gap_extend >= (1 << 29) || go < 0 || cfg.gap_extend < 0)
        return "gap costs outside the supported range";
    SwLaunch a{};
    a.match_score = cfg.match_score;
    a.mismatch = cfg.mismatch;
    a.go = go;
    a.gap_extend = cfg.gap_extend;
    a.clip_penalty = cfg.clip_penalty;
    a.bandwidth = cfg.bandwidth;
    a.adaptive = kn.adaptive ? 1 : 0;
    a.max_drift = kn.max_drift;
    a.dp_xdrop = kn.dp_xdrop >= 0 ? kn.dp_xdrop : cfg.xdrop;
    a.reads = j.reads;
    a.read_off = j.read_off;
    a.read_len = j.read_len;
    a.refs = j.refs;
    a.ref_off = j.ref_off;
    a.ref_len = j.ref_len;
    a.offset = j.offset;
    a.band = j.band;
    a.abort_score = j.abort_score;
    a.trace = d.trace;
    a.row_meta = d.meta;
    a.res = j.res;
    a.job_slot = d.job_slot;
    a.cigar = j.cigar;
    a.cigar_off = d.cigar_off;
    a.cigar_cap = d.cigar_cap;
    a.skip_flags = skip_flags;
    a.skip_src = skip_flags ? skip_src : nullptr;
    // Classes holding long jobs (packets of >= kFanMinRows rows; a class's
    // first packet is its longest) go round-robin to the fan-out's streams,
    // the rest in order to s.  The tail of a class of long jobs (its last
    // few waves, each running for milliseconds) then overlaps other classes;
    // classes of short jobs have short tails and lose more to sharing the
    // chip (1M pairs of 100-300 bp: 23.8 ms in order, 24.8 ms fanned out;
    // 82k windows of 0.5-5 kbp: 87.9 ms in order, 86.7 ms fanned out).
    constexpr int32_t kFanMinRows = 1024;
    int n_long = 0;
    for (int c = 0; c < kNumClasses; c++)
        n_long += L.class_count[c] > 0 && L.packet_rows[L.class_first[c]] >= kFanMinRows;
    const int n_aux = (fan && fan->ok) ? std::min(n_long, kSwFan) : 0;
    if (n_aux > 0) {
        hipError_t e = hipEventRecord(fan->fork, s);
        for (int k = 0; !e && k < n_aux; k++) e = hipStreamWaitEvent(fan->aux[k], fan->fork, 0);
        if (e) {
            *err = e;
            return "banded SW fan-out";
        }
    }
    int i_long = 0;
    for (int c = 0; c < kNumClasses; c++) {
        if (L.class_count[c] == 0) continue;
        hipStream_t cs = s;
        if (n_aux > 0 && L.packet_rows[L.class_first[c]] >= kFanMinRows)
            cs = fan->aux[i_long++ % n_aux];
        SwLaunch ac = a;
        int32_t f = L.class_first[c];
        ac.packet_jobs = d.packet_jobs + (int64_t)f * 64;
        ac.packet_rows = d.packet_rows + f;
        ac.packet_rowdw = d.packet_rowdw + f;
        ac.packet_trace_off = d.packet_trace_off + f;
        ac.packet_meta_off = d.packet_meta_off + f;
        ac.n_packets = L.class_count[c];
        ac.counter = d.counter + c;
        hipError_t e;
        if (class_is_coop16(c)) {
            int waves = std::min(L.class_count[c], 256 * 20);
            e = launch_banded_sw_coop16(ac, class_coop(c).p / 2, class_coop(c).g, waves, cs);
        } else if (class_is_coop(c)) {
            int waves = std::min(L.class_count[c], 256 * 20);
            e = launch_banded_sw_coop(ac, class_coop(c).p, class_coop(c).g, waves, cs);
        } else {
            int waves = std::min(L.class_count[c], 256 * 16);
            e = launch_banded_sw(ac, kLaneC[c / 2], (c & 1) != 0, waves, cs);
        }
        if (e != hipSuccess) {
            *err = e;
            return "banded SW launch";
        }
    }
    // s waits for every fan-out stream that took a class
    for (int k = 0; k < n_aux; k++) {
        hipError_t e = hipEventRecord(fan->join[k], fan->aux[k]);
        if (!e) e = hipStreamWaitEvent(s, fan->join[k], 0);
        if (e) {
            *err = e;
            return "banded SW fan-in";
        }
    }
    return nullptr;
}

}  // namespace kira

struct kira_sw_plan {
    kira::SwPlanLayout L;
    kira::SwPlanDev d{};
    void *mem = nullptr;
    int64_t device_bytes = 0;
};

extern "C" {

kira_status kira_banded_sw_plan_create(kira_ctx *ctx, int32_t n, const int32_t *read_len,
                                       const int32_t *band, const kira_sw_cfg *cfg,
                                       kira_sw_plan **out) {
    if (!ctx || !out || !cfg || n < 0 || (n > 0 && !read_len))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    *out = nullptr;
    kira_sw_plan *p = new (std::nothrow) kira_sw_plan();
    if (!p) return KIRA_ERR_OUT_OF_MEMORY;
    if (const char *why = kira::sw_plan_layout(n, read_len, band, *cfg, p->L, &ctx->tune)) {
        delete p;
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, why);
    }
    const size_t tb = kira::sw_plan_table_bytes(p->L), sb = kira::sw_plan_slab_bytes(p->L);
    p->device_bytes = (int64_t)(tb + sb);
    std::vector<uint8_t> host(tb);
    kira::sw_plan_pack_tables(p->L, host.data());
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(&p->mem, tb + sb);
    if (e != hipSuccess) {
        delete p;
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "plan allocation", e);
    }
    uint8_t *base = (uint8_t *)p->mem;
    if ((e = hipMemcpyAsync(base, host.data(), tb, hipMemcpyHostToDevice, ctx->stream)) ||
        (e = hipStreamSynchronize(ctx->stream))) {
        (void)hipFree(p->mem);
        delete p;
        return fail(ctx, KIRA_ERR_RUNTIME, "plan upload", e);
    }
    p->d = kira::sw_plan_bind(p->L, base, base + tb);
    *out = p;
    return KIRA_OK;
}

void kira_banded_sw_plan_destroy(kira_sw_plan *plan) {
    if (!plan) return;
    if (plan->mem) (void)hipFree(plan->mem);
    delete plan;
}

int64_t kira_banded_sw_plan_cigar_words(const kira_sw_plan *plan) {
    return plan ? plan->L.cigar_words : 0;
}

void kira_banded_sw_plan_route_counts(const kira_sw_plan *plan, int32_t *counts) {
    if (!counts) return;
    for (int r = 0; r < kira::kSwRouteCoop + kNumCoop; r++) counts[r] = 0;
    if (!plan) return;
    for (int c = 0; c < kNumClasses; c++) {
        int32_t jobs = 0;
        for (int32_t p = 0; p < plan->L.class_count[c]; p++) {
            const int32_t pk = plan->L.class_first[c] + p;
            for (int l = 0; l < 64; l++) jobs += plan->L.packet_jobs[(size_t)pk * 64 + l] >= 0;
        }
        counts[class_is_coop(c) ? kira::kSwRouteCoop + (c - 2 * kNumLane) : kira::kSwRouteLane] +=
            jobs;
    }
}

void kira_banded_sw_plan_job_class(const kira_sw_plan *plan, int32_t *cls) {
    if (!plan || !cls) return;
    for (int c = 0; c < kNumClasses; c++)
        for (int32_t p = 0; p < plan->L.class_count[c]; p++) {
            const int32_t pk = plan->L.class_first[c] + p;
            for (int l = 0; l < 64; l++) {
                const int32_t job = plan->L.packet_jobs[(size_t)pk * 64 + l];
                if (job >= 0) cls[job] = c;
            }
        }
}

int64_t kira_banded_sw_plan_device_bytes(const kira_sw_plan *plan) {
    return plan ? plan->device_bytes : 0;
}

const int64_t *kira_banded_sw_plan_cigar_off(const kira_sw_plan *plan) {
    return plan ? plan->L.cigar_off.data() : nullptr;
}

kira_status kira_banded_sw_run_dev(kira_ctx *ctx, kira_sw_plan *plan,
                                   const kira_sw_jobs_dev *j, const kira_sw_cfg *cfg,
                                   const kira_sw_knobs *kn) {
    if (!ctx || !plan || !j || !cfg || !kn)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    hipError_t e;
    const char *why = kira::sw_plan_launch(plan->L, plan->d, *j, *cfg, *kn, nullptr, nullptr,
                                           ctx->stream, &e, &ctx->fan);
    if (why) return fail(ctx, e ? KIRA_ERR_RUNTIME : KIRA_ERR_INVALID_ARGUMENT, why, e);
    return KIRA_OK;
}

}  // extern "C"

namespace kira {

kira_status fetch_sw_results(kira_ctx *ctx, hipStream_t s, DevMem &dm, int32_t n,
                             const int32_t *d_res, const uint32_t *d_cig, const int64_t *d_src_off,
                             int64_t cig_words, bool banded, kira_sw_result *out,
                             std::vector<uint32_t> &cigar) {
    if (n == 0) return KIRA_OK;
    hipError_t e;
    int64_t *d_dst = nullptr;
    uint32_t *d_out = nullptr;
    if ((e = dm.alloc(&d_dst, n)) || (e = dm.alloc(&d_out, (size_t)std::max<int64_t>(cig_words, 1))))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "CIGAR compaction allocation", e);
    if ((e = launch_cigar_compact(n, d_res, d_cig, d_src_off, d_dst, d_out, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "CIGAR compaction", e);
    std::vector<int32_t> res((size_t)n * kSwResWords);
    if ((e = hipMemcpyAsync(res.data(), d_res, res.size() * 4, hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
    trace_mark(".res_sync");
    int64_t total = 0;
    for (int32_t k = 0; k < n; k++) {
        const int32_t *r = &res[(size_t)k * kSwResWords];
        if (r[6] < 0) return fail(ctx, KIRA_ERR_RUNTIME, "CIGAR capacity exceeded");
        if (banded && r[7] != 0) return fail(ctx, KIRA_ERR_RUNTIME, "int16 band overflow");
        kira_sw_result &o = out[k];
        o.score = r[0];
        o.read_end = r[1];
        o.ref_end = (uint32_t)r[2];
        o.early_abort = r[3];
        o.ref_start = (uint32_t)r[4];
        o.read_start = r[5];
        o.n_cigar = r[6];
        o.cells = r[8];
        o.cigar_off = (int64_t)cigar.size() + total;
        total += r[6];
    }
    const size_t base = cigar.size();
    cigar.resize(base + (size_t)total);
    if (total > 0 &&
        ((e = hipMemcpyAsync(cigar.data() + base, d_out, (size_t)total * 4, hipMemcpyDeviceToHost,
                             s)) ||
         (e = hipStreamSynchronize(s))))
        return fail(ctx, KIRA_ERR_RUNTIME, "CIGAR download", e);
    return KIRA_OK;
}

namespace {
size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

BandedRun::~BandedRun() {
    if (ctx && stream && stream != ctx->stream) (void)hipStreamSynchronize(stream);
}

kira_status BandedRun::launch(const SwJobsHost &j, const kira_sw_cfg &cfg,
                              const kira_sw_knobs &kn, int slot, hipStream_t s_in) {
    stream = s_in ? s_in : ctx->stream;
    n = j.n;
    if (n == 0) return KIRA_OK;
    if (const char *why = sw_plan_layout(n, j.read_len, j.band, cfg, L, &ctx->tune))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, why);
    trace_mark(".bsw_plan");
    // plan tables, then the job arrays, in one pinned block and one copy
    const size_t tb = sw_plan_table_bytes(L), sb = sw_plan_slab_bytes(L);
    const size_t n8 = al256((size_t)n * 8), n4 = al256((size_t)n * 4);
    const size_t o_ro = tb, o_fo = o_ro + n8, o_rl = o_fo + n8, o_fl = o_rl + n4,
                 o_off = o_fl + n4, o_ab = o_off + n4, o_band = o_ab + n4,
                 total = o_band + (j.band ? n4 : 0);
    void *hp = nullptr;
    hipError_t e;
    if ((e = pinned_buf(ctx, slot, total, &hp)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "banded SW staging", e);
    uint8_t *h = (uint8_t *)hp;
    sw_plan_pack_tables(L, h);
    memcpy(h + o_ro, j.read_off, (size_t)n * 8);
    memcpy(h + o_fo, j.ref_off, (size_t)n * 8);
    memcpy(h + o_rl, j.read_len, (size_t)n * 4);
    memcpy(h + o_fl, j.ref_len, (size_t)n * 4);
    memcpy(h + o_off, j.offset, (size_t)n * 4);
    memcpy(h + o_ab, j.abort_score, (size_t)n * 4);
    if (j.band) memcpy(h + o_band, j.band, (size_t)n * 4);
    uint8_t *d_blob = nullptr, *d_slab = nullptr;
    if ((e = dm.alloc(&d_blob, total)) || (e = dm.alloc(&d_slab, sb)) ||
        (e = dm.alloc(&d_res, (size_t)n * kSwResWords)) ||
        (e = dm.alloc(&d_cig, (size_t)std::max<int64_t>(L.cigar_words, 1))))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "banded SW allocation", e);
    hipStream_t s = stream;
    if ((e = hipMemcpyAsync(d_blob, h, total, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    const SwPlanDev d = sw_plan_bind(L, d_blob, d_slab);
    d_cig_off = d.cigar_off;
    kira_sw_jobs_dev jd{j.d_reads,
                        (const int64_t *)(d_blob + o_ro),
                        (const int32_t *)(d_blob + o_rl),
                        j.d_refs,
                        (const int64_t *)(d_blob + o_fo),
                        (const int32_t *)(d_blob + o_fl),
                        (const int32_t *)(d_blob + o_off),
                        j.band ? (const int32_t *)(d_blob + o_band) : nullptr,
                        (const int32_t *)(d_blob + o_ab),
                        d_res,
                        d_cig};
    if (const char *why = sw_plan_launch(L, d, jd, cfg, kn, nullptr, nullptr, s, &e, &ctx->fan))
        return fail(ctx, e ? KIRA_ERR_RUNTIME : KIRA_ERR_INVALID_ARGUMENT, why, e);
    trace_mark(".bsw_launch");
    return KIRA_OK;
}

kira_status BandedRun::fetch(kira_sw_result *out, std::vector<uint32_t> &cigar) {
    const kira_status st = fetch_sw_results(ctx, stream ? stream : ctx->stream, dm, n, d_res, d_cig,
                                            d_cig_off, L.cigar_words, true, out, cigar);
    trace_mark(".bsw_fetch");
    return st;
}

kira_status banded_sw_dev(kira_ctx *ctx, const SwJobsHost &j, const kira_sw_cfg &cfg,
                          const kira_sw_knobs &kn, kira_sw_result *out,
                          std::vector<uint32_t> &cigar) {
    BandedRun run(ctx);
    kira_status st = run.launch(j, cfg, kn, kPinBanded);
    if (st != KIRA_OK) return st;
    return run.fetch(out, cigar);
}

}  // namespace kira

extern "C" {

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
    DevMem dm(ctx);
    uint8_t *d_reads, *d_refs;
    hipError_t e;
    hipStream_t s = ctx->stream;
    std::vector<int64_t> foff(n);
    for (int32_t k = 0; k < n; k++) foff[k] = ref_off[k] + kira::kRefPad;
    if ((e = dm.alloc(&d_reads, (size_t)reads_bytes + 8)) ||
        (e = dm.alloc(&d_refs, (size_t)refs_bytes + 2 * kira::kRefPad)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "batch allocation", e);
    if ((e = hipMemsetAsync(d_refs, 0, (size_t)refs_bytes + 2 * kira::kRefPad, s)) ||
        (reads_bytes && (e = hipMemcpyAsync(d_reads, reads, (size_t)reads_bytes,
                                            hipMemcpyHostToDevice, s))) ||
        (refs_bytes && (e = hipMemcpyAsync(d_refs + kira::kRefPad, refs, (size_t)refs_bytes,
                                           hipMemcpyHostToDevice, s))))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    kira::SwJobsHost j{n, d_reads, read_off, read_len, d_refs, foff.data(), ref_len, offset, band,
                       abort_score};
    std::vector<uint32_t> cig;
    kira_status st = kira::banded_sw_dev(ctx, j, *cfg, *kn, out, cig);
    if (st != KIRA_OK) return st;
    if ((int64_t)cig.size() > arena_cap || (!cig.empty() && !cigar_arena))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "cigar arena too small");
    if (!cig.empty()) memcpy(cigar_arena, cig.data(), cig.size() * 4);
    return KIRA_OK;
}

}  // extern "C"
