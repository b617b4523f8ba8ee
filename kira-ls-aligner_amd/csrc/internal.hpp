// Internal declarations shared by the HIP translation units of libkira_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <type_traits>
#include <mutex>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "kira_hip.h"

namespace kira {

// An allocator whose value-less construct() default-initialises, so that
// resize() leaves plain-data elements untouched: a large array's pages are
// then first touched (faulted in) by whoever fills it, in parallel, instead
// of being zeroed by resize() on one thread.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <class U>
    DefaultInitAlloc(const DefaultInitAlloc<U> &) {}
    template <class U>
    void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using PodVec = std::vector<T, DefaultInitAlloc<T>>;

// Genome / window text as bit planes: base p lives at bit (p & 31) of word
// (p >> 5) + kPadFront of each plane.  lo/hi = the two bits of the 2-bit code
// (A=00 C=01 G=10 T=11, bitpacked.rs:12-20); inv = not A/C/G/T; lc = lower
// case A/C/G/T (only byte-level certification needs it).
constexpr int kPadFront = 2;
constexpr int kPadBack = 16;

// 2-bit sequence as bit planes, 32 bases per word: lh[w] = (low code bit,
// high code bit) - the sieve scan reads only these, 8 bytes a word; il[w] =
// (non-ACGT flag, lower-case flag) for the skip check and certification;
// nz = one bit per word, set when the word holds a non-ACGT base, so the
// skip check usually touches a few bits of a small, cache-resident array.
// kPadFront / kPadBack zero words surround the sequence.  il / nz may be
// null (windows packed from 2-bit data carry no flags).
struct Planes {
    uint2 *lh = nullptr;
    uint2 *il = nullptr;
    uint32_t *nz = nullptr;
    int64_t n_bases = 0;
    int64_t n_words = 0;  // words including padding
};

// Read planes: 24 words per read — lo[8], hi[8], lc[8] (reads <= 256 bases).
constexpr int kReadWords = 24;
constexpr int kMaxReadBases = 256;

struct SieveLaunch {
    int32_t n_jobs;
    const uint32_t *read_planes;
    const int32_t *read_lens;
    const uint8_t *read_valid;
    const int32_t *job_read;
    int32_t jobs_per_read;
    const int64_t *win_start;
    const int32_t *win_len;
    const int32_t *limit;
    int32_t limit_all;
    int32_t *out_shift, *out_best, *out_second;
    uint8_t *out_flags;
    int32_t cert_max_mism, cert_strict;
    const uint2 *g_lh, *g_il;  // window text planes (Planes); g_il null: no flags
    const uint32_t *g_nz;
    int64_t g_bases;
};

hipError_t launch_pack_ascii_planes(const uint8_t *seq, int64_t n, Planes p, hipStream_t s);
hipError_t launch_pack_2bit_planes(const uint8_t *packed, int64_t n_bytes, Planes p,
                                   hipStream_t s);
// Read i starts at reads + (offs ? offs[i] : i * stride).
hipError_t launch_pack_reads_ascii(int32_t n, const uint8_t *reads, int32_t stride,
                                   const int64_t *offs, const int32_t *lens, uint32_t *planes,
                                   uint8_t *valid, hipStream_t s);
hipError_t launch_pack_reads_2bit(int32_t n, const uint8_t *packed, int32_t stride,
                                  const int32_t *lens, uint32_t *planes, hipStream_t s);
hipError_t launch_sieve(const SieveLaunch &a, int read_words, hipStream_t s);

// Device address of the fixed staging array of kira_sieve_run_batch_noalloc
// on `device`, looked up once per device (the first call makes `device`
// current for the lookup and restores the previous device).
hipError_t noalloc_setup(int device, void **out);

// Banded SW: results per job, kSwResWords int32 each: score, read_end,
// ref_end, early_abort, ref_start, read_start, n_cigar, int16-overflow flag,
// DP cells the reference algorithm evaluates for the job, 0.
constexpr int kSwResWords = 10;
// Reference arena padding (bytes) so band rows that poke past a window's
// ends stay inside the allocation.
constexpr int64_t kRefPad = 1024;

struct SwLaunch {
    int32_t match_score, mismatch, go, gap_extend, clip_penalty;
    int32_t bandwidth, adaptive, max_drift, dp_xdrop;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *refs;
    const int64_t *ref_off;
    const int32_t *ref_len;
    const int32_t *offset;
    const int32_t *band;  // per job, may be null (bandwidth)
    const int32_t *abort_score;
    // work packets of 64 jobs (one wave each), pulled through *counter
    const int32_t *packet_jobs;
    const int32_t *packet_rows;
    const int32_t *packet_rowdw;
    const int64_t *packet_trace_off;
    const int64_t *packet_meta_off;
    int32_t n_packets;
    int32_t *counter;
    uint32_t *trace;
    int2 *row_meta;
    int32_t *res;
    const int32_t *job_slot;
    uint32_t *cigar;
    const int64_t *cigar_off;
    const int32_t *cigar_cap;
    // optional: job k is dropped (result all zero, res[9] = 1) when
    // skip_src[k] >= 0 and skip_flags[skip_src[k]] has the certified bit (4)
    const uint8_t *skip_flags;
    const int32_t *skip_src;
};

// Result of a job dropped through SwLaunch::skip_src.
__host__ __device__ inline bool sw_job_skipped(const SwLaunch &a, int32_t job) {
    if (!a.skip_src) return false;
    const int32_t src = a.skip_src[job];
    return src >= 0 && (a.skip_flags[src] & 4u) != 0;
}
__host__ __device__ inline void sw_write_skipped(const SwLaunch &a, int32_t job) {
    int32_t *res = a.res + (int64_t)job * kSwResWords;
    for (int k = 0; k < kSwResWords; k++) res[k] = 0;
    res[9] = 1;
}

hipError_t launch_banded_sw(const SwLaunch &a, int C, bool pack, int n_waves, hipStream_t s);
// Wide bands: G lanes per job (16 or 32), P band positions per lane
// ((P, G) in (5, 16), (9, 16), (17, 16), (13, 32), (17, 32)).
hipError_t launch_banded_sw_coop(const SwLaunch &a, int P, int G, int n_waves, hipStream_t s);
// The packed int16 form: 2 P2 positions a lane as P2 registers of int16 pairs
// ((P2, G) in (3, 16), (5, 16), (9, 16), (7, 32), (9, 32), (5, 8), (9, 8)).
hipError_t launch_banded_sw_coop16(const SwLaunch &a, int P2, int G, int n_waves, hipStream_t s);

// Host-side packetisation of a banded SW batch (jobs sorted by band class
// and read length into packets, trace / meta slab offsets, CIGAR arena).
constexpr int kSwClasses = 24;
struct SwPlanLayout {
    int32_t n = 0, n_packets = 0;
    int32_t class_first[kSwClasses] = {}, class_count[kSwClasses] = {};
    std::vector<int32_t> packet_jobs, packet_rows, packet_rowdw, job_slot, cigar_cap;
    std::vector<int64_t> packet_trace_off, packet_meta_off, cigar_off;
    int64_t trace_words = 0, meta_rows = 0, cigar_words = 0;
};
// Routing of banded SW jobs chosen by the autotuner per shape key (band
// bucket, read-length bucket, job-count bucket; sw_tune_key).  A route is
// kSwRouteLane (one lane per job, int16 or int32 band by the score bound) or
// kSwRouteCoop + k (cooperative shape k of the (P, G) table: k < kSwCoop32
// the int32 form, the rest the packed int16 form); shapes without an entry
// keep the built-in routing.
constexpr int kSwRouteDefault = 0, kSwRouteLane = 1, kSwRouteCoop = 2, kSwCoop32 = 5;
struct SwTuneTable {
    std::unordered_map<uint32_t, uint8_t> route;
};
// Bucket of a band: the bands of one bucket have the same eligible routes.
int sw_band_bucket(int32_t w);
// Bucket of a read length (log2) and of a job count (log4).
int sw_len_bucket(int32_t read_len);
int sw_count_bucket(int64_t jobs);
uint32_t sw_tune_key(int band_bucket, int len_bucket, int count_bucket);
// Routes a job of band w may take under cfg (bit r set: route r eligible;
// bit kSwRouteDefault always set).
uint32_t sw_routes_for(int32_t w, int32_t read_len, const kira_sw_cfg &cfg);
// Number of cooperative shapes and shape k's (P, G).
int sw_num_coop_shapes();
void sw_coop_shape(int k, int *P, int *G);

// Packs the CIGARs a plan launch left in its arena (job k at arena_off[k],
// res[k * kSwResWords + 6] ops) one after another: dst_off[k] receives job k's
// offset in dst (an exclusive prefix sum of the op counts, computed on the
// device) and its ops are copied there.  dst may be pinned host memory.
hipError_t launch_cigar_compact(int32_t n, const int32_t *res, const uint32_t *arena,
                                const int64_t *arena_off, int64_t *dst_off, uint32_t *dst,
                                hipStream_t s);

// Auxiliary streams a plan launch fans its kernel classes out to, so that
// the tail of one class (its last, longest jobs) overlaps the next class
// instead of idling the chip; joined back into the caller's stream with
// events.  Owned by a context and used by one call at a time.
constexpr int kSwFan = 3;
struct SwFanout {
    hipStream_t aux[kSwFan] = {};
    hipEvent_t fork = nullptr, join[kSwFan] = {};
    bool ok = false;
};
hipError_t sw_fanout_create(SwFanout &f);
void sw_fanout_destroy(SwFanout &f);

// nullptr on success, else the reason the batch cannot run.  With a tuning
// table, jobs whose shape key has an entry take the tuned route (if eligible);
// force_route >= 0 sends every job that is eligible for it there (the tuner's
// candidates).
const char *sw_plan_layout(int32_t n, const int32_t *read_len, const int32_t *band,
                           const kira_sw_cfg &cfg, SwPlanLayout &out,
                           const SwTuneTable *tune = nullptr, int force_route = -1);
// The layout's tables as one block (host copy written by sw_plan_pack_tables,
// device copy bound by sw_plan_bind; both 256-byte aligned sections).
size_t sw_plan_table_bytes(const SwPlanLayout &L);
void sw_plan_pack_tables(const SwPlanLayout &L, uint8_t *dst);
// Device bytes of the trace and meta slabs plus the class counters.
size_t sw_plan_slab_bytes(const SwPlanLayout &L);
struct SwPlanDev {
    const int32_t *packet_jobs, *packet_rows, *packet_rowdw, *job_slot, *cigar_cap;
    const int64_t *packet_trace_off, *packet_meta_off, *cigar_off;
    uint32_t *trace;
    int2 *meta;
    int32_t *counter;
};
SwPlanDev sw_plan_bind(const SwPlanLayout &L, const uint8_t *dev_tables, uint8_t *dev_slabs);
// All launches of one batch, ordered on stream s (nullptr on success, else
// the error and *e the HIP status).  With a fan-out, the kernel classes are
// spread over s and the fan-out's streams, which join s again.
const char *sw_plan_launch(const SwPlanLayout &L, const SwPlanDev &d, const kira_sw_jobs_dev &j,
                           const kira_sw_cfg &cfg, const kira_sw_knobs &kn,
                           const uint8_t *skip_flags, const int32_t *skip_src, hipStream_t s,
                           hipError_t *e, SwFanout *fan = nullptr);

// Full-window lane-group SW (align_batch_simd's SIMD kernels).
struct GroupLaunch {
    int32_t match_score, mismatch, go, gap_extend, clip_penalty;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *refs;
    const int64_t *ref_off;
    const int32_t *ref_len;
    const int32_t *abort_score;
    const int32_t *group_of;  // per job
    const int32_t *neg;       // per group: row-0 E / column-0 F sentinel of the lane type
    int32_t *group_stop;      // per group
    const int32_t *packet_jobs;
    const int32_t *packet_rows;
    const int32_t *packet_words;
    const int64_t *packet_trace_off;
    const int64_t *packet_row_off;
    int32_t n_packets;
    int32_t *counter;
    uint32_t *trace;
    int2 *bnd;
    int2 *rstat;
    const int32_t *job_slot;
    int32_t *res;
    uint32_t *cigar;
    const int64_t *cigar_off;
    const int32_t *cigar_cap;
    // scores and costs keep value << 4 inside int32 (the keyed maximum)
    int32_t keyed;
    // the systolic DP (a job's strips on adjacent lanes; the trace kept by
    // step): tasks (packet, first slot) of 64 / strips jobs
    int32_t systolic;
    const int2 *tasks;
    int32_t n_tasks;
};

hipError_t launch_group_sw(const GroupLaunch &a, int32_t n_jobs, int32_t n_groups, int n_waves,
                           hipStream_t s);

// Device blocks cached on the context between calls, so an entry point's
// temporaries cost no hipMalloc / hipFree once warm: sizes round up to a
// quarter of their power of two (at least 256 B), a freed block waits in its
// size's list for the next request of that size.  Blocks above 1 GiB are not
// cached.  One call at a time per context (as for every context state).
hipError_t cached_alloc(kira_ctx *ctx, void **p, size_t bytes, size_t *size_class);
void cached_free(kira_ctx *ctx, void *p, size_t size_class);
// Grow-only pinned host buffer `slot` (PinSlot) of the context, kept until
// kira_release_buffers: staging for the large host-device copies, which run
// at full PCIe rate only from pinned memory.  Its contents are not kept when
// it grows.
hipError_t pinned_buf(kira_ctx *ctx, int slot, size_t bytes, void **out);

// The device allocations of one scope, back to the context's cache (or freed,
// without a context) when the scope ends.  A cached block may be handed to the
// next call at once, possibly on another stream, so the scope first waits for
// the context stream: an early error return can leave a copy or kernel that
// uses the block queued (hipFree, the uncached case, waits by itself).
struct DevMem {
    kira_ctx *ctx = nullptr;
    std::vector<std::pair<void *, size_t>> v;
    DevMem() = default;
    explicit DevMem(kira_ctx *c) : ctx(c) {}
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem();
    template <class T>
    hipError_t alloc(T **p, size_t count) {
        void *q = nullptr;
        size_t cls = 0;
        hipError_t e = cached_alloc(ctx, &q, std::max<size_t>(count * sizeof(T), 1), &cls);
        if (e == hipSuccess) {
            *p = (T *)q;
            v.emplace_back(q, cls);
        }
        return e;
    }
};

// SW batches over device-resident sequences with host-side job tables: job
// k's read is d_reads + read_off[k] (read_len[k] bytes, 8 readable bytes
// past the arena end), its window d_refs + ref_off[k] (ref_len[k] bytes;
// banded SW needs kRefPad readable bytes on each side).  offset / band /
// abort_score as in kira_banded_sw_batch (band may be null).
struct SwJobsHost {
    int32_t n;
    const uint8_t *d_reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *d_refs;
    const int64_t *ref_off;
    const int32_t *ref_len;
    const int32_t *offset;
    const int32_t *band;
    const int32_t *abort_score;
};

// Reads back n results (kSwResWords each) and their CIGARs, which the
// kernels left at d_src_off[k] (device array) of d_cig: a device prefix sum
// and gather pack the CIGARs back to back into scratch from dm first, and
// only the ops cross PCIe; they are appended to `cigar` and out[k].cigar_off
// indexes it.  Waits for the context's stream.
kira_status fetch_sw_results(kira_ctx *ctx, hipStream_t s, DevMem &dm, int32_t n,
                             const int32_t *d_res, const uint32_t *d_cig, const int64_t *d_src_off,
                             int64_t cig_words, bool banded, kira_sw_result *out,
                             std::vector<uint32_t> &cigar);

// A banded SW batch (banded_sw_with, src/alignment/mod.rs:1370-1650) over
// device-resident reads and windows, in two steps: launch() lays the plan
// out on the host, uploads its tables with the job arrays in one copy from
// pinned buffer `slot` and queues the kernels on `stream` (the context's
// stream when null), then returns; fetch() waits for them and reads the
// results back.  Device memory comes from the context's block cache and is
// kept until the object goes (which first waits for its stream).
struct BandedRun {
    kira_ctx *ctx;
    DevMem dm;
    SwPlanLayout L;
    hipStream_t stream = nullptr;
    int32_t n = 0;
    int32_t *d_res = nullptr;
    uint32_t *d_cig = nullptr;
    const int64_t *d_cig_off = nullptr;
    explicit BandedRun(kira_ctx *c) : ctx(c), dm(c) {}
    BandedRun(const BandedRun &) = delete;
    BandedRun &operator=(const BandedRun &) = delete;
    ~BandedRun();
    kira_status launch(const SwJobsHost &j, const kira_sw_cfg &cfg, const kira_sw_knobs &kn,
                       int slot, hipStream_t s = nullptr);
    kira_status fetch(kira_sw_result *out, std::vector<uint32_t> &cigar);
};
kira_status banded_sw_dev(kira_ctx *ctx, const SwJobsHost &j, const kira_sw_cfg &cfg,
                          const kira_sw_knobs &kn, kira_sw_result *out,
                          std::vector<uint32_t> &cigar);
// align_batch_simd's dispatch (src/alignment/mod.rs:1118-1149) of batches
// batch_off[b] .. batch_off[b+1]: group_of[k] = lane group of job k, or -1
// for the banded SW of BatchInput::scalar_sw; group_neg = each group's
// lane-type minus infinity.
void simd_groups(int32_t n_batches, const int32_t *batch_off, const uint32_t *read_len,
                 const uint32_t *win_len, kira_simd_mode mode, kira_simd_mode detected,
                 const kira_sw_cfg &cfg, std::vector<int32_t> &group_of,
                 std::vector<int32_t> &group_neg);
// Full-window lane groups (sw_batch_avx2 / sw_batch_int8 / sw_batch_neon)
// over a job table whose every job is in group group_of[k], in the same two
// steps as BandedRun.
struct GroupRun {
    kira_ctx *ctx;
    DevMem dm;
    int32_t n = 0;
    int64_t cig_words = 0;
    int32_t *d_res = nullptr;
    uint32_t *d_cig = nullptr;
    const int64_t *d_cig_off = nullptr;
    explicit GroupRun(kira_ctx *c) : ctx(c), dm(c) {}
    kira_status launch(const SwJobsHost &j, const int32_t *group_of,
                       const std::vector<int32_t> &group_neg, const kira_sw_cfg &cfg, int slot);
    kira_status fetch(kira_sw_result *out, std::vector<uint32_t> &cigar);
};
kira_status group_sw_dev(kira_ctx *ctx, const SwJobsHost &j, const int32_t *group_of,
                         const std::vector<int32_t> &group_neg, const kira_sw_cfg &cfg,
                         kira_sw_result *out, std::vector<uint32_t> &cigar);

struct PrefilterLaunch {
    int64_t n;
    const uint8_t *reads;
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *refs;
    const int64_t *ref_off;
    const int64_t *ref_len;
    const kira_span *spans;
    const uint8_t *is_top1;
    const int32_t *score_margin;
    int32_t match_score, mismatch, xdrop;
    uint16_t id_floor;  // identity floor after KIRA_UNGAP_MINID
    kira_prefilter_params p;
    kira_prefilter_out *out;
};

hipError_t launch_prefilter(const PrefilterLaunch &a, hipStream_t s);
// The prefilter's final gate (prefilter.rs:274-287, ungapped_beaten_by_gap,
// mod.rs:163-204) over a batch whose outcomes are on the host: every Accept
// with mismatches gets an exact WFA score-only probe of its read against
// contig[ref_start - read_start .. + read_len + bandwidth] with budget
// mismatches * (match + mismatch) - 1; a probe that finds an alignment turns
// the outcome into Fallback(IndelSuspect), metrics kept.  Chain k's read is
// d_reads + read_off[k] (read_len[k] bytes) and its contig d_refs +
// ref_off[k] (ref_len[k] bytes), device memory; the offset, length and span
// arrays are host arrays.  Synchronous.
kira_status prefilter_gap_check(kira_ctx *ctx, int32_t n, const uint8_t *d_reads,
                                const int64_t *read_off, const int32_t *read_len,
                                const uint8_t *d_refs, const int64_t *ref_off,
                                const int64_t *ref_len, const kira_span *spans,
                                const kira_sw_cfg &cfg, kira_prefilter_out *out);
// identity floor of the accept gate (prefilter.rs:255-256)
uint16_t prefilter_id_floor(const kira_prefilter_params &p);

// Alignment records (host): build_alignment (src/alignment/mod.rs:1200-1254)
// of one SwResult appends the soft-clipped CIGAR to `cig` and the MD text to
// `md` and returns NM; ungapped_tags (md.rs:147-151) does the same for an
// ungapped span of n columns.
uint32_t build_record(const uint8_t *read, int32_t L, const uint8_t *win, int64_t W,
                      const kira_sw_result &s, const uint32_t *ops, std::vector<uint32_t> &cig,
                      std::string &md);
uint32_t ungapped_tags(const uint8_t *read, const uint8_t *ref, int64_t n, std::string &md);
// compute_nm_md (src/alignment/mod.rs:1774-1785): replay `ops` from read
// position 0 and reference position ref_pos; MD appended to `md`, NM returned.
uint32_t nm_md(const uint8_t *read, int32_t L, const uint8_t *ref, int64_t ref_len, int64_t ref_pos,
               const uint32_t *ops, int32_t n, std::string &md);
// left_normalize_indels (src/alignment/normalize.rs:55-190) in place on a
// CIGAR over a whole contig and the strand-oriented read; true if it moved.
bool left_normalize(std::vector<uint32_t> &ops, const uint8_t *ref, int64_t ref_len,
                    uint32_t ref_start, const uint8_t *read, int64_t read_len);

// Pinned staging buffers of a context, one per use, so that uploads that are
// in flight at the same time never share one.
enum PinSlot {
    kPinOrient = 0,   // the stage-4 oriented reads
    kPinPfOut = 1,    // stage-4 prefilter outcomes
    kPinPfJobs = 2,   // stage-4 prefilter job arrays
    kPinFast = 3,     // the fast DP cascade's job lists
    kPinGroup = 4,    // lane-group SW job arrays and packet tables
    kPinBanded = 5,   // banded SW job arrays and plan tables
    kPinBanded2 = 6,  // a second banded SW batch in flight with the first
    kPinSieve = 7,    // stage-4 sieve job arrays
    kPinWfa = 8,      // WFA CIGAR slot tables
    kPinSlots = 9
};

// A typed view of pinned staging buffer `slot` (PinSlot) of the context.
template <class T>
struct Pinned {
    T *p = nullptr;
    size_t n = 0;
    T *data() const { return p; }
    size_t size() const { return n; }
    T &operator[](size_t i) const { return p[i]; }
    hipError_t resize(kira_ctx *ctx, int slot, size_t count) {
        void *q = nullptr;
        hipError_t e = pinned_buf(ctx, slot, count * sizeof(T), &q);
        if (e == hipSuccess) {
            p = (T *)q;
            n = count;
        }
        return e;
    }
};

// Host worker threads for record building: the machine's, at most 16.
inline int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

// Persistent host workers (host_threads() - 1 of them, started on first
// use) for parallel_chunks: starting and joining 16 threads per call cost
// about a millisecond, and a stage-4 batch makes a few dozen such calls.
// One call uses the pool at a time; run() returns false when it is busy
// (a call on another thread, or a nested call), and the caller then starts
// threads of its own.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool *p = new HostPool();  // never destroyed: workers idle until exit
        return *p;
    }
    // task(t) for every t in [0, n); the calling thread takes part.
    template <class F>
    bool run(int n, F &&task) {
        bool idle = false;
        if (!busy_.compare_exchange_strong(idle, true)) return false;
        start_workers();
        const std::function<void(int)> fn = std::ref(task);
        Job j;
        j.task = &fn;
        j.n = n;
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &j;
            gen_++;
        }
        cv_.notify_all();
        const int ran = drain(j);
        std::unique_lock<std::mutex> lk(m_);
        j.done += ran;
        done_.wait(lk, [&] { return j.done == j.n && j.inside == 0; });
        job_ = nullptr;
        busy_.store(false);
        return true;
    }

  private:
    struct Job {
        const std::function<void(int)> *task = nullptr;
        int n = 0;
        std::atomic<int> next{0};
        int inside = 0, done = 0;  // guarded by m_
    };
    std::atomic<bool> busy_{false};
    std::mutex m_;
    std::condition_variable cv_, done_;
    Job *job_ = nullptr;
    std::atomic<uint64_t> gen_{0};  // written under m_
    bool started_ = false;

    static int drain(Job &j) {
        int ran = 0;
        for (;;) {
            const int t = j.next.fetch_add(1);
            if (t >= j.n) return ran;
            (*j.task)(t);
            ran++;
        }
    }
    void start_workers() {
        if (started_) return;
        started_ = true;
        for (int w = 1; w < host_threads(); w++) std::thread([this] { worker(); }).detach();
    }
    void worker() {
        uint64_t seen = 0;
        for (;;) {
            // calls come in bursts: spin a little before sleeping, so the
            // next call finds its workers awake
            const auto t_spin = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen &&
                   std::chrono::steady_clock::now() - t_spin < std::chrono::microseconds(200))
                __builtin_ia32_pause();
            Job *j;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return job_ != nullptr && gen_.load() != seen; });
                seen = gen_.load();
                j = job_;
                j->inside++;
            }
            const int ran = drain(*j);
            std::lock_guard<std::mutex> lk(m_);
            j->done += ran;
            if (--j->inside == 0 && j->done == j->n) done_.notify_all();
        }
    }
};

// f(t, begin, end) on up to host_threads() threads over [0, n) in chunks of
// at least `grain`; chunk t is [n t / T, n (t + 1) / T), so chunks follow
// item order.  A std::bad_alloc in any chunk is rethrown here.
template <class F>
void parallel_chunks(int64_t n, int64_t grain, F &&f) {
    if (n <= 0) return;
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), (n + grain - 1) / grain));
    if (T == 1) {
        f(0, (int64_t)0, n);
        return;
    }
    std::vector<char> failed(T, 0);
    auto chunk = [&](int t) {
        try {
            f(t, n * t / T, n * (t + 1) / T);
        } catch (...) {
            failed[t] = 1;
        }
    };
    if (!HostPool::get().run(T, chunk)) {
        std::vector<std::thread> th;
        for (int t = 1; t < T; t++) th.emplace_back(chunk, t);
        chunk(0);
        for (auto &x : th) x.join();
    }
    for (int t = 0; t < T; t++)
        if (failed[t]) throw std::bad_alloc();
}

// The indices i in [0, n) with keep(i), in increasing order, into out:
// counted, then placed, in parallel chunks (keep is called twice per index).
template <class P>
void parallel_filter(int64_t n, int64_t grain, P &&keep, std::vector<int32_t> &out) {
    const int T = host_threads();
    std::vector<int64_t> cnt((size_t)T + 1, 0);
    parallel_chunks(n, grain, [&](int t, int64_t b, int64_t e) {
        int64_t c = 0;
        for (int64_t i = b; i < e; i++) c += keep(i) ? 1 : 0;
        cnt[(size_t)t + 1] = c;
    });
    for (int t = 0; t < T; t++) cnt[(size_t)t + 1] += cnt[(size_t)t];
    out.resize((size_t)cnt[(size_t)T]);
    parallel_chunks(n, grain, [&](int t, int64_t b, int64_t e) {
        int64_t o = cnt[(size_t)t];
        for (int64_t i = b; i < e; i++)
            if (keep(i)) out[(size_t)o++] = (int32_t)i;
    });
}

// Wall-clock marks of the steps of one entry-point call, printed to stderr
// when the environment variable KIRA_TRACE is set (a diagnostic: which host
// step or synchronisation a call's time goes to).  A StepTrace on the stack
// collects the marks trace_mark() makes on its thread while it lives.
struct StepTrace {
    using Clock = std::chrono::steady_clock;
    const char *name;
    bool on;
    Clock::time_point t0, last;
    std::vector<std::pair<const char *, double>> marks;
    StepTrace *outer;
    static StepTrace *&active() {
        static thread_local StepTrace *a = nullptr;
        return a;
    }
    explicit StepTrace(const char *n) : name(n), on(std::getenv("KIRA_TRACE") != nullptr) {
        t0 = last = Clock::now();
        outer = active();
        if (on) active() = this;
    }
    void mark(const char *what) {
        const Clock::time_point now = Clock::now();
        marks.emplace_back(what, std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    }
    ~StepTrace() {
        if (!on) return;
        active() = outer;
        std::string o = std::string("[kira trace] ") + name + ": total " +
                        std::to_string(std::chrono::duration<double, std::milli>(last - t0).count()) +
                        " ms;";
        for (auto &m : marks) o += std::string(" ") + m.first + " " + std::to_string(m.second);
        std::fprintf(stderr, "%s\n", o.c_str());
    }
};
inline void trace_mark(const char *what) {
    if (StepTrace *t = StepTrace::active()) t->mark(what);
}

// Indices 0 .. n-1 ordered by key[i] ascending, equal keys in index order
// (what std::stable_sort of the indices by key gives), by two stable
// counting passes over 16-bit digits.
inline void stable_order_by_key(const uint32_t *key, int32_t n, std::vector<int32_t> &order) {
    order.resize((size_t)n);
    if (n <= 0) return;
    std::vector<int32_t> tmp((size_t)n);
    std::vector<int32_t> cnt(65537);
    for (int pass = 0; pass < 2; pass++) {
        const int sh = 16 * pass;
        std::fill(cnt.begin(), cnt.end(), 0);
        const int32_t *src = pass == 0 ? nullptr : tmp.data();
        for (int32_t i = 0; i < n; i++) cnt[((key[src ? src[i] : i] >> sh) & 0xFFFFu) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        int32_t *dst = pass == 0 ? tmp.data() : order.data();
        for (int32_t i = 0; i < n; i++) {
            const int32_t x = src ? src[i] : i;
            dst[cnt[(key[x] >> sh) & 0xFFFFu]++] = x;
        }
    }
}

// memcpy split over the host threads when it is large (pinned staging of
// tens of MB per request)
inline void parallel_memcpy(void *dst, const void *src, size_t bytes) {
    if (bytes < ((size_t)8 << 20)) {
        if (bytes) std::memcpy(dst, src, bytes);
        return;
    }
    parallel_chunks((int64_t)bytes, (int64_t)4 << 20, [&](int, int64_t b, int64_t e) {
        std::memcpy((char *)dst + b, (const char *)src + b, (size_t)(e - b));
    });
}

// Myers / WFA on device arrays, shared by the batch entry points and the fast
// DP cascade.  WfaCosts::scored (src/alignment/wfa.rs:41-55): mismatch x,
// gap open o, insertion extend ie, deletion extend de; false when the config
// has none (no pair aligns).
struct WfaPen {
    int32_t x, o, ie, de;
};
bool wfa_pen(const kira_sw_cfg &cfg, WfaPen &pen);
// Pairs on the device (pattern / text bytes, offsets, lengths, WfaOptions)
// with host copies of the sizes the launch is planned from.
struct WfaDevJobs {
    int32_t n = 0;
    const uint8_t *pat = nullptr, *txt = nullptr;
    const int64_t *pat_off = nullptr, *txt_off = nullptr;
    const int32_t *pat_len = nullptr, *txt_len = nullptr, *max_score = nullptr, *tbf = nullptr;
    const int32_t *h_pat_len = nullptr, *h_txt_len = nullptr, *h_max_score = nullptr,
                  *h_tbf = nullptr;
};
constexpr int kWfaRes = 6;  // found, loss, text_start, text_end, n_cigar, status
// wfa_align_scored for every pair; res[6 k ..] as kWfaRes, the CIGARs (when
// want_cigar) back to back in cigar from cig_off[k].  Synchronous.
kira_status wfa_dev(kira_ctx *ctx, const WfaDevJobs &j, const WfaPen &pen, int32_t drop,
                    bool want_cigar, std::vector<int32_t> &res, std::vector<uint32_t> &cigar,
                    std::vector<int64_t> &cig_off);
// wfa_dev in two halves, so the host can work while the kernel runs:
// launch() enqueues the pairs on the context stream and returns; finish()
// waits and brings the results back.  The job arrays (device and host copies)
// must stay as they are until launch() returns, the device ones until
// finish(); no other work may be queued on the context stream in between.
struct WfaRun {
    kira_ctx *ctx;
    DevMem dm;
    int32_t n = 0;
    bool want_cigar = false;
    std::vector<int64_t> cig_off;
    std::vector<int32_t> cig_cap;
    int32_t *d_res = nullptr;
    int64_t *d_coff = nullptr;
    uint32_t *d_cig = nullptr;
    explicit WfaRun(kira_ctx *c) : ctx(c), dm(c) {}
    kira_status launch(const WfaDevJobs &j, const WfaPen &pen, int32_t drop, bool want_cigar);
    kira_status finish(std::vector<int32_t> &res, std::vector<uint32_t> &cigar,
                       std::vector<int64_t> &cig_off);
};
// bounded_edit_distance for n pairs on device arrays into device found /
// dist / end; h_pat_len is the host copy of pat_len.  Synchronous.
kira_status myers_dev(kira_ctx *ctx, int32_t n, const uint8_t *pat, const int64_t *pat_off,
                      const int32_t *pat_len, const int32_t *h_pat_len, const uint8_t *txt,
                      const int64_t *txt_off, const int32_t *txt_len, const int32_t *max_k,
                      uint8_t *found, int32_t *dist, int32_t *end);

// try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440) for n chains:
// chain k's read (oriented) at reads[read_off[k]], and at d_reads +
// read_off[k] on the device when d_reads is not NULL (else the reads are
// uploaded).  Arguments already validated.  res[k].kind 0 = not resolved;
// CIGAR words and MD text back to back in cig / md; without want_md the
// records carry no NM / MD (md_len 0).  Synchronous.
kira_status fast_dp_run(kira_ctx *ctx, const kira_genome *g, const int64_t *contig_off,
                        const int64_t *contig_len, int32_t n, const uint8_t *reads,
                        int64_t reads_bytes, const uint8_t *d_reads, const int64_t *read_off,
                        const int32_t *read_len, const kira_chain *chains,
                        const kira_sw_cfg *cfg, const kira_fast_knobs &kn,
                        PodVec<kira_fast_aln> &res, std::vector<uint32_t> &cig,
                        std::string &md, kira_fast_stats &stats, bool want_md);

}  // namespace kira

struct kira_ctx {
    int device = 0;
    hipStream_t stream = nullptr;      // stream launches go to
    hipStream_t own_stream = nullptr;  // the context's own stream
    std::string err;
    // geometric-growth device scratch for the reference-layout batch entry
    void *buf[16] = {};
    size_t cap[16] = {};
    // free cached device blocks by size class (kira::cached_alloc)
    std::unordered_map<size_t, std::vector<void *>> free_blocks;
    // pinned host staging buffers (kira::pinned_buf, kira::PinSlot)
    void *pin[kira::kPinSlots] = {};
    size_t pin_cap[kira::kPinSlots] = {};
    // banded SW routes chosen by the autotuner (kira_sw_autotune / kira_sw_tune_load)
    kira::SwTuneTable tune;
    // streams the banded SW kernel classes of one plan launch fan out to
    kira::SwFanout fan;
    // a second stream (null if it could not be made) for work that runs
    // beside the context stream's, and the event that orders it after it
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr;
};

struct kira_genome {
    kira_ctx *ctx = nullptr;
    kira::Planes planes;
    uint8_t *ascii = nullptr;  // base 0 of the ASCII copy; kRefPad zero bytes on each side
    std::vector<uint8_t> host;  // host copy, for alignment records (NM / MD)
};

inline kira::DevMem::~DevMem() {
    if (ctx && !v.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (auto &x : v) cached_free(ctx, x.first, x.second);
}
