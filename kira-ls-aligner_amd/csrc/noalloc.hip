// kira_sieve_run_batch_noalloc: the reference-layout sieve batch
// (CudaBackend::run_batch, src/cuda/backend.rs:96-249) without any device
// allocation.  Its staging (packed reads, phase-0 window bytes, job tables,
// read and window bit planes, results) lives in one file-scope device array
// of fixed size, so the call makes no hipMalloc / hipFree and no host heap
// allocation.  For callers that cannot allocate: code that must not call the
// allocator in a serving loop, or harnesses that forbid it.
//
// Every call on one device shares that device's array: calls on the same
// GPU that overlap in time (two host threads, two contexts) race and must be
// serialised by the caller.  A call has finished with the array when it
// returns (it synchronises its stream).
#include <stdio.h>

#include <mutex>

#include "internal.hpp"

// Fixed staging area of every device the library runs on (held until the
// process ends).
static __device__ __align__(256) uint8_t scratch[KIRA_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` per device (the array's name is only a
// host-side handle; kernels and copies need this address).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Phase-0 window bytes -> start base, limit -> min(limit, read_nucs)
// (backend.rs:185-188), in place.
__global__ void noalloc_prep_kernel(int32_t n, int64_t *win_start, int32_t *limit,
                                    const int32_t *read_nucs) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    win_start[i] *= 4;
    const uint32_t lim = (uint32_t)limit[i];
    const uint32_t nucs = (uint32_t)read_nucs[i];
    const uint32_t m = lim < nucs ? lim : nucs;
    limit[i] = (int32_t)(m > (uint32_t)INT32_MAX ? (uint32_t)INT32_MAX : m);
}

kira_status fail(kira_ctx *ctx, kira_status st, const char *what, hipError_t e = hipSuccess) {
    char buf[512];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "%s", what);
    ctx->err = buf;
    return st;
}

}  // namespace

namespace kira {

// Looks up the device address of `scratch` on `device` once (with that
// device current) and keeps it; later calls return the kept address.
hipError_t noalloc_setup(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess) e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && prev != device) e = hipSetDevice(prev);
        g_err[device] = e;
        g_ptr[device] = e == hipSuccess ? p : nullptr;
        g_done[device] = true;
    }
    if (out) *out = g_ptr[device];
    return g_err[device];
}

}  // namespace kira

extern "C" {

kira_status kira_sieve_run_batch_noalloc(kira_ctx *ctx, int32_t n, const uint8_t *reads_packed,
                                         int32_t read_bytes_max, const int32_t *read_packed_len,
                                         const int32_t *read_nucs,
                                         const uint8_t *ref_shifted_flat,
                                         int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                                         const int32_t *ref_nucs, const uint32_t *max_mismatches,
                                         int32_t *out_shifts, int32_t *out_mismatches) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (n == 0) return KIRA_OK;  // empty -> Ok(vec![]) (backend.rs:97-99)
    if (n < 0 || !reads_packed || !read_packed_len || !read_nucs || !ref_offsets || !ref_nucs ||
        !max_mismatches || !out_shifts || !out_mismatches || read_bytes_max <= 0 ||
        ref_bytes_per_phase < 0 || (ref_bytes_per_phase > 0 && !ref_shifted_flat))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    int32_t max_len = 0;
    for (int32_t i = 0; i < n; i++) {
        if (read_packed_len[i] > KIRA_READ_BYTES_MAX) {
            char buf[128];
            snprintf(buf, sizeof buf, "read too long for the GPU backend (%d > %d bytes packed)",
                     read_packed_len[i], KIRA_READ_BYTES_MAX);
            return fail(ctx, KIRA_ERR_READ_TOO_LONG, buf);
        }
        if (read_nucs[i] < 0 || read_nucs[i] > kira::kMaxReadBases ||
            (read_nucs[i] + 3) / 4 > read_bytes_max || ref_nucs[i] < 0 || ref_offsets[i] < 0 ||
            ref_offsets[i] + (ref_nucs[i] + 3) / 4 > ref_bytes_per_phase)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "job outside its buffers");
        max_len = max_len > read_nucs[i] ? max_len : read_nucs[i];
    }
    // Layout of the staging area: each section 256-byte aligned.
    const int64_t phase_bases = ref_bytes_per_phase * 4;
    const int64_t win_words = (phase_bases + 31) / 32 + kira::kPadFront + kira::kPadBack;
    const size_t reads_bytes = (size_t)n * (size_t)read_bytes_max;
    size_t off = 0;
    const size_t o_reads = off;
    off += up256(reads_bytes);
    const size_t o_nucs = off;
    off += up256((size_t)n * 4);
    const size_t o_ws = off;
    off += up256((size_t)n * 8);
    const size_t o_wl = off;
    off += up256((size_t)n * 4);
    const size_t o_lim = off;
    off += up256((size_t)n * 4);
    const size_t o_out = off;
    off += up256((size_t)n * 12);
    const size_t o_rp = off;
    off += up256((size_t)n * kira::kReadWords * 4);
    const size_t o_phase = off;
    off += up256((size_t)(ref_bytes_per_phase > 0 ? ref_bytes_per_phase : 1));
    const size_t o_wp = off;
    off += up256((size_t)win_words * 8);
    if (off > (size_t)KIRA_NOALLOC_SCRATCH_BYTES) {
        char buf[192];
        snprintf(buf, sizeof buf,
                 "batch needs %zu bytes of staging, more than the %lld bytes this entry point "
                 "has (use kira_sieve_run_batch or split the batch)",
                 off, (long long)KIRA_NOALLOC_SCRATCH_BYTES);
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, buf);
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    void *base = nullptr;
    if ((e = kira::noalloc_setup(ctx->device, &base)) != hipSuccess || !base)
        return fail(ctx, KIRA_ERR_DRIVER, "staging array address", e);
    uint8_t *sc = static_cast<uint8_t *>(base);
    uint8_t *d_reads = sc + o_reads;
    int32_t *d_nucs = (int32_t *)(sc + o_nucs);
    int64_t *d_ws = (int64_t *)(sc + o_ws);
    int32_t *d_wl = (int32_t *)(sc + o_wl);
    int32_t *d_lim = (int32_t *)(sc + o_lim);
    int32_t *d_out = (int32_t *)(sc + o_out);
    uint32_t *d_rp = (uint32_t *)(sc + o_rp);
    uint8_t *d_phase0 = sc + o_phase;
    kira::Planes wp;
    wp.n_bases = phase_bases;
    wp.n_words = win_words;
    wp.lh = (uint2 *)(sc + o_wp);

    hipStream_t s = ctx->stream;
    // Every staging word a kernel reads is written first by a copy or a
    // kernel of this call (the plane packers write every word, padding
    // included), so nothing needs clearing.
    if ((e = hipMemcpyAsync(d_reads, reads_packed, reads_bytes, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_nucs, read_nucs, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (ref_bytes_per_phase > 0 &&
         (e = hipMemcpyAsync(d_phase0, ref_shifted_flat, (size_t)ref_bytes_per_phase,
                             hipMemcpyHostToDevice, s))) ||
        (e = hipMemcpyAsync(d_ws, ref_offsets, (size_t)n * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_wl, ref_nucs, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_lim, max_mismatches, (size_t)n * 4, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    noalloc_prep_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, d_ws, d_lim, d_nucs);
    if ((e = hipGetLastError()) ||
        (e = kira::launch_pack_reads_2bit(n, d_reads, read_bytes_max, d_nucs, d_rp, s)) ||
        (e = kira::launch_pack_2bit_planes(d_phase0, ref_bytes_per_phase, wp, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "pack launch", e);
    kira::SieveLaunch a{};
    a.n_jobs = n;
    a.read_planes = d_rp;
    a.read_lens = d_nucs;
    a.jobs_per_read = 1;
    a.win_start = d_ws;
    a.win_len = d_wl;
    a.limit = d_lim;
    a.out_shift = d_out;
    a.out_best = d_out + n;
    a.out_second = d_out + 2 * (int64_t)n;
    a.cert_max_mism = -1;
    a.g_lh = wp.lh;
    a.g_bases = phase_bases;
    const int rw = max_len > 0 ? (max_len + 31) / 32 : 1;
    if ((e = kira::launch_sieve(a, rw, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "sieve launch", e);
    if ((e = hipMemcpyAsync(out_shifts, d_out, (size_t)n * 4, hipMemcpyDeviceToHost, s)) ||
        (e = hipMemcpyAsync(out_mismatches, d_out + n, (size_t)n * 8, hipMemcpyDeviceToHost, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
    if ((e = hipStreamSynchronize(s))) return fail(ctx, KIRA_ERR_RUNTIME, "synchronize", e);
    return KIRA_OK;
}

}  // extern "C"
