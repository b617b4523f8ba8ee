// Host runtime behind include/kira_hip.h: context, device buffers, the
// reference-layout batch entry (CudaBackend::run_batch, src/cuda/backend.rs:96-249)
// and the resident-genome entries.  No exception crosses the C ABI.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
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

// Geometric growth like ensure_capacity_* (src/cuda/backend.rs:272-308).
hipError_t ensure(kira_ctx *ctx, int slot, size_t need) {
    if (ctx->cap[slot] >= need) return hipSuccess;
    size_t nc = std::max(need, ctx->cap[slot] * 2);
    if (ctx->buf[slot]) {
        hipError_t e = hipFree(ctx->buf[slot]);
        ctx->buf[slot] = nullptr;
        ctx->cap[slot] = 0;
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&ctx->buf[slot], nc);
    if (e != hipSuccess) {
        ctx->buf[slot] = nullptr;
        return e;
    }
    ctx->cap[slot] = nc;
    return hipSuccess;
}

enum Slot {
    S_READS = 0,
    S_READ_NUCS,
    S_PHASE0,
    S_WIN_START,
    S_WIN_LEN,
    S_LIMITS,
    S_OUT,
    S_READ_PLANES,
    S_WIN_PLANES,
    S_COUNT
};

// Plane arrays carved from one block: lh, then (with flags) il and the
// word bitmap.  planes_bytes gives the block size.
kira::Planes planes_in(void *base, int64_t n_bases, bool with_flags) {
    kira::Planes p;
    p.n_bases = n_bases;
    p.n_words = (n_bases + 31) / 32 + kira::kPadFront + kira::kPadBack;
    uint8_t *b = static_cast<uint8_t *>(base);
    p.lh = reinterpret_cast<uint2 *>(b);
    if (with_flags) {
        p.il = reinterpret_cast<uint2 *>(b + p.n_words * 8);
        p.nz = reinterpret_cast<uint32_t *>(b + p.n_words * 16);
    }
    return p;
}

size_t planes_bytes(const kira::Planes &p, bool with_flags) {
    return (size_t)p.n_words * 8 + (with_flags ? (size_t)p.n_words * 8 + ((size_t)p.n_words / 32 + 2) * 4 : 0);
}

}  // namespace

extern "C" {

int kira_is_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

kira_status kira_ctx_create(int device, kira_ctx **out) {
    if (!out) return KIRA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return KIRA_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return KIRA_ERR_NO_DEVICE;
    kira_ctx *ctx = new (std::nothrow) kira_ctx();
    if (!ctx) return KIRA_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return KIRA_ERR_DRIVER;
    }
    ctx->stream = ctx->own_stream;
    // without its streams a plan launch keeps every class on one stream
    (void)kira::sw_fanout_create(ctx->fan);
    // without a side stream every launch stays on the context stream
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming) != hipSuccess) {
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        ctx->side = nullptr;
        ctx->side_fork = nullptr;
    }
    // the no-allocation entry point's staging array: its device address is
    // looked up here, once per device (a failure is reported by that entry
    // point, which needs the address)
    (void)kira::noalloc_setup(device, nullptr);
    *out = ctx;
    return KIRA_OK;
}

void kira_release_buffers(kira_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (int k = 0; k < 16; k++) {
        if (ctx->buf[k]) (void)hipFree(ctx->buf[k]);
        ctx->buf[k] = nullptr;
        ctx->cap[k] = 0;
    }
    for (auto &kv : ctx->free_blocks)
        for (void *p : kv.second) (void)hipFree(p);
    ctx->free_blocks.clear();
    for (int k = 0; k < kira::kPinSlots; k++) {
        if (ctx->pin[k]) (void)hipHostFree(ctx->pin[k]);
        ctx->pin[k] = nullptr;
        ctx->pin_cap[k] = 0;
    }
}

void kira_ctx_destroy(kira_ctx *ctx) {
    if (!ctx) return;
    kira_release_buffers(ctx);
    kira::sw_fanout_destroy(ctx->fan);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char *kira_ctx_last_error(const kira_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

void *kira_ctx_stream(kira_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

void kira_ctx_set_stream(kira_ctx *ctx, void *stream) {
    if (!ctx) return;
    ctx->stream = (hipStream_t)stream;
}

void kira_ctx_use_own_stream(kira_ctx *ctx) {
    if (!ctx) return;
    ctx->stream = ctx->own_stream;
}

kira_status kira_ctx_sync(kira_ctx *ctx) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_RUNTIME, "stream synchronize", e);
    return KIRA_OK;
}

kira_status kira_sieve_run_batch(kira_ctx *ctx, int32_t n, const uint8_t *reads_packed,
                                 int32_t read_bytes_max, const int32_t *read_packed_len,
                                 const int32_t *read_nucs, const uint8_t *ref_shifted_flat,
                                 int64_t ref_bytes_per_phase, const int64_t *ref_offsets,
                                 const int32_t *ref_nucs, const uint32_t *max_mismatches,
                                 int32_t *out_shifts, int32_t *out_mismatches) {
    if (!ctx) return KIRA_ERR_INVALID_ARGUMENT;
    if (n == 0) return KIRA_OK;  // empty -> Ok(vec![]) (backend.rs:97-99)
    if (n < 0 || !reads_packed || !read_packed_len || !read_nucs || !ref_offsets || !ref_nucs ||
        !max_mismatches || !out_shifts || !out_mismatches || read_bytes_max <= 0 ||
        ref_bytes_per_phase < 0 || (ref_bytes_per_phase > 0 && !ref_shifted_flat))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    // ReadTooLong fails the whole batch (backend.rs:101-108).
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
        max_len = std::max(max_len, read_nucs[i]);
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);

    // Host-side job arrays: window start in bases and min(limit, read_nucs)
    // (backend.rs:185-188).
    std::vector<int64_t> win_start(n);
    std::vector<int32_t> limits(n);
    for (int32_t i = 0; i < n; i++) {
        win_start[i] = ref_offsets[i] * 4;
        uint32_t lim = std::min<uint32_t>(max_mismatches[i], (uint32_t)read_nucs[i]);
        limits[i] = (int32_t)std::min<uint32_t>(lim, (uint32_t)INT32_MAX);
    }
    int64_t phase_bases = ref_bytes_per_phase * 4;
    kira::Planes probe = planes_in(nullptr, phase_bases, false);
    size_t reads_bytes = (size_t)n * read_bytes_max;
    if ((e = ensure(ctx, S_READS, reads_bytes)) != hipSuccess ||
        (e = ensure(ctx, S_READ_NUCS, (size_t)n * 4)) != hipSuccess ||
        (e = ensure(ctx, S_PHASE0, (size_t)std::max<int64_t>(ref_bytes_per_phase, 1))) !=
            hipSuccess ||
        (e = ensure(ctx, S_WIN_START, (size_t)n * 8)) != hipSuccess ||
        (e = ensure(ctx, S_WIN_LEN, (size_t)n * 4)) != hipSuccess ||
        (e = ensure(ctx, S_LIMITS, (size_t)n * 4)) != hipSuccess ||
        (e = ensure(ctx, S_OUT, (size_t)n * 12)) != hipSuccess ||
        (e = ensure(ctx, S_READ_PLANES, (size_t)n * kira::kReadWords * 4)) != hipSuccess ||
        (e = ensure(ctx, S_WIN_PLANES, planes_bytes(probe, false))) != hipSuccess)
        return fail(ctx, KIRA_ERR_RUNTIME, "device allocation", e);
    hipStream_t s = ctx->stream;
    uint8_t *d_reads = (uint8_t *)ctx->buf[S_READS];
    int32_t *d_nucs = (int32_t *)ctx->buf[S_READ_NUCS];
    uint8_t *d_phase0 = (uint8_t *)ctx->buf[S_PHASE0];
    int64_t *d_ws = (int64_t *)ctx->buf[S_WIN_START];
    int32_t *d_wl = (int32_t *)ctx->buf[S_WIN_LEN];
    int32_t *d_lim = (int32_t *)ctx->buf[S_LIMITS];
    int32_t *d_out = (int32_t *)ctx->buf[S_OUT];
    uint32_t *d_rp = (uint32_t *)ctx->buf[S_READ_PLANES];
    kira::Planes wp = planes_in(ctx->buf[S_WIN_PLANES], phase_bases, false);

    if ((e = hipMemcpyAsync(d_reads, reads_packed, reads_bytes, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_nucs, read_nucs, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (ref_bytes_per_phase > 0 &&
         (e = hipMemcpyAsync(d_phase0, ref_shifted_flat, (size_t)ref_bytes_per_phase,
                             hipMemcpyHostToDevice, s))) ||
        (e = hipMemcpyAsync(d_ws, win_start.data(), (size_t)n * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_wl, ref_nucs, (size_t)n * 4, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(d_lim, limits.data(), (size_t)n * 4, hipMemcpyHostToDevice, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "host to device copy", e);
    if ((e = kira::launch_pack_reads_2bit(n, d_reads, read_bytes_max, d_nucs, d_rp, s)) ||
        (e = kira::launch_pack_2bit_planes(d_phase0, ref_bytes_per_phase, wp, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "pack launch", e);
    kira::SieveLaunch a{};
    a.n_jobs = n;
    a.read_planes = d_rp;
    a.read_lens = d_nucs;
    a.read_valid = nullptr;
    a.job_read = nullptr;
    a.jobs_per_read = 1;
    a.win_start = d_ws;
    a.win_len = d_wl;
    a.limit = d_lim;
    a.out_shift = d_out;
    a.out_best = d_out + n;
    a.out_second = d_out + 2 * (int64_t)n;
    a.out_flags = nullptr;
    a.cert_max_mism = -1;
    a.g_lh = wp.lh;
    a.g_il = nullptr;
    a.g_nz = nullptr;
    a.g_bases = phase_bases;
    int rw = std::max(1, (max_len + 31) / 32);
    if ((e = kira::launch_sieve(a, rw, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "sieve launch", e);
    if ((e = hipMemcpyAsync(out_shifts, d_out, (size_t)n * 4, hipMemcpyDeviceToHost, s)) ||
        (e = hipMemcpyAsync(out_mismatches, d_out + n, (size_t)n * 8, hipMemcpyDeviceToHost, s)))
        return fail(ctx, KIRA_ERR_RUNTIME, "device to host copy", e);
    if ((e = hipStreamSynchronize(s))) return fail(ctx, KIRA_ERR_RUNTIME, "synchronize", e);
    return KIRA_OK;
}

kira_status kira_genome_create(kira_ctx *ctx, const uint8_t *seq, int64_t len,
                               kira_genome **out) {
    if (!ctx || !out || len < 0 || (len > 0 && !seq))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    kira::Planes probe = planes_in(nullptr, len, true);
    void *mem = nullptr;
    if ((e = hipMalloc(&mem, planes_bytes(probe, true))))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "genome allocation", e);
    // the ASCII sequence stays resident too (SW and prefilter windows read
    // bytes), with kRefPad zero bytes on each side for band rows that poke
    // past the contig ends
    uint8_t *d_seq = nullptr;
    if ((e = hipMalloc(&d_seq, (size_t)len + 2 * kira::kRefPad))) {
        (void)hipFree(mem);
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "genome allocation", e);
    }
    kira::Planes p = planes_in(mem, len, true);
    if ((e = hipMemsetAsync(d_seq, 0, (size_t)len + 2 * kira::kRefPad, ctx->stream)) ||
        (len > 0 && (e = hipMemcpyAsync(d_seq + kira::kRefPad, seq, (size_t)len,
                                        hipMemcpyHostToDevice, ctx->stream))) ||
        (e = kira::launch_pack_ascii_planes(d_seq + kira::kRefPad, len, p, ctx->stream)) ||
        (e = hipStreamSynchronize(ctx->stream))) {
        (void)hipFree(mem);
        (void)hipFree(d_seq);
        return fail(ctx, KIRA_ERR_RUNTIME, "genome packing", e);
    }
    kira_genome *g = new (std::nothrow) kira_genome();
    if (g) {
        try {
            g->host.assign(seq, seq + len);
        } catch (const std::bad_alloc &) {
            delete g;
            g = nullptr;
        }
    }
    if (!g) {
        (void)hipFree(mem);
        (void)hipFree(d_seq);
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "genome host copy");
    }
    g->ctx = ctx;
    g->planes = p;
    g->ascii = d_seq + kira::kRefPad;
    *out = g;
    return KIRA_OK;
}

void kira_genome_destroy(kira_genome *g) {
    if (!g) return;
    if (g->planes.lh) (void)hipFree(g->planes.lh);
    if (g->ascii) (void)hipFree(g->ascii - kira::kRefPad);
    delete g;
}

const uint8_t *kira_genome_ascii_dev(const kira_genome *g) { return g ? g->ascii : nullptr; }

int64_t kira_genome_len(const kira_genome *g) { return g ? g->planes.n_bases : 0; }

kira_status kira_reads_pack_dev(kira_ctx *ctx, int32_t n, const uint8_t *reads, int32_t stride,
                                const int32_t *read_lens, uint32_t *planes_out,
                                uint8_t *valid_out) {
    if (!ctx || n < 0 || (n > 0 && (!reads || !read_lens || !planes_out)))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    hipError_t e = kira::launch_pack_reads_ascii(n, reads, stride, nullptr, read_lens,
                                                 planes_out, valid_out, ctx->stream);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_RUNTIME, "read pack launch", e);
    return KIRA_OK;
}

kira_status kira_sieve_genome_dev(kira_ctx *ctx, const kira_genome *g,
                                  const kira_sieve_batch *b) {
    if (!ctx || !g || !b) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (b->n_jobs == 0) return KIRA_OK;
    if (b->n_jobs < 0 || !b->read_planes || !b->read_lens || !b->win_start || !b->win_len ||
        !b->out_shift || !b->out_best || !b->out_second ||
        (!b->job_read && b->jobs_per_read <= 0))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid batch");
    kira::SieveLaunch a{};
    a.n_jobs = b->n_jobs;
    a.read_planes = b->read_planes;
    a.read_lens = b->read_lens;
    a.read_valid = b->read_valid;
    a.job_read = b->job_read;
    a.jobs_per_read = b->jobs_per_read;
    a.win_start = b->win_start;
    a.win_len = b->win_len;
    a.limit = b->limit;
    a.limit_all = b->limit_all;
    a.out_shift = b->out_shift;
    a.out_best = b->out_best;
    a.out_second = b->out_second;
    a.out_flags = b->out_flags;
    a.cert_max_mism = b->cert_max_mism;
    a.cert_strict = b->cert_strict;
    a.g_lh = g->planes.lh;
    a.g_il = g->planes.il;
    a.g_nz = g->planes.nz;
    a.g_bases = g->planes.n_bases;
    // The read-word count is a launch-wide template parameter; the caller
    // promises every read of the batch is at most this long (256 bases).
    int rw = b->max_read_len <= 0 ? 8 : (b->max_read_len + 31) / 32;
    if (rw > 8) return fail(ctx, KIRA_ERR_READ_TOO_LONG, "read longer than 256 bases");
    hipError_t e = kira::launch_sieve(a, rw, ctx->stream);
    if (e != hipSuccess) return fail(ctx, KIRA_ERR_RUNTIME, "sieve launch", e);
    return KIRA_OK;
}

}  // extern "C"

namespace kira {

hipError_t cached_alloc(kira_ctx *ctx, void **p, size_t bytes, size_t *size_class) {
    const size_t kMaxCached = (size_t)1 << 30;
    if (!ctx || bytes > kMaxCached) {
        *size_class = 0;
        return hipMalloc(p, bytes);
    }
    size_t pow2 = 256;
    while (pow2 < bytes) pow2 <<= 1;
    const size_t step = std::max<size_t>(pow2 / 4, 64);
    const size_t cls = std::max<size_t>((bytes + step - 1) / step * step, 256);
    *size_class = cls;
    auto it = ctx->free_blocks.find(cls);
    if (it != ctx->free_blocks.end() && !it->second.empty()) {
        *p = it->second.back();
        it->second.pop_back();
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, cls);
    if (e != hipSuccess) {
        // give the cached blocks back to the device and try once more
        for (auto &kv : ctx->free_blocks)
            for (void *q : kv.second) (void)hipFree(q);
        ctx->free_blocks.clear();
        e = hipMalloc(p, cls);
    }
    return e;
}

void cached_free(kira_ctx *ctx, void *p, size_t size_class) {
    if (!p) return;
    if (!ctx || size_class == 0) {
        (void)hipFree(p);
        return;
    }
    try {
        ctx->free_blocks[size_class].push_back(p);
    } catch (const std::bad_alloc &) {
        (void)hipFree(p);
    }
}

hipError_t pinned_buf(kira_ctx *ctx, int slot, size_t bytes, void **out) {
    bytes = std::max<size_t>(bytes, 1);
    if (bytes > ctx->pin_cap[slot]) {
        if (ctx->pin[slot]) {
            // copies from or to the old buffer may still be in flight
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return e;
            (void)hipHostFree(ctx->pin[slot]);
            ctx->pin[slot] = nullptr;
            ctx->pin_cap[slot] = 0;
        }
        const size_t cap = ((bytes + bytes / 4) + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, cap, 0);
        if (e != hipSuccess) return e;
        ctx->pin[slot] = p;
        ctx->pin_cap[slot] = cap;
    }
    *out = ctx->pin[slot];
    return hipSuccess;
}

}  // namespace kira
