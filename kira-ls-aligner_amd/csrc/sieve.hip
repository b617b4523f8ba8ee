// Spectral Sieve on gfx950: exhaustive multi-shift 2-bit Hamming with best and
// second-best shift (src/alignment/bitpacked.rs:154-199, src/cuda/spectral.cu:136-213)
// plus the ungapped certification of the best hit (src/alignment/mod.rs:688-762).
//
// Layout: reads and window text are bit planes (32 bases per uint32, one plane
// per code bit), so one 32-base compare is xor, xor, or, popcount.  One lane
// evaluates one (read, window) job.  Shifts are swept in tiles of 128: the
// tile's window words are aligned once, then for each in-word offset b the
// window is funnel-shifted once and compared at the four word offsets a of
// the tile (shift = 32a + b), so the shift costs amortise over four compares.
// Best/second are kept as packed ranks (mismatches << 16 | shift-in-tile) so
// ties resolve to the lowest shift exactly as the reference's strict-< scan.
#include "internal.hpp"

namespace kira {
namespace {

__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
    return __builtin_amdgcn_alignbit(hi, lo, s);
}

// t | (a ^ b) in one v_bitop3_b32 (truth table over S0=0xF0, S1=0xCC, S2=0xAA).
__device__ __forceinline__ uint32_t or_xor(uint32_t t, uint32_t a, uint32_t b) {
    return __builtin_amdgcn_bitop3_b32(t, a, b, 0xF6);
}

// Median of three (v_med3_u32).
// popcount(x) + acc as one v_bcnt_u32_b32 (its second operand is the
// addend), keeping a word-by-word accumulation chain instead of a tree of
// separate counts and three-way adds
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

__device__ __forceinline__ uint32_t med3u(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t lo = x < y ? x : y, hi = x > y ? x : y;
    uint32_t t = lo > z ? lo : z;
    return t < hi ? t : hi;
}

// Base code of an ASCII byte and its flags (bitpacked.rs:12-20, 88-91).
__device__ __forceinline__ void classify(uint8_t c, uint32_t &lo, uint32_t &hi, uint32_t &inv,
                                         uint32_t &lc) {
    uint8_t u = c & 0xDF;  // ASCII upper
    bool acgt = (u == 'A' || u == 'C' || u == 'G' || u == 'T');
    bool letter = acgt && (c >= 'a');
    uint32_t code = (u == 'C') ? 1u : (u == 'G') ? 2u : (u == 'T') ? 3u : 0u;
    if (!acgt) code = 0;
    lo = code & 1u;
    hi = code >> 1;
    inv = acgt ? 0u : 1u;
    lc = letter ? 1u : 0u;
}

// Bit 7 of each byte of x: set when the byte is not zero (no carry leaves a
// byte, so the test is exact per byte).
__device__ __forceinline__ uint32_t nz_bytes(uint32_t x) {
    return (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// Bits 0, 8, 16 and 24 of x to bits 0-3: the multiplier places bit 8 k at
// bit 21 + k and every other partial product outside bits 21-24.
__device__ __forceinline__ uint32_t gather4(uint32_t x) {
    return (((x & 0x01010101u) * 0x00204081u) >> 21) & 0xFu;
}

// One thread per 32-base word of an ASCII sequence; each wave also writes
// the 64 bits of the non-ACGT word bitmap its words cover.
__global__ void pack_ascii_planes_kernel(const uint8_t *__restrict__ seq, int64_t n,
                                         Planes p) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t base = (w - kPadFront) * 32;
    uint32_t lo = 0, hi = 0, inv = 0, lc = 0;
    if (w >= kPadFront && w < p.n_words) {
        for (int k = 0; k < 32; k++) {
            int64_t q = base + k;
            if (q >= n) break;
            uint32_t a, b, c, d;
            classify(seq[q], a, b, c, d);
            lo |= a << k;
            hi |= b << k;
            inv |= c << k;
            lc |= d << k;
        }
    }
    if (w < p.n_words) {
        p.lh[w] = make_uint2(lo, hi);
        p.il[w] = make_uint2(inv, lc);
    }
    // blocks are whole waves starting at multiples of 64 words
    const uint64_t bits = __ballot(inv != 0);
    const int64_t w0 = w & ~(int64_t)63;
    if ((threadIdx.x & 63) == 0 && w0 < p.n_words) {
        p.nz[w0 >> 5] = (uint32_t)bits;
        if (w0 + 32 < p.n_words) p.nz[(w0 >> 5) + 1] = (uint32_t)(bits >> 32);
    }
}

// Split the even/odd bits of a 2-bit-interleaved word (16 bases) into two
// 16-bit planes.
__device__ __forceinline__ uint32_t even_bits(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

__device__ __forceinline__ uint32_t load_u32_bytes(const uint8_t *p, int64_t avail) {
    uint32_t v = 0;
    for (int k = 0; k < 4; k++)
        if (k < avail) v |= (uint32_t)p[k] << (8 * k);
    return v;
}

// One thread per 32-base word of a 2-bit packed byte stream (phase 0 of the
// reference's pre-shifted windows).
__global__ void pack_2bit_planes_kernel(const uint8_t *__restrict__ packed, int64_t n_bytes,
                                        Planes p) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= p.n_words) return;
    uint32_t lo = 0, hi = 0;
    if (w >= kPadFront) {
        int64_t b = (w - kPadFront) * 8;
        if (b < n_bytes) {
            uint32_t x0 = load_u32_bytes(packed + b, n_bytes - b);
            uint32_t x1 = b + 4 < n_bytes ? load_u32_bytes(packed + b + 4, n_bytes - b - 4) : 0;
            lo = even_bits(x0) | (even_bits(x1) << 16);
            hi = even_bits(x0 >> 1) | (even_bits(x1 >> 1) << 16);
        }
    }
    p.lh[w] = make_uint2(lo, hi);
}

// One thread per (read, plane word).
__global__ void pack_reads_ascii_kernel(int32_t n, const uint8_t *__restrict__ reads,
                                        int32_t stride, const int64_t *__restrict__ offs,
                                        const int32_t *__restrict__ lens,
                                        uint32_t *__restrict__ planes,
                                        uint8_t *__restrict__ valid) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 8) return;
    int32_t r = (int32_t)(t >> 3), k = (int32_t)(t & 7);
    int32_t L = lens[r];
    // row-major reads: a length past the row width is not read past the row
    // (and the read is marked invalid)
    const bool over = !offs && L > stride;
    const int32_t Le = over ? stride : L;
    uint32_t lo = 0, hi = 0, inv = 0, lc = 0;
    const uint8_t *src = reads + (offs ? offs[r] : (int64_t)r * stride);
    for (int b = 0; b < 32; b++) {
        int32_t q = k * 32 + b;
        if (q >= Le) break;
        uint32_t a0, b0, c0, d0;
        classify(src[q], a0, b0, c0, d0);
        lo |= a0 << b;
        hi |= b0 << b;
        inv |= c0 << b;
        lc |= d0 << b;
    }
    uint32_t *dst = planes + (int64_t)r * kReadWords;
    dst[k] = lo;
    dst[8 + k] = hi;
    dst[16 + k] = lc;
    if (valid) {
        // 8 consecutive lanes hold one read; reduce their inv flags.
        uint32_t f = inv != 0 ? 1u : 0u;
        f |= __shfl_xor(f, 1, 8);
        f |= __shfl_xor(f, 2, 8);
        f |= __shfl_xor(f, 4, 8);
        if (k == 0) valid[r] = (L <= kMaxReadBases && !over && f == 0) ? 1 : 0;
    }
}

// classify() on a 32-byte chunk, four bytes at once: with u = c & 0xDF
// (upper case), A C G T are 0x41 0x43 0x47 0x54, so code bit 1 is bit 2 of
// the byte and code bit 0 is bit 1 ^ bit 2; bit 5 marks lower case.  Bases at
// or past nb are cleared.
struct ChunkPlanes {
    uint32_t lo, hi, inv, lc;
};
__device__ __forceinline__ ChunkPlanes classify_chunk(uint4 v0, uint4 v1, int32_t nb) {
    ChunkPlanes o{0, 0, 0, 0};
    const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint32_t c = w[q], u = c & 0xDFDFDFDFu;
        const uint32_t z = nz_bytes(u ^ 0x41414141u) & nz_bytes(u ^ 0x43434343u) &
                           nz_bytes(u ^ 0x47474747u) & nz_bytes(u ^ 0x54545454u);
        const uint32_t ok = (~z & 0x80808080u) >> 7;  // bit 0 of each ACGT byte
        o.lo |= gather4(((c >> 1) ^ (c >> 2)) & ok) << (4 * q);
        o.hi |= gather4((c >> 2) & ok) << (4 * q);
        o.inv |= gather4(~ok & 0x01010101u) << (4 * q);
        o.lc |= gather4((c >> 5) & ok) << (4 * q);
    }
    const uint32_t m = nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);  // bases past L
    o.lo &= m;
    o.hi &= m;
    o.inv &= m;
    o.lc &= m;
    return o;
}

// Same as pack_reads_ascii_kernel for row-major reads with a stride that is a
// positive multiple of 32 and a 16-byte aligned base.  Thread (g, k) packs
// word k of reads g and g + h (h = ceil(n / 2)), fetching each 32-byte chunk
// with two 16-byte loads.  A chunk that starts below the stride lies inside
// the row (both are multiples of 32), so all loads of both reads, lengths
// included, are in flight together (a chunk at or past the stride loads the
// row's first chunk and is dropped by the mask), and the lengths are applied
// afterwards as bit masks.
__global__ void pack_reads_ascii_vec_kernel(int32_t n, const uint8_t *__restrict__ reads,
                                            int32_t stride, const int32_t *__restrict__ lens,
                                            uint32_t *__restrict__ planes,
                                            uint8_t *__restrict__ valid) {
    const int32_t h = (int32_t)(((int64_t)n + 1) / 2);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)h * 8) return;
    const int32_t g = (int32_t)(t >> 3), k = (int32_t)(t & 7);
    const bool two = g + h < n;
    const int32_t r[2] = {g, two ? g + h : g};
    const bool in_row = k * 32 < stride;
    int32_t L[2];
    uint4 v[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
        L[x] = lens[r[x]];
        const uint4 *src =
            (const uint4 *)(reads + (int64_t)r[x] * stride + (in_row ? k * 32 : 0));
        v[x][0] = src[0];
        v[x][1] = src[1];
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
        // a length past the row width is not read past the row (the read is
        // marked invalid)
        const bool over = L[x] > stride;
        const int32_t Le = over ? stride : L[x];
        const int32_t nb = in_row ? min(max(Le - k * 32, 0), 32) : 0;
        const ChunkPlanes o = classify_chunk(v[x][0], v[x][1], nb);
        if (x == 1 && !two) break;
        uint32_t *dst = planes + (int64_t)r[x] * kReadWords;
        dst[k] = o.lo;
        dst[8 + k] = o.hi;
        dst[16 + k] = o.lc;
        if (valid) {
            // 8 consecutive lanes hold one read; reduce their inv flags.
            uint32_t f = o.inv != 0 ? 1u : 0u;
            f |= __shfl_xor(f, 1, 8);
            f |= __shfl_xor(f, 2, 8);
            f |= __shfl_xor(f, 4, 8);
            if (k == 0) valid[r[x]] = (L[x] <= kMaxReadBases && !over && f == 0) ? 1 : 0;
        }
    }
}

// Reads already 2-bit packed (reference layout): one thread per (read, word).
__global__ void pack_reads_2bit_kernel(int32_t n, const uint8_t *__restrict__ packed,
                                       int32_t stride, const int32_t *__restrict__ lens,
                                       uint32_t *__restrict__ planes) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 8) return;
    int32_t r = (int32_t)(t >> 3), k = (int32_t)(t & 7);
    int32_t L = lens[r];
    int32_t nbytes = (L + 3) / 4;
    const uint8_t *src = packed + (int64_t)r * stride;
    int64_t b = (int64_t)k * 8;
    uint32_t lo = 0, hi = 0;
    if (b < nbytes) {
        uint32_t x0 = load_u32_bytes(src + b, nbytes - b);
        uint32_t x1 = b + 4 < nbytes ? load_u32_bytes(src + b + 4, nbytes - b - 4) : 0;
        lo = even_bits(x0) | (even_bits(x1) << 16);
        hi = even_bits(x0 >> 1) | (even_bits(x1 >> 1) << 16);
    }
    // bases at or past L are padding: keep them zero
    int32_t keep = L - k * 32;
    uint32_t m = keep >= 32 ? 0xFFFFFFFFu : (keep <= 0 ? 0u : ((1u << keep) - 1u));
    uint32_t *dst = planes + (int64_t)r * kReadWords;
    dst[k] = lo & m;
    dst[8 + k] = hi & m;
    dst[16 + k] = 0;
}

__device__ __forceinline__ uint32_t range_mask(int k, int lo, int hi) {
    // bits [lo, hi) of the 32-bit word k of a bit vector
    int a = lo - 32 * k, b = hi - 32 * k;
    a = a < 0 ? 0 : (a > 32 ? 32 : a);
    b = b < 0 ? 0 : (b > 32 ? 32 : b);
    if (b <= a) return 0u;
    uint32_t hm = b == 32 ? 0xFFFFFFFFu : ((1u << b) - 1u);
    uint32_t lm = a == 32 ? 0xFFFFFFFFu : ((1u << a) - 1u);
    return hm & ~lm;
}

// Window text words starting at absolute base position p (p >= -64): the
// low / high code planes and, with flags, the lower-case plane.
template <int NW>
__device__ __forceinline__ void window_words(const SieveLaunch &a, int64_t p, uint32_t (&lo)[NW],
                                             uint32_t (&hi)[NW], uint32_t (&lc)[NW]) {
    int64_t q = p + (int64_t)kPadFront * 32;
    int64_t w0 = q >> 5;
    uint32_t s = (uint32_t)(q & 31);
    uint2 prev = a.g_lh[w0];
    uint32_t pc = a.g_il ? a.g_il[w0].y : 0u;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        uint2 nx = a.g_lh[w0 + k + 1];
        uint32_t nc = a.g_il ? a.g_il[w0 + k + 1].y : 0u;
        lo[k] = funnel(nx.x, prev.x, s);
        hi[k] = funnel(nx.y, prev.y, s);
        lc[k] = funnel(nc, pc, s);
        prev = nx;
        pc = nc;
    }
}

// Byte mismatches between the read and the window at relative offset d from
// the best shift, restricted to read positions [lo, hi).
template <int RW>
__device__ __forceinline__ int masked_mismatch(const uint32_t (&rl)[RW], const uint32_t (&rh)[RW],
                                               const uint32_t (&rc)[RW], const SieveLaunch &a,
                                               int64_t pos, int lo, int hi) {
    uint32_t wl[RW], wh[RW], wc[RW];
    window_words<RW>(a, pos, wl, wh, wc);
    int m = 0;
#pragma unroll
    for (int k = 0; k < RW; k++) {
        uint32_t d = (rl[k] ^ wl[k]) | (rh[k] ^ wh[k]) | (rc[k] ^ wc[k]);
        m += __popc(d & range_mask(k, lo, hi));
    }
    return m;
}

// spectral_hit_is_certified (src/alignment/mod.rs:688-762) on bit planes.
// Bytes are A/C/G/T in either case here (the job was dispatched), so byte
// inequality = code inequality or case inequality.
template <int RW>
__device__ bool certify(const uint32_t (&rl)[RW], const uint32_t (&rh)[RW],
                        const uint32_t (&rc)[RW], const SieveLaunch &a, int64_t s0, int L,
                        int shift, uint32_t best, uint32_t second) {
    if ((int)best > a.cert_max_mism) return false;
    if (second != 0xFFFFFFFFu && second <= best + 1) return false;
    if (best == 0) return true;
    if (best > 2) return false;
    int64_t at = s0 + shift;
    uint32_t wl[RW], wh[RW], wc[RW];
    window_words<RW>(a, at, wl, wh, wc);
    int n = 0, p0 = -1, p1 = -1;
#pragma unroll
    for (int k = 0; k < RW; k++) {
        uint32_t d = (rl[k] ^ wl[k]) | (rh[k] ^ wh[k]) | (rc[k] ^ wc[k]);
        d &= range_mask(k, 0, L);
        while (d) {
            int b = __builtin_ctz(d);
            if (n == 0) p0 = 32 * k + b;
            else if (n == 1) p1 = 32 * k + b;
            n++;
            d &= d - 1;
        }
    }
    if (n != (int)best) return false;
    int edge = L < 8 ? L : 8;
    if (p0 < edge || p0 + edge >= L) return false;
    if (n == 2) {
        if (p1 < edge || p1 + edge >= L) return false;
        if (p1 - p0 <= 4) return false;
    }
    int p = p0;
    int suffix = L - (p + 1);
    int min_suffix = a.cert_strict ? 4 : 8;
    if (suffix >= min_suffix) {
        // read[i] vs ref[i-1] for i in [p+1, L); read[i] vs ref[i+1] for i in [p, L-1)
        if (masked_mismatch<RW>(rl, rh, rc, a, at - 1, p + 1, L) <= 1) return false;
        if (masked_mismatch<RW>(rl, rh, rc, a, at + 1, p, L - 1) <= 1) return false;
        if (a.cert_strict && suffix >= 6) {
            if (masked_mismatch<RW>(rl, rh, rc, a, at - 2, p + 2, L) <= 1) return false;
            if (masked_mismatch<RW>(rl, rh, rc, a, at + 2, p, L - 2) <= 1) return false;
        }
    }
    return true;
}

// Scan one tile of up to 128 shifts; wl / wh: the tile's window words
// (RW + 4 of them, aligned to its first shift).  FULL: all 128 valid.
// Shift aa * 32 + b of the tile is ranked (count << 16) | b against its own
// group aa's best / second (tb[aa], ts[aa]); the caller adds aa * 32 back.
template <int RW, bool FULL, bool MASKALL>
__device__ __forceinline__ void scan_tile(const uint32_t (&rl)[RW], const uint32_t (&rh)[RW],
                                          const uint32_t (&mk)[RW], const uint32_t (&wl)[RW + 4],
                                          const uint32_t (&wh)[RW + 4], uint32_t rem,
                                          uint32_t (&tb)[4], uint32_t (&ts)[4]) {
    auto step = [&](uint32_t b) {
        uint32_t xl[RW + 3], xh[RW + 3];
#pragma unroll
        for (int k = 0; k < RW + 3; k++) {
            xl[k] = funnel(wl[k + 1], wl[k], b);
            xh[k] = funnel(wh[k + 1], wh[k], b);
        }
#pragma unroll
        for (int aa = 0; aa < 4; aa++) {
            // a partial tile stops at its last shift (129 shifts = 128 + 1)
            if (!FULL && (uint32_t)aa * 32u >= rem) break;
            uint32_t m = 0;
#pragma unroll
            for (int k = 0; k < RW; k++) {
                // (rl ^ xl) | (rh ^ xh) as xor + one 3-input bit op
                uint32_t d = or_xor(rl[k] ^ xl[aa + k], rh[k], xh[aa + k]);
                if (MASKALL || k == RW - 1) d &= mk[k];
                m = k == 0 ? __popc(d) : bcnt_acc(d, m);
            }
            uint32_t rank = (m << 16) | b;
            if (!FULL && (uint32_t)aa * 32u + b >= rem) rank = 0xFFFFFFFFu;
            // tb <= ts always, so min(ts, max(tb, rank)) is the median of the three
            ts[aa] = med3u(tb[aa], rank, ts[aa]);
            tb[aa] = tb[aa] < rank ? tb[aa] : rank;
        }
    };
    if constexpr (FULL) {
        // two shifts per trip: the loop-carried best / second stay in place
#pragma unroll 2
        for (uint32_t b = 0; b < 32u; b++) step(b);
    } else {
        const uint32_t nb = rem < 32u ? rem : 32u;
        for (uint32_t b = 0; b < nb; b++) step(b);
    }
}

template <int RW, bool MASKALL>
__device__ __forceinline__ void sieve_core(const SieveLaunch &a, const uint32_t (&rl)[RW],
                                           const uint32_t (&rh)[RW], const uint32_t (&mk)[RW],
                                           int64_t s0, uint32_t nt, uint32_t &gbm,
                                           uint32_t &gbt, uint32_t &gsm) {
    gbm = 0xFFFFFFFFu;
    gbt = 0;
    gsm = 0xFFFFFFFFu;
    // Each window word is loaded once per job: a tile of 128 shifts needs the
    // NW + 1 raw words from its start, and the next tile starts 4 words on,
    // so the raw words slide by 4 and only 4 new ones are loaded.
    constexpr int NW = RW + 4;
    const int64_t q = s0 + (int64_t)kPadFront * 32;
    const int64_t w0 = q >> 5;
    const uint32_t sh = (uint32_t)(q & 31);
    uint2 raw[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; k++) raw[k] = a.g_lh[w0 + k];
    for (uint32_t T0 = 0; T0 < nt; T0 += 128) {
        if (T0 > 0) {
#pragma unroll
            for (int k = 0; k + 4 <= NW; k++) raw[k] = raw[k + 4];
#pragma unroll
            for (int k = NW - 3; k <= NW; k++) raw[k] = a.g_lh[w0 + (T0 >> 5) + k];
        }
        uint32_t wl[NW], wh[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) {
            wl[k] = funnel(raw[k + 1].x, raw[k].x, sh);
            wh[k] = funnel(raw[k + 1].y, raw[k].y, sh);
        }
        uint32_t gb[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        uint32_t gs[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        uint32_t rem = nt - T0;
        if (rem >= 128)
            scan_tile<RW, true, MASKALL>(rl, rh, mk, wl, wh, rem, gb, gs);
        else
            scan_tile<RW, false, MASKALL>(rl, rh, mk, wl, wh, rem, gb, gs);
        // merge the four groups: ranks become (count << 16) | tile shift, and
        // each group's (best, second) are its two smallest, so the union's are
        // min of the bests and min(max of the bests, the seconds)
        uint32_t tb = 0xFFFFFFFFu, ts = 0xFFFFFFFFu;
#pragma unroll
        for (int aa = 0; aa < 4; aa++) {
            const uint32_t b = gb[aa] == 0xFFFFFFFFu ? gb[aa] : gb[aa] + 32u * aa;
            const uint32_t c = gs[aa] == 0xFFFFFFFFu ? gs[aa] : gs[aa] + 32u * aa;
            const uint32_t hi = tb > b ? tb : b;
            ts = ts < c ? ts : c;
            ts = ts < hi ? ts : hi;
            tb = tb < b ? tb : b;
        }
        // a full tile followed by exactly one more shift (nt = 128 t + 1, as
        // every window of L + 128 bases gives): that shift is word offset 4,
        // bit 0 of this tile's words, ranked here (tile shift 128) instead of
        // in a tile of its own
        const bool tail = rem == 129;
        if (tail) {
            uint32_t m = 0;
#pragma unroll
            for (int k = 0; k < RW; k++) {
                uint32_t d = or_xor(rl[k] ^ wl[4 + k], rh[k], wh[4 + k]);
                if (MASKALL || k == RW - 1) d &= mk[k];
                m = k == 0 ? __popc(d) : bcnt_acc(d, m);
            }
            const uint32_t r = (m << 16) | 128u;
            const uint32_t hr = tb > r ? tb : r;
            ts = ts < hr ? ts : hr;
            tb = tb < r ? tb : r;
        }
        uint32_t tbm = tb >> 16;
        uint32_t tsm = ts == 0xFFFFFFFFu ? 0xFFFFFFFFu : (ts >> 16);
        uint32_t hi = gbm > tbm ? gbm : tbm;
        uint32_t ns = gsm < hi ? gsm : hi;
        ns = ns < tsm ? ns : tsm;
        if (tbm < gbm) {
            gbm = tbm;
            gbt = T0 + (tb & 0xFFFFu);
        }
        gsm = ns;
        if (tail) break;
    }
}

// Accepted jobs of a workgroup queued for certification (indices into the
// workgroup's 256 jobs' results), so that certification runs once per
// workgroup on a compact set of lanes instead of in every wave that holds an
// accepted job (about one job in sixteen in a read-major batch).
struct CertQueue {
    int32_t job[256], shift[256];
    uint32_t best[256], second[256];
    int32_t n;
};

template <int RW>
__global__ __launch_bounds__(256) void sieve_kernel(SieveLaunch a) {
    __shared__ CertQueue cq;
    if (threadIdx.x == 0) cq.n = 0;
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // every thread reaches the queue barrier below; `done` marks a job whose
    // results are written
    bool done = j >= a.n_jobs;
    int32_t r = 0, L = 0, W = 0, lim = 0;
    int64_t s0 = 0;
    if (!done) {
        r = a.job_read ? a.job_read[j] : (int32_t)(j / a.jobs_per_read);
        L = a.read_lens[r];
        W = a.win_len[j];
        s0 = a.win_start[j];
        lim = a.limit ? a.limit[j] : a.limit_all;
        // Skip (never dispatch, stage4_alignment.rs:1284-1288) jobs whose read
        // or window is not all ACGT, and windows that leave the text.
        bool ok = (a.read_valid ? a.read_valid[r] != 0 : true) && W >= 0 && s0 >= 0 &&
                  s0 + W <= a.g_bases && L <= 32 * RW;
        if (ok && a.g_il && W > 0) {
            const int64_t q0 = s0 + (int64_t)kPadFront * 32, q1 = q0 + W;
            const int64_t w0 = q0 >> 5, w1 = (q1 - 1) >> 5;
            // the word bitmap rules out almost every window at once: its bits
            // w0 .. w1 lie in two consecutive bitmap words when w1 - w0 < 32
            // (the bitmap has a spare word at its end)
            bool maybe = true;
            if (w1 - w0 < 32) {
                const uint64_t bits =
                    (((uint64_t)a.g_nz[(w0 >> 5) + 1] << 32) | a.g_nz[w0 >> 5]) >> (w0 & 31);
                maybe = (bits & ((2ull << (w1 - w0)) - 1)) != 0;
            }
            if (maybe) {
                uint32_t any = 0;
                for (int64_t w = w0; w <= w1; w++) {
                    if (!((a.g_nz[w >> 5] >> (w & 31)) & 1u)) continue;
                    uint32_t v = a.g_il[w].x;
                    int lo = (int)(w == w0 ? (q0 & 31) : 0);
                    int hi = (int)(w == w1 ? (((q1 - 1) & 31) + 1) : 32);
                    uint32_t m = (hi == 32 ? 0xFFFFFFFFu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
                    any |= v & m;
                }
                ok = any == 0;
            }
        }
        if (!ok || L == 0 || W < L) {
            a.out_shift[j] = -1;
            a.out_best[j] = INT32_MAX;
            a.out_second[j] = INT32_MAX;
            if (a.out_flags) a.out_flags[j] = ok ? 1 : 0;
            done = true;
        }
    }
    if (!done) {
        const uint32_t *rp = a.read_planes + (int64_t)r * kReadWords;
        uint32_t rl[RW], rh[RW], mk[RW];
#pragma unroll
        for (int k = 0; k < RW; k++) {
            rl[k] = rp[k];
            rh[k] = rp[8 + k];
            int keep = L - 32 * k;
            mk[k] = keep >= 32 ? 0xFFFFFFFFu : (keep <= 0 ? 0u : ((1u << keep) - 1u));
        }
        uint32_t nt = (uint32_t)(W - L + 1);
        uint32_t gbm, gbt, gsm;
        if (L > 32 * (RW - 1))
            sieve_core<RW, false>(a, rl, rh, mk, s0, nt, gbm, gbt, gsm);
        else
            sieve_core<RW, true>(a, rl, rh, mk, s0, nt, gbm, gbt, gsm);

        bool acc = (int64_t)gbm <= (int64_t)lim;
        a.out_shift[j] = acc ? (int32_t)gbt : -1;
        a.out_best[j] = acc ? (int32_t)gbm : INT32_MAX;
        a.out_second[j] = gsm == 0xFFFFFFFFu ? INT32_MAX : (int32_t)gsm;
        if (a.out_flags) {
            if (acc && a.cert_max_mism >= 0) {
                const int32_t at = atomicAdd(&cq.n, 1);
                cq.job[at] = (int32_t)threadIdx.x;
                cq.shift[at] = (int32_t)gbt;
                cq.best[at] = gbm;
                cq.second[at] = gsm;
            } else {
                a.out_flags[j] = acc ? 3 : 1;
            }
        }
    }
    __syncthreads();
    // certification of the queued jobs, by the first threads of the workgroup
    const int32_t nq = cq.n;
    for (int32_t q = (int32_t)threadIdx.x; q < nq; q += (int32_t)blockDim.x) {
        const int64_t jq = (int64_t)blockIdx.x * blockDim.x + cq.job[q];
        const int32_t rq = a.job_read ? a.job_read[jq] : (int32_t)(jq / a.jobs_per_read);
        const uint32_t *rp = a.read_planes + (int64_t)rq * kReadWords;
        uint32_t rl[RW], rh[RW], rc[RW];
#pragma unroll
        for (int k = 0; k < RW; k++) {
            rl[k] = rp[k];
            rh[k] = rp[8 + k];
            rc[k] = rp[16 + k];
        }
        const bool cert = certify<RW>(rl, rh, rc, a, a.win_start[jq], a.read_lens[rq],
                                      cq.shift[q], cq.best[q], cq.second[q]);
        a.out_flags[jq] = cert ? 7 : 3;
    }
}

}  // namespace

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_pack_ascii_planes(const uint8_t *seq, int64_t n, Planes p, hipStream_t s) {
    if (p.n_words == 0) return hipSuccess;
    pack_ascii_planes_kernel<<<grid_for(p.n_words, 256), 256, 0, s>>>(seq, n, p);
    return hipGetLastError();
}

hipError_t launch_pack_2bit_planes(const uint8_t *packed, int64_t n_bytes, Planes p,
                                   hipStream_t s) {
    if (p.n_words == 0) return hipSuccess;
    pack_2bit_planes_kernel<<<grid_for(p.n_words, 256), 256, 0, s>>>(packed, n_bytes, p);
    return hipGetLastError();
}

hipError_t launch_pack_reads_ascii(int32_t n, const uint8_t *reads, int32_t stride,
                                   const int64_t *offs, const int32_t *lens, uint32_t *planes,
                                   uint8_t *valid, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!offs && stride > 0 && stride % 32 == 0 && ((uintptr_t)reads & 15) == 0)
        pack_reads_ascii_vec_kernel<<<grid_for(((int64_t)n + 1) / 2 * 8, 256), 256, 0, s>>>(
            n, reads, stride, lens, planes, valid);
    else
        pack_reads_ascii_kernel<<<grid_for((int64_t)n * 8, 256), 256, 0, s>>>(
            n, reads, stride, offs, lens, planes, valid);
    return hipGetLastError();
}

hipError_t launch_pack_reads_2bit(int32_t n, const uint8_t *packed, int32_t stride,
                                  const int32_t *lens, uint32_t *planes, hipStream_t s) {
    if (n == 0) return hipSuccess;
    pack_reads_2bit_kernel<<<grid_for((int64_t)n * 8, 256), 256, 0, s>>>(n, packed, stride, lens,
                                                                         planes);
    return hipGetLastError();
}

hipError_t launch_sieve(const SieveLaunch &a, int read_words, hipStream_t s) {
    if (a.n_jobs == 0) return hipSuccess;
    unsigned g = grid_for(a.n_jobs, 256);
    switch (read_words) {
    case 1: sieve_kernel<1><<<g, 256, 0, s>>>(a); break;
    case 2: sieve_kernel<2><<<g, 256, 0, s>>>(a); break;
    case 3: sieve_kernel<3><<<g, 256, 0, s>>>(a); break;
    case 4: sieve_kernel<4><<<g, 256, 0, s>>>(a); break;
    case 5: sieve_kernel<5><<<g, 256, 0, s>>>(a); break;
    case 6: sieve_kernel<6><<<g, 256, 0, s>>>(a); break;
    case 7: sieve_kernel<7><<<g, 256, 0, s>>>(a); break;
    case 8: sieve_kernel<8><<<g, 256, 0, s>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace kira
