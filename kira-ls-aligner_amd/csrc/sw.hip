// Banded affine-gap Smith-Waterman with the reference's adaptive band
// (banded_sw_with, src/alignment/mod.rs:1370-1650) on gfx950.
//
// One lane runs one job row by row (row i+1's band centre is the arg-max of
// row i, so rows are inherently sequential).  The band lives in registers in
// diagonal coordinates: position q of row i is column j = c_i - w + q.  With
// that indexing the up neighbour of (i, j) is position q+1 and the diagonal
// neighbour position q of the previous row, both static register indices;
// when the centre moves by other than one column per row the arrays are
// shifted by the difference first.  Cells outside the band or the window
// hold pads chosen so the recurrences reproduce the reference's "absent"
// values (i32::MIN/4) exactly.  Each cell writes a 4-bit trace code
// (direction | E-extend << 2 | F-extend << 3); a packet of 64 jobs shares a
// slab laid out [row][word][lane] so every trace store is one coalesced
// 256-byte write.  After its last row the lane walks its trace back into
// the CIGAR, overlapping that latency-bound walk with other waves' rows.
#include "internal.hpp"

namespace kira {
namespace {

constexpr int32_t kMin4 = INT32_MIN / 4;
constexpr int32_t kMin8 = INT32_MIN / 8;

// The per-job logic is __host__ __device__ so the same code can be built and
// bounds-checked on the CPU.
__host__ __device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);  // v_alignbit_b32
}

__host__ __device__ __forceinline__ int32_t sat_add(int32_t a, int32_t b) {
    int64_t s = (int64_t)a + b;
    return s > INT32_MAX ? INT32_MAX : (s < INT32_MIN ? INT32_MIN : (int32_t)s);
}

struct Row {
    int32_t rb;   // row best (0 if no positive cell)
    int32_t rbq;  // its position
};

// Band storage for the previous row.  PACK keeps H (low) and E (high) as
// saturating int16 halves of one register, halving the register footprint;
// the plan routes a job there only when its scores provably fit
// (H <= L * match < 32767), so no value is ever clamped.  Pads mark cells
// outside the band: with them the recurrences give the reference's
// "absent" results (e/f = i32::MIN/4, diagonal H = 0, no extend bits).
template <int C, bool PACK>
struct Band {
    int32_t H[PACK ? 1 : C], E[PACK ? 1 : C];
    uint32_t P[PACK ? C : 1];
    int32_t padH, padE;  // stored pads
    __host__ __device__ __forceinline__ int32_t h(int q) const {
        if constexpr (PACK) return (int32_t)(int16_t)(P[q] & 0xFFFFu);
        else return H[q];
    }
    __host__ __device__ __forceinline__ int32_t e(int q) const {
        if constexpr (PACK) return ((int32_t)P[q]) >> 16;
        else return E[q];
    }
    __host__ __device__ __forceinline__ void set(int q, int32_t hv, int32_t ev) {
        if constexpr (PACK) P[q] = pack16(hv, ev);
        else {
            H[q] = hv;
            E[q] = ev;
        }
    }
    __host__ __device__ __forceinline__ void move(int dst, int src) {
        if constexpr (PACK) P[dst] = P[src];
        else {
            H[dst] = H[src];
            E[dst] = E[src];
        }
    }
    __host__ __device__ __forceinline__ void clear() {
#pragma unroll
        for (int q = 0; q < C; q++) set(q, padH, padE);
    }
    // No clamping: with the plan's bounds H lies in [0, 32767) or is the pad
    // -16384, and E in [-16384 - go, 32767) with go <= 16000.
    static __host__ __device__ __forceinline__ uint32_t pack16(int32_t hv, int32_t ev) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_perm((uint32_t)ev, (uint32_t)hv, 0x05040100u);  // v_perm_b32
#else
        return ((uint32_t)hv & 0xFFFFu) | ((uint32_t)ev << 16);
#endif
    }
};

// One row of the band.  MASK: some positions fall outside the band or the
// window and must hold pads.
// The up neighbour of the last position (q = C - 1) lies one position past the
// stored band: it is a real cell when the centre moved by less than one
// column (the band was shifted towards higher positions and its top element
// dropped), else outside the previous band.  The caller passes it in.
template <int C, bool PACK, bool MASK>
__host__ __device__ __forceinline__ Row band_row(Band<C, PACK> &bd, const uint8_t *ref,
                                                 int32_t base, int32_t js, int32_t je, uint8_t qb,
                                                 const SwLaunch &a, uint32_t *trace_row,
                                                 int lane, int32_t top_h, int32_t top_e) {
    const int32_t A = a.match_score, B = a.mismatch, go = a.go, ge = a.gap_extend;
    // the F chain lives in registers only: exact i32 absent values
    const int32_t fpadH = kMin4 + go, fpadF = kMin4 + ge;
    int32_t hl = fpadH, fl = fpadF;
    Row r{0, -1};
    // packed bands keep the row's first strict maximum as one max over
    // (h << 9 | 511 - q): h < 32767 so the key never overflows
    int32_t key = 0;
    uint32_t word = 0;
    // Interior rows fetch the band's reference bytes (column j = base + q is
    // ref[j - 1]) as aligned dwords, funnel-shifted once per row; the arena is
    // padded so the <= 8 bytes read past a window's ends stay allocated.
    constexpr int NW = (C + 3) / 4 + 1;
    const uint8_t *rp = ref + (base - 1);
    uint32_t rw[NW] = {};
    if (!MASK) {
        const uintptr_t ra = (uintptr_t)rp;
        const uint32_t *wp = (const uint32_t *)(ra & ~(uintptr_t)3);
        const uint32_t sh = (uint32_t)(ra & 3) * 8u;
        uint32_t prev = wp[0];
#pragma unroll
        for (int k = 0; k < NW - 1; k++) {
            uint32_t nx = wp[k + 1];
            rw[k] = funnel(nx, prev, sh);
            prev = nx;
        }
    }
#pragma unroll
    for (int q = 0; q < C; q++) {
        int32_t d = bd.h(q);
        int32_t uh = q + 1 < C ? bd.h(q + 1) : top_h;
        int32_t ue = q + 1 < C ? bd.e(q + 1) : top_e;
        bool valid = true;
        if (MASK) valid = (base + q >= js) && (base + q <= je);
        // edge rows may reach past the window (and the arena): guarded loads
        uint8_t rb = MASK ? (valid ? rp[q] : (uint8_t)0) : (uint8_t)(rw[q >> 2] >> (8 * (q & 3)));
        int32_t s = (qb == rb) ? A : -B;
        int32_t hm = imax(d, 0) + s;
        int32_t efh = uh - go, efe = ue - ge;
        int32_t e = imax(efh, efe);
        int32_t ffh = hl - go, fff = fl - ge;
        int32_t f = imax(ffh, fff);
        int32_t h;
        uint32_t dir;
        if constexpr (PACK) {
            // keyed maximum: value << 2 | priority with STOP (the zero
            // candidate) 3 > diagonal 2 > E 1 > F 0 resolves ties as the
            // reference does; direction = 3 - priority.  Packed bands keep
            // every value in [-16384 - go, 32767), so the keys cannot overflow.
            const int32_t K = imax(imax((hm << 2) | 2, (e << 2) | 1), imax(f << 2, 3));
            h = K >> 2;
            dir = ~(uint32_t)K & 3u;
        } else {
            h = imax(imax(hm, e), imax(f, 0));
            dir = h == 0 ? 0u : (h == hm ? 1u : (h == e ? 2u : 3u));
        }
        uint32_t code = dir | (efe > efh ? 4u : 0u) | (fff > ffh ? 8u : 0u);
        if (MASK && !valid) {
            bd.set(q, bd.padH, bd.padE);
            hl = fpadH;
            fl = fpadF;
            code = 0;
        } else {
            bd.set(q, h, e);
            hl = h;
            fl = f;
            if constexpr (PACK) {
                key = imax(key, (h << 9) | (511 - q));
            } else if (h > r.rb) {
                r.rb = h;
                r.rbq = q;
            }
        }
        word |= code << (4 * (q & 7));
        if ((q & 7) == 7 || q == C - 1) {
            trace_row[(q >> 3) * 64 + lane] = word;
            word = 0;
        }
    }
    if constexpr (PACK) {
        r.rb = key >> 9;
        r.rbq = r.rb > 0 ? 511 - (key & 511) : -1;
    }
    return r;
}

// Shift the band so that position q holds what was at q + delta.  top_h /
// top_e receive what was at C + delta (the up neighbour of position C - 1
// in the shifted band): the element the last single-step shift towards
// higher positions dropped, else the pads.
template <int C, bool PACK>
__host__ __device__ __forceinline__ void shift_band(Band<C, PACK> &bd, int32_t delta,
                                                    int32_t &top_h, int32_t &top_e) {
    top_h = bd.padH;
    top_e = bd.padE;
    if (delta >= C || delta <= -C) {
        bd.clear();
        return;
    }
    while (delta > 0) {
#pragma unroll
        for (int q = 0; q < C - 1; q++) bd.move(q, q + 1);
        bd.set(C - 1, bd.padH, bd.padE);
        delta--;
    }
    while (delta < 0) {
        top_h = bd.h(C - 1);
        top_e = bd.e(C - 1);
#pragma unroll
        for (int q = C - 1; q > 0; q--) bd.move(q, q - 1);
        bd.set(0, bd.padH, bd.padE);
        delta++;
    }
}

// One job's rows (banded_sw_with, mod.rs:1370-1579); rows <= the packet's
// row count so the slab holds every row written.
template <int C, bool PACK>
__host__ __device__ int2 sw_job(const SwLaunch &a, int32_t job, int lane, int32_t rows,
                                uint32_t *slab, int2 *meta) {
    const int32_t rowdw = (C + 7) / 8;
    const int32_t L = a.read_len[job], W = a.ref_len[job];
    const int32_t off = a.offset[job];
    const int32_t w = imax(a.band ? a.band[job] : a.bandwidth, 1);
    const int32_t abort_score = a.abort_score[job];
    const uint8_t *read = a.reads + a.read_off[job];
    const uint8_t *ref = a.refs + a.ref_off[job];
    const int32_t D = a.adaptive ? a.max_drift : 0;
    const int32_t xdrop = a.dp_xdrop;

    Band<C, PACK> bd;
    if constexpr (PACK) {
        bd.padH = -16384;
        bd.padE = -16384 - (a.go - a.gap_extend);
    } else {
        bd.padH = kMin4 + a.go;
        bd.padE = kMin4 + a.gap_extend;
    }
    bd.clear();
    int32_t ac = 1 + off;
    int32_t prev_c = 0;
    bool prev_empty = true;
    int32_t best = 0, bi = 0, bj = 0, bq = 0, bqj = 0;
    int32_t early = 0, cells = 0;
    for (int32_t i = 1; i <= L && i <= rows; i++) {
        const int32_t sc = i + off;
        const int32_t c = a.adaptive ? imin(imax(ac, sc - D), sc + D) : sc;
        const int32_t base = c - w;
        const int32_t js = imax(base, 1), je = imin(c + w, W);
        uint32_t *trow = slab + (int64_t)(i - 1) * rowdw * 64;
        if (js > je) {
            // empty row: the reference clears the previous row and skips the
            // centre update and the stop checks (mod.rs:1436-1443)
            bd.clear();
            prev_empty = true;
            meta[(int64_t)(i - 1) * 64 + lane] = make_int2(base, 1);
            continue;
        }
        int32_t top_h = bd.padH, top_e = bd.padE;
        if (prev_empty) {
            bd.clear();
        } else {
            int32_t delta = c - prev_c - 1;
            if (delta != 0) shift_band<C, PACK>(bd, delta, top_h, top_e);
        }
        prev_c = c;
        prev_empty = false;
        cells += je - js + 1;
        const uint8_t qb = read[i - 1];
        Row r;
        if (js == base && je == base + C - 1)
            r = band_row<C, PACK, false>(bd, ref, base, js, je, qb, a, trow, lane, top_h, top_e);
        else
            r = band_row<C, PACK, true>(bd, ref, base, js, je, qb, a, trow, lane, top_h, top_e);
        meta[(int64_t)(i - 1) * 64 + lane] = make_int2(base, (js - base) | ((je - base) << 16));
        const int32_t rbj = r.rbq >= 0 ? base + r.rbq : -1;
        if (r.rb > best) {
            best = r.rb;
            bi = i;
            bj = rbj;
        }
        if (i == L && r.rb > 0) {
            bq = r.rb;
            bqj = rbj;
        }
        if (a.adaptive) {
            int32_t raw = rbj >= 0 ? rbj + 1 : ac + 1;
            int32_t sn = i + 1 + off;
            ac = imin(imax(raw, sn - D), sn + D);
        }
        if (xdrop > 0 && best - r.rb > xdrop) {
            early = 1;
            break;
        }
        if (abort_score > kMin8) {
            int32_t remaining = L - i;
            if (best + remaining * a.match_score < abort_score) {
                early = 1;
                break;
            }
        }
    }
    const int32_t max_h = best;
    if (bq > 0 && sat_add(bq, a.clip_penalty) > best) {
        best = bq;
        bi = L;
        bj = bqj;
    }
    int32_t *res = a.res + (int64_t)job * kSwResWords;
    res[0] = best;
    res[1] = bi;  // read_end
    res[2] = bj;  // ref_end
    res[3] = early;
    // a packed band must never have clamped a score (the plan's bound)
    res[7] = (PACK && max_h > 32767) ? 1 : 0;
    res[8] = cells;
    res[9] = 0;
    return make_int2(bi, bj);
}

// Traceback (mod.rs:1581-1626) of a lane-per-job slab ([row][q / 8][lane]).
// Each round prefetches the meta and trace codes of the next kTbDepth cells
// on the current diagonal (cells (i - k, j - k)), so a run of diagonal
// moves costs one pair of dependent loads per kTbDepth cells.
constexpr int kTbDepth = 4;

__host__ __device__ __forceinline__ void sw_traceback_job(const SwLaunch &a, int32_t job,
                                                          int32_t bi, int32_t bj,
                                                          const uint32_t *slab, const int2 *meta,
                                                          int32_t rowdw, int32_t lane) {
    uint32_t *cig = a.cigar + a.cigar_off[job];
    const int32_t cap = a.cigar_cap[job];
    int32_t i = bi, j = bj, n = 0, last_op = -1, run = 0, state = 0;
    bool stop = false;
    auto flush = [&]() {
        if (last_op >= 0) {
            if (n < cap) cig[n] = ((uint32_t)run << 4) | (uint32_t)last_op;
            n++;
        }
    };
    while (!stop && i > 0 && j > 0) {
        uint32_t tw[kTbDepth];
        bool ok[kTbDepth];
#pragma unroll
        for (int k = 0; k < kTbDepth; k++) {
            const int32_t pi = i - k, pj = j - k;
            tw[k] = 0;
            ok[k] = false;
            if (pi >= 1 && pj >= 1) {
                const int2 m = meta[(int64_t)(pi - 1) * 64 + lane];
                const int32_t base = m.x, js = base + (m.y & 0xFFFF), je = base + (m.y >> 16);
                if (pj >= js && pj <= je) {
                    const int32_t q = pj - base;
                    tw[k] = (slab[((int64_t)(pi - 1) * rowdw + (q >> 3)) * 64 + lane] >>
                             (4 * (q & 7))) & 15u;
                    ok[k] = true;
                }
            }
        }
        const int32_t i0 = i, j0 = j;
        for (;;) {
            const int32_t k = i0 - i;
            if (k >= kTbDepth || j0 - j != k || i <= 0 || j <= 0) break;
            uint32_t tr = tw[0];
            bool okk = ok[0];
#pragma unroll
            for (int x = 1; x < kTbDepth; x++)
                if (k == x) {
                    tr = tw[x];
                    okk = ok[x];
                }
            if (!okk) {  // the cell lies outside its row's band
                stop = true;
                break;
            }
            int op = -1;
            if (state == 0) {
                const uint32_t d = tr & 3u;
                if (d == 0) {
                    stop = true;
                    break;
                }
                if (d == 1) {
                    op = 0;
                    i--;
                    j--;
                } else {
                    state = d == 2 ? 1 : 2;
                }
            } else if (state == 1) {
                op = 1;
                i--;
                state = (tr & 4u) ? 1 : 0;
            } else {
                op = 2;
                j--;
                state = (tr & 8u) ? 2 : 0;
            }
            if (op >= 0) {
                if (op == last_op) {
                    run++;
                } else {
                    flush();
                    last_op = op;
                    run = 1;
                }
            }
        }
    }
    flush();
    // ops were produced end-to-start: reverse in place
    const int32_t m = n < cap ? n : cap;
    for (int32_t x = 0, y = m - 1; x < y; x++, y--) {
        uint32_t t = cig[x];
        cig[x] = cig[y];
        cig[y] = t;
    }
    int32_t *res = a.res + (int64_t)job * kSwResWords;
    res[4] = j;  // ref_start
    res[5] = i;  // read_start
    res[6] = n <= cap ? n : -1;
}

template <int C, bool PACK>
__global__ __launch_bounds__(64) void banded_sw_kernel(SwLaunch a) {
    const int lane = threadIdx.x;
    for (;;) {
        int32_t p = 0;
        if (lane == 0) p = atomicAdd(a.counter, 1);
        p = __shfl(p, 0);
        if (p >= a.n_packets) return;
        const int32_t job = a.packet_jobs[(int64_t)p * 64 + lane];
        if (job >= 0 && sw_job_skipped(a, job)) {
            sw_write_skipped(a, job);
        } else if (job >= 0) {
            uint32_t *slab = a.trace + a.packet_trace_off[p];
            int2 *meta = a.row_meta + a.packet_meta_off[p];
            const int2 end = sw_job<C, PACK>(a, job, lane, a.packet_rows[p], slab, meta);
            // the lane reads back only the trace it wrote itself
            sw_traceback_job(a, job, end.x, end.y, slab, meta, (C + 7) / 8, lane);
        }
    }
}

__global__ void zero_counter_kernel(int32_t *c) { *c = 0; }

}  // namespace

hipError_t launch_banded_sw(const SwLaunch &a, int C, bool pack, int n_waves, hipStream_t s) {
    zero_counter_kernel<<<1, 1, 0, s>>>(a.counter);
    if (a.n_packets > 0) {
#define KIRA_SW_CASE(CC)                                                        \
    case CC:                                                                    \
        if (pack)                                                               \
            banded_sw_kernel<CC, true><<<n_waves, 64, 0, s>>>(a);               \
        else                                                                    \
            banded_sw_kernel<CC, false><<<n_waves, 64, 0, s>>>(a);              \
        break;
        switch (C) {
            KIRA_SW_CASE(9)
            KIRA_SW_CASE(17)
            KIRA_SW_CASE(33)
            KIRA_SW_CASE(65)
            KIRA_SW_CASE(129)
            KIRA_SW_CASE(257)
        default: return hipErrorInvalidValue;
        }
#undef KIRA_SW_CASE
    }
    return hipGetLastError();
}


}  // namespace kira
