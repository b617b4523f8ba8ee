// Banded adaptive Smith-Waterman (banded_sw_with, src/alignment/mod.rs:1370-1650)
// for wide bands: a group of G = 16 (one DPP row) or 32 lanes runs one job
// (8, half a row, for the narrower shapes of the packed form).
//
// A lane-per-job band of 2w + 1 > 129 cells no longer fits the register
// file, so here the band of row i is spread over the G lanes of a group:
// lane g holds positions q = g*P .. g*P + P - 1 (diagonal coordinates,
// column j = c_i - w + q, as in the lane-per-job kernel), and C = G P >=
// 2w + 1 (positions past 2w are always outside the band).  Wide bands use
// 32 lanes rather than more positions per lane, which keeps a lane's band
// slice small enough for several waves per SIMD.
//
// Per row every cell's diagonal / E term is local (the up neighbour of a
// lane's last position is the next lane's first, one DPP move).  The
// horizontal gap F is a max-plus prefix: with go = open + extend >= extend,
//   F[q] = max(Hpre[q-1] - go, F[q-1] - ge),   Hpre = max(diag + s, E, 0)
// equals the reference's F over the final H, so each lane runs its own
// chain, a 4-step DPP scan carries the chains across the lanes of a row
// (decaying by P * ge per lane; a row_bcast:15 step joins two rows for
// G = 32), and each lane continues its chain from that carry.  The row's
// first strict arg-max is a DPP rotate-reduce of (score << 10 | 1023 - q).  Everything else (pads for absent cells, trace
// codes, centre update, x-drop / abort stops, last-row clip rule) follows
// the lane-per-job kernel, so both give identical results.
//
// Trace slab per packet of 64 / G jobs: [row][word][wave lane], word k of
// lane g holding the 4-bit codes of positions g*P + 8k .. g*P + 8k + 7.
#include "internal.hpp"

namespace kira {
namespace {

constexpr int32_t kMin4 = INT32_MIN / 4;
constexpr int32_t kMin8 = INT32_MIN / 8;

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__device__ __forceinline__ int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

// DPP row controls (16-lane rows): row_shl:n, lane l reads lane l + n;
// row_shr:n, lane l reads lane l - n; row_ror:n rotates within the row.
// Lanes whose source falls outside the row keep `fill`.
template <int N>
__device__ __forceinline__ int32_t from_next(int32_t v, int32_t fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x100 + N, 0xF, 0xF, false);
}
template <int N>
__device__ __forceinline__ int32_t from_prev(int32_t v, int32_t fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x110 + N, 0xF, 0xF, false);
}
// Neighbour lanes inside a group of G lanes (g: lane within the group):
// G = 16 is one DPP row; G = 8 is half a row and G = 32 uses the wave
// shifts, both patching the group edge.  Lanes without a neighbour read `fill`.
template <int G>
__device__ __forceinline__ int32_t nb_next(int32_t v, int32_t fill, int g) {
    if constexpr (G == 16) {
        return __builtin_amdgcn_update_dpp(fill, v, 0x101, 0xF, 0xF, false);  // row_shl:1
    } else if constexpr (G == 8) {  // two groups a row
        const int32_t t = __builtin_amdgcn_update_dpp(fill, v, 0x101, 0xF, 0xF, false);
        return g == G - 1 ? fill : t;
    } else {
        const int32_t t = __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);  // wave_shl:1
        return g == G - 1 ? fill : t;
    }
}
template <int G>
__device__ __forceinline__ int32_t nb_prev(int32_t v, int32_t fill, int g) {
    if constexpr (G == 16) {
        return __builtin_amdgcn_update_dpp(fill, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    } else if constexpr (G == 8) {
        const int32_t t = __builtin_amdgcn_update_dpp(fill, v, 0x111, 0xF, 0xF, false);
        return g == 0 ? fill : t;
    } else {
        const int32_t t = __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
        return g == 0 ? fill : t;
    }
}
// In-place forms (bound_ctrl: lanes without a source read 0); the caller
// overwrites the group-edge lane, so no copy of the source array is needed.
template <int G>
__device__ __forceinline__ int32_t next_lane0(int32_t v) {
    return __builtin_amdgcn_mov_dpp(v, G <= 16 ? 0x101 : 0x130, 0xF, 0xF, true);
}
template <int G>
__device__ __forceinline__ int32_t prev_lane0(int32_t v) {
    return __builtin_amdgcn_mov_dpp(v, G <= 16 ? 0x111 : 0x138, 0xF, 0xF, true);
}
template <int N>
__device__ __forceinline__ int32_t rotate(int32_t v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + N, 0xF, 0xF, false);
}
// from_prev<N> inside groups of G lanes: lanes g < N of a group read fill
// (rows of 16 lanes hold two groups of 8)
template <int N, int G>
__device__ __forceinline__ int32_t grp_prev(int32_t v, int32_t fill, int g) {
    const int32_t t = from_prev<N>(v, fill);
    if constexpr (G == 8) return g >= N ? t : fill;
    else return t;
}
// maximum over the G lanes of a group, in every lane of it
template <int G>
__device__ __forceinline__ int32_t group_max(int32_t key) {
    if constexpr (G == 8) {
        key = imax(key, __builtin_amdgcn_update_dpp(0, key, 0xB1, 0xF, 0xF, false));  // quad [1,0,3,2]
        key = imax(key, __builtin_amdgcn_update_dpp(0, key, 0x4E, 0xF, 0xF, false));  // quad [2,3,0,1]
        key = imax(key, __builtin_amdgcn_update_dpp(0, key, 0x141, 0xF, 0xF, false)); // row_half_mirror
    } else {
        key = imax(key, rotate<8>(key));
        key = imax(key, rotate<4>(key));
        key = imax(key, rotate<2>(key));
        key = imax(key, rotate<1>(key));
        if constexpr (G == 32) key = imax(key, __shfl_xor(key, 16));
    }
    return key;
}

__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
}

template <int P, int G>
struct CoopBand {
    int32_t H[P], E[P];
    int32_t padH, padE;
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int p = 0; p < P; p++) {
            H[p] = padH;
            E[p] = padE;
        }
    }
    // new[q] = old[q + delta] over the whole band (g: lane in the group)
    __device__ __forceinline__ void shift(int32_t delta, int g) {
        if (delta >= G * P || delta <= -G * P) {
            clear();
            return;
        }
        while (delta >= P) {  // whole lanes: lane g takes lane g + 1's positions
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int32_t h = next_lane0<G>(H[p]), e = next_lane0<G>(E[p]);
                H[p] = g == G - 1 ? padH : h;
                E[p] = g == G - 1 ? padE : e;
            }
            delta -= P;
        }
        while (delta <= -P) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int32_t h = prev_lane0<G>(H[p]), e = prev_lane0<G>(E[p]);
                H[p] = g == 0 ? padH : h;
                E[p] = g == 0 ? padE : e;
            }
            delta += P;
        }
        while (delta > 0) {
            int32_t nh = nb_next<G>(H[0], padH, g), ne = nb_next<G>(E[0], padE, g);
#pragma unroll
            for (int p = 0; p < P - 1; p++) {
                H[p] = H[p + 1];
                E[p] = E[p + 1];
            }
            H[P - 1] = nh;
            E[P - 1] = ne;
            delta--;
        }
        while (delta < 0) {
            int32_t nh = nb_prev<G>(H[P - 1], padH, g), ne = nb_prev<G>(E[P - 1], padE, g);
#pragma unroll
            for (int p = P - 1; p > 0; p--) {
                H[p] = H[p - 1];
                E[p] = E[p - 1];
            }
            H[0] = nh;
            E[0] = ne;
            delta++;
        }
    }
};

// One job on a group of G lanes; g = lane within the group, wl = wave lane
// of g = 0 (the job's slot column in the packet's meta slab).
template <int P, int G>
__device__ int2 coop_job(const SwLaunch &a, int32_t job, int g, int wl, int32_t rows,
                         uint32_t *slab, int2 *meta) {
    constexpr int RW = (P + 7) / 8;
    constexpr int NW = (P + 3) / 4 + 1;
    const int32_t L = a.read_len[job], W = a.ref_len[job];
    const int32_t off = a.offset[job];
    const int32_t w = imax(a.band ? a.band[job] : a.bandwidth, 1);
    const int32_t abort_score = a.abort_score[job];
    const uint8_t *read = a.reads + a.read_off[job];
    const uint8_t *ref = a.refs + a.ref_off[job];
    const int32_t D = a.adaptive ? a.max_drift : 0;
    const int32_t xdrop = a.dp_xdrop;
    const int32_t A = a.match_score, B = a.mismatch, go = a.go, ge = a.gap_extend;
    const int32_t q0 = g * P;

    CoopBand<P, G> bd;
    bd.padH = kMin4 + go;
    bd.padE = kMin4 + ge;
    bd.clear();
    int32_t ac = 1 + off, prev_c = 0;
    bool prev_empty = true;
    int32_t best = 0, bi = 0, bj = 0, bq = 0, bqj = 0, early = 0, cells = 0;
    for (int32_t i = 1; i <= L && i <= rows; i++) {
        const int32_t sc = i + off;
        const int32_t c = a.adaptive ? imin(imax(ac, sc - D), sc + D) : sc;
        const int32_t base = c - w;
        const int32_t js = imax(base, 1), je = imin(c + w, W);
        uint32_t *trow = slab + (int64_t)(i - 1) * RW * 64 + wl + g;
        if (js > je) {
            bd.clear();
            prev_empty = true;
            if (g == 0) meta[(int64_t)(i - 1) * 64 + wl] = make_int2(base, 1);
            continue;
        }
        if (prev_empty) {
            bd.clear();
        } else {
            int32_t delta = c - prev_c - 1;
            if (delta != 0) bd.shift(delta, g);
        }
        prev_c = c;
        prev_empty = false;
        cells += je - js + 1;
        const uint8_t qb = read[i - 1];
        // reference bytes of this lane's columns base + q0 + p (ref[j - 1]);
        // every window has 1024 readable bytes on both sides and a non-empty
        // row keeps these columns within 530 of the window
        const uintptr_t ra = (uintptr_t)(ref + (base - 1 + q0));
        const uint32_t *wp = (const uint32_t *)(ra & ~(uintptr_t)3);
        const uint32_t sh = (uint32_t)(ra & 3) * 8u;
        uint32_t rw[NW - 1];
        {
            uint32_t prev = wp[0];
#pragma unroll
            for (int k = 0; k < NW - 1; k++) {
                uint32_t nx = wp[k + 1];
                rw[k] = funnel(nx, prev, sh);
                prev = nx;
            }
        }
        // valid positions of this lane: p in [vlo, vhi]
        const int32_t vlo = js - base - q0, vhi = je - base - q0;
        const int32_t upH = nb_next<G>(bd.H[0], bd.padH, g), upE = nb_next<G>(bd.E[0], bd.padE, g);
        // pass 1: diagonal and E terms, Hpre = max(diag + s, E, 0), this
        // row's E (stored over the previous row's as soon as it is read)
        // and the lane's F chain with nothing coming in
        int32_t hm[P], hp[P];
        uint64_t ebits = 0;  // E-extend bit of each position (P may exceed 32)
        int32_t f = kMin4;
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int32_t d = bd.H[p];
            const int32_t uh = p + 1 < P ? bd.H[p + 1] : upH;
            const int32_t ue = p + 1 < P ? bd.E[p + 1] : upE;
            const uint8_t rb = (uint8_t)(rw[p >> 2] >> (8 * (p & 3)));
            const int32_t s = (qb == rb) ? A : -B;
            hm[p] = imax(d, 0) + s;
            const int32_t efh = uh - go, efe = ue - ge;
            const int32_t e = imax(efh, efe);
            ebits |= (uint64_t)(efe > efh ? 1u : 0u) << p;
            const bool valid = p >= vlo && p <= vhi;
            hp[p] = valid ? imax(imax(hm[p], e), 0) : bd.padH;
            bd.E[p] = valid ? e : bd.padE;
            f = imax(hp[p] - go, f - ge);
        }
        // F at this lane's first position: max-plus scan of the chains of
        // the lanes before it (each decays by P * ge across a lane)
        const int32_t dP = P * ge;
        int32_t F = from_prev<1>(f, kMin4);
        F = imax(F, from_prev<1>(F, kMin4) - dP);
        F = imax(F, from_prev<2>(F, kMin4) - 2 * dP);
        F = imax(F, from_prev<4>(F, kMin4) - 4 * dP);
        F = imax(F, from_prev<8>(F, kMin4) - 8 * dP);
        if constexpr (G == 32) {
            // rows 1 and 3 add the carry out of rows 0 and 2: lane 15's F at
            // the next lane's start, decayed per lane of the receiving row
            const int32_t t = imax(f, F - dP);
            const int32_t y = __builtin_amdgcn_update_dpp(kMin4, t, 0x142, 0xA, 0xF, false);
            F = imax(F, y - (g & 15) * dP);
        }
        // pass 2: F along the lane, final H, trace codes, row key.  The
        // F-extend bit of position p compares F and H of position p - 1;
        // for p = 0 that is the previous lane's last position, patched into
        // word 0 after the loop.
        int32_t key = 0, prevF = kMin4, prevH = bd.padH;
        bool prevValid = false;
        uint32_t word = 0, word0 = 0;
#pragma unroll
        for (int p = 0; p < P; p++) {
            const bool valid = p >= vlo && p <= vhi;
            const uint32_t vm = valid ? 0xFFFFFFFFu : 0u;
            const int32_t e = bd.E[p];
            // Keyed maximum of the H candidates: value << 2 | priority with
            // STOP (the zero candidate) 3 > diagonal 2 > E 1 > F 0, so ties
            // resolve as the reference orders them and the low two bits name
            // the winner (direction = 3 - priority).  Every value here is
            // >= i32::MIN/4 and < 2^21, so the shifted keys do not overflow.
            const int32_t K = imax(imax((hm[p] << 2) | 2, (e << 2) | 1), imax(F << 2, 3));
            const int32_t h = K >> 2;
            uint32_t fb = (p > 0 && prevValid && (prevF - ge > prevH - go)) ? 8u : 0u;
            uint32_t code = (~(uint32_t)K & 3u) | ((uint32_t)((ebits >> p) & 1u) << 2) | fb;
            code &= vm;
            key = imax(key, (int32_t)((uint32_t)((h << 10) | (1023 - (q0 + p))) & vm));
            bd.H[p] = (int32_t)(((uint32_t)h & vm) | ((uint32_t)bd.padH & ~vm));
            const int32_t nextF = imax(hp[p] - go, F - ge);
            prevF = F;
            prevH = h;
            prevValid = valid;
            F = nextF;
            word |= code << (4 * (p & 7));
            if ((p & 7) == 7 || p == P - 1) {
                if (p < 8) word0 = word;
                else trow[(p >> 3) * 64] = word;
                word = 0;
            }
        }
        {
            const int32_t lf = nb_prev<G>(prevF, kMin4, g), lh = nb_prev<G>(prevH, bd.padH, g);
            const bool lv = (q0 - 1 >= js - base) && (q0 - 1 <= je - base);
            if (vlo <= 0 && vhi >= 0 && lv && (lf - ge > lh - go)) word0 |= 8u;
            trow[0] = word0;
        }
        // first strict row maximum over the 16 lanes
        key = imax(key, rotate<8>(key));
        key = imax(key, rotate<4>(key));
        key = imax(key, rotate<2>(key));
        key = imax(key, rotate<1>(key));
        if constexpr (G == 32) key = imax(key, __shfl_xor(key, 16));
        const int32_t rb_all = key >> 10;
        const int32_t rbj = rb_all > 0 ? base + (1023 - (key & 1023)) : -1;
        if (g == 0)
            meta[(int64_t)(i - 1) * 64 + wl] = make_int2(base, (js - base) | ((je - base) << 16));
        if (rb_all > best) {
            best = rb_all;
            bi = i;
            bj = rbj;
        }
        if (i == L && rb_all > 0) {
            bq = rb_all;
            bqj = rbj;
        }
        if (a.adaptive) {
            int32_t raw = rbj >= 0 ? rbj + 1 : ac + 1;
            int32_t sn = i + 1 + off;
            ac = imin(imax(raw, sn - D), sn + D);
        }
        if (xdrop > 0 && best - rb_all > xdrop) {
            early = 1;
            break;
        }
        if (abort_score > kMin8) {
            int32_t remaining = L - i;
            if (best + remaining * A < abort_score) {
                early = 1;
                break;
            }
        }
    }
    if (bq > 0) {
        int64_t sum = (int64_t)bq + a.clip_penalty;
        int32_t sat = sum > INT32_MAX ? INT32_MAX : (sum < INT32_MIN ? INT32_MIN : (int32_t)sum);
        if (sat > best) {
            best = bq;
            bi = L;
            bj = bqj;
        }
    }
    if (g == 0) {
        int32_t *res = a.res + (int64_t)job * kSwResWords;
        res[0] = best;
        res[1] = bi;
        res[2] = bj;
        res[3] = early;
        res[7] = 0;
        res[8] = cells;
        res[9] = 0;
    }
    return make_int2(bi, bj);
}

// Traceback (mod.rs:1581-1626) by the job's G lanes together: each round
// lane k fetches the trace code of cell (i - k, j - k), so a run of
// diagonal moves costs one dependent meta + trace load per G cells instead
// of per cell; the walk itself (the same state machine as the lane-per-job
// traceback) runs redundantly in all G lanes and lane 0 writes the CIGAR.
// SPLIT = 0: word k of a lane holds the 4-bit codes of lane positions
// 8k .. 8k + 7; SPLIT = P2 > 0 (the packed form): word k holds registers
// 4k .. 4k + 3, lane positions r (nibble r & 3) and P2 + r (nibble 4 + (r & 3)).
template <int P, int G, int SPLIT = 0>
__device__ void coop_traceback(const SwLaunch &a, int32_t job, int g, int wl, int2 end,
                               const uint32_t *slab, const int2 *meta) {
    constexpr int RW = (P + 7) / 8;
    static_assert(SPLIT == 0 || (SPLIT + 3) / 4 == RW, "split layout words");
    uint32_t *cig = a.cigar + a.cigar_off[job];
    const int32_t cap = a.cigar_cap[job];
    int32_t i = end.x, j = end.y, n = 0, last_op = -1, run = 0, state = 0;
    bool stop = false;
    auto flush = [&]() {
        if (last_op >= 0) {
            if (g == 0 && n < cap) cig[n] = ((uint32_t)run << 4) | (uint32_t)last_op;
            n++;
        }
    };
    while (!stop && i > 0 && j > 0) {
        const int32_t pi = i - g, pj = j - g;
        uint32_t tw = 0;
        int32_t ok = 0;
        if (pi >= 1 && pj >= 1) {
            const int2 m = meta[(int64_t)(pi - 1) * 64 + wl];
            const int32_t base = m.x, js = base + (m.y & 0xFFFF), je = base + (m.y >> 16);
            if (pj >= js && pj <= je) {
                const int32_t q = pj - base, lg = q / P, qq = q - lg * P;
                int32_t wd = qq >> 3, nb = qq & 7;
                if constexpr (SPLIT > 0) {
                    const int32_t hi = qq >= SPLIT, r = qq - (hi ? SPLIT : 0);
                    wd = r >> 2;
                    nb = (r & 3) + 4 * hi;
                }
                tw = (slab[((int64_t)(pi - 1) * RW + wd) * 64 + wl + lg] >> (4 * nb)) & 15u;
                ok = 1;
            }
        }
        const int32_t i0 = i, j0 = j;
        for (;;) {
            const int32_t k = i0 - i;
            if (k >= G || j0 - j != k || i <= 0 || j <= 0) break;
            const int32_t okk = __shfl(ok, wl + k);
            const uint32_t tr = (uint32_t)__shfl((int32_t)tw, wl + k);
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
    if (g == 0) {
        // ops were produced end-to-start: reverse in place
        const int32_t m = n < cap ? n : cap;
        for (int32_t x = 0, y = m - 1; x < y; x++, y--) {
            uint32_t t = cig[x];
            cig[x] = cig[y];
            cig[y] = t;
        }
        int32_t *res = a.res + (int64_t)job * kSwResWords;
        res[4] = j;
        res[5] = i;
        res[6] = n <= cap ? n : -1;
    }
}

template <int P, int G>
__global__ __launch_bounds__(64) void banded_sw_coop_kernel(SwLaunch a) {
    const int lane = threadIdx.x;
    const int g = lane & (G - 1), wl = lane & ~(G - 1);
    for (;;) {
        int32_t p = 0;
        if (lane == 0) p = atomicAdd(a.counter, 1);
        p = __shfl(p, 0);
        if (p >= a.n_packets) return;
        const int32_t job = a.packet_jobs[(int64_t)p * 64 + wl];
        if (job >= 0 && sw_job_skipped(a, job)) {
            if (g == 0) sw_write_skipped(a, job);
        } else if (job >= 0) {
            uint32_t *slab = a.trace + a.packet_trace_off[p];
            int2 *meta = a.row_meta + a.packet_meta_off[p];
            const int2 end = coop_job<P, G>(a, job, g, wl, a.packet_rows[p], slab, meta);
            // every lane reads trace words its neighbours wrote
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            coop_traceback<P, G>(a, job, g, wl, end, slab, meta);
        }
    }
}

// ---------------------------------------------------------------------------
// Packed form: two band positions per register as int16 halves, so the
// per-position arithmetic runs as v_pk_* instructions on two positions at
// once.  Register r of lane g holds lane positions r (low half) and P2 + r
// (high half), band position q = g P + lane position, P = 2 P2.  The up
// neighbour of register r is register r + 1 in both halves, and the
// horizontal gap runs as two chains (lane positions 0 .. P2 - 1 in the low
// halves, P2 .. P - 1 in the high halves) advanced together, one packed
// step per register, then joined with what comes in from the left.
//
// Values are kept x4 so that a value and a 2-bit tag share one int16:
//   H4 = 4 H (band storage; 0 outside the band),
//   E4 = 4 E + 1 (outside the band -30719); the E candidates are
//     4 (H - go) + 3 (open) and 4 (E - ge) + 1 (extend), so a tie opens, as
//     in the reference, bit 1 of the winner is clear where E extends, and
//     clearing bit 1 leaves the E key 4 E + 1,
//   the H candidates are keys 4 value + priority, STOP 3 > diagonal 2 > E 1
//     > F 0, as in the int32 form,
//   F4 = 4 F and Hg4 = 4 (Hpre - gap open), Hpre = max(diag + s, E, 0):
//     F[p + 1] = max(Hpre[p] - go, F[p] - ge) reads F4[p + 1] =
//     max(Hg4[p], F4[p]) - 4 ge, and the F-extend bit of position p,
//     F[p - 1] - ge > Hpre[p - 1] - go, is the sign of Hg4[p - 1] - F4[p - 1]
//     (with gap open > 0 this equals the reference's comparison against the
//     final H: where F is the final H, both extend).
// A job is routed here only when all of these stay inside int16: scores <=
// 8191, mismatch <= 8000, open + extend <= 4000, extend <= 500, gap open > 0.
// F values below -7000 are clamped or saturate: an F that low never wins a
// cell and never sets an F-extend bit.
//
// Trace words: word k of lane g holds registers 4k .. 4k + 3, the code of
// register 4k + m's low half in nibble m and its high half in nibble 4 + m.
typedef short s16x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kPadE4 = 0x88018801u;   // E4 = -30719 in both halves
constexpr int32_t kPadF4 = -28000;         // F4 with nothing incoming
constexpr uint32_t kPadF4x2 = 0x92A092A0u; // kPadF4 in both halves
constexpr uint32_t kLow2 = 0xFFFCFFFCu;    // clears the tag bits of both halves
static_assert((uint32_t)(kPadF4 & 0xFFFF) * 0x10001u == kPadF4x2, "F pad");

__device__ __forceinline__ s16x2 as_s16(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t as_u(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t pk2(int32_t x) { return ((uint32_t)x & 0xFFFFu) * 0x10001u; }
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return as_u(__builtin_elementwise_max(as_s16(a), as_s16(b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return as_u(as_s16(a) + as_s16(b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
    return as_u(as_s16(a) - as_s16(b));
}
__device__ __forceinline__ uint32_t pk_sub_sat(uint32_t a, uint32_t b) {
    return as_u(__builtin_elementwise_sub_sat(as_s16(a), as_s16(b)));
}
// 1 in each half that is non-zero, 0 elsewhere (kept opaque so that the
// following multiply-add stays one packed instruction)
__device__ __forceinline__ uint32_t pk_nonzero(uint32_t x) {
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(x), "s"(0x00010001u));
    return r;
}
__device__ __forceinline__ uint32_t bitsel(uint32_t m, uint32_t a, uint32_t b) {
    return (a & m) | (b & ~m);
}

template <int P2, int G>
struct SplitBand {
    uint32_t H[P2], E[P2];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int r = 0; r < P2; r++) {
            H[r] = 0u;
            E[r] = kPadE4;
        }
    }
    // new[q] = old[q + delta] over the whole band (g: lane in the group); a
    // shift by the whole band or more leaves it clear (G lane moves, each
    // filling the group edge, rather than a separate clear, which would keep
    // a copy of the pad constants live in every row)
    __device__ __forceinline__ void shift(int32_t delta, int g) {
        constexpr int P = 2 * P2;
        delta = imin(imax(delta, -G * P), G * P);
        while (delta >= P) {  // whole lanes
#pragma unroll
            for (int r = 0; r < P2; r++) {
                const uint32_t h = (uint32_t)next_lane0<G>((int32_t)H[r]);
                const uint32_t e = (uint32_t)next_lane0<G>((int32_t)E[r]);
                H[r] = g == G - 1 ? 0u : h;
                E[r] = g == G - 1 ? kPadE4 : e;
            }
            delta -= P;
        }
        while (delta <= -P) {
#pragma unroll
            for (int r = 0; r < P2; r++) {
                const uint32_t h = (uint32_t)prev_lane0<G>((int32_t)H[r]);
                const uint32_t e = (uint32_t)prev_lane0<G>((int32_t)E[r]);
                H[r] = g == 0 ? 0u : h;
                E[r] = g == 0 ? kPadE4 : e;
            }
            delta += P;
        }
        if (delta >= P2) {  // lane position p takes p + P2
#pragma unroll
            for (int r = 0; r < P2; r++) {
                H[r] = funnel((uint32_t)nb_next<G>((int32_t)H[r], 0, g), H[r], 16);
                E[r] = funnel((uint32_t)nb_next<G>((int32_t)E[r], (int32_t)kPadE4, g), E[r], 16);
            }
            delta -= P2;
        } else if (delta <= -P2) {
#pragma unroll
            for (int r = 0; r < P2; r++) {
                H[r] = funnel(H[r], (uint32_t)nb_prev<G>((int32_t)H[r], 0, g), 16);
                E[r] = funnel(E[r], (uint32_t)nb_prev<G>((int32_t)E[r], (int32_t)kPadE4, g), 16);
            }
            delta += P2;
        }
        while (delta > 0) {  // one position
            const uint32_t nh = funnel((uint32_t)nb_next<G>((int32_t)H[0], 0, g), H[0], 16);
            const uint32_t ne =
                funnel((uint32_t)nb_next<G>((int32_t)E[0], (int32_t)kPadE4, g), E[0], 16);
#pragma unroll
            for (int r = 0; r < P2 - 1; r++) {
                H[r] = H[r + 1];
                E[r] = E[r + 1];
            }
            H[P2 - 1] = nh;
            E[P2 - 1] = ne;
            delta--;
        }
        while (delta < 0) {
            const uint32_t ph =
                funnel(H[P2 - 1], (uint32_t)nb_prev<G>((int32_t)H[P2 - 1], 0, g), 16);
            const uint32_t pe = funnel(
                E[P2 - 1], (uint32_t)nb_prev<G>((int32_t)E[P2 - 1], (int32_t)kPadE4, g), 16);
#pragma unroll
            for (int r = P2 - 1; r > 0; r--) {
                H[r] = H[r - 1];
                E[r] = E[r - 1];
            }
            H[0] = ph;
            E[0] = pe;
            delta++;
        }
    }
};

template <int P2, int G>
__device__ int2 coop16_job(const SwLaunch &a, int32_t job, int g, int wl, int32_t rows,
                           uint32_t *slab, int2 *meta) {
    constexpr int P = 2 * P2;
    constexpr int RW = (P2 + 3) / 4;
    constexpr int NW = (P + 3) / 4 + 1;
    const int32_t L = a.read_len[job], W = a.ref_len[job];
    const int32_t off = a.offset[job];
    const int32_t w = imax(a.band ? a.band[job] : a.bandwidth, 1);
    const int32_t abort_score = a.abort_score[job];
    const uint8_t *read = a.reads + a.read_off[job];
    const uint8_t *ref = a.refs + a.ref_off[job];
    const int32_t D = a.adaptive ? a.max_drift : 0;
    const int32_t xdrop = a.dp_xdrop;
    const int32_t A = a.match_score, B = a.mismatch, go = a.go, ge = a.gap_extend;
    const int32_t q0 = g * P;
    const uint32_t kS4 = pk2(4 * A + 2), kNAB4 = pk2(-4 * (A + B)), kOpen = pk2(3 - 4 * go),
                   kGe4 = pk2(4 * ge), kGop4 = pk2(4 * (go - ge));
    // lane positions past 2w are outside the band in every row
    uint32_t vstat[P2];
#pragma unroll
    for (int r = 0; r < P2; r++)
        vstat[r] = (q0 + r <= 2 * w ? 0x0000FFFFu : 0u) |
                   (q0 + P2 + r <= 2 * w ? 0xFFFF0000u : 0u);

    SplitBand<P2, G> bd;
    bd.clear();
    int32_t ac = 1 + off, prev_c = 0;
    bool prev_empty = true;
    int32_t best = 0, bi = 0, bj = 0, bq = 0, bqj = 0, early = 0, cells = 0;
    for (int32_t i = 1; i <= L && i <= rows; i++) {
        const int32_t sc = i + off;
        const int32_t c = a.adaptive ? imin(imax(ac, sc - D), sc + D) : sc;
        const int32_t base = c - w;
        const int32_t js = imax(base, 1), je = imin(c + w, W);
        uint32_t *trow = slab + (int64_t)(i - 1) * RW * 64 + wl + g;
        if (js > je) {
            bd.clear();
            prev_empty = true;
            if (g == 0) meta[(int64_t)(i - 1) * 64 + wl] = make_int2(base, 1);
            continue;
        }
        // the band is clear at the job's start and after an empty row
        if (!prev_empty) {
            const int32_t delta = c - prev_c - 1;
            if (delta != 0) bd.shift(delta, g);
        }
        prev_c = c;
        prev_empty = false;
        cells += je - js + 1;
        const uint32_t qbp = (uint32_t)read[i - 1] * 0x00010001u;
        // reference bytes of this lane's columns base + q0 + p (ref[j - 1]),
        // in the bounds the int32 form states
        const uintptr_t ra = (uintptr_t)(ref + (base - 1 + q0));
        const uint32_t *wp = (const uint32_t *)(ra & ~(uintptr_t)3);
        const uint32_t sh = (uint32_t)(ra & 3) * 8u;
        uint32_t rw[NW - 1];
        {
            uint32_t prev = wp[0];
#pragma unroll
            for (int k = 0; k < NW - 1; k++) {
                uint32_t nx = wp[k + 1];
                rw[k] = funnel(nx, prev, sh);
                prev = nx;
            }
        }
        // valid halves: the static mask, narrowed in rows at a window edge
        uint32_t vm[P2];
        const bool edge = js != base || je != base + 2 * w;
        if (__any(edge)) {
            const int32_t lo = js - base - q0, hi = je - base - q0;
#pragma unroll
            for (int r = 0; r < P2; r++)
                vm[r] = (r >= lo && r <= hi ? 0x0000FFFFu : 0u) |
                        (P2 + r >= lo && P2 + r <= hi ? 0xFFFF0000u : 0u);
        } else {
#pragma unroll
            for (int r = 0; r < P2; r++) vm[r] = vstat[r];
        }
        // up neighbours of register P2 - 1: lane positions P2 (register 0's
        // high half) and P (the next lane's position 0)
        const uint32_t upH = funnel((uint32_t)nb_next<G>((int32_t)bd.H[0], 0, g), bd.H[0], 16);
        const uint32_t upE =
            funnel((uint32_t)nb_next<G>((int32_t)bd.E[0], (int32_t)kPadE4, g), bd.E[0], 16);
        // pass 1: substitution, E (stored over the previous row's as soon as
        // it is read), the non-F part of the H key, Hg4, and both F chains
        // with nothing coming in
        uint32_t khe[P2], ek[P2], hg[P2], ffl[P2 + 1];
        ffl[0] = kPadF4x2;
#pragma unroll
        for (int r = 0; r < P2; r++) {
            const uint32_t uh = r + 1 < P2 ? bd.H[r + 1] : upH;
            const uint32_t ue = r + 1 < P2 ? bd.E[r + 1] : upE;
            const int hb = P2 + r;
            const uint32_t rb = __builtin_amdgcn_perm(
                rw[hb >> 2], rw[r >> 2], 0x0C000C00u | (uint32_t)(r & 3) | ((uint32_t)(4 + (hb & 3)) << 16));
            const uint32_t mis = pk_nonzero(rb ^ qbp);
            const uint32_t s4 = as_u(as_s16(mis) * as_s16(kNAB4) + as_s16(kS4));
            const uint32_t hmK = pk_add(bd.H[r], s4);
            const uint32_t eK = pk_max(pk_add(uh, kOpen), pk_sub(ue, kGe4));
            const uint32_t eKey = eK & 0xFFFDFFFDu;
            bd.E[r] = bitsel(vm[r], eKey, kPadE4);
            ek[r] = eK;
            const uint32_t kh = pk_max(pk_max(hmK, eKey), 0x00030003u);
            khe[r] = kh;
            hg[r] = pk_sub(kh & vm[r] & kLow2, kGop4);
            ffl[r + 1] = pk_sub(pk_max(hg[r], ffl[r]), kGe4);
        }
        // the lane's F4 at lane position P with nothing coming in
        const s16x2 fend = as_s16(ffl[P2]);
        const int32_t f = imax((int32_t)fend.y, (int32_t)fend.x - P2 * 4 * ge);
        // F4 at this lane's position 0: max-plus scan over the lanes
        const int32_t dP = P * 4 * ge;
        int32_t F = grp_prev<1, G>(f, kPadF4, g);
        F = imax(F, grp_prev<1, G>(F, kPadF4, g) - dP);
        F = imax(F, grp_prev<2, G>(F, kPadF4, g) - 2 * dP);
        F = imax(F, grp_prev<4, G>(F, kPadF4, g) - 4 * dP);
        if constexpr (G >= 16) F = imax(F, from_prev<8>(F, kPadF4) - 8 * dP);
        if constexpr (G == 32) {
            const int32_t t = imax(f, F - dP);
            const int32_t y = __builtin_amdgcn_update_dpp(kPadF4, t, 0x142, 0xA, 0xF, false);
            F = imax(F, y - (g & 15) * dP);
        }
        F = imax(F, kPadF4);
        // F4 at lane positions 0 and P2, the heads of the two chains
        const int32_t FP2 = imax(imax((int32_t)fend.x, F - P2 * 4 * ge), kPadF4);
        const uint32_t fin = ((uint32_t)F & 0xFFFFu) | ((uint32_t)FP2 << 16);
        // pass 2: final F, H, trace codes, row key.  Register 0's previous
        // positions are lane position -1 (the previous lane's register
        // P2 - 1, high half) and P2 - 1 (register P2 - 1, low half).
        const uint32_t ffLast = pk_max(ffl[P2 - 1], pk_sub_sat(fin, pk2(4 * (P2 - 1) * ge)));
        uint32_t pF = funnel(ffLast, (uint32_t)nb_prev<G>((int32_t)ffLast, (int32_t)kPadF4x2, g), 16);
        uint32_t pH = funnel(hg[P2 - 1],
                             (uint32_t)nb_prev<G>((int32_t)hg[P2 - 1], (int32_t)pk2(-4 * (go - ge)), g),
                             16);
        int32_t key = 0;
        uint32_t word = 0;
#pragma unroll
        for (int r = 0; r < P2; r++) {
            const uint32_t ff =
                r == P2 - 1 ? ffLast : pk_max(ffl[r], pk_sub_sat(fin, pk2(4 * r * ge)));
            const uint32_t K = pk_max(khe[r], ff);
            const uint32_t h4 = K & kLow2 & vm[r];
            bd.H[r] = h4;
            const uint32_t t = pk_sub_sat(pH, pF);  // sign: F-extend
            pF = ff;
            pH = hg[r];
            // trace code: direction 3 - priority, E-extend (bit 1 of eK clear),
            // F-extend (sign of t); the first two are stored inverted here
            // and flipped once per word
            const uint32_t code = (K & 0x00030003u) | ((ek[r] << 1) & 0x00040004u) |
                                  ((t >> 12) & 0x00080008u);
            word |= code << (4 * (r & 3));
            if ((r & 3) == 3 || r == P2 - 1) {
                trow[(r >> 2) * 64] = word ^ 0x77777777u;
                word = 0;
            }
            // (4 h << 16) | (0xFFFF - lane position): first strict maximum
            key = imax(key, imax((int32_t)((h4 << 16) | (uint32_t)(0xFFFF - r)),
                                 (int32_t)((h4 & 0xFFFF0000u) | (uint32_t)(0xFFFF - P2 - r))));
        }
        key = group_max<G>(key - q0);
        const int32_t rb_all = key >> 18;
        const int32_t rbj = rb_all > 0 ? base + (0xFFFF - (key & 0xFFFF)) : -1;
        if (g == 0)
            meta[(int64_t)(i - 1) * 64 + wl] = make_int2(base, (js - base) | ((je - base) << 16));
        if (rb_all > best) {
            best = rb_all;
            bi = i;
            bj = rbj;
        }
        if (i == L && rb_all > 0) {
            bq = rb_all;
            bqj = rbj;
        }
        if (a.adaptive) {
            int32_t raw = rbj >= 0 ? rbj + 1 : ac + 1;
            int32_t sn = i + 1 + off;
            ac = imin(imax(raw, sn - D), sn + D);
        }
        if (xdrop > 0 && best - rb_all > xdrop) {
            early = 1;
            break;
        }
        if (abort_score > kMin8) {
            int32_t remaining = L - i;
            if (best + remaining * A < abort_score) {
                early = 1;
                break;
            }
        }
    }
    if (bq > 0) {
        int64_t sum = (int64_t)bq + a.clip_penalty;
        int32_t sat = sum > INT32_MAX ? INT32_MAX : (sum < INT32_MIN ? INT32_MIN : (int32_t)sum);
        if (sat > best) {
            best = bq;
            bi = L;
            bj = bqj;
        }
    }
    if (g == 0) {
        int32_t *res = a.res + (int64_t)job * kSwResWords;
        res[0] = best;
        res[1] = bi;
        res[2] = bj;
        res[3] = early;
        res[7] = 0;
        res[8] = cells;
        res[9] = 0;
    }
    return make_int2(bi, bj);
}

template <int P2, int G>
__global__ __launch_bounds__(64) void banded_sw_coop16_kernel(SwLaunch a) {
    const int lane = threadIdx.x;
    const int g = lane & (G - 1), wl = lane & ~(G - 1);
    for (;;) {
        int32_t p = 0;
        if (lane == 0) p = atomicAdd(a.counter, 1);
        p = __shfl(p, 0);
        if (p >= a.n_packets) return;
        const int32_t job = a.packet_jobs[(int64_t)p * 64 + wl];
        if (job >= 0 && sw_job_skipped(a, job)) {
            if (g == 0) sw_write_skipped(a, job);
        } else if (job >= 0) {
            uint32_t *slab = a.trace + a.packet_trace_off[p];
            int2 *meta = a.row_meta + a.packet_meta_off[p];
            const int2 end = coop16_job<P2, G>(a, job, g, wl, a.packet_rows[p], slab, meta);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            coop_traceback<2 * P2, G, P2>(a, job, g, wl, end, slab, meta);
        }
    }
}


__global__ void zero_coop_counter_kernel(int32_t *c) { *c = 0; }

}  // namespace

hipError_t launch_banded_sw_coop(const SwLaunch &a, int P, int G, int n_waves, hipStream_t s) {
    zero_coop_counter_kernel<<<1, 1, 0, s>>>(a.counter);
    if (a.n_packets > 0) {
        if (G == 16 && P == 5) banded_sw_coop_kernel<5, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 16 && P == 9) banded_sw_coop_kernel<9, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 16 && P == 17) banded_sw_coop_kernel<17, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 32 && P == 13) banded_sw_coop_kernel<13, 32><<<n_waves, 64, 0, s>>>(a);
        else if (G == 32 && P == 17) banded_sw_coop_kernel<17, 32><<<n_waves, 64, 0, s>>>(a);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_banded_sw_coop16(const SwLaunch &a, int P2, int G, int n_waves, hipStream_t s) {
    zero_coop_counter_kernel<<<1, 1, 0, s>>>(a.counter);
    if (a.n_packets > 0) {
        if (G == 16 && P2 == 3) banded_sw_coop16_kernel<3, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 16 && P2 == 5) banded_sw_coop16_kernel<5, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 16 && P2 == 9) banded_sw_coop16_kernel<9, 16><<<n_waves, 64, 0, s>>>(a);
        else if (G == 32 && P2 == 7) banded_sw_coop16_kernel<7, 32><<<n_waves, 64, 0, s>>>(a);
        else if (G == 32 && P2 == 9) banded_sw_coop16_kernel<9, 32><<<n_waves, 64, 0, s>>>(a);
        else if (G == 8 && P2 == 5) banded_sw_coop16_kernel<5, 8><<<n_waves, 64, 0, s>>>(a);
        else if (G == 8 && P2 == 9) banded_sw_coop16_kernel<9, 8><<<n_waves, 64, 0, s>>>(a);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace kira
