// Full-window lane-group Smith-Waterman: the SIMD batch kernels of the
// reference (sw_batch_avx2_inner src/alignment/mod.rs:1914-2135, sw_batch_int8
// src/alignment/sw_int8_vnni.rs:141-371, sw_batch_neon mod.rs:2139-2324).
//
// The reference advances a lane group row by row and stops the whole group
// after the first row at which every lane is past its abort bound; each
// lane's answer is taken from the rows processed.  Here each job is computed
// over all rows by one lane, recording per row the maximum and its first
// column; a second kernel derives every lane's first "done" row, the group's
// stop row S is their maximum (or the last row if some lane never finishes),
// and a third kernel takes the best cell over rows <= S (first strict maximum
// in row-major order), applies the last-row clip rule when row L was
// processed, and walks the trace back.  Identical answers, no cross-lane
// synchronisation in the DP.
//
// Scores are exact int32.  With the reference's cost gate (mod.rs:1854-1868)
// the saturating int16 / int8 lanes never clamp a value that changes a
// comparison, so exact arithmetic with the same row-0 / column-0 sentinels
// (-16384, -100 or i32::MIN/4) gives the same H, E, F and trace codes.
//
// Layout: one lane sweeps its matrix in vertical strips of 16 columns; the
// strip's previous-row H/E live in registers, the strip's left boundary
// (H, F per row) and the per-row statistics live in per-packet global slabs
// laid out [row][lane] so every access of a wave is one coalesced line.
#include "internal.hpp"

namespace kira {
namespace {

constexpr int kStrip = 16;

__host__ __device__ __forceinline__ int32_t gmax(int32_t a, int32_t b) { return a > b ? a : b; }

template <bool KEYED>
__host__ __device__ __forceinline__ void group_dp_job(const GroupLaunch &a, int32_t job,
                                                      int lane, int32_t rows, int32_t words,
                                                      uint32_t *trace, int2 *bnd, int2 *rstat) {
    const int32_t L = a.read_len[job], W = a.ref_len[job];
    const uint8_t *read = a.reads + a.read_off[job];
    const uint8_t *ref = a.refs + a.ref_off[job];
    const int32_t neg = a.neg[a.group_of[job]];
    const int32_t A = a.match_score, B = a.mismatch, go = a.go, ge = a.gap_extend;
    (void)rows;
    const int32_t nstrips = (W + kStrip - 1) / kStrip;
    for (int32_t st = 0; st < nstrips; st++) {
        const int32_t j0 = st * kStrip + 1;
        uint8_t rb[kStrip];
#pragma unroll
        for (int c = 0; c < kStrip; c++) rb[c] = (j0 + c <= W) ? ref[j0 - 1 + c] : (uint8_t)0;
        int32_t Hup[kStrip], Eup[kStrip];
#pragma unroll
        for (int c = 0; c < kStrip; c++) {
            Hup[c] = 0;
            Eup[c] = neg;
        }
        int32_t diag_row_prev = 0;  // H[i-1][j0-1]
        // row i's left boundary (H, F), its statistics so far and its read
        // byte, loaded two rows ahead: the loop would otherwise wait on one
        // dependent global load per row (each row stores only its own entry,
        // so the early loads read what the previous strip wrote)
        auto ld = [&](int32_t i, int2 &b, int2 &r, uint8_t &q) {
            if (i > L) return;
            if (st == 0) {
                b = make_int2(0, neg);  // column 0 is H = 0, F starts at the sentinel
            } else {
                b = bnd[(int64_t)(i - 1) * 64 + lane];
                r = rstat[(int64_t)(i - 1) * 64 + lane];
            }
            q = read[i - 1];
        };
        int2 b1 = make_int2(0, neg), b2 = b1, r1 = make_int2(0, 0), r2 = r1;
        uint8_t q1 = 0, q2 = 0;
        ld(1, b1, r1, q1);
        ld(2, b2, r2, q2);
        for (int32_t i = 1; i <= L; i++) {
            int32_t hl = b1.x, fl = b1.y;
            const int2 prev_stat = r1;
            const uint8_t qb = q1;
            b1 = b2;
            r1 = r2;
            q1 = q2;
            ld(i + 2, b2, r2, q2);
            int32_t diag = diag_row_prev;
            diag_row_prev = hl;
            int32_t rmax = 0, rarg = 0, rkey = 0;
            uint32_t w0 = 0, w1 = 0;
#pragma unroll
            for (int c = 0; c < kStrip; c++) {
                if (j0 + c <= W) {
                    int32_t s = (qb == rb[c]) ? A : -B;
                    int32_t hm = diag + s;
                    int32_t efh = Hup[c] - go, efe = Eup[c] - ge;
                    int32_t e = gmax(efh, efe);
                    int32_t ffh = hl - go, fff = fl - ge;
                    int32_t f = gmax(ffh, fff);
                    int32_t h;
                    uint32_t dir;
                    if constexpr (KEYED) {
                        // keyed maximum: value << 2 | priority, STOP (the
                        // zero candidate) 3 > diagonal 2 > E 1 > F 0, the
                        // reference's tie order; direction = 3 - priority
                        const int32_t K = gmax(gmax((hm << 2) | 2, (e << 2) | 1), gmax(f << 2, 3));
                        h = K >> 2;
                        dir = ~(uint32_t)K & 3u;
                        // first strict row maximum of the strip
                        rkey = gmax(rkey, (h << 4) | (kStrip - 1 - c));
                    } else {
                        h = gmax(gmax(hm, e), gmax(f, 0));
                        dir = h == 0 ? 0u : (h == hm ? 1u : (h == e ? 2u : 3u));
                        if (h > rmax) {
                            rmax = h;
                            rarg = j0 + c;
                        }
                    }
                    uint32_t code = dir | (efe > efh ? 4u : 0u) | (fff > ffh ? 8u : 0u);
                    diag = Hup[c];
                    Hup[c] = h;
                    Eup[c] = e;
                    hl = h;
                    fl = f;
                    if (c < 8) w0 |= code << (4 * c);
                    else w1 |= code << (4 * (c - 8));
                }
            }
            if constexpr (KEYED) {
                rmax = rkey >> 4;
                rarg = rmax > 0 ? j0 + (kStrip - 1) - (rkey & (kStrip - 1)) : 0;
            }
            uint32_t *tr = trace + ((int64_t)(i - 1) * words + 2 * st) * 64 + lane;
            tr[0] = w0;
            tr[64] = w1;
            bnd[(int64_t)(i - 1) * 64 + lane] = make_int2(hl, fl);
            int2 *rs = rstat + (int64_t)(i - 1) * 64 + lane;
            if (st == 0 || rmax > prev_stat.x) *rs = make_int2(rmax, rarg);
        }
    }
}

template <bool KEYED>
__global__ __launch_bounds__(64) void group_dp_kernel(GroupLaunch a) {
    const int lane = threadIdx.x;
    for (;;) {
        int32_t p = 0;
        if (lane == 0) p = atomicAdd(a.counter, 1);
        p = __shfl(p, 0);
        if (p >= a.n_packets) return;
        const int32_t job = a.packet_jobs[(int64_t)p * 64 + lane];
        if (job >= 0)
            group_dp_job<KEYED>(a, job, lane, a.packet_rows[p], a.packet_words[p],
                         a.trace + a.packet_trace_off[p], a.bnd + a.packet_row_off[p],
                         a.rstat + a.packet_row_off[p]);
    }
}

// First row after which the lane is done (abort bound, mod.rs:2058-2082);
// L when it never is.  The group stops at the maximum over its lanes.
__host__ __device__ __forceinline__ int32_t lane_done_row(const GroupLaunch &a, int32_t job,
                                                           bool &done) {
    const int32_t L = a.read_len[job];
    const int32_t pl = a.job_slot[job];
    const int32_t p = pl >> 6, lane = pl & 63;
    const int2 *rs = a.rstat + a.packet_row_off[p];
    const int32_t ab = a.abort_score[job];
    done = false;
    if (ab <= INT32_MIN / 8) return L;
    int32_t best = 0;
    for (int32_t i = 1; i <= L; i++) {
        best = gmax(best, rs[(int64_t)(i - 1) * 64 + lane].x);
        if (best + (L - i) * a.match_score < ab) {
            done = true;
            return i;
        }
    }
    return L;
}

__global__ void group_stop_kernel(GroupLaunch a, int32_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    bool done;
    int32_t r = lane_done_row(a, (int32_t)t, done);
    atomicMax(&a.group_stop[a.group_of[t]], r);
}

__host__ __device__ void group_finish_job(const GroupLaunch &a, int32_t job) {
    const int32_t L = a.read_len[job];
    const int32_t pl = a.job_slot[job];
    const int32_t p = pl >> 6, lane = pl & 63;
    const int2 *rs = a.rstat + a.packet_row_off[p];
    const uint32_t *trace = a.trace + a.packet_trace_off[p];
    const int32_t words = a.packet_words[p];
    const int32_t S = a.group_stop[a.group_of[job]];
    bool done;
    int32_t dr = lane_done_row(a, job, done);
    (void)dr;
    int32_t best = 0, bi = 0, bj = 0;
    for (int32_t i = 1; i <= S; i++) {
        int2 v = rs[(int64_t)(i - 1) * 64 + lane];
        if (v.x > best) {
            best = v.x;
            bi = i;
            bj = v.y;
        }
    }
    int32_t bq = 0, bqj = 0;
    if (S == L && L > 0) {
        int2 v = rs[(int64_t)(L - 1) * 64 + lane];
        bq = v.x;
        bqj = v.y;
    }
    int32_t si = bi, sj = bj, bs = best;
    int64_t qc = (int64_t)bq + a.clip_penalty;
    if (qc > INT32_MAX) qc = INT32_MAX;
    if (bq > 0 && qc > best) {
        si = L;
        sj = bqj;
        bs = bq;
    }
    // traceback_interleaved (mod.rs:1667-1716)
    uint32_t *cig = a.cigar + a.cigar_off[job];
    const int32_t cap = a.cigar_cap[job];
    int32_t i = si, j = sj, n = 0, last = -1, state = 0;
    while (i > 0 && j > 0) {
        int32_t c = j - 1;
        uint32_t word = trace[((int64_t)(i - 1) * words + 2 * (c / kStrip) + ((c % kStrip) >> 3)) *
                                  64 + lane];
        uint32_t tr = (word >> (4 * (c & 7))) & 15u;
        int op = -1;
        if (state == 0) {
            uint32_t d = tr & 3u;
            if (d == 0) break;
            if (d == 1) {
                op = 0;
                i--;
                j--;
            } else if (d == 2) {
                state = 1;
            } else {
                state = 2;
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
            if (op == last) {
                if (n - 1 < cap) cig[n - 1] += 16u;
            } else {
                if (n < cap) cig[n] = (1u << 4) | (uint32_t)op;
                n++;
                last = op;
            }
        }
    }
    int32_t m = n < cap ? n : cap;
    for (int32_t x = 0, y = m - 1; x < y; x++, y--) {
        uint32_t tmp = cig[x];
        cig[x] = cig[y];
        cig[y] = tmp;
    }
    int32_t *res = a.res + (int64_t)job * kSwResWords;
    res[0] = bs;
    res[1] = si;
    res[2] = sj;
    res[3] = done ? 1 : 0;
    res[4] = j;
    res[5] = i;
    res[6] = n <= cap ? n : -1;
    res[7] = 0;
    res[8] = S * a.ref_len[job];  // the reference evaluates S full rows
    res[9] = 0;
}

__global__ void group_finish_kernel(GroupLaunch a, int32_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) group_finish_job(a, (int32_t)t);
}

__global__ void group_reset_kernel(int32_t *counter, int32_t *group_stop, int32_t n_groups) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *counter = 0;
    if (t < n_groups) group_stop[t] = 0;
}

}  // namespace

hipError_t launch_group_sw(const GroupLaunch &a, int32_t n_jobs, int32_t n_groups, int n_waves,
                           hipStream_t s) {
    if (n_jobs == 0) return hipSuccess;
    int32_t m = n_groups > 1 ? n_groups : 1;
    group_reset_kernel<<<(m + 255) / 256, 256, 0, s>>>(a.counter, a.group_stop, n_groups);
    if (a.keyed)
        group_dp_kernel<true><<<n_waves, 64, 0, s>>>(a);
    else
        group_dp_kernel<false><<<n_waves, 64, 0, s>>>(a);
    group_stop_kernel<<<(n_jobs + 255) / 256, 256, 0, s>>>(a, n_jobs);
    group_finish_kernel<<<(n_jobs + 255) / 256, 256, 0, s>>>(a, n_jobs);
    return hipGetLastError();
}

}  // namespace kira
