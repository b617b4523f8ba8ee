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
// Two layouts of the DP.  Systolic (windows up to 1024 columns): a job's
// strips of 16 columns go to adjacent lanes that advance one row a step,
// skewed by one step each, passing boundaries lane to lane; 64 / strips jobs
// a wave, so a batch of a few thousand jobs still fills the chip.  One lane a
// job (wider windows): the lane sweeps its matrix strip after strip, the
// strip's previous-row H/E in registers, its left boundary (H, F per row) in a
// per-packet global slab laid out [row][lane].  Both keep the per-row
// statistics in a [row][slot] slab per packet.
#include "internal.hpp"

namespace kira {
namespace {

constexpr int kStrip = 16;

__host__ __device__ __forceinline__ int32_t gmax(int32_t a, int32_t b) { return a > b ? a : b; }

// One row of one strip: cells (i, j0 .. j0 + kStrip - 1) from the row's read
// byte qb, the strip's reference bytes rb, H[i-1][j0-1] (diag) and the left
// boundary H[i][j0-1] / F[i][j0-1] (hl / fl, which leave as the strip's right
// boundary); Hup / Eup hold the strip's row i-1 and are advanced to row i.
// Out: the 16 four-bit trace codes in w0 / w1 and the strip's row maximum
// with its first column (0, 0 when no cell is positive).
template <bool KEYED>
__host__ __device__ __forceinline__ void strip_row(uint8_t qb, const uint8_t (&rb)[kStrip],
                                                   int32_t j0, int32_t W, int32_t A, int32_t B,
                                                   int32_t go, int32_t ge, int32_t diag,
                                                   int32_t (&Hup)[kStrip],
                                                   int32_t (&Eup)[kStrip], int32_t &hl,
                                                   int32_t &fl, uint32_t &w0, uint32_t &w1,
                                                   int32_t &rmax, int32_t &rarg) {
    int32_t rkey = 0;
    rmax = 0;
    rarg = 0;
    w0 = 0;
    w1 = 0;
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
                // keyed maximum: value << 2 | priority, STOP (the zero
                // candidate) 3 > diagonal 2 > E 1 > F 0, the reference's tie
                // order; direction = 3 - priority
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
}

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
    if (nstrips == 0)  // an empty window: no positive cell in any row
        for (int32_t i = 1; i <= L; i++) rstat[(int64_t)(i - 1) * 64 + lane] = make_int2(0, 0);
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
            const int32_t diag = diag_row_prev;
            diag_row_prev = hl;
            int32_t rmax, rarg;
            uint32_t w0, w1;
            strip_row<KEYED>(qb, rb, j0, W, A, B, go, ge, diag, Hup, Eup, hl, fl, w0, w1, rmax,
                             rarg);
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

// The same cells with each job spread over G lanes of a wave, G = the
// packet's strip count (words / 2), 64 / G jobs a wave: lane t of a job owns
// strip t and computes its row i at step i + t, taking the row's left
// boundary, read byte and the row statistics of strips < t from lane t - 1
// (which computed them at the previous step) by a one-lane shift.  The last
// strip's lane writes the row statistics ([row][slot] per packet, as the
// one-lane form does); the trace is kept per job by step, so the
// G lanes of a job write adjacent words at each step: word
// ((slot * (rows + G - 1) + s - 1) * words + 2 t + half) of the packet.
// A task is (packet, first slot) of 64 / G jobs.
template <bool KEYED>
__global__ __launch_bounds__(64) void group_dp_sys_kernel(GroupLaunch a) {
    const int lane = threadIdx.x;
    for (;;) {
        int32_t task = 0;
        if (lane == 0) task = atomicAdd(a.counter, 1);
        task = __shfl(task, 0);
        if (task >= a.n_tasks) return;
        const int2 tk = a.tasks[task];
        const int32_t p = tk.x;
        const int32_t words = a.packet_words[p], rows = a.packet_rows[p];
        const int32_t G = words >> 1 > 0 ? words >> 1 : 1;
        const int32_t J = 64 / G;
        const int32_t jl = lane / G, t = lane - jl * G, l = tk.y + jl;
        int32_t job = -1;
        if (jl < J && l < 64) job = a.packet_jobs[(int64_t)p * 64 + l];
        int32_t L = 0, W = 0, nst = 0, neg = 0;
        const uint8_t *read = a.reads;
        const int32_t j0 = t * kStrip + 1;
        uint8_t rb[kStrip];
#pragma unroll
        for (int c = 0; c < kStrip; c++) rb[c] = 0;
        if (job >= 0) {
            L = a.read_len[job];
            W = a.ref_len[job];
            nst = (W + kStrip - 1) / kStrip;
            neg = a.neg[a.group_of[job]];
            read = a.reads + a.read_off[job];
            const uint8_t *ref = a.refs + a.ref_off[job];
#pragma unroll
            for (int c = 0; c < kStrip; c++)
                if (j0 + c <= W) rb[c] = ref[j0 - 1 + c];
        }
        int2 *rs = a.rstat + a.packet_row_off[p] + (l < 64 ? l : 0);
        if (job >= 0 && nst == 0 && t == 0)  // an empty window: no positive cell in any row
            for (int32_t i = 1; i <= L; i++) rs[(int64_t)(i - 1) * 64] = make_int2(0, 0);
        const bool live = job >= 0 && t < nst;
        uint32_t *tr = a.trace + a.packet_trace_off[p] +
                       (int64_t)(live ? l : 0) * (rows + G - 1) * words + 2 * t;
        int32_t Hup[kStrip], Eup[kStrip];
#pragma unroll
        for (int c = 0; c < kStrip; c++) {
            Hup[c] = 0;
            Eup[c] = neg;
        }
        int32_t out_h = 0, out_f = neg, out_m = 0, out_a = 0, out_q = 0;
        int32_t diag_next = 0;  // H[i-1][j0-1] for this lane's next row
        // strip 0's lanes read the row bytes two rows ahead
        int32_t q1 = 0, q2 = 0;
        if (live && t == 0) {
            if (L >= 1) q1 = read[0];
            if (L >= 2) q2 = read[1];
        }
        const int32_t steps = rows + G - 1;
        for (int32_t s = 1; s <= steps; s++) {
            const int32_t in_h = __shfl_up(out_h, 1), in_f = __shfl_up(out_f, 1);
            const int32_t in_m = __shfl_up(out_m, 1), in_a = __shfl_up(out_a, 1);
            const int32_t in_q = __shfl_up(out_q, 1);
            const int32_t i = s - t;
            if (live && i >= 1 && i <= L) {
                int32_t hl, fl, qb;
                if (t == 0) {
                    hl = 0;  // column 0 is H = 0, F starts at the sentinel
                    fl = neg;
                    qb = q1;
                    q1 = q2;
                    q2 = i + 2 <= L ? read[i + 1] : 0;
                } else {
                    hl = in_h;
                    fl = in_f;
                    qb = in_q;
                }
                const int32_t diag = diag_next;
                diag_next = hl;
                int32_t rmax, rarg;
                uint32_t w0, w1;
                strip_row<KEYED>((uint8_t)qb, rb, j0, W, a.match_score, a.mismatch, a.go,
                                 a.gap_extend, diag, Hup, Eup, hl, fl, w0, w1, rmax, rarg);
                *(uint2 *)(tr + (int64_t)(s - 1) * words) = make_uint2(w0, w1);
                // strips < t first: a later strip replaces them only when
                // strictly greater (the first strict maximum of the row)
                if (t > 0 && !(rmax > in_m)) {
                    rmax = in_m;
                    rarg = in_a;
                }
                if (t == nst - 1) rs[(int64_t)(i - 1) * 64] = make_int2(rmax, rarg);
                out_h = hl;
                out_f = fl;
                out_m = rmax;
                out_a = rarg;
                out_q = qb;
            }
        }
    }
}

// Row i's statistics of the job in slot `lane` of packet p ([row][slot]).
__host__ __device__ __forceinline__ const int2 *row_stat(const GroupLaunch &a, int32_t p,
                                                        int32_t lane, int32_t i) {
    return a.rstat + a.packet_row_off[p] + (int64_t)(i - 1) * 64 + lane;
}

// First row after which the lane is done (abort bound, mod.rs:2058-2082);
// L when it never is.  The group stops at the maximum over its lanes.
__host__ __device__ __forceinline__ int32_t lane_done_row(const GroupLaunch &a, int32_t job,
                                                           bool &done) {
    const int32_t L = a.read_len[job];
    const int32_t pl = a.job_slot[job];
    const int32_t p = pl >> 6, lane = pl & 63;
    const int32_t ab = a.abort_score[job];
    done = false;
    if (ab <= INT32_MIN / 8) return L;
    int32_t best = 0;
    for (int32_t i = 1; i <= L; i++) {
        best = gmax(best, row_stat(a, p, lane, i)->x);
        if (best + (L - i) * a.match_score < ab) {
            done = true;
            return i;
        }
    }
    return L;
}

// One thread per packet slot (n = packets x 64), so a wave reads each row's
// statistics of 64 adjacent slots at once.
__global__ void group_stop_kernel(GroupLaunch a, int32_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t job = a.packet_jobs[t];
    if (job < 0) return;
    bool done;
    int32_t r = lane_done_row(a, job, done);
    atomicMax(&a.group_stop[a.group_of[job]], r);
}

__host__ __device__ void group_finish_job(const GroupLaunch &a, int32_t job) {
    const int32_t L = a.read_len[job];
    const int32_t pl = a.job_slot[job];
    const int32_t p = pl >> 6, lane = pl & 63;
    const uint32_t *trace = a.trace + a.packet_trace_off[p];
    const int32_t words = a.packet_words[p];
    const int32_t S = a.group_stop[a.group_of[job]];
    bool done;
    int32_t dr = lane_done_row(a, job, done);
    (void)dr;
    // best row (the first of the maximum) over rows 1..S: eight rows' loads
    // are issued before their compares, which keep row order
    int32_t best = 0, bi = 0, bj = 0;
    int32_t r0 = 1;
    for (; r0 + 7 <= S; r0 += 8) {
        int2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = *row_stat(a, p, lane, r0 + u);
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (v[u].x > best) {
                best = v[u].x;
                bi = r0 + u;
                bj = v[u].y;
            }
    }
    for (; r0 <= S; r0++) {
        int2 v = *row_stat(a, p, lane, r0);
        if (v.x > best) {
            best = v.x;
            bi = r0;
            bj = v.y;
        }
    }
    int32_t bq = 0, bqj = 0;
    if (S == L && L > 0) {
        int2 v = *row_stat(a, p, lane, L);
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
    // the systolic layout: this job's words by step (step = row + strip)
    const int32_t G = words >> 1 > 0 ? words >> 1 : 1;
    const uint32_t *jtr = trace + (int64_t)lane * (a.packet_rows[p] + G - 1) * words;
    // the current op's run length stays in a register and is written once,
    // when the op changes or the walk ends (an extension written back to
    // the CIGAR would put a global load and store on every step's path)
    int32_t i = si, j = sj, n = 0, last = -1, state = 0;
    uint32_t run = 0;
    while (i > 0 && j > 0) {
        int32_t c = j - 1;
        const int32_t st = c / kStrip, half = (c % kStrip) >> 3;
        uint32_t word =
            a.systolic ? jtr[(int64_t)(i + st - 1) * words + 2 * st + half]
                       : trace[((int64_t)(i - 1) * words + 2 * st + half) * 64 + lane];
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
                run++;
            } else {
                if (last >= 0 && n - 1 < cap) cig[n - 1] = (run << 4) | (uint32_t)last;
                n++;
                last = op;
                run = 1;
            }
        }
    }
    if (last >= 0 && n - 1 < cap) cig[n - 1] = (run << 4) | (uint32_t)last;
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

// One thread per packet slot, as group_stop_kernel.
__global__ void group_finish_kernel(GroupLaunch a, int32_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t job = a.packet_jobs[t];
    if (job >= 0) group_finish_job(a, job);
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
    if (a.systolic && a.keyed)
        group_dp_sys_kernel<true><<<n_waves, 64, 0, s>>>(a);
    else if (a.systolic)
        group_dp_sys_kernel<false><<<n_waves, 64, 0, s>>>(a);
    else if (a.keyed)
        group_dp_kernel<true><<<n_waves, 64, 0, s>>>(a);
    else
        group_dp_kernel<false><<<n_waves, 64, 0, s>>>(a);
    const int32_t n_slots = a.n_packets * 64;
    group_stop_kernel<<<(n_slots + 255) / 256, 256, 0, s>>>(a, n_slots);
    group_finish_kernel<<<(n_slots + 255) / 256, 256, 0, s>>>(a, n_slots);
    return hipGetLastError();
}

}  // namespace kira
