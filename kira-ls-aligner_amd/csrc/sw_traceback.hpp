// Banded SW traceback shared by the lane-per-job and 16-lane kernels.
#pragma once
#include "internal.hpp"

namespace kira {

// Traceback (mod.rs:1581-1626) over a packet's trace slab: lane-per-job
// slabs hold position q of a job at [row][q / 8][lane]; 16-lane jobs
// (P positions a lane) at [row][(q % P) / 8][lane + q / P]; P = 0 selects
// the lane-per-job layout.  lane: the job's column (its first lane).
__host__ __device__ inline void sw_traceback_job(const SwLaunch &a, int32_t job,
                                                 const uint32_t *slab, const int2 *meta,
                                                 int32_t rowdw, int32_t P, int32_t lane) {
    int32_t *res = a.res + (int64_t)job * kSwResWords;
    int32_t i = res[1], j = res[2];
    uint32_t *cig = a.cigar + a.cigar_off[job];
    const int32_t cap = a.cigar_cap[job];
    int32_t n = 0;
    int32_t last_op = -1;
    int state = 0;
    while (i > 0 && j > 0) {
        int2 m = meta[(int64_t)(i - 1) * 64 + lane];
        int32_t base = m.x, js = base + (m.y & 0xFFFF), je = base + (m.y >> 16);
        if (j < js || j > je) break;
        int32_t q = j - base;
        int32_t col = lane;
        if (P > 0) {  // 16-lane job: position q sits in lane q / P at q % P
            col += q / P;
            q -= (q / P) * P;
        }
        uint32_t tr = (slab[((int64_t)(i - 1) * rowdw + (q >> 3)) * 64 + col] >> (4 * (q & 7))) & 15u;
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
            if (op == last_op) {
                if (n - 1 < cap) cig[n - 1] += 16u;
            } else {
                if (n < cap) cig[n] = (1u << 4) | (uint32_t)op;
                n++;
                last_op = op;
            }
        }
    }
    // ops were produced end-to-start: reverse in place
    int32_t m = n < cap ? n : cap;
    for (int32_t x = 0, y = m - 1; x < y; x++, y--) {
        uint32_t tmp = cig[x];
        cig[x] = cig[y];
        cig[y] = tmp;
    }
    res[4] = j;  // ref_start
    res[5] = i;  // read_start
    res[6] = n <= cap ? n : -1;
}

}  // namespace kira
