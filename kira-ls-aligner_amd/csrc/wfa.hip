// Myers bounded edit distance and the scored wavefront aligner on the GPU:
// the CPU fast path of the reference (src/alignment/myers.rs,
// src/alignment/wfa.rs), over batches of (pattern, text) pairs.
//
// Myers: one lane per pair; the pattern's match masks, VP and VN live in
// registers for patterns of up to 512 bases (B = ceil(m / 64) blocks of 64
// bits, one kernel per B), in a per-pair global scratch above that.
//
// WFA: one wave per pair, one lane per diagonal.  A persistent grid of
// single-wave work-groups pulls pairs from an atomic queue; each keeps the
// whole wavefront history of its current pair (M, I and D offsets of every
// score, plus a descriptor per layer) in its own global slab, sized on the
// host for the batch's worst case, so the traceback can walk it afterwards.
// The pair's pattern and text are staged in LDS as 32-bit words; extension
// compares four bytes a step.
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

using kira::DevMem;

// ------------------------------------------------------------------ Myers

// base_to_idx (myers.rs:8-16): A/C/G/T either case, else no match mask
__device__ __forceinline__ int base_idx(uint32_t b) {
    const uint32_t u = b & 0xDFu;
    return u == 'A' ? 0 : u == 'C' ? 1 : u == 'G' ? 2 : u == 'T' ? 3 : -1;
}

struct MyersArgs {
    const uint8_t *pat, *txt;
    const int64_t *pat_off, *txt_off;
    const int32_t *pat_len, *txt_len, *max_k;
    uint8_t *found;
    int32_t *dist, *end;
};

// Trivial pairs (myers.rs:56-64): empty pattern, or empty text.
__device__ __forceinline__ bool myers_trivial(const MyersArgs &a, int32_t k, int32_t m,
                                              int32_t n, int32_t K) {
    if (m == 0) {
        a.found[k] = 1;
        a.dist[k] = 0;
        a.end[k] = 0;
        return true;
    }
    if (n == 0) {
        const bool ok = m <= K;
        a.found[k] = ok;
        a.dist[k] = ok ? m : 0;
        a.end[k] = 0;
        return true;
    }
    return false;
}

// One text column of the multi-block recurrence (myers.rs:243-330): the
// blocks' add carries and horizontal deltas chain from block 0 upwards.
struct ColState {
    uint64_t hin_p, hin_n;
    uint32_t carry;
};

__device__ __forceinline__ void myers_block(uint64_t eq, uint64_t &vp, uint64_t &vn, uint64_t mask,
                                            ColState &cs, uint64_t &hp_out, uint64_t &hn_out) {
    const uint64_t x = eq | vn;
    const uint64_t av = eq & vp;
    const uint64_t s1 = av + vp;
    const uint32_t c1 = s1 < av;
    const uint64_t sum = s1 + cs.carry;
    const uint32_t c2 = sum < s1;
    cs.carry = c1 | c2;
    const uint64_t d0 = (sum ^ vp) | x;
    const uint64_t hp = vn | ~(d0 | vp);
    const uint64_t hn = d0 & vp;
    const uint64_t hps = (hp << 1) | cs.hin_p, hns = (hn << 1) | cs.hin_n;
    cs.hin_p = hp >> 63;
    cs.hin_n = hn >> 63;
    vp = (hns | ~(d0 | hps)) & mask;
    vn = (d0 & hps) & mask;
    hp_out = hp;
    hn_out = hn;
}

// Patterns of 64 (B - 1) + 1 .. 64 B bases, state in registers.
template <int B>
__global__ __launch_bounds__(256) void myers_kernel(int32_t cnt, const int32_t *__restrict__ jobs,
                                                    MyersArgs a) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int32_t k = jobs[t];
    const int32_t m = a.pat_len[k], n = a.txt_len[k], K = a.max_k[k];
    if (myers_trivial(a, k, m, n, K)) return;
    const uint8_t *p = a.pat + a.pat_off[k];
    const uint8_t *tx = a.txt + a.txt_off[k];
    uint64_t pa[B], pc[B], pg[B], pt[B], vp[B], vn[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
        pa[b] = pc[b] = pg[b] = pt[b] = 0;
        const int32_t lim = min(64, m - 64 * b);
#pragma unroll 4
        for (int i = 0; i < lim; i++) {
            const int c = base_idx(p[64 * b + i]);
            const uint64_t bit = 1ull << i;
            pa[b] |= c == 0 ? bit : 0;
            pc[b] |= c == 1 ? bit : 0;
            pg[b] |= c == 2 ? bit : 0;
            pt[b] |= c == 3 ? bit : 0;
        }
    }
    const int32_t top_used = m - (B - 1) * 64;
    const uint64_t top_mask = top_used == 64 ? ~0ull : ((1ull << top_used) - 1);
    const uint64_t high_bit = 1ull << (top_used - 1);
#pragma unroll
    for (int b = 0; b < B; b++) {
        vp[b] = b == B - 1 ? top_mask : ~0ull;
        vn[b] = 0;
    }
    int32_t err = m, best = m, best_pos = 0;
    bool rejected = false;
    for (int32_t j = 0; j < n; j++) {
        const int c = base_idx(tx[j]);
        ColState cs = {0, 0, 0};
        uint64_t hp = 0, hn = 0;
#pragma unroll
        for (int b = 0; b < B; b++) {
            const uint64_t eq = c == 0 ? pa[b] : c == 1 ? pc[b] : c == 2 ? pg[b] : c == 3 ? pt[b] : 0;
            myers_block(eq, vp[b], vn[b], b == B - 1 ? top_mask : ~0ull, cs, hp, hn);
        }
        err += (hp & high_bit) ? 1 : 0;
        err -= (hn & high_bit) ? 1 : 0;
        if (err <= best) {
            best = err;
            best_pos = j + 1;
        }
        // exact early reject (myers.rs:331-345): even a run of matches over
        // the rest of the text cannot bring the score back under max_k
        const int32_t rest = n - 1 - j;
        if (best > K && (err > rest ? err - rest : 0) > K) {
            rejected = true;
            break;
        }
    }
    const bool ok = !rejected && best <= K;
    a.found[k] = ok;
    a.dist[k] = ok ? best : 0;
    a.end[k] = ok ? best_pos : 0;
}

// Patterns above 512 bases: the same recurrence with the masks and vectors
// in a global scratch of 6 words a block (peq A C G T, VP, VN).
__global__ __launch_bounds__(256) void myers_long_kernel(int32_t cnt, const int32_t *__restrict__ jobs,
                                                         const int64_t *__restrict__ scratch_off,
                                                         uint64_t *__restrict__ scratch,
                                                         MyersArgs a) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int32_t k = jobs[t];
    const int32_t m = a.pat_len[k], n = a.txt_len[k], K = a.max_k[k];
    if (myers_trivial(a, k, m, n, K)) return;
    const uint8_t *p = a.pat + a.pat_off[k];
    const uint8_t *tx = a.txt + a.txt_off[k];
    const int32_t blocks = (m + 63) / 64;
    uint64_t *peq = scratch + scratch_off[t];
    uint64_t *vp = peq + 4 * (int64_t)blocks, *vn = vp + blocks;
    for (int32_t b = 0; b < blocks; b++) {
        uint64_t q[4] = {0, 0, 0, 0};
        const int32_t lim = min(64, m - 64 * b);
        for (int i = 0; i < lim; i++) {
            const int c = base_idx(p[64 * b + i]);
            if (c >= 0) q[c] |= 1ull << i;
        }
        for (int c = 0; c < 4; c++) peq[c * (int64_t)blocks + b] = q[c];
        vp[b] = ~0ull;
        vn[b] = 0;
    }
    const int32_t top_used = m - (blocks - 1) * 64;
    const uint64_t top_mask = top_used == 64 ? ~0ull : ((1ull << top_used) - 1);
    const uint64_t high_bit = 1ull << (top_used - 1);
    vp[blocks - 1] &= top_mask;
    int32_t err = m, best = m, best_pos = 0;
    bool rejected = false;
    for (int32_t j = 0; j < n; j++) {
        const int c = base_idx(tx[j]);
        ColState cs = {0, 0, 0};
        uint64_t hp = 0, hn = 0;
        for (int32_t b = 0; b < blocks; b++) {
            const uint64_t eq = c >= 0 ? peq[c * (int64_t)blocks + b] : 0;
            uint64_t vpb = vp[b], vnb = vn[b];
            myers_block(eq, vpb, vnb, b == blocks - 1 ? top_mask : ~0ull, cs, hp, hn);
            vp[b] = vpb;
            vn[b] = vnb;
        }
        err += (hp & high_bit) ? 1 : 0;
        err -= (hn & high_bit) ? 1 : 0;
        if (err <= best) {
            best = err;
            best_pos = j + 1;
        }
        const int32_t rest = n - 1 - j;
        if (best > K && (err > rest ? err - rest : 0) > K) {
            rejected = true;
            break;
        }
    }
    const bool ok = !rejected && best <= K;
    a.found[k] = ok;
    a.dist[k] = ok ? best : 0;
    a.end[k] = ok ? best_pos : 0;
}

constexpr int kMyersRegBlocks = 8;

template <int B>
hipError_t launch_myers(int32_t cnt, const int32_t *jobs, const MyersArgs &a, hipStream_t st) {
    if (cnt == 0) return hipSuccess;
    hipLaunchKernelGGL(myers_kernel<B>, dim3((cnt + 255) / 256), dim3(256), 0, st, cnt, jobs, a);
    return hipGetLastError();
}

hipError_t launch_myers_class(int b, int32_t cnt, const int32_t *jobs, const MyersArgs &a,
                              hipStream_t st) {
    switch (b) {
    case 1: return launch_myers<1>(cnt, jobs, a, st);
    case 2: return launch_myers<2>(cnt, jobs, a, st);
    case 3: return launch_myers<3>(cnt, jobs, a, st);
    case 4: return launch_myers<4>(cnt, jobs, a, st);
    case 5: return launch_myers<5>(cnt, jobs, a, st);
    case 6: return launch_myers<6>(cnt, jobs, a, st);
    case 7: return launch_myers<7>(cnt, jobs, a, st);
    default: return launch_myers<8>(cnt, jobs, a, st);
    }
}

// ------------------------------------------------------------------ WFA

constexpr int32_t kNone = INT32_MIN / 4;  // NONE offset, wfa.rs
constexpr int32_t kWfaMaxLen = 60000;     // pattern + text bytes staged in LDS
constexpr int32_t kWfaMaxScore = 65535;
using kira::kWfaRes;

using Pen = kira::WfaPen;

struct WfaArgs {
    const uint8_t *pat, *txt;
    const int64_t *pat_off, *txt_off;
    const int32_t *pat_len, *txt_len, *max_score, *tbf;
    int32_t drop, want_cigar;
    Pen pen;
    int32_t *res;
    uint32_t *cigar;
    const int64_t *cig_off;
    const int32_t *cig_cap;
    int32_t *slab;
    int64_t slot_ints, val_cap;
    int32_t desc_layers;
    int32_t *queue;
    int32_t n;
};

// Layer descriptor: diagonal k of a non-empty layer (lo <= k <= hi) is at
// values[org + k].
struct Desc {
    int32_t lo, hi, org, pad;
};

__device__ __forceinline__ int32_t desc_get(const int32_t *vals, const Desc &d, int32_t k) {
    return (d.hi >= d.lo && k >= d.lo && k <= d.hi) ? vals[d.org + k] : kNone;
}

__device__ __forceinline__ Desc load_desc(const int4 *descs, int32_t s, int layer) {
    if (s < 0) return Desc{0, -1, 0, 0};
    const int4 v = descs[s * 3 + layer];
    return Desc{__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                __builtin_amdgcn_readfirstlane(v.z), 0};
}

// The forward pass reads the descriptors of the last max(x, o + ie, o + de)
// scores; with a look-back below kDescRing they also live in an LDS ring
// (slot s % kDescRing), so a score step does not wait on a global round trip
// for them.  The global copies stay for the traceback.
constexpr int32_t kDescRing = 32;

// Four bytes of an LDS byte string at byte position pos (two aligned words).
__device__ __forceinline__ uint32_t ld4(const uint32_t *w, int32_t pos) {
    const int32_t q = pos >> 2;
    const uint64_t pair = ((uint64_t)w[q + 1] << 32) | w[q];
    return (uint32_t)(pair >> (8 * (pos & 3)));
}

__device__ __forceinline__ uint32_t ld1(const uint32_t *w, int32_t pos) {
    return (w[pos >> 2] >> ((pos & 3) * 8)) & 0xFFu;
}

// extend_diagonal (wfa.rs:207-224): matching bytes from pattern position off
// and text position off + k, never past either end.
__device__ __forceinline__ int32_t extend_diag(const uint32_t *P, int32_t m, const uint32_t *T,
                                               int32_t n, int32_t k, int32_t off) {
    const int32_t j = off + k;
    if (off >= m || j < 0 || j >= n) return off;
    const int32_t lim = min(m - off, n - j);
    int32_t e = 0;
    while (e < lim) {
        const uint32_t x = ld4(P, off + e) ^ ld4(T, j + e);
        if (x) {
            e += __builtin_ctz(x) >> 3;
            break;
        }
        e += 4;
    }
    return off + min(e, lim);
}

__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Matching bytes going back from pattern position i_end - 1 and text position
// j_end - 1, at most lim (<= both positions), four a compare.
__device__ __forceinline__ int32_t back_match(const uint32_t *P, int32_t i_end,
                                              const uint32_t *T, int32_t j_end, int32_t lim) {
    int32_t e = 0;
    while (e + 4 <= lim) {
        const uint32_t x = ld4(P, i_end - 4 - e) ^ ld4(T, j_end - 4 - e);
        if (x) return e + (__builtin_clz(x) >> 3);
        e += 4;
    }
    while (e < lim && ld1(P, i_end - 1 - e) == ld1(T, j_end - 1 - e)) e++;
    return e;
}

// Appends one op to a CIGAR written backwards from the top of its slot
// (the traceback emits ops last to first); an op of the kind just written
// lengthens it (wfa.rs:954-966).
struct RevCigar {
    uint32_t *top;  // one past the slot's last word
    int32_t cap, cnt;
    bool overflow;
    __device__ void push(uint32_t op, uint32_t len) {
        if (cnt > 0 && (top[-cnt] & 15u) == op) {
            top[-cnt] += len << 4;
            return;
        }
        if (cnt == cap) {
            overflow = true;
            return;
        }
        cnt++;
        top[-cnt] = (len << 4) | op;
    }
};

// build_cigar (wfa.rs:762-952) by one lane: from (s, k_end, off = m) in the
// M layer back to (0, k0, 0).  Returns k0 (the start diagonal).
__device__ int32_t wfa_traceback(const uint32_t *P, int32_t m, const uint32_t *T, int32_t n,
                                 const int32_t *vals, const int4 *descs, int32_t s, int32_t k,
                                 const Pen pen, RevCigar &cg) {
    auto get = [&](int layer, int32_t ss, int32_t kk) -> int32_t {
        if (ss < 0) return kNone;
        const int4 d = descs[ss * 3 + layer];
        return (d.y >= d.x && kk >= d.x && kk <= d.y) ? vals[d.z + kk] : kNone;
    };
    int32_t off = m;
    int layer = 0;
    while (!(s == 0 && layer == 0 && off == 0)) {
        if (cg.overflow) break;
        if (layer == 0) {
            const int32_t iv = get(1, s, k), dv = get(2, s, k);
            const int32_t pv = s >= pen.x ? get(0, s - pen.x, k) : kNone;
            const int32_t mv = pv != kNone ? pv + 1 : kNone;
            // the walk back over matches stops at the highest of: 0, the
            // text start (walk = -k), an offset of I, D or the mismatch source
            // at or below off, or just above the first mismatch going back;
            // cells always satisfy off <= m and off + k <= n
            int32_t lo = max(0, -k);
            if (iv != kNone && iv <= off && iv > lo) lo = iv;
            if (dv != kNone && dv <= off && dv > lo) lo = dv;
            if (mv != kNone && mv <= off && mv > lo) lo = mv;
            const int32_t walk = off - back_match(P, off, T, off + k, off - lo);
            if (off - walk > 0) cg.push(0, (uint32_t)(off - walk));
            off = walk;
            if (s == 0 && off == 0) break;
            if (mv != kNone && mv == off) {
                cg.push(0, 1);
                off--;
                s -= pen.x;
            } else if (iv == off) {
                layer = 1;
            } else if (dv == off) {
                layer = 2;
            } else {
                break;
            }
        } else if (layer == 1) {
            const int32_t so = s - pen.o - pen.ie, se = s - pen.ie;
            const int32_t po = so >= 0 ? get(0, so, k + 1) : kNone;
            const int32_t pe = se >= 0 ? get(1, se, k + 1) : kNone;
            if (po != kNone && po + 1 == off) {
                cg.push(1, 1);
                off--;
                k++;
                s = so;
                layer = 0;
            } else if (pe != kNone && pe + 1 == off) {
                cg.push(1, 1);
                off--;
                k++;
                s = se;
            } else {
                break;
            }
        } else {
            const int32_t so = s - pen.o - pen.de, se = s - pen.de;
            const int32_t po = so >= 0 ? get(0, so, k - 1) : kNone;
            const int32_t pe = se >= 0 ? get(2, se, k - 1) : kNone;
            if (po != kNone && po == off) {
                cg.push(2, 1);
                k--;
                s = so;
                layer = 0;
            } else if (pe != kNone && pe == off) {
                cg.push(2, 1);
                k--;
                s = se;
            } else {
                break;
            }
        }
    }
    return k;
}

// wfa_align_with_costs (wfa.rs:549-648) for one pair per work-group
// iteration.  The score-s step (wfa.rs:374-525) runs over the M range 64
// diagonals a pass: I from M[s - o - ie] / I[s - ie] at k + 1, D from
// M[s - o - de] / D[s - de] at k - 1, M from M[s - x] at k, I and D, then
// extension, the optional pruning, and the done check.
__global__ __launch_bounds__(64) void wfa_kernel(WfaArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ int32_t job_sh;
    __shared__ int4 dring[kDescRing * 3];
    const int lane = threadIdx.x;
    int32_t *vals = a.slab + (int64_t)blockIdx.x * a.slot_ints;
    int4 *descs = (int4 *)(vals + a.val_cap);
    const Pen pen = a.pen;
    const bool ring = max(pen.x, max(pen.o + pen.ie, pen.o + pen.de)) < kDescRing;
    auto desc_at = [&](int32_t ss, int layer) -> Desc {
        if (!ring) return load_desc(descs, ss, layer);
        if (ss < 0) return Desc{0, -1, 0, 0};
        const int4 v = dring[(ss & (kDescRing - 1)) * 3 + layer];
        return Desc{__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                    __builtin_amdgcn_readfirstlane(v.z), 0};
    };
    for (;;) {
        __syncthreads();
        if (lane == 0) job_sh = atomicAdd(a.queue, 1);
        __syncthreads();
        const int32_t job = job_sh;
        if (job >= a.n) break;
        int32_t *res = a.res + (int64_t)job * kWfaRes;
        const int32_t m = a.pat_len[job], n = a.txt_len[job];
        const int32_t S = a.max_score[job];
        if (m == 0) {
            if (lane == 0) {
                res[0] = 1;
                res[1] = res[2] = res[3] = res[4] = res[5] = 0;
            }
            continue;
        }
        int32_t tbf = max(a.tbf[job], 0);
        tbf = min(tbf, max(n - 1, 0));
        // stage pattern and text as words (zero padded by two words)
        const int32_t pw = (m + 3) / 4 + 2, tw = (n + 3) / 4 + 2;
        uint32_t *P = lds, *T = lds + pw;
        {
            const uint8_t *p = a.pat + a.pat_off[job], *t = a.txt + a.txt_off[job];
            for (int32_t w = lane; w < pw + tw; w += 64) {
                const bool isp = w < pw;
                const uint8_t *src = isp ? p : t;
                const int32_t len = isp ? m : n;
                const int32_t b0 = (isp ? w : w - pw) * 4;
                uint32_t v = 0;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (b0 + i < len) v |= (uint32_t)src[b0 + i] << (8 * i);
                lds[w] = v;
            }
        }
        __syncthreads();
        int64_t cursor = 0;
        bool overflow = false;
        int32_t k_done = kNone, s_done = -1;
        for (int32_t s = 0; s <= max(S, 0) && s_done < 0; s++) {
            Desc dm, di, dd;
            Desc mx, moi, mod, ie, de;
            if (s == 0) {
                dm = Desc{0, tbf, 0, 0};
                di = dd = Desc{0, -1, 0, 0};
            } else {
                mx = desc_at(s - pen.x, 0);
                moi = desc_at(s - pen.o - pen.ie, 0);
                mod = desc_at(s - pen.o - pen.de, 0);
                ie = desc_at(s - pen.ie, 1);
                de = desc_at(s - pen.de, 2);
                int32_t ilo = INT32_MAX, ihi = INT32_MIN, dlo = INT32_MAX, dhi = INT32_MIN;
                int32_t mlo = INT32_MAX, mhi = INT32_MIN;
                if (moi.hi >= moi.lo) { ilo = min(ilo, moi.lo - 1); ihi = max(ihi, moi.hi - 1); }
                if (ie.hi >= ie.lo) { ilo = min(ilo, ie.lo - 1); ihi = max(ihi, ie.hi - 1); }
                if (mod.hi >= mod.lo) { dlo = min(dlo, mod.lo + 1); dhi = max(dhi, mod.hi + 1); }
                if (de.hi >= de.lo) { dlo = min(dlo, de.lo + 1); dhi = max(dhi, de.hi + 1); }
                if (mx.hi >= mx.lo) { mlo = min(mlo, mx.lo); mhi = max(mhi, mx.hi); }
                if (ilo <= ihi) { mlo = min(mlo, ilo); mhi = max(mhi, ihi); }
                if (dlo <= dhi) { mlo = min(mlo, dlo); mhi = max(mhi, dhi); }
                // diagonals outside [-m, n] only ever hold NONE
                di = Desc{max(ilo, -m), min(ihi, n), 0, 0};
                dd = Desc{max(dlo, -m), min(dhi, n), 0, 0};
                dm = Desc{max(mlo, -m), min(mhi, n), 0, 0};
            }
            // allocate the three layers from the slot
            Desc *lay[3] = {&dm, &di, &dd};
            for (int l = 0; l < 3; l++) {
                Desc &d = *lay[l];
                if (d.hi < d.lo) {
                    d = Desc{0, -1, 0, 0};
                    continue;
                }
                const int64_t w = (int64_t)d.hi - d.lo + 1;
                if (cursor + w > a.val_cap) {
                    overflow = true;
                    break;
                }
                d.org = (int32_t)(cursor - d.lo);
                cursor += w;
            }
            if (overflow) break;
            int32_t first_done = INT32_MAX, amax = INT32_MIN;
            if (dm.hi >= dm.lo) {
                for (int32_t base = dm.lo; base <= dm.hi; base += 64) {
                    const int32_t k = base + lane;
                    if (k > dm.hi) continue;
                    int32_t mv = kNone;
                    if (s == 0) {
                        mv = 0;
                    } else {
                        int32_t iv = kNone, dv = kNone;
                        if (k >= di.lo && k <= di.hi) {
                            int32_t v = desc_get(vals, moi, k + 1);
                            if (v != kNone && v < m && v + 1 > iv) iv = v + 1;
                            v = desc_get(vals, ie, k + 1);
                            if (v != kNone && v < m && v + 1 > iv) iv = v + 1;
                            vals[di.org + k] = iv;
                        }
                        if (k >= dd.lo && k <= dd.hi) {
                            int32_t v = desc_get(vals, mod, k - 1);
                            if (v != kNone && v + k <= n && v > dv) dv = v;
                            v = desc_get(vals, de, k - 1);
                            if (v != kNone && v + k <= n && v > dv) dv = v;
                            vals[dd.org + k] = dv;
                        }
                        const int32_t v = desc_get(vals, mx, k);
                        if (v != kNone && v < m && v + k < n && v + 1 > mv) mv = v + 1;
                        if (iv != kNone && iv > mv) mv = iv;
                        if (dv != kNone && dv > mv) mv = dv;
                    }
                    if (mv != kNone) {
                        mv = extend_diag(P, m, T, n, k, mv);
                        amax = max(amax, 2 * mv + k);
                        if (mv >= m) first_done = min(first_done, k);
                    }
                    vals[dm.org + k] = mv;
                }
            }
            // prune_wavefront (wfa.rs:293-346)
            if (a.drop > 0 && dm.hi >= dm.lo) {
                const int32_t max_ad = wave_max(amax);
                if (max_ad != INT32_MIN) {
                    const int64_t th = (int64_t)max_ad - a.drop;
                    int32_t nlo = INT32_MAX, nhi = INT32_MIN;
                    for (int32_t base = dm.lo; base <= dm.hi; base += 64) {
                        const int32_t k = base + lane;
                        if (k > dm.hi) continue;
                        const int32_t v = vals[dm.org + k];
                        if (v != kNone && (int64_t)2 * v + k >= th) {
                            nlo = min(nlo, k);
                            nhi = max(nhi, k);
                        }
                    }
                    nlo = -wave_max(-nlo);
                    nhi = wave_max(nhi);
                    first_done = INT32_MAX;
                    if (nlo > nhi) {
                        dm = Desc{0, -1, 0, 0};
                    } else {
                        // the same diagonal-to-lane mapping as the pass that
                        // stored the values: each lane rereads its own
                        for (int32_t base = dm.lo; base <= dm.hi; base += 64) {
                            const int32_t k = base + lane;
                            if (k < nlo || k > nhi) continue;
                            const int32_t v = vals[dm.org + k];
                            if (v == kNone) continue;
                            if ((int64_t)2 * v + k < th)
                                vals[dm.org + k] = kNone;
                            else if (v >= m)
                                first_done = min(first_done, k);
                        }
                        dm.lo = nlo;
                        dm.hi = nhi;
                    }
                }
            }
            if (lane == 0) {
                descs[s * 3 + 0] = make_int4(dm.lo, dm.hi, dm.org, 0);
                descs[s * 3 + 1] = make_int4(di.lo, di.hi, di.org, 0);
                descs[s * 3 + 2] = make_int4(dd.lo, dd.hi, dd.org, 0);
                if (ring) {
                    int4 *r = dring + (s & (kDescRing - 1)) * 3;
                    r[0] = make_int4(dm.lo, dm.hi, dm.org, 0);
                    r[1] = make_int4(di.lo, di.hi, di.org, 0);
                    r[2] = make_int4(dd.lo, dd.hi, dd.org, 0);
                }
            }
            // check_done (wfa.rs:348-357): the first diagonal from lo
            int32_t fd = INT32_MAX;
            if (__ballot(first_done != INT32_MAX)) fd = -wave_max(-first_done);
            __syncthreads();
            if (fd != INT32_MAX) {
                k_done = fd;
                s_done = s;
            }
        }
        if (lane == 0) {
            res[5] = overflow ? 1 : 0;
            if (overflow || s_done < 0) {
                res[0] = 0;
                res[1] = res[2] = res[3] = res[4] = 0;
            } else {
                res[0] = 1;
                res[1] = s_done;
                res[3] = m + k_done;
                res[2] = 0;
                res[4] = 0;
                if (a.want_cigar) {
                    const int32_t cap = a.cig_cap[job];
                    uint32_t *slot = a.cigar + a.cig_off[job];
                    RevCigar cg{slot + cap, cap, 0, false};
                    const int32_t k0 = wfa_traceback(P, m, T, n, vals, descs, s_done, k_done, pen, cg);
                    if (cg.overflow) {
                        res[5] = 2;
                    } else {
                        for (int32_t i = 0; i < cg.cnt; i++) slot[i] = slot[cap - cg.cnt + i];
                        res[4] = cg.cnt;
                        res[2] = k0 > 0 ? k0 : 0;
                    }
                }
            }
        }
    }
}

__global__ void wfa_compact_kernel(int32_t n, const int32_t *__restrict__ res,
                                   const uint32_t *__restrict__ src,
                                   const int64_t *__restrict__ src_off,
                                   const int64_t *__restrict__ dst_off,
                                   uint32_t *__restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t c = res[k * kWfaRes + 4];
    const uint32_t *s = src + src_off[k];
    uint32_t *d = dst + dst_off[k];
    for (int32_t i = 0; i < c; i++) d[i] = s[i];
}

// Host mirror of the kernel's range recurrence without pruning, for the
// batch's largest pattern, text, free lead and score: the value words a
// slot can need (pruning only narrows ranges).
int64_t wfa_slot_values(const Pen &pen, int32_t S, int32_t m, int32_t n, int32_t tbf) {
    std::vector<int32_t> lo(3 * (size_t)(S + 1)), hi(3 * (size_t)(S + 1));
    auto get = [&](int32_t s, int l, int32_t &a, int32_t &b) {
        if (s < 0) return false;
        a = lo[3 * (size_t)s + l];
        b = hi[3 * (size_t)s + l];
        return b >= a;
    };
    int64_t total = 0;
    for (int32_t s = 0; s <= S; s++) {
        int32_t ml, mh, il = 0, ih = -1, dl = 0, dh = -1;
        if (s == 0) {
            ml = 0;
            mh = tbf;
        } else {
            int32_t ilo = INT32_MAX, ihi = INT32_MIN, dlo = INT32_MAX, dhi = INT32_MIN;
            int32_t mlo = INT32_MAX, mhi = INT32_MIN, a, b;
            if (get(s - pen.o - pen.ie, 0, a, b)) { ilo = std::min(ilo, a - 1); ihi = std::max(ihi, b - 1); }
            if (get(s - pen.ie, 1, a, b)) { ilo = std::min(ilo, a - 1); ihi = std::max(ihi, b - 1); }
            if (get(s - pen.o - pen.de, 0, a, b)) { dlo = std::min(dlo, a + 1); dhi = std::max(dhi, b + 1); }
            if (get(s - pen.de, 2, a, b)) { dlo = std::min(dlo, a + 1); dhi = std::max(dhi, b + 1); }
            if (get(s - pen.x, 0, a, b)) { mlo = std::min(mlo, a); mhi = std::max(mhi, b); }
            if (ilo <= ihi) { mlo = std::min(mlo, ilo); mhi = std::max(mhi, ihi); }
            if (dlo <= dhi) { mlo = std::min(mlo, dlo); mhi = std::max(mhi, dhi); }
            il = std::max(ilo, -m);
            ih = std::min(ihi, n);
            dl = std::max(dlo, -m);
            dh = std::min(dhi, n);
            ml = std::max(mlo, -m);
            mh = std::min(mhi, n);
        }
        const int32_t L[3][2] = {{ml, mh}, {il, ih}, {dl, dh}};
        for (int l = 0; l < 3; l++) {
            lo[3 * (size_t)s + l] = L[l][0];
            hi[3 * (size_t)s + l] = L[l][1];
            if (L[l][1] >= L[l][0]) total += (int64_t)L[l][1] - L[l][0] + 1;
        }
    }
    return total;
}

// The history slab stays with the context between calls (the last of its
// scratch slots; kira_release_buffers frees it), regrown to the need.
constexpr int kSlabSlot = 15;
hipError_t slab_for(kira_ctx *ctx, size_t need, int32_t **p) {
    if (ctx->cap[kSlabSlot] < need) {
        if (ctx->buf[kSlabSlot]) {
            hipError_t e = hipFree(ctx->buf[kSlabSlot]);
            ctx->buf[kSlabSlot] = nullptr;
            ctx->cap[kSlabSlot] = 0;
            if (e != hipSuccess) return e;
        }
        // a quarter more than asked, so batches of slowly growing shapes do
        // not free and allocate (each hipFree waits for the device) every time
        need += need / 4;
        hipError_t e = hipMalloc(&ctx->buf[kSlabSlot], need);
        if (e != hipSuccess) {
            ctx->buf[kSlabSlot] = nullptr;
            return e;
        }
        ctx->cap[kSlabSlot] = need;
    }
    *p = (int32_t *)ctx->buf[kSlabSlot];
    return hipSuccess;
}

bool check_ranges(int32_t n, const int64_t *off, const int32_t *len, int64_t bytes) {
    for (int32_t k = 0; k < n; k++)
        if (off[k] < 0 || len[k] < 0 || off[k] > bytes || len[k] > bytes - off[k]) return false;
    return true;
}

}  // namespace

namespace kira {

bool wfa_pen(const kira_sw_cfg &cfg, WfaPen &pen) {
    const int64_t x = (int64_t)cfg.match_score + cfg.mismatch;
    const int64_t ie = (int64_t)cfg.match_score + cfg.gap_extend;
    if (cfg.match_score <= 0 || cfg.mismatch <= 0 || cfg.gap_open < 0 || cfg.gap_extend <= 0 ||
        x > INT32_MAX || ie > INT32_MAX)
        return false;
    pen = WfaPen{(int32_t)x, cfg.gap_open, (int32_t)ie, cfg.gap_extend};
    return true;
}

kira_status myers_dev(kira_ctx *ctx, int32_t n, const uint8_t *pat, const int64_t *pat_off,
                      const int32_t *pat_len, const int32_t *h_pat_len, const uint8_t *txt,
                      const int64_t *txt_off, const int32_t *txt_len, const int32_t *max_k,
                      uint8_t *found, int32_t *dist, int32_t *end) {
    if (n == 0) return KIRA_OK;
    // classes by block count; 0 = above 512 bases
    std::vector<int32_t> order, cls_start(kMyersRegBlocks + 2, 0), long_off;
    try {
        order.resize(n);
        std::vector<int32_t> cls(n);
        for (int32_t k = 0; k < n; k++) {
            const int32_t b = (h_pat_len[k] + 63) / 64;
            cls[k] = b > kMyersRegBlocks ? 0 : std::max(b, 1);
            cls_start[cls[k] + 1]++;
        }
        for (int c = 0; c <= kMyersRegBlocks; c++) cls_start[c + 1] += cls_start[c];
        std::vector<int32_t> pos(cls_start.begin(), cls_start.end() - 1);
        for (int32_t k = 0; k < n; k++) order[pos[cls[k]]++] = k;
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    const int32_t n_long = cls_start[1];
    std::vector<int64_t> loff(n_long + 1, 0);
    for (int32_t t = 0; t < n_long; t++)
        loff[t + 1] = loff[t] + 6 * (int64_t)((h_pat_len[order[t]] + 63) / 64);
    DevMem dm(ctx);
    int32_t *d_order = nullptr;
    int64_t *d_loff = nullptr;
    uint64_t *d_scratch = nullptr;
    hipStream_t st = ctx->stream;
    hipError_t e;
    if ((e = dm.alloc(&d_order, n)) || (e = dm.alloc(&d_loff, n_long + 1)) ||
        (e = dm.alloc(&d_scratch, loff[n_long])))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "Myers allocation", e);
    if ((e = hipMemcpyAsync(d_order, order.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_loff, loff.data(), 8 * (size_t)(n_long + 1), hipMemcpyHostToDevice,
                            st)))
        return fail(ctx, KIRA_ERR_DRIVER, "Myers upload", e);
    MyersArgs a{pat, txt, pat_off, txt_off, pat_len, txt_len, max_k, found, dist, end};
    if (n_long) {
        hipLaunchKernelGGL(myers_long_kernel, dim3((n_long + 255) / 256), dim3(256), 0, st, n_long,
                           d_order, d_loff, d_scratch, a);
        if ((e = hipGetLastError())) return fail(ctx, KIRA_ERR_RUNTIME, "Myers launch", e);
    }
    for (int b = 1; b <= kMyersRegBlocks; b++) {
        const int32_t c0 = cls_start[b], cnt = cls_start[b + 1] - c0;
        if ((e = launch_myers_class(b, cnt, d_order + c0, a, st)))
            return fail(ctx, KIRA_ERR_RUNTIME, "Myers launch", e);
    }
    if ((e = hipStreamSynchronize(st))) return fail(ctx, KIRA_ERR_RUNTIME, "Myers", e);
    return KIRA_OK;
}

kira_status WfaRun::launch(const WfaDevJobs &j, const WfaPen &pen, int32_t drop, bool want) {
    n = j.n;
    want_cigar = want;
    if (n == 0) return KIRA_OK;
    int32_t m_max = 0, n_max = 0, tbf_max = 0, s_max = 0;
    for (int32_t k = 0; k < n; k++) {
        if (j.h_pat_len[k] + (int64_t)j.h_txt_len[k] > kWfaMaxLen)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "WFA pair longer than 60000 bytes");
        if (j.h_max_score[k] > kWfaMaxScore)
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "WFA max_score above 65535");
        m_max = std::max(m_max, j.h_pat_len[k]);
        n_max = std::max(n_max, j.h_txt_len[k]);
        const int32_t t = std::min(std::max(j.h_tbf[k], 0), std::max(j.h_txt_len[k] - 1, 0));
        tbf_max = std::max(tbf_max, t);
        s_max = std::max(s_max, j.h_max_score[k]);
    }
    // CIGAR slots: every op has length >= 1 (<= m + n ops), and op runs
    // alternate between M and gaps costing >= o + min(ie, de) each
    int64_t cig_total = 0;
    try {
        cig_off.resize(n);
        cig_cap.resize(n);
        const int64_t gap_min = (int64_t)pen.o + std::min(pen.ie, pen.de);
        for (int32_t k = 0; k < n; k++) {
            int64_t c = 0;
            if (want_cigar) {
                c = (int64_t)j.h_pat_len[k] + j.h_txt_len[k];
                c = std::min<int64_t>(c, 2 * (std::max(j.h_max_score[k], 0) / gap_min) + 1) + 1;
            }
            cig_off[k] = cig_total;
            cig_cap[k] = (int32_t)c;
            cig_total += c;
        }
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    trace_mark(".wfa_loops");
    int64_t val_cap = wfa_slot_values(pen, s_max, m_max, n_max, tbf_max);
    val_cap = (val_cap + 3) & ~(int64_t)3;
    const int64_t slot_ints = val_cap + 12 * (int64_t)(s_max + 1);
    const int64_t slot_bytes = 4 * slot_ints;
    // one slot per resident wave: up to 32 single-wave work-groups on each
    // of the 256 CUs, within a 4 GiB slab
    const int64_t kSlabBudget = (int64_t)4 << 30;
    int64_t slots = std::min<int64_t>({(int64_t)n, 8192, kSlabBudget / slot_bytes});
    if (slots < 1) {
        if (slot_bytes > ((int64_t)16 << 30))
            return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "WFA history of one pair above 16 GiB");
        slots = 1;
    }
    int32_t *d_ccap = nullptr, *d_slab = nullptr, *d_queue = nullptr;
    hipStream_t st = ctx->stream;
    hipError_t e;
    if ((e = dm.alloc(&d_res, (size_t)n * kWfaRes)) || (e = dm.alloc(&d_coff, n)) ||
        (e = dm.alloc(&d_ccap, n)) || (e = dm.alloc(&d_cig, cig_total)) ||
        (e = slab_for(ctx, 4 * (size_t)(slots * slot_ints), &d_slab)) ||
        (e = dm.alloc(&d_queue, 1)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "WFA allocation", e);
    if ((e = hipMemcpyAsync(d_coff, cig_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_ccap, cig_cap.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemsetAsync(d_queue, 0, sizeof(int32_t), st)))
        return fail(ctx, KIRA_ERR_DRIVER, "WFA upload", e);
    WfaArgs a{};
    a.pat = j.pat;
    a.txt = j.txt;
    a.pat_off = j.pat_off;
    a.txt_off = j.txt_off;
    a.pat_len = j.pat_len;
    a.txt_len = j.txt_len;
    a.max_score = j.max_score;
    a.tbf = j.tbf;
    a.drop = drop > 0 ? drop : 0;
    a.want_cigar = want_cigar ? 1 : 0;
    a.pen = pen;
    a.res = d_res;
    a.cigar = d_cig;
    a.cig_off = d_coff;
    a.cig_cap = d_ccap;
    a.slab = d_slab;
    a.slot_ints = slot_ints;
    a.val_cap = val_cap;
    a.desc_layers = s_max + 1;
    a.queue = d_queue;
    a.n = n;
    const size_t lds = 4 * (size_t)((m_max + 3) / 4 + 2 + (n_max + 3) / 4 + 2);
    trace_mark(".wfa_prep");
    hipLaunchKernelGGL(wfa_kernel, dim3((uint32_t)slots), dim3(64), lds, st, a);
    if ((e = hipGetLastError())) return fail(ctx, KIRA_ERR_RUNTIME, "WFA launch", e);
    return KIRA_OK;
}

kira_status WfaRun::finish(std::vector<int32_t> &res, std::vector<uint32_t> &cigar,
                           std::vector<int64_t> &cig_out) {
    try {
        res.assign((size_t)n * kWfaRes, 0);
        cig_out.assign(n, 0);
        cigar.clear();
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    if (n == 0) return KIRA_OK;
    hipStream_t st = ctx->stream;
    hipError_t e;
    if ((e = hipMemcpyAsync(res.data(), d_res, 4 * res.size(), hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return fail(ctx, KIRA_ERR_RUNTIME, "WFA", e);
    trace_mark(".wfa_sync");
    int64_t used = 0;
    std::vector<int64_t> dst(n);
    for (int32_t k = 0; k < n; k++) {
        const int32_t *r = &res[(size_t)k * kWfaRes];
        if (r[5] != 0)
            return fail(ctx, KIRA_ERR_RUNTIME,
                        r[5] == 1 ? "WFA history slot overflow" : "WFA CIGAR slot overflow");
        dst[k] = cig_out[k] = used;
        used += r[4];
    }
    if (!want_cigar || used == 0) return KIRA_OK;
    try {
        cigar.resize(used);
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    int64_t *d_dst = nullptr;
    uint32_t *d_cig2 = nullptr;
    if ((e = dm.alloc(&d_dst, n)) || (e = dm.alloc(&d_cig2, used)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "WFA allocation", e);
    if ((e = hipMemcpyAsync(d_dst, dst.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st)))
        return fail(ctx, KIRA_ERR_DRIVER, "WFA upload", e);
    hipLaunchKernelGGL(wfa_compact_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, d_res, d_cig,
                       d_coff, d_dst, d_cig2);
    if ((e = hipGetLastError())) return fail(ctx, KIRA_ERR_RUNTIME, "WFA compact launch", e);
    if ((e = hipMemcpyAsync(cigar.data(), d_cig2, 4 * (size_t)used, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return fail(ctx, KIRA_ERR_RUNTIME, "WFA", e);
    return KIRA_OK;
}

kira_status wfa_dev(kira_ctx *ctx, const WfaDevJobs &j, const WfaPen &pen, int32_t drop,
                    bool want_cigar, std::vector<int32_t> &res, std::vector<uint32_t> &cigar,
                    std::vector<int64_t> &cig_out) {
    WfaRun run(ctx);
    kira_status st = run.launch(j, pen, drop, want_cigar);
    if (st != KIRA_OK) return st;
    return run.finish(res, cigar, cig_out);
}

}  // namespace kira

extern "C" kira_status kira_myers_batch(kira_ctx *ctx, int32_t n, const uint8_t *patterns,
                                        int64_t patterns_bytes, const int64_t *pat_off,
                                        const int32_t *pat_len, const uint8_t *texts,
                                        int64_t texts_bytes, const int64_t *text_off,
                                        const int32_t *text_len, const int32_t *max_k,
                                        uint8_t *found, int32_t *dist, int32_t *end) {
    if (!ctx || n < 0 || patterns_bytes < 0 || texts_bytes < 0)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (n == 0) return KIRA_OK;
    if ((patterns_bytes && !patterns) || (texts_bytes && !texts) || !pat_off || !pat_len ||
        !text_off || !text_len || !max_k || !found || !dist || !end)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (!check_ranges(n, pat_off, pat_len, patterns_bytes) ||
        !check_ranges(n, text_off, text_len, texts_bytes))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "pattern or text range outside its buffer");
    for (int32_t k = 0; k < n; k++)
        if (max_k[k] < 0) return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "max_k < 0");
    hipError_t e = hipSetDevice(ctx->device);
    if (e) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    DevMem dm(ctx);
    uint8_t *d_pat = nullptr, *d_txt = nullptr, *d_found = nullptr;
    int64_t *d_poff = nullptr, *d_toff = nullptr;
    int32_t *d_plen = nullptr, *d_tlen = nullptr, *d_k = nullptr, *d_dist = nullptr,
            *d_end = nullptr;
    hipStream_t st = ctx->stream;
    if ((e = dm.alloc(&d_pat, patterns_bytes)) || (e = dm.alloc(&d_txt, texts_bytes)) ||
        (e = dm.alloc(&d_poff, n)) || (e = dm.alloc(&d_toff, n)) || (e = dm.alloc(&d_plen, n)) ||
        (e = dm.alloc(&d_tlen, n)) || (e = dm.alloc(&d_k, n)) || (e = dm.alloc(&d_found, n)) ||
        (e = dm.alloc(&d_dist, n)) || (e = dm.alloc(&d_end, n)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "Myers batch allocation", e);
    if ((patterns_bytes &&
         (e = hipMemcpyAsync(d_pat, patterns, patterns_bytes, hipMemcpyHostToDevice, st))) ||
        (texts_bytes && (e = hipMemcpyAsync(d_txt, texts, texts_bytes, hipMemcpyHostToDevice, st))) ||
        (e = hipMemcpyAsync(d_poff, pat_off, 8 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_toff, text_off, 8 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_plen, pat_len, 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_tlen, text_len, 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_k, max_k, 4 * (size_t)n, hipMemcpyHostToDevice, st)))
        return fail(ctx, KIRA_ERR_DRIVER, "Myers batch upload", e);
    kira_status rs = kira::myers_dev(ctx, n, d_pat, d_poff, d_plen, pat_len, d_txt, d_toff, d_tlen,
                                     d_k, d_found, d_dist, d_end);
    if (rs != KIRA_OK) return rs;
    if ((e = hipMemcpyAsync(found, d_found, (size_t)n, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(dist, d_dist, 4 * (size_t)n, hipMemcpyDeviceToHost, st)) ||
        (e = hipMemcpyAsync(end, d_end, 4 * (size_t)n, hipMemcpyDeviceToHost, st)) ||
        (e = hipStreamSynchronize(st)))
        return fail(ctx, KIRA_ERR_RUNTIME, "Myers batch", e);
    return KIRA_OK;
}

extern "C" kira_status kira_wfa_batch(kira_ctx *ctx, int32_t n, const uint8_t *patterns,
                                      int64_t patterns_bytes, const int64_t *pat_off,
                                      const int32_t *pat_len, const uint8_t *texts,
                                      int64_t texts_bytes, const int64_t *text_off,
                                      const int32_t *text_len, const int32_t *max_score,
                                      const int32_t *text_begin_free, int32_t adaptive_drop,
                                      const kira_sw_cfg *cfg, int32_t want_cigar,
                                      kira_wfa_result *out, uint32_t *cigar_arena,
                                      int64_t arena_cap) {
    if (!ctx || n < 0 || patterns_bytes < 0 || texts_bytes < 0 || !cfg)
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (n == 0) return KIRA_OK;
    if ((patterns_bytes && !patterns) || (texts_bytes && !texts) || !pat_off || !pat_len ||
        !text_off || !text_len || !max_score || !out || (want_cigar && arena_cap > 0 && !cigar_arena))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (!check_ranges(n, pat_off, pat_len, patterns_bytes) ||
        !check_ranges(n, text_off, text_len, texts_bytes))
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "pattern or text range outside its buffer");
    Pen pen;
    if (!kira::wfa_pen(*cfg, pen)) {
        for (int32_t k = 0; k < n; k++) out[k] = kira_wfa_result{0, 0, 0, 0, 0, 0};
        return KIRA_OK;
    }
    std::vector<int32_t> tbf;
    try {
        tbf.assign(n, 0);
        if (text_begin_free) tbf.assign(text_begin_free, text_begin_free + n);
    } catch (const std::bad_alloc &) {
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "host allocation");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e) return fail(ctx, KIRA_ERR_DRIVER, "set device", e);
    DevMem dm(ctx);
    uint8_t *d_pat = nullptr, *d_txt = nullptr;
    int64_t *d_poff = nullptr, *d_toff = nullptr;
    int32_t *d_plen = nullptr, *d_tlen = nullptr, *d_ms = nullptr, *d_tbf = nullptr;
    hipStream_t st = ctx->stream;
    if ((e = dm.alloc(&d_pat, patterns_bytes)) || (e = dm.alloc(&d_txt, texts_bytes)) ||
        (e = dm.alloc(&d_poff, n)) || (e = dm.alloc(&d_toff, n)) || (e = dm.alloc(&d_plen, n)) ||
        (e = dm.alloc(&d_tlen, n)) || (e = dm.alloc(&d_ms, n)) || (e = dm.alloc(&d_tbf, n)))
        return fail(ctx, KIRA_ERR_OUT_OF_MEMORY, "WFA batch allocation", e);
    if ((patterns_bytes &&
         (e = hipMemcpyAsync(d_pat, patterns, patterns_bytes, hipMemcpyHostToDevice, st))) ||
        (texts_bytes && (e = hipMemcpyAsync(d_txt, texts, texts_bytes, hipMemcpyHostToDevice, st))) ||
        (e = hipMemcpyAsync(d_poff, pat_off, 8 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_toff, text_off, 8 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_plen, pat_len, 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_tlen, text_len, 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_ms, max_score, 4 * (size_t)n, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(d_tbf, tbf.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st)))
        return fail(ctx, KIRA_ERR_DRIVER, "WFA batch upload", e);
    kira::WfaDevJobs j;
    j.n = n;
    j.pat = d_pat;
    j.txt = d_txt;
    j.pat_off = d_poff;
    j.txt_off = d_toff;
    j.pat_len = d_plen;
    j.txt_len = d_tlen;
    j.max_score = d_ms;
    j.tbf = d_tbf;
    j.h_pat_len = pat_len;
    j.h_txt_len = text_len;
    j.h_max_score = max_score;
    j.h_tbf = tbf.data();
    std::vector<int32_t> res;
    std::vector<uint32_t> cig;
    std::vector<int64_t> coff;
    kira_status rs = kira::wfa_dev(ctx, j, pen, adaptive_drop, want_cigar != 0, res, cig, coff);
    if (rs != KIRA_OK) return rs;
    if ((int64_t)cig.size() > arena_cap && !cig.empty())
        return fail(ctx, KIRA_ERR_INVALID_ARGUMENT, "CIGAR arena too small");
    for (int32_t k = 0; k < n; k++) {
        const int32_t *r = &res[(size_t)k * kWfaRes];
        out[k] = kira_wfa_result{r[0], r[1], r[2], r[3], r[4], coff[k]};
    }
    if (!cig.empty()) memcpy(cigar_arena, cig.data(), 4 * cig.size());
    return KIRA_OK;
}
