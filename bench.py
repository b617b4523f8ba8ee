#!/usr/bin/env python3
"""Benchmark of the sieve + SW hot path on one MI355X.

Workload (the headline step): the batched 2-bit Hamming sieve with
certification over a 100 Mbp synthetic genome (80% uniform ACGT, 20% repeat
families of 300 bp - 6 kbp copies at 85-99% identity), 2,000,000 reads of
150 bp at 1% substitution, 16 candidate windows per read (true locus, then
copies of the same repeat family, then random loci), window = L + 4w = 278
(w = 32), mismatch limit 8.  One step = pack the 2M ASCII reads into bit
planes on the device + one sieve launch over the 32M (read, window) jobs.
Inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0).  Only --gpus 1 is supported.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "kira-ls-aligner_amd")

# MI355X peaks: HBM3E 8 TB/s (published); int32 VALU = 256 CU x 64 lanes x 2.4 GHz.
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


def load_pkg():
    name = "kira_ls_aligner_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


# --------------------------------------------------------------- workloads
ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def make_genome(rng, total=100_000_000, repeat_frac=0.20):
    """80% uniform ACGT + 20% repeat-family copies.  Returns (seq uint8, copies)
    with copies = int64 array [n, 3] of (start, length, family)."""
    target = int(total * repeat_frac)
    copies_seq, fam_of = [], []
    placed = 0
    fam = 0
    while placed < target:
        clen = int(rng.integers(300, 6001))
        cons = rng.integers(0, 4, clen, dtype=np.uint8)
        for _ in range(int(rng.integers(2, 11))):
            ident = rng.uniform(0.85, 0.99)
            c = cons.copy()
            sub = rng.random(clen) < (1 - ident)
            c[sub] = (c[sub] + rng.integers(1, 4, int(sub.sum()), dtype=np.uint8)) % 4
            # 1-bp indels at a tenth of the substitution rate, applied in one pass
            n_indel = int(rng.poisson(clen * (1 - ident) * 0.1))
            if n_indel:
                where = np.unique(rng.integers(0, clen, n_indel))
                is_del = rng.random(len(where)) < 0.5
                keep = np.ones(clen, dtype=bool)
                keep[where[is_del]] = False
                ins_at = where[~is_del]
                c = np.insert(c, ins_at, rng.integers(0, 4, len(ins_at), dtype=np.uint8))
                keep = np.insert(keep, ins_at, True)
                c = c[keep]
            copies_seq.append(c)
            fam_of.append(fam)
            placed += len(c)
            if placed >= target:
                break
        fam += 1
    n_bg = total - placed
    cuts = np.sort(rng.integers(0, n_bg, len(copies_seq)))
    bg = rng.integers(0, 4, n_bg, dtype=np.uint8)
    order = rng.permutation(len(copies_seq))
    parts, meta, prev, pos = [], [], 0, 0
    for k, ci in enumerate(order):
        seg = bg[prev:cuts[k]]
        parts.append(seg)
        pos += len(seg)
        parts.append(copies_seq[ci])
        meta.append((pos, len(copies_seq[ci]), fam_of[ci]))
        pos += len(copies_seq[ci])
        prev = cuts[k]
    parts.append(bg[prev:])
    codes = np.concatenate(parts)
    return ACGT[codes], np.array(meta, dtype=np.int64)


def make_sieve_workload(n_reads=2_000_000, read_len=150, windows=16, band=32, sub_rate=0.01,
                        genome_len=100_000_000, seed=2):
    rng = np.random.default_rng(seed)
    genome, copies = make_genome(rng, genome_len)
    G = len(genome)
    flank = 2 * band
    W = read_len + 4 * band
    pos = rng.integers(flank, G - read_len - flank - 1, n_reads)
    reads = np.empty((n_reads, read_len), dtype=np.uint8)
    chunk = 250_000
    for s in range(0, n_reads, chunk):
        e = min(n_reads, s + chunk)
        idx = pos[s:e, None] + np.arange(read_len)[None, :]
        r = genome[idx]
        m = rng.integers(0, 10000, r.shape, dtype=np.uint16) < int(sub_rate * 10000)
        codes = np.searchsorted(ACGT, r)
        codes[m] = (codes[m] + rng.integers(1, 4, int(m.sum()), dtype=np.uint8)) % 4
        reads[s:e] = ACGT[codes]
    # windows: true locus, then copies of the same repeat family, then random loci
    win = rng.integers(0, G - W, (n_reads, windows)).astype(np.int64)
    win[:, 0] = pos - flank
    starts = copies[:, 0]
    ci = np.searchsorted(starts, pos, side="right") - 1
    inside = (ci >= 0) & (pos + read_len <= starts[np.maximum(ci, 0)] + copies[np.maximum(ci, 0), 1])
    fam_members = {}
    for k, f in enumerate(copies[:, 2]):
        fam_members.setdefault(int(f), []).append(k)
    for r in np.nonzero(inside)[0]:
        c = ci[r]
        off = pos[r] - starts[c]
        slot = 1
        for other in fam_members[int(copies[c, 2])]:
            if other == c or slot >= windows:
                continue
            s0 = starts[other] + off - flank
            if 0 <= s0 <= G - W:
                win[r, slot] = s0
                slot += 1
    return dict(genome=genome, reads=reads, win_start=win.reshape(-1),
                read_len=read_len, W=W, windows=windows, n_reads=n_reads,
                repeat_reads=int(inside.sum()))


def make_sw_workload(genome, n_pairs=1_000_000, seed=3, pad=1024, chunk=100_000):
    """Uneven banded SW pairs: read length U{100..300}, band U{16,32,64,128},
    window = clamp_window geometry (flank max(2w, 50) around the read's locus,
    src/alignment/mod.rs:1766-1771), read drawn from the locus with a per-pair
    substitution rate U[1%, 5%] and 0.5% indel events (half insertions, half
    deletions, 1 bp).  Windows are slices of the resident genome: refs is the
    genome with `pad` zero bytes on both sides and ref_off points into it."""
    rng = np.random.default_rng(seed)
    G = len(genome)
    L = rng.integers(100, 301, n_pairs).astype(np.int64)
    band = rng.choice(np.array([16, 32, 64, 128], dtype=np.int64), n_pairs)
    flank = np.maximum(2 * band, 50)
    sub = rng.uniform(0.01, 0.05, n_pairs)
    pos = rng.integers(300, G - 700, n_pairs).astype(np.int64)
    codes_g = np.searchsorted(ACGT, genome)  # genome is ACGT only
    reads, read_len = [], np.empty(n_pairs, dtype=np.int32)
    for s in range(0, n_pairs, chunk):
        e = min(n_pairs, s + chunk)
        n = e - s
        slen = L[s:e] + 16
        start = np.repeat(pos[s:e], slen)
        base_off = np.arange(int(slen.sum())) - np.repeat(np.cumsum(slen) - slen, slen)
        src = codes_g[start + base_off].astype(np.uint8)
        tot = len(src)
        u = rng.random(tot)
        is_del = u < 0.0025
        is_ins = (u >= 0.0025) & (u < 0.005)
        subm = rng.random(tot) < np.repeat(sub[s:e], slen)
        src[subm] = (src[subm] + rng.integers(1, 4, int(subm.sum()), dtype=np.uint8)) % 4
        cnt = 1 - is_del.astype(np.int64) + is_ins.astype(np.int64)
        out = np.repeat(src, cnt)
        first = np.cumsum(cnt) - cnt
        out[first[is_ins]] = rng.integers(0, 4, int(is_ins.sum()), dtype=np.uint8)
        pair = np.repeat(np.repeat(np.arange(n), slen), cnt)
        emitted = np.bincount(pair, minlength=n)
        local = np.arange(len(out)) - np.repeat(np.cumsum(emitted) - emitted, emitted)
        keep = local < np.repeat(L[s:e], emitted)
        reads.append(ACGT[out[keep]])
        read_len[s:e] = np.minimum(L[s:e], emitted)
    reads = np.concatenate(reads)
    read_off = np.concatenate([[0], np.cumsum(read_len[:-1], dtype=np.int64)])
    win_start = pos - flank
    win_end = np.minimum(pos + L + flank, G)
    refs = np.zeros(G + 2 * pad, dtype=np.uint8)
    refs[pad:pad + G] = genome
    return dict(reads=reads, read_off=read_off.astype(np.int64), read_len=read_len,
                refs=refs, ref_off=(win_start + pad).astype(np.int64),
                ref_len=(win_end - win_start).astype(np.int32),
                offset=(pos - win_start).astype(np.int32), band=band.astype(np.int32),
                abort=np.full(n_pairs, -(2 ** 31) // 8, dtype=np.int32), n=n_pairs)


def make_long_workload(genome, n_windows=82_000, seed=4, pad=1024, chunk=8_000):
    """Long-read chain-gap windows: source length U[500, 5000] bp, band
    ceil(5% of the length) (25-250), query drawn from the source with 5%
    substitutions, 2.5% insertions and 2.5% deletions (1 bp each); the window
    is the source plus `band` bases on each side, offset = band.  82k windows
    ~ 20k reads of median 10 kbp cut into chain gaps."""
    rng = np.random.default_rng(seed)
    G = len(genome)
    S = rng.integers(500, 5001, n_windows).astype(np.int64)
    band = np.ceil(0.05 * S).astype(np.int64)
    pos = rng.integers(300, G - 5600, n_windows).astype(np.int64)
    codes_g = np.searchsorted(ACGT, genome)
    reads, read_len = [], np.empty(n_windows, dtype=np.int32)
    for s in range(0, n_windows, chunk):
        e = min(n_windows, s + chunk)
        n = e - s
        slen = S[s:e]
        start = np.repeat(pos[s:e], slen)
        base_off = np.arange(int(slen.sum())) - np.repeat(np.cumsum(slen) - slen, slen)
        src = codes_g[start + base_off].astype(np.uint8)
        tot = len(src)
        u = rng.random(tot)
        is_del = u < 0.025
        is_ins = (u >= 0.025) & (u < 0.05)
        subm = rng.random(tot) < 0.05
        src[subm] = (src[subm] + rng.integers(1, 4, int(subm.sum()), dtype=np.uint8)) % 4
        cnt = 1 - is_del.astype(np.int64) + is_ins.astype(np.int64)
        out = np.repeat(src, cnt)
        first = np.cumsum(cnt) - cnt
        out[first[is_ins]] = rng.integers(0, 4, int(is_ins.sum()), dtype=np.uint8)
        pair = np.repeat(np.repeat(np.arange(n), slen), cnt)
        read_len[s:e] = np.bincount(pair, minlength=n)
        reads.append(ACGT[out])
    reads = np.concatenate(reads)
    read_off = np.concatenate([[0], np.cumsum(read_len[:-1], dtype=np.int64)])
    win_start = pos - band
    refs = np.zeros(G + 2 * pad, dtype=np.uint8)
    refs[pad:pad + G] = genome
    return dict(reads=reads, read_off=read_off.astype(np.int64), read_len=read_len,
                refs=refs, ref_off=(win_start + pad).astype(np.int64),
                ref_len=(S + 2 * band).astype(np.int32), offset=band.astype(np.int32),
                band=band.astype(np.int32),
                abort=np.full(n_windows, -(2 ** 31) // 8, dtype=np.int32), n=n_windows)


# --------------------------------------------------------------- CPU baseline
_LUT = np.zeros(256, dtype=np.uint8)
for _c, _v in zip(b"ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3]):
    _LUT[_c] = _v


def pack_rows(rows):
    """Vectorised PackedDna::pack of equal-length rows (bitpacked.rs:57-86)."""
    n, L = rows.shape
    nb = (L + 3) // 4
    codes = np.zeros((n, nb * 4), dtype=np.uint8)
    codes[:, :L] = _LUT[rows]
    c = codes.reshape(n, nb, 4)
    return (c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)).astype(np.uint8)


def pre_shift_rows(bits):
    """Vectorised pre_shift_into (bitpacked.rs:95-119): [n, 4, nb]."""
    n, nb = bits.shape
    out = np.empty((n, 4, nb), dtype=np.uint8)
    out[:, 0] = bits
    nxt = np.zeros_like(bits)
    nxt[:, :-1] = bits[:, 1:]
    for p in range(1, 4):
        out[:, p] = (bits >> (2 * p)) | ((nxt << (8 - 2 * p)) & 0xFF)
    return out


def cpu_env():
    """The host the CPU baselines run on: lscpu's model and core counts, and
    the CPUs this process may use (its affinity mask)."""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key, default=0):
        try:
            return int(info.get(key, default))
        except ValueError:
            return default
    sockets, cps, tpc = num("Socket(s)", 1), num("Core(s) per socket"), num("Thread(s) per core", 1)
    return {"model": info.get("Model name"), "sockets": sockets, "cores_per_socket": cps,
            "threads_per_core": tpc, "physical_cores": sockets * cps,
            "logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def median_of_runs(fn, runs=5):
    """One warm-up call, then `runs` timed calls: (median seconds, last result)."""
    fn()
    ts, r = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), r


def sieve_layout(work, read_idx, limit):
    """The jobs of the given reads (all their windows) in the reference's SoA
    launch layout (CudaBackend::run_batch, src/cuda/backend.rs:118-192):
    packed reads of READ_BYTES_MAX bytes, four pre-shifted window phases."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    L, W, K = work["read_len"], work["W"], work["windows"]
    reads = work["reads"][read_idx]
    n = len(read_idx) * K
    starts = work["win_start"].reshape(-1, K)[read_idx].reshape(-1)
    texts = work["genome"][starts[:, None] + np.arange(W)[None, :]]
    rb = pack_rows(reads)
    tb = pre_shift_rows(pack_rows(texts))
    nb = tb.shape[2]
    rp = np.zeros((n, O.READ_BYTES_MAX), dtype=np.uint8)
    rp[:, :rb.shape[1]] = np.repeat(rb, K, axis=0)
    return dict(n=n, reads=rp.reshape(-1), read_lens=np.full(n, L, dtype=np.int32),
                ref_lens=np.full(n, W, dtype=np.int32),
                limits=np.full(n, limit, dtype=np.uint32),
                ref_phase_bytes=np.full(n, nb, dtype=np.int64),
                ref_offsets=np.arange(n, dtype=np.int64) * nb,
                ref_flat=np.ascontiguousarray(tb.transpose(1, 0, 2)).reshape(-1),
                ref_bytes_per_phase=n * nb, texts=texts, read_rows=reads)


def sieve_oracle_flags(lay, shifts, best, second, K, cert_max):
    """Flags of the resident sieve for jobs the oracle scanned: dispatched
    (the genome and reads are all ACGT), accepted, certified
    (spectral_hit_is_certified, src/alignment/mod.rs:688-762)."""
    from oracle import oracle as O
    L = lay["read_rows"].shape[1]
    flags = np.ones(lay["n"], dtype=np.uint8)
    for j in np.nonzero(shifts >= 0)[0]:
        flags[j] |= 2
        sh = int(shifts[j])
        sec = None if second[j] == 2 ** 31 - 1 else int(second[j])
        if O.certified(lay["read_rows"][j // K].tobytes(), lay["texts"][j, sh:sh + L].tobytes(),
                       int(best[j]), sec, cert_max, False):
            flags[j] |= 4
    return flags


# --------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reads", type=int, default=2_000_000)
    ap.add_argument("--cpu-sample-reads", type=int, default=131_072)
    ap.add_argument("--pcie-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sw-pairs", type=int, default=1_000_000)
    ap.add_argument("--sw-steps", type=int, default=3)
    ap.add_argument("--sw-tune-cache", default=os.environ.get("KIRA_SW_TUNE_CACHE", ""),
                    help="file the SW autotuner writes its routes to (default: none)")
    ap.add_argument("--no-sw-tune", action="store_true",
                    help="SW legs on the built-in routing only (no autotuner pass)")
    ap.add_argument("--cpu-sample-pairs", type=int, default=65_536)
    ap.add_argument("--long-windows", type=int, default=82_000)
    ap.add_argument("--stage4-reads", type=int, default=200_000)
    ap.add_argument("--stage4-steps", type=int, default=3)
    ap.add_argument("--c1-reads", type=int, default=10_000)
    ap.add_argument("--stage4-only", action="store_true",
                    help="run only the stage-4 driver section (for profiling)")
    ap.add_argument("--fastpath-chains", type=int, default=200_000)
    ap.add_argument("--fastpath-steps", type=int, default=3)
    ap.add_argument("--sw-only", action="store_true",
                    help="run only the banded SW section (for profiling)")
    ap.add_argument("--sw-genome-mbp", type=int, default=100)
    ap.add_argument("--fastpath-only", action="store_true",
                    help="run only the fast DP cascade section (for profiling)")
    ap.add_argument("--serve-requests", type=int, default=24)
    # 8 requests in flight in a 192 GiB pool (two thirds of the 288 GB of HBM):
    # 4 slots in 96 GiB gave 0.87-0.91 M reads/s, 8 in 192 GiB 1.0-1.06 M
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gb", type=int, default=192, help="GiB")
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--calib-copy-gb", type=float, default=0.0,
                    help="after the timed region, run one plain device copy of this many "
                         "GiB (the reference kernel for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU; --gpus must be 1 (batches shard across GPUs "
              "with no exchange, each GPU holding its own copy of the packed genome)",
              file=sys.stderr)
        sys.exit(2)
    run_bench(args)


def run_bench(args):
    import torch

    pkg = load_pkg()
    if not pkg.is_available():
        raise SystemExit("no GPU visible to libkira_hip")
    if args.stage4_only:
        ctx = pkg.Context(0)
        print(json.dumps({"stage4": bench_stage4(pkg, ctx, args)}), flush=True)
        return
    if args.fastpath_only:
        ctx = pkg.Context(0)
        print(json.dumps({"fastpath": bench_fastpath(pkg, ctx, args)}), flush=True)
        return
    if args.sw_only:
        # the SW legs alone (for profiling): genome as in the headline workload
        ctx = pkg.Context(0)
        genome = make_genome(np.random.default_rng(2), args.sw_genome_mbp * 1_000_000)[0]
        lw = make_long_workload(genome, n_windows=args.long_windows) \
            if args.long_windows > 0 else None
        stream = torch.cuda.Stream(torch.device("cuda:0"))
        torch.cuda.set_stream(stream)
        print(json.dumps({"sw": bench_sw(pkg, ctx, genome, args, stream, lw)}), flush=True)
        return
    t_gen = time.time()
    work = make_sieve_workload(n_reads=args.reads)
    t_gen = time.time() - t_gen
    print(f"[bench] workload generated in {t_gen:.1f}s: {work['n_reads']} reads, "
          f"{work['repeat_reads']} in repeats", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    L, W, K, n_reads = work["read_len"], work["W"], work["windows"], work["n_reads"]
    n_jobs = n_reads * K
    limit = 8
    ctx = pkg.Context(0)
    genome = pkg.Genome(ctx, work["genome"].tobytes())
    stride = 160
    rd = np.zeros((n_reads, stride), dtype=np.uint8)
    rd[:, :L] = work["reads"]
    reads = torch.from_numpy(rd).to(dev)
    lens = torch.full((n_reads,), L, dtype=torch.int32, device=dev)
    win_start = torch.from_numpy(work["win_start"]).to(dev)
    win_len = torch.full((n_jobs,), W, dtype=torch.int32, device=dev)
    planes = torch.empty((n_reads, 24), dtype=torch.int32, device=dev)
    valid = torch.empty(n_reads, dtype=torch.uint8, device=dev)
    out = pkg.SieveOutputs(*(torch.empty(n_jobs, dtype=torch.int32, device=dev)
                             for _ in range(3)),
                           torch.empty(n_jobs, dtype=torch.uint8, device=dev))
    # one explicit stream for every launch and event of the timed region
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    cert_max = 4  # spectral_max_mismatches(150) = max(150*3/100, 2)

    def step(ev=None):
        pkg.sieve_genome(ctx, genome, reads, lens, win_start, win_len, jobs_per_read=K,
                         limit_all=limit, cert_max_mism=cert_max, max_read_len=L, planes=planes,
                         valid=valid, out=out, stream=stream)

    def step_timed(evs):
        # pack reads, then the sieve launch bracketed by events on the same stream
        pkg.sieve_genome(ctx, genome, reads, lens, win_start[:0], win_len[:0],
                         jobs_per_read=K, limit_all=limit, max_read_len=L, planes=planes,
                         valid=valid, out=out, stream=stream)
        evs[0].record(stream)
        pkg.sieve_genome(ctx, genome, reads, lens, win_start, win_len, jobs_per_read=K,
                         limit_all=limit, cert_max_mism=cert_max, max_read_len=L,
                         planes=planes, valid=valid, out=out, stream=stream, skip_pack=True)
        evs[1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kern_ms = []
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step_timed(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [e[0].elapsed_time(e[1]) for e in evs]
    ms_per_step = wall * 1000 / args.steps
    kernel_ms = float(np.mean(kern_ms))
    print(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step, sieve kernel "
          f"{kernel_ms:.3f} ms", file=sys.stderr, flush=True)

    if args.calib_copy_gb > 0:
        nbytes = int(args.calib_copy_gb * (1 << 30))
        src = torch.ones(nbytes // 4, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy: {nbytes} bytes read + written", file=sys.stderr,
              flush=True)
        del src, dst
    flags = out.flags.cpu().numpy()
    shifts = out.shift.cpu().numpy()
    best = out.best.cpu().numpy()
    second = out.second.cpu().numpy()
    accepted = int(((flags & 2) != 0).sum())
    certified = int(((flags & 4) != 0).sum())
    true_hit = float((shifts.reshape(n_reads, K)[:, 0] == 2 * 32).mean())

    # the drop-in form: the same workload's host arrays in the reference layout
    # through kira_sieve_run_batch (PCIe both ways), and the resident path with
    # its host transfers included
    pcie = bench_resident_pcie(pkg, ctx, genome, rd, lens, win_start, win_len, out, planes,
                               valid, stream, K, L, limit, cert_max, args.pcie_steps)
    print(f"[bench] resident path with PCIe: {pcie['ms_per_step']:.2f} ms/step, "
          f"{pcie['reads_per_s']:.3g} reads/s", file=sys.stderr, flush=True)
    host = cpu_env()
    cpu = None
    parity = None
    plain = None
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        threads = host["affinity_cpus"]
        ns = min(args.cpu_sample_reads, n_reads)
        # reads drawn from the whole batch (seeded), every window of each
        idx = np.sort(np.random.default_rng(17).choice(n_reads, ns, replace=False))
        lay = sieve_layout(work, idx, limit)
        t_all, (cs, cb, csec) = median_of_runs(
            lambda: O.backend_run_batch_layout(lay, threads=threads))
        n1 = min(ns, 8192)
        lay1 = sieve_layout(work, idx[:n1], limit)
        t_one, _ = median_of_runs(lambda: O.backend_run_batch_layout(lay1, threads=1))
        jobs = (idx[:, None] * K + np.arange(K)[None, :]).reshape(-1)
        cflags = sieve_oracle_flags(lay, cs, cb, csec, K, cert_max)
        parity = bool(np.array_equal(cs, shifts[jobs]) and np.array_equal(cb, best[jobs]) and
                      np.array_equal(csec, second[jobs]) and np.array_equal(cflags, flags[jobs]))
        cpu = {"value": ns / t_all, "unit": "reads/s", "cores": threads, "kind": "port",
               "single_thread_value": n1 / t_one,
               "sample": f"{ns} reads drawn from the whole batch x {K} windows = {ns * K} "
                         f"scans of 150 bp vs 278 bp (mismatch_count_u64 + best/second, packing "
                         f"excluded), {threads} threads (the CPUs this process may use); "
                         f"{n1} of them on 1 thread; 1 warm-up + median of 5",
               "host": host}
        print(f"[bench] cpu baseline {ns / t_all:.0f} reads/s on {threads} threads, "
              f"{n1 / t_one:.0f} on 1; parity (shift, best, second, flags)={parity}",
              file=sys.stderr, flush=True)
        plain = bench_plain_sieve(pkg, ctx, lay, ns, cs, cb, csec)
        print(f"[bench] plain drop-in form (kira_sieve_run_batch, host arrays): "
              f"{plain['ms_per_call']:.1f} ms per {ns}-read call, {plain['reads_per_s']:.3g} "
              f"reads/s, parity={plain['parity_vs_oracle']}", file=sys.stderr, flush=True)
        del lay, lay1

    n_shifts = W - L + 1
    words = (L + 31) // 32
    bytes_per_launch = n_reads * ((L + 3) // 4) + n_jobs * (((W + 3) // 4) + 12)
    ops_per_launch = n_jobs * n_shifts * words * 4
    gbs = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    # on this workload (scaled by a plain 1 GiB copy), kept beside the profiles
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sieve_kernel_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("workload_reads") == n_reads and t.get("window") == W:
            traffic = (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024
    tops = ops_per_launch / (kernel_ms * 1e-3) / 1e12
    line = {
        "metric": "reads/s + SW GCUPS on 1 MI355X vs host CPU; p50/p99 batch latency",
        "value": n_reads / (ms_per_step * 1e-3),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "batched 2-bit sieve + certification: 100 Mbp genome (20% "
                               "repeat families), 2M x 150 bp reads at 1% subs, 16 windows "
                               "per read, W=278, limit 8",
                   "reads": n_reads, "jobs": n_jobs, "window": W, "read_len": L,
                   "parallelism": "1 GPU"},
        "jobs_per_s": n_jobs / (ms_per_step * 1e-3),
        "base_compares_per_s": n_jobs * n_shifts * L / (ms_per_step * 1e-3),
        "sieve_kernel_ms": kernel_ms,
        "accepted_jobs": accepted,
        "certified_jobs": certified,
        "true_locus_recall": true_hit,
        "parity_vs_oracle_sample": parity,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": traffic},
        "roofline_valu": {"bound": "valu", "achieved": tops, "peak": VALU_PEAK_TOPS,
                          "unit": "Tops/s", "frac": tops / VALU_PEAK_TOPS,
                          "ops_per_launch": ops_per_launch},
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "plain_gpu_form": plain,
        "batched_vs_plain": (n_reads / (ms_per_step * 1e-3)) / plain["reads_per_s"]
        if plain else None,
    }
    lw = None
    if args.long_windows > 0 and (args.sw_pairs > 0 or args.serve_requests > 0):
        t0 = time.time()
        lw = make_long_workload(work["genome"], n_windows=args.long_windows)
        print(f"[bench] long windows: {lw['n']} generated in {time.time() - t0:.1f}s",
              file=sys.stderr, flush=True)
    if args.sw_pairs > 0:
        line["sw"] = bench_sw(pkg, ctx, work["genome"], args, stream, lw)
        line["sw_gcups"] = line["sw"]["gcups"]
    if args.serve_requests > 0 and lw is not None:
        del reads, lens, win_start, win_len, planes, valid, out
        torch.cuda.empty_cache()
        line["serving"] = bench_serving(pkg, ctx, genome, work, lw, args)
    if args.stage4_reads > 0:
        line["stage4"] = bench_stage4(pkg, ctx, args)
    if args.fastpath_chains > 0:
        line["fastpath"] = bench_fastpath(pkg, ctx, args)
    print(json.dumps(line), flush=True)


def bench_resident_pcie(pkg, ctx, genome, rd, lens, win_start, win_len, out, planes, valid,
                        stream, K, L, limit, cert_max, steps):
    """The headline step with its host transfers: ASCII reads from pinned host
    memory to the device, packing + sieve, the four outputs back to pinned
    host memory; HIP events on the launch stream around each step."""
    import torch
    h_reads = torch.from_numpy(rd).pin_memory()
    d_reads = torch.empty(rd.shape, dtype=torch.uint8, device="cuda:0")
    h_out = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
             for t in (out.shift, out.best, out.second, out.flags)]

    def one():
        d_reads.copy_(h_reads, non_blocking=True)
        pkg.sieve_genome(ctx, genome, d_reads, lens, win_start, win_len, jobs_per_read=K,
                         limit_all=limit, cert_max_mism=cert_max, max_read_len=L, planes=planes,
                         valid=valid, out=out, stream=stream)
        for h, d in zip(h_out, (out.shift, out.best, out.second, out.flags)):
            h.copy_(d, non_blocking=True)

    one()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(steps):
        one()
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    n_reads = rd.shape[0]
    moved = rd.nbytes + sum(t.numel() * t.element_size() for t in h_out)
    del h_reads, d_reads, h_out
    return {"ms_per_step": ms, "reads_per_s": n_reads / (ms * 1e-3), "steps": steps,
            "host_bytes_moved": moved,
            "note": "reads H2D from pinned memory + pack + sieve + shift/best/second/flags "
                    "D2H to pinned memory, per step"}


def bench_plain_sieve(pkg, ctx, lay, n_reads, cs, cb, csec, runs=5):
    """The drop-in form of the reference's --gpu sieve (CudaBackend::run_batch,
    src/cuda/backend.rs:96-249): host arrays in the reference's SoA layout
    (packed reads, four pre-shifted window phases) through kira_sieve_run_batch,
    which uploads, packs the window planes, scans and downloads per call."""
    import ctypes as C
    lib = pkg.lib()
    n = lay["n"]
    plen = np.full(n, (int(lay["read_lens"][0]) + 3) // 4, dtype=np.int32)
    out_s = np.empty(n, dtype=np.int32)
    out_m = np.empty(2 * n, dtype=np.int32)
    P = lambda a: a.ctypes.data_as(C.c_void_p)

    def call():
        st = lib.kira_sieve_run_batch(ctx.h, n, P(lay["reads"]), pkg.READ_BYTES_MAX, P(plen),
                                      P(lay["read_lens"]), P(lay["ref_flat"]),
                                      int(lay["ref_bytes_per_phase"]), P(lay["ref_offsets"]),
                                      P(lay["ref_lens"]), P(lay["limits"]), P(out_s), P(out_m))
        ctx.check(st, "kira_sieve_run_batch")

    t, _ = median_of_runs(call, runs)
    parity = bool(np.array_equal(out_s, cs) and np.array_equal(out_m[:n], cb) and
                  np.array_equal(out_m[n:], csec))
    return {"ms_per_call": t * 1e3, "reads_per_s": n_reads / t, "jobs_per_s": n / t,
            "reads": n_reads, "jobs": n, "parity_vs_oracle": parity,
            "host_bytes_per_call": int(lay["reads"].nbytes + lay["ref_bytes_per_phase"] +
                                       n * 28 + 12 * n),
            "note": "1 warm-up + median of 5 calls; host arrays built once outside the timing"}


# the read kinds of tests/fastpath_data.py weighted like the chains that reach
# the reference's fast path on real data (about half certified ungapped, half
# gapped): exact, substitutions, insertion, deletion, 5' deletion, foreign
# tail, unrelated, N, lower case, contig end, over-long, across an N run
FAST_MIX = [0.18, 0.30, 0.16, 0.16, 0.08, 0.02, 0.02, 0.02, 0.0, 0.02, 0.02, 0.02]


def bench_fastpath(pkg, ctx, args):
    """The fast DP cascade (kira_fast_dp_batch: packed certificate, Myers,
    WFA with the origin retry) over one batch of chains, host arrays in and
    records out, plus the WFA and Myers batch entry points alone on the
    cascade's gapped pairs; the CPU restatement on a sample with parity."""
    from tests.fastpath_data import make_fast_batch
    from oracle import fastpath as F
    from oracle import oracle as O
    t0 = time.time()
    n = args.fastpath_chains
    b = make_fast_batch(31, n, lens=[2_500_000] * 8, mix=FAST_MIX)
    print(f"[bench] fastpath batch: {n} chains generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    cfg = pkg.sw_cfg()
    kn = pkg.fast_knobs()
    # host arrays as a caller hands them over (marshalled once, outside the timing)
    x = pkg.FastInputs(b["contig_off"], b["contig_len"], b["reads"], b["chains"])
    pkg.fast_dp_run(ctx, genome, x, cfg, kn)
    times, runs = [], []
    for _ in range(args.fastpath_steps):
        t0 = time.perf_counter()
        r = pkg.fast_dp_run(ctx, genome, x, cfg, kn)
        times.append(time.perf_counter() - t0)
        runs.append(r)
    k = int(np.argsort(times)[len(times) // 2])
    arr, cig, md, stats = runs[k]
    ms = 1e3 * times[k]
    out = {"workload": f"fast DP cascade: 8 x 2.5 Mbp contigs, {n} chains of 76-250 bp reads "
                       "(exact, substitutions, indels, 5' deletions, foreign tails, N), "
                       "router defaults",
           "chains": n, "ms_per_batch": ms, "chains_per_s": n / (ms * 1e-3), "stats": stats}
    # the gapped pairs alone through the WFA and Myers entry points
    pats, texts = [], []
    for read, ch in zip(b["reads"], b["chains"]):
        c = b["contigs"][ch[0]]
        ws = ch[1] - ch[3]
        if len(read) > 300 or ws < 0 or ws + len(read) + 32 > len(c):
            continue
        pats.append(read)
        texts.append(c[ws:ws + len(read) + 32])
    budget = [max((len(r) * 25 // 100) * 5, 42) for r in pats]
    wx = pkg.PairInputs(pats, texts, budget)
    mx = pkg.PairInputs(pats, texts, [len(r) * 15 // 100 + 4 for r in pats])
    pkg.wfa_run(ctx, pkg.PairInputs(pats[:1000], texts[:1000], budget[:1000]), cfg)
    t0 = time.perf_counter()
    wres, _ = pkg.wfa_run(ctx, wx, cfg)
    t_wfa = time.perf_counter() - t0
    t0 = time.perf_counter()
    pkg.myers_run(ctx, mx)
    t_myers = time.perf_counter() - t0
    out["wfa_pairs"] = len(pats)
    out["wfa_pairs_per_s"] = len(pats) / t_wfa
    out["wfa_found"] = int(wres["found"].sum())
    out["myers_pairs_per_s"] = len(pats) / t_myers
    print(f"[bench] fastpath: {n} chains, {ms:.1f} ms/batch, {out['chains_per_s']:.0f} chains/s, "
          f"stats {stats}; WFA {out['wfa_pairs_per_s']:.0f} pairs/s, Myers "
          f"{out['myers_pairs_per_s']:.0f} pairs/s (host arrays in/out)", file=sys.stderr,
          flush=True)
    if not args.no_cpu:
        m = 4000
        ocfg, router = O.cfg(), F.Router()
        t0 = time.perf_counter()
        want = [F.fast_dp(r, b["contigs"][ch[0]], ch[1], ch[3], ch[6], ocfg, router)
                for r, ch in zip(b["reads"][:m], b["chains"][:m])]
        cpu_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for r, t, bud in zip(pats[:m], texts[:m], budget[:m]):
            O.wfa_align_scored(r, t, ocfg, bud)
        cpu_wfa = time.perf_counter() - t0
        norm = [None if w is None else dict(kind=1 if w[0] == "packed" else 2,
                                            ref_start=w[1]["ref_start"], ref_end=w[1]["ref_end"],
                                            score=w[1]["score"], nm=w[1]["nm"], md=w[1]["md"],
                                            cigar=w[1]["cigar"])
                for w in want]
        gotm = []
        for o in arr[:m]:
            if o["kind"] == 0:
                gotm.append(None)
                continue
            c0, m0 = int(o["cigar_off"]), int(o["md_off"])
            gotm.append(dict(kind=int(o["kind"]), ref_start=int(o["ref_start"]),
                             ref_end=int(o["ref_end"]), score=int(o["score"]), nm=int(o["nm"]),
                             md=md[m0:m0 + int(o["md_len"])].decode(),
                             cigar=[int(v) for v in cig[c0:c0 + int(o["n_cigar"])]]))
        parity = norm == gotm
        # the C restatement of the cascade on every CPU of the process and on
        # one thread (1 warm-up + median of 5), with parity on its sample
        mc = min(n, 65536)
        threads = len(os.sched_getaffinity(0))
        prep = F.fast_dp_prepare(b["reads"][:mc], b["chains"][:mc], b["genome"],
                                 b["contig_off"], b["contig_len"])
        t_all, cres = median_of_runs(lambda: F.fast_dp_run_mt(prep, ocfg, router, threads))
        m1 = max(1, mc // 16)
        prep1 = F.fast_dp_prepare(b["reads"][:m1], b["chains"][:m1], b["genome"],
                                  b["contig_off"], b["contig_len"])
        t_one, _ = median_of_runs(lambda: F.fast_dp_run_mt(prep1, ocfg, router, 1))
        g = arr[:mc]
        keys = ("kind", "ref_start", "ref_end", "score", "nm", "n_cigar")
        live = g["kind"] != 0
        parity_c = bool(np.array_equal(cres["kind"], g["kind"]) and all(
            np.array_equal(cres[k][live], g[k][live]) for k in keys[1:]))
        out["cpu_baseline"] = {
            "value": mc / t_all, "unit": "chains/s", "cores": threads, "kind": "port",
            "single_thread_value": m1 / t_one,
            "sample": f"first {mc} chains of the batch through the C restatement of the "
                      f"cascade (oracle/kira_fast_oracle.c, pthreads over chains, {threads} "
                      f"threads = the CPUs this process may use; the first {m1} on 1 "
                      f"thread), 1 warm-up + median of 5",
            "python_driver_chains_per_s": m / cpu_s}
        out["cpu_wfa_pairs_per_s"] = m / cpu_wfa
        out["parity_vs_oracle_sample"] = parity and parity_c
        print(f"[bench] fastpath cpu baseline {mc / t_all:.0f} chains/s on {threads} threads, "
              f"{m1 / t_one:.0f} on 1 (C restatement); Python driver {m / cpu_s:.0f} chains/s; "
              f"WFA {m / cpu_wfa:.0f} pairs/s (1 thread); parity={parity} (MD / CIGAR, "
              f"{m} chains), {parity_c} (C, {mc} chains)", file=sys.stderr, flush=True)
    genome.close()
    return out


def bench_stage4(pkg, ctx, args):
    """The stage-4 driver (kira_stage4_run) over one batch of single-end reads
    with their chains: reads/s per call and host-wall phase split in the
    --gpu and CPU-mode configurations; the restated reference CPU path (the
    C driver, every CPU) over the same batch, with parity of the CPU-mode
    results on every read and of the --gpu results on the first 2000."""
    from tests.stage4_data import make_batch
    from oracle import stage4 as S
    t0 = time.time()
    b = make_batch(21, args.stage4_reads, n_contigs=8, contig_len=2_500_000)
    print(f"[bench] stage4 batch: {len(b['reads'])} reads, "
          f"{sum(len(c) for c in b['chains'])} chains, generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    cfg = pkg.stage4_cfg()
    # host arrays as a caller hands them over (marshalled once, outside the timing)
    x = pkg.Stage4Inputs(b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                         b["repeat_min_occ"])
    n = len(b["reads"])

    def timed(c):
        pkg.stage4_run(ctx, genome, x, cfg=c).close()
        times, phases, stats = [], [], []
        for _ in range(args.stage4_steps):
            t0 = time.perf_counter()
            res = pkg.stage4_run(ctx, genome, x, cfg=c)
            times.append(time.perf_counter() - t0)
            phases.append(res.phase_ms)
            stats.append(res.stats)
            res.close()
        k = int(np.argsort(times)[len(times) // 2])
        ms = 1e3 * times[k]
        return {"ms_per_batch": ms, "reads_per_s": n / (ms * 1e-3),
                "phase_ms": dict(zip(["triage", "sieve", "lane_sw", "banded_sw", "records"],
                                     phases[k])),
                "stats": {kk: v for kk, v in stats[k].items() if kk != "reason_counts"}}

    out = {"workload": "stage-4 driver: 8 x 2.5 Mbp contigs (20% repeat families), "
                       f"{n} single-end reads of 100-250 bp (subs, indels, N, chimeras), "
                       "1-3 chains each, mem defaults (two-tier reorder on), AVX2 lane groups",
           "reads": n}
    out.update(timed(cfg))
    ms = out["ms_per_batch"]
    print(f"[bench] stage4 (--gpu): {n} reads, {ms:.1f} ms/batch, {out['reads_per_s']:.0f} "
          f"reads/s, phases {out['phase_ms']}", file=sys.stderr, flush=True)
    cpu_cfg = pkg.stage4_cfg(cpu_mode=True)
    out["cpu_mode"] = timed(cpu_cfg)
    print(f"[bench] stage4 (cpu mode, fast DP cascade): {out['cpu_mode']['ms_per_batch']:.1f} "
          f"ms/batch, {out['cpu_mode']['reads_per_s']:.0f} reads/s, phases "
          f"{out['cpu_mode']['phase_ms']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        # the --gpu configuration against the Python restatement on the first
        # reads (it has no faster form)
        m = 2000
        sub = {key: b[key][:m] for key in ("reads", "chains", "repeat_min_occ")}
        want, _ = S.run(b["contigs"], sub["reads"], sub["chains"], S.default_cfg(),
                        sub["repeat_min_occ"])
        res = pkg.stage4_run(ctx, genome, b["contig_off"], b["contig_len"], sub["reads"],
                             sub["chains"], cfg, sub["repeat_min_occ"])
        out["parity_vs_oracle_sample"] = res.alignments() == want
        res.close()
        # the restated reference CPU path: the stage-4 driver in C without the
        # GPU dispatcher (oracle/kira_stage4_oracle.c, pthreads over reads,
        # lane groups and jobs) over the whole batch on every CPU of this
        # process, and a slice of it on one thread; 1 warm-up + median of 5
        T = len(os.sched_getaffinity(0))
        ccfg = S.default_cfg(cpu_mode=True)
        cb = S.CBatch(b["genome"], b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                      b["repeat_min_occ"])
        t_all, _ = median_of_runs(lambda: S.run_c(cb, ccfg, T, decode=False))
        m1 = min(n, 10_000)
        cb1 = S.CBatch(b["genome"], b["contig_off"], b["contig_len"], b["reads"][:m1],
                       b["chains"][:m1], b["repeat_min_occ"][:m1])
        t_one, _ = median_of_runs(lambda: S.run_c(cb1, ccfg, 1, decode=False), runs=3)
        # the device in the same configuration against it, every read
        want_cpu, _ = S.run_c(cb, ccfg, T)
        res = pkg.stage4_run(ctx, genome, x, cfg=cpu_cfg)
        out["cpu_mode"]["parity_vs_c_driver_all_reads"] = res.alignments() == want_cpu
        res.close()
        del want_cpu
        out["cpu_baseline"] = {
            "value": n / t_all, "unit": "reads/s", "cores": T, "kind": "port",
            "label": "restated reference CPU path",
            "single_thread_value": m1 / t_one,
            "sample": f"the whole batch ({n} reads) through the stage-4 driver restated in C "
                      "without the GPU dispatcher (oracle/kira_stage4_oracle.c: fast DP "
                      "cascade before SW, AVX2 16-lane groups, pthreads over reads, groups "
                      f"and jobs), {T} threads = the CPUs this process may use; the first "
                      f"{m1} reads on 1 thread; 1 warm-up + median of 5 (3 on 1 thread)"}
        cb_ = out["cpu_baseline"]
        print(f"[bench] stage4 cpu baseline {cb_['value']:.0f} reads/s on {T} threads, "
              f"{cb_['single_thread_value']:.0f} on 1 (C driver); parity: --gpu "
              f"{out['parity_vs_oracle_sample']} (first {m} reads vs the Python driver), "
              f"cpu mode {out['cpu_mode']['parity_vs_c_driver_all_reads']} (all {n} reads "
              "vs the C driver)", file=sys.stderr, flush=True)
    genome.close()
    out["plain_mem_c1"] = bench_c1(pkg, ctx, args)
    return out


def bench_c1(pkg, ctx, args):
    """The plain `mem` configuration's reads through the stage-4 driver: a
    1 Mbp contig from the reference's test LCG, 10,000 single-end 150 bp reads
    (1% substitutions, 0.1% 1-3 bp indels, truth in the name), one chain per
    read from its truth span, band 32; reads/s per batch, truth recall of the
    first record, parity with the Python restatement, and the restated
    reference CPU path (the C driver) on the same reads."""
    from tests.stage4_data import make_c1
    from oracle import stage4 as S
    t0 = time.time()
    b = make_c1(args.c1_reads)
    print(f"[bench] C1 (plain mem): {len(b['reads'])} reads generated in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    cfg = pkg.stage4_cfg(bandwidth=32)
    x = pkg.Stage4Inputs(b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                         b["repeat_min_occ"])

    def run():
        return pkg.stage4_run(ctx, genome, x, cfg=cfg)

    t, res = median_of_runs(run, 5)
    got = res.alignments()
    res.close()
    n = len(b["reads"])
    hit = 0
    for name, recs in zip(b["names"], got):
        s0, e0 = map(int, name.rsplit(":", 1)[1].split("-"))
        hit += bool(recs) and recs[0]["ref_start"] < e0 and recs[0]["ref_end"] > s0
    out = {"workload": "plain mem: 1 Mbp contig (reference test LCG), "
                       f"{n} x 150 bp single-end reads, 1% subs + 0.1% 1-3 bp indels, "
                       "truth-span chains, band 32, mem defaults, --gpu configuration",
           "reads": n, "ms_per_batch": t * 1e3, "reads_per_s": n / t,
           "truth_recall": hit / n}
    if not args.no_cpu:
        want, _ = S.run(b["contigs"], b["reads"], b["chains"], S.default_cfg(bandwidth=32),
                        b["repeat_min_occ"])
        out["parity_vs_oracle"] = got == want
        # the restated reference CPU path (the C driver without the GPU
        # dispatcher) on every CPU of this process, with the device in the
        # same configuration checked against it
        T = len(os.sched_getaffinity(0))
        ccfg = S.default_cfg(bandwidth=32, cpu_mode=True)
        cb = S.CBatch(b["genome"], b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                      b["repeat_min_occ"])
        tc, _ = median_of_runs(lambda: S.run_c(cb, ccfg, T, decode=False))
        tc1, _ = median_of_runs(lambda: S.run_c(cb, ccfg, 1, decode=False))
        want_cpu, _ = S.run_c(cb, ccfg, T)
        res = pkg.stage4_run(ctx, genome, x, cfg=pkg.stage4_cfg(bandwidth=32, cpu_mode=True))
        out["cpu_mode_parity_vs_c_driver"] = res.alignments() == want_cpu
        res.close()
        out["cpu_baseline"] = {"value": n / tc, "unit": "reads/s", "cores": T, "kind": "port",
                               "label": "restated reference CPU path",
                               "single_thread_value": n / tc1,
                               "sample": f"all {n} reads through the stage-4 driver restated "
                                         "in C without the GPU dispatcher "
                                         f"(oracle/kira_stage4_oracle.c), {T} threads and 1 "
                                         "thread; 1 warm-up + median of 5"}
    print(f"[bench] C1: {out['ms_per_batch']:.1f} ms per batch, {out['reads_per_s']:.0f} "
          f"reads/s, recall {out['truth_recall']:.4f}, parity {out.get('parity_vs_oracle')} "
          f"/ cpu mode {out.get('cpu_mode_parity_vs_c_driver')}",
          file=sys.stderr, flush=True)
    genome.close()
    return out


def make_requests(work, lw, n_req, seed=5, max_reads=50_000):
    """Request stream: batch sizes U{1..50k} reads, 80% short (150 bp, 16
    candidate windows each; one banded SW job (w = 32) on the true window,
    dropped when that window is certified) and 20% long reads (each 4
    chain-gap windows of the long-window set, one banded SW job per window,
    band 5%)."""
    rng = np.random.default_rng(seed)
    L, W, K = work["read_len"], work["W"], work["windows"]
    n_short_pool = work["n_reads"]
    reqs = []
    s_cur, l_cur = 0, 0
    pad = 1024
    for _ in range(n_req):
        n = int(rng.integers(1, max_reads + 1))
        ns = max(1, int(round(0.8 * n)))
        nl = n - ns
        idx = (s_cur + np.arange(ns)) % n_short_pool
        s_cur += ns
        short = work["reads"][idx]                        # [ns, L]
        nwin = 4 * nl
        widx = (l_cur + np.arange(nwin)) % lw["n"]
        l_cur += nwin
        lens_l = lw["read_len"][widx]
        offs_l = lw["read_off"][widx]
        long_reads = [lw["reads"][o:o + m] for o, m in zip(offs_l, lens_l)]
        flat = np.concatenate([short.reshape(-1)] + long_reads) if nwin else short.reshape(-1)
        lens = np.concatenate([np.full(ns, L, np.int32), lens_l.astype(np.int32)])
        offs = np.concatenate([[0], np.cumsum(lens[:-1], dtype=np.int64)]).astype(np.int64)
        sieve_read = np.repeat(np.arange(ns, dtype=np.int32), K)
        win = work["win_start"].reshape(-1, K)[idx].reshape(-1)
        sieve = (sieve_read, win, np.full(ns * K, W, np.int32))
        true_w = win.reshape(ns, K)[:, 0]
        sw_read = np.concatenate([np.arange(ns, dtype=np.int32),
                                  ns + np.arange(nwin, dtype=np.int32)])
        sw_start = np.concatenate([true_w, lw["ref_off"][widx] - pad]).astype(np.int64)
        sw_len = np.concatenate([np.full(ns, W, np.int32), lw["ref_len"][widx]])
        sw_off = np.concatenate([np.full(ns, 64, np.int32), lw["offset"][widx]])
        sw_band = np.concatenate([np.full(ns, 32, np.int32), lw["band"][widx]])
        sw_abort = np.full(ns + nwin, -(2 ** 31) // 8, np.int32)
        sw_skip = np.concatenate([np.arange(ns, dtype=np.int32) * K,
                                  np.full(nwin, -1, np.int32)])
        reqs.append(dict(flat=flat, offs=offs, lens=lens, sieve=sieve,
                         sw=(sw_read, sw_start, sw_len, sw_off, sw_band, sw_abort, sw_skip),
                         n_reads=ns + nl, cells_hint=int(nwin)))
    return reqs


def serve_stream(pkg, ctx, genome, reqs, arrivals, *, slots, budget, graph):
    """Submit each request at its arrival time (seconds from the start; None =
    all at once) and wait for all; returns latencies (s) and the span."""
    srv = pkg.Server(ctx, genome, slots=slots, hbm_budget_bytes=budget,
                     host_staging_bytes=8 << 30, use_graph=graph,
                     cfg=pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5), sieve_limit=8, cert_max_mism=4)
    t0 = time.perf_counter()
    tickets = []
    for k, q in enumerate(reqs):
        if arrivals is not None:
            dt = t0 + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        tickets.append(srv.submit(q["flat"], q["offs"], q["lens"], sieve=q["sieve"],
                                  sw=q["sw"]))
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    lat = np.array([r["done_s"] - r["submit_s"] for r in res])
    span = max(r["done_s"] for r in res) - min(r["submit_s"] for r in res)
    skipped = int(sum(int(r["sw_skipped"].sum()) for r in res))
    return lat, span, st, skipped, res


def bench_serving(pkg, ctx, genome, work, lw, args):
    """Poisson arrivals of read batches into the continuous-batching server;
    p50 / p99 batch latency (submit -> results on the host) and reads/s,
    against the plain form (one batch on the device at a time)."""
    n_req = args.serve_requests
    t0 = time.time()
    reqs = make_requests(work, lw, n_req)
    total_reads = sum(q["n_reads"] for q in reqs)
    print(f"[bench] serving: {n_req} requests, {total_reads} reads generated in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    budget = args.serve_budget_gb << 30
    # warm-up (first launches, pool first touch)
    serve_stream(pkg, ctx, genome, reqs[:2], None, slots=4, budget=budget, graph=False)
    # saturated capacity: every request queued at once, continuous (slots in
    # flight) and plain (one request on the device at a time)
    lat, span, st, _, _ = serve_stream(pkg, ctx, genome, reqs, None, slots=args.serve_slots,
                                       budget=budget, graph=False)
    capacity = total_reads / span
    _, span1, _, _, _ = serve_stream(pkg, ctx, genome, reqs, None, slots=1, budget=budget,
                                     graph=False)
    capacity_plain = total_reads / span1
    print(f"[bench] serving capacity (all queued at once): continuous {capacity:.0f} reads/s "
          f"(peak in flight {st['peak_in_flight']}), plain {capacity_plain:.0f} reads/s, "
          f"ratio {capacity / capacity_plain:.2f}", file=sys.stderr, flush=True)
    rate = args.serve_load * capacity / (total_reads / n_req)  # requests per second
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    arrivals -= arrivals[0]
    out = {"workload": "Poisson arrivals of 1-50k-read batches (80% 150 bp short reads x 16 "
                       "windows + banded SW on the true window unless certified; 20% long "
                       "reads x 4 chain-gap windows, band 5%), resident genome, HBM-budgeted "
                       "admission",
           "requests": n_req, "reads": total_reads, "capacity_reads_per_s": capacity,
           "capacity_plain_reads_per_s": capacity_plain,
           "capacity_continuous_vs_plain": capacity / capacity_plain,
           "load": args.serve_load, "arrival_rate_per_s": rate,
           "hbm_budget_gib": args.serve_budget_gb, "slots": args.serve_slots}
    modes = [("continuous", args.serve_slots, False), ("continuous_graph", args.serve_slots, True),
             ("plain_one_batch_at_a_time", 1, False)]
    results = {}
    for name, slots, graph in modes:
        lat, span, st, skipped, res = serve_stream(pkg, ctx, genome, reqs, arrivals,
                                                   slots=slots, budget=budget, graph=graph)
        results[name] = res
        out[name] = {"p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "mean_ms": float(lat.mean() * 1e3), "reads_per_s": total_reads / span,
                     "peak_in_flight": st["peak_in_flight"],
                     "peak_device_gib": st["peak_device_bytes"] / 2 ** 30,
                     "sw_jobs_dropped_as_certified": skipped}
        print(f"[bench] serving {name}: p50 {out[name]['p50_ms']:.1f} ms, p99 "
              f"{out[name]['p99_ms']:.1f} ms, {out[name]['reads_per_s']:.0f} reads/s, peak in "
              f"flight {st['peak_in_flight']}", file=sys.stderr, flush=True)
    # the same request gives the same results in every mode
    same = True
    base = results["plain_one_batch_at_a_time"]
    for name in ("continuous", "continuous_graph"):
        for a, b in zip(base, results[name]):
            for k in ("shift", "best", "second", "flags", "sw_skipped"):
                same &= bool(np.array_equal(a[k], b[k]))
            same &= a["sw"] == b["sw"]
    out["results_identical_across_modes"] = same
    return out


# kira_banded_sw_plan_job_class ids: lane-per-job band sizes (int32 / int16
# storage), then the cooperative (P, G) shapes
SW_CLASS_NAMES = [f"lane C={c}{' i16' if p else ''}" for c in (9, 17, 33, 65, 129, 257)
                  for p in (0, 1)] + ["coop (5,16)", "coop (9,16)", "coop (17,16)",
                                      "coop (13,32)", "coop (17,32)", "coop (6,16) i16",
                                      "coop (10,16) i16", "coop (18,16) i16",
                                      "coop (14,32) i16", "coop (18,32) i16",
                                      "coop (10,8) i16", "coop (18,8) i16"]
SW_OPS_PER_CELL = 16  # E, F, H max chains + substitution select, without trace bookkeeping


def run_sw_batch(pkg, ctx, sw, args, stream, steps, cpu_pairs, label, sample_from=0,
                 cpu_timing=True):
    """One device-resident banded SW batch (plan + timed run_dev launches);
    GCUPS over the cells the reference algorithm evaluates; oracle parity
    (score, end, cells) on `cpu_pairs` jobs from job `sample_from`, and, with
    cpu_timing, the oracle's CPU rate on them (all the process's CPUs and one
    thread, 1 warm-up + median of 5)."""
    import torch
    dev = torch.device("cuda:0")
    n = sw["n"]
    cfg = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    kn = pkg.sw_knobs()
    t0 = time.time()
    plan = pkg.BandedSwPlan(ctx, sw["read_len"], sw["band"], cfg)
    t_plan = time.time() - t0
    routes = plan.route_counts()
    job_class = plan.job_class()
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d = dict(reads=T(sw["reads"]), read_off=T(sw["read_off"]), read_len=T(sw["read_len"]),
             refs=T(sw["refs"]), ref_off=T(sw["ref_off"]), ref_len=T(sw["ref_len"]),
             offset=T(sw["offset"]), band=T(sw["band"]), abort_score=T(sw["abort"]))
    res = torch.empty((n, 10), dtype=torch.int32, device=dev)
    cig = torch.empty(max(plan.cigar_words, 1), dtype=torch.int32, device=dev)

    def run():
        pkg.banded_sw_run_dev(ctx, plan, res=res, cigar=cig, cfg=cfg, knobs=kn, stream=stream,
                              **d)

    run()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(stream)
        run()
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000 / steps
    ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    r = res.cpu().numpy()
    cells = int(r[:, 8].astype(np.int64).sum())
    # cells and jobs per kernel class, in launch order (one dispatch each)
    per_class = {}
    for c in np.unique(job_class):
        m = job_class == c
        per_class[SW_CLASS_NAMES[int(c)]] = {"jobs": int(m.sum()),
                                             "cells": int(r[m, 8].astype(np.int64).sum())}
    gcups = cells / (ms * 1e-3) / 1e9
    ops = cells * SW_OPS_PER_CELL / (ms * 1e-3) / 1e12
    print(f"[bench] {label}: {n} jobs, {ms:.2f} ms/batch ({wall_ms:.2f} wall), "
          f"{cells / n:.0f} cells/job, {gcups:.1f} GCUPS, plan {plan.device_bytes / 2 ** 30:.1f} "
          f"GiB, routes {routes}",
          file=sys.stderr, flush=True)
    out = {"jobs": n, "cells": cells, "ms_per_batch": ms, "wall_ms_per_batch": wall_ms,
           "gcups": gcups, "plan_ms": t_plan * 1000,
           "plan_device_gib": plan.device_bytes / 2 ** 30, "routes": routes,
           "classes": per_class,
           "steps": steps,
           "roofline": {"bound": "valu", "achieved": ops, "peak": VALU_PEAK_TOPS,
                        "unit": "Tops/s", "frac": ops / VALU_PEAK_TOPS,
                        "ops_per_cell": SW_OPS_PER_CELL},
           "parity_vs_oracle_sample": None, "cpu_baseline": None}
    del res, cig, d
    plan.close()
    torch.cuda.empty_cache()
    if not args.no_cpu and cpu_pairs > 0:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        m0 = min(sample_from, n)
        m = min(cpu_pairs, n - m0)
        threads = len(os.sched_getaffinity(0))
        oc = O.cfg(1, 4, 6, 1, 32, 100, 5)

        def job_args(a, b):
            sl = slice(a, b)
            return [sw["reads"], sw["read_off"][sl], sw["read_len"][sl], sw["refs"],
                    sw["ref_off"][sl], sw["ref_len"][sl], sw["offset"][sl], sw["abort"][sl]]

        full = job_args(m0, m0 + m)
        run_all = lambda: O.banded_sw_batch_mt(*full, oc, bands=sw["band"][m0:m0 + m],
                                               threads=threads)
        if cpu_timing:
            t, (sc, en, cc) = median_of_runs(run_all)
            m1 = max(1, m // 16)
            one = job_args(m0, m0 + m1)
            t1, (_, _, c1) = median_of_runs(
                lambda: O.banded_sw_batch_mt(*one, oc, bands=sw["band"][m0:m0 + m1], threads=1))
            out["cpu_baseline"] = {
                "value": float(cc.sum()) / t / 1e9, "unit": "GCUPS", "cores": threads,
                "kind": "port", "single_thread_value": float(c1.sum()) / t1 / 1e9,
                "sample": f"jobs {m0}..{m0 + m} of the same batch (scalar banded_sw_with, "
                          f"pthreads over jobs, {threads} threads = the CPUs this process may "
                          f"use; the first {m1} on 1 thread), 1 warm-up + median of 5"}
            print(f"[bench] {label} cpu baseline {cc.sum() / t / 1e9:.2f} GCUPS on {threads} "
                  f"threads, {c1.sum() / t1 / 1e9:.2f} on 1", file=sys.stderr, flush=True)
        else:
            sc, en, cc = run_all()
        out["parity_vs_oracle_sample"] = bool(
            np.array_equal(sc, r[m0:m0 + m, 0]) and np.array_equal(en, r[m0:m0 + m, 2]) and
            np.array_equal(cc, r[m0:m0 + m, 8].astype(np.int64)))
        print(f"[bench] {label} parity on jobs {m0}..{m0 + m}: "
              f"{out['parity_vs_oracle_sample']}", file=sys.stderr, flush=True)
    return out


def concat_sw(a, b):
    """One queue holding the jobs of a then b (both windows into the same refs)."""
    out = {"n": a["n"] + b["n"], "refs": a["refs"]}
    for k in ("read_len", "ref_off", "ref_len", "offset", "band", "abort"):
        out[k] = np.concatenate([a[k], b[k]])
    out["reads"] = np.concatenate([a["reads"], b["reads"]])
    out["read_off"] = np.concatenate([a["read_off"], b["read_off"] + len(a["reads"])])
    return out


# Shape buckets of the banded SW autotuner (include/kira_hip.h): band buckets
# end at these bands, read lengths bucket by log2, group sizes by log4.
SW_BAND_EDGES = [4, 8, 16, 32, 39, 47, 64, 71, 79, 128, 135, 143, 207, 223, 271]


def sw_tune_shapes(*workloads):
    """One (band, read_len, n_jobs) shape per (band bucket, length bucket) group
    of each workload's jobs (each workload is one batch): its largest band, its
    median read length, and the number of the batch's jobs in its band bucket."""
    shapes = []
    for sw in workloads:
        bb = np.searchsorted(np.array(SW_BAND_EDGES), sw["band"], side="left")
        lb = np.floor(np.log2(np.maximum(sw["read_len"], 1))).astype(np.int64)
        key = bb * 64 + lb
        for k in np.unique(key):
            m = key == k
            shapes.append((int(sw["band"][m].max()), int(np.median(sw["read_len"][m])),
                           int((bb == k // 64).sum())))
    return shapes


def bench_sw(pkg, ctx, genome, args, stream, lw):
    """Uneven banded SW (1M pairs, L U{100..300}, band U{16,32,64,128}, adaptive,
    x-drop 100, no abort) in one device-resident batch; long-read chain-gap
    windows (500-5000 bp, band 5%); and both in one queue.  Each runs first on
    the built-in routing, then on the routes the autotuner picks for the
    workloads' shapes (the headline figures); with --no-sw-tune only on the
    built-in routing."""
    t0 = time.time()
    sw = make_sw_workload(genome, n_pairs=args.sw_pairs)
    print(f"[bench] SW workload: {sw['n']} pairs generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    pkg.sw_tune_clear(ctx)
    mixed = concat_sw(sw, lw) if lw is not None else None
    tune = not args.no_sw_tune
    if tune:
        builtin = {"short": run_sw_batch(pkg, ctx, sw, args, stream, args.sw_steps, 0,
                                         "SW short pairs (built-in routing)")}
        if lw is not None:
            builtin["long"] = run_sw_batch(pkg, ctx, lw, args, stream, args.sw_steps, 0,
                                           "SW long windows (built-in routing)")
        shapes = sw_tune_shapes(*[x for x in (sw, lw, mixed) if x is not None])
        t0 = time.time()
        cache = args.sw_tune_cache or None
        # one shape a call, so that each finished shape is reported as it ends
        tuned = []
        for k, shape in enumerate(shapes):
            t, = pkg.sw_autotune(ctx, [shape], pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5),
                                 pkg.sw_knobs(), cache_path=cache)
            tuned.append(t)
            print(f"[bench] SW autotune {k + 1}/{len(shapes)}: band {shape[0]}, read "
                  f"{shape[1]}, {shape[2]} jobs -> {t['route']} ({t['ms_best']:.2f} ms, "
                  f"built-in {t['ms_builtin']:.2f} ms, {t['candidates']} timed)",
                  file=sys.stderr, flush=True)
        t_tune = time.time() - t0
        chosen = [dict(band=b, read_len=l, jobs=n, route=t["route"], candidates=t["candidates"],
                       ms_best=t["ms_best"], ms_builtin=t["ms_builtin"])
                  for (b, l, n), t in zip(shapes, tuned)]
        print(f"[bench] SW autotune: {len(shapes)} shapes in {t_tune:.1f}s; "
              f"{sum(c['route'] != 'built-in' for c in chosen)} re-routed", file=sys.stderr,
              flush=True)
    out = run_sw_batch(pkg, ctx, sw, args, stream, args.sw_steps, args.cpu_sample_pairs,
                       "SW short pairs")
    if tune:
        out["builtin_routing"] = {k: builtin["short"][k] for k in ("ms_per_batch", "gcups",
                                                                    "routes")}
        out["tuned_vs_builtin"] = out["gcups"] / builtin["short"]["gcups"]
        out["autotune"] = {"shapes": chosen, "seconds": t_tune, "cache": cache}
    else:
        out["autotune"] = None
    out["workload"] = ("uneven banded SW: 1M pairs, L U{100..300}, band U{16,32,64,128}, "
                       "1-5% subs + 0.5% indels, adaptive band, x-drop 100, int16/int32 "
                       "routed per job")
    out["pairs"] = out["jobs"]
    if lw is not None:
        lo = run_sw_batch(pkg, ctx, lw, args, stream, args.sw_steps, 256, "SW long windows")
        if tune:
            lo["builtin_routing"] = {k: builtin["long"][k] for k in ("ms_per_batch", "gcups",
                                                                      "routes")}
            lo["tuned_vs_builtin"] = lo["gcups"] / builtin["long"]["gcups"]
        lo["workload"] = ("long-read chain-gap windows: 500-5000 bp, band ceil(5% len) "
                          "(25-250), 5% subs, 2.5% ins, 2.5% del, exact int32 (int16 storage only where proven)")
        out["long_windows"] = lo
        mx = run_sw_batch(pkg, ctx, mixed, args, stream, args.sw_steps, 512, "SW mixed queue",
                          sample_from=sw["n"] - 256, cpu_timing=False)
        mx["workload"] = "the short pairs and the long windows in one plan / one queue"
        out["mixed_queue"] = mx
    return out


if __name__ == "__main__":
    main()
