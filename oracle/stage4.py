"""CPU restatement of the stage-4 alignment driver (test infrastructure only).

Follows src/pipeline/stage4_alignment.rs run() (:959-1268) and
process_read_prefilter (:535-957) in the configuration the GPU dispatcher is
active (`--gpu`), or with cpu_mode the one without it (fallback chains try
the fast DP cascade of oracle/fastpath.py before SW; no sieve dispatch);
single-end, with the two-tier candidate reorder of uncertain short reads
(:346-451, KIRA_TWOTIER, on by default; off with twotier=False) and
KIRA_TWOTIER_MAPQ unset (its default).  Supplementary alignments follow
src/pipeline/split_read.rs:22-151.  Every numeric step goes through the C
oracle (oracle.py): prefilter_chain, the 2-bit sieve + certification,
align_batch_simd's SW, banded SW and build_alignment.

Pinned by: the reference's split-read unit tests (tests/unit/
pipeline_split_read.rs) for the supplementary picks, hand-computed rule
cases (chain ratio filter, top-k, exact shortcut, XS proxy, job routing) in
tests/test_oracle_stage4.py, and the pinned building blocks above.
"""
import math

from oracle import fastpath as F
from oracle import oracle as O

I32_MIN = -(2 ** 31)
I32_MAX = 2 ** 31 - 1
LANES = {"avxvnni": 32, "avx2": 16, "neon": 4, "scalar": 1}

# split_read.rs:10-19
SUPP_MIN_SCORE_RATIO = 0.5
SUPP_MIN_READ_SPAN = 50
SUPP_MAX_OVERLAP_PCT = 50
MAX_SUPP_PER_READ = 4


def default_cfg(**kw):
    """PipelineConfig defaults of `mem` (src/cli/commands/mem.rs:183-205) plus
    the environment knobs stage 4 reads (defaults of their getters)."""
    c = dict(match_score=1, mismatch=4, gap_open=6, gap_extend=1, bandwidth=32, xdrop=100,
             clip_penalty=5, adaptive=True, max_drift=16, dp_xdrop=-1,
             min_chain_ratio=0.4, accept_enable=True, accept_only_top1=True,
             accept_span_slack=15, accept_min_identity=98.5, accept_max_mismatches=5,
             accept_require_score_margin=0, dp_topk=1, dp_abort_margin=20,
             long_read_threshold=500, max_alignments=1, tail_ext=True, ungap_min_id=-1,
             short_dp_topk=0, ambig_div=5, xs_min_ratio_pct=0, spectral_pct=3,
             spectral_floor=2, cert_strict=False, simd="avx2", supplementary=True,
             cpu_mode=False, fast={}, twotier=True, twotier_k=5, ungap_gap_check=False)
    c.update(kw)
    return c


def _f32(x):
    import numpy as np
    return float(np.float32(x))


def reverse_complement(s: bytes) -> bytes:
    """src/seq.rs:1-17: A/C/G/T either case -> upper complement, else N."""
    comp = {65: 84, 97: 84, 67: 71, 99: 71, 71: 67, 103: 67, 84: 65, 116: 65}
    return bytes(comp.get(b, 78) for b in reversed(s))


def _sat_i32(v):
    return max(I32_MIN, min(I32_MAX, v))


def _wrap_i32(v):
    v &= 0xFFFFFFFF
    return v - 2 ** 32 if v & 0x80000000 else v


def original_interval(start, end, is_rev, read_len):
    """split_read.rs:128-134 (u32 saturating_sub)."""
    if is_rev:
        return max(read_len - end, 0), max(read_len - start, 0)
    return start, end


def read_overlap_pct(a0, a1, b0, b1):
    """split_read.rs:137-148."""
    al, bl = max(a1 - a0, 0), max(b1 - b0, 0)
    if al == 0 or bl == 0:
        return 0
    s, e = max(a0, b0), min(a1, b1)
    if e <= s:
        return 0
    return min((e - s) * 100, 0xFFFFFFFF) // al


def pick_supplementary_chains(primary, chains):
    """split_read.rs:60-125.  primary: record dict (read_start, read_end,
    is_rev, cigar text); chains: the read's chain tuples.  Returns indices."""
    import re
    primary_score = max(chains[0][6], 1)
    min_score = math.ceil(_f32(_f32(primary_score) * _f32(SUPP_MIN_SCORE_RATIO)))
    read_len = sum(int(n) for n, op in re.findall(r"(\d+)([MIDNS])", primary["cigar"])
                   if op in "MIS")
    prs, pre = original_interval(primary["read_start"], primary["read_end"], primary["is_rev"],
                                 read_len)
    picks = []
    for ci in range(1, len(chains)):
        c = chains[ci]
        if c[6] < min_score:
            break
        if max(c[4] - c[3], 0) < SUPP_MIN_READ_SPAN:
            continue
        crs, cre = original_interval(c[3], c[4], c[5] == 1, read_len)
        if read_overlap_pct(prs, pre, crs, cre) > SUPP_MAX_OVERLAP_PCT:
            continue
        dup = False
        for pi in picks:
            p = chains[pi]
            ps, pe = original_interval(p[3], p[4], p[5] == 1, read_len)
            if read_overlap_pct(ps, pe, crs, cre) > SUPP_MAX_OVERLAP_PCT:
                dup = True
                break
        if dup:
            continue
        picks.append(ci)
        if len(picks) >= MAX_SUPP_PER_READ:
            break
    return picks


def select_chains(read_len, chains, repeat_min_occ, cfg, ranker=None):
    """The chain loop of process_read_prefilter (:549-708), mate guidance off.
    ranker(passing, min_score) runs two_tier_reorder (:346-451) over the
    passing chain indices and returns (scores dict, cached dict, dp_limit) or
    None.  Returns (selected chain indices, chain_best, chain_second,
    score_margin, cached fast results of the selected chains, competitor
    score or None)."""
    if not chains:
        return [], 0, 0, I32_MAX, {}, None
    best = chains[0][6]
    # (best as f32 * min_chain_ratio).ceil() as i32 (:567-571)
    min_score = int(math.ceil(_f32(_f32(best) * _f32(cfg["min_chain_ratio"])))) if best > 0 \
        else best
    second = chains[1][6] if len(chains) > 1 else 0
    margin = best - second if len(chains) > 1 else I32_MAX
    short_read = read_len <= 300
    div = cfg["ambig_div"]
    q = int(best / div)  # i32 division truncates toward zero
    ambiguous = len(chains) > 1 and (margin <= max(q, 8) or repeat_min_occ > 1)
    weak = len(chains) > 1 and best < int(read_len * 2 / 3)
    twotier = cfg["twotier"]
    if short_read and weak and twotier:
        min_score = min(min_score, int(best / 5))
    override = cfg["short_dp_topk"] if cfg["short_dp_topk"] >= 1 else None
    if short_read:
        topk = override if override is not None else (2 if ambiguous else 1)
    else:
        topk = max(cfg["dp_topk"], 1)
    order = list(range(len(chains)))
    cached, competitor = {}, None
    if twotier and short_read and (ambiguous or weak) and ranker is not None:
        passing = [ci for ci, c in enumerate(chains) if c[6] >= min_score][:cfg["twotier_k"]]
        if len(passing) >= 2:
            scores, cached, dp_limit, order = ranker(passing)
            if override is None:
                topk = dp_limit
            if topk == 1:
                others = [s for ci, s in scores.items() if ci != order[0]]
                competitor = max(others) if others else None
    sel = []
    for ci in order:
        if chains[ci][6] < min_score:
            continue
        if len(sel) >= topk:
            break
        sel.append(ci)
    return sel, best, second, margin, {ci: cached[ci] for ci in sel if ci in cached}, competitor


def two_tier_reorder(read_fwd, chains, contigs, c, router, passing):
    """two_tier_reorder (:346-451) on the passing chains: exact copies score
    A * L, the rest the fast cascade's alignment score (its result is kept);
    order by score (ties by chain index), then the other chains in order."""
    rev = None
    scored, cached = [], {}
    for ci in passing:
        ch = chains[ci]
        is_rev = ch[5] == 1
        if is_rev and rev is None:
            rev = reverse_complement(read_fwd)
        read = rev if is_rev else read_fwd
        contig = contigs[ch[0]]
        if exact_match_exists(read, contig, (ch[1], ch[2], ch[3], ch[4])):
            scored.append((ci, _wrap_i32(c.match_score * len(read))))
            continue
        hit = F.fast_dp(read, contig, ch[1], ch[3], ch[6], c, router)
        if hit is not None:
            cached[ci] = hit
        scored.append((ci, hit[1]["score"] if hit is not None else I32_MIN))
    scored.sort(key=lambda t: (-t[1], t[0]))
    scores = {ci: s for ci, s in scored if s != I32_MIN}
    margin = max(_sat_i32((c.match_score + c.mismatch) * 2), 1)
    if any(s == I32_MIN for _, s in scored):
        dp_limit = 2
    elif len(scored) >= 3 and scored[2][1] == scored[0][1]:
        dp_limit = 2
    else:
        top = _sat_i32(scored[0][1] - margin)
        n_close = 0
        for _, s in scored:
            if s < top:
                break
            n_close += 1
        dp_limit = min(max(n_close, 1), 3)
    order = [ci for ci, _ in scored] + [ci for ci in range(len(chains))
                                        if ci not in {x for x, _ in scored}]
    return scores, cached, dp_limit, order


def exact_match_exists(read, contig, span):
    """mod.rs:818-832."""
    ref_start, ref_end, read_start, read_end = span
    n = len(read)
    if read_start != 0 or read_end != n or ref_end - ref_start != n:
        return False
    if ref_start + n > len(contig):
        return False
    return contig[ref_start:ref_start + n] == read


def _record(kind, ref_id, ref_start, ref_end, read_start, read_end, cigar, score, mapq, is_rev,
            nm, md, supplementary=False):
    return dict(kind=kind, ref_id=ref_id, ref_start=ref_start, ref_end=ref_end,
                read_start=read_start, read_end=read_end, cigar=cigar, score=score, mapq=mapq,
                is_rev=bool(is_rev), is_supplementary=supplementary, nm=nm, md=md,
                xs_score=None)


def _identical_nm(bases):
    return sum(1 for b in bases if chr(b).upper() not in "ACGT=")


def _sw_cfg(cfg):
    return O.cfg(cfg["match_score"], cfg["mismatch"], cfg["gap_open"], cfg["gap_extend"],
                 cfg["bandwidth"], cfg["xdrop"], cfg["clip_penalty"])


def _knobs(cfg):
    return O.knobs(cfg["adaptive"], cfg["max_drift"], cfg["dp_xdrop"])


def _exact_record(read, span, ref_id, is_rev, cfg):
    """exact_match_alignment (mod.rs:778-815)."""
    n = len(read)
    return _record("accepted_ungapped", ref_id, span[0], span[0] + n, 0, n, f"{n}M",
                   cfg["match_score"] * n, 0, is_rev, _identical_nm(read), str(n))


def _clamp_window(ref_len, ref_start, ref_end, bandwidth):
    """mod.rs:1766-1771."""
    flank = max(max(bandwidth, 1) * 2, 50)
    start = max(ref_start - flank, 0)
    end = min(ref_end + flank, ref_len)
    return start, max(end, start + 1)


def _scalar_alignments(jobs, contigs, cfg, supplementary=False):
    """align_chain_with_meta (mod.rs:835-883) for (read_seq, is_rev, ref_id,
    span, abort) jobs: exact shortcut, else banded SW on the clamped window."""
    c, kn = _sw_cfg(cfg), _knobs(cfg)
    out = []
    for read, is_rev, ref_id, span, abort in jobs:
        contig = contigs[ref_id]
        if exact_match_exists(read, contig, span):
            r = _exact_record(read, span, ref_id, is_rev, cfg)
            r["is_supplementary"] = supplementary
            out.append((r, False))
            continue
        ws, we = _clamp_window(len(contig), span[0], span[1], cfg["bandwidth"])
        win = contig[ws:we]
        off = span[0] - ws - span[2]
        sw = O.banded_sw(read, win, off, c, abort, kn)
        b = O.build_alignment(read, win, ws, sw)
        out.append((_record("dp_aligned", ref_id, b["ref_start"], b["ref_end"],
                            b["read_start"], b["read_end"], b["cigar"], b["score"], 0, is_rev,
                            b["nm"], b["md"], supplementary), sw["early_abort"]))
    return out


def run(contigs, reads, chains, cfg=None, repeat_min_occ=None):
    """stage4_alignment::run restated.  contigs: list of bytes; reads: list of
    bytes; chains: per read, list of (ref_id, ref_start, ref_end, read_start,
    read_end, strand 0/1, score) in stage-3 order.  Returns (alignments per
    read as lists of record dicts, stats dict)."""
    cfg = cfg or default_cfg()
    n = len(reads)
    repeat_min_occ = repeat_min_occ or [0] * n
    lanes = LANES[cfg["simd"]]
    c = _sw_cfg(cfg)
    accept_allowed = cfg["accept_enable"] and cfg["max_alignments"] <= 1
    multi = cfg["max_alignments"] > 1
    st = dict(reads=n, chains_total=0, chains_used=0, exact_matches=0, prefilter_accept=0,
              prefilter_reject=0, prefilter_fallback=0, reason_counts=[0] * 10,
              gpu_spectral_resolved=0, dp_simd=0, dp_scalar=0, dp_early_abort=0, dp_reads=0,
              supplementary=0, unmapped=0, packed_spectral_resolved=0, wfa_resolved=0)
    alns = [[] for _ in range(n)]
    simd_jobs, scalar_jobs = [], []
    router = F.Router(**cfg["fast"])
    cb, cs = [0] * n, [0] * n
    competitor = [None] * n
    for r in range(n):
        read_fwd = reads[r]
        L = len(read_fwd)
        bucket = 0 if L <= 120 else (1 if L <= 150 else 2)
        cl = chains[r]
        st["chains_total"] += len(cl)
        ranker = (lambda passing, rf=read_fwd, cl=cl:
                  two_tier_reorder(rf, cl, contigs, c, router, passing))
        sel, cb[r], cs[r], margin, cached, competitor[r] = select_chains(
            L, cl, repeat_min_occ[r], cfg, ranker)
        dp_used = False
        for rank, ci in enumerate(sel):
            ch = cl[ci]
            st["chains_used"] += 1
            ref_id, is_rev = ch[0], ch[5] == 1
            span = (ch[1], ch[2], ch[3], ch[4])
            contig = contigs[ref_id]
            read = reverse_complement(read_fwd) if is_rev else read_fwd
            if exact_match_exists(read, contig, span):
                st["exact_matches"] += 1
                alns[r].append(_exact_record(read, span, ref_id, is_rev, cfg))
                continue
            if ci in cached:
                # a complete alignment the candidate search verified (:736-747)
                kind, a = cached[ci]
                st["packed_spectral_resolved" if kind == "packed" else "wfa_resolved"] += 1
                alns[r].append(_record("dp_aligned", ref_id, a["ref_start"], a["ref_end"], 0,
                                       len(read), O.cigar_str(a["cigar"]), a["score"], 0,
                                       is_rev, a["nm"], a["md"]))
                dp_used = True
                continue
            pf = O.prefilter_chain(
                read, contig, span, c, is_top1=rank == 0, accept_enable=accept_allowed,
                accept_only_top1=cfg["accept_only_top1"],
                accept_span_slack=cfg["accept_span_slack"],
                accept_min_identity=cfg["accept_min_identity"],
                accept_max_mismatches=cfg["accept_max_mismatches"],
                accept_require_score_margin=cfg["accept_require_score_margin"],
                score_margin=margin, long_read_threshold=cfg["long_read_threshold"],
                multi_alignments_enabled=multi, tail_ext=cfg["tail_ext"],
                ungap_min_id=cfg["ungap_min_id"], ungap_gap_check=cfg["ungap_gap_check"])
            reason = pf["reason"]
            st["reason_counts"][O.REASONS.index(reason) if reason in O.REASONS else 9] += 1
            m = pf.get("metrics")
            abort = _sat_i32(m["score"] - cfg["dp_abort_margin"]) if m else I32_MIN // 4
            if pf["result"] == "accept":
                st["prefilter_accept"] += 1
                a = pf["alignment"]
                alns[r].append(_record("accepted_ungapped", ref_id, a["ref_start"], a["ref_end"],
                                       a["read_start"], a["read_end"], a["cigar"], a["score"],
                                       60, is_rev, a["nm"], a["md"]))
                continue
            if pf["result"] == "reject":
                st["prefilter_reject"] += 1
                continue
            st["prefilter_fallback"] += 1
            dp_used = True
            if cfg["cpu_mode"]:
                # without the GPU dispatcher: try_fast_dp_alignment first
                # (:868-882), accept_fast_result (:289-299)
                fr = F.fast_dp(read, contig, ch[1], ch[3], ch[6], c, router)
                if fr is not None:
                    kind, a = fr
                    st["packed_spectral_resolved" if kind == "packed" else "wfa_resolved"] += 1
                    alns[r].append(_record("dp_aligned", ref_id, a["ref_start"], a["ref_end"], 0,
                                           len(read), O.cigar_str(a["cigar"]), a["score"], 0,
                                           is_rev, a["nm"], a["md"]))
                    continue
            if bucket < 2 and lanes > 1:
                w = O.build_simd_window(len(contig), ch[1], ch[3], L, cfg["bandwidth"])
                if w is not None:
                    simd_jobs.append(dict(read_idx=r, read=read, ref_id=ref_id, is_rev=is_rev,
                                          span=span, win_start=w[0],
                                          window=contig[w[0]:w[0] + w[1]], abort=abort))
                    continue
            scalar_jobs.append((r, read, is_rev, ref_id, span, abort))
        st["dp_reads"] += int(dp_used)

    # gpu_dispatch_spectral (:1272-1358), only with the dispatcher active
    kept = []
    for j in simd_jobs:
        if cfg["cpu_mode"]:
            kept.append(j)
            continue
        read, win = j["read"], j["window"]
        if not (O.pack(read)[1] and O.pack(win)[1]):
            kept.append(j)
            continue
        L = len(read)
        mx = O.spectral_max_mismatches(L, cfg["spectral_pct"], cfg["spectral_floor"])
        hit = O.scan_best_with_second(read, win)
        cert = False
        if hit is not None:
            shift, best, second = hit
            if best <= mx:
                seg = win[shift:shift + L]
                cert = len(seg) == L and O.certified(read, seg, best, second, mx,
                                                     cfg["cert_strict"])
        if cert:
            nm, md = O.ungapped_tags(read, seg)
            score = (L - best) * cfg["match_score"] - best * cfg["mismatch"]
            rs = j["win_start"] + shift
            alns[j["read_idx"]].append(_record("dp_aligned", j["ref_id"], rs, rs + L, 0, L,
                                               f"{L}M", score, 0, j["is_rev"], nm, md))
            st["gpu_spectral_resolved"] += 1
        else:
            kept.append(j)

    # lane groups: BTreeMap over (read_len, window_len), chunks of `lanes` (:1153-1186)
    buckets = {}
    for k, j in enumerate(kept):
        buckets.setdefault((len(j["read"]), len(j["window"])), []).append(k)
    kn = _knobs(cfg)
    simd_out = []
    for shape in sorted(buckets):
        idx = buckets[shape]
        for s in range(0, len(idx), lanes):
            chunk = [kept[k] for k in idx[s:s + lanes]]
            jobs = [O.Job(j["read"], j["window"], j["win_start"], j["span"][0], j["span"][1],
                          j["span"][2], j["span"][3], j["abort"]) for j in chunk]
            res = O.align_batch_sw(jobs, c, cfg["simd"], cfg["simd"], kn)
            for j, sw in zip(chunk, res):
                b = O.build_alignment(j["read"], j["window"], j["win_start"], sw)
                simd_out.append((j["read_idx"], _record(
                    "dp_aligned", j["ref_id"], b["ref_start"], b["ref_end"], b["read_start"],
                    b["read_end"], b["cigar"], b["score"], 0, j["is_rev"], b["nm"], b["md"]),
                    sw["early_abort"]))
    st["dp_simd"] = len(simd_out)
    st["dp_early_abort"] += sum(1 for x in simd_out if x[2])
    for r, rec, _ in simd_out:
        alns[r].append(rec)

    sc = _scalar_alignments([(j[1], j[2], j[3], j[4], j[5]) for j in scalar_jobs], contigs, cfg)
    st["dp_scalar"] = len(sc)
    st["dp_early_abort"] += sum(1 for x in sc if x[1])
    for j, (rec, _) in zip(scalar_jobs, sc):
        alns[j[0]].append(rec)

    # XS: the two-tier competitor's affine score, else the chain-score proxy
    # (:1224-1250)
    for r in range(n):
        if len(alns[r]) != 1:
            continue
        if competitor[r] is not None:
            alns[r][0]["xs_score"] = min(max(competitor[r], 0), max(alns[r][0]["score"], 0))
            continue
        if cb[r] <= 0 or cs[r] <= 0:
            continue
        first = alns[r][0]
        # (score as i64 * cs as i64 / cb as i64) as i32: truncating division,
        # then a wrapping cast
        prod = first["score"] * cs[r]
        q = abs(prod) // cb[r]
        sub = _wrap_i32(-q if prod < 0 else q)
        if sub > 0 and cs[r] * 100 >= cb[r] * cfg["xs_min_ratio_pct"]:
            first["xs_score"] = sub

    # emit_supplementary_alignments (split_read.rs:22-57)
    if cfg["supplementary"]:
        jobs, owner = [], []
        for r in range(n):
            if not alns[r] or len(chains[r]) < 2:
                continue
            for ci in pick_supplementary_chains(alns[r][0], chains[r]):
                ch = chains[r][ci]
                is_rev = ch[5] == 1
                read = reverse_complement(reads[r]) if is_rev else reads[r]
                jobs.append((read, is_rev, ch[0], (ch[1], ch[2], ch[3], ch[4]), I32_MIN // 8))
                owner.append(r)
        for r, (rec, _) in zip(owner, _scalar_alignments(jobs, contigs, cfg, True)):
            alns[r].append(rec)
            st["supplementary"] += 1

    st["unmapped"] = sum(1 for a in alns if not a)
    return alns, st


# ---------------------------------------------------------------- C driver
# kira_stage4_oracle.c: the same steps with cpu_mode, in C on pthreads (the
# CPU baseline timed beside the device driver).
import ctypes as _C

import numpy as _np

_CHAIN = _np.dtype([("ref_id", _np.uint32), ("ref_start", _np.uint32), ("ref_end", _np.uint32),
                    ("read_start", _np.uint32), ("read_end", _np.uint32),
                    ("strand", _np.int32), ("score", _np.int32)])
_SIMD = {"scalar": 0, "avx2": 1, "avxvnni": 2, "neon": 3}
_ALN_KINDS = {"accepted_ungapped": 0, "dp_aligned": 1}
_STATS = ("reads", "chains_total", "chains_used", "exact_matches", "prefilter_accept",
          "prefilter_reject", "prefilter_fallback")
_STATS_TAIL = ("gpu_spectral_resolved", "dp_simd", "dp_scalar", "dp_early_abort", "dp_reads",
               "supplementary", "unmapped", "packed_spectral_resolved", "wfa_resolved")


class _S4Cfg(_C.Structure):
    _fields_ = ([("sw", O.Cfg)] + [(k, _C.c_int32) for k in ("adaptive", "max_drift", "dp_xdrop")]
                + [("min_chain_ratio", _C.c_float)]
                + [(k, _C.c_int32) for k in ("accept_enable", "accept_only_top1",
                                              "accept_span_slack")]
                + [("accept_min_identity", _C.c_float)]
                + [(k, _C.c_int32) for k in (
                    "accept_max_mismatches", "accept_require_score_margin", "dp_topk",
                    "dp_abort_margin", "long_read_threshold", "max_alignments", "tail_ext",
                    "ungap_min_id", "short_dp_topk", "ambig_div", "xs_min_ratio_pct", "simd",
                    "supplementary", "twotier", "twotier_k", "ungap_gap_check")]
                + [("router", F._Router)])


class _S4In(_C.Structure):
    _fields_ = [("n_reads", _C.c_int32), ("genome", _C.c_void_p), ("contig_off", _C.c_void_p),
                ("contig_len", _C.c_void_p), ("reads", _C.c_void_p), ("read_off", _C.c_void_p),
                ("read_len", _C.c_void_p), ("repeat_min_occ", _C.c_void_p),
                ("chain_off", _C.c_void_p), ("chains", _C.c_void_p)]


class _S4Rec(_C.Structure):
    _fields_ = [(k, _C.c_int32) for k in ("read", "kind")] + \
        [(k, _C.c_uint32) for k in ("ref_id", "ref_start", "ref_end", "read_start", "read_end")] + \
        [(k, _C.c_int32) for k in ("score", "mapq", "is_rev", "is_supplementary")] + \
        [("nm", _C.c_uint32)] + \
        [(k, _C.c_int32) for k in ("has_xs", "xs_score", "n_cigar", "md_len")] + \
        [("cigar_off", _C.c_int64), ("md_off", _C.c_int64)]


class _S4Stats(_C.Structure):
    _fields_ = [(k, _C.c_int64) for k in _STATS] + [("reason_counts", _C.c_int64 * 10)] + \
        [(k, _C.c_int64) for k in _STATS_TAIL]


class _S4View(_C.Structure):
    _fields_ = [("n_reads", _C.c_int32), ("aln_off", _C.POINTER(_C.c_int64)),
                ("alns", _C.POINTER(_S4Rec)), ("cigar", _C.POINTER(_C.c_uint32)),
                ("md", _C.c_void_p), ("n_alns", _C.c_int64), ("cigar_words", _C.c_int64),
                ("md_bytes", _C.c_int64), ("stats", _S4Stats)]


def _c_lib():
    L = O.lib()
    if not getattr(L, "_s4_ready", False):
        L.ko_stage4_run_mt.restype = _C.c_void_p
        L.ko_stage4_run_mt.argtypes = [_C.POINTER(_S4In), _C.POINTER(_S4Cfg), _C.c_int]
        L.ko_stage4_view.argtypes = [_C.c_void_p, _C.POINTER(_S4View)]
        L.ko_stage4_free.argtypes = [_C.c_void_p]
        L._s4_ready = True
    return L


def c_cfg(cfg):
    """A default_cfg dict as the C driver's configuration (cpu_mode is the
    only mode it has)."""
    s = _S4Cfg()
    s.sw = _sw_cfg(cfg)
    for k in ("adaptive", "max_drift", "dp_xdrop", "min_chain_ratio", "accept_enable",
              "accept_only_top1", "accept_span_slack", "accept_min_identity",
              "accept_max_mismatches", "accept_require_score_margin", "dp_topk",
              "dp_abort_margin", "long_read_threshold", "max_alignments", "tail_ext",
              "ungap_min_id", "short_dp_topk", "ambig_div", "xs_min_ratio_pct", "supplementary",
              "twotier", "twotier_k", "ungap_gap_check"):
        setattr(s, k, type(getattr(s, k))(cfg[k]) if isinstance(getattr(s, k), float)
                else int(cfg[k]))
    s.simd = _SIMD[cfg["simd"]]
    s.router = F._router(F.Router(**cfg["fast"]))
    return s


class CBatch:
    """A batch as the flat arrays the C driver reads (built once, outside any
    timing): genome bytes with contig offsets / lengths, reads, per-read
    chain lists in stage-3 order, repeat_min_occ."""

    def __init__(self, genome, contig_off, contig_len, reads, chains, repeat_min_occ=None):
        n = len(reads)
        self.n = n
        self.g = _np.frombuffer(genome, _np.uint8)
        self.co = _np.ascontiguousarray(contig_off, _np.int64)
        self.cl = _np.ascontiguousarray(contig_len, _np.int64)
        self.rl = _np.array([len(r) for r in reads], _np.int32)
        self.ro = _np.zeros(n + 1, _np.int64)
        self.ro[1:] = _np.cumsum(self.rl, dtype=_np.int64)
        self.rb = _np.frombuffer(b"".join(reads) + b"\0" * 8, _np.uint8)
        self.coff = _np.zeros(n + 1, _np.int64)
        self.coff[1:] = _np.cumsum([len(c) for c in chains], dtype=_np.int64)
        flat = [c for cs in chains for c in cs]
        self.ch = _np.array(flat, _CHAIN) if flat else _np.zeros(1, _CHAIN)
        self.rmo = _np.ascontiguousarray(repeat_min_occ if repeat_min_occ is not None
                                         else [0] * n, _np.int32)
        P = lambda a: a.ctypes.data_as(_C.c_void_p)
        self.inp = _S4In(n, P(self.g), P(self.co), P(self.cl), P(self.rb), P(self.ro),
                         P(self.rl), P(self.rmo), P(self.coff), P(self.ch))


def run_c(batch: CBatch, cfg=None, threads=1, decode=True):
    """The C driver over a CBatch on `threads` pthreads.  Returns (alignments
    per read as run() gives them, stats dict), or None with decode=False
    (timing)."""
    cfg = cfg or default_cfg(cpu_mode=True)
    if not cfg["cpu_mode"]:
        raise ValueError("the C driver restates the configuration without the GPU "
                         "dispatcher (cpu_mode=True)")
    L = _c_lib()
    c = c_cfg(cfg)
    h = L.ko_stage4_run_mt(_C.byref(batch.inp), _C.byref(c), int(threads))
    try:
        if not decode:
            return None
        v = _S4View()
        L.ko_stage4_view(h, _C.byref(v))
        n = v.n_reads
        off = _np.ctypeslib.as_array(v.aln_off, shape=(n + 1,)).copy()
        cig = _np.ctypeslib.as_array(v.cigar, shape=(max(v.cigar_words, 1),))[:v.cigar_words] \
            .copy()
        md = _C.string_at(v.md, v.md_bytes) if v.md_bytes else b""
        kinds = {k: name for name, k in _ALN_KINDS.items()}
        alns = []
        for r in range(n):
            recs = []
            for k in range(int(off[r]), int(off[r + 1])):
                a = v.alns[k]
                rec = _record(kinds[a.kind], a.ref_id, a.ref_start, a.ref_end, a.read_start,
                              a.read_end, O.cigar_str(cig[a.cigar_off:a.cigar_off + a.n_cigar]),
                              a.score, a.mapq, a.is_rev, a.nm,
                              md[a.md_off:a.md_off + a.md_len].decode(), bool(a.is_supplementary))
                rec["xs_score"] = a.xs_score if a.has_xs else None
                recs.append(rec)
            alns.append(recs)
        s = v.stats
        st = {k: getattr(s, k) for k in _STATS + _STATS_TAIL}
        st["reason_counts"] = list(s.reason_counts)
        return alns, st
    finally:
        L.ko_stage4_free(h)
