"""CPU restatement of the reference's fast DP cascade,
try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440): packed spectral
certificate -> Myers reject -> scored WFA (with the suspect-origin retry) ->
partial-read check -> wfa_result_to_alignment (mod.rs:481-575).

Test infrastructure only (the product never imports it).  It composes the
oracle's pinned pieces (scan_best_with_second / certified for the packed
scan, myers_bounded, wfa_align_scored, nm_md_from_cigar) with the cascade's
own control flow; tests/test_oracle_fastpath.py pins that flow on hand-made
cases whose outcome is fixed by the mathematics (exact reads, a lone
substitution, a 5' deletion that needs the retry, a partial read)."""
from dataclasses import dataclass
from typing import Optional

from oracle import oracle as O

OPS = "MIDNS"


@dataclass
class Router:
    """The router knobs (src/alignment/router.rs:48-215) with their defaults."""
    wfa_max_read_len: int = 300   # KIRA_WFA_MAX_LEN
    algo: str = "packed"          # KIRA_ALGO: packed | spectral | wfa | sw
    quality_gate: bool = False    # KIRA_CHAIN_QUALITY_GATE
    quality_pct: int = 70         # KIRA_CHAIN_QUALITY_PCT
    wfa_budget_pct: int = 25      # KIRA_WFA_BUDGET_PCT
    myers_pct: int = 15           # KIRA_MYERS_BOUND_PCT
    myers_floor: int = 4          # KIRA_MYERS_BOUND_FLOOR
    spectral_pct: int = 3         # KIRA_SPECTRAL_MISM_PCT
    spectral_floor: int = 2       # KIRA_SPECTRAL_MISM_FLOOR
    wfa_ends_free: int = 0        # KIRA_WFA_ENDS_FREE
    wfa_lead: int = 0             # KIRA_WFA_LEAD
    wfa_adaptive: int = 0         # KIRA_WFA_ADAPTIVE (<= 0: exact)
    cert_strict: bool = False     # KIRA_CERT_STRICT


def worth_attempting(chain_score: int, read_len: int, r: Router) -> bool:
    """fast_path_worth_attempting (router.rs:169-189)."""
    if not r.quality_gate:
        return True
    if chain_score <= 0 or read_len == 0:
        return False
    return min(chain_score * 100, 2 ** 32 - 1) // read_len >= r.quality_pct


def wfa_score_budget(read_len: int, mismatch: int, gap_open: int, gap_extend: int,
                     r: Router) -> int:
    """wfa_score_budget (router.rs:160-166)."""
    return max((read_len * r.wfa_budget_pct) // 100 * mismatch, (gap_open + gap_extend * 4) * 3)


def common_prefix_len(a: bytes, b: bytes) -> int:
    """seq.rs:29-45 (byte equality)."""
    n = min(len(a), len(b))
    i = 0
    while i < n and a[i] == b[i]:
        i += 1
    return i


def clipping_can_improve(read: bytes, text: bytes, w: dict, c) -> bool:
    """mod.rs:442-479."""
    qi, ti, score, min_prefix, best_prefix = 0, w["text_start"], 0, 0, 0
    for word in w["cigar"]:
        n, op = word >> 4, word & 15
        if op == 0:
            for _ in range(n):
                same = qi < len(read) and ti < len(text) and read[qi] == text[ti]
                score += c.match_score if same else -c.mismatch
                qi += 1
                ti += 1
                min_prefix = min(min_prefix, score)
                best_prefix = max(best_prefix, score)
        elif op in (1, 2):
            score -= c.gap_open + c.gap_extend * n
            if op == 1:
                qi += n
            else:
                ti += n
            min_prefix = min(min_prefix, score)
        else:
            return True
    return min_prefix < 0 or best_prefix > score + c.clip_penalty


def wfa_result_to_alignment(read: bytes, text: bytes, win_start: int, w: dict, c) -> Optional[dict]:
    """mod.rs:481-575: reject a CIGAR that does not span the read or leaves
    the text; the DP score with byte-equality substitutions; NM / MD."""
    cref = sum(x >> 4 for x in w["cigar"] if (x & 15) in (0, 2, 3))
    cqry = sum(x >> 4 for x in w["cigar"] if (x & 15) in (0, 1, 4))
    if cqry != len(read) or w["text_start"] + cref > len(text) or w["text_end"] > len(text):
        return None
    score, qi, ti = 0, 0, w["text_start"]
    for word in w["cigar"]:
        n, op = word >> 4, word & 15
        if op == 0:
            for _ in range(n):
                score += c.match_score if read[qi] == text[ti] else -c.mismatch
                qi += 1
                ti += 1
        elif op in (1, 2):
            score -= c.gap_open + c.gap_extend * n
            if op == 1:
                qi += n
            else:
                ti += n
    cig = "".join(f"{x >> 4}{OPS[x & 15]}" for x in w["cigar"])
    nm, md = O.nm_md_from_cigar(read, text, 0, w["text_start"], cig)
    return dict(ref_start=win_start + w["text_start"], ref_end=win_start + w["text_end"],
                read_start=0, read_end=len(read), cigar=list(w["cigar"]), score=score, nm=nm,
                md=md)


def spectral(read: bytes, text: bytes, win_start: int, c, r: Router) -> Optional[dict]:
    """try_packed_spectral_alignment (mod.rs:624-686): both sequences must
    pack (A/C/G/T only), then the certified best shift becomes {L}M."""
    if not O.pack(read)[1] or not O.pack(text)[1]:
        return None
    hit = O.scan_best_with_second(read, text)
    if hit is None:
        return None
    shift, mism, second = hit
    L = len(read)
    ref = text[shift:shift + L]
    if len(ref) != L:
        return None
    max_mism = O.spectral_max_mismatches(L, r.spectral_pct, r.spectral_floor)
    if not O.certified(read, ref, mism, second, max_mism, r.cert_strict):
        return None
    nm, md = O.ungapped_tags(read, ref)
    score = (L - mism) * c.match_score - mism * c.mismatch
    return dict(ref_start=win_start + shift, ref_end=win_start + shift + L, read_start=0,
                read_end=L, cigar=[L << 4], score=score, nm=nm, md=md)


def fast_dp(read: bytes, contig: bytes, ref_start: int, read_start: int, chain_score: int, c,
            r: Router = Router()):
    """try_fast_dp_alignment_inner (mod.rs:285-440).  Returns (kind, record)
    with kind 'packed' or 'wfa', or None (the caller falls back to SW)."""
    L = len(read)
    # parse_algo (router.rs:28-37): anything unrecognised is packed
    algo = r.algo if r.algo in ("packed", "spectral", "wfa", "sw") else "packed"
    if L == 0 or L > r.wfa_max_read_len or algo == "sw":
        return None
    if not worth_attempting(chain_score, L, r):
        return None
    win_start = ref_start - read_start
    if win_start < 0 or win_start >= len(contig):
        return None
    band = max(c.bandwidth, 1)
    text_len = min(L + band, len(contig) - win_start)
    if text_len < L:
        return None
    text = contig[win_start:win_start + text_len]
    if algo in ("packed", "spectral"):
        a = spectral(read, text, win_start, c, r)
        if a is not None:
            return "packed", a
    if O.myers_bounded(read, text, L * r.myers_pct // 100 + r.myers_floor) is None:
        return None
    budget = wfa_score_budget(L, c.match_score + c.mismatch, c.gap_open,
                              c.match_score + c.gap_extend, r)
    lead = min(max(r.wfa_lead, 0), win_start, band)
    if lead > 0:
        wtext, wstart, tbf = contig[win_start - lead:win_start + text_len], win_start - lead, lead
    else:
        wtext, wstart, tbf = text, win_start, r.wfa_ends_free
    drop = r.wfa_adaptive if r.wfa_adaptive > 0 else 0
    w = O.wfa_align_scored(read, wtext, c, budget, drop, tbf)
    if w is None:
        return None
    ftext, fstart = wtext, wstart
    cg = w["cigar"]
    suspect = bool(cg) and ((cg[0] & 15) != 0 or ((cg[0] >> 4) < 19 and len(cg) > 1))
    if suspect and lead == 0 and r.wfa_ends_free == 0:
        slack = min(band, 32, win_start)
        if slack > 0:
            origin, prefix = win_start, 0
            for start in range(win_start - slack, min(win_start + slack, len(contig) - 1) + 1):
                m = common_prefix_len(read, contig[start:start + L])
                if m > prefix:
                    prefix, origin = m, start
            if prefix >= 4:
                exp = contig[origin:min(origin + L + band, len(contig))]
                estart, free = origin, 0
            else:
                exp = contig[win_start - slack:win_start + text_len]
                estart, free = win_start - slack, 2 * slack
            retry = O.wfa_align_scored(read, exp, c, w["score"], drop, free)
            if retry is not None and retry["score"] < w["score"]:
                w, ftext, fstart = retry, exp, estart
    perfect = c.match_score * L
    if w["score"] * 10 > perfect * 3 and clipping_can_improve(read, ftext, w, c):
        return None
    a = wfa_result_to_alignment(read, ftext, fstart, w, c)
    return None if a is None else ("wfa", a)


# ---------------------------------------------------------------- C cascade
import ctypes as _C

import numpy as _np


class _Router(_C.Structure):
    _fields_ = [(n, _C.c_int32) for n in (
        "wfa_max_read_len", "algo", "quality_gate", "quality_pct", "wfa_budget_pct", "myers_pct",
        "myers_floor", "spectral_pct", "spectral_floor", "wfa_ends_free", "wfa_lead",
        "wfa_adaptive", "cert_strict")]


FAST_OUT = _np.dtype([("kind", _np.int32), ("ref_start", _np.uint32), ("ref_end", _np.uint32),
                      ("read_start", _np.uint32), ("read_end", _np.uint32),
                      ("score", _np.int32), ("nm", _np.uint32), ("n_cigar", _np.int32)])
_ALGO = {"packed": 0, "spectral": 1, "wfa": 2, "sw": 3}


def _router(r: Router) -> _Router:
    return _Router(r.wfa_max_read_len, _ALGO.get(r.algo, 0), int(r.quality_gate), r.quality_pct,
                   r.wfa_budget_pct, r.myers_pct, r.myers_floor, r.spectral_pct,
                   r.spectral_floor, r.wfa_ends_free, r.wfa_lead, r.wfa_adaptive,
                   int(r.cert_strict))


def fast_dp_prepare(reads, chains, genome: bytes, contig_off, contig_len):
    """The flat arrays ko_fast_dp_batch_mt reads (built once, outside any
    timing).  chains: (contig, ref_start, ref_end, read_start, read_end,
    strand, score)."""
    n = len(reads)
    lens = _np.array([len(x) for x in reads], _np.int32)
    off = _np.zeros(n, _np.int64)
    if n:
        off[1:] = _np.cumsum(lens[:-1], dtype=_np.int64)
    ch = _np.array([(x[0], x[1], x[3], x[6]) for x in chains], _np.int64).reshape(-1, 4)
    return dict(n=n, lens=lens, off=off,
                flat=_np.frombuffer(b"".join(reads) + b"\0" * 8, _np.uint8),
                cc=_np.ascontiguousarray(ch[:, 0], _np.int32),
                rs=_np.ascontiguousarray(ch[:, 1], _np.uint32),
                qs=_np.ascontiguousarray(ch[:, 2], _np.uint32),
                sc=_np.ascontiguousarray(ch[:, 3], _np.int32),
                g=_np.frombuffer(genome, _np.uint8),
                co=_np.ascontiguousarray(contig_off, _np.int64),
                cl=_np.ascontiguousarray(contig_len, _np.int64))


def fast_dp_run_mt(prep, c, r: Router = Router(), threads: int = 1):
    """The cascade (kira_fast_oracle.c, the same steps as fast_dp) over every
    prepared chain on `threads` pthreads.  Returns a FAST_OUT array (kind 0:
    no result, 1 packed, 2 WFA)."""
    out = _np.zeros(prep["n"], FAST_OUT)
    P = lambda a: a.ctypes.data_as(_C.c_void_p)
    rt = _router(r)
    O.lib().ko_fast_dp_batch_mt(prep["n"], P(prep["flat"]), P(prep["off"]), P(prep["lens"]),
                                P(prep["g"]), P(prep["co"]), P(prep["cl"]), P(prep["cc"]),
                                P(prep["rs"]), P(prep["qs"]), P(prep["sc"]), _C.byref(c),
                                _C.byref(rt), P(out), int(threads))
    return out


def fast_dp_batch_mt(reads, chains, genome: bytes, contig_off, contig_len, c, r: Router = Router(),
                     threads: int = 1):
    """fast_dp_prepare + fast_dp_run_mt."""
    return fast_dp_run_mt(fast_dp_prepare(reads, chains, genome, contig_off, contig_len), c, r,
                          threads)
