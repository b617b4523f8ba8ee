"""ctypes binding of the CPU oracle (liboracle.so built from kira_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU
baseline.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

CIGAR_OPS = "MIDNS"
OP = {c: k for k, c in enumerate(CIGAR_OPS)}


class Cfg(C.Structure):
    """AlignmentConfig — src/alignment/mod.rs:26-35."""

    _fields_ = [(n, C.c_int32) for n in (
        "match_score", "mismatch", "gap_open", "gap_extend", "bandwidth", "xdrop",
        "clip_penalty")]


class Knobs(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("adaptive", "max_drift", "dp_xdrop", "cert_strict")]


class SwResult(C.Structure):
    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_int32),
                ("read_end", C.c_int32), ("score", C.c_int32), ("early_abort", C.c_int32),
                ("n_cigar", C.c_int32), ("cigar", C.POINTER(C.c_uint32)),
                ("cigar_cap", C.c_int32), ("cells", C.c_int64)]


class BatchInput(C.Structure):
    _fields_ = [("read", C.c_char_p), ("read_len", C.c_uint32), ("window", C.c_char_p),
                ("window_len", C.c_uint32), ("win_start", C.c_uint32),
                ("chain_ref_start", C.c_uint32), ("chain_ref_end", C.c_uint32),
                ("chain_read_start", C.c_uint32), ("chain_read_end", C.c_uint32),
                ("abort_score", C.c_int32)]


class Span(C.Structure):
    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_uint32),
                ("read_end", C.c_uint32)]


class PrefilterOut(C.Structure):
    _fields_ = [("result", C.c_int32), ("reason", C.c_int32), ("has_metrics", C.c_int32),
                ("len", C.c_uint16), ("mism", C.c_uint16), ("matches", C.c_uint16),
                ("identity_x10000", C.c_uint16), ("score", C.c_int32),
                ("read_start", C.c_uint64), ("ref_start", C.c_uint64),
                ("aln_ref_start", C.c_uint32), ("aln_ref_end", C.c_uint32),
                ("aln_read_start", C.c_uint32), ("aln_read_end", C.c_uint32),
                ("aln_score", C.c_int32), ("aln_nm", C.c_uint32), ("n_cigar", C.c_int32),
                ("cigar", C.c_uint32 * 3)]


class Alignment(C.Structure):
    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_uint32),
                ("read_end", C.c_uint32), ("score", C.c_int32), ("nm", C.c_uint32),
                ("n_cigar", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.ko_pack.restype = C.c_int
        L.ko_pack.argtypes = [C.c_char_p, C.c_size_t, P]
        L.ko_pre_shift.argtypes = [C.c_char_p, C.c_size_t, P]
        L.ko_mismatch_count.restype = C.c_uint32
        L.ko_mismatch_count.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_size_t]
        L.ko_scan_best_with_second.restype = C.c_int
        L.ko_scan_best_with_second.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, C.c_int,
                                               C.c_char_p, C.c_size_t, C.c_size_t,
                                               C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64)]
        for fn in ("ko_backend_run_batch", "ko_backend_run_batch_mt"):
            f = getattr(L, fn)
            f.argtypes = [C.c_int32, P, C.c_int32, P, P, C.c_int64, P, P, P, P, P, P] + (
                [C.c_int] if fn.endswith("_mt") else [])
        L.ko_spectral_max_mismatches.restype = C.c_uint32
        L.ko_spectral_max_mismatches.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.ko_spectral_hit_is_certified.restype = C.c_int
        L.ko_spectral_hit_is_certified.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t,
                                                   C.c_size_t, C.c_uint64, C.c_int, C.c_uint64,
                                                   C.c_uint64, C.c_int]
        L.ko_ungapped_tags.restype = C.c_uint32
        L.ko_ungapped_tags.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_char_p,
                                       C.c_size_t]
        L.ko_build_simd_window.restype = C.c_int
        L.ko_build_simd_window.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_int32, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]
        L.ko_identity_x10000.restype = C.c_uint16
        L.ko_identity_x10000.argtypes = [C.c_size_t, C.c_size_t]
        L.ko_min_len_required.restype = C.c_size_t
        L.ko_min_len_required.argtypes = [C.c_size_t, C.c_uint16, C.c_size_t, C.c_size_t]
        L.ko_prefilter_chain.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                         C.POINTER(Span), C.POINTER(Cfg), C.c_int, C.c_int,
                                         C.c_int, C.c_size_t, C.c_float, C.c_size_t, C.c_int32,
                                         C.c_int32, C.c_size_t, C.c_int, C.c_int, C.c_int32,
                                         C.POINTER(PrefilterOut), C.c_char_p, C.c_size_t]
        L.ko_banded_sw.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.c_int32,
                                   C.POINTER(Cfg), C.c_int32, C.POINTER(Knobs),
                                   C.POINTER(SwResult)]
        L.ko_sw_group.argtypes = [C.POINTER(BatchInput), C.c_int32, C.POINTER(Cfg), C.c_int,
                                  C.POINTER(Knobs), C.POINTER(SwResult)]
        L.ko_align_batch_sw.argtypes = [C.POINTER(BatchInput), C.c_int32, C.POINTER(Cfg),
                                        C.c_int, C.c_int, C.POINTER(Knobs), C.POINTER(SwResult)]
        L.ko_build_alignment.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                         C.c_uint32, C.POINTER(SwResult), C.POINTER(Alignment),
                                         C.POINTER(C.c_uint32), C.c_int32, C.c_char_p,
                                         C.c_size_t]
        L.ko_nm_md_from_cigar.restype = C.c_uint32
        L.ko_nm_md_from_cigar.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                          C.c_size_t, C.c_size_t, C.POINTER(C.c_uint32),
                                          C.c_int32, C.c_char_p, C.c_size_t]
        L.ko_banded_sw_batch_mt.argtypes = [C.c_int32, P, P, P, P, P, P, P, P, P,
                                            C.POINTER(Cfg), C.POINTER(Knobs), P, P, P, C.c_int]
        L.ko_myers_bounded.restype = C.c_int
        L.ko_myers_bounded.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                       C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.ko_wfa_align_scored.restype = C.c_int
        L.ko_wfa_align_scored.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32,
                                          C.POINTER(Cfg), C.c_int32, C.c_int32, C.c_int32,
                                          C.POINTER(WfaAln)]
        L.ko_wfa_score_only_scored.restype = C.c_int
        L.ko_wfa_score_only_scored.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32,
                                               C.POINTER(Cfg), C.c_int32, C.c_int32, C.c_int32,
                                               C.POINTER(C.c_int32)]
        _lib = L
    return _lib


class WfaAln(C.Structure):
    _fields_ = [("score", C.c_int32), ("text_start", C.c_int32), ("text_end", C.c_int32),
                ("n_cigar", C.c_int32), ("cigar", C.POINTER(C.c_uint32)),
                ("cigar_cap", C.c_int32)]


def myers_bounded(pattern: bytes, text: bytes, max_k: int):
    """bounded_edit_distance (src/alignment/myers.rs:47-367): (dist, end) or None."""
    d, p = C.c_size_t(), C.c_size_t()
    ok = lib().ko_myers_bounded(pattern, len(pattern), text, len(text), max_k, C.byref(d),
                                C.byref(p))
    return (d.value, p.value) if ok else None


def wfa_align_scored(pattern: bytes, text: bytes, c: "Cfg", max_score: int, adaptive_drop=0,
                     text_begin_free=0):
    """wfa_align_scored (src/alignment/wfa.rs:58-67): dict(score (the WFA loss),
    cigar (words), text_start, text_end) or None."""
    cap = 2 * (len(pattern) + len(text)) + 8
    buf = (C.c_uint32 * cap)()
    a = WfaAln(0, 0, 0, 0, C.cast(buf, C.POINTER(C.c_uint32)), cap)
    r = lib().ko_wfa_align_scored(pattern, len(pattern), text, len(text), C.byref(c), max_score,
                                  adaptive_drop, text_begin_free, C.byref(a))
    if r < 0:
        raise RuntimeError("CIGAR capacity")
    if r == 0:
        return None
    return dict(score=a.score, cigar=[buf[k] for k in range(a.n_cigar)],
                text_start=a.text_start, text_end=a.text_end)


def wfa_score_only_scored(pattern: bytes, text: bytes, c: "Cfg", max_score: int,
                          adaptive_drop=0, text_begin_free=0):
    """wfa_score_only_scored (wfa.rs:69-81): the affine score or None."""
    v = C.c_int32()
    ok = lib().ko_wfa_score_only_scored(pattern, len(pattern), text, len(text), C.byref(c),
                                        max_score, adaptive_drop, text_begin_free, C.byref(v))
    return v.value if ok else None


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def cfg(match_score=1, mismatch=4, gap_open=6, gap_extend=1, bandwidth=32, xdrop=100,
        clip_penalty=5) -> Cfg:
    return Cfg(match_score, mismatch, gap_open, gap_extend, bandwidth, xdrop, clip_penalty)


def knobs(adaptive=True, max_drift=16, dp_xdrop=-1, cert_strict=False) -> Knobs:
    return Knobs(int(adaptive), max_drift, dp_xdrop, int(cert_strict))


# ---------------------------------------------------------------- sieve
def pack(seq: bytes):
    out = np.zeros((len(seq) + 3) // 4, dtype=np.uint8)
    valid = lib().ko_pack(seq, len(seq), _p(out))
    return out.tobytes(), bool(valid)


def pre_shift(bits: bytes) -> List[bytes]:
    out = np.zeros(4 * len(bits), dtype=np.uint8)
    lib().ko_pre_shift(bits, len(bits), _p(out))
    n = len(bits)
    return [out[k * n:(k + 1) * n].tobytes() for k in range(4)]


def mismatch_count(read_bits: bytes, ref_bits: bytes, n_nucs: int) -> int:
    return lib().ko_mismatch_count(read_bits, ref_bits, len(read_bits), n_nucs)


def scan_best_with_second(read: bytes, text: bytes):
    """pack read and text, pre-shift text, scan every shift (bitpacked.rs:138-199).

    Returns (shift, best, second_or_None) or None."""
    rb, rv = pack(read)
    tb, _ = pack(text)
    ph = b"".join(pre_shift(tb))
    sh, be, se = C.c_uint64(), C.c_uint64(), C.c_uint64()
    ok = lib().ko_scan_best_with_second(rb, len(rb), len(read), int(rv), ph, len(tb), len(text),
                                        C.byref(sh), C.byref(be), C.byref(se))
    if not ok:
        return None
    second = None if se.value == 2 ** 64 - 1 else se.value
    return sh.value, be.value, second


READ_BYTES_MAX = 64


def backend_layout(jobs):
    """Flatten (read, text, max_mismatches) jobs into the CudaBackend::run_batch
    SoA layout (src/cuda/backend.rs:118-192)."""
    n = len(jobs)
    reads = np.zeros(n * READ_BYTES_MAX, dtype=np.uint8)
    read_lens = np.zeros(n, dtype=np.int32)
    ref_lens = np.zeros(n, dtype=np.int32)
    limits = np.zeros(n, dtype=np.uint32)
    phase_bytes = np.zeros(n, dtype=np.int64)
    offsets = np.zeros(n, dtype=np.int64)
    packed = []
    for i, (read, text, mm) in enumerate(jobs):
        rb, _ = pack(read)
        reads[i * READ_BYTES_MAX:i * READ_BYTES_MAX + len(rb)] = np.frombuffer(rb, np.uint8)
        read_lens[i] = len(read)
        ref_lens[i] = len(text)
        limits[i] = mm
        tb, _ = pack(text)
        packed.append(pre_shift(tb))
        phase_bytes[i] = len(tb)
    total = int(phase_bytes.sum())
    flat = np.zeros(4 * total + 8, dtype=np.uint8)
    acc = 0
    for i, ph in enumerate(packed):
        for p in range(4):
            flat[p * total + acc:p * total + acc + len(ph[p])] = np.frombuffer(ph[p], np.uint8)
        offsets[i] = acc
        acc += len(ph[0])
    return dict(n=n, reads=reads, read_lens=read_lens, ref_flat=flat, ref_bytes_per_phase=total,
                ref_offsets=offsets, ref_phase_bytes=phase_bytes, ref_lens=ref_lens,
                limits=limits)


def backend_run_batch_layout(lay, threads: int = 0):
    n = lay["n"]
    out_s = np.zeros(n, dtype=np.int32)
    out_m = np.zeros(2 * n, dtype=np.int32)
    args = [n, _p(lay["reads"]), READ_BYTES_MAX, _p(lay["read_lens"]), _p(lay["ref_flat"]),
            int(lay["ref_bytes_per_phase"]), _p(lay["ref_offsets"]), _p(lay["ref_phase_bytes"]),
            _p(lay["ref_lens"]), _p(lay["limits"]), _p(out_s), _p(out_m)]
    if threads:
        lib().ko_backend_run_batch_mt(*args, threads)
    else:
        lib().ko_backend_run_batch(*args)
    return out_s, out_m[:n], out_m[n:]


def backend_run_batch(jobs):
    """CudaBackend::run_batch semantics on the CPU: list of (shift, mism, second|None)."""
    if not jobs:
        return []
    s, m, sec = backend_run_batch_layout(backend_layout(jobs))
    I32MAX = 2 ** 31 - 1
    return [(int(s[i]), int(m[i]), None if sec[i] == I32MAX else int(sec[i]))
            for i in range(len(jobs))]


def spectral_max_mismatches(read_len: int, pct: int = 3, floor: int = 2) -> int:
    return lib().ko_spectral_max_mismatches(read_len, pct, floor)


def certified(read: bytes, ref_aligned: bytes, mism: int, second: Optional[int], max_mism: int,
              strict: bool = False) -> bool:
    return bool(lib().ko_spectral_hit_is_certified(
        read, ref_aligned, len(read), len(ref_aligned), mism, int(second is not None),
        second if second is not None else 0, max_mism, int(strict)))


def ungapped_tags(read: bytes, ref: bytes):
    buf = C.create_string_buffer(4 * len(read) + 32)
    nm = lib().ko_ungapped_tags(read, ref, len(read), buf, len(buf))
    return nm, buf.value.decode()


def build_simd_window(ref_len, chain_ref_start, chain_read_start, read_len, bandwidth):
    s, l = C.c_uint32(), C.c_uint32()
    ok = lib().ko_build_simd_window(ref_len, chain_ref_start, chain_read_start, read_len,
                                    bandwidth, C.byref(s), C.byref(l))
    return (s.value, l.value) if ok else None


# ---------------------------------------------------------------- prefilter
REASONS = ["accept", "disabled", "not_short_preset", "not_top1", "max_alignments_not_one",
           "span_too_small", "mism_too_high", "id_too_low", "indel_suspect"]
OTHER = {1: "read_len0", 2: "non_monotonic", 3: "anchor0", 4: "span0", 5: "neg_start",
         6: "span_oob", 7: "margin_fail"}
RESULTS = ["accept", "reject", "fallback"]


def identity_x10000(length: int, mism: int) -> int:
    return lib().ko_identity_x10000(length, mism)


def min_len_required(read_len, id_x10000, mism, span_slack):
    return lib().ko_min_len_required(read_len, id_x10000, mism, span_slack)


def ungapped_beaten_by_gap(read: bytes, ref: bytes, span, ungapped_mism: int, c: Cfg) -> bool:
    """ungapped_beaten_by_gap — src/alignment/mod.rs:163-204: an exact WFA
    score-only probe of the read against ref[ref_start - read_start .. + L +
    bandwidth] with budget mism * (match + mismatch) - 1 finds an alignment."""
    if ungapped_mism == 0 or len(read) == 0:
        return False
    ws = span[0] - span[2]  # chain ref_start - read_start
    if ws < 0 or ws >= len(ref):
        return False
    band = max(c.bandwidth, 1)
    tl = min(len(read) + band, len(ref) - ws)
    if tl < len(read):
        return False
    cost = ungapped_mism * (c.match_score + c.mismatch)
    if cost <= 0:
        return False
    return wfa_score_only_scored(read, ref[ws:ws + tl], c, cost - 1) is not None


def prefilter_chain(read: bytes, ref: bytes, span, c: Cfg, is_top1=True, accept_enable=True,
                    accept_only_top1=True, accept_span_slack=5, accept_min_identity=98.0,
                    accept_max_mismatches=2, accept_require_score_margin=0,
                    score_margin=2 ** 31 - 1, long_read_threshold=300,
                    multi_alignments_enabled=False, tail_ext=True, ungap_min_id=-1,
                    ungap_gap_check=False):
    """prefilter_chain (src/alignment/prefilter.rs:93-295); ungap_gap_check is
    KIRA_UNGAP_GAP_CHECK, the final gate of :274-287."""
    o = PrefilterOut()
    md = C.create_string_buffer(4 * len(read) + 32)
    sp = Span(*span)
    lib().ko_prefilter_chain(read, len(read), ref, len(ref), C.byref(sp), C.byref(c),
                             int(is_top1), int(accept_enable), int(accept_only_top1),
                             accept_span_slack, accept_min_identity, accept_max_mismatches,
                             accept_require_score_margin, score_margin, long_read_threshold,
                             int(multi_alignments_enabled), int(tail_ext), ungap_min_id,
                             C.byref(o), md, len(md))
    if ungap_gap_check and o.result == 0 and o.mism > 0 and \
            ungapped_beaten_by_gap(read, ref, span, o.mism, c):
        o.result = 2   # Fallback
        o.reason = 8   # IndelSuspect
    r = o.reason
    reason = REASONS[r] if r < 9 else OTHER[(r - 9) // 16]
    d = dict(result=RESULTS[o.result], reason=reason)
    if o.has_metrics:
        d["metrics"] = dict(len=o.len, mism=o.mism, matches=o.matches,
                            identity_x10000=o.identity_x10000, score=o.score,
                            read_start=o.read_start, ref_start=o.ref_start)
    if o.result == 0:
        d["alignment"] = dict(ref_start=o.aln_ref_start, ref_end=o.aln_ref_end,
                              read_start=o.aln_read_start, read_end=o.aln_read_end,
                              score=o.aln_score, nm=o.aln_nm, md=md.value.decode(),
                              cigar=cigar_str([o.cigar[k] for k in range(o.n_cigar)]))
    return d


# ---------------------------------------------------------------- SW
def cigar_str(ops: Sequence[int]) -> str:
    return "".join(f"{o >> 4}{CIGAR_OPS[o & 15]}" for o in ops)


def _sw_dict(r: SwResult):
    return dict(ref_start=r.ref_start, ref_end=r.ref_end, read_start=r.read_start,
                read_end=r.read_end, score=r.score, early_abort=bool(r.early_abort),
                cigar=[r.cigar[k] for k in range(r.n_cigar)], cells=int(r.cells))


def _new_result(cap):
    buf = (C.c_uint32 * cap)()
    r = SwResult()
    r.cigar = C.cast(buf, C.POINTER(C.c_uint32))
    r.cigar_cap = cap
    return r, buf


def banded_sw(read: bytes, ref: bytes, offset: int, c: Cfg, abort_score: int = -(2 ** 31) // 4,
              kn: Optional[Knobs] = None):
    kn = kn or knobs()
    r, buf = _new_result(len(read) + len(ref) + 4)
    lib().ko_banded_sw(read, len(read), ref, len(ref), offset, C.byref(c), abort_score,
                       C.byref(kn), C.byref(r))
    return _sw_dict(r)


@dataclass
class Job:
    """BatchInput — src/alignment/mod.rs:48-56."""

    read: bytes
    window: bytes
    win_start: int = 0
    chain_ref_start: int = 0
    chain_ref_end: int = 0
    chain_read_start: int = 0
    chain_read_end: int = 0
    abort_score: int = -(2 ** 31) // 4


def _inputs(jobs: Sequence[Job]):
    arr = (BatchInput * len(jobs))()
    keep = []
    for k, j in enumerate(jobs):
        rb = C.create_string_buffer(j.read, len(j.read))
        wb = C.create_string_buffer(j.window, len(j.window))
        keep += [rb, wb]
        arr[k] = BatchInput(C.cast(rb, C.c_char_p), len(j.read), C.cast(wb, C.c_char_p),
                            len(j.window), j.win_start, j.chain_ref_start, j.chain_ref_end,
                            j.chain_read_start, j.chain_read_end, j.abort_score)
    return arr, keep


MODES = {"scalar": 0, "avx2": 1, "avxvnni": 2, "neon": 3}


def align_batch_sw(jobs: Sequence[Job], c: Cfg, mode="avx2", detected=None,
                   kn: Optional[Knobs] = None):
    """align_batch_simd's SW results (src/alignment/mod.rs:1111-1149)."""
    if not jobs:
        return []
    kn = kn or knobs()
    detected = mode if detected is None else detected
    arr, keep = _inputs(jobs)
    res = (SwResult * len(jobs))()
    bufs = []
    for k, j in enumerate(jobs):
        r, b = _new_result(len(j.read) + len(j.window) + 4)
        res[k] = r
        bufs.append(b)
    lib().ko_align_batch_sw(arr, len(jobs), C.byref(c), MODES[mode], MODES[detected],
                            C.byref(kn), res)
    return [_sw_dict(res[k]) for k in range(len(jobs))]


def sw_group(jobs: Sequence[Job], c: Cfg, bits=16, kn: Optional[Knobs] = None):
    kn = kn or knobs()
    arr, keep = _inputs(jobs)
    res = (SwResult * len(jobs))()
    bufs = []
    for k, j in enumerate(jobs):
        r, b = _new_result(len(j.read) + len(j.window) + 4)
        res[k] = r
        bufs.append(b)
    lib().ko_sw_group(arr, len(jobs), C.byref(c), bits, C.byref(kn), res)
    return [_sw_dict(res[k]) for k in range(len(jobs))]


def build_alignment(read: bytes, window: bytes, win_start: int, sw: dict):
    r, buf = _new_result(max(1, len(sw["cigar"])))
    for k, o in enumerate(sw["cigar"]):
        buf[k] = o
    r.n_cigar = len(sw["cigar"])
    r.ref_start, r.ref_end = sw["ref_start"], sw["ref_end"]
    r.read_start, r.read_end, r.score = sw["read_start"], sw["read_end"], sw["score"]
    out = Alignment()
    cap = len(sw["cigar"]) + 2
    cig = (C.c_uint32 * cap)()
    md = C.create_string_buffer(4 * (len(read) + len(window)) + 32)
    lib().ko_build_alignment(read, len(read), window, len(window), win_start, C.byref(r),
                             C.byref(out), cig, cap, md, len(md))
    return dict(ref_start=out.ref_start, ref_end=out.ref_end, read_start=out.read_start,
                read_end=out.read_end, score=out.score, nm=out.nm, md=md.value.decode(),
                cigar=cigar_str([cig[k] for k in range(out.n_cigar)]))


def nm_md_from_cigar(read: bytes, ref: bytes, read_start: int, ref_start: int, cigar: str):
    import re
    ops = [(int(n) << 4) | OP[o] for n, o in re.findall(r"(\d+)([MIDNS])", cigar)]
    arr = (C.c_uint32 * max(1, len(ops)))(*ops)
    md = C.create_string_buffer(4 * (len(read) + len(ref)) + 32)
    nm = lib().ko_nm_md_from_cigar(read, len(read), ref, len(ref), read_start, ref_start, arr,
                                   len(ops), md, len(md))
    return nm, md.value.decode()


def banded_sw_batch_mt(reads, read_off, read_len, refs, ref_off, ref_len, offsets, aborts,
                       c: Cfg, kn: Knobs = None, bands=None, threads=1):
    """banded_sw_with over many jobs on `threads` pthreads (flat uint8 buffers,
    int64 offsets, int32 lengths).  Returns (score, ref_end, cells) arrays."""
    kn = kn or knobs()
    n = len(read_len)
    score = np.zeros(n, dtype=np.int32)
    end = np.zeros(n, dtype=np.int32)
    cells = np.zeros(n, dtype=np.int64)
    arrs = [np.ascontiguousarray(reads, dtype=np.uint8),
            np.ascontiguousarray(read_off, dtype=np.int64),
            np.ascontiguousarray(read_len, dtype=np.int32),
            np.ascontiguousarray(refs, dtype=np.uint8),
            np.ascontiguousarray(ref_off, dtype=np.int64),
            np.ascontiguousarray(ref_len, dtype=np.int32),
            np.ascontiguousarray(offsets, dtype=np.int32),
            np.ascontiguousarray(aborts, dtype=np.int32)]
    b = None if bands is None else np.ascontiguousarray(bands, dtype=np.int32)
    lib().ko_banded_sw_batch_mt(n, *[_p(a) for a in arrs], None if b is None else _p(b),
                                C.byref(c), C.byref(kn), _p(score), _p(end), _p(cells), threads)
    return score, end, cells
