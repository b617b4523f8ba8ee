"""Python host mirror of kira-ls-aligner's GPU backend API over libkira_hip.so.

The shapes follow the reference's `crate::cuda` module (src/cuda/mod.rs,
backend.rs, dispatcher.rs): `is_available()`, `HipBackend` (CudaBackend),
`HipJob` / `HipResult` (CudaJob / CudaResult), `HipError` (CudaError) and the
`dispatcher` start / dispatch / is_active / stop functions.  Everything runs
through the C ABI declared in include/kira_hip.h; there is no CPU fallback:
a missing library or device raises.

The directory name is not a valid Python identifier, so it is loaded with
importlib from this file's path (see bench.py / tests/kira_pkg.py).
"""
from __future__ import annotations

import ctypes as C
import os
import queue
import threading
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkira_hip.so")
READ_BYTES_MAX = 64  # src/cuda/mod.rs:37
I32_MAX = 2 ** 31 - 1

KIRA_OK = 0
_STATUS = {1: "NotBuilt", 2: "Driver", 3: "Runtime", 4: "NoDevice", 5: "ReadTooLong",
           6: "InvalidArgument", 7: "OutOfMemory"}


class HipError(RuntimeError):
    """CudaError — src/cuda/mod.rs:17-33 (kind is the variant name)."""

    def __init__(self, kind: str, msg: str = ""):
        super().__init__(f"{kind}: {msg}" if msg else kind)
        self.kind = kind


class SieveBatch(C.Structure):
    """kira_sieve_batch (include/kira_hip.h); every pointer is a device pointer."""

    _fields_ = [("n_jobs", C.c_int32), ("max_read_len", C.c_int32),
                ("read_planes", C.c_void_p), ("read_lens", C.c_void_p),
                ("read_valid", C.c_void_p), ("job_read", C.c_void_p),
                ("jobs_per_read", C.c_int32), ("win_start", C.c_void_p),
                ("win_len", C.c_void_p), ("limit", C.c_void_p), ("limit_all", C.c_int32),
                ("out_shift", C.c_void_p), ("out_best", C.c_void_p),
                ("out_second", C.c_void_p), ("out_flags", C.c_void_p),
                ("cert_max_mism", C.c_int32), ("cert_strict", C.c_int32)]


class SwCfg(C.Structure):
    """kira_sw_cfg = AlignmentConfig (src/alignment/mod.rs:26-35)."""

    _fields_ = [(n, C.c_int32) for n in ("match_score", "mismatch", "gap_open", "gap_extend",
                                         "bandwidth", "xdrop", "clip_penalty")]


class SwKnobs(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("adaptive", "max_drift", "dp_xdrop")]


class SwResult(C.Structure):
    """kira_sw_result = SwPublicResult (src/alignment/mod.rs:904-915)."""

    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_int32),
                ("read_end", C.c_int32), ("score", C.c_int32), ("early_abort", C.c_int32),
                ("n_cigar", C.c_int32), ("cells", C.c_int32), ("cigar_off", C.c_int64)]


class WfaResult(C.Structure):
    """kira_wfa_result = WfaAlignment (src/alignment/wfa.rs:20-39); score is
    the WFA loss."""

    _fields_ = [("found", C.c_int32), ("score", C.c_int32), ("text_start", C.c_int32),
                ("text_end", C.c_int32), ("n_cigar", C.c_int32), ("cigar_off", C.c_int64)]


class FastKnobs(C.Structure):
    """kira_fast_knobs = the router knobs of the fast path (src/alignment/router.rs:48-215)."""

    _fields_ = [(n, C.c_int32) for n in (
        "wfa_max_read_len", "algo", "quality_gate", "quality_pct", "wfa_budget_pct", "myers_pct",
        "myers_floor", "spectral_pct", "spectral_floor", "wfa_ends_free", "wfa_lead",
        "wfa_adaptive", "cert_strict")]


class FastAln(C.Structure):
    """kira_fast_aln: the Alignment try_fast_dp_alignment returns."""

    _fields_ = [("kind", C.c_int32), ("ref_start", C.c_uint32), ("ref_end", C.c_uint32),
                ("read_start", C.c_uint32), ("read_end", C.c_uint32), ("score", C.c_int32),
                ("nm", C.c_uint32), ("n_cigar", C.c_int32), ("md_len", C.c_int32),
                ("cigar_off", C.c_int64), ("md_off", C.c_int64)]


class FastStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("attempted", "spectral", "myers_pass", "origin_retries",
                                         "clip_declined", "wfa")] + [("phase_ms", C.c_double * 4)]


class BatchInput(C.Structure):
    """kira_batch_input = BatchInput (src/alignment/mod.rs:48-56)."""

    _fields_ = [("read", C.c_char_p), ("read_len", C.c_uint32), ("window", C.c_char_p),
                ("window_len", C.c_uint32), ("win_start", C.c_uint32),
                ("chain_ref_start", C.c_uint32), ("chain_ref_end", C.c_uint32),
                ("chain_read_start", C.c_uint32), ("chain_read_end", C.c_uint32),
                ("abort_score", C.c_int32)]


SIMD_MODES = {"scalar": 0, "avx2": 1, "avxvnni": 2, "neon": 3}


class Span(C.Structure):
    """kira_span = AnchorSpan (src/alignment/mod.rs:38-46)."""

    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_uint32),
                ("read_end", C.c_uint32)]


class PrefilterParams(C.Structure):
    """kira_prefilter_params: prefilter_chain's per-batch arguments
    (src/alignment/prefilter.rs:93-109) and KIRA_UNGAP_TAIL / KIRA_UNGAP_MINID /
    KIRA_UNGAP_GAP_CHECK."""

    _fields_ = [(n, C.c_int32) for n in (
        "accept_enable", "accept_only_top1", "multi_alignments_enabled", "tail_ext",
        "accept_span_slack", "accept_max_mismatches", "accept_require_score_margin",
        "long_read_threshold")] + [("accept_min_identity", C.c_float), ("ungap_min_id", C.c_int32),
                                   ("ungap_gap_check", C.c_int32)]


class PrefilterOut(C.Structure):
    """kira_prefilter_out = PrefilterOutcome (src/alignment/prefilter.rs:62-80)."""

    _fields_ = [("result", C.c_int32), ("reason", C.c_int32), ("has_metrics", C.c_int32),
                ("len", C.c_uint16), ("mism", C.c_uint16), ("matches", C.c_uint16),
                ("identity_x10000", C.c_uint16), ("score", C.c_int32),
                ("read_start", C.c_uint64), ("ref_start", C.c_uint64),
                ("aln_ref_start", C.c_uint32), ("aln_ref_end", C.c_uint32),
                ("aln_read_start", C.c_uint32), ("aln_read_end", C.c_uint32),
                ("aln_score", C.c_int32), ("aln_nm", C.c_uint32), ("n_cigar", C.c_int32),
                ("cigar", C.c_uint32 * 3)]


class AlignmentRec(C.Structure):
    """kira_alignment: one record of build_alignment (src/alignment/mod.rs:1200-1254)."""

    _fields_ = [("ref_start", C.c_uint32), ("ref_end", C.c_uint32), ("read_start", C.c_uint32),
                ("read_end", C.c_uint32), ("score", C.c_int32), ("nm", C.c_uint32),
                ("n_cigar", C.c_int32), ("md_len", C.c_int32), ("cigar_off", C.c_int64),
                ("md_off", C.c_int64)]


class ServerConfig(C.Structure):
    """kira_server_config (include/kira_hip.h)."""

    _fields_ = [("slots", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("host_staging_bytes", C.c_int64), ("use_graph", C.c_int32), ("cfg", SwCfg),
                ("knobs", SwKnobs), ("sieve_limit", C.c_int32), ("cert_max_mism", C.c_int32),
                ("cert_strict", C.c_int32), ("group_max_reads", C.c_int64)]


class Request(C.Structure):
    """kira_request: one batch of reads with its sieve and banded SW jobs."""

    _fields_ = [("n_reads", C.c_int32), ("reads", C.c_void_p), ("read_off", C.c_void_p),
                ("read_len", C.c_void_p), ("n_sieve", C.c_int32), ("sieve_read", C.c_void_p),
                ("sieve_win_start", C.c_void_p), ("sieve_win_len", C.c_void_p),
                ("sieve_limit", C.c_void_p), ("n_sw", C.c_int32), ("sw_read", C.c_void_p),
                ("sw_win_start", C.c_void_p), ("sw_win_len", C.c_void_p),
                ("sw_offset", C.c_void_p), ("sw_band", C.c_void_p), ("sw_abort", C.c_void_p),
                ("sw_skip_sieve", C.c_void_p)]


class Response(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_sieve", C.c_int32), ("n_sw", C.c_int32),
                ("shift", C.c_void_p), ("best", C.c_void_p), ("second", C.c_void_p),
                ("flags", C.c_void_p), ("sw", C.POINTER(SwResult)), ("sw_skipped", C.c_void_p),
                ("cigar", C.c_void_p), ("cigar_words", C.c_int64), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("peak_in_flight", C.c_int32), ("peak_device_bytes", C.c_int64),
                ("grouped", C.c_int64)]


# PrefilterResult, and PrefilterReason::idx() -> as_str() (src/alignment/
# prefilter.rs:31-58); Other(..) carries the sub-reason as 9 + 16 * k.
PREFILTER_RESULTS = ("accept", "reject", "fallback")
PREFILTER_REASONS = ("accept", "disabled", "not_short_preset", "not_top1",
                     "max_alignments_not_one", "span_too_small", "mism_too_high", "id_too_low",
                     "indel_suspect")
PREFILTER_OTHER = (None, "read_len0", "non_monotonic", "anchor0", "span0", "neg_start",
                   "span_oob", "margin_fail")


class Chain(C.Structure):
    """kira_chain = Chain (src/types.rs:131-140) without anchors."""

    _fields_ = [(n, C.c_uint32) for n in ("ref_id", "ref_start", "ref_end", "read_start",
                                          "read_end")] + [("strand", C.c_int32),
                                                          ("score", C.c_int32)]


class Stage4Cfg(C.Structure):
    """kira_stage4_cfg = AlignmentStageConfig (src/pipeline/stage4_alignment.rs:33-54)
    plus the environment knobs stage 4 reads."""

    _fields_ = [("cfg", SwCfg), ("knobs", SwKnobs), ("min_chain_ratio", C.c_float),
                ("accept_enable", C.c_int32), ("accept_only_top1", C.c_int32),
                ("accept_span_slack", C.c_int32), ("accept_min_identity", C.c_float)] + [
        (n, C.c_int32) for n in (
            "accept_max_mismatches", "accept_require_score_margin", "dp_topk", "dp_abort_margin",
            "long_read_threshold", "max_alignments", "tail_ext", "ungap_min_id", "short_dp_topk",
            "ambig_div", "xs_min_ratio_pct", "spectral_pct", "spectral_floor", "cert_strict",
            "simd", "supplementary", "cpu_mode")] + [("fast", FastKnobs), ("twotier", C.c_int32),
                                                     ("twotier_k", C.c_int32),
                                                     ("ungap_gap_check", C.c_int32)]


class Stage4Aln(C.Structure):
    _fields_ = [("kind", C.c_int32)] + [(n, C.c_uint32) for n in (
        "ref_id", "ref_start", "ref_end", "read_start", "read_end")] + [
        (n, C.c_int32) for n in ("score", "mapq", "is_rev", "is_supplementary")] + [
        ("nm", C.c_uint32), ("has_xs", C.c_int32), ("xs_score", C.c_int32),
        ("n_cigar", C.c_int32), ("md_len", C.c_int32), ("cigar_off", C.c_int64),
        ("md_off", C.c_int64)]


STAGE4_STATS = ["reads", "chains_total", "chains_used", "exact_matches", "prefilter_accept",
                "prefilter_reject", "prefilter_fallback"]
STAGE4_STATS_TAIL = ["gpu_spectral_resolved", "dp_simd", "dp_scalar", "dp_early_abort",
                     "dp_reads", "supplementary", "unmapped", "packed_spectral_resolved",
                     "wfa_resolved"]


class Stage4Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in STAGE4_STATS] + [("reason_counts", C.c_int64 * 10)] + [
        (n, C.c_int64) for n in STAGE4_STATS_TAIL]


class Stage4View(C.Structure):
    _fields_ = [("n_reads", C.c_int32), ("aln_off", C.POINTER(C.c_int64)),
                ("alns", C.POINTER(Stage4Aln)), ("cigar", C.POINTER(C.c_uint32)),
                ("md", C.c_void_p), ("n_alns", C.c_int64), ("cigar_words", C.c_int64),
                ("md_bytes", C.c_int64), ("stats", Stage4Stats), ("phase_ms", C.c_double * 5)]


class SamCfg(C.Structure):
    """kira_sam_cfg: the OutputConfig fields of single-end records (src/io/mod.rs:797-850)."""

    _fields_ = [(n, C.c_int32) for n in ("write_nm", "write_md", "write_as", "write_xs",
                                          "write_rg", "split_as_secondary",
                                          "soft_clip_supplementary")] + [
        ("read_group", C.c_char_p)]


class SwShape(C.Structure):
    _fields_ = [("band", C.c_int32), ("read_len", C.c_int32), ("n_jobs", C.c_int32)]


class SwTuneResult(C.Structure):
    _fields_ = [("key", C.c_uint32), ("route", C.c_int32), ("n_candidates", C.c_int32),
                ("ms_best", C.c_float), ("ms_builtin", C.c_float)]


class SwJobsDev(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("reads", "read_off", "read_len", "refs", "ref_off",
                                          "ref_len", "offset", "band", "abort_score", "res",
                                          "cigar")]


# Every symbol include/kira_hip.h declares, with (restype, argtypes).
_P = C.c_void_p
SIGNATURES = {
    "kira_is_available": (C.c_int, []),
    "kira_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "kira_ctx_destroy": (None, [_P]),
    "kira_ctx_last_error": (C.c_char_p, [_P]),
    "kira_ctx_stream": (C.c_void_p, [_P]),
    "kira_ctx_set_stream": (None, [_P, _P]),
    "kira_ctx_use_own_stream": (None, [_P]),
    "kira_release_buffers": (None, [_P]),
    "kira_sieve_run_batch": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, _P, C.c_int64, _P,
                                       _P, _P, _P, _P]),
    "kira_sieve_run_batch_noalloc": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, _P,
                                               C.c_int64, _P, _P, _P, _P, _P]),
    "kira_genome_create": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_void_p)]),
    "kira_genome_destroy": (None, [_P]),
    "kira_genome_len": (C.c_int64, [_P]),
    "kira_genome_ascii_dev": (C.c_void_p, [_P]),
    "kira_reads_pack_dev": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, _P]),
    "kira_sieve_genome_dev": (C.c_int, [_P, _P, C.POINTER(SieveBatch)]),
    "kira_ctx_sync": (C.c_int, [_P]),
    "kira_banded_sw_batch": (C.c_int, [_P, C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int64, _P,
                                       _P, _P, _P, _P, C.POINTER(SwCfg), C.POINTER(SwKnobs),
                                       _P, _P, C.c_int64]),
    "kira_align_batch_simd": (C.c_int, [_P, C.c_int32, _P, C.POINTER(BatchInput), C.c_int,
                                        C.c_int, C.POINTER(SwCfg), C.POINTER(SwKnobs), _P, _P,
                                        C.c_int64]),
    "kira_prefilter_batch": (C.c_int, [_P, C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int64, _P,
                                       _P, C.POINTER(Span), _P, _P, C.POINTER(SwCfg),
                                       C.POINTER(PrefilterParams), C.POINTER(PrefilterOut)]),
    "kira_build_alignments": (C.c_int, [C.c_int32, _P, _P, _P, _P, _P, _P, _P,
                                        C.POINTER(SwResult), _P, C.POINTER(AlignmentRec), _P,
                                        C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64)]),
    "kira_server_create": (C.c_int, [_P, _P, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "kira_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "kira_server_wait": (C.c_int, [_P, C.c_int64, C.POINTER(Response)]),
    "kira_server_release": (None, [_P, C.c_int64]),
    "kira_server_get_stats": (None, [_P, C.POINTER(ServerStats)]),
    "kira_server_last_error": (C.c_char_p, [_P]),
    "kira_server_destroy": (None, [_P]),
    "kira_banded_sw_plan_create": (C.c_int, [_P, C.c_int32, _P, _P, C.POINTER(SwCfg),
                                             C.POINTER(C.c_void_p)]),
    "kira_banded_sw_plan_destroy": (None, [_P]),
    "kira_banded_sw_plan_cigar_words": (C.c_int64, [_P]),
    "kira_banded_sw_plan_device_bytes": (C.c_int64, [_P]),
    "kira_banded_sw_plan_cigar_off": (C.POINTER(C.c_int64), [_P]),
    "kira_banded_sw_run_dev": (C.c_int, [_P, _P, C.POINTER(SwJobsDev), C.POINTER(SwCfg),
                                         C.POINTER(SwKnobs)]),
    "kira_sw_autotune": (C.c_int, [_P, C.POINTER(SwCfg), C.POINTER(SwKnobs), C.c_int32,
                                   C.POINTER(SwShape), C.c_char_p, C.POINTER(SwTuneResult)]),
    "kira_sw_tune_load": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int32)]),
    "kira_sw_tune_clear": (None, [_P]),
    "kira_banded_sw_plan_route_counts": (None, [_P, _P]),
    "kira_banded_sw_plan_job_class": (None, [_P, _P]),
    "kira_stage4_run": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_int32, _P, _P, _P, _P, _P,
                                  C.POINTER(Chain), C.POINTER(Stage4Cfg),
                                  C.POINTER(C.c_void_p)]),
    "kira_stage4_result_view": (C.c_int, [_P, C.POINTER(Stage4View)]),
    "kira_stage4_result_free": (None, [_P]),
    "kira_left_normalize_indels": (C.c_int32, [_P, C.c_int32, _P, C.c_int64, C.c_uint32, _P,
                                               C.c_int32, _P, C.c_int32,
                                               C.POINTER(C.c_int32)]),
    "kira_stage4_left_normalize": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_int32, _P, _P, _P,
                                             _P, _P]),
    "kira_stage4_sam": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P, _P, _P, _P,
                                  C.POINTER(SamCfg), _P, C.c_int64, C.POINTER(C.c_int64)]),
    "kira_myers_batch": (C.c_int, [_P, C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int64, _P, _P,
                                   _P, _P, _P, _P]),
    "kira_wfa_batch": (C.c_int, [_P, C.c_int32, _P, C.c_int64, _P, _P, _P, C.c_int64, _P, _P,
                                 _P, _P, C.c_int32, C.POINTER(SwCfg), C.c_int32, _P, _P,
                                 C.c_int64]),
    "kira_fast_knobs_default": (None, [C.POINTER(FastKnobs)]),
    "kira_fast_dp_batch": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_int32, _P, C.c_int64, _P, _P,
                                     _P, C.POINTER(SwCfg), C.POINTER(FastKnobs), _P, _P,
                                     C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64), C.POINTER(FastStats)]),
}

_lib = None


def lib():
    """Load libkira_hip.so or raise (never falls back to a CPU path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipError("NotBuilt", f"{LIB_PATH} missing; build with __graft_entry__.build()")
        # torch ships its own libamdhip64.so.7.  Loading it first makes this
        # library bind to the same HIP runtime (same soname), so device
        # pointers and streams from torch are valid here and torch still sees
        # the GPU after this library initialises HIP.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def is_available() -> bool:
    """is_available() — src/cuda/mod.rs:40-49."""
    try:
        return bool(lib().kira_is_available())
    except (HipError, OSError):
        return False


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class HipJob:
    """CudaJob — src/cuda/backend.rs:13-26."""

    read_packed: bytes
    read_nucs: int
    ref_shifted: Sequence[bytes]
    ref_nucs: int
    max_mismatches: int


@dataclass
class HipResult:
    """CudaResult — src/cuda/backend.rs:28-35."""

    shift: int
    mismatches: int
    second_mismatches: Optional[int]


class Context:
    """One kira_ctx (device + stream + buffers)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().kira_ctx_create(device, C.byref(h))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "Driver"), "kira_ctx_create failed")
        self.h = h

    def check(self, st: int, what: str):
        if st != KIRA_OK:
            msg = lib().kira_ctx_last_error(self.h)
            raise HipError(_STATUS.get(st, "Runtime"), f"{what}: {msg.decode() if msg else ''}")

    def set_stream(self, stream_handle: int):
        lib().kira_ctx_set_stream(self.h, C.c_void_p(stream_handle))

    def sync(self):
        self.check(lib().kira_ctx_sync(self.h), "sync")

    def release_buffers(self):
        lib().kira_release_buffers(self.h)

    def close(self):
        if self.h:
            lib().kira_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipBackend:
    """CudaBackend — src/cuda/backend.rs:42-262, on libkira_hip."""

    def __init__(self, device: int = 0):
        self.ctx = Context(device)

    def run_batch(self, jobs: Sequence[HipJob], noalloc: bool = False) -> List[HipResult]:
        """run_batch — src/cuda/backend.rs:96-249 (same errors, same result order).

        noalloc=True goes through kira_sieve_run_batch_noalloc, which stages
        the batch in the library's fixed per-device array instead of
        allocating (batches beyond its size raise InvalidArgument)."""
        n = len(jobs)
        if n == 0:
            return []
        for j in jobs:
            if len(j.read_packed) > READ_BYTES_MAX:
                raise HipError("ReadTooLong", f"{len(j.read_packed)} > {READ_BYTES_MAX}")
            l0 = len(j.ref_shifted[0])
            if any(len(p) != l0 for p in j.ref_shifted):
                raise HipError("Runtime", "ref_shifted phases have different lengths")
        reads = np.zeros(n * READ_BYTES_MAX, dtype=np.uint8)
        plen = np.empty(n, dtype=np.int32)
        nucs = np.empty(n, dtype=np.int32)
        ref_nucs = np.empty(n, dtype=np.int32)
        limits = np.empty(n, dtype=np.uint32)
        offsets = np.empty(n, dtype=np.int64)
        sizes = np.array([len(j.ref_shifted[0]) for j in jobs], dtype=np.int64)
        total = int(sizes.sum())
        flat = np.zeros(4 * total + 8, dtype=np.uint8)
        acc = 0
        for i, j in enumerate(jobs):
            rb = np.frombuffer(bytes(j.read_packed), dtype=np.uint8)
            reads[i * READ_BYTES_MAX:i * READ_BYTES_MAX + len(rb)] = rb
            plen[i] = len(rb)
            nucs[i] = j.read_nucs
            ref_nucs[i] = j.ref_nucs
            limits[i] = min(j.max_mismatches, 2 ** 32 - 1)
            for p in range(4):
                ph = np.frombuffer(bytes(j.ref_shifted[p]), dtype=np.uint8)
                flat[p * total + acc:p * total + acc + len(ph)] = ph
            offsets[i] = acc
            acc += int(sizes[i])
        out_s = np.empty(n, dtype=np.int32)
        out_m = np.empty(2 * n, dtype=np.int32)
        fn = lib().kira_sieve_run_batch_noalloc if noalloc else lib().kira_sieve_run_batch
        st = fn(self.ctx.h, n, _np_ptr(reads), READ_BYTES_MAX, _np_ptr(plen), _np_ptr(nucs),
                _np_ptr(flat), total, _np_ptr(offsets), _np_ptr(ref_nucs), _np_ptr(limits),
                _np_ptr(out_s), _np_ptr(out_m))
        self.ctx.check(st, "run_batch")
        return [HipResult(int(out_s[i]), int(out_m[i]),
                          None if out_m[n + i] == I32_MAX else int(out_m[n + i]))
                for i in range(n)]

    def release_buffers(self):
        self.ctx.release_buffers()


class Genome:
    """A packed reference resident on the device (kira_genome_create)."""

    def __init__(self, ctx: Context, seq: bytes):
        self.ctx = ctx
        buf = np.frombuffer(seq, dtype=np.uint8)
        h = C.c_void_p()
        ctx.check(lib().kira_genome_create(ctx.h, _np_ptr(buf), len(seq), C.byref(h)),
                  "genome_create")
        self.h = h
        self.length = len(seq)
        # device address of base 0 of the resident ASCII copy (1024 readable
        # zero bytes on each side)
        self.ascii_ptr = int(lib().kira_genome_ascii_dev(h) or 0)

    def close(self):
        if self.h:
            lib().kira_genome_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_reads_dev(ctx: Context, n: int, reads_ptr: int, stride: int, lens_ptr: int,
                   planes_ptr: int, valid_ptr: int):
    """kira_reads_pack_dev on device pointers (ints)."""
    ctx.check(lib().kira_reads_pack_dev(ctx.h, n, C.c_void_p(reads_ptr), stride,
                                        C.c_void_p(lens_ptr), C.c_void_p(planes_ptr),
                                        C.c_void_p(valid_ptr) if valid_ptr else None),
              "reads_pack")


def sieve_genome_dev(ctx: Context, genome: Genome, batch: SieveBatch):
    ctx.check(lib().kira_sieve_genome_dev(ctx.h, genome.h, C.byref(batch)), "sieve_genome")


# ------------------------------------------------------------ dispatcher
class _Dispatcher:
    """dispatcher — src/cuda/dispatcher.rs:13-104: one worker thread owns the
    backend; callers block on a private reply queue; any error -> None."""

    def __init__(self):
        self.lock = threading.Lock()
        self.tx: Optional[queue.Queue] = None
        self.thread: Optional[threading.Thread] = None

    def start(self, device: int = 0):
        with self.lock:
            if self.tx is not None:
                return
            tx: queue.Queue = queue.Queue()

            def worker():
                try:
                    backend = HipBackend(device)
                except HipError as e:
                    err = HipError("Driver", f"GPU worker thread is dead: {e}")
                    while True:
                        item = tx.get()
                        if item is None:
                            return
                        item[1].put(err)
                while True:
                    item = tx.get()
                    if item is None:
                        backend.release_buffers()
                        return
                    jobs, reply = item
                    try:
                        reply.put(backend.run_batch(jobs))
                    except HipError as e:
                        reply.put(e)

            self.tx = tx
            self.thread = threading.Thread(target=worker, name="kira-hip-dispatch", daemon=True)
            self.thread.start()

    def dispatch(self, jobs: List[HipJob]) -> Optional[List[HipResult]]:
        if not jobs:
            return []
        with self.lock:
            if self.tx is None:
                return None
            reply: queue.Queue = queue.Queue()
            self.tx.put((jobs, reply))
        res = reply.get()
        return None if isinstance(res, HipError) else res

    def is_active(self) -> bool:
        with self.lock:
            return self.tx is not None

    def stop(self):
        with self.lock:
            tx, th = self.tx, self.thread
            self.tx, self.thread = None, None
        if tx is not None:
            tx.put(None)
            th.join()


dispatcher = _Dispatcher()


# ------------------------------------------------------------ torch plumbing
def _dev_ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


class SieveOutputs:
    def __init__(self, shift, best, second, flags):
        self.shift, self.best, self.second, self.flags = shift, best, second, flags


def sieve_genome(ctx: Context, genome: Genome, reads, read_lens, win_start, win_len, *,
                 job_read=None, jobs_per_read: int = 1, limit=None, limit_all: int = 0,
                 cert_max_mism: int = -1, cert_strict: bool = False, max_read_len=None,
                 planes=None, valid=None, out: Optional[SieveOutputs] = None,
                 stream=None, skip_pack: bool = False) -> SieveOutputs:
    """Batched sieve + certification against a resident genome.

    reads: uint8 CUDA tensor [n_reads, stride] (ASCII); read_lens int32 [n_reads];
    win_start int64 [n_jobs]; win_len int32 [n_jobs]; job_read int32 [n_jobs] or
    None (job // jobs_per_read); limit int32 [n_jobs] or None (limit_all).
    Launches on `stream` (a torch.cuda.Stream, default: the current stream).
    skip_pack reuses `planes`/`valid` packed by an earlier call."""
    import torch
    dev = reads.device
    n_reads, stride = reads.shape
    n_jobs = int(win_start.shape[0])
    if stream is None:
        stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    if planes is None:
        planes = torch.empty((n_reads, 24), dtype=torch.int32, device=dev)
    if valid is None:
        valid = torch.empty(n_reads, dtype=torch.uint8, device=dev)
    if not skip_pack:
        pack_reads_dev(ctx, n_reads, _dev_ptr(reads), stride, _dev_ptr(read_lens),
                       _dev_ptr(planes), _dev_ptr(valid))
    if out is None:
        out = SieveOutputs(torch.empty(n_jobs, dtype=torch.int32, device=dev),
                           torch.empty(n_jobs, dtype=torch.int32, device=dev),
                           torch.empty(n_jobs, dtype=torch.int32, device=dev),
                           torch.empty(n_jobs, dtype=torch.uint8, device=dev))
    if max_read_len is None:
        max_read_len = int(read_lens.max().item()) if n_reads else 1
    b = SieveBatch()
    b.n_jobs = n_jobs
    b.max_read_len = int(max_read_len)
    b.read_planes = _dev_ptr(planes)
    b.read_lens = _dev_ptr(read_lens)
    b.read_valid = _dev_ptr(valid)
    b.job_read = _dev_ptr(job_read) if job_read is not None else None
    b.jobs_per_read = jobs_per_read
    b.win_start = _dev_ptr(win_start)
    b.win_len = _dev_ptr(win_len)
    b.limit = _dev_ptr(limit) if limit is not None else None
    b.limit_all = limit_all
    b.out_shift = _dev_ptr(out.shift)
    b.out_best = _dev_ptr(out.best)
    b.out_second = _dev_ptr(out.second)
    b.out_flags = _dev_ptr(out.flags)
    b.cert_max_mism = cert_max_mism
    b.cert_strict = int(cert_strict)
    sieve_genome_dev(ctx, genome, b)
    return out


# ------------------------------------------------------------ banded SW
ABORT_NONE = -(2 ** 31) // 8


def sw_cfg(match_score=1, mismatch=4, gap_open=6, gap_extend=1, bandwidth=32, xdrop=100,
           clip_penalty=5) -> SwCfg:
    return SwCfg(match_score, mismatch, gap_open, gap_extend, bandwidth, xdrop, clip_penalty)


def sw_knobs(adaptive=True, max_drift=16, dp_xdrop=-1) -> SwKnobs:
    return SwKnobs(int(adaptive), max_drift, dp_xdrop)


def banded_sw_batch(ctx: Context, jobs, cfg: SwCfg, knobs: Optional[SwKnobs] = None):
    """banded_sw_public (src/alignment/mod.rs:891-920) over a batch.

    jobs: sequence of (read: bytes, ref: bytes, offset: int, abort_score: int,
    band: Optional[int]).  Returns dicts with ref_start, ref_end, read_start,
    read_end, score, early_abort and cigar (list of len << 4 | op words)."""
    knobs = knobs or sw_knobs()
    n = len(jobs)
    if n == 0:
        return []
    reads = b"".join(j[0] for j in jobs)
    refs = b"".join(j[1] for j in jobs)
    rl = np.array([len(j[0]) for j in jobs], dtype=np.int32)
    fl = np.array([len(j[1]) for j in jobs], dtype=np.int32)
    ro = np.concatenate([[0], np.cumsum(rl[:-1], dtype=np.int64)]).astype(np.int64)
    fo = np.concatenate([[0], np.cumsum(fl[:-1], dtype=np.int64)]).astype(np.int64)
    off = np.array([j[2] for j in jobs], dtype=np.int32)
    ab = np.array([j[3] for j in jobs], dtype=np.int32)
    bands = [j[4] if len(j) > 4 else None for j in jobs]
    band = None
    if any(b is not None for b in bands):
        band = np.array([cfg.bandwidth if b is None else b for b in bands], dtype=np.int32)
    rbuf = np.frombuffer(reads, dtype=np.uint8) if reads else np.zeros(1, np.uint8)
    fbuf = np.frombuffer(refs, dtype=np.uint8) if refs else np.zeros(1, np.uint8)
    out = (SwResult * n)()
    cap = int(3 * rl.sum() + 2 * n + 2)
    arena = np.zeros(cap, dtype=np.uint32)
    st = lib().kira_banded_sw_batch(ctx.h, n, _np_ptr(rbuf), len(reads), _np_ptr(ro), _np_ptr(rl),
                                    _np_ptr(fbuf), len(refs), _np_ptr(fo), _np_ptr(fl),
                                    _np_ptr(off), _np_ptr(band) if band is not None else None,
                                    _np_ptr(ab), C.byref(cfg), C.byref(knobs), out,
                                    _np_ptr(arena), cap)
    ctx.check(st, "banded_sw_batch")
    res = []
    for k in range(n):
        o = out[k]
        res.append(dict(ref_start=o.ref_start, ref_end=o.ref_end, read_start=o.read_start,
                        read_end=o.read_end, score=o.score, early_abort=bool(o.early_abort),
                        cigar=[int(x) for x in arena[o.cigar_off:o.cigar_off + o.n_cigar]],
                        cells=int(o.cells)))
    return res


class BandedSwPlan:
    """kira_sw_plan: the packet schedule and trace slabs of one batch shape
    (read lengths and bands); reusable for every batch of that shape."""

    def __init__(self, ctx: Context, read_len: np.ndarray, band: Optional[np.ndarray],
                 cfg: SwCfg):
        rl = np.ascontiguousarray(read_len, dtype=np.int32)
        bd = None if band is None else np.ascontiguousarray(band, dtype=np.int32)
        h = C.c_void_p()
        st = lib().kira_banded_sw_plan_create(ctx.h, len(rl), _np_ptr(rl),
                                              None if bd is None else _np_ptr(bd),
                                              C.byref(cfg), C.byref(h))
        ctx.check(st, "banded_sw_plan_create")
        self.h, self.n = h, len(rl)
        self.cigar_words = int(lib().kira_banded_sw_plan_cigar_words(h))
        self.device_bytes = int(lib().kira_banded_sw_plan_device_bytes(h))
        ptr = lib().kira_banded_sw_plan_cigar_off(h)
        self.cigar_off = (np.ctypeslib.as_array(ptr, shape=(self.n,)).copy() if self.n
                          else np.zeros(0, dtype=np.int64))

    def route_counts(self) -> dict:
        """Jobs per route (kira_banded_sw_plan_route_counts), by SW_ROUTES name."""
        c = np.zeros(len(SW_ROUTES), dtype=np.int32)
        lib().kira_banded_sw_plan_route_counts(self.h, _np_ptr(c))
        return {SW_ROUTES[r]: int(c[r]) for r in range(1, len(SW_ROUTES)) if c[r]}

    def job_class(self) -> np.ndarray:
        """Kernel class per job (kira_banded_sw_plan_job_class)."""
        c = np.zeros(max(self.n, 1), dtype=np.int32)
        lib().kira_banded_sw_plan_job_class(self.h, _np_ptr(c))
        return c[:self.n]

    def close(self):
        if self.h:
            lib().kira_banded_sw_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def banded_sw_run_dev(ctx: Context, plan: BandedSwPlan, *, reads, read_off, read_len, refs,
                      ref_off, ref_len, offset, abort_score, res, cigar, cfg: SwCfg,
                      band=None, knobs: Optional[SwKnobs] = None, stream=None):
    """banded_sw_public over device-resident jobs (torch CUDA tensors): reads /
    refs uint8, *_off int64, *_len / offset / band / abort_score int32, res int32
    [n, 10], cigar int32 or uint32 [plan.cigar_words].  refs must hold 1024
    readable bytes before and after every window.  Launches on `stream`
    (default: the current stream) and does not synchronise."""
    import torch
    knobs = knobs or sw_knobs()
    if stream is None:
        stream = torch.cuda.current_stream(reads.device)
    ctx.set_stream(stream.cuda_stream)
    j = SwJobsDev(*[_dev_ptr(t) if t is not None else None for t in (
        reads, read_off, read_len, refs, ref_off, ref_len, offset, band, abort_score, res,
        cigar)])
    ctx.check(lib().kira_banded_sw_run_dev(ctx.h, plan.h, C.byref(j), C.byref(cfg),
                                           C.byref(knobs)), "banded_sw_run_dev")


# (P, G): P band positions per lane, G lanes per job; "i16" = packed int16 pairs
SW_ROUTES = {0: "built-in", 1: "lane per job", 2: "(5, 16)", 3: "(9, 16)", 4: "(17, 16)",
             5: "(13, 32)", 6: "(17, 32)", 7: "(6, 16) i16", 8: "(10, 16) i16",
             9: "(18, 16) i16", 10: "(14, 32) i16", 11: "(18, 32) i16", 12: "(10, 8) i16",
             13: "(18, 8) i16"}


def sw_autotune(ctx: Context, shapes, cfg: SwCfg, knobs: Optional[SwKnobs] = None,
                cache_path: Optional[str] = None):
    """kira_sw_autotune: shapes = [(band, read_len, n_jobs)].  Returns one dict per
    shape: key, route (SW_ROUTES), candidates, ms_best, ms_builtin."""
    knobs = knobs or sw_knobs()
    n = len(shapes)
    arr = (SwShape * max(n, 1))(*[SwShape(*map(int, sh)) for sh in shapes])
    out = (SwTuneResult * max(n, 1))()
    ctx.check(lib().kira_sw_autotune(ctx.h, C.byref(cfg), C.byref(knobs), n, arr,
                                     cache_path.encode() if cache_path else None, out),
              "sw_autotune")
    return [dict(key=int(o.key), route=SW_ROUTES[o.route], route_id=int(o.route),
                 candidates=int(o.n_candidates), ms_best=float(o.ms_best),
                 ms_builtin=float(o.ms_builtin)) for o in out[:n]]


def sw_tune_load(ctx: Context, cache_path: str) -> int:
    """kira_sw_tune_load: this device's routes from a cache file into the context."""
    k = C.c_int32()
    ctx.check(lib().kira_sw_tune_load(ctx.h, cache_path.encode(), C.byref(k)), "sw_tune_load")
    return int(k.value)


def sw_tune_clear(ctx: Context):
    lib().kira_sw_tune_clear(ctx.h)


def align_batch_simd(ctx: Context, batches, cfg: SwCfg, mode: str = "avx2",
                     detected: Optional[str] = None, knobs: Optional[SwKnobs] = None):
    """align_batch_simd's SW results (src/alignment/mod.rs:1111-1149) for several
    calls at once.  batches: list of lists of BatchInput-like tuples
    (read, window, win_start, chain_ref_start, chain_ref_end, chain_read_start,
    chain_read_end, abort_score).  Returns one list of result dicts per batch."""
    knobs = knobs or sw_knobs()
    detected = mode if detected is None else detected
    flat = [j for b in batches for j in b]
    off = np.zeros(len(batches) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(b) for b in batches])
    n = len(flat)
    if n == 0:
        return [[] for _ in batches]
    arr = (BatchInput * n)()
    keep = []
    for k, j in enumerate(flat):
        rb = C.create_string_buffer(bytes(j[0]), len(j[0]))
        wb = C.create_string_buffer(bytes(j[1]), len(j[1]))
        keep += [rb, wb]
        arr[k] = BatchInput(C.cast(rb, C.c_char_p), len(j[0]), C.cast(wb, C.c_char_p), len(j[1]),
                            *j[2:8])
    out = (SwResult * n)()
    cap = int(sum(3 * len(j[0]) + 2 for j in flat) + 2)
    arena = np.zeros(cap, dtype=np.uint32)
    st = lib().kira_align_batch_simd(ctx.h, len(batches), _np_ptr(off), arr, SIMD_MODES[mode],
                                     SIMD_MODES[detected], C.byref(cfg), C.byref(knobs), out,
                                     _np_ptr(arena), cap)
    ctx.check(st, "align_batch_simd")
    res = []
    for k in range(n):
        o = out[k]
        res.append(dict(ref_start=o.ref_start, ref_end=o.ref_end, read_start=o.read_start,
                        read_end=o.read_end, score=o.score, early_abort=bool(o.early_abort),
                        cigar=[int(x) for x in arena[o.cigar_off:o.cigar_off + o.n_cigar]],
                        cells=int(o.cells)))
    return [res[off[b]:off[b + 1]] for b in range(len(batches))]


# ------------------------------------------------------------ ungapped prefilter
def prefilter_params(accept_enable=True, accept_only_top1=True, multi_alignments_enabled=False,
                     tail_ext=True, accept_span_slack=5, accept_max_mismatches=2,
                     accept_require_score_margin=0, long_read_threshold=300,
                     accept_min_identity=98.0, ungap_min_id=-1,
                     ungap_gap_check=False) -> PrefilterParams:
    return PrefilterParams(int(accept_enable), int(accept_only_top1),
                           int(multi_alignments_enabled), int(tail_ext), accept_span_slack,
                           accept_max_mismatches, accept_require_score_margin,
                           long_read_threshold, accept_min_identity, ungap_min_id,
                           int(ungap_gap_check))


def prefilter_batch(ctx: Context, chains, cfg: SwCfg, params: Optional[PrefilterParams] = None):
    """prefilter_chain (src/alignment/prefilter.rs:93-295) for a batch of chains.

    chains: sequence of (read: bytes, ref: bytes, span: (ref_start, ref_end,
    read_start, read_end), is_top1: bool = True, score_margin: int = i32::MAX).
    Returns dicts {result, reason, metrics?, alignment?}; the alignment's MD
    string is left to the caller (md_from_cigar)."""
    params = params or prefilter_params()
    n = len(chains)
    if n == 0:
        return []
    reads = b"".join(c[0] for c in chains)
    refs = b"".join(c[1] for c in chains)
    rl = np.array([len(c[0]) for c in chains], dtype=np.int32)
    fl = np.array([len(c[1]) for c in chains], dtype=np.int64)
    ro = np.concatenate([[0], np.cumsum(rl[:-1], dtype=np.int64)]).astype(np.int64)
    fo = np.concatenate([[0], np.cumsum(fl[:-1])]).astype(np.int64)
    spans = (Span * n)(*[Span(*c[2]) for c in chains])
    top = np.array([int(c[3]) if len(c) > 3 else 1 for c in chains], dtype=np.uint8)
    margin = np.array([c[4] if len(c) > 4 else I32_MAX for c in chains], dtype=np.int32)
    rbuf = np.frombuffer(reads, dtype=np.uint8) if reads else np.zeros(1, np.uint8)
    fbuf = np.frombuffer(refs, dtype=np.uint8) if refs else np.zeros(1, np.uint8)
    out = (PrefilterOut * n)()
    st = lib().kira_prefilter_batch(ctx.h, n, _np_ptr(rbuf), len(reads), _np_ptr(ro),
                                    _np_ptr(rl), _np_ptr(fbuf), len(refs), _np_ptr(fo),
                                    _np_ptr(fl), spans, _np_ptr(top), _np_ptr(margin),
                                    C.byref(cfg), C.byref(params), out)
    ctx.check(st, "prefilter_batch")
    res = []
    for k in range(n):
        o = out[k]
        r = o.reason
        d = dict(result=PREFILTER_RESULTS[o.result],
                 reason=PREFILTER_REASONS[r] if r < 9 else PREFILTER_OTHER[(r - 9) // 16])
        if o.has_metrics:
            d["metrics"] = dict(len=o.len, mism=o.mism, matches=o.matches,
                                identity_x10000=o.identity_x10000, score=o.score,
                                read_start=o.read_start, ref_start=o.ref_start)
        if o.result == 0:
            d["alignment"] = dict(ref_start=o.aln_ref_start, ref_end=o.aln_ref_end,
                                  read_start=o.aln_read_start, read_end=o.aln_read_end,
                                  score=o.aln_score, nm=o.aln_nm,
                                  cigar=[int(o.cigar[i]) for i in range(o.n_cigar)])
        res.append(d)
    return res


# ------------------------------------------------------------ serving
def _arr(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class Server:
    """kira_server: continuous batching of read batches over a resident genome.

    submit() copies a request and returns a ticket at once; wait(ticket)
    blocks until its results are back on the host and returns them as numpy
    arrays with the submit / admit / done times (steady clock, seconds)."""

    def __init__(self, ctx: Context, genome: Genome, *, slots: int = 4,
                 hbm_budget_bytes: int = 32 << 30, host_staging_bytes: int = 4 << 30,
                 use_graph: bool = False, cfg: Optional[SwCfg] = None,
                 knobs: Optional[SwKnobs] = None, sieve_limit: int = 8,
                 cert_max_mism: int = -1, cert_strict: bool = False,
                 group_max_reads: int = 0):
        self.genome = genome  # keeps the genome alive
        c = ServerConfig(slots, hbm_budget_bytes, host_staging_bytes, int(use_graph),
                         cfg or sw_cfg(), knobs or sw_knobs(), sieve_limit, cert_max_mism,
                         int(cert_strict), int(group_max_reads))
        h = C.c_void_p()
        st = lib().kira_server_create(ctx.h, genome.h, C.byref(c), C.byref(h))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "Runtime"), "kira_server_create")
        self.h = h

    def submit(self, reads, read_off, read_len, *, sieve=None, sw=None) -> int:
        """reads: flat uint8 (ASCII); read_off int64, read_len int32.
        sieve: (read, win_start, win_len[, limit]) arrays; sw: (read,
        win_start, win_len, offset, band, abort[, skip_sieve]) arrays."""
        keep = []

        def P(a, dt):
            a = _arr(a, dt)
            if a is None:
                return None
            keep.append(a)
            return _np_ptr(a) if a.size else None

        reads = np.ascontiguousarray(np.frombuffer(reads, np.uint8) if isinstance(reads, bytes)
                                     else reads, dtype=np.uint8)
        keep.append(reads)
        r = Request()
        r.n_reads = len(read_len)
        r.reads = _np_ptr(reads) if reads.size else None
        r.read_off = P(read_off, np.int64)
        r.read_len = P(read_len, np.int32)
        if sieve is not None:
            r.n_sieve = len(sieve[0])
            r.sieve_read = P(sieve[0], np.int32)
            r.sieve_win_start = P(sieve[1], np.int64)
            r.sieve_win_len = P(sieve[2], np.int32)
            r.sieve_limit = P(sieve[3], np.int32) if len(sieve) > 3 else None
        if sw is not None:
            r.n_sw = len(sw[0])
            r.sw_read = P(sw[0], np.int32)
            r.sw_win_start = P(sw[1], np.int64)
            r.sw_win_len = P(sw[2], np.int32)
            r.sw_offset = P(sw[3], np.int32)
            r.sw_band = P(sw[4], np.int32) if sw[4] is not None else None
            r.sw_abort = P(sw[5], np.int32)
            r.sw_skip_sieve = P(sw[6], np.int32) if len(sw) > 6 and sw[6] is not None else None
        t = C.c_int64()
        st = lib().kira_server_submit(self.h, C.byref(r), C.byref(t))
        if st != KIRA_OK:
            msg = lib().kira_server_last_error(self.h)
            raise HipError(_STATUS.get(st, "Runtime"), msg.decode() if msg else "submit")
        return t.value

    def wait(self, ticket: int) -> dict:
        o = Response()
        st = lib().kira_server_wait(self.h, ticket, C.byref(o))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "Runtime"), "wait")
        try:
            if o.status != KIRA_OK:
                msg = lib().kira_server_last_error(self.h)
                raise HipError(_STATUS.get(o.status, "Runtime"), msg.decode() if msg else "")

            def view(ptr, n, dt):
                if n == 0 or not ptr:
                    return np.zeros(0, dtype=dt)
                return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                             shape=(n,)).copy()

            ns, nw = o.n_sieve, o.n_sw
            cig = view(o.cigar, o.cigar_words, np.uint32)
            sw = []
            for k in range(nw):
                x = o.sw[k]
                sw.append(dict(ref_start=x.ref_start, ref_end=x.ref_end, read_start=x.read_start,
                               read_end=x.read_end, score=x.score,
                               early_abort=bool(x.early_abort), cells=int(x.cells),
                               cigar=[int(v) for v in cig[x.cigar_off:x.cigar_off + x.n_cigar]]))
            return dict(shift=view(o.shift, ns, np.int32), best=view(o.best, ns, np.int32),
                        second=view(o.second, ns, np.int32), flags=view(o.flags, ns, np.uint8),
                        sw=sw, sw_skipped=view(o.sw_skipped, nw, np.uint8),
                        submit_s=o.submit_s, start_s=o.start_s, done_s=o.done_s)
        finally:
            lib().kira_server_release(self.h, ticket)

    def stats(self) -> dict:
        s = ServerStats()
        lib().kira_server_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().kira_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ records
CIGAR_OPS = "MIDNS"


def cigar_string(ops) -> str:
    """CigarOp Display order (src/types.rs:184-197)."""
    return "".join(f"{o >> 4}{CIGAR_OPS[o & 15]}" for o in ops)


def build_alignments(jobs, sw_results):
    """build_alignment (src/alignment/mod.rs:1200-1254) for SW results.

    jobs: sequence of (read: bytes, window: bytes, win_start: int); sw_results:
    the dicts returned by banded_sw_batch / align_batch_simd.  Returns dicts
    {ref_start, ref_end, read_start, read_end, score, nm, md, cigar (text)}."""
    n = len(jobs)
    if n == 0:
        return []
    reads = b"".join(j[0] for j in jobs)
    wins = b"".join(j[1] for j in jobs)
    rl = np.array([len(j[0]) for j in jobs], np.int32)
    wl = np.array([len(j[1]) for j in jobs], np.int32)
    ro = np.concatenate([[0], np.cumsum(rl[:-1], dtype=np.int64)]).astype(np.int64)
    wo = np.concatenate([[0], np.cumsum(wl[:-1], dtype=np.int64)]).astype(np.int64)
    ws = np.array([j[2] for j in jobs], np.uint32)
    sw = (SwResult * n)()
    ops = []
    for k, r in enumerate(sw_results):
        sw[k] = SwResult(r["ref_start"], r["ref_end"], r["read_start"], r["read_end"],
                         r["score"], int(r["early_abort"]), len(r["cigar"]), 0, len(ops))
        ops.extend(r["cigar"])
    opa = np.array(ops if ops else [0], np.uint32)
    rb = np.frombuffer(reads, np.uint8) if reads else np.zeros(1, np.uint8)
    wb = np.frombuffer(wins, np.uint8) if wins else np.zeros(1, np.uint8)
    cu, mu = C.c_int64(), C.c_int64()
    args = [n, _np_ptr(rb), _np_ptr(ro), _np_ptr(rl), _np_ptr(wb), _np_ptr(wo), _np_ptr(wl),
            _np_ptr(ws), sw, _np_ptr(opa)]
    st = lib().kira_build_alignments(*args, None, None, 0, None, 0, C.byref(cu), C.byref(mu))
    if st != KIRA_OK:
        raise HipError(_STATUS.get(st, "InvalidArgument"), "build_alignments")
    out = (AlignmentRec * n)()
    cig = np.zeros(max(cu.value, 1), np.uint32)
    md = C.create_string_buffer(max(mu.value, 1))
    st = lib().kira_build_alignments(*args, out, _np_ptr(cig), cu.value, md, mu.value,
                                     C.byref(cu), C.byref(mu))
    if st != KIRA_OK:
        raise HipError(_STATUS.get(st, "InvalidArgument"), "build_alignments")
    raw = md.raw
    return [dict(ref_start=a.ref_start, ref_end=a.ref_end, read_start=a.read_start,
                 read_end=a.read_end, score=a.score, nm=a.nm,
                 md=raw[a.md_off:a.md_off + a.md_len].decode(),
                 cigar=cigar_string(cig[a.cigar_off:a.cigar_off + a.n_cigar]))
            for a in out]


# ------------------------------------------------------------ stage 4
SIMD_MODES = {"scalar": 0, "avx2": 1, "avxvnni": 2, "neon": 3}
ALN_KINDS = ["accepted_ungapped", "dp_aligned"]


def stage4_cfg(*, match_score=1, mismatch=4, gap_open=6, gap_extend=1, bandwidth=32, xdrop=100,
               clip_penalty=5, adaptive=True, max_drift=16, dp_xdrop=-1, min_chain_ratio=0.4,
               accept_enable=True, accept_only_top1=True, accept_span_slack=15,
               accept_min_identity=98.5, accept_max_mismatches=5, accept_require_score_margin=0,
               dp_topk=1, dp_abort_margin=20, long_read_threshold=500, max_alignments=1,
               tail_ext=True, ungap_min_id=-1, short_dp_topk=0, ambig_div=5, xs_min_ratio_pct=0,
               spectral_pct=3, spectral_floor=2, cert_strict=False, simd="avx2",
               supplementary=True, cpu_mode=False, fast=None, twotier=True,
               twotier_k=5, ungap_gap_check=False) -> Stage4Cfg:
    """Stage-4 configuration; defaults are `mem`'s (src/cli/commands/mem.rs:183-205).
    cpu_mode: the reference's configuration without --gpu (fallback chains try
    the fast DP cascade first); fast: dict of router knob overrides; twotier /
    twotier_k: KIRA_TWOTIER (on) / KIRA_TWOTIER_K (5); ungap_gap_check:
    KIRA_UNGAP_GAP_CHECK (off)."""
    fk = FastKnobs(300, 0, 0, 70, 25, 15, 4, 3, 2, 0, 0, 0, 0)  # router.rs defaults
    for name, v in (fast or {}).items():
        setattr(fk, name, FAST_ALGOS[v] if name == "algo" and isinstance(v, str) else int(v))
    return Stage4Cfg(
        SwCfg(match_score, mismatch, gap_open, gap_extend, bandwidth, xdrop, clip_penalty),
        SwKnobs(int(adaptive), max_drift, dp_xdrop), min_chain_ratio, int(accept_enable),
        int(accept_only_top1), accept_span_slack, accept_min_identity, accept_max_mismatches,
        accept_require_score_margin, dp_topk, dp_abort_margin, long_read_threshold,
        max_alignments, int(tail_ext), ungap_min_id, short_dp_topk, ambig_div, xs_min_ratio_pct,
        spectral_pct, spectral_floor, int(cert_strict), SIMD_MODES[simd], int(supplementary),
        int(cpu_mode), fk, int(twotier), twotier_k, int(ungap_gap_check))


class Stage4Result:
    """Owner of a kira_stage4_result; `alignments()` decodes it."""

    def __init__(self, h):
        self.h = h
        v = Stage4View()
        st = lib().kira_stage4_result_view(h, C.byref(v))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "InvalidArgument"), "stage4_result_view")
        self.view = v
        s = v.stats
        self.stats = {f: getattr(s, f) for f in STAGE4_STATS + STAGE4_STATS_TAIL}
        self.stats["reason_counts"] = list(s.reason_counts)
        self.phase_ms = list(v.phase_ms)

    def alignments(self):
        v = self.view
        n = v.n_reads
        off = np.ctypeslib.as_array(v.aln_off, shape=(n + 1,)).copy()
        cig = np.ctypeslib.as_array(v.cigar, shape=(v.cigar_words,)).copy() \
            if v.cigar_words else np.zeros(0, np.uint32)
        md = C.string_at(v.md, v.md_bytes) if v.md_bytes else b""
        out = []
        for r in range(n):
            recs = []
            for k in range(int(off[r]), int(off[r + 1])):
                a = v.alns[k]
                recs.append(dict(
                    kind=ALN_KINDS[a.kind], ref_id=a.ref_id, ref_start=a.ref_start,
                    ref_end=a.ref_end, read_start=a.read_start, read_end=a.read_end,
                    cigar=cigar_string(cig[a.cigar_off:a.cigar_off + a.n_cigar]),
                    score=a.score, mapq=a.mapq, is_rev=bool(a.is_rev),
                    is_supplementary=bool(a.is_supplementary), nm=a.nm,
                    md=md[a.md_off:a.md_off + a.md_len].decode(),
                    xs_score=a.xs_score if a.has_xs else None))
            out.append(recs)
        return out

    def _refresh(self):
        v = Stage4View()
        st = lib().kira_stage4_result_view(self.h, C.byref(v))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "InvalidArgument"), "stage4_result_view")
        self.view = v

    def left_normalize(self, genome: "Genome", x: "Stage4Inputs", original=None):
        """normalize_alignments (src/alignment/normalize.rs:197-249) in place.
        original: per read the IUPAC-bearing original sequence or None."""
        ob = hb = None
        if original is not None:
            ob = np.frombuffer(b"".join(o if o is not None else r
                                        for o, r in zip(original, x.reads)), np.uint8) \
                if x.rl.sum() else np.zeros(1, np.uint8)
            hb = np.array([o is not None for o in original], np.uint8)
        st = lib().kira_stage4_left_normalize(self.h, genome.h, len(x.co), _np_ptr(x.co),
                                              _np_ptr(x.cl), x.n, _np_ptr(x.rb), _np_ptr(x.ro),
                                              _np_ptr(x.rl),
                                              _np_ptr(ob) if ob is not None else None,
                                              _np_ptr(hb) if hb is not None else None)
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "InvalidArgument"), "stage4_left_normalize")
        self._refresh()

    def sam(self, contig_names, read_ids, x: "Stage4Inputs", quals=None,
            cfg: Optional[SamCfg] = None) -> bytes:
        """SamFormatter::append_alignment / append_unmapped text (src/io/mod.rs:971-1140)
        of every read, single-end.  quals: list of bytes per read or None."""
        cfg = cfg or sam_cfg()
        names = (C.c_char_p * max(len(contig_names), 1))(*[n.encode() for n in contig_names])
        ids = (C.c_char_p * max(x.n, 1))(*[i.encode() for i in read_ids])
        qb = None
        if quals is not None:
            qb = np.frombuffer(b"".join(quals), np.uint8) if x.rl.sum() else np.zeros(1, np.uint8)
        used = C.c_int64()
        args = [self.h, len(contig_names), names, x.n, ids, _np_ptr(x.rb), _np_ptr(x.ro),
                _np_ptr(x.rl), _np_ptr(qb) if qb is not None else None, C.byref(cfg)]
        st = lib().kira_stage4_sam(*args, None, 0, C.byref(used))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "InvalidArgument"), "stage4_sam")
        buf = C.create_string_buffer(max(used.value, 1))
        st = lib().kira_stage4_sam(*args, buf, used.value, C.byref(used))
        if st != KIRA_OK:
            raise HipError(_STATUS.get(st, "InvalidArgument"), "stage4_sam")
        return buf.raw[:used.value]

    def close(self):
        if self.h:
            lib().kira_stage4_result_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CHAIN_DTYPE = np.dtype([("ref_id", "<u4"), ("ref_start", "<u4"), ("ref_end", "<u4"),
                        ("read_start", "<u4"), ("read_end", "<u4"), ("strand", "<i4"),
                        ("score", "<i4")])
assert CHAIN_DTYPE.itemsize == C.sizeof(Chain)


class Stage4Inputs:
    """A stage-4 batch as the flat host arrays kira_stage4_run reads.

    contig i = genome bases contig_off[i] .. +contig_len[i]; reads: list of
    bytes (forward strand); chains: per read, a list of (ref_id, ref_start,
    ref_end, read_start, read_end, strand 0/1, score) in stage-3 order."""

    def __init__(self, contig_off, contig_len, reads, chains, repeat_min_occ=None):
        n = len(reads)
        self.n = n
        self.co = _arr(contig_off, np.int64)
        self.cl = _arr(contig_len, np.int64)
        self.rl = np.array([len(r) for r in reads], np.int32)
        self.ro = np.zeros(n, np.int64)
        if n > 1:
            self.ro[1:] = np.cumsum(self.rl[:-1], dtype=np.int64)
        self.rb = np.frombuffer(b"".join(reads), np.uint8) if self.rl.sum() \
            else np.zeros(1, np.uint8)
        self.coff = np.zeros(n + 1, np.int64)
        self.coff[1:] = np.cumsum([len(c) for c in chains], dtype=np.int64)
        flat = [c for cs in chains for c in cs]
        self.ch = np.array(flat, dtype=CHAIN_DTYPE) if flat else np.zeros(1, CHAIN_DTYPE)
        self.reads = reads
        self.rmo = _arr(repeat_min_occ, np.int32) if repeat_min_occ is not None else None


def stage4_run(ctx: Context, genome: Genome, contig_off, contig_len=None, reads=None,
               chains=None, cfg: Optional[Stage4Cfg] = None,
               repeat_min_occ=None) -> Stage4Result:
    """stage4_alignment::run (src/pipeline/stage4_alignment.rs:959-1268) on the device.

    Either pass a Stage4Inputs as contig_off, or the contig table, reads and
    chains as Stage4Inputs takes them."""
    x = contig_off if isinstance(contig_off, Stage4Inputs) else \
        Stage4Inputs(contig_off, contig_len, reads, chains, repeat_min_occ)
    cfg = cfg or stage4_cfg()
    h = C.c_void_p()
    st = lib().kira_stage4_run(ctx.h, genome.h, len(x.co), _np_ptr(x.co), _np_ptr(x.cl), x.n,
                               _np_ptr(x.rb), _np_ptr(x.ro), _np_ptr(x.rl),
                               _np_ptr(x.rmo) if x.rmo is not None else None, _np_ptr(x.coff),
                               x.ch.ctypes.data_as(C.POINTER(Chain)), C.byref(cfg), C.byref(h))
    ctx.check(st, "stage4_run")
    return Stage4Result(h)


def sam_cfg(write_nm=True, write_md=True, write_as=True, write_xs=True, write_rg=True,
            split_as_secondary=False, soft_clip_supplementary=False,
            read_group: Optional[str] = None) -> SamCfg:
    """OutputConfig::full() defaults (src/io/mod.rs:833-850)."""
    return SamCfg(int(write_nm), int(write_md), int(write_as), int(write_xs), int(write_rg),
                  int(split_as_secondary), int(soft_clip_supplementary),
                  read_group.encode() if read_group is not None else None)


def _concat(seqs):
    """Back-to-back bytes of seqs: (buffer, offsets int64, lengths int32)."""
    ln = np.array([len(x) for x in seqs], dtype=np.int32)
    off = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        off[1:] = np.cumsum(ln[:-1], dtype=np.int64)
    blob = b"".join(seqs)
    buf = np.frombuffer(blob, dtype=np.uint8) if blob else np.zeros(1, np.uint8)
    return buf, len(blob), off, ln


class PairInputs:
    """(pattern, text) pairs as flat host arrays, with a per-pair int (Myers
    max_k or WFA max_score) and an optional second one (WFA
    text_begin_free)."""

    def __init__(self, patterns, texts, a, b=None):
        self.n = len(patterns)
        self.pb, self.pn, self.po, self.pl = _concat(list(patterns))
        self.tb, self.tn, self.to, self.tl = _concat(list(texts))
        self.a = np.ascontiguousarray(a, dtype=np.int32)
        self.b = np.ascontiguousarray(b if b is not None else np.zeros(self.n), dtype=np.int32)


WFA_RES_DTYPE = np.dtype([("found", "<i4"), ("score", "<i4"), ("text_start", "<i4"),
                          ("text_end", "<i4"), ("n_cigar", "<i4"), ("cigar_off", "<i8")],
                         align=True)
assert WFA_RES_DTYPE.itemsize == C.sizeof(WfaResult)


def myers_run(ctx: Context, x: PairInputs):
    """kira_myers_batch on marshalled pairs (a = max_k): (found, dist, end) arrays."""
    n = x.n
    found = np.zeros(max(n, 1), np.uint8)
    dist = np.zeros(max(n, 1), np.int32)
    end = np.zeros(max(n, 1), np.int32)
    st = lib().kira_myers_batch(ctx.h, n, _np_ptr(x.pb), x.pn, _np_ptr(x.po), _np_ptr(x.pl),
                                _np_ptr(x.tb), x.tn, _np_ptr(x.to), _np_ptr(x.tl), _np_ptr(x.a),
                                _np_ptr(found), _np_ptr(dist), _np_ptr(end))
    ctx.check(st, "myers_batch")
    return found[:n], dist[:n], end[:n]


def myers_batch(ctx: Context, pairs):
    """bounded_edit_distance (src/alignment/myers.rs:47-367) over a batch.
    pairs: sequence of (pattern, text, max_k).  Returns (dist, end) or None
    per pair."""
    if not pairs:
        return []
    x = PairInputs([p[0] for p in pairs], [p[1] for p in pairs], [p[2] for p in pairs])
    found, dist, end = myers_run(ctx, x)
    return [(int(dist[k]), int(end[k])) if found[k] else None for k in range(x.n)]


def wfa_run(ctx: Context, x: PairInputs, cfg: SwCfg, adaptive_drop: int = 0,
            want_cigar: bool = True):
    """kira_wfa_batch on marshalled pairs (a = max_score, b =
    text_begin_free): (WFA_RES_DTYPE records, CIGAR words)."""
    n = x.n
    out = np.zeros(max(n, 1), dtype=WFA_RES_DTYPE)
    cap = int((x.pl.astype(np.int64) + x.tl).sum() + 2 * n + 2) if want_cigar else 0
    arena = np.zeros(max(cap, 1), dtype=np.uint32)
    st = lib().kira_wfa_batch(ctx.h, n, _np_ptr(x.pb), x.pn, _np_ptr(x.po), _np_ptr(x.pl),
                              _np_ptr(x.tb), x.tn, _np_ptr(x.to), _np_ptr(x.tl), _np_ptr(x.a),
                              _np_ptr(x.b), int(adaptive_drop), C.byref(cfg),
                              int(bool(want_cigar)), _np_ptr(out), _np_ptr(arena), cap)
    ctx.check(st, "wfa_batch")
    return out[:n], arena


def wfa_batch(ctx: Context, pairs, cfg: SwCfg, adaptive_drop: int = 0, want_cigar: bool = True):
    """wfa_align_scored (src/alignment/wfa.rs:58-67) over a batch.  pairs:
    sequence of (pattern, text, max_score, text_begin_free).  Returns per
    pair None or dict(score (the loss), text_start, text_end, cigar)."""
    if not pairs:
        return []
    x = PairInputs([p[0] for p in pairs], [p[1] for p in pairs], [p[2] for p in pairs],
                   [p[3] if len(p) > 3 else 0 for p in pairs])
    out, arena = wfa_run(ctx, x, cfg, adaptive_drop, want_cigar)
    res = []
    for o in out:
        if not o["found"]:
            res.append(None)
            continue
        c0 = int(o["cigar_off"])
        res.append(dict(score=int(o["score"]), text_start=int(o["text_start"]),
                        text_end=int(o["text_end"]),
                        cigar=[int(v) for v in arena[c0:c0 + int(o["n_cigar"])]]))
    return res


FAST_ALGOS = {"packed": 0, "spectral": 1, "wfa": 2, "sw": 3}


def fast_knobs(**kw) -> FastKnobs:
    """The fast-path router knobs, reference defaults overridden by kw
    (algo may be given by name)."""
    k = FastKnobs()
    lib().kira_fast_knobs_default(C.byref(k))
    for name, v in kw.items():
        if name == "algo" and isinstance(v, str):
            v = FAST_ALGOS[v]
        setattr(k, name, int(v))
    return k


FAST_ALN_DTYPE = np.dtype([("kind", "<i4"), ("ref_start", "<u4"), ("ref_end", "<u4"),
                           ("read_start", "<u4"), ("read_end", "<u4"), ("score", "<i4"),
                           ("nm", "<u4"), ("n_cigar", "<i4"), ("md_len", "<i4"),
                           ("cigar_off", "<i8"), ("md_off", "<i8")], align=True)
assert FAST_ALN_DTYPE.itemsize == C.sizeof(FastAln)


class FastInputs:
    """A fast-path batch as the flat host arrays kira_fast_dp_batch reads:
    contig i = genome bases contig_off[i] .. +contig_len[i]; reads: list of
    strand-oriented bytes; chains: one (ref_id, ref_start, ref_end,
    read_start, read_end, strand, score) per read."""

    def __init__(self, contig_off, contig_len, reads, chains):
        self.n = len(reads)
        self.co = np.ascontiguousarray(contig_off, dtype=np.int64)
        self.cl = np.ascontiguousarray(contig_len, dtype=np.int64)
        self.rb, self.rn, self.ro, self.rl = _concat(list(reads))
        self.ch = (np.array([tuple(c) for c in chains], dtype=CHAIN_DTYPE) if self.n
                   else np.zeros(1, dtype=CHAIN_DTYPE))


def fast_dp_run(ctx: Context, genome: "Genome", x: FastInputs, cfg: SwCfg,
                knobs: Optional[FastKnobs] = None):
    """kira_fast_dp_batch on marshalled inputs.  Returns (records as a
    FAST_ALN_DTYPE array, CIGAR words, MD bytes, stats dict)."""
    n = x.n
    out = np.zeros(max(n, 1), dtype=FAST_ALN_DTYPE)
    band = max(cfg.bandwidth, 1)
    cap = int((2 * x.rl.astype(np.int64) + band + 66).sum()) + 1
    cig = np.zeros(cap, dtype=np.uint32)
    md = np.zeros(4 * cap, dtype=np.uint8)
    cu, mu, stats = C.c_int64(), C.c_int64(), FastStats()
    st = lib().kira_fast_dp_batch(ctx.h, genome.h, len(x.co), _np_ptr(x.co), _np_ptr(x.cl), n,
                                  _np_ptr(x.rb), x.rn, _np_ptr(x.ro), _np_ptr(x.rl), _np_ptr(x.ch),
                                  C.byref(cfg), C.byref(knobs or fast_knobs()), _np_ptr(out),
                                  _np_ptr(cig), cap, _np_ptr(md), len(md), C.byref(cu),
                                  C.byref(mu), C.byref(stats))
    ctx.check(st, "fast_dp_batch")
    sd = {f: getattr(stats, f) for f, _ in FastStats._fields_ if f != "phase_ms"}
    sd["phase_ms"] = dict(zip(["spectral", "myers", "wfa", "records"], list(stats.phase_ms)))
    return out[:n], cig[:cu.value], md[:mu.value].tobytes(), sd


def fast_dp_batch(ctx: Context, genome: "Genome", contig_off, contig_len, reads, chains,
                  cfg: SwCfg, knobs: Optional[FastKnobs] = None):
    """try_fast_dp_alignment_inner (src/alignment/mod.rs:285-440) over a batch.

    Returns (records, stats): per read None or dict(kind 'packed' | 'wfa',
    ref_start, ref_end, score, nm, md, cigar (words))."""
    x = FastInputs(contig_off, contig_len, reads, chains)
    out, cig, md, stats = fast_dp_run(ctx, genome, x, cfg, knobs)
    recs = []
    for o in out:
        if o["kind"] == 0:
            recs.append(None)
            continue
        c0, m0 = int(o["cigar_off"]), int(o["md_off"])
        recs.append(dict(kind="packed" if o["kind"] == 1 else "wfa",
                         ref_start=int(o["ref_start"]), ref_end=int(o["ref_end"]),
                         score=int(o["score"]), nm=int(o["nm"]),
                         md=md[m0:m0 + int(o["md_len"])].decode(),
                         cigar=[int(v) for v in cig[c0:c0 + int(o["n_cigar"])]]))
    return recs, stats


def left_normalize_indels(cigar, contig: bytes, ref_start: int, read: bytes):
    """left_normalize_indels (src/alignment/normalize.rs:55-190), host code.
    cigar: list of len << 4 | op words.  Returns the new CIGAR or None."""
    n = len(cigar)
    ops = np.array(cigar if cigar else [0], np.uint32)
    out = np.zeros(2 * n + 1, np.uint32)
    cb = np.frombuffer(contig, np.uint8) if contig else np.zeros(1, np.uint8)
    rb = np.frombuffer(read, np.uint8) if read else np.zeros(1, np.uint8)
    on = C.c_int32()
    r = lib().kira_left_normalize_indels(_np_ptr(ops), n, _np_ptr(cb), len(contig), ref_start,
                                         _np_ptr(rb), len(read), _np_ptr(out), len(out),
                                         C.byref(on))
    if r < 0:
        raise HipError("InvalidArgument", "left_normalize_indels")
    return [int(v) for v in out[:on.value]] if r == 1 else None
