"""GPU choice among equal-score alignments: the banded SW kernels and the
full-window lane groups on the tie-heavy corpus of test_oracle_sw_ties (low
complexity windows, indels inside runs, cost sets where open ties extend and a
mismatch ties an insertion plus a deletion), compared field by field with the
oracle and with the alignment the path rule selects."""
import zlib

import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.test_oracle_sw_ties import (ABORT_NONE, TooMany, as_result, co_optimal_paths,
                                       select, tie_corpus, values_banded, values_full_window)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    p = load()
    assert p.is_available(), "no GPU visible"
    return p


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.Context(0)
    yield c
    c.close()


def _key(c):
    return (c.match_score, c.mismatch, c.gap_open, c.gap_extend, c.xdrop, c.clip_penalty)


def _groups(cases):
    """Cases by cost set (one device batch per configuration)."""
    out = {}
    for case in cases:
        out.setdefault(_key(case[2]), []).append(case)
    return out


def _strip(r):
    return {k: r[k] for k in ("read_start", "ref_start", "read_end", "ref_end", "score",
                              "cigar")}


@pytest.mark.parametrize("adaptive", [False, True])
def test_banded_ties_match_oracle_and_rule(pkg, ctx, adaptive):
    cases = tie_corpus(zlib.crc32(b"gpu-banded") + adaptive, 3000, 12, 14)
    n_rule = 0
    for key, group in sorted(_groups(cases).items()):
        a, b, o, e, xd, clip = key
        jobs = [(read, ref, off, ABORT_NONE, c.bandwidth) for read, ref, c, off in group]
        got = pkg.banded_sw_batch(ctx, jobs, pkg.sw_cfg(a, b, o, e, 8, xd, clip),
                                  pkg.sw_knobs(adaptive=adaptive, max_drift=2))
        for (read, ref, c, off), g in zip(group, got):
            want = O.banded_sw(read, ref, off, c, ABORT_NONE,
                               O.knobs(adaptive=adaptive, max_drift=2))
            assert g == want, (read, ref, off, c.bandwidth, key)
            v = values_banded(read, ref, off, c, adaptive=adaptive, max_drift=2)
            try:
                paths = co_optimal_paths(read, ref, c, v)
            except TooMany:
                continue
            assert _strip(g) == as_result(v, select(paths))
            n_rule += len(paths) > 1
    assert n_rule > 250


@pytest.mark.parametrize("mode", ["avx2", "avxvnni", "neon"])
def test_full_window_ties_match_oracle_and_rule(pkg, ctx, mode):
    cases = tie_corpus(zlib.crc32(mode.encode()), 2000, 12, 14)
    n_rule = 0
    for key, group in sorted(_groups(cases).items()):
        a, b, o, e, xd, clip = key
        # one align_batch_simd call per job: each job is its own lane group
        batches = [[(read, ref, 0, 0, len(ref), 0, len(read), ABORT_NONE)]
                   for read, ref, c, off in group]
        got = pkg.align_batch_simd(ctx, batches, pkg.sw_cfg(a, b, o, e, 8, xd, clip), mode)
        for (read, ref, c, off), (g,) in zip(group, got):
            want = O.align_batch_sw([O.Job(read, ref, 0, 0, len(ref), 0, len(read),
                                           ABORT_NONE)], c, mode)[0]
            assert g == want, (read, ref, key, mode)
            v = values_full_window(read, ref, c)
            try:
                paths = co_optimal_paths(read, ref, c, v)
            except TooMany:
                continue
            assert _strip(g) == as_result(v, select(paths))
            n_rule += len(paths) > 1
    assert n_rule > 150
