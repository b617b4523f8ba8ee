#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the CPU oracle.

Every fixture is data only (inputs and expected outputs, JSON).  The inputs
are the reference's own test inputs where it has them (the GPU corpus of
tests/cuda_spectral.rs:13-47, the prefilter / MD cases of
tests/prefilter_accept.rs and tests/md_format.rs, the banded SW cases of
src/alignment/mod.rs:2368-2500) plus small seeded random cases; the expected
outputs come from the oracle after it has been checked against brute force
(sieve) or a textbook Gotoh local alignment (SW, full band) in this script.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.json
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests.refgen import cuda_corpus, gen_ref  # noqa: E402

ABORT_NONE = -(2 ** 31) // 8


def brute_best_second(read, text):
    if not read or len(text) < len(read):
        return None
    m = [sum(a != b for a, b in zip(read, text[t:t + len(read)]))
         for t in range(len(text) - len(read) + 1)]
    best = min(m)
    shift = m.index(best)
    s = sorted(m)
    return shift, best, (s[1] if len(s) > 1 else None)


def sieve_fixture():
    corpus = cuda_corpus(64)
    got = O.backend_run_batch(corpus)
    for (read, text, mm), (sh, m, sec) in zip(corpus, got):
        bb = brute_best_second(read, text)
        lim = min(mm, len(read))
        assert (sec if sec is not None else None) == bb[2]
        if bb[1] <= lim:
            assert (sh, m) == (bb[0], bb[1])
        else:
            assert (sh, m) == (-1, 2 ** 31 - 1)
    return {"source": "tests/cuda_spectral.rs:13-47 corpus (xorshift 0xCAFE_F00D, 64 jobs)",
            "jobs": [{"read": r.decode(), "text": t.decode(), "max_mismatches": mm,
                      "shift": sh, "mismatches": m, "second": sec}
                     for (r, t, mm), (sh, m, sec) in zip(corpus, got)]}


def gotoh_local(read, ref, A, B, O_, E):
    """Textbook local affine alignment score (Gotoh), full matrix."""
    NEG = -10 ** 9
    n, m = len(read), len(ref)
    H = [[0] * (m + 1) for _ in range(n + 1)]
    Ev = [[NEG] * (m + 1) for _ in range(n + 1)]
    F = [[NEG] * (m + 1) for _ in range(n + 1)]
    best = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            Ev[i][j] = max(H[i - 1][j] - (O_ + E), Ev[i - 1][j] - E)
            F[i][j] = max(H[i][j - 1] - (O_ + E), F[i][j - 1] - E)
            s = A if read[i - 1] == ref[j - 1] else -B
            H[i][j] = max(0, H[i - 1][j - 1] + s, Ev[i][j], F[i][j])
            best = max(best, H[i][j])
    return best


def sw_fixture():
    cases = []
    # the reference's adaptive-band regression inputs (mod.rs:2344-2500)
    ref = gen_ref(120)
    cases.append((ref[0:30] + ref[33:60] + ref[63:120], ref, 0, (1, 4, 6, 1, 4, 1000, 5)))
    cases.append((ref[10:90], ref, 10, (1, 4, 6, 1, 6, 100, 5)))
    r90 = gen_ref(90)
    cases.append((r90[:40] + b"TTGCA" + r90[40:], r90, 0, (1, 4, 6, 1, 20, 1000, 5)))
    rng = random.Random(77)
    for k in range(40):
        L = rng.randrange(5, 60)
        W = rng.randrange(5, 90)
        read = bytes(rng.choice(b"ACGT") for _ in range(L))
        refb = bytearray(rng.choice(b"ACGT") for _ in range(W))
        if W > L and rng.random() < 0.7:
            st = rng.randrange(W - L)
            refb[st:st + L] = read
            for _ in range(rng.randrange(0, 4)):
                refb[rng.randrange(W)] = rng.choice(b"ACGT")
        w = rng.choice([2, 5, 16, 40, 100])
        cfgv = rng.choice([(1, 4, 6, 1, w, 100, 5), (2, 3, 5, 2, w, 0, 0)])
        cases.append((read, bytes(refb), rng.randrange(-5, W), cfgv))
    out, checked = [], []
    for read, refb, off, cfgv in cases:
        c = O.cfg(*cfgv)
        r = O.banded_sw(read, refb, off, c, ABORT_NONE, O.knobs())
        if cfgv[4] >= max(len(read), len(refb)) + abs(off) + 17 and cfgv[5] == 0:
            # a band covering the whole matrix (centre drift <= 16) without
            # x-drop and clip bonus is plain local SW
            assert r["score"] == gotoh_local(read, refb, cfgv[0], cfgv[1], cfgv[2], cfgv[3])
            checked.append(1)
        out.append({"read": read.decode(), "ref": refb.decode(), "offset": off,
                    "cfg": list(cfgv), "abort_score": ABORT_NONE, "adaptive": True,
                    "expect": {k: r[k] for k in ("ref_start", "ref_end", "read_start",
                                                 "read_end", "score", "early_abort", "cells",
                                                 "cigar")}})
    assert len(checked) >= 5
    return {"source": "src/alignment/mod.rs:2344-2500 inputs + seeded random (Random(77))",
            "cases": out}


def prefilter_fixture():
    base = b"ACGT" * 37 + b"AC"
    one = bytearray(base)
    one[10] = ord("T")
    two = bytearray(base)
    two[5], two[99] = ord("T"), ord("A")
    amb = bytearray(b"ACGT" * 38)
    amb[70] = ord("N")
    cases = [  # (read, ref, span, kwargs)
        (base, bytes(one), (0, 150, 0, 150), {}),
        (b"A" * 400, b"T" * 400, (0, 44, 0, 44), {}),
        (b"A" * 400, b"T" * 400, (0, 400, 0, 400), {"is_top1": False}),
        (b"A" * 150, b"T" * 150, (0, 150, 0, 150), {}),
        (b"A" * 400, b"T" * 400, (0, 400, 0, 400), {}),
        (base, bytes(two), (0, 150, 0, 150), {"accept_span_slack": 15, "accept_max_mismatches": 5}),
        (bytes(amb), bytes(amb), (0, 152, 0, 152), {"accept_span_slack": 15,
                                                    "accept_max_mismatches": 5}),
    ]
    c = O.cfg(1, 4, 6, 1, 50, 50, 5)
    out = []
    for read, ref, span, kw in cases:
        kw2 = dict(accept_span_slack=5, accept_max_mismatches=2)
        kw2.update(kw)
        r = O.prefilter_chain(read, ref, span, c, **kw2)
        out.append({"read": read.decode(), "ref": ref.decode(), "span": list(span),
                    "args": kw2, "expect": r})
    return {"source": "tests/prefilter_accept.rs:18-198, tests/md_format.rs:76-186",
            "cfg": [1, 4, 6, 1, 50, 50, 5], "cases": out}


def main():
    for name, fn in (("sieve_cuda_corpus.json", sieve_fixture), ("banded_sw_cases.json", sw_fixture),
                     ("prefilter_cases.json", prefilter_fixture)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
            f.write("\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
