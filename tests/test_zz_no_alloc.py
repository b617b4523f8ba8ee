"""kira_sieve_run_batch_noalloc against kira_sieve_run_batch through the C ABI:
two calls in a row on the same inputs (the second batch smaller, so stale
staging from the first call would show), and the size limit."""
import random

import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.refgen import cuda_corpus

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    p = load()
    assert p.is_available(), "no GPU visible"
    return p


def _jobs(pkg, corpus):
    jobs = []
    for read, text, mm in corpus:
        rb, _ = O.pack(read)
        tb, _ = O.pack(text)
        jobs.append(pkg.HipJob(rb, len(read), O.pre_shift(tb), len(text), mm))
    return jobs


def _triples(res):
    return [(r.shift, r.mismatches, r.second_mismatches) for r in res]


def _corpus(seed, n):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        L = rng.choice([0, 1, 31, 32, 33, 64, 100, 150, 151, 200, 255, 256])
        W = L + rng.choice([-1, 0, 1, 31, 128, 129, 300])
        if W < 0:
            W = 0
        read = bytes(rng.choice(b"ACGT") for _ in range(L))
        text = bytearray(rng.choice(b"ACGT") for _ in range(W))
        if W >= L and rng.random() < 0.6:
            st = rng.randrange(W - L + 1)
            text[st:st + L] = read
            for _ in range(rng.randrange(5)):
                if L:
                    p = st + rng.randrange(L)
                    text[p] = ord("C") if text[p] != ord("C") else ord("G")
        out.append((read, bytes(text), rng.choice([0, 2, 5, 8, 300, 2 ** 32 - 1])))
    return out


def test_noalloc_matches_run_batch_twice(pkg):
    be = pkg.HipBackend(0)
    big = cuda_corpus(64) + _corpus(11, 3000)
    small = _corpus(12, 700)
    for corpus in (big, small):
        jobs = _jobs(pkg, corpus)
        want = _triples(be.run_batch(jobs))
        got = _triples(be.run_batch(jobs, noalloc=True))
        assert got == want
        assert want == O.backend_run_batch(corpus)
    assert be.run_batch([], noalloc=True) == []


def test_noalloc_errors_like_run_batch(pkg):
    be = pkg.HipBackend(0)
    long_read = pkg.HipJob(b"\x00" * 65, 260, [b"\x00" * 70] * 4, 280, 5)
    with pytest.raises(pkg.HipError) as ei:
        be.run_batch([long_read], noalloc=True)
    assert ei.value.kind == "ReadTooLong"
    # a batch whose staging exceeds the fixed array is refused before any
    # device work, and the entry point still works afterwards
    n = (64 << 20) // 192 + 5000
    job = pkg.HipJob(b"\x1b" * 8, 32, [b"\x1b" * 9] * 4, 36, 4)
    with pytest.raises(pkg.HipError) as ei:
        be.run_batch([job] * n, noalloc=True)
    assert ei.value.kind == "InvalidArgument"
    assert "staging" in str(ei.value)
    corpus = cuda_corpus(64)
    assert _triples(be.run_batch(_jobs(pkg, corpus), noalloc=True)) == \
        O.backend_run_batch(corpus)
