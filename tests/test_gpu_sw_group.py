"""GPU parity of align_batch_simd's SW (lane groups + scalar fallbacks) vs the oracle."""
import random
import zlib

import pytest

from oracle import oracle as O
from tests.kira_pkg import load

pytestmark = pytest.mark.gpu
NEG = -(2 ** 31) // 4
AN = -(2 ** 31) // 8


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


def make_batches(seed, n_batches, A):
    rng = random.Random(seed)
    batches = []
    for b in range(n_batches):
        size = rng.randrange(0, 40)
        L = rng.randrange(1, 151)
        w = rng.choice([8, 16, 32, 50])
        W = L + 4 * w
        hetero = rng.random() < 0.15
        base = bytes(rng.choice(b"ACGT") for _ in range(W + 300))
        jobs = []
        for k in range(size):
            st = rng.randrange(0, 200)
            Lk = L if not hetero else rng.randrange(1, 151)
            window = base[st:st + (W if not hetero else Lk + 4 * w)]
            p = 2 * w + rng.randrange(-3, 4)
            read = bytearray(window[max(p, 0):max(p, 0) + Lk])
            while len(read) < Lk:
                read.append(rng.choice(b"ACGT"))
            for _ in range(rng.randrange(0, 5)):
                q = rng.randrange(Lk)
                read[q] = rng.choice(b"ACGTN")
            if rng.random() < 0.1:
                read = bytearray(rng.choice(b"ACGT") for _ in range(Lk))
            win_start = 1000 * b
            chain_ref_start = win_start + 2 * w
            abort = rng.choice([AN, NEG, 0, Lk // 2, Lk - 3 if Lk > 3 else 0, Lk * A + 1])
            jobs.append((bytes(read), window, win_start, chain_ref_start, chain_ref_start + Lk,
                         0, Lk, abort))
        batches.append(jobs)
    return batches


@pytest.mark.parametrize("mode,detected", [("avx2", "avx2"), ("avxvnni", "avxvnni"),
                                           ("neon", "neon"), ("scalar", "avx2"),
                                           ("avx2", "avxvnni"), ("neon", "avx2")])
@pytest.mark.parametrize("cfgv", [(1, 4, 6, 1, 32, 100, 5), (2, 4, 6, 1, 32, 40, 0),
                                  (5000, 4, 6, 1, 32, 100, 5)])
def test_align_batch_matches_oracle(pkg, ctx, mode, detected, cfgv):
    # crc32 of the parameters: the same batches in every process (str hashes are salted)
    batches = make_batches(zlib.crc32(repr((mode, detected, cfgv)).encode()) & 0xFFFF, 40,
                           cfgv[0])
    got = pkg.align_batch_simd(ctx, batches, pkg.sw_cfg(*cfgv), mode, detected)
    c = O.cfg(*cfgv)
    for b, (jobs, g) in enumerate(zip(batches, got)):
        want = O.align_batch_sw([O.Job(*j) for j in jobs], c, mode, detected)
        assert g == want, (b, len(jobs))


def _jobs_of(rng, n, L, W, A):
    jobs = []
    base = bytes(rng.choice(b"ACGT") for _ in range(W + 64))
    for _ in range(n):
        st = rng.randrange(0, 32)
        window = base[st:st + W]
        p = rng.randrange(0, max(W - L, 0) + 1)
        read = bytearray(window[p:p + L])
        while len(read) < L:
            read.append(rng.choice(b"ACGT"))
        for _ in range(rng.randrange(0, 6)):
            read[rng.randrange(L)] = rng.choice(b"ACGTN")
        abort = rng.choice([AN, NEG, 0, L // 2, L * A + 1])
        jobs.append((bytes(read), window, 5000, 5000 + p, 5000 + p + L, 0, L, abort))
    return jobs


@pytest.mark.parametrize("wide", [False, True])
def test_align_batch_window_layouts_and_empty_windows(pkg, ctx, wide):
    """Windows up to 1024 columns run the systolic DP (a job's 16-column
    strips on adjacent lanes), wider ones the one-lane-per-job DP; a call
    with one window over 1024 columns takes the second for every job.  Both
    against the oracle, with empty windows and groups of mixed read lengths
    in one call."""
    rng = random.Random(77 + wide)
    cfgv = (1, 4, 6, 1, 32, 100, 5)
    batches = [_jobs_of(rng, 20, 150, 278, 1), _jobs_of(rng, 5, 40, 0, 1),
               _jobs_of(rng, 33, 120, 248, 1), _jobs_of(rng, 3, 16, 16, 1),
               _jobs_of(rng, 7, 150, 1000, 1)]
    if wide:
        batches.append(_jobs_of(rng, 4, 100, 1100, 1))
    got = pkg.align_batch_simd(ctx, batches, pkg.sw_cfg(*cfgv), "avx2", "avx2")
    c = O.cfg(*cfgv)
    for b, (jobs, g) in enumerate(zip(batches, got)):
        want = O.align_batch_sw([O.Job(*j) for j in jobs], c, "avx2", "avx2")
        assert g == want, b
