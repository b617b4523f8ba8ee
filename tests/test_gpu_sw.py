"""GPU parity of the adaptive banded SW (banded_sw_with) vs the CPU oracle."""
import random

import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.refgen import gen_ref

pytestmark = pytest.mark.gpu
NEG = -(2 ** 31) // 4
ABORT_NONE = -(2 ** 31) // 8


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


def mutate(rng, s, sub, indel):
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < indel / 2:
            continue  # deletion
        if r < indel:
            out.append(rng.choice(b"ACGT"))  # insertion before the base
        out.append(rng.choice(b"ACGT".replace(bytes([b]), b"")) if rng.random() < sub else b)
    return bytes(out)


def make_jobs(seed, n, bands=(16, 32, 64, 128), lmin=20, lmax=300):
    rng = random.Random(seed)
    jobs = []
    for k in range(n):
        w = rng.choice(bands)
        L = rng.randrange(lmin, lmax + 1)
        flank = max(2 * w, 50)
        g = bytes(rng.choice(b"ACGT") for _ in range(L + 2 * flank + 20))
        src = g[flank:flank + L]
        read = mutate(rng, src, rng.uniform(0.01, 0.05), 0.005)[:lmax] or b"A"
        ref = g
        off = flank + rng.randrange(-3, 4)
        abort = rng.choice([ABORT_NONE, NEG, len(read) // 2, len(read) + 5, 0])
        jobs.append((read, ref, off, abort, w))
    return jobs


def check(pkg, ctx, jobs, c, kn_gpu, kn_or):
    import time
    t0 = time.time()
    print(f"[sw] batch of {len(jobs)} jobs, bands {sorted(set(j[4] for j in jobs))}",
          flush=True)
    got = pkg.banded_sw_batch(ctx, jobs, c, kn_gpu)
    print(f"[sw] gpu done in {time.time() - t0:.2f}s", flush=True)
    for k, (read, ref, off, abort, w) in enumerate(jobs):
        cc = O.cfg(c.match_score, c.mismatch, c.gap_open, c.gap_extend, w, c.xdrop,
                   c.clip_penalty)
        want = O.banded_sw(read, ref, off, cc, abort, kn_or)
        assert got[k] == want, (k, len(read), len(ref), off, abort, w)


@pytest.mark.parametrize("adaptive", [True, False])
def test_random_jobs_match_oracle(pkg, ctx, adaptive):
    jobs = make_jobs(1 if adaptive else 2, 400)
    c = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    check(pkg, ctx, jobs, c, pkg.sw_knobs(adaptive, 16, -1), O.knobs(adaptive, 16, -1))


def test_edges_offsets_and_odd_bands(pkg, ctx):
    """Bands off the class sizes, centres far off the window, tiny windows,
    empty rows, x-drop aborts and a large match score."""
    rng = random.Random(5)
    jobs = []
    for k in range(300):
        w = rng.choice([1, 2, 3, 5, 7, 8, 11, 20, 33, 40, 100])
        L = rng.randrange(1, 120)
        W = rng.randrange(1, 200)
        read = bytes(rng.choice(b"ACGTN") for _ in range(L))
        ref = bytes(rng.choice(b"ACGT") for _ in range(W))
        if W > L and rng.random() < 0.5:
            st = rng.randrange(W - L)
            ref = ref[:st] + read + ref[st + L:]
            ref = ref[:W]
        off = rng.randrange(-L - 5, W + 5)
        abort = rng.choice([ABORT_NONE, NEG, 5, L])
        jobs.append((read, ref, off, abort, w))
    for cfgv in [(1, 4, 6, 1, 32, 100, 5), (2, 3, 0, 2, 32, 0, 0), (200, 4, 6, 1, 32, 30, 5)]:
        c = pkg.sw_cfg(*cfgv)
        for adaptive in (True, False):
            drift = rng.choice([0, 4, 16])
            check(pkg, ctx, jobs, c, pkg.sw_knobs(adaptive, drift, -1),
                  O.knobs(adaptive, drift, -1))


def test_reference_known_answers(pkg, ctx):
    """src/alignment/mod.rs:2368-2417 and :2592-2635 on the GPU."""
    ref = gen_ref(120)
    read = ref[0:30] + ref[33:60] + ref[63:120]
    c = pkg.sw_cfg(1, 4, 6, 1, 4, 1000, 5)
    ad, = pkg.banded_sw_batch(ctx, [(read, ref, 0, ABORT_NONE)], c, pkg.sw_knobs(True))
    assert sum(o >> 4 for o in ad["cigar"] if o & 15 == 2) == 6
    assert ad["read_end"] - ad["read_start"] >= 113
    ref = gen_ref(90)
    read = ref[:40] + b"TTGCA" + ref[40:]
    r, = pkg.banded_sw_batch(ctx, [(read, ref, 0, ABORT_NONE)], pkg.sw_cfg(1, 4, 6, 1, 20, 1000, 5))
    want = O.banded_sw(read, ref, 0, O.cfg(1, 4, 6, 1, 20, 1000, 5), ABORT_NONE, O.knobs())
    assert r == want


def test_wide_bands_and_long_windows(pkg, ctx):
    """Bands 65..271 (16 and 32 lanes per job) on long windows at long-read error
    rates (5% substitutions, 5% indels), next to narrow-band jobs."""
    rng = random.Random(11)
    jobs = []
    for k in range(96):
        w = rng.choice([65, 70, 71, 72, 100, 128, 129, 135, 136, 199, 200, 207, 208, 250, 271, 16, 40])
        L = rng.randrange(50, 2500)
        flank = max(2 * w, 50)
        g = bytes(rng.choice(b"ACGT") for _ in range(L + 2 * flank + 20))
        read = mutate(rng, g[flank:flank + L], 0.05, 0.05) or b"A"
        off = flank + rng.randrange(-20, 21)
        abort = rng.choice([ABORT_NONE, ABORT_NONE, len(read) // 3])
        jobs.append((read, g, off, abort, w))
    for adaptive in (True, False):
        c = pkg.sw_cfg(1, 4, 6, 1, 32, rng.choice([100, 1000]), 5)
        check(pkg, ctx, jobs, c, pkg.sw_knobs(adaptive, 16, -1), O.knobs(adaptive, 16, -1))


def test_wide_band_limits(pkg, ctx):
    ref = gen_ref(400)
    read = ref[100:200]
    with pytest.raises(pkg.HipError) as e:
        pkg.banded_sw_batch(ctx, [(read, ref, 100, ABORT_NONE, 272)], pkg.sw_cfg())
    assert e.value.kind == "InvalidArgument"
    # scores past the 16-lane kernel's exact range: bands up to 128 still run
    # (lane per job, int32), wider ones are refused
    big = pkg.sw_cfg(30000, 4, 6, 1, 32, 100, 5)
    with pytest.raises(pkg.HipError):
        pkg.banded_sw_batch(ctx, [(read, ref, 100, ABORT_NONE, 200)], big)
    r, = pkg.banded_sw_batch(ctx, [(read, ref, 100, ABORT_NONE, 100)], big)
    want = O.banded_sw(read, ref, 100, O.cfg(30000, 4, 6, 1, 100, 100, 5), ABORT_NONE, O.knobs())
    assert r == want


@pytest.mark.parametrize("widths", [(33, 48, 64), (17, 24, 32)])
def test_large_batch_keeps_lane_per_job_classes(pkg, ctx, widths):
    """A class of >= 65536 jobs of band 17-32 runs lane per job (smaller ones
    move to 16-lane jobs), bands 33-64 run as packed 16-lane jobs; both must
    agree with the oracle."""
    import numpy as np
    rng = random.Random(23 + widths[0])
    jobs = []
    for k in range(66000):
        w = rng.choice(widths)
        L = rng.randrange(40, 80)
        flank = 2 * w
        g = bytes(rng.choice(b"ACGT") for _ in range(L + 2 * flank))
        read = mutate(rng, g[flank:flank + L], 0.03, 0.01) or b"A"
        jobs.append((read, g, flank, ABORT_NONE, w))
    c = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    print(f"[sw] large batch {len(jobs)} jobs, bands {widths}", flush=True)
    got = pkg.banded_sw_batch(ctx, jobs, c, pkg.sw_knobs())
    reads = np.frombuffer(b"".join(j[0] for j in jobs), np.uint8)
    refs = np.frombuffer(b"".join(j[1] for j in jobs), np.uint8)
    rl = np.array([len(j[0]) for j in jobs], np.int32)
    fl = np.array([len(j[1]) for j in jobs], np.int32)
    ro = np.concatenate([[0], np.cumsum(rl[:-1])]).astype(np.int64)
    fo = np.concatenate([[0], np.cumsum(fl[:-1])]).astype(np.int64)
    off = np.array([j[2] for j in jobs], np.int32)
    ab = np.array([j[3] for j in jobs], np.int32)
    bands = np.array([j[4] for j in jobs], np.int32)
    sc, en, cells = O.banded_sw_batch_mt(reads, ro, rl, refs, fo, fl, off, ab,
                                         O.cfg(1, 4, 6, 1, 32, 100, 5), bands=bands, threads=8)
    assert [g["score"] for g in got] == sc.tolist()
    assert [g["ref_end"] for g in got] == en.tolist()
    assert [g["cells"] for g in got] == cells.tolist()
    for k in range(0, 66000, 330):
        read, ref, o, a, w = jobs[k]
        want = O.banded_sw(read, ref, o, O.cfg(1, 4, 6, 1, w, 100, 5), a, O.knobs())
        assert got[k] == want, k
