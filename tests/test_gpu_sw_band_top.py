"""Regression: the lane-per-job band keeps exactly C = 2w + 1 positions when the
band fills its class (w = 4, 8, 16, 32, 64).  When the adaptive centre moves
by less than one column, the band shifts towards higher positions and the
previous row's last cell, the up neighbour of the new last position, must
still feed that position's E.  These jobs (high indel rates, tight bands) hit
that case; the GPU must match the oracle on every one."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kira_pkg import load

pytestmark = pytest.mark.gpu
ABORT_NONE = -(2 ** 31) // 8
M64 = (1 << 64) - 1


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


def _tuner_job(w, L, n, key, want):
    """Job `want` of the SW autotuner's seeded synthetic batch for a shape
    (xorshift64; window = clamp_window geometry; 3% substitutions, 2.5% one-base
    deletions and insertions)."""
    s = (key * 0x9E3779B97F4A7C15 + 0x1234567) & M64

    def nxt():
        nonlocal s
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        return s
    flank = max(2 * w, 50)
    W = L + 2 * flank
    tl = max(4 * W, 1 << 22)
    text = bytes(b"ACGT"[nxt() & 3] for _ in range(tl))
    for k in range(n):
        ws = nxt() % (tl - W)
        src = text[ws + flank:]
        read, got, p = bytearray(), 0, 0
        while got < L and p < L + flank:
            u = nxt() % 1000
            if u < 25:
                p += 1
                continue
            if u < 50:
                read.append(b"ACGT"[nxt() & 3])
                got += 1
                if got == L:
                    break
            b = src[p]
            if nxt() % 100 < 3:
                b = b"ACGT"[nxt() & 3]
            read.append(b)
            got += 1
            p += 1
        if k == want:
            return bytes(read), text[ws:ws + W], flank, w


def test_tuner_batch_jobs(pkg, ctx):
    # band 16: shape keys (bucket 2, log2 length, log4 count) of a 300 bp /
    # 20,000-job and a 200 bp / 83,000-job batch
    jobs = [_tuner_job(16, 300, 20000, (2 << 16) | (8 << 8) | 7, 6315),
            _tuner_job(16, 200, 83000, (2 << 16) | (7 << 8) | 8, 17011)]
    got = pkg.banded_sw_batch(ctx, [(r, f, o, ABORT_NONE, w) for r, f, o, w in jobs],
                              pkg.sw_cfg(1, 4, 6, 1, 16, 100, 5), pkg.sw_knobs())
    for (r, f, o, w), g in zip(jobs, got):
        assert g == O.banded_sw(r, f, o, O.cfg(1, 4, 6, 1, w, 100, 5), ABORT_NONE, O.knobs())


def test_high_indel_tight_bands(pkg, ctx):
    rng = np.random.default_rng(2024)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    jobs = []
    for k in range(20000):
        w = int(rng.choice([4, 8, 16, 32, 64]))
        L = int(rng.integers(150, 301))
        flank = max(2 * w, 50)
        g = acgt[rng.integers(0, 4, L + 2 * flank)]
        src = g[flank:]
        u = rng.random(len(src))
        keep = u >= 0.025                         # 2.5% deletions
        ins = (u >= 0.025) & (u < 0.05)           # 2.5% one-base insertions
        base = src.copy()
        sub = rng.random(len(src)) < 0.04
        base[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
        cnt = keep.astype(np.int64) + ins
        out = np.repeat(base, cnt)
        first = np.cumsum(cnt) - cnt
        out[first[ins]] = acgt[rng.integers(0, 4, int(ins.sum()))]
        read = out[:L].tobytes()
        jobs.append((read, g.tobytes(), flank + int(rng.integers(-3, 4)), ABORT_NONE, w))
    c = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
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
    for k in range(0, len(jobs), 40):
        read, ref, o, a, w = jobs[k]
        assert got[k] == O.banded_sw(read, ref, o, O.cfg(1, 4, 6, 1, w, 100, 5), a, O.knobs()), k
