"""The banded SW autotuner: every route it times returns identical results (the
tuner checks that on its own synthetic batches), the cache file round-trips,
and batches routed by a tuned table, or with every job forced onto each
route in turn, still match the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kira_pkg import load

pytestmark = pytest.mark.gpu
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


def _jobs(seed, shapes):
    """len(shapes) groups of jobs: (band, read_len, n) each, clamp_window windows."""
    rng = np.random.default_rng(seed)
    out = []
    for w, L, n in shapes:
        flank = max(2 * w, 50)
        for _ in range(n):
            g = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), L + 2 * flank))
            read = bytearray(g[flank:flank + L])
            for p in rng.integers(0, L, max(1, L // 40)):
                read[p] = b"ACGT"[(b"ACGT".index(read[p]) + 1) % 4]
            if rng.random() < 0.3:
                k = int(rng.integers(5, L - 5))
                del read[k:k + int(rng.integers(1, 4))]
            out.append((bytes(read), g, flank, ABORT_NONE, w))
    return out


SHAPES = [(16, 150, 600), (32, 120, 300), (48, 200, 300), (100, 300, 80), (180, 700, 16)]


def test_autotune_cache_and_tuned_batches(pkg, ctx, tmp_path):
    cfg = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    path = str(tmp_path / "sw_routes.txt")
    res = pkg.sw_autotune(ctx, SHAPES, cfg, cache_path=path)
    print("[tune]", res, flush=True)
    assert len(res) == len(SHAPES)
    assert all(r["candidates"] >= 1 and r["ms_best"] > 0 for r in res)
    assert all(r["ms_best"] <= r["ms_builtin"] * 1.0001 for r in res)
    lines = [l for l in open(path).read().splitlines() if l and not l.startswith("#")]
    assert len(lines) == len(SHAPES)
    fresh = pkg.Context(0)
    try:
        assert pkg.sw_tune_load(fresh, path) == len(SHAPES)
        jobs = _jobs(3, SHAPES)
        got = pkg.banded_sw_batch(fresh, jobs, cfg, pkg.sw_knobs())
        for k, (read, ref, off, ab, w) in enumerate(jobs):
            want = O.banded_sw(read, ref, off, O.cfg(1, 4, 6, 1, w, 100, 5), ab, O.knobs())
            assert got[k] == want, k
    finally:
        fresh.close()
    # tuning again keeps one line per key (replaced, not appended)
    pkg.sw_autotune(ctx, SHAPES[:2], cfg, cache_path=path)
    lines = [l for l in open(path).read().splitlines() if l and not l.startswith("#")]
    assert len(lines) == len(SHAPES)


@pytest.mark.parametrize("route", list(range(1, 14)))
def test_every_route_matches_oracle(pkg, ctx, tmp_path, route):
    """Every job forced onto one route (where the route holds its band) through
    a cache file whose lines name each shape key of the batch."""
    cfg = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    shapes = [(8, 100, 200), (30, 150, 200), (60, 180, 120), (120, 250, 40), (250, 400, 8)]
    probe = str(tmp_path / "probe.txt")
    keys = [r["key"] for r in pkg.sw_autotune(ctx, shapes, cfg, cache_path=probe)]
    pkg.sw_tune_clear(ctx)
    dev = [l.split()[0] for l in open(probe) if not l.startswith("#")][0]
    path = str(tmp_path / f"forced{route}.txt")
    with open(path, "w") as f:
        for k in keys:
            f.write(f"{dev} {k} {route} 0 0\n")
    local = pkg.Context(0)
    try:
        assert pkg.sw_tune_load(local, path) == len(keys)
        jobs = _jobs(10 + route, shapes)
        rl = np.array([len(j[0]) for j in jobs], np.int32)
        bands = np.array([j[4] for j in jobs], np.int32)
        plan = pkg.BandedSwPlan(local, rl, bands, cfg)
        counts = plan.route_counts()
        plan.close()
        name = pkg.SW_ROUTES[route]
        assert counts.get(name, 0) > 0, counts
        got = pkg.banded_sw_batch(local, jobs, cfg, pkg.sw_knobs())
        for k, (read, ref, off, ab, w) in enumerate(jobs):
            want = O.banded_sw(read, ref, off, O.cfg(1, 4, 6, 1, w, 100, 5), ab, O.knobs())
            assert got[k] == want, (route, k, w)
    finally:
        local.close()


def _forced_context(pkg, ctx, tmp_path, route):
    """A context whose table sends every shape key (band bucket 0..15, read
    length bucket 0..31, count bucket 0..15) to `route`."""
    cfg = pkg.sw_cfg(1, 4, 6, 1, 8, 100, 5)
    probe = str(tmp_path / "probe_dev.txt")
    pkg.sw_autotune(ctx, [(8, 50, 64)], cfg, cache_path=probe)
    pkg.sw_tune_clear(ctx)
    dev = [l.split()[0] for l in open(probe) if not l.startswith("#")][0]
    path = str(tmp_path / f"all{route}.txt")
    with open(path, "w") as f:
        for bb in range(16):
            for lb in range(32):
                for cb in range(16):
                    f.write(f"{dev} {(bb << 16) | (lb << 8) | cb} {route} 0 0\n")
    local = pkg.Context(0)
    assert pkg.sw_tune_load(local, path) == 16 * 32 * 16
    return local


PACKED = [7, 8, 9, 10, 11, 12, 13]


@pytest.mark.parametrize("route", [2, 3, 4, 5, 6] + PACKED)
@pytest.mark.parametrize("adaptive", [False, True])
def test_cooperative_routes_on_tie_corpus(pkg, ctx, tmp_path, route, adaptive):
    """The tie-heavy corpus (low-complexity windows, indels inside runs) with
    every job forced onto one cooperative shape, against the oracle; the packed
    int16 shapes take the cost sets with gap open > 0."""
    import zlib
    from tests.test_oracle_sw_ties import tie_corpus
    cases = tie_corpus(zlib.crc32(b"coop-ties") + route, 2500, 14, 16)
    if route in PACKED:
        cases = [c for c in cases if c[2].gap_open > 0]
    local = _forced_context(pkg, ctx, tmp_path, route)
    try:
        groups = {}
        for case in cases:
            c = case[2]
            groups.setdefault((c.match_score, c.mismatch, c.gap_open, c.gap_extend, c.xdrop,
                               c.clip_penalty), []).append(case)
        n = 0
        for key, group in sorted(groups.items()):
            a, b, o, e, xd, clip = key
            cfg = pkg.sw_cfg(a, b, o, e, 8, xd, clip)
            jobs = [(read, ref, off, ABORT_NONE, c.bandwidth) for read, ref, c, off in group]
            rl = np.array([len(j[0]) for j in jobs], np.int32)
            plan = pkg.BandedSwPlan(local, rl, np.array([j[4] for j in jobs], np.int32), cfg)
            counts = plan.route_counts()
            plan.close()
            assert counts == {pkg.SW_ROUTES[route]: len(jobs)}, counts
            kn = pkg.sw_knobs(adaptive=adaptive, max_drift=2)
            got = pkg.banded_sw_batch(local, jobs, cfg, kn)
            for (read, ref, c, off), g in zip(group, got):
                want = O.banded_sw(read, ref, off, c, ABORT_NONE,
                                   O.knobs(adaptive=adaptive, max_drift=2))
                assert g == want, (route, read, ref, off, c.bandwidth, key)
                n += 1
        assert n > 600
    finally:
        local.close()


# bands that fill each packed shape (2w + 1 = 2 P2 G - 1; 271, the widest
# band a plan takes, for the (18, 32) shape) and the cost bounds of the
# packed form
PACKED_EDGES = [(7, 47), (8, 79), (9, 143), (10, 223), (11, 271), (12, 39), (13, 71)]


@pytest.mark.parametrize("route,w", PACKED_EDGES)
@pytest.mark.parametrize("costs", [(1, 4, 6, 1), (20, 8000, 3500, 500), (2, 3, 1, 1)])
def test_packed_routes_at_band_and_cost_limits(pkg, ctx, tmp_path, route, w, costs):
    """Packed shapes with the band filling the shape, adaptive drift, and costs
    at the packed form's bounds (read_len x match up to 8180)."""
    a, b, o, e = costs
    L = min(400, 8191 // a - 16)  # room for insertions within the score bound
    rng = np.random.default_rng(route * 131 + a)
    jobs = []
    for k in range(96):
        flank = 2 * w + 8
        g = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), L + 2 * flank))
        read = bytearray(g[flank:flank + L])
        for p in rng.integers(0, L, max(1, L // 30)):
            read[p] = b"ACGT"[(b"ACGT".index(read[p]) + 1 + int(rng.integers(0, 3))) % 4]
        for _ in range(int(rng.integers(0, 4))):
            q = int(rng.integers(5, len(read) - 5))
            if rng.random() < 0.5:
                del read[q:q + int(rng.integers(1, 6))]
            else:
                read[q:q] = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8),
                                             int(rng.integers(1, 6))))
        off = flank + int(rng.integers(-w // 2, w // 2 + 1))
        jobs.append((bytes(read), g, off, ABORT_NONE, w - int(k % 3 == 1)))
    cfg = pkg.sw_cfg(a, b, o, e, w, 100, 5)
    local = _forced_context(pkg, ctx, tmp_path, route)
    try:
        rl = np.array([len(j[0]) for j in jobs], np.int32)
        plan = pkg.BandedSwPlan(local, rl, np.array([j[4] for j in jobs], np.int32), cfg)
        counts = plan.route_counts()
        plan.close()
        assert counts == {pkg.SW_ROUTES[route]: len(jobs)}, counts
        for adaptive in (False, True):
            kn = pkg.sw_knobs(adaptive=adaptive, max_drift=8)
            got = pkg.banded_sw_batch(local, jobs, cfg, kn)
            for k, (read, ref, off, ab, wk) in enumerate(jobs):
                want = O.banded_sw(read, ref, off, O.cfg(a, b, o, e, wk, 100, 5), ab,
                                   O.knobs(adaptive=adaptive, max_drift=8))
                assert got[k] == want, (route, k, wk, adaptive, costs)
    finally:
        local.close()
