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


@pytest.mark.parametrize("route", [1, 2, 3, 4, 5, 6])
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
