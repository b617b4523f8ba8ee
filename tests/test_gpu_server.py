"""The continuous-batching server: every request's results equal the oracle's,
and are the same whether it runs alone or among concurrent requests, with
and without one hipGraph per request."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.test_gpu_sieve import I32MAX, _oracle_genome

pytestmark = pytest.mark.gpu
ABORT_NONE = -(2 ** 31) // 8
G = 200_000


@pytest.fixture(scope="module")
def pkg():
    p = load()
    assert p.is_available(), "no GPU visible"
    return p


@pytest.fixture(scope="module")
def world(pkg):
    rng = np.random.default_rng(21)
    g = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=G)
    g[rng.integers(0, G, 40)] = ord("N")
    gseq = bytes(g)
    ctx = pkg.Context(0)
    genome = pkg.Genome(ctx, gseq)
    yield pkg, ctx, genome, gseq
    genome.close()
    ctx.close()


def mutate(rng, s, sub, indel):
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(rng.choice(b"ACGT"))
        out.append(rng.choice(b"ACGT".replace(bytes([b]), b"")) if rng.random() < sub else b)
    return bytes(out)


def make_request(seed, gseq, n_short, n_long):
    """Short reads (20-256 bp) with 8 sieve windows each (true locus first)
    and one banded SW job on the true window dropped if that window is
    certified; long segments (500-3000 bp, 5% subs + 5% indels) with one
    banded SW job each (band 5%)."""
    rng = random.Random(seed)
    reads, sieve, sw = [], ([], [], []), ([], [], [], [], [], [], [])
    for r in range(n_short):
        L = rng.choice([150, 150, rng.randrange(20, 257)])
        pos = rng.randrange(64, G - L - 64)
        read = bytearray(gseq[pos:pos + L])
        for _ in range(rng.choice([0, 0, 1, 2, 6])):
            p = rng.randrange(L)
            read[p] = rng.choice(b"ACGT")
        ri = len(reads)
        reads.append(bytes(read))
        first = len(sieve[0])
        for k in range(8):
            W = L + 128
            s = pos - 64 if k == 0 else rng.randrange(0, G - W)
            sieve[0].append(ri)
            sieve[1].append(s)
            sieve[2].append(W)
        for lst, v in zip(sw, (ri, pos - 64, L + 128, 64, 32, ABORT_NONE, first)):
            lst.append(v)
    for _ in range(n_long):
        S = rng.randrange(500, 3000)
        band = -(-S * 5 // 100)
        pos = rng.randrange(band, G - S - band)
        ri = len(reads)
        reads.append(mutate(rng, gseq[pos:pos + S], 0.05, 0.05) or b"A")
        for lst, v in zip(sw, (ri, pos - band, S + 2 * band, band, band,
                               rng.choice([ABORT_NONE, 20]), -1)):
            lst.append(v)
    lens = np.array([len(r) for r in reads], np.int32)
    offs = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.int64)
    flat = np.frombuffer(b"".join(reads), np.uint8) if reads else np.zeros(0, np.uint8)
    return dict(reads=reads, flat=flat, offs=offs, lens=lens,
                sieve=tuple(np.array(a) for a in sieve), sw=tuple(np.array(a) for a in sw))


def submit(srv, q):
    return srv.submit(q["flat"], q["offs"], q["lens"], sieve=q["sieve"] if len(q["sieve"][0]) else None,
                      sw=q["sw"] if len(q["sw"][0]) else None)


def strip(res):
    return {k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in res.items()
            if not k.endswith("_s")}


def requests(gseq):
    rng = random.Random(5)
    return [make_request(100 + k, gseq, rng.randrange(1, 300), rng.randrange(0, 12))
            for k in range(10)]


def oracle_check(q, res, gseq, cfg):
    sj, ss, sl = q["sieve"]
    want = _oracle_genome(gseq, q["reads"], q["lens"], ss.tolist(), sl.tolist(), sj.tolist(),
                          8, 4, False)
    got = list(zip(res["shift"].tolist(), res["best"].tolist(), res["second"].tolist(),
                   res["flags"].tolist()))
    assert got == want
    rd, ws, wl, off, band, ab, skip = q["sw"]
    for k in range(len(rd)):
        skipped = skip[k] >= 0 and (res["flags"][skip[k]] & 4) != 0
        assert bool(res["sw_skipped"][k]) == skipped
        if skipped:
            continue
        c = O.cfg(cfg.match_score, cfg.mismatch, cfg.gap_open, cfg.gap_extend, int(band[k]),
                  cfg.xdrop, cfg.clip_penalty)
        w = O.banded_sw(q["reads"][rd[k]], gseq[ws[k]:ws[k] + wl[k]], int(off[k]), c, int(ab[k]),
                        O.knobs())
        assert res["sw"][k] == w, (k, int(band[k]))


def test_results_match_oracle_and_are_the_same_alone_and_mixed(world):
    pkg, ctx, genome, gseq = world
    qs = requests(gseq)
    cfg = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    kw = dict(cfg=cfg, sieve_limit=8, cert_max_mism=4, host_staging_bytes=256 << 20)
    print("[server] alone", flush=True)
    alone = []
    srv = pkg.Server(ctx, genome, slots=1, hbm_budget_bytes=1 << 30, **kw)
    for q in qs:
        alone.append(srv.wait(submit(srv, q)))
    srv.close()
    for q, r in zip(qs, alone):
        oracle_check(q, r, gseq, cfg)
    for graph in (False, True):
        print(f"[server] mixed graph={graph}", flush=True)
        # a budget of ~3 requests: admission waits for memory as well as slots
        srv = pkg.Server(ctx, genome, slots=4, hbm_budget_bytes=96 << 20, use_graph=graph, **kw)
        tickets = [submit(srv, q) for q in qs]
        mixed = [srv.wait(t) for t in tickets]
        st = srv.stats()
        srv.close()
        assert st["completed"] == len(qs) and st["peak_in_flight"] >= 2, st
        assert st["peak_device_bytes"] <= 96 << 20
        for a, b in zip(alone, mixed):
            assert strip(a) == strip(b)


def test_request_larger_than_budget_is_refused(world):
    pkg, ctx, genome, gseq = world
    q = make_request(7, gseq, 50, 20)
    srv = pkg.Server(ctx, genome, slots=2, hbm_budget_bytes=1 << 20,
                     host_staging_bytes=64 << 20)
    t = submit(srv, q)
    with pytest.raises(pkg.HipError) as e:
        srv.wait(t)
    assert e.value.kind == "OutOfMemory"
    # the server keeps serving smaller requests
    small = make_request(8, gseq, 3, 0)
    r = srv.wait(submit(srv, small))
    assert len(r["shift"]) == 24
    srv.close()


def test_invalid_requests_fail_at_submit(world):
    pkg, ctx, genome, gseq = world
    srv = pkg.Server(ctx, genome, slots=1, hbm_budget_bytes=64 << 20,
                     host_staging_bytes=64 << 20)
    read = np.frombuffer(b"ACGT" * 40, np.uint8)
    with pytest.raises(pkg.HipError):
        srv.submit(read, [0], [160], sieve=([0], [G - 10], [300]))
    with pytest.raises(pkg.HipError):
        srv.submit(read, [0], [160], sieve=([1], [0], [300]))
    srv.close()


@pytest.mark.parametrize("graph", [False, True])
def test_coalesced_requests_give_the_same_results(world, graph):
    """Waiting requests admitted together as one device batch
    (group_max_reads): each request's results equal its results alone.  The
    budget holds about three requests, so some groups do not fit and hand
    their requests back (the first then runs alone)."""
    pkg, ctx, genome, gseq = world
    qs = requests(gseq)
    cfg = pkg.sw_cfg(1, 4, 6, 1, 32, 100, 5)
    kw = dict(cfg=cfg, sieve_limit=8, cert_max_mism=4, host_staging_bytes=256 << 20)
    srv = pkg.Server(ctx, genome, slots=1, hbm_budget_bytes=1 << 30, **kw)
    alone = [srv.wait(submit(srv, q)) for q in qs]
    srv.close()
    for budget in (1 << 30, 96 << 20):
        print(f"[server] coalesced graph={graph} budget={budget}", flush=True)
        srv = pkg.Server(ctx, genome, slots=2, hbm_budget_bytes=budget, use_graph=graph,
                         group_max_reads=700, **kw)
        tickets = [submit(srv, q) for q in qs]
        grouped = [srv.wait(t) for t in tickets]
        st = srv.stats()
        srv.close()
        assert st["completed"] == len(qs) and st["admitted"] == len(qs), st
        assert st["peak_device_bytes"] <= budget
        if budget == 1 << 30:
            assert st["grouped"] >= 2, st
        for a, b in zip(alone, grouped):
            assert strip(a) == strip(b)
