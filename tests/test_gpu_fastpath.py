"""GPU parity of kira_fast_dp_batch with the cascade restatement
(oracle/fastpath.py): for every chain, whether the fast path resolves it,
by which step, and the record (span, CIGAR, score, NM, MD)."""
import pytest

from oracle import fastpath as F
from oracle import oracle as O
from tests.fastpath_data import make_fast_batch
from tests.kira_pkg import load

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


@pytest.fixture(scope="module")
def batch():
    return make_fast_batch(11, 1200)


def oracle_records(b, c, router):
    out = []
    for read, ch in zip(b["reads"], b["chains"]):
        r = F.fast_dp(read, b["contigs"][ch[0]], ch[1], ch[3], ch[6], c, router)
        if r is None:
            out.append(None)
            continue
        kind, a = r
        out.append(dict(kind=kind, ref_start=a["ref_start"], ref_end=a["ref_end"],
                        score=a["score"], nm=a["nm"], md=a["md"], cigar=a["cigar"]))
    return out


CASES = [
    ("defaults", {}, {}),
    ("wfa-only", dict(algo="wfa"), dict(algo="wfa")),
    ("lead", dict(wfa_lead=8), dict(wfa_lead=8)),
    ("ends-free", dict(wfa_ends_free=6), dict(wfa_ends_free=6)),
    ("adaptive", dict(wfa_adaptive=12), dict(wfa_adaptive=12)),
    ("gate-strict", dict(quality_gate=1, cert_strict=1), dict(quality_gate=True, cert_strict=True)),
    ("algo-unknown", dict(algo=7), dict(algo="bogus")),
    ("tight", dict(myers_pct=5, myers_floor=2, wfa_budget_pct=10, spectral_pct=1,
                   spectral_floor=1), dict(myers_pct=5, myers_floor=2, wfa_budget_pct=10,
                                           spectral_pct=1, spectral_floor=1)),
]


@pytest.mark.parametrize("name,gk,ok", CASES, ids=[c[0] for c in CASES])
def test_fast_dp_matches_oracle(pkg, ctx, batch, name, gk, ok):
    b = batch
    gcfg, ocfg = pkg.sw_cfg(), O.cfg()
    print(f"[fastpath] {name}: {len(b['reads'])} chains", flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    got, stats = pkg.fast_dp_batch(ctx, genome, b["contig_off"], b["contig_len"], b["reads"],
                                   b["chains"], gcfg, pkg.fast_knobs(**gk))
    genome.close()
    want = oracle_records(b, ocfg, F.Router(**ok))
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, [(i, b["chains"][i], got[i], want[i]) for i in bad[:3]]
    kinds = [w["kind"] if w else None for w in want]
    print(f"[fastpath] {name}: stats {stats}, packed {kinds.count('packed')}, "
          f"wfa {kinds.count('wfa')}, none {kinds.count(None)}", flush=True)
    assert stats["spectral"] == kinds.count("packed") and stats["wfa"] == kinds.count("wfa")
    if name == "defaults":
        assert kinds.count("packed") > 100 and kinds.count("wfa") > 100
        assert stats["origin_retries"] > 10 and stats["clip_declined"] > 10


def test_fast_dp_other_costs(pkg, ctx, batch):
    b = batch
    for costs in [(2, 3, 2, 1), (1, 2, 0, 2)]:
        gcfg, ocfg = pkg.sw_cfg(*costs, bandwidth=20), O.cfg(*costs, bandwidth=20)
        genome = pkg.Genome(ctx, b["genome"])
        got, _ = pkg.fast_dp_batch(ctx, genome, b["contig_off"], b["contig_len"], b["reads"],
                                   b["chains"], gcfg)
        genome.close()
        want = oracle_records(b, ocfg, F.Router())
        bad = [i for i in range(len(want)) if got[i] != want[i]]
        assert not bad, [(costs, i, got[i], want[i]) for i in bad[:3]]


def test_fast_dp_empty_and_errors(pkg, ctx, batch):
    b = batch
    genome = pkg.Genome(ctx, b["genome"])
    assert pkg.fast_dp_batch(ctx, genome, b["contig_off"], b["contig_len"], [], [],
                             pkg.sw_cfg())[0] == []
    with pytest.raises(pkg.HipError):
        pkg.fast_dp_batch(ctx, genome, b["contig_off"], b["contig_len"], [b"ACGT"],
                          [(7, 0, 4, 0, 4, 0, 4)], pkg.sw_cfg())
    genome.close()
