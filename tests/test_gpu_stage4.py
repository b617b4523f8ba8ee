"""GPU parity of the stage-4 driver (kira_stage4_run) with its CPU
restatement (oracle/stage4.py): every record and every counter."""
import time

import pytest

from oracle import stage4 as S
from tests.kira_pkg import load
from tests.stage4_data import make_batch

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


def run_both(pkg, ctx, b, **cfg):
    print(f"[stage4] {len(b['reads'])} reads, cfg {cfg}", flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    t0 = time.time()
    res = pkg.stage4_run(ctx, genome, b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                         pkg.stage4_cfg(**cfg), b["repeat_min_occ"])
    got, gst = res.alignments(), res.stats
    print(f"[stage4] gpu done in {time.time() - t0:.2f}s, phases {res.phase_ms}", flush=True)
    res.close()
    genome.close()
    want, wst = S.run(b["contigs"], b["reads"], b["chains"], S.default_cfg(**cfg),
                      b["repeat_min_occ"])
    return got, gst, want, wst


@pytest.mark.parametrize("cfg", [dict(), dict(simd="avxvnni"), dict(simd="neon", bandwidth=20),
                                 dict(simd="scalar", supplementary=False, twotier_k=2),
                                 dict(cert_strict=True, min_chain_ratio=0.2, dp_abort_margin=5,
                                      twotier=False),
                                 dict(twotier=False, short_dp_topk=2),
                                 dict(ungap_gap_check=True)])
def test_records_and_counters_match_oracle(pkg, ctx, cfg):
    b = make_batch(11, 2500)
    got, gst, want, wst = run_both(pkg, ctx, b, **cfg)
    for r in range(len(want)):
        assert got[r] == want[r], (r, b["chains"][r])
    assert gst == wst
    # every branch of the driver ran
    if not cfg:
        for k in ("exact_matches", "prefilter_accept", "prefilter_reject",
                  "gpu_spectral_resolved", "dp_simd", "dp_scalar", "supplementary",
                  "unmapped", "dp_early_abort"):
            assert gst[k] > 0, k


@pytest.mark.parametrize("cfg", [dict(cpu_mode=True), dict(cpu_mode=True, simd="scalar"),
                                 dict(cpu_mode=True, simd="avxvnni", twotier_k=3),
                                 dict(cpu_mode=True, simd="neon", twotier=False, bandwidth=24),
                                 dict(cpu_mode=True, fast=dict(wfa_lead=6, wfa_adaptive=10)),
                                 dict(cpu_mode=True, fast=dict(algo="wfa", cert_strict=1))])
def test_cpu_mode_fast_path_matches_oracle(pkg, ctx, cfg):
    """The configuration without the GPU dispatcher: fallback chains try the
    fast DP cascade before SW (stage4_alignment.rs:868-882)."""
    b = make_batch(14, 2500)
    got, gst, want, wst = run_both(pkg, ctx, b, **cfg)
    for r in range(len(want)):
        assert got[r] == want[r], (r, b["chains"][r])
    assert gst == wst
    assert gst["gpu_spectral_resolved"] == 0 and gst["wfa_resolved"] > 0
    if cfg.get("fast", {}).get("algo") != "wfa":
        assert gst["packed_spectral_resolved"] > 0


def test_empty_batch_and_reads_without_chains(pkg, ctx):
    b = make_batch(12, 20)
    genome = pkg.Genome(ctx, b["genome"])
    res = pkg.stage4_run(ctx, genome, b["contig_off"], b["contig_len"], [], [])
    assert res.alignments() == [] and res.stats["reads"] == 0
    res = pkg.stage4_run(ctx, genome, b["contig_off"], b["contig_len"], b["reads"][:5],
                         [[]] * 5)
    assert res.alignments() == [[]] * 5 and res.stats["unmapped"] == 5
    bad = [[(7, 0, 10, 0, 10, 0, 10)]]  # contig 7 does not exist
    with pytest.raises(pkg.HipError):
        pkg.stage4_run(ctx, genome, b["contig_off"], b["contig_len"], b["reads"][:1], bad)
    genome.close()


@pytest.mark.parametrize("sam_kw", [dict(), dict(split_as_secondary=True, write_md=False),
                                    dict(soft_clip_supplementary=True,
                                         read_group="@RG\tID:g7\tSM:s")])
def test_left_normalized_sam_matches_oracle(pkg, ctx, sam_kw):
    # the SW tracebacks already place most gaps leftmost, so few records move
    # here; the moving cases are pinned on the host by tests/test_oracle_sam.py
    import random
    from oracle import sam as OS
    b = make_batch(13, 1500)
    rng = random.Random(3)
    quals = [bytes(rng.randrange(33, 74) for _ in r) for r in b["reads"]]
    ids = [f"r{k}:x" for k in range(len(b["reads"]))]
    names = [f"chr{k + 1}" for k in range(len(b["contigs"]))]
    print(f"[stage4] sam {sam_kw}", flush=True)
    genome = pkg.Genome(ctx, b["genome"])
    x = pkg.Stage4Inputs(b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                         b["repeat_min_occ"])
    res = pkg.stage4_run(ctx, genome, x)
    res.left_normalize(genome, x)
    got_recs = res.alignments()
    got = res.sam(names, ids, x, quals, pkg.sam_cfg(**sam_kw))
    res.close()
    genome.close()
    want_recs, _ = S.run(b["contigs"], b["reads"], b["chains"], S.default_cfg(),
                         b["repeat_min_occ"])
    OS.normalize_alignments(b["contigs"], b["reads"], want_recs)
    assert got_recs == want_recs
    want = OS.sam_text(ids, b["reads"], quals, want_recs, names, OS.default_cfg(**sam_kw))
    assert got == want
    assert got.count(b"\n") == sum(max(len(r), 1) for r in want_recs)
