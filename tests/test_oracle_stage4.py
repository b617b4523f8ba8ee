"""The stage-4 driver restatement (oracle/stage4.py) on rule-level cases with
hand-computed answers, and the supplementary picks on the reference's own
split-read unit tests (tests/unit/pipeline_split_read.rs)."""
from oracle import stage4 as S

I32_MIN = -(2 ** 31)


def chain(score, rs, re, ref_id, ref_start=1000, strand=0):
    # make_chain (tests/unit/pipeline_split_read.rs:4-15)
    return (ref_id, ref_start, ref_start + (re - rs), rs, re, strand, score)


def primary(rs, re):
    # make_aln (tests/unit/pipeline_split_read.rs:17-40)
    return dict(read_start=rs, read_end=re, is_rev=False, cigar=f"{re - rs}M")


def test_split_read_known_answers():
    p = S.pick_supplementary_chains
    assert p(primary(0, 80), [chain(100, 0, 80, 0), chain(70, 80, 150, 1)]) == [1]
    assert p(primary(0, 150), [chain(100, 0, 150, 0), chain(90, 10, 150, 1)]) == []
    assert p(primary(0, 80), [chain(100, 0, 80, 0), chain(30, 80, 150, 1)]) == []
    assert p(primary(0, 80), [chain(100, 0, 80, 0), chain(80, 100, 130, 1)]) == []
    cl = [chain(100, 0, 80, 0), chain(80, 80, 150, 1), chain(75, 82, 150, 2)]
    picks = p(primary(0, 80), cl)
    assert len(picks) == 1 and cl[picks[0]][0] == 1
    cl = [chain(200, 0, 200, 0)] + [chain(150, 200 + k * 60, 255 + k * 60, k + 1)
                                    for k in range(6)]
    assert len(p(primary(0, 200), cl)) == S.MAX_SUPP_PER_READ
    assert p(primary(0, 80), [chain(100, 0, 80, 0)]) == []


def test_overlap_pct_known_answers():
    o = S.read_overlap_pct
    assert o(0, 50, 50, 100) == 0
    assert o(0, 100, 0, 100) == 100
    assert o(0, 100, 50, 100) == 50
    assert o(0, 100, 50, 150) == 50
    assert o(50, 100, 0, 100) == 100
    assert o(50, 50, 0, 100) == 0


def test_original_interval_flips_reverse_reads():
    assert S.original_interval(10, 40, True, 150) == (110, 140)
    assert S.original_interval(10, 40, False, 150) == (10, 40)
    assert S.original_interval(0, 200, True, 150) == (0, 150)  # saturating


def test_chain_selection_rules():
    cfg = S.default_cfg()
    c = lambda s: (0, 0, 10, 0, 10, 0, s)  # noqa: E731
    # min_chain_ratio 0.4: ceil(100 * 0.4) = 40, so 39 is skipped but 40 kept
    sel, best, second, margin, *_ = S.select_chains(150, [c(100), c(39), c(40)], 0, cfg)
    assert (best, second, margin) == (100, 39, 61)
    assert sel == [0]  # unambiguous (61 > max(100 / 5, 8)): top-1 only
    # ambiguous: margin 15 <= 20 -> two DP placements, skipping the low one
    sel, *_ = S.select_chains(150, [c(100), c(30), c(85)], 0, cfg)
    assert sel == [0]  # margin = 100 - 30 = 70: not ambiguous
    sel, *_ = S.select_chains(150, [c(100), c(85), c(30), c(80)], 0, cfg)
    assert sel == [0, 1]
    # repeat_min_occ > 1 makes any multi-chain read ambiguous
    sel, *_ = S.select_chains(150, [c(100), c(50)], 2, cfg)
    assert sel == [0, 1]
    # margin floor 8: best 20 -> max(4, 8) = 8
    sel, *_ = S.select_chains(150, [c(20), c(12)], 0, cfg)
    assert sel == [0, 1]
    sel, *_ = S.select_chains(150, [c(20), c(11)], 0, cfg)
    assert sel == [0]
    # long reads use dp_topk
    sel, *_ = S.select_chains(301, [c(100), c(99), c(98)], 0, S.default_cfg(dp_topk=2))
    assert sel == [0, 1]
    # KIRA_SHORT_DPTOPK overrides the short-read top-k
    sel, *_ = S.select_chains(150, [c(100), c(99), c(98)], 0, S.default_cfg(short_dp_topk=3))
    assert sel == [0, 1, 2]
    # non-positive best: min_score = best, so only chains scoring as well pass
    sel, *_ = S.select_chains(150, [c(0), c(-5)], 0, cfg)
    assert sel == [0]
    sel, *_ = S.select_chains(150, [c(-3), c(-3)], 0, cfg)
    assert sel == [0, 1]


def _genome(n, seed=3):
    import random
    rng = random.Random(seed)
    return bytes(rng.choice(b"ACGT") for _ in range(n))


def test_exact_match_shortcut_and_xs_proxy():
    g = _genome(5000)
    read = g[1000:1150]
    chains = [[(0, 1000, 1150, 0, 150, 0, 150), (0, 3000, 3100, 0, 100, 0, 60)]]
    alns, st = S.run([g], [read], chains)
    a, = alns[0]
    assert (a["kind"], a["cigar"], a["ref_start"], a["ref_end"], a["nm"], a["md"], a["mapq"],
            a["score"]) == ("accepted_ungapped", "150M", 1000, 1150, 0, "150", 0, 150)
    # XS = score * second / best = 150 * 60 / 150
    assert a["xs_score"] == 60
    assert st["exact_matches"] == 1 and st["chains_used"] == 1 and st["unmapped"] == 0


def test_reverse_strand_read_and_prefilter_accept():
    g = bytearray(_genome(5000, 4))
    seg = bytes(g[2000:2150])
    read = S.reverse_complement(seg)
    mut = bytearray(seg)
    mut[75] = ord("A") if mut[75] != ord("A") else ord("C")
    read = S.reverse_complement(bytes(mut))
    alns, st = S.run([bytes(g)], [read], [[(0, 2000, 2150, 0, 150, 1, 150)]])
    a, = alns[0]
    # one mismatch: the ungapped prefilter accepts with mapq 60
    assert (a["kind"], a["is_rev"], a["cigar"], a["nm"], a["mapq"]) == \
        ("accepted_ungapped", True, "150M", 1, 60)
    assert a["score"] == 149 - 4
    assert st["prefilter_accept"] == 1


def test_indel_read_goes_to_lane_sw_and_long_read_to_banded():
    g = _genome(20000, 5)
    src = g[5000:5150]
    # an anchor shorter than half the read: the prefilter rejects (span_too_small)
    short = src[:70] + src[73:]
    alns, st = S.run([g], [short], [[(0, 5000, 5070, 0, 70, 0, 70)]])
    assert alns == [[]] and st["prefilter_reject"] == 1 and st["reason_counts"][5] == 1
    assert st["unmapped"] == 1
    read = src[:100] + src[103:]  # 3-bp deletion: the prefilter falls back
    alns, st = S.run([g], [read], [[(0, 5000, 5100, 0, 100, 0, 100)]])
    a, = alns[0]
    assert st["prefilter_fallback"] == 1 and st["dp_simd"] == 1 and st["dp_scalar"] == 0
    assert (a["kind"], a["cigar"], a["ref_start"], a["nm"], a["md"]) == \
        ("dp_aligned", "100M3D47M", 5000, 3, "100^" + src[100:103].decode() + "47")
    long_read = g[8000:8400]
    alns, st = S.run([g], [long_read[:200] + long_read[202:]],
                     [[(0, 8000, 8200, 0, 200, 0, 200)]])
    assert st["dp_scalar"] == 1 and st["dp_simd"] == 0
    assert "2D" in alns[0][0]["cigar"]
    # scalar SIMD mode: no lane jobs at all
    alns, st = S.run([g], [read], [[(0, 5000, 5100, 0, 100, 0, 100)]],
                     S.default_cfg(simd="scalar"))
    assert st["dp_simd"] == 0 and st["dp_scalar"] == 1


def test_sieve_certifies_a_not_top1_chain():
    g = bytearray(_genome(20000, 6))
    src = bytes(g[6000:6150])
    g[12000:12150] = src  # a perfect repeat copy
    mut = bytearray(src)
    mut[60] = ord("A") if mut[60] != ord("A") else ord("G")
    read = bytes(mut)
    # two equal-score chains: the read is ambiguous, both get DP; the second
    # is not top-1 so the prefilter falls back and the sieve certifies it
    chains = [[(0, 6000, 6150, 0, 150, 0, 150), (0, 12000, 12150, 0, 150, 0, 150)]]
    alns, st = S.run([bytes(g)], [read], chains, S.default_cfg(twotier=False))
    assert st["prefilter_accept"] == 1 and st["gpu_spectral_resolved"] == 1
    kinds = [(a["kind"], a["ref_start"], a["mapq"]) for a in alns[0]]
    assert kinds == [("accepted_ungapped", 6000, 60), ("dp_aligned", 12000, 0)]
    assert alns[0][1]["cigar"] == "150M" and alns[0][1]["score"] == 149 - 4
    # with the two-tier reorder (the default) the candidate search certifies
    # both copies (equal scores: dp_limit 2) and stage 4 keeps those results
    alns, st = S.run([bytes(g)], [read], chains)
    assert st["packed_spectral_resolved"] == 2 and st["gpu_spectral_resolved"] == 0
    assert [(a["kind"], a["ref_start"], a["score"]) for a in alns[0]] == \
        [("dp_aligned", 6000, 145), ("dp_aligned", 12000, 145)]


def test_supplementary_alignment_of_a_chimeric_read():
    g = _genome(40000, 7)
    read = g[3000:3080] + g[20000:20070]
    chains = [[(0, 3000, 3080, 0, 80, 0, 80), (0, 20000, 20070, 80, 150, 0, 70)]]
    alns, st = S.run([g], [read], chains)
    assert st["supplementary"] == 1
    assert [a["is_supplementary"] for a in alns[0]] == [False, True]
    sup = alns[0][1]
    assert sup["ref_start"] == 20000 and sup["cigar"] == "80S70M"


def test_cpu_mode_resolves_fallbacks_on_the_fast_path():
    """Without the GPU dispatcher (stage4_alignment.rs:868-882) the deletion read
    that goes to lane SW above is resolved by WFA instead: same CIGAR, the
    full-query DP score, no SW job; a read the cascade declines (foreign tail)
    still reaches SW, after the records of the chains resolved in order."""
    g = _genome(20000, 5)
    src = g[5000:5150]
    read = src[:100] + src[103:]
    ch = [[(0, 5000, 5100, 0, 100, 0, 100)]]
    alns, st = S.run([g], [read], ch, S.default_cfg(cpu_mode=True))
    a, = alns[0]
    assert st["wfa_resolved"] == 1 and st["dp_simd"] == 0 and st["dp_scalar"] == 0
    assert st["prefilter_fallback"] == 1 and st["dp_reads"] == 1
    # src[100] == src[103] == G: the deletion may sit after base 100 or 101;
    # WFA's traceback takes the later place, SW's the earlier one
    assert src[100] == src[103] and src[101] != src[104]
    assert (a["kind"], a["cigar"], a["ref_start"], a["ref_end"], a["nm"], a["mapq"]) == \
        ("dp_aligned", "101M3D46M", 5000, 5150, 3, 0)
    assert a["score"] == 147 - (6 + 3)
    # a read with a foreign 40-base tail: WFA's loss is above 30% and clipping
    # improves it, so the chain falls through to lane SW (which soft-clips)
    tail = _genome(40, 9)
    partial = g[9000:9110] + tail
    alns, st = S.run([g], [partial], [[(0, 9000, 9100, 0, 100, 0, 100)]],
                     S.default_cfg(cpu_mode=True))
    assert st["wfa_resolved"] == 0 and st["dp_simd"] == 1
    assert alns[0][0]["cigar"].endswith("S")
    # the sieve dispatch does not run in cpu mode: a substitution-only read is
    # certified by the fast path's packed scan instead
    sub = bytearray(src)
    sub[60] = ord("A") if sub[60] != ord("A") else ord("C")
    sub[100] = ord("A") if sub[100] != ord("A") else ord("C")
    sub[140] = ord("A") if sub[140] != ord("A") else ord("C")
    ch3 = [[(0, 5000, 5050, 0, 50, 0, 50)]]
    _, st_gpu = S.run([g], [bytes(sub)], ch3)
    _, st_cpu = S.run([g], [bytes(sub)], ch3, S.default_cfg(cpu_mode=True))
    assert st_cpu["gpu_spectral_resolved"] == 0
    assert st_gpu["prefilter_fallback"] == st_cpu["prefilter_fallback"] == 1


def _repeat_genome():
    """A 20 kbp genome with a second copy of 2000..2150 at 8000 that differs
    at three spaced positions."""
    g = bytearray(_genome(20000, 11))
    g[8000:8150] = g[2000:2150]
    for p in (30, 75, 120):
        g[8000 + p] = ord("A") if g[8000 + p] != ord("A") else ord("C")
    return bytes(g)


def test_two_tier_reorders_an_ambiguous_read_and_scores_the_competitor():
    """two_tier_reorder (stage4_alignment.rs:346-451): the read is the second
    copy exactly, but its chain ranks second.  The candidate search scores the
    exact copy A * L = 150 and the first copy by the fast cascade (three
    mismatches: WFA loss 15, score 135); the 15-point gap exceeds the 2 (A + B)
    = 10 margin, so one DP placement (dp_limit 1) on the exact copy, and the
    first copy's affine score becomes XS."""
    g = _repeat_genome()
    read = g[8000:8150]
    chains = [[(0, 2000, 2150, 0, 150, 0, 100), (0, 8000, 8150, 0, 150, 0, 95)]]
    alns, st = S.run([g], [read], chains)
    a, = alns[0]
    assert (a["ref_start"], a["cigar"], a["score"], a["xs_score"]) == (8000, "150M", 150, 135)
    assert st["chains_used"] == 1 and st["exact_matches"] == 1
    assert st["prefilter_accept"] + st["prefilter_fallback"] + st["prefilter_reject"] == 0
    # with the reorder off: ambiguous (margin 5 <= 20), both chains get DP; the
    # exact copy's record comes first, the first copy's lane-SW record after
    alns, st = S.run([g], [read], chains, S.default_cfg(twotier=False))
    assert [x["ref_start"] for x in alns[0]] == [8000, 2000] and st["chains_used"] == 2
    assert st["prefilter_fallback"] == 1 and st["dp_simd"] == 1


def test_two_tier_reuses_the_verified_alignment():
    """A 2-base deletion read of the second copy: both candidates go through the
    cascade (WFA scores 142 and 122, 20 apart: dp_limit 1); the winner's WFA
    alignment is taken as it is, with no prefilter, and the loser's score is XS.
    The weak chain list (best 60 < 2 L / 3) lowers the ratio floor from
    ceil(60 * 0.4) = 24 to 60 / 5 = 12."""
    g = _repeat_genome()
    read = g[8000:8060] + g[8062:8152]
    c = lambda s: [(0, 2000, 2060, 0, 60, 0, 60), (0, 8000, 8060, 0, 60, 0, 58),  # noqa: E731
                   (0, 15000, 15020, 0, 20, 0, s)]
    seen = []
    S.select_chains(150, c(12), 0, S.default_cfg(),
                    lambda passing: seen.append(passing) or ({}, {}, 2, [0, 1, 2]))
    S.select_chains(150, c(11), 0, S.default_cfg(),
                    lambda passing: seen.append(passing) or ({}, {}, 2, [0, 1, 2]))
    assert seen == [[0, 1, 2], [0, 1]]
    sel, _, _, _, _, comp = S.select_chains(
        150, c(11), 0, S.default_cfg(), lambda passing: ({0: 1, 1: 2}, {}, 1, [1, 0, 2]))
    assert sel == [1] and comp == 1          # the ranker's order and dp_limit are used
    cfg = S.default_cfg(supplementary=False)
    alns, st = S.run([g], [read], [c(11)], cfg)
    a, = alns[0]
    assert (a["ref_start"], a["ref_end"], a["kind"], a["mapq"], a["cigar"], a["score"],
            a["xs_score"]) == (8000, 8152, "dp_aligned", 0, "60M2D90M", 142, 122)
    assert st["wfa_resolved"] == 1 and st["chains_used"] == 1
    assert sum(st["reason_counts"]) == 0 and st["dp_simd"] == 0 and st["dp_reads"] == 1
    # a candidate the cascade cannot align is unknown, not poor: dp_limit 2
    alns, st = S.run([g], [read], [c(12)], cfg)
    assert [x["ref_start"] for x in alns[0]] == [8000, 2000] and st["wfa_resolved"] == 2


def test_c_driver_gives_the_records_of_the_python_driver():
    """kira_stage4_oracle.c (the CPU baseline's driver, on pthreads) against
    this restatement with cpu_mode, record for record and counter for
    counter, over configurations that move every branch (lane widths, the
    two-tier reorder, the cascade's algorithms, the gap gate, top-k, XS),
    and independent of the thread count."""
    import pytest
    from tests.stage4_data import make_batch
    b = make_batch(11, 600, n_contigs=3, contig_len=200_000)
    cb = S.CBatch(b["genome"], b["contig_off"], b["contig_len"], b["reads"], b["chains"],
                  b["repeat_min_occ"])
    variants = [dict(), dict(twotier=False), dict(simd="scalar"), dict(simd="avxvnni"),
                dict(simd="neon"), dict(supplementary=False), dict(ungap_gap_check=True),
                dict(fast={"algo": "wfa"}), dict(fast={"algo": "sw"}), dict(short_dp_topk=2),
                dict(xs_min_ratio_pct=50), dict(dp_topk=3, long_read_threshold=100),
                dict(bandwidth=16), dict(twotier_k=3, ambig_div=3),
                dict(fast={"wfa_lead": 8, "quality_gate": True})]
    for k, v in enumerate(variants):
        cfg = S.default_cfg(cpu_mode=True, **v)
        want = S.run(b["contigs"], b["reads"], b["chains"], cfg, b["repeat_min_occ"])
        got = S.run_c(cb, cfg, threads=1 + k % 4)
        assert got[1] == want[1], v
        assert got[0] == want[0], v
    with pytest.raises(ValueError):
        S.run_c(cb, S.default_cfg())
