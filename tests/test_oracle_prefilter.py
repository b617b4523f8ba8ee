"""Pins of the oracle's ungapped prefilter (prefilter_chain) against the
reference's own tests and against properties fixed by its definition."""
import random

import pytest

from oracle import oracle as O

# tests/prefilter_accept.rs:6-16 and tests/md_format.rs:13-23
CFG = O.cfg(1, 4, 6, 1, 50, 50, 5)


def acgt150():
    return b"ACGT" * 37 + b"AC"


def full_span(n):
    return (0, n, 0, n)


def run(read, ref, span, **kw):
    kw.setdefault("accept_span_slack", 5)
    kw.setdefault("accept_max_mismatches", 2)
    return O.prefilter_chain(read, ref, span, CFG, **kw)


def canonical_md(read, ref):
    """SAM-spec MD / NM for an ungapped alignment (tests/md_format.rs:26-43)."""
    md, run_, nm = "", 0, 0
    for q, r in zip(read, ref):
        if q == r:
            run_ += 1
        else:
            nm += 1
            md += f"{run_}{chr(r)}"
            run_ = 0
    return md + str(run_), nm


# ---------------------------------------------------------------- reference tests
def test_accepts_short_top1_low_mism():
    # tests/prefilter_accept.rs:18-56
    read = acgt150()
    ref = bytearray(read)
    ref[10] = ord("T")
    o = run(read, bytes(ref), full_span(150))
    assert (o["result"], o["reason"]) == ("accept", "accept")
    assert o["alignment"]["cigar"] == "150M"


def test_rejects_short_span():
    # tests/prefilter_accept.rs:58-93: 44-bp anchor of all-mismatch sequence
    o = run(b"A" * 400, b"T" * 400, (0, 44, 0, 44))
    assert (o["result"], o["reason"]) == ("reject", "span_too_small")


def test_not_top1_fallback():
    # tests/prefilter_accept.rs:95-128
    o = run(b"A" * 400, b"T" * 400, full_span(400), is_top1=False)
    assert (o["result"], o["reason"]) == ("fallback", "not_top1")


def test_mism_too_high():
    # tests/prefilter_accept.rs:130-163
    o = run(b"A" * 150, b"T" * 150, full_span(150))
    assert (o["result"], o["reason"]) == ("fallback", "mism_too_high")


def test_not_short_preset():
    # tests/prefilter_accept.rs:165-198 (400 bp > long_read_threshold 300)
    o = run(b"A" * 400, b"T" * 400, full_span(400))
    assert (o["result"], o["reason"]) == ("fallback", "not_short_preset")


@pytest.mark.parametrize("length,mism,expect", [(150, 0, 10000), (150, 1, 9933), (100, 2, 9800)])
def test_identity_known_answers(length, mism, expect):
    # tests/ungapped_stats.rs:1-8
    assert O.identity_x10000(length, mism) == expect


def md_case(edits):
    read = acgt150()
    ref = bytearray(read)
    for pos, b in edits:
        ref[pos] = b
    return read, bytes(ref)


@pytest.mark.parametrize("edits,nm", [
    ([], 0),                                   # md_format.rs:76-88
    ([(10, ord("T"))], 1),                     # :90-102
    ([(5, ord("T")), (99, ord("A"))], 2),      # :104-118
    ([(149, ord("A"))], 1),                    # :120-133 (last base is C)
    ([(0, ord("T"))], 1),                      # :135-147
])
def test_md_matches_canonical(edits, nm):
    read, ref = md_case(edits)
    o = run(read, ref, full_span(150), accept_span_slack=15, accept_max_mismatches=5)
    assert o["result"] == "accept", o
    md, enm = canonical_md(read, ref)
    assert o["alignment"]["md"] == md
    assert o["alignment"]["nm"] == enm == nm


def test_ambiguous_identical_base_counts_in_nm():
    # tests/md_format.rs:183-185: an N matched against N is an edit in NM only
    read = bytearray(b"ACGT" * 38)
    read[70] = ord("N")
    o = run(bytes(read), bytes(read), full_span(152), accept_span_slack=15,
            accept_max_mismatches=5)
    assert o["result"] == "accept"
    assert (o["alignment"]["md"], o["alignment"]["nm"]) == ("152", 1)


# ---------------------------------------------------------------- definitional properties
def rand_seq(rng, n):
    return bytes(rng.choice(b"ACGT") for _ in range(n))


def test_exact_copy_extends_to_whole_read():
    rng = random.Random(3)
    for _ in range(50):
        L = rng.randint(40, 300)
        ref = rand_seq(rng, L + 400)
        off = rng.randint(0, 400)
        read = ref[off:off + L]
        a = rng.randint(0, L - 1)
        b = rng.randint(a + 1, L)
        o = run(read, ref, (off + a, off + b, a, b))
        m = o["metrics"]
        # full-length ungapped copy: the extension reaches both read ends, every
        # base matches, score = L * match
        assert (m["len"], m["mism"], m["read_start"], m["ref_start"]) == (L, 0, 0, off)
        assert m["score"] == L and m["identity_x10000"] == 10000
        assert o["result"] == ("accept" if L <= 300 else "fallback")


def test_left_extension_is_block_granular_without_tail():
    # tail extension off: the left extension is a multiple of 32 bases
    rng = random.Random(4)
    for _ in range(40):
        L = rng.randint(60, 200)
        ref = rand_seq(rng, L + 100)
        read = ref[50:50 + L]
        a = rng.randint(1, L - 10)
        o = run(read, ref, (50 + a, 50 + a + 5, a, a + 5), tail_ext=False)
        m = o["metrics"]
        left = a - m["read_start"]
        right = m["len"] - 5 - left
        assert left % 32 == 0 and right % 32 == 0
        assert left == (a // 32) * 32 and right == ((L - a - 5) // 32) * 32


def test_xdrop_stops_extension_at_a_bad_block():
    # one 32-base block of pure mismatches costs 32 * 4 = 128 > xdrop 50: the
    # extension stops before it and never reaches the good sequence beyond
    rng = random.Random(5)
    ref = rand_seq(rng, 300)
    read = bytearray(ref)
    for k in range(100, 132):
        read[k] = ord("A") if ref[k] != ord("A") else ord("C")
    o = run(bytes(read), ref, (200, 260, 200, 260))
    m = o["metrics"]
    assert m["read_start"] == 132 + ((200 - 132) % 32)
    assert m["mism"] == 0


def test_accept_gates_follow_their_order_and_thresholds():
    rng = random.Random(6)
    ref = rand_seq(rng, 150)
    one = bytearray(ref)
    one[75] = ord("A") if ref[75] != ord("A") else ord("C")
    o = run(bytes(one), ref, full_span(150), accept_enable=False)
    assert (o["result"], o["reason"]) == ("fallback", "disabled")
    o = run(bytes(one), ref, full_span(150), multi_alignments_enabled=True)
    assert o["reason"] == "max_alignments_not_one"
    o = run(bytes(one), ref, full_span(150), accept_require_score_margin=10, score_margin=9)
    assert o["reason"] == "margin_fail"
    o = run(bytes(one), ref, full_span(150), accept_require_score_margin=10, score_margin=10)
    assert o["reason"] == "accept"
    # co-linearity: ref span 150, read span 149
    o = run(bytes(one), ref, (0, 150, 0, 149))
    assert o["reason"] == "indel_suspect"
    # identity 149/150 = 9933 < floor 9950
    o = run(bytes(one), ref, full_span(150), ungap_min_id=9950)
    assert o["reason"] == "id_too_low"
    o = run(bytes(one), ref, full_span(150), accept_min_identity=99.34)
    assert o["reason"] == "id_too_low"
    o = run(bytes(one), ref, full_span(150), accept_min_identity=99.33)
    assert o["reason"] == "accept"
    # hard mismatch cap ceil(5 L / 150): 3 mismatches at L = 150 allowed only
    # up to 5, while accept_max_mismatches = 2 caps it first
    three = bytearray(one)
    for k in (20, 120):
        three[k] = ord("A") if ref[k] != ord("A") else ord("C")
    o = run(bytes(three), ref, full_span(150), accept_max_mismatches=9, ungap_min_id=0)
    assert o["reason"] == "accept"
    o = run(bytes(three), ref, full_span(150), ungap_min_id=0)
    assert o["reason"] == "mism_too_high"


def test_early_exits():
    o = run(b"", b"ACGT", (0, 0, 0, 0))
    assert o["reason"] == "read_len0"
    o = run(b"ACGT", b"ACGT", (2, 1, 0, 1))
    assert o["reason"] == "non_monotonic"
    o = run(b"ACGT", b"ACGT", (1, 1, 1, 1))
    assert o["reason"] == "anchor0"


def test_accepted_alignment_replays():
    rng = random.Random(7)
    for _ in range(200):
        L = rng.randint(20, 250)
        ref = rand_seq(rng, L + 80)
        off = rng.randint(0, 80)
        read = bytearray(ref[off:off + L])
        for _ in range(rng.randint(0, 3)):
            p = rng.randrange(L)
            read[p] = rng.choice(b"ACGT")
        a = rng.randint(0, L - 1)
        b = rng.randint(a + 1, L)
        o = run(bytes(read), ref, (off + a, off + b, a, b), accept_max_mismatches=5,
                ungap_min_id=rng.choice([-1, 0, 9800]))
        if o["result"] != "accept":
            continue
        al = o["alignment"]
        s, e, fs = al["read_start"], al["read_end"], al["ref_start"]
        q, r = bytes(read[s:e]), ref[fs:fs + (e - s)]
        md, nm = canonical_md(q, r)
        assert (al["md"], al["nm"]) == (md, nm)
        assert al["score"] == (e - s - nm) - 4 * nm
        clip = (f"{s}S" if s else "") + f"{e - s}M" + (f"{L - e}S" if e < L else "")
        assert al["cigar"] == clip


# ---- KIRA_UNGAP_GAP_CHECK: ungapped_beaten_by_gap (mod.rs:163-204) and the
# final gate of prefilter_chain (prefilter.rs:274-287)

def _gap_cfg(bw=50, a=1, b=4):
    return O.cfg(a, b, 6, 1, bw, 1000, 5)


def test_gap_gate_known_answers():
    """mod.rs:2502-2550 (gap_gate_flags_indel_not_snps): two substitutions are
    never beaten by a gap; a 2-bp deletion at read position 48 (the ungapped
    placement smears the tail into 6 mismatches) is; an exact read never is."""
    from tests.refgen import gen_ref
    ref = gen_ref(160)
    flip = lambda x: ord("C") if x == ord("A") else ord("A")
    span = lambda n: (10, 10 + n, 0, n)
    snp = bytearray(ref[10:70])
    snp[20] = flip(snp[20])
    snp[40] = flip(snp[40])
    assert not O.ungapped_beaten_by_gap(bytes(snp), ref, span(60), 2, _gap_cfg())
    dele = ref[10:58] + ref[60:72]
    assert O.ungapped_beaten_by_gap(dele, ref, span(len(dele)), 6, _gap_cfg())
    exact = ref[10:70]
    assert not O.ungapped_beaten_by_gap(exact, ref, span(60), 0, _gap_cfg())


def test_gap_gate_budget_is_strict():
    """The probe's budget is one below the ungapped cost, mism * (A + B) - 1
    (mod.rs:193-201), so a gapped alignment must be strictly cheaper.  A 1-bp
    deletion costs gap open 6 + deletion extend 1 = 7 in WFA terms; with
    A + B = 7 one mismatch ties it (not beaten) and two do not (beaten)."""
    from tests.refgen import gen_ref
    ref = gen_ref(160)
    read = ref[10:80] + ref[81:100]
    sp = (10, 10 + len(read), 0, len(read))
    c = _gap_cfg(a=1, b=6)
    assert not O.ungapped_beaten_by_gap(read, ref, sp, 1, c)
    assert O.ungapped_beaten_by_gap(read, ref, sp, 2, c)
    # windows: a chain implying a negative or past-the-end window start is
    # never probed (mod.rs:177-184)
    assert not O.ungapped_beaten_by_gap(read, ref, (5, 100, 6, 100), 2, c)
    assert not O.ungapped_beaten_by_gap(read, ref, (170, 180, 0, 10), 2, c)


def _end_indel_read(rng, ref, st, L):
    """A read from ref[st:] with one indel a few bases before a run of A that
    reaches the read end, laid out so the ungapped placement shows exactly two
    mismatches (u/v before the run), which one gap explains: a 1-bp deletion of
    u, or a 1-bp insertion w before v."""
    ref = bytearray(ref)
    q = rng.randrange(L - 40, L - 4)
    u, v, w = rng.sample(b"CGT", 3)
    ref[st + q] = u
    ref[st + q + 1] = v
    ref[st + q + 2:st + L + 2] = b"A" * (L - q)
    if rng.random() < 0.5:
        read = ref[st:st + q] + ref[st + q + 1:st + L + 1]       # deletion of u
    else:
        read = ref[st:st + q] + bytes([w]) + ref[st + q:st + L - 1]  # insertion of w
    return bytes(read), bytes(ref)


def test_gap_gate_turns_end_indels_into_fallback():
    """Reads whose ungapped accept hides an indel near the read end as
    mismatches: with the gate on the same chain falls back with IndelSuspect,
    metrics unchanged, exactly when the probe finds a cheaper gapped
    alignment; other accepts are untouched."""
    rng = random.Random(77)
    flips = kept = 0
    for k in range(400):
        ref = bytes(rng.choice(b"ACGT") for _ in range(400))
        st, L = rng.randrange(100, 200), 150
        if k % 4:
            read, ref = _end_indel_read(rng, ref, st, L)
        else:
            read = bytearray(ref[st:st + L])
            read[rng.randrange(L)] = ord("T")
            read = bytes(read)
        span = (st + 20, st + 80, 20, 80)
        c = O.cfg(1, 4, 6, 1, 32, 100, 5)
        kw = dict(accept_max_mismatches=5, accept_min_identity=96.0)
        off = O.prefilter_chain(read, ref, span, c, **kw)
        on = O.prefilter_chain(read, ref, span, c, ungap_gap_check=True, **kw)
        if off["result"] != "accept":
            assert on == off
            continue
        mism = off["metrics"]["mism"]
        if mism > 0 and O.ungapped_beaten_by_gap(read, ref, span, mism, c):
            assert (on["result"], on["reason"]) == ("fallback", "indel_suspect")
            assert on["metrics"] == off["metrics"] and "alignment" not in on
            flips += 1
        else:
            assert on == off
            kept += 1
    assert flips > 150 and kept > 50
