"""Alignment records (build_alignment + NM/MD) of libkira_hip's host code
against the oracle and the SAM-spec MD construction (no GPU needed)."""
import random

from oracle import oracle as O
from tests.kira_pkg import load
from tests.refgen import gen_ref
from tests.test_oracle_prefilter import canonical_md

ABORT_NONE = -(2 ** 31) // 8


def mutate(rng, s, sub, indel):
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(rng.choice(b"ACGTNR"))
        out.append(rng.choice(b"ACGTacgtN".replace(bytes([b]), b"")) if rng.random() < sub else b)
    return bytes(out)


def test_records_match_oracle_build_alignment():
    pkg = load()
    rng = random.Random(9)
    jobs, sws, want = [], [], []
    for k in range(300):
        L = rng.randrange(1, 200)
        flank = rng.randrange(0, 80)
        g = bytes(rng.choice(b"ACGTacgtNRY") for _ in range(L + 2 * flank + 10))
        read = mutate(rng, g[flank:flank + L].upper(), 0.05, 0.04) or b"A"
        c = O.cfg(1, 4, 6, 1, rng.choice([4, 16, 32]), 100, rng.choice([0, 5]))
        sw = O.banded_sw(read, g, flank, c, ABORT_NONE, O.knobs())
        ws = rng.randrange(0, 10 ** 6)
        jobs.append((read, g, ws))
        sws.append(sw)
        want.append(O.build_alignment(read, g, ws, sw))
    got = pkg.build_alignments(jobs, sws)
    assert got == want


def test_known_answers():
    pkg = load()
    # gapless 8M alignment, NM 0 (tests/alignment.rs:4-40)
    ref = gen_ref(60)
    sw = O.banded_sw(ref[10:18], ref, 10, O.cfg(1, 4, 6, 1, 8, 100, 5), ABORT_NONE, O.knobs())
    r, = pkg.build_alignments([(ref[10:18], ref, 1000)], [sw])
    assert (r["cigar"], r["nm"], r["md"], r["ref_start"]) == ("8M", 0, "8", 1010)
    # certified sieve hit: {L}M at the shift (mod.rs:636-686); MD as SAM spec
    read = b"ACGT" * 37 + b"AC"
    window = bytearray(b"TT" + read + b"GG")
    window[2 + 10] = ord("T")
    hit = dict(ref_start=2, ref_end=152, read_start=0, read_end=150, score=144,
               early_abort=False, cigar=[(150 << 4) | 0])
    r, = pkg.build_alignments([(read, bytes(window), 500)], [hit])
    md, nm = canonical_md(read, bytes(window[2:152]))
    assert (r["md"], r["nm"], r["cigar"], r["ref_start"], r["ref_end"]) == (md, nm, "150M", 502,
                                                                              652)
    # deletion and insertion tags (md.rs:206-295 style)
    read = b"ACGTACGTAC" + b"GGTTAACC"
    window = b"ACGTACGTAC" + b"TTT" + b"GGTTAACC"
    del_sw = dict(ref_start=0, ref_end=21, read_start=0, read_end=18, score=0, early_abort=False,
                  cigar=[(10 << 4) | 0, (3 << 4) | 2, (8 << 4) | 0])
    r, = pkg.build_alignments([(read, window, 0)], [del_sw])
    assert (r["md"], r["nm"], r["cigar"]) == ("10^TTT8", 3, "10M3D8M")
    ins_sw = dict(ref_start=0, ref_end=10, read_start=0, read_end=13, score=0, early_abort=False,
                  cigar=[(5 << 4) | 0, (3 << 4) | 1, (5 << 4) | 0])
    r, = pkg.build_alignments([(b"ACGTAGGGCGTAC", b"ACGTACGTAC", 0)], [ins_sw])
    assert (r["md"], r["nm"], r["cigar"]) == ("10", 3, "5M3I5M")


def test_records_on_long_clean_runs_with_case_and_n():
    """Long reads that mostly copy an upper-case A/C/G/T window (so whole
    8-byte groups match), with lower-case stretches, N / IUPAC bytes on either
    side and the occasional substitution or indel: the records equal the
    oracle's build_alignment."""
    pkg = load()
    rng = random.Random(21)
    jobs, sws, want = [], [], []
    for k in range(200):
        L = rng.randrange(40, 300)
        flank = rng.randrange(0, 60)
        g = bytearray(rng.choice(b"ACGT") for _ in range(L + 2 * flank + 10))
        for _ in range(rng.randrange(0, 4)):  # lower-case and ambiguous stretches
            a = rng.randrange(len(g))
            for t in range(a, min(len(g), a + rng.randrange(1, 12))):
                g[t] = rng.choice([g[t] | 0x20, ord("N"), ord("R"), g[t]])
        g = bytes(g)
        read = bytearray(g[flank:flank + L].upper())
        for _ in range(rng.randrange(0, 5)):
            i = rng.randrange(len(read))
            read[i] = rng.choice(b"ACGTNY")
        if rng.random() < 0.4:
            i = rng.randrange(1, len(read) - 1)
            del read[i:i + rng.randrange(1, 4)]
        read = bytes(read)
        c = O.cfg(1, 4, 6, 1, rng.choice([16, 32]), 100, rng.choice([0, 5]))
        sw = O.banded_sw(read, g, flank, c, ABORT_NONE, O.knobs())
        ws = rng.randrange(0, 10 ** 6)
        jobs.append((read, g, ws))
        sws.append(sw)
        want.append(O.build_alignment(read, g, ws, sw))
    got = pkg.build_alignments(jobs, sws)
    assert got == want
