"""Indel left-normalization (library host code and its CPU restatement) on
the reference's own unit tests (src/alignment/normalize.rs:253-420) and a
random property check; single-end SAM lines of the restatement on the
reference's SAM flag tests (tests/sam_flags.rs)."""
import random

from oracle import sam as S
from tests.kira_pkg import load

OPC = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4}

# (reference, read, cigar, ref_start, expected or None) — normalize.rs:262-395
KNOWN = [
    (b"ACGTAAAATGCA", b"ACGTAAAAATGCA", "7M1I5M", 0, "4M1I8M"),
    (b"TTTTGG", b"TTTGG", "3M1D2M", 0, "1D5M"),
    (b"ACGTAA", b"ACGAA", "3M1D2M", 0, None),
    (b"AAAA", b"ATA", "2M1D1M", 0, None),
    (b"AAATTT", b"AAAATT", "3M1I2M", 0, "1I5M"),
    (b"TTTCCAAAAAGGGGG", b"TTTAAAAGGGGG", "3M2D4M1D5M", 0, "3M3D9M"),
    (b"AAACACACACAA", b"AAACACACAA", "7M2D3M", 0, "2M2D8M"),
    (b"AAATTT", b"AAAATT", "1I5M", 0, None),
    (b"ACGTACGT", b"ACGTACGT", "8M", 0, None),
    (b"GGAAAATT", b"CCAAAAATT", "2S4M1I2M", 2, "2S1I6M"),
    (b"CCCCAAAATTTT", b"AAAAATTTT", "3M1I5M", 4, "1I8M"),
    (b"GAATT", b"GAACATT", "3M2I2M", 0, "2M2I3M"),
]


def words(text):
    return [(n << 4) | OPC[op] for n, op in S.parse_cigar(text)]


def text(ws):
    return "".join(f"{w >> 4}{'MIDNS'[w & 15]}" for w in ws)


def test_known_answers_oracle_and_library():
    pkg = load()
    for ref, read, cig, rs, want in KNOWN:
        got = S.left_normalize_indels(S.parse_cigar(cig), ref, rs, read)
        assert (S.cigar_text(got) if got else None) == want, (ref, read, cig)
        lib = pkg.left_normalize_indels(words(cig), ref, rs, read)
        assert (text(lib) if lib else None) == want, (ref, read, cig)


def _random_alignment(rng):
    """A random reference, a read with indels / substitutions drawn from it,
    and the CIGAR that generated it."""
    ref = bytes(rng.choice(b"ACGT" if rng.random() < 0.5 else b"AAC") for _ in range(60))
    ops, read, r = [], bytearray(), 5
    while r < 50:
        k = rng.randrange(1, 6)
        kind = rng.choice("MMMID")
        if kind == "M":
            seg = bytearray(ref[r:r + k])
            if rng.random() < 0.2 and seg:
                seg[rng.randrange(len(seg))] = rng.choice(b"ACGT")
            read += seg
            r += len(seg)
            k = len(seg)
        elif kind == "I":
            read += bytes(rng.choice(b"ACGT") for _ in range(k))
        else:
            r += k
        if k:
            if ops and ops[-1][1] == kind:
                ops[-1] = (ops[-1][0] + k, kind)
            else:
                ops.append((k, kind))
    return ref, bytes(read), ops


def _columns(ref, rs, read, ops):
    """(aligned column mismatches, query bases consumed)."""
    q, r, mm = 0, rs, 0
    for n, op in ops:
        if op == "M":
            mm += sum(1 for t in range(n) if read[q + t] != ref[r + t])
            q += n
            r += n
        elif op in "IS":
            q += n
        else:
            r += n
    return mm, q


def test_random_normalization_is_exact_and_library_matches():
    pkg = load()
    rng = random.Random(17)
    moved = 0
    for _ in range(2000):
        ref, read, ops = _random_alignment(rng)
        new = S.left_normalize_indels(ops, ref, 5, read)
        lib = pkg.left_normalize_indels(words(S.cigar_text(ops)), ref, 5, read)
        assert (text(lib) if lib else None) == (S.cigar_text(new) if new else None)
        if new is None:
            continue
        moved += 1
        # exact transformation: same mismatching columns, same query length
        assert _columns(ref, 5, read, new) == _columns(ref, 5, read, ops)
        # idempotent: a normalized CIGAR is canonical
        assert S.left_normalize_indels(new, ref, 5, read) is None
    assert moved > 100


def _aln(ref_id, start, end, is_rev, **kw):
    a = dict(kind="dp_aligned", ref_id=ref_id, ref_start=start, ref_end=end, read_start=0,
             read_end=end - start, cigar=f"{end - start}M", score=100, mapq=60, is_rev=is_rev,
             is_supplementary=False, nm=0, md=str(end - start), xs_score=None)
    a.update(kw)
    return a


def test_sam_flag_known_answers():
    # tests/sam_flags.rs: unpaired_mapped_emits_no_paired_bits
    line = S.append_alignment("test_read", b"A" * 150, b"I" * 150, _aln(0, 100, 250, False),
                              ["chr1", "chr2"])
    cols = line.rstrip("\n").split("\t")
    assert cols[:9] == ["test_read", "0", "chr1", "101", "60", "150M", "*", "0", "0"]
    assert cols[11:] == ["NM:i:0", "MD:Z:150", "AS:i:100"]
    # reverse strand: 0x10, SEQ reverse-complemented, QUAL reversed
    line = S.append_alignment("r", b"AACG", b"1234", _aln(1, 0, 4, True, md="4"), ["c1", "c2"])
    cols = line.rstrip("\n").split("\t")
    assert (cols[1], cols[2], cols[9], cols[10]) == ("16", "c2", "CGTT", "4321")
    # supplementary: 0x800 and hard clips (bwa-mem default), -M and -Y variants
    sup = _aln(0, 10, 14, False, cigar="2S4M", read_start=2, is_supplementary=True, md="4")
    cols = S.append_alignment("r", b"GGACGT", None, sup, ["c"]).split("\t")
    assert (cols[1], cols[5], cols[9], cols[10]) == ("2048", "2H4M", "ACGT", "*")
    cols = S.append_alignment("r", b"GGACGT", None, sup, ["c"],
                              S.default_cfg(split_as_secondary=True,
                                            soft_clip_supplementary=True)).split("\t")
    assert (cols[1], cols[5], cols[9]) == ("256", "2S4M", "GGACGT")
    # XS and RG tags
    line = S.append_alignment("r", b"ACGT", None, _aln(0, 0, 4, False, md="4", xs_score=7),
                              ["c"], S.default_cfg(read_group="@RG\tID:grp1\tSM:x"))
    assert line.endswith("\tXS:i:7\tRG:Z:grp1\n")
    # unmapped single-end record
    assert S.append_unmapped("u", b"ACGT", None) == "u\t4\t*\t0\t0\t*\t*\t0\t0\tACGT\t*\n"
