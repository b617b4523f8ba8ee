"""Pins the CPU oracle's sieve functions against the reference's own known
answers and against brute force on the raw bytes."""
import random

import pytest

from oracle import oracle as O
from tests.refgen import XorShift, cuda_corpus, naive_best_second, naive_mismatches

I32MAX = 2 ** 31 - 1


# --- pack / pre-shift: tests/bitpacked_scan.rs:44-59, tests/unit/alignment_bitpacked.rs:4-57
def test_pack_roundtrip_short():
    bits, valid = O.pack(b"ACGTACGT")
    assert bits == bytes([0xE4, 0xE4]) and valid


def test_pack_handles_padding():
    bits, _ = O.pack(b"ACGTA")
    assert len(bits) == 2 and bits[1] & 0b11 == 0 and bits[1] >> 2 == 0


def test_shift_aligns_one_nucleotide():
    assert O.pre_shift(O.pack(b"ACGT")[0])[1][0] == 0x39


def test_ambiguous_bases_invalidate():
    assert not O.pack(b"ACNT")[1]
    assert not O.pack(b"acgu")[1]
    assert O.pack(b"acgt")[1]


def test_pack_closed_form():
    rng = random.Random(7)
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    for n in range(0, 70):
        s = "".join(rng.choice("ACGT") for _ in range(n))
        bits, valid = O.pack(s.encode())
        want = bytearray((n + 3) // 4)
        for i, ch in enumerate(s):
            want[i // 4] |= code[ch] << (2 * (i % 4))
        assert bits == bytes(want) and valid


def test_pre_shift_closed_form():
    """phase p holds the packed sequence starting p bases later."""
    rng = random.Random(3)
    for n in range(1, 60):
        s = "".join(rng.choice("ACGT") for _ in range(n)).encode()
        bits, _ = O.pack(s)
        ph = O.pre_shift(bits)
        for p in range(4):
            shifted, _ = O.pack(s[p:])
            assert ph[p][:len(shifted)] == shifted
            assert len(ph[p]) == len(bits)


# --- mismatch count and scan
def test_scan_exact_match_at_zero():
    r = b"ACGTACGTACGT"
    assert O.scan_best_with_second(r, r)[:2] == (0, 0)


def test_scan_exact_match_in_middle():
    assert O.scan_best_with_second(b"ACGTACGT", b"AAAAACGTACGTAAAA")[:2] == (4, 0)


def test_scan_threshold_semantics():
    sh, best, _ = O.scan_best_with_second(b"AAAAAAAA", b"TTTTTTTTTTTT")
    assert best == 8  # scan(.., 3) is None, scan(.., 8) hits with 8


def test_scan_global_best_not_first_acceptable():
    assert O.scan_best_with_second(b"ACGT", b"TTTTACGT")[:2] == (4, 0)


def test_scan_reports_tied_second():
    sh, best, second = O.scan_best_with_second(b"ACGT", b"ACGTACGT")
    assert (sh, best, second) == (0, 0, 0)


def test_scan_long_read_avx_path():
    read = bytes(b"ACGT"[i % 4] for i in range(150))
    text = bytearray(b"ACGT"[(i + 1) % 4] for i in range(200))
    text[25:175] = read
    text[32] = ord("T") if read[7] == ord("A") else ord("A")
    text[125] = ord("G") if read[100] == ord("C") else ord("C")
    sh, best, _ = O.scan_best_with_second(read, bytes(text))
    assert naive_mismatches(read, bytes(text), sh) == best and best <= 5


def test_random_corpus_vs_brute_force():
    """tests/bitpacked_scan.rs:110-149, strengthened to the exact best/second."""
    rng = XorShift(0xC0FFEEBAD)
    for _ in range(200):
        r = 16 + rng.next() % 180
        w = r + rng.next() % 100
        read = bytes(b"ACGT"[rng.next() % 4] for _ in range(r))
        text = bytearray(b"ACGT"[rng.next() % 4] for _ in range(w))
        if rng.next() % 2 == 0:
            start = rng.next() % (w - r + 1)
            text[start:start + r] = read
        got = O.scan_best_with_second(read, bytes(text))
        assert got == naive_best_second(read, bytes(text))


def test_mismatch_count_tail_masking():
    """Partial last byte: padding pairs never count (spectral.cu:78-92)."""
    rng = random.Random(11)
    for n in range(1, 40):
        a = bytes(rng.choice(b"ACGT") for _ in range(n))
        b = bytes(rng.choice(b"ACGT") for _ in range(n + 3))
        pa, _ = O.pack(a)
        pb, _ = O.pack(b)
        assert O.mismatch_count(pa, pb[:len(pa)], n) == naive_mismatches(a, b, 0)


# --- batch backend: the reference's GPU corpus, tests/cuda_spectral.rs:20-137
def test_backend_corpus_matches_brute_force():
    corpus = cuda_corpus(64)
    res = O.backend_run_batch(corpus)
    accepted = 0
    for (read, text, mm), (shift, mism, second) in zip(corpus, res):
        nshift, nbest, nsecond = naive_best_second(read, text)
        assert second == nsecond
        limit = min(mm, len(read))
        if nbest <= limit:
            accepted += 1
            assert (shift, mism) == (nshift, nbest)
            assert naive_mismatches(read, text, shift) == mism
        else:
            assert (shift, mism) == (-1, I32MAX)
    assert accepted > 10


def test_backend_sentinels():
    res = O.backend_run_batch([(b"ACGTACGT", b"ACG", 5), (b"", b"ACGT", 5),
                               (b"ACGT", b"ACGT", 0), (b"ACGT", b"TTTTT", 1)])
    assert res[0] == (-1, I32MAX, None)
    assert res[1] == (-1, I32MAX, None)
    assert res[2] == (0, 0, None)  # a single shift has no second
    assert res[3] == (-1, I32MAX, 3)  # rejected best (3 > 1) keeps its second


# --- limits, certification, tags, windows
def test_spectral_max_mismatches():
    assert O.spectral_max_mismatches(150) == 4
    assert O.spectral_max_mismatches(50) == 2
    assert O.spectral_max_mismatches(100) == 3


def _flip(b):
    return ord("A") if b != ord("A") else ord("C")


def test_certification_rules():
    rng = random.Random(5)
    read = bytes(rng.choice(b"ACGT") for _ in range(150))
    # exact: certified unless the second is within one of the best
    assert O.certified(read, read, 0, None, 4)
    assert O.certified(read, read, 0, 2, 4)
    assert not O.certified(read, read, 0, 1, 4)
    # > max mismatches
    ref = bytearray(read)
    for p in (40, 80, 120):
        ref[p] = _flip(ref[p])
    assert not O.certified(read, bytes(ref), 3, None, 4)  # > 2 mismatches
    # terminal mismatch within 8 of an end
    ref = bytearray(read)
    ref[3] = _flip(ref[3])
    assert not O.certified(read, bytes(ref), 1, None, 4)
    ref = bytearray(read)
    ref[145] = _flip(ref[145])
    assert not O.certified(read, bytes(ref), 1, None, 4)
    # two mismatches <= 4 apart
    ref = bytearray(read)
    ref[50] = _flip(ref[50])
    ref[53] = _flip(ref[53])
    assert not O.certified(read, bytes(ref), 2, None, 4)
    # count disagreeing with the hit -> reject
    assert not O.certified(read, bytes(ref), 1, None, 4)


def test_certification_indel_proxy():
    """A 1-bp insertion placed as mismatches must not certify
    (tests/algo_router.rs:116-134), a clean SNP must."""
    rng = random.Random(9)
    base = bytes(rng.choice(b"ACGT") for _ in range(200))
    snp = bytearray(base[:150])
    snp[75] = _flip(snp[75])
    assert O.certified(bytes(snp), base[:150], 1, None, 4)
    ins = base[:75] + b"G" + base[75:149]
    ham = naive_mismatches(ins, base, 0)
    # the suffix after the first mismatch re-aligns after a 1-base shift
    assert not O.certified(ins, base[:150], min(ham, 2), None, 4)


def test_ungapped_tags():
    assert O.ungapped_tags(b"ACGT", b"ACGT") == (0, "4")
    assert O.ungapped_tags(b"ACGT", b"AGGT") == (1, "1G2")
    assert O.ungapped_tags(b"TTGA", b"ACGT") == (3, "0A0C1T0")


def test_build_simd_window():
    # W = L + 4w starting 2w before the diagonal (stage4_alignment.rs:1426-1445)
    assert O.build_simd_window(10_000, 1000, 0, 150, 32) == (936, 278)
    assert O.build_simd_window(10_000, 60, 0, 150, 32) is None
    assert O.build_simd_window(1100, 1000, 0, 150, 32) is None
