"""Deterministic generators used by the reference's own tests, restated so the
same inputs (and therefore the same known answers) can be produced here."""
from __future__ import annotations

M64 = (1 << 64) - 1


class XorShift:
    """xorshift64 as in tests/cuda_spectral.rs:13-18 and tests/bitpacked_scan.rs:36-41."""

    def __init__(self, seed: int):
        self.s = seed & M64

    def next(self) -> int:
        s = self.s
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        self.s = s
        return s


def cuda_corpus(n: int):
    """make_corpus — tests/cuda_spectral.rs:20-47."""
    bases = b"ACGT"
    rng = XorShift(0xCAFE_F00D)
    out = []
    for _ in range(n):
        r = 16 + rng.next() % 200
        w = r + rng.next() % 80
        read = bytes(bases[rng.next() % 4] for _ in range(r))
        text = bytearray(bases[rng.next() % 4] for _ in range(w))
        if rng.next() % 2 == 0 and w >= r:
            start = rng.next() % (w - r + 1)
            text[start:start + r] = read
            n_mism = rng.next() % 4
            for _ in range(n_mism):
                pos = start + rng.next() % r
                text[pos] = ord("T") if text[pos] == ord("A") else ord("A")
        max_mism = max(r * 8 // 100, 5)
        out.append((read, bytes(text), max_mism))
    return out


def gen_ref(length: int) -> bytes:
    """gen_ref — src/alignment/mod.rs:2344-2355 (xorshift, seed 0x1234_5678_9abc_def0)."""
    s = 0x1234_5678_9ABC_DEF0
    out = bytearray()
    for _ in range(length):
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        out.append(b"ACGT"[s & 3])
    return bytes(out)


def lcg_dna(length: int, state: int) -> bytes:
    """rand_dna — tests/unit/alignment_sw_int8_vnni.rs:28-38."""
    out = bytearray()
    for _ in range(length):
        state = (state * 6364136223846793005 + 1442695040888963407) & M64
        out.append(b"ACGT"[(state >> 60) & 3])
    return bytes(out)


def naive_mismatches(read: bytes, text: bytes, shift: int) -> int:
    """tests/bitpacked_scan.rs:13-19."""
    return sum(1 for a, b in zip(read, text[shift:shift + len(read)]) if a != b)


def naive_best_second(read: bytes, text: bytes):
    """Brute-force best shift (lowest on ties) and the second order statistic
    of the per-shift Hamming distances (None with a single shift)."""
    counts = [naive_mismatches(read, text, t) for t in range(len(text) - len(read) + 1)]
    best = min(counts)
    shift = counts.index(best)
    rest = sorted(counts)
    second = rest[1] if len(rest) > 1 else None
    return shift, best, second
