"""Seeded corpora for the Myers and WFA batch tests: the random shapes of the
reference's own tests (tests/myers_bounded_edit.rs, tests/scored_wfa.rs) and
read-like pairs (a read against its reference window with substitutions and
small indels, the shapes the fast path sees)."""
import numpy as np

BASES = b"ACGT"


def _seq(rng, n, alphabet=BASES):
    return bytes(rng.choice(np.frombuffer(alphabet, np.uint8), size=n).tolist())


def mutate(rng, s: bytes, subs: int, indels: int, alphabet=BASES) -> bytes:
    b = bytearray(s)
    for _ in range(subs):
        if b:
            i = int(rng.integers(len(b)))
            b[i] = alphabet[(alphabet.index(b[i]) + 1 + int(rng.integers(3))) % 4] \
                if b[i] in alphabet else alphabet[0]
    for _ in range(indels):
        i = int(rng.integers(len(b) + 1))
        ln = int(rng.integers(1, 5))
        if rng.random() < 0.5:
            b[i:i] = _seq(rng, ln)
        else:
            del b[i:i + ln]
    return bytes(b)


def read_like_pairs(seed: int, n: int, read_len: int = 150, pad: int = 32, max_edits: int = 6):
    """(read, window) pairs: the window is the reference text of the fast
    path, the read a mutated copy of its first read_len bases."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        win = _seq(rng, read_len + pad)
        read = mutate(rng, win[:read_len], int(rng.integers(0, max_edits)),
                      int(rng.integers(0, 3)))
        if rng.random() < 0.05:  # an N in the read and one in the window
            r = bytearray(read)
            r[int(rng.integers(len(r)))] = ord("N")
            read = bytes(r)
            w = bytearray(win)
            w[int(rng.integers(len(w)))] = ord("N")
            win = bytes(w)
        out.append((read, win))
    return out


def myers_pairs(seed: int, n: int, max_len: int = 600):
    """Patterns over every block-count class (1 .. 8 blocks and longer),
    planted in their texts with a few edits, plus unrelated texts."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        m = int(rng.integers(1, max_len + 1)) if t % 4 else int(rng.integers(1, 70))
        p = _seq(rng, m)
        if rng.random() < 0.1:
            pb = bytearray(p)
            pb[int(rng.integers(m))] = ord("N")
            p = bytes(pb)
        slack = int(rng.integers(0, 40))
        text = bytearray(_seq(rng, m + slack))
        if rng.random() < 0.8:
            at = int(rng.integers(0, slack + 1))
            planted = mutate(rng, p, int(rng.integers(0, 5)), int(rng.integers(0, 2)))
            text[at:at + len(planted)] = planted
        out.append((p, bytes(text)))
    return out
