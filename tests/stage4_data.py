"""Synthetic stage-4 input: contigs with repeat families, lower-case and N
runs; single-end reads drawn from them on both strands with substitutions,
indels, N bases and chimeric joins; and per-read chains as stage 3 would hand
them over (best first: the true anchor plus repeat-copy and random decoys)."""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)
COMP = np.zeros(256, np.uint8) + ord("N")
for a, b in zip(b"ACGTacgt", b"TGCATGCA"):
    COMP[a] = b


def revcomp(s: bytes) -> bytes:
    return COMP[np.frombuffer(s, np.uint8)[::-1]].tobytes()


def make_contigs(rng, n_contigs, length, repeat_frac=0.2, lower_frac=0.002, n_runs=4):
    contigs = []
    fam = [ACGT[rng.integers(0, 4, rng.integers(300, 1500))] for _ in range(6)]
    for _ in range(n_contigs):
        g = ACGT[rng.integers(0, 4, length)].copy()
        covered = 0
        while covered < repeat_frac * length:
            f = fam[rng.integers(0, len(fam))].copy()
            sub = rng.random(len(f)) < rng.uniform(0.002, 0.03)
            f[sub] = ACGT[rng.integers(0, 4, int(sub.sum()))]
            p = int(rng.integers(0, length - len(f)))
            g[p:p + len(f)] = f
            covered += len(f)
        low = rng.random(length) < lower_frac
        g[low] = g[low] + 32
        for _ in range(n_runs):
            p = int(rng.integers(0, length - 40))
            g[p:p + int(rng.integers(1, 30))] = ord("N")
        contigs.append(g.tobytes())
    return contigs


def _mutate(rng, s: np.ndarray, n_sub, n_indel, n_n):
    s = s.copy()
    L = len(s)
    for _ in range(n_sub):
        p = int(rng.integers(0, L))
        s[p] = ACGT[(int(np.searchsorted(ACGT, s[p] & 0xDF)) + 1 + int(rng.integers(0, 3))) % 4]
    out = s
    for _ in range(n_indel):
        p = int(rng.integers(10, max(11, len(out) - 10)))
        k = int(rng.integers(1, 4))
        if rng.random() < 0.5:
            out = np.concatenate([out[:p], out[p + k:]])
        else:
            out = np.concatenate([out[:p], ACGT[rng.integers(0, 4, k)], out[p:]])
    for _ in range(n_n):
        out[int(rng.integers(0, len(out)))] = ord("N")
    return out


def make_batch(seed=1, n_reads=2000, n_contigs=3, contig_len=200_000,
               lengths=(100, 120, 150, 150, 150, 180, 250)):
    """Returns dict(genome, contig_off, contig_len, contigs, reads, chains,
    repeat_min_occ)."""
    rng = np.random.default_rng(seed)
    contigs = make_contigs(rng, n_contigs, contig_len)
    cv = [np.frombuffer(c, np.uint8) for c in contigs]
    reads, chains, rmo = [], [], []
    for _ in range(n_reads):
        L = int(lengths[int(rng.integers(0, len(lengths)))])
        cid = int(rng.integers(0, n_contigs))
        pos = int(rng.integers(200, contig_len - L - 200))
        src = cv[cid][pos:pos + L]
        kind = rng.random()
        chimera = kind > 0.93
        if chimera:
            # read[0:cut] from here, read[cut:] from elsewhere
            cut = int(rng.integers(L // 3, 2 * L // 3))
            cid2 = int(rng.integers(0, n_contigs))
            pos2 = int(rng.integers(200, contig_len - L - 200))
            seq = np.concatenate([src[:cut], cv[cid2][pos2:pos2 + L - cut]])
            seq = _mutate(rng, seq, int(rng.integers(0, 3)), 0, 0)
        elif kind < 0.25:
            seq = src.copy()
        elif kind < 0.60:
            seq = _mutate(rng, src, int(rng.integers(1, 4)), 0, 0)
        elif kind < 0.75:
            seq = _mutate(rng, src, int(rng.integers(4, 12)), 0, 0)
        elif kind < 0.88:
            seq = _mutate(rng, src, int(rng.integers(0, 3)), int(rng.integers(1, 3)), 0)
        else:
            seq = _mutate(rng, src, int(rng.integers(0, 2)), 0, int(rng.integers(1, 3)))
        seq = bytes(seq & 0xDF)  # reads arrive upper-cased (src/io/mod.rs:740-776)
        Lr = len(seq)
        rev = rng.random() < 0.5
        read = revcomp(seq) if rev else seq
        strand = 1 if rev else 0
        ch = []
        if chimera:
            ch.append((cid, pos, pos + cut, 0, cut, strand, cut))
            ch.append((cid2, pos2, pos2 + Lr - cut, cut, Lr, strand, Lr - cut))
        else:
            a = rng.random()
            if a < 0.6:  # anchor over the whole read
                ch.append((cid, pos, pos + Lr, 0, Lr, strand, Lr))
            else:  # partial anchor
                s0 = int(rng.integers(0, Lr // 3))
                e0 = int(rng.integers(s0 + Lr // 3, Lr + 1))
                ch.append((cid, pos + s0, pos + e0, s0, e0, strand, e0 - s0))
        # decoys, some close in score (ambiguous reads)
        for _ in range(int(rng.integers(0, 3))):
            dc = int(rng.integers(0, n_contigs))
            dp = int(rng.integers(0, contig_len - Lr))
            dl = int(rng.integers(Lr // 4, Lr))
            ds = int(rng.integers(0, Lr - dl + 1))
            score = int(ch[0][6] - rng.integers(0, max(2, ch[0][6] // 2)))
            ch.append((dc, dp, min(dp + dl, contig_len), ds, ds + dl, int(rng.integers(0, 2)),
                       max(score, 1)))
        if rng.random() < 0.02:
            ch = []
        first, rest = ch[:1], sorted(ch[1:], key=lambda c: -c[6])
        chains.append(first + rest)
        reads.append(read)
        rmo.append(int(rng.integers(1, 3)) if rng.random() < 0.1 else 0)
    off = np.zeros(n_contigs, np.int64)
    off[1:] = np.cumsum([len(c) for c in contigs[:-1]])
    return dict(genome=b"".join(contigs), contig_off=off,
                contig_len=np.array([len(c) for c in contigs], np.int64), contigs=contigs,
                reads=reads, chains=chains, repeat_min_occ=rmo)


class Lcg:
    """The seeded 64-bit LCG of the reference's integration tests
    (tests/common/mod.rs:27-50, Knuth's MMIX constants)."""

    M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.M

    def next_u64(self) -> int:
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & self.M
        return self.s

    def below(self, n: int) -> int:
        return (self.next_u64() >> 33) % n

    def base(self) -> int:
        return b"ACGT"[(self.next_u64() >> 33) & 3]


def synth_contig(rng: Lcg, length: int) -> bytes:
    """synth_contig — tests/common/mod.rs:52-54."""
    return bytes(rng.base() for _ in range(length))


def make_c1(n_reads=10_000, contig_len=1_000_000, read_len=150, seed=0x1234_5678_9ABC_DEF0):
    """The plain `mem` configuration: one uniform-ACGT contig chr1 from the
    reference's test LCG, single-end reads from uniform loci on either strand,
    each base substituted with probability 1/100 and an indel event with
    probability 1/1000 per base (length U{1..3}, insertion or deletion 50/50),
    truth locus in the read name (`r{i}:chr1:{start}-{end}`, 0-based half-open,
    the eval format of src/eval/mod.rs:47).  Each read gets the chain stage 3
    would find for an error-free seed: its truth span, whole read, score L.
    Returns the dict make_batch returns plus `names`."""
    rng = Lcg(seed)
    contig = synth_contig(rng, contig_len)
    reads, chains, names = [], [], []
    for i in range(n_reads):
        start = rng.below(contig_len - read_len - 16)
        rev = rng.below(2) == 1
        out = bytearray()
        p = start
        while len(out) < read_len:
            if rng.below(1000) == 0:
                k = 1 + rng.below(3)
                if rng.below(2) == 0:
                    p += k  # deletion from the read
                else:
                    out += bytes(rng.base() for _ in range(k))  # insertion
                continue
            b = contig[p]
            p += 1
            if rng.below(100) == 0:
                b = b"ACGT"[(b"ACGT".index(b) + 1 + rng.below(3)) % 4]
            out.append(b)
        seq = bytes(out[:read_len])
        reads.append(revcomp(seq) if rev else seq)
        names.append(f"r{i}:chr1:{start}-{p}")
        chains.append([(0, start, p, 0, read_len, int(rev), read_len)])
    return dict(genome=contig, contig_off=np.zeros(1, np.int64),
                contig_len=np.array([contig_len], np.int64), contigs=[contig], reads=reads,
                chains=chains, repeat_min_occ=[0] * n_reads, names=names)
