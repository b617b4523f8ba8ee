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
