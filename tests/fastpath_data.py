"""A seeded fast-path batch: a small multi-contig genome (with an N run) and
chains whose reads cover every branch of the cascade: exact and
substitution-only reads (the packed certificate), indel reads (WFA), reads
with a 5' deletion before the seed (the origin retry), reads with a foreign
tail (the partial-read check), unrelated reads (Myers reject), reads with N
or lower-case bases, windows at the contig ends, and reads over the WFA
length limit."""
import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)


def _rand(rng, n):
    return bytes(rng.choice(ACGT, size=n).tolist())


def _sub(rng, s: bytes, k: int) -> bytes:
    b = bytearray(s)
    for _ in range(k):
        i = int(rng.integers(len(b)))
        b[i] = b"ACGT"[(b"ACGT".index(b[i]) + 1 + int(rng.integers(3))) % 4] \
            if b[i] in b"ACGT" else ord("A")
    return bytes(b)


def make_fast_batch(seed: int, n: int, lens=(20000, 15000, 5000), mix=None):
    """mix: optional weights over the 12 read kinds (default: round robin)."""
    rng = np.random.default_rng(seed)
    lens = list(lens)
    contigs = [bytearray(_rand(rng, L)) for L in lens]
    contigs[0][7000:7040] = b"N" * 40
    contigs = [bytes(c) for c in contigs]
    genome = b"".join(contigs)
    contig_off = np.cumsum([0] + lens[:-1]).astype(np.int64)
    contig_len = np.array(lens, dtype=np.int64)
    reads, chains = [], []
    kinds = (np.arange(n) % 12 if mix is None
             else rng.choice(12, size=n, p=np.asarray(mix, float) / np.sum(mix)))
    for t in range(n):
        kind = int(kinds[t])
        L = int(rng.choice([150, 150, 150, 100, 250, 76]))
        cid = int(rng.integers(len(lens)))
        clen = lens[cid]
        pos = int(rng.integers(40, clen - L - 60))
        ref = contigs[cid]
        read_start = int(rng.integers(0, 30))
        seed_shift = 0
        if kind == 0:
            read = ref[pos:pos + L]
        elif kind == 1:
            read = _sub(rng, ref[pos:pos + L], int(rng.integers(1, 4)))
        elif kind in (2, 3):
            i = int(rng.integers(10, L - 10))
            k = int(rng.integers(1, 6))
            if kind == 2:
                read = ref[pos:pos + i] + _rand(rng, k) + ref[pos + i:pos + L - k]
            else:
                read = ref[pos:pos + i] + ref[pos + i + k:pos + L + k]
            read = _sub(rng, read, int(rng.integers(0, 3)))
        elif kind == 4:
            # 5' deletion after a few bases: the seed implies a start shifted by k
            k = int(rng.integers(2, 8))
            head = int(rng.integers(3, 14))
            read = ref[pos:pos + head] + ref[pos + head + k:pos + L + k]
            seed_shift = k
        elif kind == 5:
            read = ref[pos:pos + L - 30] + _rand(rng, 30)
        elif kind == 6:
            read = _rand(rng, L)
        elif kind == 7:
            r = bytearray(_sub(rng, ref[pos:pos + L], 1))
            r[int(rng.integers(L))] = ord("N")
            read = bytes(r)
        elif kind == 8:
            r = bytearray(ref[pos:pos + L])
            for i in rng.integers(0, L, size=3):
                r[int(i)] = r[int(i)] | 0x20
            read = bytes(r)
        elif kind == 9:
            # window at a contig end, or past it
            pos = clen - L - int(rng.integers(-5, 20))
            pos = max(0, min(pos, clen - 1))
            read = (ref[pos:pos + L] + _rand(rng, L))[:L]
        elif kind == 10:
            L = 320
            pos = int(rng.integers(40, clen - L - 60))
            read = ref[pos:pos + L]
        else:
            pos = 6950 if cid == 0 else pos  # across the N run
            read = _sub(rng, ref[pos:pos + L], 1)
        ref_start = pos + read_start + seed_shift
        if kind == 9 and rng.random() < 0.3:
            read_start = ref_start + 5  # negative window start
        score = int(rng.integers(60, L + 1))
        reads.append(read)
        chains.append((cid, ref_start, ref_start + 20, read_start, read_start + 20,
                       int(rng.integers(2)), score))
    return dict(genome=genome, contigs=contigs, contig_off=contig_off, contig_len=contig_len,
                reads=reads, chains=chains)
