"""GPU parity of the sieve: libkira_hip.so through the C ABI vs the CPU oracle."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.refgen import cuda_corpus, naive_best_second, naive_mismatches

pytestmark = pytest.mark.gpu
I32MAX = 2 ** 31 - 1


@pytest.fixture(scope="module")
def pkg():
    p = load()
    assert p.is_available(), "no GPU visible"
    return p


@pytest.fixture(scope="module")
def backend(pkg):
    return pkg.HipBackend(0)


def _jobs(pkg, corpus):
    jobs = []
    for read, text, mm in corpus:
        rb, _ = O.pack(read)
        tb, _ = O.pack(text)
        jobs.append(pkg.HipJob(rb, len(read), O.pre_shift(tb), len(text), mm))
    return jobs


def _triples(res):
    return [(r.shift, r.mismatches, r.second_mismatches) for r in res]


def test_reference_gpu_corpus(pkg, backend):
    """tests/cuda_spectral.rs:49-137 on the same 64-job corpus."""
    corpus = cuda_corpus(64)
    got = backend.run_batch(_jobs(pkg, corpus))
    assert _triples(got) == O.backend_run_batch(corpus)
    for i, ((read, text, mm), g) in enumerate(zip(corpus, got)):
        nshift, nbest, nsecond = naive_best_second(read, text)
        assert g.second_mismatches == nsecond, i
        if g.shift >= 0:
            assert g.mismatches <= mm and (g.shift, g.mismatches) == (nshift, nbest), i
            assert naive_mismatches(read, text, g.shift) == g.mismatches, i
        else:
            assert nbest > min(mm, len(read)), i


def test_ragged_edge_cases(pkg, backend):
    rng = random.Random(1234)
    corpus = []
    for L in [0, 1, 2, 3, 4, 5, 15, 16, 17, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 150,
              160, 191, 192, 200, 224, 255, 256]:
        for extra in [-1, 0, 1, 3, 31, 32, 33, 127, 128, 129, 300]:
            W = L + extra
            if W < 0:
                continue
            read = bytes(rng.choice(b"ACGT") for _ in range(L))
            text = bytearray(rng.choice(b"ACGT") for _ in range(W))
            if W >= L and rng.random() < 0.6:
                st = rng.randrange(W - L + 1)
                text[st:st + L] = read
                for _ in range(rng.randrange(4)):
                    if L:
                        p = st + rng.randrange(L)
                        text[p] = ord("C") if text[p] != ord("C") else ord("G")
            corpus.append((read, bytes(text), rng.choice([0, 1, 2, 5, 8, 300, 2 ** 32 - 1])))
    got = backend.run_batch(_jobs(pkg, corpus))
    assert _triples(got) == O.backend_run_batch(corpus)


def test_empty_batch_and_errors(pkg, backend):
    assert backend.run_batch([]) == []
    long_read = pkg.HipJob(b"\x00" * 65, 260, [b"\x00" * 70] * 4, 280, 5)
    with pytest.raises(pkg.HipError) as e:
        backend.run_batch([long_read])
    assert e.value.kind == "ReadTooLong"
    bad = pkg.HipJob(b"\x00", 4, [b"\x00" * 3, b"\x00" * 3, b"\x00" * 2, b"\x00" * 3], 12, 5)
    with pytest.raises(pkg.HipError) as e:
        backend.run_batch([bad])
    assert e.value.kind == "Runtime"
    # the backend is still usable after a failed batch
    corpus = cuda_corpus(8)
    assert _triples(backend.run_batch(_jobs(pkg, corpus))) == O.backend_run_batch(corpus)


def test_dispatcher_roundtrip(pkg):
    d = pkg.dispatcher
    assert not d.is_active()
    assert d.dispatch([]) == []
    d.start()
    try:
        assert d.is_active()
        corpus = cuda_corpus(16)
        res = d.dispatch(_jobs(pkg, corpus))
        assert _triples(res) == O.backend_run_batch(corpus)
        # a failing batch maps to None, like dispatcher.rs:69-76
        assert d.dispatch([pkg.HipJob(b"\x00" * 65, 260, [b"\x00" * 70] * 4, 280, 5)]) is None
    finally:
        d.stop()
    assert not d.is_active() and d.dispatch(_jobs(pkg, cuda_corpus(2))) is None


def _genome_case(seed, G=20000, n_reads=300, per_read=8, with_noise=True):
    rng = np.random.default_rng(seed)
    g = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=G)
    if with_noise:
        # sprinkle N's and a soft-masked (lower case) stretch
        g[rng.integers(0, G, 25)] = ord("N")
        g[5000:5400] = np.char.lower(g[5000:5400].view("S1")).view(np.uint8)
    reads, lens, wstart, wlen, jread = [], [], [], [], []
    for r in range(n_reads):
        L = int(rng.integers(20, 257)) if r % 3 else 150
        pos = int(rng.integers(0, G - L))
        read = g[pos:pos + L].copy()
        read = np.frombuffer(bytes(read).upper(), np.uint8).copy()
        for p in rng.integers(0, L, int(rng.integers(0, 4))):
            read[p] = ord("ACGT"[(("ACGT".find(chr(read[p])) + 1) % 4)])
        if r % 29 == 0:
            read[int(rng.integers(0, L))] = ord("N")
        reads.append(bytes(read))
        lens.append(L)
        w = 32
        for k in range(per_read):
            W = L + 4 * w if k else L + 4 * w
            s = pos - 2 * w if k == 0 else int(rng.integers(-10, G - W + 10))
            if k == per_read - 1:
                W = int(rng.integers(0, L + 40))
            wstart.append(s)
            wlen.append(W)
            jread.append(r)
    return bytes(g), reads, lens, wstart, wlen, jread


def _oracle_genome(gseq, reads, lens, wstart, wlen, jread, limit, cert_max, strict):
    out = []
    for s, W, r in zip(wstart, wlen, jread):
        read = reads[r]
        if s < 0 or s + W > len(gseq):
            out.append((-1, I32MAX, I32MAX, 0))
            continue
        text = gseq[s:s + W]
        if not (O.pack(read)[1] and O.pack(text)[1]):
            out.append((-1, I32MAX, I32MAX, 0))
            continue
        (sh, m, sec), = O.backend_run_batch([(read, text, limit)])
        flags = 1
        if sh >= 0:
            flags |= 2
            if cert_max >= 0 and O.certified(read, text[sh:sh + len(read)], m, sec, cert_max,
                                             strict):
                flags |= 4
        out.append((sh, m, I32MAX if sec is None else sec, flags))
    return out


@pytest.mark.parametrize("strict,stride", [(False, 264), (True, 264), (False, 288)])
def test_genome_batch_with_certification(pkg, strict, stride):
    # stride 288 (a multiple of 32) takes the 16-byte-load packing kernel
    import torch
    gseq, reads, lens, wstart, wlen, jread = _genome_case(7 if strict else 8)
    ctx = pkg.Context(0)
    genome = pkg.Genome(ctx, gseq)
    rd = np.zeros((len(reads), stride), np.uint8)
    for i, r in enumerate(reads):
        rd[i, :len(r)] = np.frombuffer(r, np.uint8)
    dev = torch.device("cuda:0")
    t = lambda a, dt: torch.tensor(np.asarray(a), dtype=dt, device=dev)
    out = pkg.sieve_genome(ctx, genome, t(rd, torch.uint8), t(lens, torch.int32),
                           t(wstart, torch.int64), t(wlen, torch.int32),
                           job_read=t(jread, torch.int32), limit_all=8, cert_max_mism=4,
                           cert_strict=strict)
    torch.cuda.synchronize()
    got = list(zip(out.shift.tolist(), out.best.tolist(), out.second.tolist(),
                   out.flags.tolist()))
    want = _oracle_genome(gseq, reads, lens, wstart, wlen, jread, 8, 4, strict)
    assert got == want
    assert sum(1 for w in want if w[3] & 4) > 20  # certification was exercised
    genome.close()
    ctx.close()


def _expected_planes(read: bytes, L: int):
    """Bit planes of a read from the oracle's 2-bit packing (bitpacked.rs:12-20):
    code bit 0 / bit 1 of base q at bit q of the low / high plane; the lower-case
    plane marks a/c/g/t bytes."""
    packed, valid = O.pack(read[:L])
    lo = [0] * 8
    hi = [0] * 8
    lc = [0] * 8
    for q in range(L):
        code = (packed[q // 4] >> (2 * (q % 4))) & 3
        lo[q // 32] |= (code & 1) << (q % 32)
        hi[q // 32] |= (code >> 1) << (q % 32)
        if read[q] in b"acgt":
            lc[q // 32] |= 1 << (q % 32)
    return lo + hi + lc, valid


@pytest.mark.parametrize("stride,odd", [(264, False), (288, False), (288, True), (160, True),
                                        (160, False)])
def test_read_packing_every_byte_value(pkg, stride, odd):
    """All 256 byte values (upper and lower case A/C/G/T, 'n', bytes >= 0x80) through
    both device packers (strides 288 and 160 take the four-bytes-at-once kernel, which
    packs two reads a thread; 160 leaves word slots past the row, an odd count a thread
    with one read), plus reads longer than the row, which must come back invalid."""
    import torch
    rng = np.random.default_rng(stride + odd)
    alphabet = np.frombuffer(b"ACGTacgtnN", np.uint8)
    reads, lens = [], []
    for k in range(256 + odd):
        L = int(rng.integers(1, min(257, stride + 1)))
        row = rng.choice(alphabet, size=L) if k % 2 else \
            rng.choice(np.frombuffer(b"ACGTacgt", np.uint8), size=L)
        row = row.copy()
        row[k % L] = k % 256  # byte value k somewhere in read k
        reads.append(bytes(row))
        lens.append(L)
    for L in (stride + 1, stride + 40):  # longer than the row: clamped, invalid
        reads.append(bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=stride)))
        lens.append(L)
    n = len(reads)
    rd = np.zeros((n, stride), np.uint8)
    for i, r in enumerate(reads):
        rd[i, :len(r)] = np.frombuffer(r, np.uint8)
    ctx = pkg.Context(0)
    dev = torch.device("cuda:0")
    d_reads = torch.tensor(rd, device=dev)
    d_lens = torch.tensor(np.asarray(lens, np.int32), device=dev)
    d_planes = torch.zeros((n, 24), dtype=torch.int32, device=dev)
    d_valid = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    pkg.pack_reads_dev(ctx, n, d_reads.data_ptr(), stride, d_lens.data_ptr(),
                       d_planes.data_ptr(), d_valid.data_ptr())
    torch.cuda.synchronize()
    planes = d_planes.cpu().numpy().view(np.uint32)
    valid = d_valid.cpu().numpy()
    for i in range(256 + odd):
        want, ok = _expected_planes(reads[i], lens[i])
        assert planes[i].tolist() == want, i
        assert valid[i] == int(ok), i
    assert valid[256 + odd:].tolist() == [0, 0]
    ctx.close()
