"""GPU parity of kira_myers_batch and kira_wfa_batch with the oracle's Myers
and scored-WFA restatements (oracle/kira_wfa_oracle.c, pinned by
tests/test_oracle_wfa.py): distances, end positions, losses, text spans and
CIGARs, bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.kira_pkg import load
from tests.test_oracle_wfa import xorshift
from tests.wfa_data import myers_pairs, read_like_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    p = load()
    assert p.is_available(), "no GPU visible"
    return p


@pytest.fixture(scope="module")
def ctx(pkg):
    c = pkg.Context(0)
    yield c
    c.close()


def test_myers_known_answers_and_edges(pkg, ctx):
    pairs = [(b"ACGTACGTACGT", b"NNACGTACGTACGTNN", 5), (b"ACGTACGTACGT", b"NNACGTAXGTACGTNN", 3),
             (b"ACGTACGT", b"NNACGTAAACGTNN", 3), (b"ACGTACGT", b"NNACGACGTNN", 3),
             (b"AAAAAAAA", b"TTTTTTTT", 3), (b"AAAAAAAA", b"TTTTTTTT", 8), (b"", b"ACGT", 0),
             (b"ACGT", b"", 4), (b"ACGT", b"", 3), (b"ACNT", b"ACNT", 3), (b"acgt", b"TTACGT", 0)]
    print(f"[myers] {len(pairs)} known-answer pairs", flush=True)
    got = pkg.myers_batch(ctx, pairs)
    want = [O.myers_bounded(p, t, k) for p, t, k in pairs]
    assert got == want


@pytest.mark.parametrize("seed", [1, 2])
def test_myers_random_corpus_every_block_class(pkg, ctx, seed):
    base = myers_pairs(seed, 700)
    pairs = []
    for p, t in base:
        truth = O.myers_bounded(p, t, len(p))
        d = truth[0] if truth else len(p)
        for k in {max(d - 1, 0), d, d + 3}:
            pairs.append((p, t, k))
    print(f"[myers] seed {seed}: {len(pairs)} pairs, max pattern {max(len(p) for p, _, _ in pairs)}",
          flush=True)
    got = pkg.myers_batch(ctx, pairs)
    want = [O.myers_bounded(p, t, k) for p, t, k in pairs]
    bad = [i for i in range(len(pairs)) if got[i] != want[i]]
    assert not bad, (bad[:5], [(len(pairs[i][0]), got[i], want[i]) for i in bad[:5]])


def _check_wfa(pkg, ctx, pairs, a, b, o, e, drop=0, label=""):
    gcfg = pkg.sw_cfg(a, b, o, e)
    ocfg = O.cfg(a, b, o, e)
    print(f"[wfa] {label}: {len(pairs)} pairs cfg {(a, b, o, e)} drop {drop}", flush=True)
    got = pkg.wfa_batch(ctx, pairs, gcfg, drop)
    want = [O.wfa_align_scored(p, t, ocfg, ms, drop, tbf) for p, t, ms, tbf in pairs]
    bad = [i for i in range(len(pairs)) if got[i] != want[i]]
    assert not bad, [(i, pairs[i], got[i], want[i]) for i in bad[:3]]
    scores = pkg.wfa_batch(ctx, pairs, gcfg, drop, want_cigar=False)
    for g, w in zip(scores, want):
        assert (g is None) == (w is None)
        if g is not None:
            assert (g["score"], g["text_end"], g["cigar"]) == (w["score"], w["text_end"], [])
    return got


def test_wfa_reference_test_shapes(pkg, ctx):
    """tests/scored_wfa.rs:84-120: three cost sets, ample and loss - 1 budgets."""
    s = 0x6A09E667F3BCC909
    cfgs = [(1, 4, 6, 1), (2, 3, 2, 1), (3, 2, 0, 2)]
    by_cfg = {c: [] for c in cfgs}
    for trial in range(1800):
        c = cfgs[trial % 3]
        vals = []
        for _ in range(2):
            s = xorshift(s)
            vals.append(s >> 32)
        qlen, rlen = 1 + vals[0] % 28, 1 + vals[1] % 35
        q, r = bytearray(), bytearray()
        for _ in range(qlen):
            s = xorshift(s)
            q.append(b"ACGT"[(s >> 32) % 4])
        for _ in range(rlen):
            s = xorshift(s)
            r.append(b"ACGT"[(s >> 32) % 4])
        s = xorshift(s)
        lead = (s >> 32) % 5
        by_cfg[c].append((bytes(q), bytes(r), 300, lead))
    for c, pairs in by_cfg.items():
        got = _check_wfa(pkg, ctx, pairs, *c, label="scored_wfa shapes")
        bounded = [(q, r, g["score"] - 1, lead) for (q, r, _, lead), g in zip(pairs, got)
                   if g["score"] > 0]
        assert all(x is None for x in pkg.wfa_batch(ctx, bounded, pkg.sw_cfg(*c)))


@pytest.mark.parametrize("read_len,pad", [(150, 32), (100, 16), (250, 48), (300, 32)])
def test_wfa_read_like_pairs(pkg, ctx, read_len, pad):
    pairs = []
    for k, (read, win) in enumerate(read_like_pairs(read_len + pad, 600, read_len, pad)):
        budget = max((len(read) * 25 // 100) * 5, (6 + 2 * 4) * 3)  # wfa_score_budget
        pairs.append((read, win, budget if k % 7 else 20, k % 3 == 0 and 4 or 0))
    _check_wfa(pkg, ctx, pairs, 1, 4, 6, 1, label=f"reads {read_len}")


def test_wfa_adaptive_drop(pkg, ctx):
    pairs = [(r, w, 185, 0) for r, w in read_like_pairs(7, 400, 150, 32, max_edits=10)]
    for drop in (3, 10, 40):
        _check_wfa(pkg, ctx, pairs, 1, 4, 6, 1, drop=drop, label="adaptive")


def test_wfa_edges(pkg, ctx):
    pairs = [(b"", b"ACGT", 10, 0), (b"ACGT", b"", 10, 0), (b"ACGT", b"", 100, 3),
             (b"ACGT", b"ACGT", 0, 0), (b"ACGT", b"TTACGT", 0, 5), (b"ACGT", b"TTACGT", 0, 1),
             (b"NNNN", b"NNNN", 0, 0), (b"ACGTTT", b"ACGT", 50, 0), (b"A" * 40, b"C" * 40, -3, 0)]
    _check_wfa(pkg, ctx, pairs, 1, 4, 6, 1, label="edges")
    # WfaCosts::scored rejects non-positive match / mismatch / extend
    assert pkg.wfa_batch(ctx, pairs, pkg.sw_cfg(0, 4, 6, 1)) == [None] * len(pairs)
    with pytest.raises(pkg.HipError):
        pkg.wfa_batch(ctx, [(b"A", b"A", 70000, 0)], pkg.sw_cfg())


def test_long_pairs_near_the_size_limits(pkg, ctx):
    """Patterns of thousands of bases: Myers' global-scratch kernel (above 512
    bases) and WFA with 28 KB of pattern + text staged in LDS."""
    rng = np.random.default_rng(17)
    pairs_m, pairs_w = [], []
    for m in (700, 3000, 14000):
        text = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=m + 40).tolist())
        pat = bytearray(text[20:20 + m])
        for i in rng.integers(0, m, size=4):
            pat[int(i)] = ord("A") if pat[int(i)] != ord("A") else ord("C")
        del pat[m // 2:m // 2 + 3]
        pairs_m.append((bytes(pat), text, 12))
        pairs_w.append((bytes(pat), text[20:], 60, 0))
    print(f"[long] patterns {[len(p[0]) for p in pairs_m]}", flush=True)
    assert pkg.myers_batch(ctx, pairs_m) == [O.myers_bounded(p, t, k) for p, t, k in pairs_m]
    c = O.cfg()
    got = pkg.wfa_batch(ctx, pairs_w, pkg.sw_cfg())
    want = [O.wfa_align_scored(p, t, c, ms, 0, tbf) for p, t, ms, tbf in pairs_w]
    assert got == want and all(g is not None for g in got)
    with pytest.raises(pkg.HipError):
        pkg.wfa_batch(ctx, [(b"A" * 40000, b"A" * 30000, 10, 0)], pkg.sw_cfg())
