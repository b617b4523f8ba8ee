"""The Myers and WFA restatements (oracle/kira_wfa_oracle.c) against the
reference's own independent oracles: the naive semi-global edit distance of
tests/myers_bounded_edit.rs:9-31 and the affine DP + CIGAR replay of
tests/scored_wfa.rs:17-78, on the same known answers and random corpora."""
from oracle import oracle as O

M64 = (1 << 64) - 1


def naive_semi_global(p, t):
    """tests/myers_bounded_edit.rs:9-31 (first minimal end position).  A base
    outside ACGT matches nothing, itself included: the bit-vector algorithm
    sets no match mask for it (src/alignment/myers.rs:8-16, 243-257)."""
    m, n = len(p), len(t)
    prev = [0] * (n + 1)
    for i in range(1, m + 1):
        cur = [i] + [0] * n
        for j in range(1, n + 1):
            cost = 0 if p[i - 1] == t[j - 1] and p[i - 1] in b"ACGT" else 1
            cur[j] = min(prev[j - 1] + cost, prev[j] + 1, cur[j - 1] + 1)
        prev = cur
    best, pos = prev[0], 0
    for j, v in enumerate(prev):
        if v < best:
            best, pos = v, j
    return best, pos


def xorshift(s):
    s ^= (s << 13) & M64
    s ^= s >> 7
    s ^= (s << 17) & M64
    return s


def test_myers_known_answers():
    assert O.myers_bounded(b"", b"ACGT", 0) == (0, 0)
    assert O.myers_bounded(b"ACGT", b"", 4) == (4, 0)
    for p, t, k in [(b"ACGTACGTACGT", b"NNACGTACGTACGTNN", 5),
                    (b"ACGTACGTACGT", b"NNACGTAXGTACGTNN", 3),
                    (b"ACGTACGT", b"NNACGTAAACGTNN", 3),
                    (b"ACGTACGT", b"NNACGACGTNN", 3)]:
        assert O.myers_bounded(p, t, k) == naive_semi_global(p, t), (p, t)
    assert O.myers_bounded(b"AAAAAAAA", b"TTTTTTTT", 3) is None
    assert O.myers_bounded(b"AAAAAAAA", b"TTTTTTTT", 8)[0] == 8
    assert O.myers_bounded(b"ACNT", b"ACNT", 3) == naive_semi_global(b"ACNT", b"ACNT") == (1, 4)
    # block boundaries (64, 65) and a 150-base pattern (myers_bounded_edit.rs:102-135)
    p = bytes(b"ACGT"[i % 4] for i in range(64))
    t = bytearray(p)
    t[20] = ord("A")
    assert O.myers_bounded(p, bytes(t), 5)[0] == naive_semi_global(p, bytes(t))[0]
    p = bytes(b"ACGT"[i % 4] for i in range(65))
    t = bytearray(p)
    t[64] = ord("C") if p[64] == ord("A") else ord("A")
    assert O.myers_bounded(p, bytes(t), 5)[0] == naive_semi_global(p, bytes(t))[0]
    p = bytes(b"ACGT"[(i * 7 + 3) % 4] for i in range(150))
    t = bytearray(p)
    for i, x, y in [(10, "A", "T"), (77, "C", "G"), (140, "G", "A")]:
        t[i] = ord(y) if t[i] == ord(x) else ord(x)
    assert O.myers_bounded(p, bytes(t), 10)[0] == naive_semi_global(p, bytes(t))[0]


def test_myers_random_and_accept_boundary():
    s = 0x5EED_1234
    for trial in range(70):
        s = xorshift(s)
        m = 1 + s % (330 if trial % 5 == 0 else 90)
        s = xorshift(s)
        n = m + s % 40
        p = bytearray()
        for _ in range(m):
            s = xorshift(s)
            p.append(b"ACGTN"[s % 5] if s % 97 == 0 else b"ACGT"[s % 4])
        t = bytearray()
        for _ in range(n):
            s = xorshift(s)
            t.append(b"ACGT"[s % 4])
        # plant the pattern with a few edits so distances stay small
        s = xorshift(s)
        at = s % (n - m + 1)
        t[at:at + m] = p
        for _ in range(trial % 6):
            s = xorshift(s)
            t[at + s % m] = b"ACGT"[(s >> 8) % 4]
        want, _ = naive_semi_global(bytes(p), bytes(t))
        for k in sorted({0, max(want - 1, 0), want, want + 1, m}):
            got = O.myers_bounded(bytes(p), bytes(t), k)
            if want <= k:
                assert got is not None and got[0] == want, (trial, k)
            else:
                assert got is None, (trial, k)


def test_myers_both_sides_of_the_256_boundary():
    """tests/myers_bounded_edit.rs:175-212: multi-block patterns around 256."""
    s = 0x0BADC0DE
    for m in [200, 255, 256, 257, 300, 400]:
        for trial in range(3):
            s = xorshift(s)
            n = m + s % 60
            p = bytearray()
            for _ in range(m):
                s = xorshift(s)
                p.append(b"ACGT"[s % 4])
            t = bytearray()
            for _ in range(n):
                s = xorshift(s)
                t.append(b"ACGT"[s % 4])
            if trial % 2 == 0:
                s = xorshift(s)
                at = s % (n - m + 1)
                t[at:at + m] = p
                s = xorshift(s)
                q = at + s % m
                t[q] = ord("T") if t[q] == ord("A") else ord("A")
            truth, _ = naive_semi_global(bytes(p), bytes(t))
            assert O.myers_bounded(bytes(p), bytes(t), m)[0] == truth, (m, trial)
            if truth > 0:
                assert O.myers_bounded(bytes(p), bytes(t), truth - 1) is None


def affine_oracle(q, r, a, b, o, e, lead):
    """tests/scored_wfa.rs:17-45: best full-query score with `lead` free text bases."""
    width = len(r) + 1
    neg = -(2 ** 31) // 8
    h = [neg] * ((len(q) + 1) * width)
    ins, dele = h[:], h[:]
    lead = min(lead, max(len(r) - 1, 0))
    for j in range(width):
        h[j] = 0 if j <= lead else -o - e * (j - lead)
    for i in range(1, len(q) + 1):
        h[i * width] = -o - e * i
        ins[i * width] = h[i * width]
        for j in range(1, width):
            p = i * width + j
            ins[p] = max(h[p - width] - o - e, ins[p - width] - e)
            dele[p] = max(h[p - 1] - o - e, dele[p - 1] - e)
            sub = a if q[i - 1] == r[j - 1] else -b
            h[p] = max(h[p - width - 1] + sub, ins[p], dele[p])
    return max(h[len(q) * width:])


def replay(q, r, aln, a, b, o, e):
    """tests/scored_wfa.rs:47-78."""
    qi, ri, score = 0, aln["text_start"], 0
    for w in aln["cigar"]:
        n, op = w >> 4, w & 15
        if op == 0:
            for _ in range(n):
                score += a if q[qi] == r[ri] else -b
                qi += 1
                ri += 1
        elif op == 1:
            qi += n
            score -= o + e * n
        elif op == 2:
            ri += n
            score -= o + e * n
        else:
            raise AssertionError("unexpected CIGAR op")
    assert qi == len(q) and ri == aln["text_end"]
    assert a * len(q) - score == aln["score"]
    return score


def test_scored_wfa_matches_affine_oracle_replay_and_score_only():
    s = 0x6A09E667F3BCC909

    def nxt():
        nonlocal s
        s = xorshift(s)
        return s >> 32

    cfgs = [(1, 4, 6, 1), (2, 3, 2, 1), (3, 2, 0, 2)]
    for trial in range(900):
        a, b, o, e = cfgs[trial % 3]
        c = O.cfg(a, b, o, e, 50, 0, 5)
        qlen = 1 + nxt() % 28
        rlen = 1 + nxt() % 35
        q = bytes(b"ACGT"[nxt() % 4] for _ in range(qlen))
        r = bytes(b"ACGT"[nxt() % 4] for _ in range(rlen))
        lead = nxt() % 5
        expected = affine_oracle(q, r, a, b, o, e, lead)
        aln = O.wfa_align_scored(q, r, c, 300, 0, lead)
        assert aln is not None
        assert replay(q, r, aln, a, b, o, e) == expected, trial
        assert O.wfa_score_only_scored(q, r, c, 300, 0, lead) == expected
        loss = len(q) * a - expected
        if loss > 0:
            assert O.wfa_align_scored(q, r, c, loss - 1, 0, lead) is None
            assert O.wfa_score_only_scored(q, r, c, loss - 1, 0, lead) is None


def test_adaptive_drop_never_beats_the_exact_optimum():
    s = 0xBB67AE8584CAA73B
    c = O.cfg(1, 4, 6, 1, 50, 0, 5)
    for _ in range(200):
        s = xorshift(s)
        q = bytes(b"ACGT"[(s >> (2 * k)) % 4] for k in range(24))
        s = xorshift(s)
        r = bytes(b"ACGT"[(s >> (2 * k)) % 4] for k in range(30))
        exact = O.wfa_align_scored(q, r, c, 400)
        pruned = O.wfa_align_scored(q, r, c, 400, 3)
        if pruned is not None:
            assert pruned["score"] >= exact["score"]
            replay(q, r, pruned, 1, 4, 6, 1)
