"""The fast DP cascade restatement (oracle/fastpath.py) on cases whose outcome
the mathematics fixes: exact and one-substitution reads take the packed
certificate, a read with an indel goes through Myers and WFA to a CIGAR whose
replay gives the score, a 5' deletion before the seed needs the suspect-origin
retry, a read whose second half is foreign fails the partial-read check, and
the router gates (length, algorithm, chain-quality gate, window checks)."""
import numpy as np

from oracle import fastpath as F
from oracle import oracle as O

C = O.cfg(1, 4, 6, 1, 32, 100, 5)


def genome(seed=3, n=4000):
    rng = np.random.default_rng(seed)
    return bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=n).tolist())


def replay_score(read, contig, rec, c=C):
    qi, ti, score = 0, rec["ref_start"], 0
    for w in rec["cigar"]:
        n, op = w >> 4, w & 15
        if op == 0:
            for _ in range(n):
                score += c.match_score if read[qi] == contig[ti] else -c.mismatch
                qi, ti = qi + 1, ti + 1
        elif op == 1:
            qi += n
            score -= c.gap_open + c.gap_extend * n
        else:
            ti += n
            score -= c.gap_open + c.gap_extend * n
    assert qi == len(read) and ti == rec["ref_end"]
    return score


def test_exact_and_substitution_reads_take_the_certificate():
    g = genome()
    read = g[1000:1150]
    kind, rec = F.fast_dp(read, g, 1000 + 20, 20, 150, C)
    assert kind == "packed" and rec["cigar"] == [150 << 4] and rec["ref_start"] == 1000
    assert rec["score"] == 150 and rec["nm"] == 0 and rec["md"] == "150"
    r = bytearray(read)
    r[70] = ord("A") if r[70] != ord("A") else ord("C")
    kind, rec = F.fast_dp(bytes(r), g, 1000, 0, 150, C)
    assert kind == "packed" and rec["score"] == 149 - 4 and rec["nm"] == 1


def test_insertion_goes_to_wfa():
    g = genome()
    read = g[2000:2075] + b"TTT" + g[2075:2147]  # 150 bases, 3-base insertion
    kind, rec = F.fast_dp(read, g, 2000, 0, 100, C)
    assert kind == "wfa"
    assert rec["ref_start"] == 2000 and rec["ref_end"] == 2147
    assert sum(w >> 4 for w in rec["cigar"] if w & 15 == 1) == 3
    assert rec["score"] == replay_score(read, g, rec) == 147 - (6 + 3)
    assert rec["nm"] == 3


def test_five_prime_deletion_needs_the_origin_retry():
    g = genome(5)
    # the seed implies the window starts at 1000, but the read starts at 995
    # and lacks 5 reference bases right after its first 10 bases
    read = g[995:1005] + g[1010:1150]
    kind, rec = F.fast_dp(read, g, 1000, 0, 100, C)
    assert kind == "wfa"
    assert rec["ref_start"] == 995
    assert rec["cigar"][:3] == [10 << 4, (5 << 4) | 2, (140 << 4)]
    assert rec["score"] == replay_score(read, g, rec) == 150 - (6 + 5)


def test_partial_read_falls_back():
    g = genome(7)
    rng = np.random.default_rng(9)
    tail = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=30).tolist())
    read = g[3000:3120] + tail
    text = g[3000:3182]
    # Myers (bound 150 * 15 / 100 + 4 = 26) lets it through and WFA aligns it,
    # but the loss is above 30% of the perfect score and clipping the foreign
    # tail would score more: the cascade declines
    assert O.myers_bounded(read, text, 26) is not None
    w = O.wfa_align_scored(read, text, C, F.wfa_score_budget(150, 5, 6, 2, F.Router()))
    assert w is not None and w["score"] * 10 > 150 * 3
    assert F.clipping_can_improve(read, text, w, C)
    assert F.fast_dp(read, g, 3000, 0, 100, C) is None


def test_router_gates():
    g = genome()
    read = g[500:650]
    assert F.fast_dp(b"", g, 500, 0, 1, C) is None
    assert F.fast_dp(g[0:301], g, 0, 0, 1, C) is None                        # > 300 bases
    assert F.fast_dp(read, g, 500, 0, 1, C, F.Router(algo="sw")) is None
    assert F.fast_dp(read, g, 500, 0, 1, C, F.Router(algo="bogus"))[0] == "packed"
    assert F.fast_dp(read, g, 10, 20, 1, C) is None                          # negative start
    assert F.fast_dp(read, g, len(g), 0, 1, C) is None                       # past the contig
    assert F.fast_dp(read, g, len(g) - 100, 0, 1, C) is None                 # window too short
    gate = F.Router(quality_gate=True)
    assert F.fast_dp(read, g, 500, 0, 104, C, gate) is None                  # 69% < 70%
    assert F.fast_dp(read, g, 500, 0, 105, C, gate)[0] == "packed"
    kind, rec = F.fast_dp(read, g, 500, 0, 1, C, F.Router(algo="wfa"))      # skips the sieve
    assert kind == "wfa" and rec["cigar"] == [150 << 4] and rec["score"] == 150


def test_n_in_window_skips_the_certificate():
    g = bytearray(genome())
    g[1160] = ord("N")  # inside the L + band window, outside the read's span
    g = bytes(g)
    kind, rec = F.fast_dp(g[1000:1150], g, 1000, 0, 1, C)
    assert kind == "wfa" and rec["cigar"] == [150 << 4]


def test_myers_reject():
    g = genome()
    rng = np.random.default_rng(1)
    read = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=150).tolist())
    assert F.fast_dp(read, g, 1000, 0, 1, C) is None


def test_c_cascade_matches_the_python_cascade():
    """The C restatement used as the multi-core CPU baseline
    (oracle/kira_fast_oracle.c) gives the Python cascade's kind, coordinates,
    score and NM on every chain of a mixed batch (all twelve read kinds: exact,
    substitutions, indels, 5' deletions, foreign tails, N, contig ends, ...),
    on one thread and on several, with the default and non-default routers."""
    from tests.fastpath_data import make_fast_batch
    b = make_fast_batch(5, 1200)
    for router in (F.Router(), F.Router(wfa_lead=8, cert_strict=True, wfa_budget_pct=15)):
        want = [F.fast_dp(r, b["contigs"][ch[0]], ch[1], ch[3], ch[6], C, router)
                for r, ch in zip(b["reads"], b["chains"])]
        for threads in (1, 4):
            got = F.fast_dp_batch_mt(b["reads"], b["chains"], b["genome"], b["contig_off"],
                                     b["contig_len"], C, router, threads)
            kinds = {0: 0, 1: 0, 2: 0}
            for k, (w, g) in enumerate(zip(want, got)):
                kinds[int(g["kind"])] += 1
                if w is None:
                    assert g["kind"] == 0, k
                    continue
                kind, rec = w
                assert int(g["kind"]) == (1 if kind == "packed" else 2), k
                assert (int(g["ref_start"]), int(g["ref_end"]), int(g["score"]), int(g["nm"]),
                        int(g["n_cigar"])) == (rec["ref_start"], rec["ref_end"], rec["score"],
                                               rec["nm"], len(rec["cigar"])), k
            assert min(kinds.values()) > 50, kinds
