"""Pins the CPU oracle's Smith-Waterman rows against the reference's own known
answers, a textbook Gotoh local alignment on tiny inputs, and the CIGAR-replay
invariant."""
import random

import pytest

from oracle import oracle as O
from tests.refgen import gen_ref, lcg_dna

NEG = -(2 ** 31) // 4
ABORT_NONE = -(2 ** 31) // 8


def clamp_window(ref_len, ref_start, ref_end, bw):
    """clamp_window — src/alignment/mod.rs:1766-1771."""
    flank = max(max(bw, 1) * 2, 50)
    start = max(ref_start - flank, 0)
    end = min(ref_end + flank, ref_len)
    return start, max(end, start + 1)


def align_chain(read, ref, span, c, abort=NEG):
    """align_oriented_chain_with_meta without the exact-match shortcut
    (src/alignment/mod.rs:851-877)."""
    rs, re_, qs, qe = span
    ws, we = clamp_window(len(ref), rs, re_, c.bandwidth)
    win = ref[ws:we]
    sw = O.banded_sw(read, win, rs - ws - qs, c, abort)
    return O.build_alignment(read, win, ws, sw)


def replay_score(read, ref, sw, c):
    """Affine score of the CIGAR replayed over read/ref (mod.rs:2592-2635)."""
    qi, ri, s = sw["read_start"], sw["ref_start"], 0
    for op in sw["cigar"]:
        n, k = op >> 4, O.CIGAR_OPS[op & 15]
        if k == "M":
            for _ in range(n):
                s += c.match_score if read[qi] == ref[ri] else -c.mismatch
                qi += 1
                ri += 1
        elif k == "I":
            s -= c.gap_open + c.gap_extend * n
            qi += n
        elif k == "D":
            s -= c.gap_open + c.gap_extend * n
            ri += n
    assert qi == sw["read_end"] and ri == sw["ref_end"]
    return s


def gotoh_local(a, b, c):
    """Textbook local affine alignment score (gap of length k costs O + k*E)."""
    A, B, O_, E = c.match_score, c.mismatch, c.gap_open, c.gap_extend
    n, m = len(a), len(b)
    H = [[0] * (m + 1) for _ in range(n + 1)]
    Ei = [[-10 ** 9] * (m + 1) for _ in range(n + 1)]
    F = [[-10 ** 9] * (m + 1) for _ in range(n + 1)]
    best = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            Ei[i][j] = max(H[i - 1][j] - O_ - E, Ei[i - 1][j] - E)
            F[i][j] = max(H[i][j - 1] - O_ - E, F[i][j - 1] - E)
            s = A if a[i - 1] == b[j - 1] else -B
            H[i][j] = max(0, H[i - 1][j - 1] + s, Ei[i][j], F[i][j])
            best = max(best, H[i][j])
    return best


# ---- tests/alignment.rs
def test_alignment_finds_full_match():
    c = O.cfg(2, 4, 6, 1, 50, 40, 5)
    aln = align_chain(b"ACGTACGT", b"TTACGTACGTGG", (2, 10, 0, 8), c)
    assert "8M" in aln["cigar"] and aln["nm"] == 0


def _end_indel_fixture():
    base = bytes(b"CGTA"[i % 4] for i in range(150))
    ref = b"TT" + base + b"TT"
    read = base[:145] + b"G" + base[145:149]
    return ref, read


def test_end_of_read_insertion_emits_i_op():
    ref, read = _end_indel_fixture()
    aln = align_chain(read, ref, (2, 147, 0, 145), O.cfg(1, 4, 6, 1, 50, 50, 5))
    assert "I" in aln["cigar"]
    trailing = aln["cigar"].rsplit("M", 1)[-1]
    assert not trailing.endswith("S") or int(trailing.rstrip("S").split("I")[-1] or 0) <= 1


def test_end_of_read_insertion_softclips_without_penalty():
    ref, read = _end_indel_fixture()
    aln = align_chain(read, ref, (2, 147, 0, 145), O.cfg(1, 4, 6, 1, 50, 50, 0))
    assert "I" not in aln["cigar"]


def test_simd_api_cardinality_and_order():
    """tests/alignment.rs:124-171."""
    bases = b"ACGTTGCTAATCGTGCATGACCTG"
    c = O.cfg(2, 4, 6, 1, 50, 40, 5)
    jobs = [O.Job(bases, bases, 100 * i, 100 * i, 100 * i + len(bases), 0, len(bases), 0)
            for i in range(33)]
    for mode in ["avx2", "neon", "avxvnni"]:
        for count in [0, 1, 4, 5, 16, 17, 33]:
            res = O.align_batch_sw(jobs[:count], c, mode)
            assert len(res) == count
            for j, r in zip(jobs, res):
                aln = O.build_alignment(j.read, j.window, j.win_start, r)
                assert (aln["ref_start"], aln["ref_end"]) == (j.win_start, j.chain_ref_end)
                assert aln["nm"] == 0


def test_scalar_and_mixed_batches_follow_the_seed_diagonal():
    """tests/alignment.rs:173-224."""
    read = b"GATTACAGCTTGAGTCCACG"
    window = b"C" * 100 + read
    c = O.cfg(2, 4, 6, 1, 8, 40, 5)
    full = O.Job(read, window, 200, 303, 320, 3, 20, ABORT_NONE // 2 * 2 // 2)
    full.abort_score = NEG
    shorter = O.Job(read[:19], window, 200, 303, 320, 3, 20, NEG)
    for batch in ([full], [full, shorter]):
        for mode in ["scalar", "avx2", "neon"]:
            r = O.align_batch_sw(batch, c, mode)[0]
            aln = O.build_alignment(full.read, window, 200, r)
            assert (aln["ref_start"], aln["ref_end"], aln["score"], aln["nm"]) == (300, 320, 40, 0)
            assert aln["cigar"] == "20M"


@pytest.mark.parametrize("length,reward,offset", [(7, 4681, 0), (8, 4096, 0), (150, 300, 0),
                                                  (8, 1, 32768)])
def test_i16_boundary(length, reward, offset):
    """tests/alignment.rs:226-274: inside the i16 range the SIMD path is used,
    outside it the scalar path; both give the exact answer."""
    read = b"G" * length
    ref = b"C" * offset + read
    c = O.cfg(reward, 4, 6, 1, 8, 40, 5)
    job = O.Job(read, ref, 0, offset, offset + length, 0, length, NEG)
    for mode in ["avx2", "avxvnni"]:
        r = O.align_batch_sw([job], c, mode)[0]
        aln = O.build_alignment(read, ref, 0, r)
        assert aln["score"] == reward * length
        assert (aln["ref_start"], aln["ref_end"]) == (offset, offset + length)
        assert aln["cigar"] == f"{length}M"


# ---- src/alignment/mod.rs:2326-2635 (adaptive band)
def _cfg(bw):
    return O.cfg(1, 4, 6, 1, bw, 1000, 5)


def _del_len(cig):
    return sum(o >> 4 for o in cig if O.CIGAR_OPS[o & 15] == "D")


def test_adaptive_recovers_drifted_indels():
    ref = gen_ref(120)
    read = ref[0:30] + ref[33:60] + ref[63:120]
    ad = O.banded_sw(read, ref, 0, _cfg(4), ABORT_NONE, O.knobs(adaptive=True))
    fx = O.banded_sw(read, ref, 0, _cfg(4), ABORT_NONE, O.knobs(adaptive=False))
    al = lambda r: r["read_end"] - r["read_start"]
    assert _del_len(ad["cigar"]) == 6 and al(ad) >= 113
    assert _del_len(fx["cigar"]) < 6
    assert ad["score"] > fx["score"] + 15 and al(ad) > al(fx) + 15


def test_gapless_adaptive_equals_fixed():
    ref = gen_ref(120)
    read = ref[10:90]
    ad = O.banded_sw(read, ref, 10, _cfg(6), ABORT_NONE, O.knobs(adaptive=True))
    fx = O.banded_sw(read, ref, 10, _cfg(6), ABORT_NONE, O.knobs(adaptive=False))
    # same alignment; the two bands cover different cells on some rows, so
    # only the evaluated cell counts differ
    ad.pop("cells")
    fx.pop("cells")
    assert ad == fx and _del_len(ad["cigar"]) == 0 and ad["read_end"] - ad["read_start"] >= 79


def test_first_base_match_is_not_softclipped():
    ref = gen_ref(80)
    read = ref[0:60]
    for adaptive in (False, True):
        r = O.banded_sw(read, ref, 0, _cfg(8), ABORT_NONE, O.knobs(adaptive=adaptive))
        assert (r["read_start"], r["read_end"], r["score"]) == (0, 60, 60)
        assert O.CIGAR_OPS[r["cigar"][0] & 15] == "M"


def test_affine_traceback_score_matches_replayed_cigar():
    ref = gen_ref(90)
    read = ref[:40] + b"TTGCA" + ref[40:]
    c = _cfg(20)
    r = O.banded_sw(read, ref, 0, c, ABORT_NONE, O.knobs())
    assert replay_score(read, ref, r, c) == r["score"]
    assert any(O.CIGAR_OPS[o & 15] == "I" and (o >> 4) >= 5 for o in r["cigar"])


def test_leading_softclip_is_not_counted_twice_in_md():
    sw = dict(ref_start=0, ref_end=4, read_start=2, read_end=6, score=4, cigar=[(4 << 4) | 0])
    aln = O.build_alignment(b"TTACGT", b"ACGT", 0, sw)
    assert (aln["nm"], aln["md"], aln["cigar"]) == (0, "4", "2S4M")


# ---- md.rs:214-285
def test_md_known_answers():
    assert O.nm_md_from_cigar(b"ACGT", b"ACGT", 0, 0, "4M") == (0, "4")
    assert O.nm_md_from_cigar(b"ACGT", b"AGGT", 0, 0, "4M") == (1, "1G2")
    assert O.nm_md_from_cigar(b"TTGA", b"ACGT", 0, 0, "4M") == (3, "0A0C1T0")
    assert O.nm_md_from_cigar(b"ACGTACGT", b"ACGTTTACGT", 0, 0, "3M2D5M") == (2, "3^TT5")
    assert O.nm_md_from_cigar(b"ACGAAGTA", b"ACGGTA", 0, 0, "3M2I3M") == (2, "6")
    assert O.nm_md_from_cigar(b"NNACGN", b"XXXXACG", 0, 4, "2S3M1S") == (0, "3")
    ref = bytearray(b"T" * 104)
    ref[0:2] = b"AC"
    ref[102:104] = b"GA"
    assert O.nm_md_from_cigar(b"ACGA", bytes(ref), 0, 0, "2M100N2M") == (0, "4")


# ---- textbook / brute-force pins
def _rand_pair(rng, n, m):
    a = bytes(rng.choice(b"ACGT") for _ in range(n))
    b = bytes(rng.choice(b"ACGT") for _ in range(m))
    return a, b


def test_full_band_banded_sw_equals_textbook_local():
    """With a band covering the matrix, no x-drop, no abort and clip 0, the
    banded DP is plain local affine SW."""
    rng = random.Random(42)
    for _ in range(150):
        a, b = _rand_pair(rng, rng.randrange(1, 14), rng.randrange(1, 18))
        if rng.random() < 0.5 and len(b) > 4:
            st = rng.randrange(len(b) - 3)
            a = b[st:st + rng.randrange(3, len(b) - st + 1)]
            a = a[:2] + a[3:] if len(a) > 5 and rng.random() < 0.5 else a
        c = O.cfg(rng.choice([1, 2]), rng.choice([1, 4]), rng.choice([0, 6]), rng.choice([1, 2]),
                  64, 0, 0)
        for adaptive in (False, True):
            r = O.banded_sw(a, b, 0, c, ABORT_NONE, O.knobs(adaptive=adaptive, max_drift=64))
            assert r["score"] == gotoh_local(a, b, c)
            if r["score"] > 0:
                assert replay_score(a, b, r, c) == r["score"]


@pytest.mark.parametrize("bits", [16, 8, 32])
def test_full_window_group_equals_textbook_local(bits):
    rng = random.Random(bits)
    for _ in range(40):
        L, W = rng.randrange(1, 12), rng.randrange(1, 16)
        jobs, cs = [], O.cfg(1, 4, 6, 1, 8, 40, 0)
        for _k in range(rng.randrange(1, 5)):
            a, b = _rand_pair(rng, L, W)
            jobs.append(O.Job(a, b, 0, 0, 0, 0, L, ABORT_NONE))
        res = O.sw_group(jobs, cs, bits=bits)
        for j, r in zip(jobs, res):
            assert r["score"] == gotoh_local(j.read, j.window, cs)
            if r["score"] > 0:
                assert replay_score(j.read, j.window, r, cs) == r["score"]


def test_group_early_stop_semantics():
    """Rows stop only once every lane of the group is past its abort bound
    (mod.rs:2058-2085): a lane that could still reach its bound keeps the
    whole group running, and an aborted lane reports early_abort."""
    rng = random.Random(3)
    ref = lcg_dna(200, 11)
    good = ref[20:170]
    bad = lcg_dna(150, 99)
    c = O.cfg(1, 4, 6, 1, 32, 100, 5)
    J = lambda read, ab: O.Job(read, ref[:278] if len(ref) >= 278 else ref, 0, 20, 170, 0, 150, ab)
    alone = O.sw_group([J(bad, 100)], c)[0]
    assert alone["early_abort"]
    mixed = O.sw_group([J(bad, 100), J(good, 100)], c)
    assert not mixed[1]["early_abort"] and mixed[1]["score"] == 150
    assert mixed[0]["early_abort"]  # done, but the group ran on
    # the group ran every row, so the aborted lane's result is its no-abort
    # result (last-row clip rule included), only the flag differs
    free = O.sw_group([J(bad, ABORT_NONE)], c)[0]
    assert {**mixed[0], "early_abort": False} == free
    assert alone["score"] != free["score"]  # the stop changed the answer here


def test_int8_matches_truth_on_clean_reads():
    """tests/unit/alignment_sw_int8_vnni.rs:78-124 (rand_dna generator)."""
    c = O.cfg(1, 4, 6, 1, 50, 50, 5)
    jobs = []
    for k in range(32):
        r = lcg_dna(130, 0xCAFE_BABE_DEAD ^ k)
        jobs.append(O.Job(r[15:115], r, 0, 0, 100, 0, 100, ABORT_NONE))
    for r in O.sw_group(jobs, c, bits=8):
        assert r["score"] == 100


def test_avx2_group_form_equals_the_generic_lane_loop():
    """align_batch_sw's AVX2 mode runs the 16-lane group in AVX2 registers;
    the generic lane loop (sw_group, bits=16) must give the same results on
    groups of 1-16 lanes with N bases, case, indels, unrelated reads, early
    stops and several cost sets (gap open 0 included)."""
    rng = random.Random(5)

    def seq(n, alpha=b"ACGT"):
        return bytes(rng.choice(alpha) for _ in range(n))

    for _ in range(120):
        L = rng.choice([20, 50, 100, 150])
        W = L + rng.choice([0, 10, 64, 128])
        n = rng.randint(1, 16)
        jobs = []
        for _ in range(n):
            w = bytearray(seq(W, b"ACGTN" if rng.random() < 0.2 else b"ACGT"))
            st = rng.randint(0, W - L)
            rd = bytearray(w[st:st + L])
            for _ in range(rng.randint(0, 12)):
                rd[rng.randrange(L)] = rng.choice(b"ACGTNa")
            if rng.random() < 0.3:
                del rd[rng.randrange(1, L - 1)]
                rd.append(rng.choice(b"ACGT"))
            if rng.random() < 0.2:
                rd = bytearray(seq(L))
            ab = rng.choice([-(2 ** 31) // 4, -(2 ** 31) // 8, rng.randint(-20, L)])
            jobs.append(O.Job(bytes(rd), bytes(w), 1000, 1000 + st, 1000 + st + L, 0, L, ab))
        c = O.cfg(rng.choice([1, 2]), rng.choice([3, 4]), rng.choice([0, 4, 6]),
                  rng.choice([1, 2]), 32, 100, 5)
        assert O.align_batch_sw(jobs, c, "avx2") == O.sw_group(jobs, c, bits=16)
