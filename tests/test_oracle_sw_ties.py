"""Pins the oracle's choice among equal-score Smith-Waterman alignments.

The reference picks one alignment out of all co-optimal ones through three
pieces of code: the per-cell priority of the banded DP (src/alignment/mod.rs:
1505-1523), the trace codes of the AVX2 full-window DP (mod.rs:2022-2035) and
the traceback walk (mod.rs:1667-1716, banded variant :1581-1626).  Read
together they select, among every alignment path that reaches the chosen end
cell with the optimal score, the one whose moves, read backwards from the end
cell, are lexicographically greatest under this order:

* in the H state: stop (the prefix scores exactly 0) > match/mismatch column
  (DIAG) > the cell's value comes from an insertion (E) > from a deletion (F);
* inside a gap: close the gap (the cell was opened from H) > extend it (the
  extend bit is set only when extending is strictly better, mod.rs:1516-1521,
  2003/2009).

This module restates the DP values independently (a plain three-state Gotoh
recursion over the cells the reference evaluates), enumerates EVERY co-optimal
path explicitly, picks one by the rule above written over whole paths (no trace
codes), and requires the oracle to return exactly that alignment on tie-heavy
tiny inputs.  A mutation check shows that each plausible tie mistake (E/F
priority swapped, extend on ">=", gaps before matches, no stop at a zero
prefix) picks a different alignment on this corpus, so the pin has teeth.  It
also pins, with hand derivations, the exact CIGARs of the reference's own
gapped test inputs."""
import random

import pytest

from oracle import oracle as O
from tests.refgen import gen_ref

ABORT_NONE = -(2 ** 31) // 8
NEG_A = -16384  # the i16 lanes' minus infinity for row-0 E / column-0 F (mod.rs:1948)
ABSENT = -(2 ** 31) // 4  # the banded DP's absent E / F input

# move priorities (higher wins), H state and gap states
P_H = {"stop": 3, "diag": 2, "toI": 1, "toD": 0}
P_GAP = {"open": 1, "ext": 0}


class Values:
    """DP values of one job: H, E (insertion, vertical) and F (deletion,
    horizontal) per evaluated cell, the evaluated rows and the end cell."""

    def __init__(self):
        self.H, self.E, self.F = {}, {}, {}
        self.band = {}  # row -> (j_start, j_end) or None for an empty row
        self.end = (0, 0)
        self.score = 0
        self.early_abort = False


def values_full_window(read, ref, c, abort=ABORT_NONE):
    """Semantics A (sw_batch_avx2_inner, mod.rs:1914-2135) for one lane group of
    one job: every column of every row, zero row 0 and column 0."""
    L, W = len(read), len(ref)
    go, ge = c.gap_open + c.gap_extend, c.gap_extend
    v = Values()
    for j in range(W + 1):
        v.H[(0, j)] = 0
        v.E[(0, j)] = NEG_A
    best = bi = bj = qbest = qj = 0
    for i in range(1, L + 1):
        if abort > ABORT_NONE and best + (L - i + 1) * c.match_score < abort:
            v.early_abort = True
            break
        v.band[i] = (1, W)
        v.H[(i, 0)] = 0
        v.F[(i, 0)] = NEG_A
        for j in range(1, W + 1):
            s = c.match_score if read[i - 1] == ref[j - 1] else -c.mismatch
            e = max(v.H[(i - 1, j)] - go, v.E[(i - 1, j)] - ge)
            f = max(v.H[(i, j - 1)] - go, v.F[(i, j - 1)] - ge)
            h = max(v.H[(i - 1, j - 1)] + s, e, f, 0)
            v.H[(i, j)], v.E[(i, j)], v.F[(i, j)] = h, e, f
            if h > best:
                best, bi, bj = h, i, j
            if i == L and h > qbest:
                qbest, qj = h, j
    v.score, v.end = best, (bi, bj)
    if not v.early_abort and qbest > 0 and qbest + c.clip_penalty > best:
        v.score, v.end = qbest, (L, qj)
    return v


def values_banded(read, ref, off, c, abort=ABORT_NONE, adaptive=True, max_drift=16):
    """Semantics B (banded_sw_with, mod.rs:1370-1650): per-row bands that follow
    the previous row's best column when adaptive; a cell's E / F input outside
    the previous / current band is absent; the diagonal outside the previous
    band reads H = 0 but the pair still scores."""
    L, W = len(read), len(ref)
    go, ge = c.gap_open + c.gap_extend, c.gap_extend
    band = max(c.bandwidth, 1)
    md = max_drift if adaptive else 0
    v = Values()
    best = bi = bj = qbest = qj = 0
    centre_next = 1 + off
    prev = None
    for i in range(1, L + 1):
        static = i + off
        centre = min(max(centre_next, static - md), static + md) if adaptive else static
        js, je = max(centre - band, 1), min(centre + band, W)
        if js > je:
            v.band[i] = None
            prev = None
            continue
        v.band[i] = (js, je)
        inprev = lambda j: prev is not None and prev[0] <= j <= prev[1]
        row_best, row_best_j = 0, -1
        for j in range(js, je + 1):
            s = c.match_score if read[i - 1] == ref[j - 1] else -c.mismatch
            hd = v.H[(i - 1, j - 1)] if inprev(j - 1) else 0
            # an input outside the band is absent: i32::MIN / 4 (mod.rs:1487-1502)
            e = max(v.H[(i - 1, j)] - go, v.E[(i - 1, j)] - ge) if inprev(j) else ABSENT
            f = max(v.H[(i, j - 1)] - go, v.F[(i, j - 1)] - ge) if j > js else ABSENT
            h = max(hd + s, e, f, 0)
            v.H[(i, j)], v.E[(i, j)], v.F[(i, j)] = h, e, f
            if h > row_best:
                row_best, row_best_j = h, j
            if h > best:
                best, bi, bj = h, i, j
            if i == L and h > qbest:
                qbest, qj = h, j
        prev = (js, je)
        if adaptive:
            raw = row_best_j + 1 if row_best_j >= 0 else centre_next + 1
            centre_next = min(max(raw, static + 1 - md), static + 1 + md)
        if c.xdrop > 0 and best - row_best > c.xdrop:
            v.early_abort = True
            break
        if abort > ABORT_NONE and best + (L - i) * c.match_score < abort:
            v.early_abort = True
            break
    v.score, v.end = best, (bi, bj)
    if qbest > 0 and qbest + c.clip_penalty > best:
        v.score, v.end = qbest, (L, qj)
    return v


class TooMany(Exception):
    pass


def co_optimal_paths(read, ref, c, v, cap=20000):
    """Every path that reaches v.end in the H state with score v.score, as
    (backward move list, CIGAR op letters backwards, start cell).  A prefix must
    score exactly what the rest of the path leaves, and no prefix ending at a
    cell can beat that cell's value, so the search follows cells whose value
    equals the score still needed."""
    go, ge = c.gap_open + c.gap_extend, c.gap_extend
    out = []

    def hval(i, j):
        if i == 0 or j == 0:
            return 0
        return v.H.get((i, j), 0)  # a diagonal from outside the band reads 0

    def walk(i, j, state, need, moves, ops):
        if len(out) > cap:
            raise TooMany
        if state == "H":
            inside = i > 0 and j > 0 and (i, j) in v.H
            if not inside:
                # the traceback ends at the matrix edge or on leaving the band
                if need == 0:
                    out.append((moves, ops, (i, j)))
                return
            if need == 0:
                out.append((moves + ["stop"], ops, (i, j)))
            s = c.match_score if read[i - 1] == ref[j - 1] else -c.mismatch
            # moves past a zero prefix are listed too (the rule ranks stop first)
            if hval(i - 1, j - 1) == need - s:
                walk(i - 1, j - 1, "H", need - s, moves + ["diag"], ops + ["M"])
            if v.E.get((i, j)) == need:
                walk(i, j, "I", need, moves + ["toI"], ops)
            if v.F.get((i, j)) == need:
                walk(i, j, "D", need, moves + ["toD"], ops)
        elif state == "I":
            if (i - 1, j) in v.H and v.H[(i - 1, j)] - go == need:
                walk(i - 1, j, "H", need + go, moves + ["open"], ops + ["I"])
            if v.E.get((i - 1, j)) is not None and v.E[(i - 1, j)] - ge == need:
                walk(i - 1, j, "I", need + ge, moves + ["ext"], ops + ["I"])
        else:
            if (i, j - 1) in v.H and v.H[(i, j - 1)] - go == need:
                walk(i, j - 1, "H", need + go, moves + ["open"], ops + ["D"])
            if v.F.get((i, j - 1)) is not None and v.F[(i, j - 1)] - ge == need:
                walk(i, j - 1, "D", need + ge, moves + ["ext"], ops + ["D"])

    i, j = v.end
    walk(i, j, "H", v.score, [], [])
    return out


def rank(moves, p_h=P_H, p_gap=P_GAP):
    return [p_h[m] if m in p_h else p_gap[m] for m in moves]


def select(paths, p_h=P_H, p_gap=P_GAP):
    """The rule over whole paths: lexicographically greatest backward move list."""
    return max(paths, key=lambda p: rank(p[0], p_h, p_gap))


def as_result(v, path):
    moves, ops, (si, sj) = path
    runs = []
    for op in reversed(ops):
        if runs and runs[-1][1] == op:
            runs[-1][0] += 1
        else:
            runs.append([1, op])
    code = {"M": 0, "I": 1, "D": 2}
    return dict(read_start=si, ref_start=sj, read_end=v.end[0], ref_end=v.end[1],
                score=v.score, cigar=[(n << 4) | code[o] for n, o in runs])


def brute_force_paths(read, ref, c, v):
    """Every local alignment path of the full matrix ending at v.end, scored by
    replaying its operations (no DP), for the smallest inputs: pins the pruned
    enumeration above.  A path = start cell + operation string."""
    go, ge = c.gap_open + c.gap_extend, c.gap_extend
    ei, ej = v.end
    found = []

    def ext(i, j, ops, score, last):
        if (i, j) == (ei, ej) and ops and ops[-1] == "M":
            found.append((list(ops), score))
        if i < ei and j < ej:
            s = c.match_score if read[i] == ref[j] else -c.mismatch
            ext(i + 1, j + 1, ops + ["M"], score + s, "M")
        if i < ei:
            ext(i + 1, j, ops + ["I"], score - (ge if last == "I" else go), "I")
        if j < ej:
            ext(i, j + 1, ops + ["D"], score - (ge if last == "D" else go), "D")

    for si in range(ei + 1):
        for sj in range(ej + 1):
            ext(si, sj, [], 0, None)
    return found


# Cost sets that make different alignments score the same: gap open 0 (open
# ties extend), a mismatch costing exactly one insertion plus one deletion
# (1, 2, 0, 1), (2, 4, 1, 1), (2, 2, 0, 1), and the reference defaults.
TIE_COSTS = [(1, 2, 0, 1), (2, 4, 1, 1), (1, 4, 0, 1), (1, 4, 6, 1), (1, 1, 0, 1),
             (2, 2, 0, 1), (2, 3, 0, 1)]


def _motif_seq(rng, n):
    """Low-complexity sequence: homopolymer runs and dinucleotide repeats, where
    an indel has several equivalent placements."""
    out = bytearray()
    while len(out) < n:
        k = rng.random()
        if k < 0.35:
            out += bytes([rng.choice(b"ACGT")]) * rng.randrange(2, 5)
        elif k < 0.6:
            out += bytes(rng.choice(b"ACGT") for _ in range(2)) * rng.randrange(2, 4)
        else:
            out.append(rng.choice(b"ACGT"))
    return bytes(out[:n])


def tie_corpus(seed, n, maxL, maxW):
    """(read, window, cfg, band offset): the read is a window slice with one or
    two edits (deletion, duplicated base, substitution) and the band offset
    follows the slice, give or take 3."""
    rng = random.Random(seed)
    cases = []
    for _ in range(n):
        W = rng.randrange(2, maxW + 1)
        ref = _motif_seq(rng, W)
        L = rng.randrange(1, maxL + 1)
        st = rng.randrange(max(1, W - L + 1))
        read = bytearray(ref[st:st + L])
        for _ in range(rng.randrange(1, 3)):
            if not read:
                break
            k = rng.randrange(len(read))
            r = rng.random()
            if r < 0.35:
                del read[k]
            elif r < 0.7:
                read.insert(k, read[k])
            else:
                read[k] = rng.choice(b"ACGT")
        read = bytes(read) or b"A"
        a, b, o, e = rng.choice(TIE_COSTS)
        c = O.cfg(a, b, o, e, rng.choice([1, 2, 3, 8]), rng.choice([0, 0, 3, 100]),
                  rng.choice([0, 5]))
        cases.append((read, ref, c, st + rng.randrange(-3, 4)))
    return cases


def _strip(r):
    return {k: r[k] for k in ("read_start", "ref_start", "read_end", "ref_end", "score",
                              "cigar")}


def test_enumeration_equals_brute_force_on_the_smallest_inputs():
    """The pruned enumeration finds exactly the optimal paths a full enumeration
    of every path (start cell x operation string, scored by replay) finds."""
    checked = 0
    for read, ref, c, _off in tie_corpus(1, 400, 4, 5):
        c.clip_penalty = 0
        v = values_full_window(read, ref, c)
        if v.score <= 0:
            continue
        everything = brute_force_paths(read, ref, c, v)
        assert max(s for _, s in everything) == v.score
        opt = sorted("".join(o) for o, s in everything if s == v.score)
        got = sorted("".join(reversed(p[1])) for p in co_optimal_paths(read, ref, c, v))
        # the enumeration ends at the matrix edge; brute force also starts
        # paths inside the matrix, each of which is a suffix of a listed path
        # or listed itself (a stop at a zero prefix)
        assert set(got) <= set(opt)
        for o in opt:
            assert any(g.endswith(o) for g in got), (read, ref, o, got)
        checked += 1
    assert checked > 150


def test_full_window_ties_follow_the_path_rule():
    """Semantics A (one job per lane group) on tie-heavy inputs."""
    n_ties = 0
    for read, ref, c, _off in tie_corpus(2, 2500, 10, 12):
        v = values_full_window(read, ref, c)
        try:
            paths = co_optimal_paths(read, ref, c, v)
        except TooMany:
            continue
        want = as_result(v, select(paths))
        got = O.sw_group([O.Job(read, ref, 0, 0, 0, 0, len(read), ABORT_NONE)], c, bits=16)[0]
        assert _strip(got) == want, (read, ref, c.gap_open, len(paths))
        n_ties += len(paths) > 1
    assert n_ties > 250  # hundreds of inputs with several co-optimal alignments


@pytest.mark.parametrize("adaptive", [False, True])
def test_banded_ties_follow_the_path_rule(adaptive):
    """Semantics B, fixed and adaptive bands (narrow bands, odd offsets, x-drop)."""
    n_ties = 0
    for read, ref, c, off in tie_corpus(3 + adaptive, 2500, 10, 12):
        v = values_banded(read, ref, off, c, adaptive=adaptive, max_drift=2)
        try:
            paths = co_optimal_paths(read, ref, c, v)
        except TooMany:
            continue
        want = as_result(v, select(paths))
        got = O.banded_sw(read, ref, off, c, ABORT_NONE,
                          O.knobs(adaptive=adaptive, max_drift=2))
        assert _strip(got) == want, (read, ref, off, c.bandwidth, c.gap_open)
        assert got["early_abort"] == v.early_abort
        n_ties += len(paths) > 1
    assert n_ties > 250


MUTANTS = {
    "E/F priority swapped": (dict(P_H, toI=0, toD=1), P_GAP),
    "extend preferred on ties (>= for the extend bit)": (P_H, {"open": 0, "ext": 1}),
    "gaps before matches": (dict(P_H, diag=0, toI=2, toD=1), P_GAP),
    "no stop at a zero prefix": (dict(P_H, stop=-1), P_GAP),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_each_plausible_tie_mistake_is_caught(name):
    """A selection rule with one plausible mistake disagrees with the oracle on
    this corpus (so the tests above would catch that mistake in the oracle)."""
    p_h, p_gap = MUTANTS[name]
    for read, ref, c, off in tie_corpus(4, 2500, 10, 12):
        v = values_banded(read, ref, off, c, adaptive=True, max_drift=2)
        try:
            paths = co_optimal_paths(read, ref, c, v)
        except TooMany:
            continue
        got = O.banded_sw(read, ref, off, c, ABORT_NONE, O.knobs(adaptive=True, max_drift=2))
        if as_result(v, select(paths, p_h, p_gap)) != _strip(got):
            return
    pytest.fail(f"mutation '{name}' is not distinguished by the corpus")


# ---- exact CIGARs of the reference's own gapped test inputs, derived by hand

def _cfg(bw, clip=5):
    return O.cfg(1, 4, 6, 1, bw, 1000, clip)


def _left_aligned_gap(longer: bytes, start: int, n: int) -> int:
    """Leftmost start of an n-base gap in `longer` that leaves the same shorter
    sequence: slide left while the base before the gap equals its last base."""
    while start > 0 and longer[start - 1] == longer[start + n - 1]:
        start -= 1
    return start


def test_two_deletions_exact_cigar():
    """mod.rs:2368-2417: read = ref[0:30] + ref[33:60] + ref[63:120].  Both 3-bp
    deletions survive in the adaptive band; reading back from the end the walk
    prefers match columns, so each deletion sits at its leftmost equivalent
    position; every read base aligns (the first-row fix), so the CIGAR is
    {d1}M3D{d2-d1-3}M3D{rest}M with score 114 - 2 * (6 + 3)."""
    ref = gen_ref(120)
    read = ref[0:30] + ref[33:60] + ref[63:120]
    d1 = _left_aligned_gap(ref, 30, 3)
    d2 = _left_aligned_gap(ref, 60, 3)
    m1, m2 = d1, (d2 - 3) - d1
    want = f"{m1}M3D{m2}M3D{114 - m1 - m2}M"
    r = O.banded_sw(read, ref, 0, _cfg(4), ABORT_NONE, O.knobs(adaptive=True))
    assert O.cigar_str(r["cigar"]) == want
    assert (r["read_start"], r["read_end"], r["ref_start"], r["ref_end"], r["score"]) == \
        (0, 114, 0, 120, 96)
    assert want == "30M3D27M3D57M"  # gen_ref's bases allow no slide at either junction


def test_five_base_insertion_exact_cigar():
    """mod.rs:2592-2635: read = ref[:40] + TTGCA + ref[40:]; the insertion is
    left-aligned in the read and costs 6 + 5."""
    ref = gen_ref(90)
    read = ref[:40] + b"TTGCA" + ref[40:]
    p = _left_aligned_gap(read, 40, 5)
    r = O.banded_sw(read, ref, 0, _cfg(20), ABORT_NONE, O.knobs())
    assert O.cigar_str(r["cigar"]) == f"{p}M5I{95 - 5 - p}M"
    assert (r["read_start"], r["read_end"], r["ref_start"], r["ref_end"], r["score"]) == \
        (0, 95, 0, 90, 90 - 11)


def _end_indel():
    base = bytes(b"CGTA"[i % 4] for i in range(150))
    return b"TT" + base + b"TT", base[:145] + b"G" + base[145:149]


def test_end_insertion_exact_cigars():
    """tests/alignment.rs:35-122.  The read is base[:145] + G + base[145:149]
    (tail GGTAC) against TT + base + TT.  Down the seed diagonal read[145] = G
    still matches base[145] = G (146 matches, score 146); read[146:150] = GTAC
    then mismatches TACG.  Aligning the whole read needs one inserted G: 149
    matches - 7 = 142.  With clip 5 the last row wins (142 + 5 > 146); reading
    back from the end, GTAC matches base[145:149] and the walk prefers the match
    column at read[146] = G / base[145] = G, so the insertion is read[145]:
    145M1I4M from reference 2.  With clip 0 the best cell (146 at read 146)
    stands: 146M and four soft-clipped bases."""
    ref, read = _end_indel()
    span = (2, 147, 0, 145)
    from tests.test_oracle_sw import align_chain
    a5 = align_chain(read, ref, span, O.cfg(1, 4, 6, 1, 50, 50, 5))
    assert (a5["cigar"], a5["ref_start"], a5["ref_end"], a5["score"]) == \
        ("145M1I4M", 2, 151, 142)
    a0 = align_chain(read, ref, span, O.cfg(1, 4, 6, 1, 50, 50, 0))
    assert (a0["cigar"], a0["ref_start"], a0["ref_end"], a0["score"]) == \
        ("146M4S", 2, 148, 146)


def test_full_match_exact_cigar():
    """tests/alignment.rs:4-33: ACGTACGT inside TTACGTACGTGG at 2 (match 2)."""
    from tests.test_oracle_sw import align_chain
    a = align_chain(b"ACGTACGT", b"TTACGTACGTGG", (2, 10, 0, 8), O.cfg(2, 4, 6, 1, 50, 40, 5))
    assert (a["cigar"], a["ref_start"], a["ref_end"], a["score"], a["nm"]) == \
        ("8M", 2, 10, 16, 0)
