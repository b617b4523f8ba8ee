"""GPU parity of the ungapped prefilter (prefilter_chain) vs the CPU oracle."""
import random

import pytest

from oracle import oracle as O
from tests.kira_pkg import load

pytestmark = pytest.mark.gpu
I32_MAX = 2 ** 31 - 1


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


PARAM_SETS = [
    dict(),
    dict(tail_ext=False),
    dict(accept_span_slack=15, accept_max_mismatches=5),
    dict(ungap_min_id=9800, accept_max_mismatches=4),
    dict(accept_min_identity=95.0, accept_require_score_margin=3),
    dict(accept_only_top1=False, long_read_threshold=1000, accept_max_mismatches=20),
    dict(multi_alignments_enabled=True),
    dict(accept_enable=False),
    dict(ungap_gap_check=True, accept_max_mismatches=5, accept_min_identity=96.0),
]


def oracle_one(read, ref, span, c, top1, margin, prm):
    kw = dict(prm)
    return O.prefilter_chain(read, ref, span, c, is_top1=top1, score_margin=margin, **kw)


def strip_md(d):
    d = dict(d)
    if "alignment" in d:
        a = dict(d["alignment"])
        a.pop("md")
        d["alignment"] = a
    return d


def gpu_view(pkg, d):
    d = dict(d)
    if "alignment" in d:
        a = dict(d["alignment"])
        a["cigar"] = O.cigar_str(a["cigar"])
        d["alignment"] = a
    return d


def random_chain(rng):
    L = rng.choice([0, 1, 5, 31, 32, 33, 64, 100, 150, 151, 250, 300, 301, 450])
    L = max(0, L + rng.randint(-2, 2)) if L else 0
    flank = rng.randint(0, 120)
    ref = bytearray(rng.choice(b"ACGT") for _ in range(L + 2 * flank + rng.randint(0, 50)))
    off = rng.randint(0, max(0, len(ref) - L))
    read = bytearray(ref[off:off + L])
    kind = rng.random()
    for _ in range(rng.choice([0, 0, 1, 2, 3, 5, 40]) if L else 0):
        p = rng.randrange(L)
        read[p] = rng.choice(b"ACGTNacgtR")
    if kind < 0.1 and L:  # lowercase / ambiguity in the reference too
        for _ in range(3):
            p = rng.randrange(len(ref))
            ref[p] = rng.choice(b"acgtN")
    if L == 0:
        span = (0, 0, 0, 0)
    else:
        a = rng.randint(0, L - 1)
        b = rng.randint(a, L)
        d = 0
        if rng.random() < 0.1:
            d = rng.choice([-1, 1])  # indel-suspect anchor
        rs, re_ = off + a, off + b + d
        re_ = max(rs, min(re_, len(ref)))
        if rng.random() < 0.03:
            span = (rs, re_, b, a)  # non-monotonic
        else:
            span = (rs, re_, a, b)
    return bytes(read), bytes(ref), span


def check_batch(pkg, ctx, chains, prm, c):
    got = pkg.prefilter_batch(ctx, chains, pkg.sw_cfg(*c), pkg.prefilter_params(**prm))
    oc = O.cfg(*c)
    for k, (ch, g) in enumerate(zip(chains, got)):
        want = strip_md(oracle_one(ch[0], ch[1], ch[2], oc, ch[3], ch[4], prm))
        assert gpu_view(pkg, g) == want, (k, prm, ch[2], len(ch[0]))


def test_reference_cases(pkg, ctx):
    base = b"ACGT" * 37 + b"AC"
    one = bytearray(base)
    one[10] = ord("T")
    cases = [
        (base, bytes(one), (0, 150, 0, 150), True, I32_MAX),            # accept
        (b"A" * 400, b"T" * 400, (0, 44, 0, 44), True, I32_MAX),        # span_too_small
        (b"A" * 400, b"T" * 400, (0, 400, 0, 400), False, I32_MAX),     # not_top1
        (b"A" * 150, b"T" * 150, (0, 150, 0, 150), True, I32_MAX),      # mism_too_high
        (b"A" * 400, b"T" * 400, (0, 400, 0, 400), True, I32_MAX),      # not_short_preset
    ]
    c = (1, 4, 6, 1, 50, 50, 5)
    got = pkg.prefilter_batch(ctx, cases, pkg.sw_cfg(*c), pkg.prefilter_params())
    assert [(g["result"], g["reason"]) for g in got] == [
        ("accept", "accept"), ("reject", "span_too_small"), ("fallback", "not_top1"),
        ("fallback", "mism_too_high"), ("fallback", "not_short_preset")]
    check_batch(pkg, ctx, cases, {}, c)


@pytest.mark.parametrize("pi", range(len(PARAM_SETS)))
def test_random_chains(pkg, ctx, pi):
    rng = random.Random(100 + pi)
    chains = []
    for _ in range(600):
        read, ref, span = random_chain(rng)
        chains.append((read, ref, span, rng.random() < 0.8,
                       rng.choice([I32_MAX, 0, 2, 3, 5, -4])))
    c = rng.choice([(1, 4, 6, 1, 50, 50, 5), (1, 4, 6, 1, 32, 100, 5), (2, 3, 5, 2, 32, 20, 5)])
    print(f"prefilter random params={PARAM_SETS[pi]} cfg={c}", flush=True)
    check_batch(pkg, ctx, chains, PARAM_SETS[pi], c)


def test_rejects_span_outside_read(pkg, ctx):
    with pytest.raises(pkg.HipError) as e:
        pkg.prefilter_batch(ctx, [(b"ACGT", b"ACGTACGT", (0, 9, 0, 9))], pkg.sw_cfg())
    assert e.value.kind == "InvalidArgument"


def test_empty_batch(pkg, ctx):
    assert pkg.prefilter_batch(ctx, [], pkg.sw_cfg()) == []


@pytest.mark.parametrize("cfgv", [(1, 4, 6, 1, 32, 100, 5), (1, 6, 6, 1, 50, 100, 5),
                                  (2, 3, 5, 2, 16, 40, 5)])
def test_gap_check_gate(pkg, ctx, cfgv):
    """KIRA_UNGAP_GAP_CHECK on the device (WFA score-only probe per accepted
    chain with mismatches): reads with an indel just before a run that reaches
    the read end, plain substitution reads, and random chains."""
    from tests.test_oracle_prefilter import _end_indel_read
    rng = random.Random(sum(cfgv))
    chains = []
    for k in range(900):
        if k % 3 == 2:
            read, ref, span = random_chain(rng)
            chains.append((read, ref, span, True, I32_MAX))
            continue
        ref = bytes(rng.choice(b"ACGT") for _ in range(400))
        st, L = rng.randrange(100, 200), rng.choice([100, 150])
        if k % 3 == 0:
            read, ref = _end_indel_read(rng, ref, st, L)
        else:
            read = bytearray(ref[st:st + L])
            for _ in range(rng.randrange(1, 3)):
                read[rng.randrange(L)] = rng.choice(b"ACGT")
            read = bytes(read)
        chains.append((read, ref, (st + 20, st + 80, 20, 80), True, I32_MAX))
    prm = dict(ungap_gap_check=True, accept_max_mismatches=5, accept_min_identity=96.0)
    check_batch(pkg, ctx, chains, prm, cfgv)
    got = pkg.prefilter_batch(ctx, chains, pkg.sw_cfg(*cfgv), pkg.prefilter_params(**prm))
    assert sum(g["reason"] == "indel_suspect" for g in got) > 100
