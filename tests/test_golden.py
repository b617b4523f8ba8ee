"""Committed golden fixtures (tests/golden/*.json, written by make_golden.py):
the oracle must reproduce them on the CPU and the HIP path on the GPU."""
import json
import os

import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def sw_expect_from_oracle(c):
    r = O.banded_sw(c["read"].encode(), c["ref"].encode(), c["offset"], O.cfg(*c["cfg"]),
                    c["abort_score"], O.knobs(adaptive=c["adaptive"]))
    return {k: r[k] for k in c["expect"]}


def test_oracle_reproduces_sieve_fixture():
    f = load("sieve_cuda_corpus.json")
    jobs = [(j["read"].encode(), j["text"].encode(), j["max_mismatches"]) for j in f["jobs"]]
    got = O.backend_run_batch(jobs)
    assert got == [(j["shift"], j["mismatches"], j["second"]) for j in f["jobs"]]


def test_oracle_reproduces_sw_fixture():
    for c in load("banded_sw_cases.json")["cases"]:
        assert sw_expect_from_oracle(c) == c["expect"]


def test_oracle_reproduces_prefilter_fixture():
    f = load("prefilter_cases.json")
    c = O.cfg(*f["cfg"])
    for case in f["cases"]:
        r = O.prefilter_chain(case["read"].encode(), case["ref"].encode(), tuple(case["span"]), c,
                              **case["args"])
        assert r == case["expect"]


@pytest.mark.gpu
def test_gpu_reproduces_fixtures():
    from tests.kira_pkg import load as load_pkg
    pkg = load_pkg()
    ctx = pkg.Context(0)
    f = load("sieve_cuda_corpus.json")
    jobs = []
    for j in f["jobs"]:
        rb, _ = O.pack(j["read"].encode())
        tb, _ = O.pack(j["text"].encode())
        jobs.append(pkg.HipJob(rb, len(j["read"]), O.pre_shift(tb), len(j["text"]),
                               j["max_mismatches"]))
    got = pkg.HipBackend().run_batch(jobs)
    assert [(r.shift, r.mismatches, r.second_mismatches) for r in got] == \
        [(j["shift"], j["mismatches"], j["second"]) for j in f["jobs"]]
    for c in load("banded_sw_cases.json")["cases"]:
        cfg = pkg.sw_cfg(*c["cfg"])
        r, = pkg.banded_sw_batch(ctx, [(c["read"].encode(), c["ref"].encode(), c["offset"],
                                        c["abort_score"])], cfg, pkg.sw_knobs(c["adaptive"]))
        assert {k: r[k] for k in c["expect"]} == c["expect"]
    f = load("prefilter_cases.json")
    chains = [(c["read"].encode(), c["ref"].encode(), tuple(c["span"]),
               c["args"].get("is_top1", True)) for c in f["cases"]]
    params = [pkg.prefilter_params(**{k: v for k, v in c["args"].items() if k != "is_top1"})
              for c in f["cases"]]
    for ch, p, c in zip(chains, params, f["cases"]):
        r, = pkg.prefilter_batch(ctx, [ch], pkg.sw_cfg(*f["cfg"]), p)
        want = dict(c["expect"])
        if "alignment" in want:
            a = dict(want["alignment"])
            a.pop("md")
            want["alignment"] = a
            r["alignment"]["cigar"] = O.cigar_str(r["alignment"]["cigar"])
        assert r == want
    ctx.close()
