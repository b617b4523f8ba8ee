"""A gpu-server session end to end on the GPU: FASTA + FASTQ + chains in,
SAM file out, compared byte for byte with the CPU restatement run on the
same read batches."""
import io
import os
import importlib.util

import numpy as np
import pytest

from oracle import sam as OS
from oracle import stage4 as S
from tests.kira_pkg import load
from tests.stage4_data import make_batch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod():
    spec = importlib.util.spec_from_file_location(
        "kira_gpu_server", os.path.join(ROOT, "kira-ls-aligner_amd", "gpu_server.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_session_writes_the_oracle_sam(tmp_path):
    G = _mod()
    pkg = load()
    assert pkg.is_available(), "no GPU visible"
    b = make_batch(31, 900)
    names = [f"chr{k + 1}" for k in range(len(b["contigs"]))]
    rng = np.random.default_rng(2)
    with open(tmp_path / "ref.fa", "wb") as f:
        for n, c in zip(names, b["contigs"]):
            f.write(b">" + n.encode() + b" synthetic\n")
            for k in range(0, len(c), 70):
                f.write(c[k:k + 70] + b"\n")
    ids, fq_reads, quals = [], [], []
    for k, r in enumerate(b["reads"]):
        r = bytearray(r)
        if k % 50 == 7 and b"N" in r:
            r[r.index(b"N")] = ord("R")  # an IUPAC letter: kept in SEQ, aligned as N
        ids.append(f"read{k}")
        fq_reads.append(bytes(r))
        quals.append(bytes(rng.integers(33, 74, len(r)).astype(np.uint8)))
    with open(tmp_path / "reads.fq", "wb") as f:
        for i, r, q in zip(ids, fq_reads, quals):
            f.write(b"@" + i.encode() + b" c\n" + r + b"\n+\n" + q + b"\n")
    with open(tmp_path / "chains.tsv", "w") as f:
        for i, cl in zip(ids, b["chains"]):
            for c in cl:
                f.write(f"{i}\t{names[c[0]]}\t{c[1]}\t{c[2]}\t{c[3]}\t{c[4]}\t"
                        f"{'-' if c[5] else '+'}\t{c[6]}\n")
    out, log = io.StringIO(), io.StringIO()
    srv = G.Server(pkg, stdout=out, log=log)
    srv.start()
    batch = 40_000
    script = (f"ref {tmp_path}/ref.fa\nreads {tmp_path}/reads.fq\n"
              f"chains {tmp_path}/chains.tsv\noutput {tmp_path}/out.sam\nbatch {batch}\n"
              "read-group @RG\\tID:grpA\\tSM:s1\n\nquit\n")
    print("[session] running", flush=True)
    srv.serve(io.StringIO(script))
    assert out.getvalue().startswith("[KIRA_GPU_OK] elapsed="), out.getvalue()
    got = (tmp_path / "out.sam").read_bytes()

    # the same job through the CPU restatement, batch by batch
    fa_names, contigs = G.read_fasta(str(tmp_path / "ref.fa"))
    rid, seqs, origs, qs = G.read_fastq(str(tmp_path / "reads.fq"))
    chains = G.read_chains(str(tmp_path / "chains.tsv"), rid, fa_names)
    rg = "@RG\tID:grpA\tSM:s1"
    want = ["@HD\tVN:1.6\tSO:unsorted"] + [f"@SQ\tSN:{n}\tLN:{len(c)}"
                                           for n, c in zip(fa_names, contigs)] + [rg]
    want = ("\n".join(want) + "\n").encode()
    k = 0
    while k < len(seqs):
        e, bases = k, 0
        while e < len(seqs) and (e == k or bases + len(seqs[e]) <= batch):
            bases += len(seqs[e])
            e += 1
        recs, _ = S.run(contigs, seqs[k:e], chains[k:e], S.default_cfg(), [0] * (e - k))
        OS.normalize_alignments(contigs, seqs[k:e], recs, origs[k:e])
        out_seqs = [o if o is not None else s for o, s in zip(origs[k:e], seqs[k:e])]
        want += OS.sam_text(rid[k:e], out_seqs, qs[k:e], recs, fa_names,
                            OS.default_cfg(read_group=rg))
        k = e
    assert sum(o is not None for o in origs) > 0
    assert got == want
