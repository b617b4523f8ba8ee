"""The gpu-server session front end (kira-ls-aligner_amd/gpu_server.py) on
the CPU: the protocol of src/cuda/server.rs:43-175 (keys, aliases, error
texts, blank-line submission, quit) with the job stubbed out, and the input
readers (reference FASTA normalization, read ingestion, chain lines)."""
import io
import os
import importlib.util

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod():
    spec = importlib.util.spec_from_file_location(
        "kira_gpu_server", os.path.join(ROOT, "kira-ls-aligner_amd", "gpu_server.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


G = _mod()


def test_keys_aliases_and_errors():
    s = G.Session()
    assert s.is_empty()
    for k, v in [("ref", "a.fa"), ("reads", "r1.fq"), ("reads", "r2.fq"), ("idx", "x"),
                 ("o", "out.sam"), ("t", "+4"), ("K", "1000"), ("R", "@RG\\tID:x")]:
        s.set(k, v)
    assert (s.reference, s.reads, s.index, s.output, s.threads, s.batch_bases, s.read_group) == \
        ("a.fa", ["r1.fq", "r2.fq"], "x", "out.sam", 4, 1000, "@RG\\tID:x")
    msgs = {}
    for v in ["", "abc", "-1", "+", "18446744073709551616"]:
        with pytest.raises(G.SessionError) as e:
            G.Session().set("threads", v)
        msgs[v] = str(e.value)
    assert msgs[""] == "threads must be a positive integer: cannot parse integer from empty string"
    assert msgs["abc"] == "threads must be a positive integer: invalid digit found in string"
    assert msgs["-1"] == msgs["+"] == msgs["abc"]
    assert msgs["18446744073709551616"].endswith("number too large to fit in target type")
    assert G.Session().set("batch", "18446744073709551615") is None
    with pytest.raises(G.SessionError) as e:
        G.Session().set("reffy", "x")
    assert str(e.value) == ("unknown key `reffy` (use ref / reads / index / output / threads / "
                            "batch / read-group)")
    # output / threads / batch alone leave a session empty (server.rs:111-116)
    s = G.Session()
    s.set("threads", "3")
    s.set("output", "o")
    assert s.is_empty()


class StubServer(G.Server):
    def __init__(self, out, log):
        super().__init__(pkg=type("P", (), {"HipError": RuntimeError}), stdout=out, log=log)
        self.jobs = []

    def run_job(self, s):
        if s.reference is None:
            raise G.SessionError("`ref <path>` is required before submission")
        if not s.reads:
            raise G.SessionError("`reads <path>` is required")
        self.jobs.append((s.reference, list(s.reads)))


def test_session_loop_markers():
    out, log = io.StringIO(), io.StringIO()
    srv = StubServer(out, log)
    script = ("\n\n"                     # blank lines with nothing entered: keep waiting
              "bogus 1\n"                # unknown key: error, session goes on
              "ref  /x/ref.fa  \n"       # key, then the rest of the line trimmed
              "reads r.fq\n"
              "\n"                       # submit
              "reads only.fq\n"
              "\n"                       # submit without ref: error
              "threads 4\n\n"            # not a job: still waiting
              "QUIT\n"
              "ref never.fa\n\n")
    assert srv.serve(io.StringIO(script)) == 0
    assert srv.jobs == [("/x/ref.fa", ["r.fq"])]
    lines = out.getvalue().splitlines()
    assert lines[0].startswith("[KIRA_GPU_ERR] unknown key `bogus`")
    assert lines[1].startswith("[KIRA_GPU_OK] elapsed=") and lines[1].endswith("s")
    assert lines[2] == "[KIRA_GPU_ERR] `ref <path>` is required before submission"
    assert len(lines) == 3
    assert log.getvalue().count("[KIRA_GPU_READY]") == 2
    assert log.getvalue().endswith("[KIRA_GPU] shutting down.\n")


def test_input_readers(tmp_path):
    fa = tmp_path / "r.fa"
    fa.write_bytes(b">chr1 desc\nacgtXX\nRYn\n>chr2\nTTTT\n")
    names, seqs = G.read_fasta(str(fa))
    assert names == ["chr1", "chr2"] and seqs == [b"ACGTNNRYN", b"TTTT"]
    fa.write_bytes(b">a\nA\n>a\nC\n")
    with pytest.raises(G.SessionError, match="duplicate reference contig name"):
        G.read_fasta(str(fa))
    # ingest_read: upper-case, IUPAC kept as the original, aligned as N
    assert G.ingest(b"acgtu") == (b"ACGTN", None)
    assert G.ingest(b"acRy") == (b"ACNN", b"ACRY")
    fq = tmp_path / "r.fq"
    fq.write_bytes(b"@q1 x\nACGT\n+\nIIII\n@q2\nAC\n+\n##\n")
    ids, seqs, origs, quals = G.read_fastq(str(fq))
    assert (ids, seqs, origs, quals) == (["q1", "q2"], [b"ACGT", b"AC"], [None, None],
                                         [b"IIII", b"##"])
    ch = tmp_path / "c.tsv"
    ch.write_text("q1\tchr2\t0\t4\t0\t4\t-\t4\n# comment\nq1\tchr1\t1\t3\t1\t3\t+\t2\n")
    assert G.read_chains(str(ch), ids, ["chr1", "chr2"]) == [
        [(1, 0, 4, 0, 4, 1, 4), (0, 1, 3, 1, 3, 0, 2)], []]
    ch.write_text("q9\tchr1\t0\t1\t0\t1\t+\t1\n")
    with pytest.raises(G.SessionError, match="malformed chain line"):
        G.read_chains(str(ch), ids, ["chr1"])
