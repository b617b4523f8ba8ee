"""`gpu-server` session front end (src/cuda/server.rs:15-193) over libkira_hip.

The session protocol is the reference's: one `key value` line per job field
(`ref`, `reads`, `index`, `output`, `threads`, `batch`, `read-group`, with
their aliases), a blank line submits the job, `quit` / `exit` (any case) ends
the session.  Each job answers `[KIRA_GPU_OK] elapsed=S.mmms` or
`[KIRA_GPU_ERR] <error>` on stdout; `[KIRA_GPU_READY]` and the other
`[KIRA_GPU]` lines go to stderr.

A reference job runs the whole `mem` pipeline.  This library holds stage 4
and the record output only, so a job here also names the chains stage 3 hands
to stage 4 (`chains <path>`, one tab-separated line per chain:
`read_id contig ref_start ref_end read_start read_end strand(+/-) score`,
0-based half-open, each read's chains best first) and runs, per batch of
reads of at most `batch` bases: the stage-4 driver on the GPU
(kira_stage4_run, `mem` defaults), indel left-normalization and single-end
SAM lines (header @HD / @SQ / @RG, then every record, or an unmapped line for
a read with none).  `threads` is accepted and not used (host worker threads
are the library's), `index` is accepted and not used (no seeding here).
Single-end only: a second `reads` file fails the job.  The genome of a
reference stays resident across jobs that name the same path.

Run: python kira-ls-aligner_amd/gpu_server.py [-t THREADS] [-K BATCH_BASES]
"""
from __future__ import annotations

import os
import sys
import time


class SessionError(Exception):
    pass


def _parse_usize(value: str, what: str) -> int:
    """`value.parse::<usize>()` with the context message of Session::set
    (server.rs:135-147); the cause text is Rust's ParseIntError display."""
    body = value[1:] if value.startswith("+") else value
    if value == "":
        cause = "cannot parse integer from empty string"
    elif body == "" or not body.isascii() or not body.isdigit():
        cause = "invalid digit found in string"
    elif int(body) > 2 ** 64 - 1:
        cause = "number too large to fit in target type"
    else:
        return int(body)
    raise SessionError(f"{what} must be a positive integer: {cause}")


class Session:
    """A pending job accumulated across lines (server.rs:100-175)."""

    def __init__(self):
        self.reference = None
        self.reads = []
        self.index = None
        self.output = None
        self.threads = None
        self.batch_bases = None
        self.read_group = None
        self.chains = None

    def is_empty(self) -> bool:
        return (self.reference is None and not self.reads and self.index is None
                and self.read_group is None and self.chains is None)

    def set(self, key: str, value: str):
        if key in ("ref", "reference"):
            self.reference = value
        elif key == "reads":
            self.reads.append(value)
        elif key in ("index", "idx", "i"):
            self.index = value
        elif key in ("output", "out", "o"):
            self.output = value
        elif key in ("threads", "t"):
            self.threads = _parse_usize(value, "threads")
        elif key in ("batch", "K"):
            self.batch_bases = _parse_usize(value, "batch")
        elif key in ("read-group", "rg", "R"):
            self.read_group = value
        elif key == "chains":
            self.chains = value
        else:
            raise SessionError(f"unknown key `{key}` (use ref / reads / index / output / "
                               "threads / batch / read-group)")


# ------------------------------------------------------------------ inputs
IUPAC = set(b"RYSWKMBDHV")


def read_fasta(path):
    """visit_reference (src/io/mod.rs:27-59): name = header up to the first
    whitespace, bases upper-cased, anything not IUPAC becomes N."""
    keep = set(b"ACGTRYSWKMBDHVN")
    names, seqs, seen = [], [], set()
    name, chunks = None, []

    def flush():
        if name is None:
            return
        seq = bytes(b if b in keep else 78 for b in b"".join(chunks).upper())
        names.append(name)
        seqs.append(seq)

    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                flush()
                name = line[1:].split()[0].decode() if line[1:].split() else ""
                if not name:
                    raise SessionError("empty reference contig name")
                if name in seen:
                    raise SessionError(f'duplicate reference contig name "{name}"')
                seen.add(name)
                chunks = []
            elif name is not None:
                chunks.append(line.strip())
    flush()
    return names, seqs


def ingest(seq: bytes):
    """ingest_read (src/io/mod.rs:737-776): upper-case, keep ACGTN and IUPAC,
    other bytes become N; a read with IUPAC letters keeps them as its
    original sequence and aligns with them as N."""
    up = bytes(b if b in b"ACGTN" or b in IUPAC else 78 for b in seq.upper())
    if any(b in IUPAC for b in up):
        return bytes(b if b in b"ACGT" else 78 for b in up), up
    return up, None


def read_fastq(path):
    ids, seqs, origs, quals = [], [], [], []
    with open(path, "rb") as f:
        while True:
            h = f.readline()
            if not h:
                break
            if not h.strip():
                continue
            s = f.readline().rstrip(b"\r\n")
            f.readline()
            q = f.readline().rstrip(b"\r\n")
            if not h.startswith(b"@"):
                raise SessionError(f"{path}: malformed FASTQ record")
            parts = h[1:].split()
            ids.append(parts[0].decode() if parts else "")
            seq, orig = ingest(s)
            seqs.append(seq)
            origs.append(orig)
            quals.append(q if len(q) else None)
    return ids, seqs, origs, quals


def read_chains(path, ids, names):
    """Chains per read, in file order, as stage 3 hands them to stage 4."""
    by_id = {r: k for k, r in enumerate(ids)}
    contig = {n: k for k, n in enumerate(names)}
    chains = [[] for _ in ids]
    with open(path) as f:
        for ln, line in enumerate(f, 1):
            if not line.strip() or line.startswith("#"):
                continue
            c = line.rstrip("\n").split("\t")
            if len(c) != 8 or c[0] not in by_id or c[1] not in contig or c[6] not in ("+", "-"):
                raise SessionError(f"{path}:{ln}: malformed chain line")
            chains[by_id[c[0]]].append((contig[c[1]], int(c[2]), int(c[3]), int(c[4]),
                                        int(c[5]), 1 if c[6] == "-" else 0, int(c[7])))
    return chains


# ------------------------------------------------------------------ server
class Server:
    def __init__(self, pkg, default_threads=8, default_batch_bases=10_000_000, device=0,
                 stdout=None, log=None):
        self.pkg = pkg
        self.default_threads = default_threads
        self.default_batch_bases = default_batch_bases
        self.out = stdout or sys.stdout
        self.log = log or sys.stderr
        self.ctx = None
        self.device = device
        self.genome = None  # (path, names, seqs, offsets, lengths, Genome)

    def info(self, msg):
        print(msg, file=self.log, flush=True)

    def start(self):
        """CudaBackend::new + warmup_kernel + dispatcher::start (server.rs:15-32)."""
        self.info("[KIRA_GPU] Initializing HIP backend...")
        t0 = time.perf_counter()
        pkg = self.pkg
        self.ctx = pkg.Context(self.device)
        try:
            # one 4-bp read matching itself (server.rs:178-192)
            phase = bytes([0x1B, 0])
            pkg.HipBackend(self.device).run_batch(
                [pkg.HipJob(bytes([0x1B]), 4, [phase] * 4, 4, 0)])
        except Exception as e:  # noqa: BLE001 - the reference logs and continues
            self.info(f"[KIRA_GPU] main-thread warmup launch failed: {e}")
        dt = time.perf_counter() - t0
        self.info(f"[KIRA_GPU] HIP ready ({int(dt)}.{int(dt * 1000) % 1000:03d} s warmup).")
        self.info("[KIRA_GPU] Enter job parameters (one per line: `ref`, `reads`, `chains`, "
                  "`index`, `output`, `threads`, `batch`, `read-group`). Blank line submits. "
                  "`quit` to exit.")
        self.info("[KIRA_GPU_READY]")

    def _genome(self, path):
        if self.genome is None or self.genome[0] != path:
            if self.genome is not None:
                self.genome[5].close()
            names, seqs = read_fasta(path)
            if not names:
                raise SessionError(f"{path}: no reference sequences")
            import numpy as np
            lens = np.array([len(s) for s in seqs], np.int64)
            offs = np.zeros(len(seqs), np.int64)
            offs[1:] = np.cumsum(lens[:-1])
            self.genome = (path, names, seqs, offs, lens,
                           self.pkg.Genome(self.ctx, b"".join(seqs)))
        return self.genome

    def run_job(self, s: Session):
        if s.reference is None:
            raise SessionError("`ref <path>` is required before submission")
        if not s.reads:
            raise SessionError("`reads <path>` is required")
        if len(s.reads) > 1:
            raise SessionError("paired-end input is not supported: stage 4 here is single-end")
        if s.chains is None:
            raise SessionError("`chains <path>` is required: this build starts at stage 4 "
                               "(no seeding or chaining)")
        pkg = self.pkg
        _, names, seqs, offs, lens, genome = self._genome(s.reference)
        ids, reads, origs, quals = read_fastq(s.reads[0])
        chains = read_chains(s.chains, ids, names)
        batch = s.batch_bases if s.batch_bases is not None else self.default_batch_bases
        rg = s.read_group.replace("\\t", "\t") if s.read_group is not None else None
        sink = open(s.output, "wb") if s.output else None
        try:
            w = sink.write if sink else (lambda b: (self.out.write(b.decode()), self.out.flush()))
            head = ["@HD\tVN:1.6\tSO:unsorted"]
            head += [f"@SQ\tSN:{n}\tLN:{len(q)}" for n, q in zip(names, seqs)]
            if rg is not None:
                head.append(rg if rg.startswith("@RG") else "@RG\t" + rg)
            w(("\n".join(head) + "\n").encode())
            cfg = pkg.stage4_cfg()
            scfg = pkg.sam_cfg(read_group=rg)
            b = 0
            while b < len(reads):
                e, bases = b, 0
                while e < len(reads) and (e == b or bases + len(reads[e]) <= batch):
                    bases += len(reads[e])
                    e += 1
                x = pkg.Stage4Inputs(offs, lens, reads[b:e], chains[b:e])
                res = pkg.stage4_run(self.ctx, genome, x, cfg=cfg)
                try:
                    res.left_normalize(genome, x, origs[b:e])
                    out_seqs = [o if o is not None else r for o, r in zip(origs[b:e], reads[b:e])]
                    xo = pkg.Stage4Inputs(offs, lens, out_seqs, [[]] * (e - b))
                    # FASTQ gives every read a QUAL of its length; FASTA none
                    q = quals[b:e]
                    ok = all(qq is not None and len(qq) == len(sq) for qq, sq in zip(q, out_seqs))
                    w(res.sam(names, ids[b:e], xo, q if ok else None, scfg))
                finally:
                    res.close()
                b = e
        finally:
            if sink:
                sink.close()

    def serve(self, stdin=None) -> int:
        """The session loop of run_gpu_server (server.rs:43-95)."""
        stdin = stdin or sys.stdin
        session = Session()
        for line in stdin:
            trimmed = line.strip()
            if trimmed.lower() in ("quit", "exit"):
                self.info("[KIRA_GPU] shutting down.")
                break
            if not trimmed:
                if session.is_empty():
                    continue
                t0 = time.perf_counter()
                try:
                    self.run_job(session)
                    dt = time.perf_counter() - t0
                    print(f"[KIRA_GPU_OK] elapsed={int(dt)}.{int(dt * 1000) % 1000:03d}s",
                          file=self.out, flush=True)
                except (SessionError, OSError, ValueError, self.pkg.HipError) as e:
                    print(f"[KIRA_GPU_ERR] {e}", file=self.out, flush=True)
                session = Session()
                self.info("[KIRA_GPU_READY]")
                continue
            parts = trimmed.split(None, 1)
            key = parts[0] if parts else ""
            value = parts[1].strip() if len(parts) > 1 else ""
            try:
                session.set(key, value)
            except SessionError as e:
                print(f"[KIRA_GPU_ERR] {e}", file=self.out, flush=True)
        if self.genome is not None:
            self.genome[5].close()
        return 0


def _load_pkg():
    import importlib.util
    name = "kira_ls_aligner_amd"
    if name in sys.modules:
        return sys.modules[name]
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(name, os.path.join(here, "__init__.py"),
                                                  submodule_search_locations=[here])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="kira gpu-server session front end")
    ap.add_argument("-t", "--threads", type=int, default=8)
    ap.add_argument("-K", "--batch", type=int, default=10_000_000)
    a = ap.parse_args(argv)
    srv = Server(_load_pkg(), a.threads, a.batch)
    srv.start()
    return srv.serve()


if __name__ == "__main__":
    sys.exit(main())
