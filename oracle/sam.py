"""CPU restatement of record output (test infrastructure only): indel
left-normalization (src/alignment/normalize.rs:55-249) and single-end SAM
lines (SamFormatter::append_alignment / append_unmapped, src/io/mod.rs:
971-1140) over the record dicts oracle/stage4.py produces.

Pinned by the reference's normalization unit tests (normalize.rs:253-420)
and SAM flag tests (tests/sam_flags.rs) in tests/test_oracle_sam.py.
"""
import re

from oracle import oracle as O
from oracle.stage4 import reverse_complement

OPS = "MIDNS"


def parse_cigar(text):
    return [(int(n), op) for n, op in re.findall(r"(\d+)([MIDNS])", text)]


def cigar_text(ops):
    return "".join(f"{n}{op}" for n, op in ops)


def left_normalize_indels(cigar, ref, ref_start, read):
    """normalize.rs:55-190.  cigar: list of (len, op).  Returns the new list
    or None when nothing moved."""
    if not any(op in "ID" for _, op in cigar):
        return None
    ops = [list(o) for o in cigar]
    rs = ref_start

    def rget(i):
        return ref[i] if 0 <= i < len(ref) else None

    def qget(i):
        return read[i] if 0 <= i < len(read) else None

    changed = False
    i = 0
    while i < len(ops):
        kind = ops[i][1]
        if kind not in "ID":
            i += 1
            continue
        g = ops[i][0]
        bq = sum(n for n, op in ops[:i] if op in "MIS")
        br = sum(n for n, op in ops[:i] if op in "MDN")
        m_run = ops[i - 1][0] if i > 0 and ops[i - 1][1] == "M" else 0
        shift_max = 0
        while (shift_max < m_run and br - shift_max >= 1 and bq - shift_max >= 1):
            r = rget(rs + br - shift_max - 1)
            if (r if r is not None else ord("N")) not in b"ACGT":
                break
            if r != qget(bq - shift_max - 1):
                break
            shift_max += 1
        shift = 0
        while shift < shift_max:
            before = rget(rs + br - shift - 1)
            after = rget(rs + br + g - shift - 1) if kind == "D" else qget(bq + g - shift - 1)
            if before != after:
                break
            shift += 1
        if shift == 0:
            i += 1
            continue
        changed = True
        if shift == m_run:
            ops[i - 1], ops[i] = ops[i], ops[i - 1]
            i -= 1
            if i > 0 and ops[i - 1][1] == kind:
                ops[i - 1][0] += ops[i][0]
                del ops[i]
                i -= 1
            continue
        ops[i - 1][0] -= shift
        ops.insert(i + 1, [shift, "M"])
        i += 1
    if not changed:
        return None
    out = []
    for n, op in ops:
        if out and out[-1][1] == op:
            out[-1] = (out[-1][0] + n, op)
        else:
            out.append((n, op))
    return out


def normalize_alignments(contigs, reads, alns, original=None):
    """normalize_alignments (normalize.rs:197-249) on oracle record dicts, in
    place.  original: per read its IUPAC-bearing original_seq or None; such
    reads get NM / MD from it for every record."""
    original = original or [None] * len(reads)
    for read, orig, recs in zip(reads, original, alns):
        for a in recs:
            oriented = reverse_complement(read) if a["is_rev"] else read
            ref = contigs[a["ref_id"]]
            new = left_normalize_indels(parse_cigar(a["cigar"]), ref, a["ref_start"], oriented)
            if new is not None:
                a["cigar"] = cigar_text(new)
            if new is not None or orig is not None:
                tags = oriented if orig is None else (_rc_iupac(orig) if a["is_rev"] else orig)
                a["nm"], a["md"] = O.nm_md_from_cigar(tags, ref, 0, a["ref_start"], a["cigar"])


IUPAC_COMP = {ord(a): ord(b) for a, b in zip("AaCcGgTUtuRYSWKMBDHV", "TtGgCcAAaaYRSWMKVHDB")}


def _rc_iupac(s: bytes) -> bytes:
    """append_reverse_complement (io/mod.rs:1509-1533)."""
    return bytes(IUPAC_COMP.get(b, ord("N")) for b in reversed(s))


def default_cfg(**kw):
    c = dict(write_nm=True, write_md=True, write_as=True, write_xs=True, write_rg=True,
             split_as_secondary=False, soft_clip_supplementary=False, read_group=None)
    c.update(kw)
    return c


def _rg_id(rg: str) -> str:
    """extract_rg_id (io/mod.rs:1788-1799)."""
    text = rg[4:] if rg.startswith("@RG\t") else rg
    for field in text.split("\t"):
        if field.startswith("ID:"):
            return field[3:]
    return text


def append_unmapped(read_id, seq: bytes, qual):
    """append_unmapped without a mate (io/mod.rs:971-1033)."""
    s = seq.decode() if seq else "*"
    q = qual.decode() if qual else "*"
    return f"{read_id}\t4\t*\t0\t0\t*\t*\t0\t0\t{s}\t{q}\n"


def append_alignment(read_id, seq: bytes, qual, a, contig_names, cfg=None):
    """append_alignment (io/mod.rs:1036-1140) for a single-end record."""
    cfg = cfg or default_cfg()
    ops = parse_cigar(a["cigar"])
    n = len(seq)
    if not ops or sum(k for k, op in ops if op in "MIS") != n:
        raise ValueError("invalid mapped CIGAR")
    flag = 0x10 if a["is_rev"] else 0
    if a["is_supplementary"]:
        flag |= 0x100 if cfg["split_as_secondary"] else 0x800
    hard = a["is_supplementary"] and not cfg["soft_clip_supplementary"]
    lead = ops[0][0] if hard and ops[0][1] == "S" else 0
    trail = ops[-1][0] if hard and len(ops) > 1 and ops[-1][1] == "S" else 0
    ke = max(n - trail, 0)
    ks = min(lead, ke)
    cig = "".join(f"{k}{'H' if op == 'S' and hard and (i == 0 or i + 1 == len(ops)) else op}"
                  for i, (k, op) in enumerate(ops))
    if ks == ke:
        sq = "*"
    elif a["is_rev"]:
        sq = _rc_iupac(seq[n - ke:n - ks]).decode()
    else:
        sq = seq[ks:ke].decode()
    if qual is None or len(qual) != n or ks == ke:
        ql = "*"
    elif a["is_rev"]:
        ql = qual[n - ke:n - ks][::-1].decode()
    else:
        ql = qual[ks:ke].decode()
    line = (f"{read_id}\t{flag}\t{contig_names[a['ref_id']]}\t{a['ref_start'] + 1}\t{a['mapq']}"
            f"\t{cig}\t*\t0\t0\t{sq}\t{ql}")
    if cfg["write_nm"]:
        line += f"\tNM:i:{a['nm']}"
    if cfg["write_md"]:
        line += f"\tMD:Z:{a['md']}"
    if cfg["write_as"]:
        line += f"\tAS:i:{a['score']}"
    if cfg["write_xs"] and a.get("xs_score") is not None:
        line += f"\tXS:i:{a['xs_score']}"
    if cfg["write_rg"] and cfg["read_group"] is not None:
        line += f"\tRG:Z:{_rg_id(cfg['read_group'])}"
    return line + "\n"


def sam_text(read_ids, reads, quals, alns, contig_names, cfg=None):
    out = []
    for k, recs in enumerate(alns):
        q = quals[k] if quals is not None else None
        if not recs:
            out.append(append_unmapped(read_ids[k], reads[k], q))
        for a in recs:
            out.append(append_alignment(read_ids[k], reads[k], q, a, contig_names, cfg))
    return "".join(out).encode()
