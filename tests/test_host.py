"""Host-side logic of the drop-in (CPU only, no device calls).

Known answers from pyPaSWAS's own unit tests where they still describe the
current reference code (test_hit.py, test_formatters.py), plus the pieces
of the Python host mirror: scores, readers, CLI, starting-point records and
the conversion of library results into Hits, Hit lists and TXT/SAM records.
"""
import logging
import os
import struct
import tempfile

import numpy as np
import pytest

from pypaswas_amd.Core.Exceptions import CudaException, InvalidOptionException, ReaderException
from pypaswas_amd.Core.Formatters import FASTA, DefaultFormatter, SamFormatter, TrimmerFormatter
from pypaswas_amd.Core.Hit import Hit
from pypaswas_amd.Core.HitList import HitList

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LOG = logging.getLogger("test")


class _Seq:
    def __init__(self, ident, seq):
        self.id = ident
        self.seq = seq
        self.original_length = len(seq)
        self.distance = None


def _hit(seq_id="sequence", target_id="target"):
    h = Hit(LOG, _Seq(seq_id, "GGACTGTAGC"), _Seq(target_id, "GGGGACTGTAGCCC"), (0, 9), (2, 11))
    h.set_sequence_match("GGACTGTAGC")
    h.set_target_match("GGACTGTAGC")
    h.set_alignment("||||||||||")
    h.set_scores(100.0, 42, 12)
    return h


def _rev_hit():
    h = Hit(LOG, _Seq("sequenceRC", "ACTGACTGACTGACTGCC"), _Seq("targetRC", "CCGCTGTCCGATTCTGTGTGGG"), (1, 15),
            (3, 17))
    h.set_sequence_match("CTGACTGA--CTGACTG")
    h.set_target_match("CTGTCCGATTCTG--TG")
    h.set_alignment("|||.|.||  |||  ||")
    h.set_scores(100, 42, 12)
    return h


# ---- Hit: pyPaSWAS/Core/test/test_hit.py ------------------------------------------

def test_hit_invalid_locations():
    s, t = _Seq("s", "GGACTGTAGC"), _Seq("t", "GGGGACTGTAGCCC")
    assert Hit(LOG, s, t, (1, 9), (2, 11)).seq_location == (1, 9)
    for bad in [(1), (1, 1, 9), (-1, 9), (1, 10), (9, 9), ("a", 9), (1, "a")]:
        with pytest.raises(CudaException):
            Hit(LOG, s, t, bad, (2, 11))
    for bad in [(2), (0, 0, 11), (-1, 9), (0, 14), (7, 7), ("a", 9), (0, "a")]:
        with pytest.raises(CudaException):
            Hit(LOG, s, t, (0, 9), bad)


def test_hit_set_scores_validation():
    h = _hit()
    for args in [(-100.0, 42, 12), (100.0, -42, 12), (100.0, 42.0, 12), (100.0, 42, -12), (100.0, 42, 12.0)]:
        with pytest.raises(CudaException):
            h.set_scores(*args)
    assert len(h.keys()) == 3


def test_sam_line_known_answers():
    f = _hit().get_sam_line().split("\t")
    # 16 fields in the current record layout (Hit.py:266-272)
    assert len(f) == 16
    assert f[0] == "sequence" and f[2] == "target"
    assert int(f[1]) == 0 and int(f[3]) == 3 and int(f[4]) == 255
    assert f[5] == "10="
    assert f[6] == "*" and int(f[7]) == 0 and int(f[8]) == 0
    assert f[9] == "GGACTGTAGC" and f[10] == "*"
    assert f[11] == "AS:i:100"
    # relative score = score / shortest length (Hit.py:112-115)
    assert f[12] == "RS:f:10.0"
    assert f[13] == "BS:f:10.0" and f[14] == "QC:f:1.0" and f[15] == "QI:f:4.2"


def test_sam_line_reverse_complement():
    f = _rev_hit().get_sam_line().split("\t")
    assert int(f[1]) == 16
    assert f[0] == "sequenc" and f[2] == "targe"   # ids ending in RC lose three characters
    # the RC record reverses the alignment in place before the CIGAR is built
    assert f[5] == "2S2=2I3=2D2=1X1=1X3=1S"
    assert int(f[3]) == 22 - 17
    assert f[9] == "GGCAGTCAGTCAGTCAGT"


def test_sam_line_requires_strings():
    s, t = _Seq("s", "GGACTGTAGC"), _Seq("t", "GGGGACTGTAGCCC")
    for missing in ("seq", "tgt", "aln"):
        h = Hit(LOG, s, t, (0, 9), (2, 11))
        if missing != "seq":
            h.set_sequence_match("GGACTGTAGC")
        if missing != "tgt":
            h.set_target_match("GGACTGTAGC")
        if missing != "aln":
            h.set_alignment("||||||||||")
        else:
            h.set_alignment("|")  # set_scores divides by the alignment length
        h.set_scores(100.0, 1, 0)
        if missing == "aln":
            h.set_alignment("")
        with pytest.raises(CudaException):
            h.get_sam_line()


def test_sam_sq_and_trimmed():
    h = _hit()
    assert h.get_sam_sq() == "@SQ\tSN:target\tLN:14"
    # trimmed FASTA cuts the query at the target coordinates (Hit.py:285-286)
    assert h.get_trimmed_line() == ">target sequence\n" + "GG" + ""
    assert h.get_full_fasta() == ">sequence\nGGACTGTAGC"


# ---- Formatters: pyPaSWAS/Core/test/test_formatters.py -----------------------------

def _many_hits(n=500):
    hl = HitList(LOG)
    for k in range(n):
        s = _Seq("sequence-{}".format(k), "GCTGACTGACTG")
        t = _Seq("target-{}".format(k), "ACTGACTGACTG")
        h = Hit(LOG, s, t, (0, 1), (0, 1))
        h.set_alignment("-CTGACTGACTG")
        h.set_sequence_match(s.seq)
        h.set_target_match(t.seq)
        h.set_scores(1.0, 1, 0)
        hl.append(h)
    return hl


def test_formatter_line_counts():
    hl = _many_hits()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "o")
        DefaultFormatter(LOG, hl, out).print_results()
        assert len(open(out).readlines()) == 4 * 500
        SamFormatter(LOG, hl, out).print_results()
        assert len(open(out).readlines()) == 2 * 500 + 2
        TrimmerFormatter(LOG, hl, out).print_results()
        assert len(open(out).readlines()) == 2 * 500
        FASTA(LOG, hl, out).print_results()
        assert len(open(out).readlines()) == 2 * 500


def test_hitlist_dedupe():
    hl = HitList(LOG)
    s, t = _Seq("s", "ACGTACGTACGT"), _Seq("t", "ACGTACGTACGT")

    def mk(a, b, score):
        h = Hit(LOG, s, t, (a, b), (a, b))
        h.set_alignment("|" * (b - a + 1))
        h.set_sequence_match("A" * (b - a + 1))
        h.set_target_match("A" * (b - a + 1))
        h.set_scores(score, b - a + 1, 0)
        return h

    h1 = mk(0, 5, 6.0)
    hl.append(h1)
    hl.append(mk(0, 4, 5.0))      # same start, lower score: rejected
    assert list(hl.real_hits.values()) == [h1]
    h3 = mk(1, 5, 7.0)            # same end, higher score: replaces h1
    hl.append(h3)
    assert list(hl.real_hits.values()) == [h3]


# ---- scores ------------------------------------------------------------------------

class _Settings:
    gap_score = "-5"
    gap_extension = "0"
    lower_limit_score = "1.0"
    minimum_score = "30"
    match_score = "5"
    mismatch_score = "-3"
    other_score = "-1"
    any_score = "0"
    custom_matrix = None


def test_dna_matrix_rules():
    from pypaswas_amd.Core.Scores import DnaRnaScore
    sc = DnaRnaScore(LOG, _Settings())
    S = lambda a, b: sc.matrix[ord(a) - 65][ord(b) - 65]
    assert S("A", "A") == "5" and S("A", "C") == "-3" and S("U", "T") == "-3"
    assert S("N", "N") == "5"   # identity comes first
    assert S("N", "A") == "0" and S("A", "N") == "0"
    assert S("K", "G") == "5" and S("T", "K") == "5" and S("R", "C") == "-3"
    assert S("B", "D") == "-1" and S("Y", "R") == "-1"
    assert sc.gap_extension is None and sc.score_type == "DNA_RNA"
    m = sc.to_float32()
    assert m.dtype == np.float32 and m[0, 0] == 5.0 and (m == m.T).all()


def test_blosum62_published_entries():
    from pypaswas_amd.Core.Scores import Blosum62Score
    sc = Blosum62Score(LOG, _Settings())
    B = lambda a, b: sc.matrix[ord(a) - 65][ord(b) - 65]
    # published BLOSUM62 values
    assert (B("A", "A"), B("W", "W"), B("C", "C"), B("A", "R"), B("W", "C"), B("L", "I"), B("K", "R")) == \
        (4, 11, 9, -1, -2, 2, 2)
    m = np.array(sc.matrix)
    assert (m == m.T).all()
    # the reference takes max over the lexicographically largest row (Scores.py:243)
    assert sc.highest_score == 4


def test_gap_extension_zero_is_linear():
    from pypaswas_amd.Core.Scores import BasicScore
    s = _Settings()
    s.gap_extension = "-1"
    sc = BasicScore(LOG, s)
    assert sc.gap_extension == -1.0 and sc.dimensions == 255 and sc.char_offset == "\1"


# ---- readers, CLI, starting points -------------------------------------------------

def test_readers_fasta_and_genbank():
    from pypaswas_amd.Core.Readers import BioPythonReader
    r = BioPythonReader(LOG, os.path.join(GOLDEN, "target_fasta.faa"), "fasta", 100000)
    r.read_records()
    ids = [x.id for x in r.get_records()]
    assert ids[0] == "full_match_first_A_block_200_NUCL" and len(r.get_records()[0].seq) == 200
    r.complement_records()
    r.sort_records()
    assert any(x.id.endswith("_RC") for x in r.get_records())
    lens = [len(x.seq) for x in r.get_records()]
    assert lens == sorted(lens, reverse=True)
    g = BioPythonReader(LOG, os.path.join(GOLDEN, "query_genbank.gb"), "genbank", 100000)
    g.read_records()
    assert g.get_records()[0].id == "ATBQ01000001.1" and len(g.get_records()[0].seq) == 1907
    with pytest.raises(ReaderException):
        BioPythonReader(LOG, os.path.join(GOLDEN, "query_fasta.faa"), "fasta", 100000).read_records(5, 6)
    short = BioPythonReader(LOG, os.path.join(GOLDEN, "target_fasta.faa"), "fasta", 50)
    short.read_records()
    assert all(len(x.seq) <= 50 for x in short.get_records())


def test_fastq_reader(tmp_path):
    from pypaswas_amd.Core.Readers import BioPythonReader
    p = tmp_path / "r.fastq"
    p.write_text("@r1 desc\nACGTN\n+\nIIIII\n@r2\nacgt\n+r2\nIIII\n")
    r = BioPythonReader(LOG, str(p), "fastq", 0)
    r.read_records()
    assert [(x.id, str(x.seq)) for x in r.get_records()] == [("r1", "ACGTN"), ("r2", "ACGT")]


def test_cli_defaults_and_override(tmp_path):
    from pypaswas_amd.cli import parse_cli
    s, a = parse_cli(argv=["q.fa", "t.fa"])
    assert (s.gap_score, s.match_score, s.minimum_score, s.framework, s.relative_score) == ("-5", "5", "30", "HIP",
                                                                                          "2.0")
    cfg = tmp_path / "c.cfg"
    cfg.write_text("[Aligner]\nmatch_score=2\n")
    s, a = parse_cli(argv=["-c", str(cfg), "q.fa", "t.fa", "-G", "-7"])
    assert s.match_score == "2" and s.gap_score == "-7"
    with pytest.raises(InvalidOptionException):
        parse_cli(argv=["only_one.fa"])


def test_starting_point_record():
    from pypaswas_amd.Core.StartingPoint import StartingPoint
    raw = struct.pack("<6I3f", 3, 4, 5, 6, 7, 1, 12.0, 15.0, 2.5) * 2
    sp = StartingPoint(LOG)
    sp.parse_byte_string(raw, 1)
    assert (sp.sequence, sp.target, sp.block_x, sp.block_y, sp.value_x, sp.value_y) == (3, 4, 5, 6, 7, 1)
    assert sp.score == 12.0 and isinstance(sp.score, float) and sp.pos_score == 2.5


def test_min_start_scores_match_reference_arithmetic():
    from pypaswas_amd.Core.SmithWatermanHIP import min_start_scores
    got = min_start_scores("5", "0.2", [150, 40], [60, 300])
    want = np.array([np.float32(float("5") * 60 * float("0.2")), np.float32(float("5") * 40 * float("0.2"))])
    assert np.array_equal(got, want)


# ---- library results -> Hits -> records (oracle output stands in for the device) ----

def _fake_result(O, br, queries, targets, qids, tids, minimum):
    """An AlignResult built from the oracle's walk, laid out as pw_align_pairs lays it out."""
    from pypaswas_amd._native import AlignResult, HIT_DTYPE, SP_DTYPE
    ws = O.walk_c(br.directions.copy(), br.starting_points, queries, targets, qids, tids, minimum)
    per_pair = {}
    for w in ws:
        per_pair.setdefault((w.qi, w.ti), []).append(w)
    hits, ops, aligned = [], bytearray(), bytearray()
    for key in sorted(per_pair):
        for w in per_pair[key]:
            codes = []
            if w.shown:
                for a, b in zip(w.sequence_match, w.target_match):
                    codes.append("L" if b == "-" else ("U" if a == "-" else "D"))
            off = len(ops)
            ops += "".join(reversed(codes)).encode()
            # the three rows at 3 * ops_offset, as pw_result.aligned holds them
            aligned += (w.sequence_match + w.alignment + w.target_match).encode() if w.shown else b""
            hits.append((w.qi, w.ti, int(w.shown), w.s_start, w.s_end, w.t_start, w.t_end, w.matches,
                         w.mismatches, len(codes), w.score, 0, off))
    N, T = br.maxima.shape
    return AlignResult(br.maxima.reshape(-1),
                       np.zeros(0, SP_DTYPE), np.array(hits, dtype=HIT_DTYPE), bytes(ops), 0, 0, 0.0, 0, 0, 0, 0, 0,
                       0, aligned=memoryview(bytes(aligned)))


def test_collect_hits_reproduces_genbank_fixture(oracle):
    """The HIP backend's Hit construction, filtering and TXT formatting on the fixture's walk."""
    import re
    from pypaswas_amd.Core.SmithWatermanHIP import SmithWatermanHIP
    from pypaswas_amd.Core.SWSeqRecord import SWSeqRecord
    from pypaswas_amd.Core.SWSeq import Seq
    O = oracle

    def gb(path):
        txt = open(path).read()
        return (re.search(r"^VERSION\s+(\S+)", txt, re.M).group(1),
                re.sub(r"[^a-zA-Z]", "", txt.split("ORIGIN", 1)[1].split("//")[0]).upper())

    q = gb(os.path.join(GOLDEN, "query_genbank.gb"))
    t = gb(os.path.join(GOLDEN, "target_genbank.gb"))
    rq, rt = O.prepare_records([q], [t], True)
    p = O.Params(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, minimum=30.0, highest=1.0)
    br = O.sched_batch([s for _, s in rq], [s for _, s in rt], p)
    res = _fake_result(O, br, [s for _, s in rq], [s for _, s in rt], [i for i, _ in rq], [i for i, _ in rt], 30.0)

    class S:
        relative_score = query_identity = query_coverage = base_score = "0"

    eng = SmithWatermanHIP.__new__(SmithWatermanHIP)
    eng.logger, eng.settings = LOG, S()
    qrecs = [SWSeqRecord(Seq(s), i) for i, s in rq]
    trecs = [SWSeqRecord(Seq(s), i) for i, s in rt]
    hl = HitList(LOG)
    eng._collect_hits(res, qrecs, trecs, hl)
    lines = DefaultFormatter(LOG, hl, None).format_lines()
    want = open(os.path.join(GOLDEN, "aligner_basic_genbank_output.txt")).read().splitlines()
    got = "\n".join(lines).splitlines()
    assert got[1:] == want[1:]
    assert got[0].split(", ")[:11] == want[0].split(", ")[:11]


def test_collect_hits_matches_oracle_records(oracle):
    """Random DNA batch with many starts: TXT records of the host path == the oracle's."""
    from pypaswas_amd.Core.SmithWatermanHIP import SmithWatermanHIP
    from pypaswas_amd.Core.SWSeqRecord import SWSeqRecord
    from pypaswas_amd.Core.SWSeq import Seq
    from tests._parity import DNA, planted_set
    O = oracle
    rng = np.random.default_rng(8)
    qs, ts = planted_set(rng, 5, 4, (40, 160), (20, 90), DNA, plant=1.0, sub=0.04)
    rq = [("q%d" % i, s) for i, s in enumerate(qs)]
    rt = [("t%d" % i, s) for i, s in enumerate(ts)]
    p = O.Params(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, gap_ext=-1.0, minimum=8.0, lower_limit=0.7,
                 filter_factor=0.1, highest=1.0)
    hl_o, _, _ = O.align_all(rq, rt, p)
    br = O.sched_batch(qs, ts, p)  # align_all's walk consumed its direction matrices
    res = _fake_result(O, br, qs, ts, [i for i, _ in rq], [i for i, _ in rt], p.minimum)

    class S:
        relative_score = query_identity = query_coverage = base_score = "0"

    eng = SmithWatermanHIP.__new__(SmithWatermanHIP)
    eng.logger, eng.settings = LOG, S()
    hl = HitList(LOG)
    eng._collect_hits(res, [SWSeqRecord(Seq(s), i) for i, s in rq], [SWSeqRecord(Seq(s), i) for i, s in rt], hl)
    got = "".join(l + "\n" for l in DefaultFormatter(LOG, hl, None).format_lines())
    assert got == O.txt_records(hl_o)
    assert len(hl.real_hits) > 3
