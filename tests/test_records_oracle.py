"""The oracle's record restatements (SAM, trimmed FASTA, FASTA), pinned before use (CPU only).

* SAM fields: the still-valid known answers of pyPaSWAS/Core/test/test_hit.py
  (:265-372) on the same two hits, plus the reverse-complement CIGAR worked out
  by hand from Hit._get_sam_cigar (Hit.py:338-372; see the comment below);
* trimmed FASTA: the query cut at the target coordinates (Hit.py:285-286),
  worked out by hand;
* then the product's formatters (pypaswas_amd/Core/Formatters.py) against the
  oracle's records on the oracle's own hits of configuration 1 (64 reads x 16
  adapters + reverse complements), for all four output formats.
"""
import logging

import pytest

from oracle import oracle as O

LOG = logging.getLogger("test")


def ohit(qid, qseq, tid, tlen, s, t, sm, aln, tm, score=100.0, matches=42, mismatches=12):
    h = O.OHit(qid, tid, len(qseq), tlen, s, t, score, matches, mismatches, sm, aln, tm, qseq=qseq)
    shortest = tlen if len(qseq) > tlen else len(qseq)
    h.relative_score = float(score) / float(shortest)
    h.query_identity = float(matches) / float(s[1] - s[0] + 1)
    h.query_coverage = float(s[1] - s[0] + 1) / float(len(qseq))
    h.base_score = float(score) / float(len(aln))
    return h


def fwd():
    # test_hit.py:33-51
    return ohit("sequence", "GGACTGTAGC", "target", 14, (0, 9), (2, 11), "GGACTGTAGC", "||||||||||", "GGACTGTAGC")


def rev():
    # test_hit.py:53-71
    return ohit("sequenceRC", "ACTGACTGACTGACTGCC", "targetRC", 22, (1, 15), (3, 17), "CTGACTGA--CTGACTG",
                "|||.|.||  |||  ||", "CTGTCCGATTCTG--TG", score=100)


def test_oracle_sam_line_known_answers():
    f = O.sam_line(fwd()).split("\t")
    assert len(f) == 16
    assert f[0] == "sequence" and int(f[1]) == 0 and f[2] == "target" and int(f[3]) == 3 and int(f[4]) == 255
    assert f[5] == "10=" and f[6] == "*" and int(f[7]) == 0 and int(f[8]) == 0
    assert f[9] == "GGACTGTAGC" and f[10] == "*" and f[11] == "AS:i:100"
    assert f[12] == "RS:f:10.0" and f[13] == "BS:f:10.0" and f[14] == "QC:f:1.0" and f[15] == "QI:f:4.2"


def test_oracle_sam_line_reverse_complement():
    # By hand from Hit.py:338-372 with rc set: leading clip len(seq) - s_end - 1 =
    # 18 - 15 - 1 = 2 -> 2S; the reversed alignment ||  |||  ||.|.||| gives runs
    # 2 '|', 2 ' ' (query char at the run's last column is 'A': I), 3 '|',
    # 2 ' ' (query char '-': D), 2 '|', '.', '|', '.', 3 '|'; trailing clip
    # s_start = 1 -> 1S.  POS = 22 - 17 = 5, SEQ = reverse complement.
    f = O.sam_line(rev()).split("\t")
    assert f[0] == "sequenc" and int(f[1]) == 16 and f[2] == "targe"
    assert f[3] == "5"
    assert f[5] == "2S2=2I3=2D2=1X1=1X3=1S"
    assert f[9] == "GGCAGTCAGTCAGTCAGT"


def test_oracle_sam_rc_always_clips():
    # a reverse-complement hit ending at the last query position still opens with 0S
    h = ohit("qRC", "ACGT", "t", 4, (0, 3), (0, 3), "ACGT", "||||", "ACGT")
    assert O.sam_line(h).split("\t")[5] == "0S4="


def test_oracle_sam_records_header():
    hl = O.OHitList()
    hl.append(fwd())
    text = O.sam_records(hl).splitlines()
    assert text[0] == "@HD\tVN:1.4\tSO:unknown"
    assert text[1] == "@SQ\tSN:target\tLN:14"
    assert text[2] == "@PG\tID:0\tPN:paswas\tVN:3.0"
    assert len(text) == 4


def test_oracle_trimmed_and_fasta():
    hl = O.OHitList()
    hl.append(fwd())
    # query GGACTGTAGC cut at target columns 2..11: keeps [0, 2) and [12, ...)
    assert O.trimmed_records(hl) == ">target sequence\nGG\n"
    assert O.fasta_records(hl) == ">sequence\nGGACTGTAGC\n"


def _product_hitlist(ohl):
    """Product Hit objects carrying the oracle's hits, in the oracle's insertion order."""
    from pypaswas_amd.Core.Hit import Hit
    from pypaswas_amd.Core.HitList import HitList
    from pypaswas_amd.Core.SWSeqRecord import SWSeqRecord
    hl = HitList(LOG)
    for h in ohl.real_hits.values():
        qrec = SWSeqRecord(h.qseq, h.qid)
        trec = SWSeqRecord("N" * h.tlen, h.tid)
        ph = Hit(LOG, qrec, trec, h.s, h.t)
        ph.set_sequence_match(h.sequence_match)
        ph.set_alignment(h.alignment)
        ph.set_target_match(h.target_match)
        ph.set_scores(h.score, h.matches, h.mismatches)
        hl.add(ph)
    return hl


@pytest.mark.parametrize("affine", [False, True])
def test_product_formatters_match_oracle_records(tmp_path, affine):
    from pypaswas_amd import workloads as W
    from pypaswas_amd.Core.Formatters import FASTA, DefaultFormatter, SamFormatter, TrimmerFormatter
    reads, adapters, _ = W.config1()
    q, t = O.prepare_records(reads, adapters, True)
    p = O.Params(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, gap_ext=(-1.0 if affine else None),
                 lower_limit=1.0, minimum=10.0, filter_factor=0.2, highest=1.0)
    ohl, _, _ = O.align_all(q, t, p)
    assert len(ohl.real_hits) >= 32
    assert any(h.tid.endswith("_RC") for h in ohl.real_hits.values())
    for cls, want in ((DefaultFormatter, O.txt_records), (SamFormatter, O.sam_records),
                      (TrimmerFormatter, O.trimmed_records), (FASTA, O.fasta_records)):
        out = tmp_path / ("out_" + cls.__name__)
        cls(LOG, _product_hitlist(ohl), str(out)).print_results()
        assert out.read_text() == want(ohl), cls.__name__
