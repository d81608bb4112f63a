"""Pins the oracle before it is trusted (CPU only).

* the reference's own fixture: test_pypaswas_aligner_basic_genbank
  (pyPaSWAS/Core/test/test_pypaswasall.py:62-83) reproduced end to end;
* exhaustive enumeration of local alignments on tiny inputs (maxima);
* closed forms (identical sequences, homopolymers, one substitution);
* hand-checked direction tie-breaks;
* the two independent restatements (reference schedule vs serial row-major)
  agree on maxima, post-traceback directions and starting points;
* the C walk agrees with the pure-Python walk.
"""
import itertools
import math
import os
import re

import numpy as np
import pytest

from tests._parity import AMINO, DNA, planted_set

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = float("inf")


def read_gb(path):
    txt = open(path).read()
    ver = re.search(r"^VERSION\s+(\S+)", txt, re.M).group(1)
    seq = re.sub(r"[^a-zA-Z]", "", txt.split("ORIGIN", 1)[1].split("//")[0])
    return ver, seq.upper()


def dna_p(O, affine=False, **kw):
    base = dict(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, gap_ext=(-1.0 if affine else None),
                lower_limit=1.0, minimum=10.0, filter_factor=0.2, highest=1.0)
    base.update(kw)
    return O.Params(**base)


# --------------------------------------------------------------------------
# the reference's fixture
# --------------------------------------------------------------------------

def test_genbank_fixture(oracle):
    """-r 1 -q -3 -G -5, filters 0: one hit 7..179 / 7..179, score 165, 171 matches, 2 mismatches."""
    O = oracle
    q = read_gb(os.path.join(GOLDEN, "query_genbank.gb"))
    t = read_gb(os.path.join(GOLDEN, "target_genbank.gb"))
    p = dna_p(O, minimum=30.0)
    rq, rt = O.prepare_records([q], [t], True)
    hl, _, _ = O.align_all(rq, rt, p)
    got = O.txt_records(hl).splitlines()
    want = open(os.path.join(GOLDEN, "aligner_basic_genbank_output.txt")).read().splitlines()
    assert len(got) == len(want) == 4
    gf, wf = got[0].split(", "), want[0].split(", ")
    assert gf[:11] == wf[:11] and gf[12:14] == wf[12:14] and gf[16] == wf[16]
    # the fixture was written by Python 2 (12 significant digits)
    for k in (11, 14, 15):
        assert float(gf[k]) == pytest.approx(float(wf[k]), rel=1e-11)
    assert got[1:] == want[1:]


# --------------------------------------------------------------------------
# exhaustive enumeration
# --------------------------------------------------------------------------

def brute_max(q, t, S, G, E, affine):
    """Best local alignment score by enumerating every alignment of every substring pair.

    Gap model of the reference's CPU kernels: linear, every gap column costs G;
    affine, the first column of a gap run costs G and each further one E, and a
    run of query gaps cannot touch a run of target gaps (default_main_cpu.cl:283-330).
    """
    best = 0.0

    def rec(a, b, i, j, score, last):
        nonlocal best
        if i == len(a) and j == len(b):
            best = max(best, score)
            return
        if i < len(a) and j < len(b):
            rec(a, b, i + 1, j + 1, score + S(a[i], b[j]), "P")
        if i < len(a) and not (affine and last == "Y"):
            rec(a, b, i + 1, j, score + (E if affine and last == "X" else G), "X")
        if j < len(b) and not (affine and last == "X"):
            rec(a, b, i, j + 1, score + (E if affine and last == "Y" else G), "Y")

    for i1, i2 in itertools.combinations(range(len(q) + 1), 2):
        for j1, j2 in itertools.combinations(range(len(t) + 1), 2):
            rec(q[i1:i2], t[j1:j2], 0, 0, 0.0, "P")
    return best


@pytest.mark.parametrize("affine", [False, True])
def test_maxima_exhaustive(oracle, affine):
    O = oracle
    rng = np.random.default_rng(5 + affine)
    m = O.dna_rna_matrix(2.0, -1.0, -1.0, 0.0)
    S = lambda a, b: float(m[ord(a) - 65][ord(b) - 65])
    p = O.Params(matrix=m, gap=-2.0, gap_ext=(-1.0 if affine else None), minimum=1e9, highest=2.0)
    for _ in range(60):
        q = "".join(rng.choice(list("ACG"), size=int(rng.integers(1, 6))))
        t = "".join(rng.choice(list("ACG"), size=int(rng.integers(1, 6))))
        mx, _, _ = O.serial_pair(q, t, p)
        assert mx == brute_max(q, t, S, -2.0, -1.0, affine), (q, t)


def test_affine_gap_open_is_G_not_G_plus_E(oracle):
    """One gap of length 1 costs G, of length k costs G + (k-1)E (CPU variant)."""
    O = oracle
    p = O.Params(matrix=O.dna_rna_matrix(5.0, -4.0, -1.0, 0.0), gap=-6.0, gap_ext=-1.0, minimum=1e9, highest=5.0)
    left, right = "ACGTACGTAC", "GATTACAGGT"
    for k in (1, 2, 3):
        mx, _, _ = O.serial_pair(left + right, left + "T" * k + right, p)
        # 20 matches vs gap of k: 100 - 6 - (k-1); inserted Ts could also match nothing better
        assert mx == 100.0 - 6.0 - (k - 1)


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------

@pytest.mark.parametrize("affine", [False, True])
def test_closed_forms(oracle, affine):
    O = oracle
    p = dna_p(O, affine, minimum=1e9)
    rng = np.random.default_rng(1)
    s = "".join(rng.choice(list(DNA), size=57))
    assert O.serial_pair(s, s, p)[0] == 57.0
    assert O.serial_pair("A" * 31, "A" * 12, p)[0] == 12.0
    mutated = s[:30] + ("C" if s[30] != "C" else "G") + s[31:]
    assert O.serial_pair(s, mutated, p)[0] == 56.0 - 3.0
    # no common residue: maximum 0
    assert O.serial_pair("AAAA", "CCCC", p)[0] == 0.0


def test_direction_tie_breaks(oracle):
    """LEFT (x-1) wins ties over UP (y-1), which wins over UL; NONE exactly when H = 0."""
    O = oracle
    # match 2, mismatch -1, gap -2: q = "AA", t = "A":  H[1][1] = 2 (UL); H[2][1]: d = 0+2 = 2,
    # l = H[1][1] - 2 = 0 -> UL (2 > 0)
    m = O.dna_rna_matrix(2.0, -1.0, -1.0, 0.0)
    p = O.Params(matrix=m, gap=-2.0, minimum=1e9, highest=2.0)
    _, d, _ = O.serial_pair("AA", "A", p)
    assert d.tolist() == [[1], [1]]
    # gap 0: q = "AC", t = "A": H[2][1]: d = 0 + (-1) = -1, l = H[1][1] + 0 = 2 -> LEFT (3)
    p0 = O.Params(matrix=m, gap=0.0, minimum=1e9, highest=2.0)
    _, d, _ = O.serial_pair("AC", "A", p0)
    assert d.tolist() == [[1], [3]]
    # q = "A", t = "AC": H[1][2]: u = H[1][1] + 0 = 2 -> UP (2)
    _, d, _ = O.serial_pair("A", "AC", p0)
    assert d.tolist() == [[1, 2]]
    # tie l == u == d == 2 at (2,2) of q = "AA", t = "AA" with gap 0: d = H[1][1] + 2 = 4 wins;
    # at (1,2) of q = "A", t = "AA": u = 2, d = 0 + 2 = 2, tie -> UP before UL
    _, d, _ = O.serial_pair("A", "AA", p0)
    assert d.tolist() == [[1, 2]]
    # mismatch everywhere: NONE (32)
    _, d, _ = O.serial_pair("AA", "CC", p)
    assert (d == 32).all()


# --------------------------------------------------------------------------
# two restatements agree
# --------------------------------------------------------------------------

CASES = [
    ("dna linear", DNA, False, (5, 90), (3, 40), dict(minimum=6.0)),
    ("dna affine", DNA, True, (5, 90), (3, 40), dict(minimum=6.0)),
    ("dna linear llimit", DNA, False, (20, 70), (10, 50), dict(minimum=4.0, lower_limit=0.6, filter_factor=0.0)),
    ("dna affine llimit", DNA, True, (20, 70), (10, 50), dict(minimum=4.0, lower_limit=0.6, filter_factor=0.0)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_schedule_invariance(oracle, case):
    O = oracle
    name, alpha, affine, ql, tl, kw = case
    rng = np.random.default_rng(len(name))
    qs, ts = planted_set(rng, 4, 3, ql, tl, alpha, plant=0.8, sub=0.05)
    p = dna_p(O, affine, **kw)
    br = O.sched_batch(qs, ts, p)
    n_sp = 0
    for qi, q in enumerate(qs):
        for ti, t in enumerate(ts):
            mx, d, sps = O.serial_pair(q, t, p)
            assert mx == br.maxima[qi, ti]
            assert np.array_equal(d, br.directions[qi, ti, :len(q), :len(t)])
            mine = [(s.blockX, s.blockY, s.valueX, s.valueY, s.score, s.maxScore, s.posScore) for s in sps]
            theirs = [(s.blockX, s.blockY, s.valueX, s.valueY, s.score, s.maxScore, s.posScore)
                      for s in br.starting_points if s.sequence == qi and s.target == ti]
            assert mine == theirs
            n_sp += len(mine)
    assert n_sp > 0


@pytest.mark.parametrize("affine", [False, True])
def test_walk_c_matches_python(oracle, affine):
    O = oracle
    rng = np.random.default_rng(17)
    qs, ts = planted_set(rng, 4, 4, (40, 120), (20, 80), DNA, plant=1.0, sub=0.04)
    p = dna_p(O, affine, minimum=4.0, lower_limit=0.5, filter_factor=0.0)
    br = O.sched_batch(qs, ts, p)
    qid = ["q%d" % i for i in range(4)]
    tid = ["t%d" % i for i in range(4)]
    a = O.walk_py(br.directions.copy(), br.starting_points, qs, ts, qid, tid, p.minimum)
    b = O.walk_c(br.directions.copy(), br.starting_points, qs, ts, qid, tid, p.minimum)
    def norm(w):
        d = dict(vars(w))
        if not w.shown:  # no Hit is built: the strings of a suppressed walk are never used
            d.update(sequence_match="", alignment="", target_match="")
        return d
    assert [norm(x) for x in a] == [norm(x) for x in b]
    assert sum(not x.shown for x in a) > 0 and sum(x.shown for x in a) > 0


def test_protein_blosum_restatements(oracle):
    O = oracle
    rng = np.random.default_rng(3)
    qs, ts = planted_set(rng, 3, 3, (30, 150), (30, 150), AMINO, plant=1.0, sub=0.4, indel=0.05)
    p = O.Params(matrix=O.BLOSUM62, gap=-11.0, gap_ext=-1.0, minimum=20.0, filter_factor=0.05,
                 highest=float(max(max(O.BLOSUM62))), dna=False)
    br = O.sched_batch(qs, ts, p)
    for qi, q in enumerate(qs):
        for ti, t in enumerate(ts):
            mx, d, _ = O.serial_pair(q, t, p)
            assert mx == br.maxima[qi, ti]
            assert np.array_equal(d, br.directions[qi, ti, :len(q), :len(t)])


def _fake_result(O, qs, ts, p):
    """The schedule-faithful oracle's batch in the shape of a HIP result (maxima,
    36-byte records, hits with op codes), to exercise the per-pair checker."""
    from types import SimpleNamespace
    from pypaswas_amd._native import HIT_DTYPE, SP_DTYPE
    br = O.sched_batch(qs, ts, p)
    ws = O.walk_c(br.directions.copy(), br.starting_points, qs, ts, ["q%d" % i for i in range(len(qs))],
                  ["t%d" % i for i in range(len(ts))], p.minimum)
    sps = np.zeros(len(br.starting_points), SP_DTYPE)
    for k, s in enumerate(br.starting_points):
        sps[k] = (s.sequence, s.target, s.blockX, s.blockY, s.valueX, s.valueY, s.score, s.maxScore, s.posScore)
    hits = np.zeros(len(ws), HIT_DTYPE)
    ops, aligned = bytearray(), bytearray()
    for k, w in enumerate(ws):
        off = len(ops)
        if w.shown:
            for a, b in reversed(list(zip(w.sequence_match, w.target_match))):
                ops += b"L" if b == "-" else (b"U" if a == "-" else b"D")
            aligned += (w.sequence_match + w.alignment + w.target_match).encode()
        hits[k] = (w.qi, w.ti, int(w.shown), w.s_start, w.s_end, w.t_start, w.t_end, w.matches, w.mismatches,
                   len(w.alignment) if w.shown else 0, w.score, 0, off)
    return SimpleNamespace(maxima=br.maxima.reshape(-1), starting_points=sps, hits=hits, ops=bytes(ops),
                           aligned=bytes(aligned))


@pytest.mark.parametrize("affine", [False, True])
def test_per_pair_checker_accepts_the_batch_and_rejects_a_change(oracle, affine):
    """tests._parity.compare_pairs (serial restatement pair by pair) accepts the
    schedule-faithful batch and notices a changed walk."""
    from tests._parity import compare_pairs
    O = oracle
    rng = np.random.default_rng(23 + affine)
    qs, ts = planted_set(rng, 5, 4, (30, 120), (20, 70), DNA, plant=1.0, sub=0.05)
    p = dna_p(O, affine, minimum=6.0, lower_limit=0.7, filter_factor=0.0)
    res = _fake_result(O, qs, ts, p)
    pq = np.repeat(np.arange(5), 4)
    pt = np.tile(np.arange(4), 5)
    assert compare_pairs(O, qs, ts, pq, pt, p, res, threads=4) > 0
    k = int(np.flatnonzero(res.hits["shown"])[0])
    res.hits["t_end"][k] += 1
    with pytest.raises(AssertionError):
        compare_pairs(O, qs, ts, pq, pt, p, res, threads=4)
