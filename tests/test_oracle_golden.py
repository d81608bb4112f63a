"""Hand-derived golden cases for the oracle's traceback and walk (CPU only).

Each case below was worked on paper from the reference's cell and traceback
rules, not computed by any restatement:

* scoring, linear: pyPaSWAS/Core/ocl/default_main_cpu.cl:154-181
  (d = H[x-1][y-1] + S, l = H[x-1][y] + G, u = H[x][y-1] + G; the first strict
  improvement over 0 in the order l, u, d wins: LEFT(3) > UP(2) > UL(1) on
  ties, NONE(32) when H = 0);
* scoring, affine: default_main_cpu.cl:275-346 (M from the diagonal maximum of
  I, J, M plus S; I from I[x][y-1] + E or M[x][y-1] + G; J from J[x-1][y] + E
  or M[x-1][y] + G; I and J stored as -1e10 when negative; the byte is the
  source of c = max(M, I, J) in the order M (5), I (10), J (15), NONE (32)
  when c = 0);
* traceback: default_main_cpu.cl:424-485 (linear: emit a UL cell that reaches
  both thresholds, negate it, negate the predecessor along the code or write
  STOP(64) when that predecessor is 0) and :535-596 (affine: the same on M for
  main-matrix cells only, every visited byte rewritten to `byte & 3`, A(1) or
  STOP), cells in reverse topological order;
* walk: pyPaSWAS/Core/SmithWaterman.py:689-841 (starting points ascending by
  score, stable; follow the codes, IN_ALIGNMENT(96) marks persist, a STOP cell
  is emitted and ends the walk, code 0 / NONE ends it without emitting).

Tables are indexed by 1-based query position x (rows) and target position y
(columns); starting points and walk coordinates are 0-based as the records
hold them.  Every case is checked against both restatements of the oracle
(schedule-faithful and serial) and both walks (C and Python), and the planted
mistakes at the end of this file show that the cases catch a swapped tie rule,
a wrong STOP rule, a wrong walk order and the other rules they exercise.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_C = os.path.join(os.path.dirname(HERE), "oracle", "sw_oracle.c")

UL, UP, LF, NO, ST = 1, 2, 3, 32, 64
M_, I_, J_ = 5, 10, 15  # affine bytes: code | source << 2


def custom_matrix(diag=5.0, off=-4.0, extra=None):
    """26 x 26 over 'A'..'Z': diag on the diagonal of A, C, G, T, off elsewhere."""
    m = [[off] * 26 for _ in range(26)]
    for c in "ACGT":
        m[ord(c) - 65][ord(c) - 65] = diag
    for (a, b), v in (extra or {}).items():
        m[ord(a) - 65][ord(b) - 65] = v
        m[ord(b) - 65][ord(a) - 65] = v
    return m


def _pad(rows, left_cols=0, top_rows=0, width=None, fill=NO):
    """Places a table below top_rows rows and right of left_cols columns of `fill`."""
    w = (width if width is not None else len(rows[0]) + left_cols)
    out = [[fill] * w for _ in range(top_rows)]
    for r in rows:
        out.append([fill] * left_cols + list(r))
    return out


# The 5 x 5 table of ACGTA against itself, match 2, mismatch -1, gap -1:
#          A      C      G      T      A
#   A    2 UL   1 UP   0 NO   0 NO   2 UL
#   C    1 LF   4 UL   3 UP   2 UP   1 LF   (C/A: l = u = 1, LEFT first)
#   G    0 NO   3 LF   6 UL   5 UP   4 UP
#   T    0 NO   2 LF   5 LF   8 UL   7 UP
#   A    2 UL   1 LF   4 LF   7 LF  10 UL   (A/C: l = u = 1, LEFT first)
ACGTA_SELF = [[UL, UP, NO, NO, UL], [LF, UL, UP, UP, LF], [NO, LF, UL, UP, UP], [NO, LF, LF, UL, UP],
              [UL, LF, LF, LF, UL]]

CASES = [
    # L1 -- q ACGT, t AGT; +2/-1, gap -1.  H (x rows, y cols):
    #        A      G      T
    #   A  2 UL   1 UP   0 NO
    #   C  1 LF   1 UL   0 NO     (C/G: d = 1 > l = u = 0)
    #   G  0 NO   3 UL   2 UP
    #   T  0 NO   2 LF   5 UL
    # Start (4,3) = 5; marks (3,2), (2,1) [LEFT], (1,1) whose diagonal
    # predecessor is the corner (0): STOP.  Walk: T/T, G/G, C/-, A/A (STOP
    # emitted) -> ACGT / | || / A-GT, s 0..3, t 0..2.
    dict(name="L1 gap in the target, path to the corner", q="ACGT", t="AGT", m=("dna", 2.0, -1.0), G=-1.0, E=None,
         ll=1.0, mx=5.0,
         d_score=[[UL, UP, NO], [LF, UL, NO], [NO, UL, UP], [NO, LF, UL]],
         d_after=[[ST, UP, NO], [LF, UL, NO], [NO, UL, UP], [NO, LF, UL]],
         sps=[(3, 2, 5.0)],
         walks=[(True, 0, 3, 0, 2, 3, 0, "ACGT", "| ||", "A-GT")]),
    # L2 -- q ACGT, t AGCT; +2/-1, gap -1.  H:
    #        A      G      C      T
    #   A  2 UL   1 UP   0 NO   0 NO
    #   C  1 LF   1 UL   3 UL   2 UP
    #   G  0 NO   3 UL   2 LF   2 UL   (G/C: l = u = 2 > d = 0: LEFT wins the tie)
    #   T  0 NO   2 LF   2 UL   4 UL
    # Start (4,4) = 4 -> (3,3) LEFT -> (2,3) UL -> (1,2) UP -> (1,1) STOP.
    # With UP winning the tie at (3,3) the path and strings would differ.
    dict(name="L2 gaps on both axes, LEFT over UP on the path", q="ACGT", t="AGCT", m=("dna", 2.0, -1.0), G=-1.0,
         E=None, ll=1.0, mx=4.0,
         d_score=[[UL, UP, NO, NO], [LF, UL, UL, UP], [NO, UL, LF, UL], [NO, LF, UL, UL]],
         d_after=[[ST, UP, NO, NO], [LF, UL, UL, UP], [NO, UL, LF, UL], [NO, LF, UL, UL]],
         sps=[(3, 3, 4.0)],
         walks=[(True, 0, 3, 0, 3, 3, 0, "A-CGT", "| | |", "AGC-T")]),
    # L3 -- q ACGAT, t ACGTCA; +3/-3, gap -1, lower limit 0.8 (threshold 8.8).  H:
    #        A      C      G      T      C      A
    #   A  3 UL   2 UP   1 UP   0 NO   0 NO   3 UL
    #   C  2 LF   6 UL   5 UP   4 UP   3 UP   2 LF   (C/C: u = d = 3, UP first; C/A: l = u = 2)
    #   G  1 LF   5 LF   9 UL   8 UP   7 UP   6 UP
    #   A  3 UL   4 LF   8 LF   7 LF   6 LF  10 UL   (A/T: l = u = 7)
    #   T  2 LF   3 LF   7 LF  11 UL  10 UP   9 LF
    # Candidates: UL cells >= 8.8 are (5,4) = 11, (4,6) = 10 and (3,3) = 9.  The
    # sweep emits (5,4) and marks (4,3) [LEFT]; emits (4,6) and marks (3,5),
    # (3,4) [UP, UP]; (4,3) and (3,4) both negate (3,3), so (3,3) is visited
    # marked (-9 < 8.8) and is not a starting point; (1,1) gets STOP.
    # Walk, ascending: (4,6) = 10 first: A/A, -/C, -/T, G/G, C/C, A/A (STOP)
    # -> ACG--A / |||  | / ACGTCA.  Then (5,4) = 11: T/T, A/-, then (3,3) is
    # IN_ALIGNMENT: suppressed.  The best hit of the pair is not shown.
    dict(name="L3 merging paths, marked candidate, suppressed walk", q="ACGAT", t="ACGTCA", m=("dna", 3.0, -3.0),
         G=-1.0, E=None, ll=0.8, mx=11.0,
         d_score=[[UL, UP, UP, NO, NO, UL], [LF, UL, UP, UP, UP, LF], [LF, LF, UL, UP, UP, UP],
                  [UL, LF, LF, LF, LF, UL], [LF, LF, LF, UL, UP, LF]],
         d_after=[[ST, UP, UP, NO, NO, UL], [LF, UL, UP, UP, UP, LF], [LF, LF, UL, UP, UP, UP],
                  [UL, LF, LF, LF, LF, UL], [LF, LF, LF, UL, UP, LF]],
         sps=[(4, 3, 11.0), (3, 5, 10.0)],
         walks=[(True, 0, 3, 0, 5, 4, 0, "ACG--A", "|||  |", "ACGTCA"), (False,)]),
    # L4 -- q XXXXXXXACGTA (X mismatches every base: rows 1-7 stay 0), t ACGTA;
    # +2/-1, gap -1.  Rows 8-12 hold ACGTA_SELF.  Start (12,5) = 10 in block
    # (1, 0); the diagonal reaches (8,1), whose predecessor lies in column 0:
    # STOP at (8,1).  Walk crosses the block edge: s 7..11, t 0..4.
    dict(name="L4 path to column 0 across an 8x8 block edge", q="XXXXXXXACGTA", t="ACGTA", m=("dna", 2.0, -1.0),
         G=-1.0, E=None, ll=1.0, mx=10.0,
         d_score=_pad(ACGTA_SELF, top_rows=7, width=5),
         d_after=_pad([[ST] + ACGTA_SELF[0][1:]] + ACGTA_SELF[1:], top_rows=7, width=5),
         sps=[(11, 4, 10.0)],
         walks=[(True, 7, 11, 0, 4, 5, 0, "ACGTA", "|||||", "ACGTA")]),
    # L4T -- the same with the X run on the target: columns 8-12 hold
    # ACGTA_SELF, the path reaches row 0 at (1,8): STOP there.
    dict(name="L4T path to row 0 across an 8x8 block edge", q="ACGTA", t="XXXXXXXACGTA", m=("dna", 2.0, -1.0),
         G=-1.0, E=None, ll=1.0, mx=10.0,
         d_score=_pad(ACGTA_SELF, left_cols=7),
         d_after=_pad([[ST] + ACGTA_SELF[0][1:]] + ACGTA_SELF[1:], left_cols=7),
         sps=[(4, 11, 10.0)],
         walks=[(True, 0, 4, 7, 11, 5, 0, "ACGTA", "|||||", "ACGTA")]),
    # L5 -- q ACGAT, t ACGTA; +3/-3, gap -1.  H:
    #        A      C      G      T      A
    #   A  3 UL   2 UP   1 UP   0 NO   3 UL
    #   C  2 LF   6 UL   5 UP   4 UP   3 UP
    #   G  1 LF   5 LF   9 UL   8 UP   7 UP
    #   A  3 UL   4 LF   8 LF   7 LF  11 UL
    #   T  2 LF   3 LF   7 LF  11 UL  10 LF
    # Two equal maxima, (4,5) and (5,4), whose paths merge at (3,3).  Both
    # are starting points; the one whose walk is shown is the one emitted
    # first -- in the reference a race of concurrent work-items (equal-score
    # order is parity unpinned), here the canonical order, which takes work-item
    # (1,0) before (0,1): (5,4) walks first and is shown, (4,5) is suppressed.
    dict(name="L5 equal maxima merging (canonical tie order)", q="ACGAT", t="ACGTA", m=("dna", 3.0, -3.0), G=-1.0,
         E=None, ll=1.0, mx=11.0,
         d_score=[[UL, UP, UP, NO, UL], [LF, UL, UP, UP, UP], [LF, LF, UL, UP, UP], [UL, LF, LF, LF, UL],
                  [LF, LF, LF, UL, LF]],
         d_after=[[ST, UP, UP, NO, UL], [LF, UL, UP, UP, UP], [LF, LF, UL, UP, UP], [UL, LF, LF, LF, UL],
                  [LF, LF, LF, UL, LF]],
         sps=[(4, 3, 11.0), (3, 4, 11.0)],
         walks=[(True, 0, 4, 0, 3, 4, 0, "ACGAT", "||| |", "ACG-T"), (False,)]),
    # A1 -- q ACGT, t ACCGT; +5/-4, gap open -3, extension -1.  (M, I, J; c):
    #   x1 A: (5,-,-;5) M | (0,2,-;2) I | (0,1,-;1) I | (0,0,-;0) NO | (0,-,-;0) NO
    #   x2 C: (0,-,2;2) J | (10,-,-;10) M | (7,7,-;7) M (tie M = I: M) | (0,6,-;6) I | (0,5,-;5) I
    #   x3 G: (0,-,1;1) J | (0,-,7;7) J | (6,-,4;6) M | (12,3,-;12) M | (2,9,-;9) I
    #   x4 T: (0,-,0;0) NO | (0,-,6;6) J | (3,-,3;3) M (tie M = J: M) | (2,0,9;9) J | (17,-,-;17) M
    # ("-" = -1e10).  Start (4,5) = 17; M chain (3,4) -> (2,3); the diagonal
    # predecessor of (2,3) is (1,2) with M = 0 (its c = 2 comes from I): STOP.
    # Every visited byte becomes byte & 3 (NONE -> 0).  The walk stops at the
    # STOP: CGT / ||| / CGT, s 1..3, t 2..4, while the starting point keeps 17.
    dict(name="A1 affine ties M>I and M>J, STOP on M = 0 under I > 0", q="ACGT", t="ACCGT", m=("dna", 5.0, -4.0),
         G=-3.0, E=-1.0, ll=1.0, mx=17.0,
         d_score=[[M_, I_, I_, NO, NO], [J_, M_, M_, I_, I_], [J_, J_, M_, M_, I_], [NO, J_, M_, J_, M_]],
         d_after=[[1, 2, 2, 0, 0], [3, 1, ST, 2, 2], [3, 3, 1, 1, 2], [0, 3, 1, 3, 1]],
         sps=[(3, 4, 17.0)],
         walks=[(True, 1, 3, 2, 4, 3, 0, "CGT", "|||", "CGT")]),
    # A2 -- q ACGT, t AGT; diagonal of ACGT +5, S(A,C) = S(C,A) = +1, else -4;
    # gap open -2, extension -1.  (M, I, J; c):
    #   x1 A: (5,-,-;5) M | (0,3,-;3) I | (0,2,-;2) I
    #   x2 C: (1,-,3;3) J | (1,-,-;1) M | (0,-,-;0) NO
    #   x3 G: (0,-,2;2) J | (8,-,-;8) M (from J(2,1) + 5) | (0,6,-;6) I
    #   x4 T: (0,-,1;1) J | (0,-,6;6) J | (13,-,-;13) M
    # Start (4,3) = 13; chain (3,2) -> (2,1), whose M = 1 is negated although
    # its byte is J: the chain stops there and (2,1) keeps code 3.  The walk
    # follows code 3 (C) past the chain and a plain main-matrix code at (1,1)
    # until the coordinates leave the matrix: ACGT / | || / A-GT.
    dict(name="A2 affine chain ends on a J cell, walk follows code 3", q="ACGT", t="AGT",
         m=("custom", {("A", "C"): 1.0}), G=-2.0, E=-1.0, ll=1.0, mx=13.0,
         d_score=[[M_, I_, I_], [J_, M_, NO], [J_, M_, I_], [J_, J_, M_]],
         d_after=[[1, 2, 2], [3, 1, 0], [3, 1, 2], [3, 3, 1]],
         sps=[(3, 2, 13.0)],
         walks=[(True, 0, 3, 0, 2, 3, 0, "ACGT", "| ||", "A-GT")]),
    # A2T -- A2 with query and target exchanged (the matrix is symmetric, so
    # M transposes and I and J exchange roles): the walk follows code 2 (B).
    dict(name="A2T affine chain ends on an I cell, walk follows code 2", q="AGT", t="ACGT",
         m=("custom", {("A", "C"): 1.0}), G=-2.0, E=-1.0, ll=1.0, mx=13.0,
         d_score=[[M_, I_, I_, I_], [J_, M_, M_, I_], [J_, NO, J_, M_]],
         d_after=[[1, 2, 2, 2], [3, 1, 1, 2], [3, 0, 3, 1]],
         sps=[(2, 3, 13.0)],
         walks=[(True, 0, 2, 0, 3, 3, 0, "A-GT", "| ||", "ACGT")]),
    # A3 -- A2 behind a first row (W) and column (Y) that mismatch everything:
    # they score 0 (NONE) and act as the old boundary, so rows/cols 2.. are A2
    # shifted by one.  The walk now reaches (1,1), whose byte & 3 is 0, and
    # ends there without emitting: s 1..4, t 1..3.
    dict(name="A3 affine walk ends on a code-0 cell", q="WACGT", t="YAGT", m=("custom", {("A", "C"): 1.0}),
         G=-2.0, E=-1.0, ll=1.0, mx=13.0,
         d_score=[[NO] * 4] + [[NO] + r for r in [[M_, I_, I_], [J_, M_, NO], [J_, M_, I_], [J_, J_, M_]]],
         d_after=[[0] * 4] + [[0] + r for r in [[1, 2, 2], [3, 1, 0], [3, 1, 2], [3, 3, 1]]],
         sps=[(4, 3, 13.0)],
         walks=[(True, 1, 4, 1, 3, 3, 0, "ACGT", "| ||", "A-GT")]),
]


def params(O, c, minimum=1.0):
    if c["m"][0] == "dna":
        mat = O.dna_rna_matrix(c["m"][1], c["m"][2], -1.0, 0.0)
    else:
        mat = custom_matrix(extra=c["m"][1])
    return O.Params(matrix=mat, gap=c["G"], gap_ext=c["E"], lower_limit=c["ll"], minimum=minimum, filter_factor=0.0,
                    highest=5.0)


def check_case(O, c, L=None):
    """Returns the list of mismatches between the oracle (library L) and case c."""
    bad = []
    q, t = c["q"], c["t"]
    n, m = len(q), len(t)
    # scoring directions: the traceback gate closed (maximum < MINIMUM_SCORE)
    closed = O.sched_batch([q], [t], params(O, c, minimum=1e9), L=L)
    if float(closed.maxima[0, 0]) != c["mx"]:
        bad.append("maximum %r" % float(closed.maxima[0, 0]))
    if closed.directions[0, 0, :n, :m].tolist() != c["d_score"]:
        bad.append("scoring directions %r" % closed.directions[0, 0, :n, :m].tolist())
    p = params(O, c)
    br = O.sched_batch([q], [t], p, L=L)
    if br.directions[0, 0, :n, :m].tolist() != c["d_after"]:
        bad.append("directions after the traceback %r" % br.directions[0, 0, :n, :m].tolist())
    sps = [(s.blockX * 8 + s.valueX, s.blockY * 8 + s.valueY, float(s.score)) for s in br.starting_points]
    if sps != c["sps"]:
        bad.append("starting points %r" % sps)
    if any(float(s.maxScore) != c["mx"] or float(s.posScore) != 0.0 for s in br.starting_points):
        bad.append("maxScore / posScore")
    # the serial restatement on the same case
    smx, sd, ssp = O.serial_pair(q, t, p, L=L)
    if smx != c["mx"] or sd.tolist() != c["d_after"] or \
            [(s.blockX * 8 + s.valueX, s.blockY * 8 + s.valueY, float(s.score)) for s in ssp] != c["sps"]:
        bad.append("serial restatement")
    for walk_name, walk in (("C walk", lambda d: O.walk_c(d, br.starting_points, [q], [t], ["q"], ["t"], 1.0, L=L)),
                            ("Python walk", lambda d: O.walk_py(d, br.starting_points, [q], [t], ["q"], ["t"], 1.0))):
        ws = walk(br.directions.copy())
        got = []
        for w in ws:
            if w.shown:
                got.append((True, w.s_start, w.s_end, w.t_start, w.t_end, w.matches, w.mismatches, w.sequence_match,
                            w.alignment, w.target_match))
            else:
                got.append((False,))
        if got != c["walks"]:
            bad.append("%s %r" % (walk_name, got))
    return bad


@pytest.mark.parametrize("case", CASES, ids=[c["name"].split()[0] for c in CASES])
def test_golden_case(oracle, case):
    assert check_case(oracle, case) == []


def test_golden_cases_are_consistent():
    """The literals agree with themselves: STOP cells only where a code was, walk strings of equal length."""
    for c in CASES:
        for w in c["walks"]:
            if w[0]:
                assert len(w[7]) == len(w[8]) == len(w[9])
                assert w[7].replace("-", "") == c["q"][w[1]:w[2] + 1]
                assert w[9].replace("-", "") == c["t"][w[3]:w[4] + 1]
                assert w[5] == w[8].count("|") and w[6] == w[8].count(".")


# --------------------------------------------------------------------------
# Linear invariant: a shown hit's strings re-score to its starting-point score
# --------------------------------------------------------------------------

def rescore_linear(sm, tm, S, G):
    return sum(G if a == "-" or b == "-" else S(a, b) for a, b in zip(sm, tm))


def test_linear_hits_rescore_to_their_score(oracle):
    """Linear gaps: a shown walk runs from its start to the STOP of a marked path, so
    the sum over its columns (S per pair, G per gap column) is the starting point's
    score; that score never exceeds the pair maximum, which the brute force pins."""
    import numpy as np
    from tests._parity import DNA, planted_set
    O = oracle
    rng = np.random.default_rng(41)
    mat = O.dna_rna_matrix(2.0, -3.0, -1.0, 0.0)
    S = lambda a, b: float(mat[ord(a) - 65][ord(b) - 65])
    for ll, G in ((1.0, -2.0), (0.5, -1.0), (0.7, -3.0)):
        qs, ts = planted_set(rng, 5, 4, (20, 90), (10, 60), DNA, plant=0.9, sub=0.08, indel=0.05)
        p = O.Params(matrix=mat, gap=G, lower_limit=ll, minimum=4.0, filter_factor=0.0, highest=2.0)
        br = O.sched_batch(qs, ts, p)
        ws = O.walk_c(br.directions, br.starting_points, qs, ts, ["q%d" % i for i in range(5)],
                      ["t%d" % i for i in range(4)], p.minimum)
        shown = [w for w in ws if w.shown]
        assert len(shown) >= 5
        for w in shown:
            assert rescore_linear(w.sequence_match, w.target_match, S, G) == w.score
            assert w.score <= float(br.maxima[w.qi, w.ti])


# --------------------------------------------------------------------------
# Planted mistakes: each changes one rule of sw_oracle.c (in every place the
# rule is written); the golden cases must notice every one of them
# --------------------------------------------------------------------------

MUTANTS = [
    ("linear tie UP before LEFT",
     "    if (cur < lS) { cur = lS; d = DIR_LEFT; }\n    if (cur < uS) { cur = uS; d = DIR_UP; }\n",
     "    if (cur < uS) { cur = uS; d = DIR_UP; }\n    if (cur < lS) { cur = lS; d = DIR_LEFT; }\n"),
    ("linear tie UL first",
     "    float cur = 0.0f;\n    uint8_t d = DIR_NONE;\n    if (cur < lS)",
     "    float cur = 0.0f;\n    uint8_t d = DIR_NONE;\n    if (cur < ulS) { cur = ulS; d = DIR_UL; }\n    if (cur < lS)"),
    ("linear STOP not written",
     "        if (score == 0.0f) { direction = DIR_STOP; DIRC(S, q, t, aXM1, aYM1) = direction; }\n"
     "        else MAT(S, S->H, q, t, aIDx - 1, aIDy - 1) = as_neg(score);\n",
     "        if (score == 0.0f) { direction = DIR_STOP; }\n"
     "        else MAT(S, S->H, q, t, aIDx - 1, aIDy - 1) = as_neg(score);\n"),
    ("linear marked candidates emitted",
     "    if (direction == DIR_UL && score >= ll_max && score >= mp) {",
     "    if (direction == DIR_UL && fabsf(score) >= ll_max && fabsf(score) >= mp) {"),
    ("affine tie I before M",
     "        if (cur == MAT(S, S->H, q, t, aIDx, aIDy)) d = d;\n"
     "        else if (cur == MAT(S, S->I, q, t, aIDx, aIDy)) d = di;\n",
     "        if (cur == MAT(S, S->I, q, t, aIDx, aIDy)) d = di;\n"
     "        else if (cur == MAT(S, S->H, q, t, aIDx, aIDy)) d = d;\n"),
    ("affine STOP tested on max(M, I, J)",
     "            score = MAT(S, S->H, q, t, aIDx - 1, aIDy - 1);\n            if (score == 0.0f) direction = DIR_STOP;\n",
     "            score = MAT(S, S->H, q, t, aIDx - 1, aIDy - 1);\n"
     "            if (fmaxf(score, fmaxf(MAT(S, S->I, q, t, aIDx - 1, aIDy - 1), MAT(S, S->J, q, t, aIDx - 1, aIDy - 1)))"
     " == 0.0f) direction = DIR_STOP;\n"),
    ("affine I and J source bits exchanged",
     "    if (ci < mI) { ci = mI; di = DIR_UP | I_MATRIX; }\n    if (ci < mM) { ci = mM; di = DIR_UP | I_MATRIX; }\n",
     "    if (ci < mI) { ci = mI; di = DIR_LEFT | J_MATRIX; }\n    if (ci < mM) { ci = mM; di = DIR_LEFT | J_MATRIX; }\n"),
    ("affine gap open G + E",
     "    mM = p->gap + MAT(S, S->H, q, t, aIDx, aYM1);\n",
     "    mM = p->gap + p->gap_ext + MAT(S, S->H, q, t, aIDx, aYM1);\n"),
    ("walk in descending score order",
     "    if (x->rec.score != y->rec.score) return x->rec.score < y->rec.score ? -1 : 1;\n",
     "    if (x->rec.score != y->rec.score) return x->rec.score > y->rec.score ? -1 : 1;\n"),
    ("walk does not emit the STOP cell",
     "            } else if (direction == DIR_UL || direction == DIR_STOP) {\n",
     "            } else if (direction == DIR_UL) {\n"),
    ("walk ignores IN_ALIGNMENT",
     "            if (direction == DIR_IN_ALIGNMENT) {\n                show = 0;\n",
     "            if (direction == DIR_IN_ALIGNMENT) {\n                bx = -1;\n"),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(ORACLE_C).read()
    d = tmp_path_factory.mktemp("mutants")
    out = []
    for k, (name, old, new) in enumerate(MUTANTS):
        # every occurrence: a mistake shared by both restatements must be caught too
        assert src.count(old) >= 1, "mutation site not found: " + name
        c = d / ("m%02d.c" % k)
        so = d / ("m%02d.so" % k)
        c.write_text(src.replace(old, new))
        subprocess.run(["gcc", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", str(c),
                        "-o", str(so), "-lm"], check=True)
        out.append((name, str(so)))
    return out


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_golden_cases_catch_planted_mistake(oracle, mutant_libs, k):
    name, path = mutant_libs[k]
    L = oracle.bind(path)
    failing = [c["name"] for c in CASES if check_case(oracle, c, L=L)]
    assert failing, "no golden case notices the planted mistake: " + name
