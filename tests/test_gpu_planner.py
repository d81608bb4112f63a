"""The planner's all-vs-all path against its general path.

A pair list that is every query with every target in query-major order (what
the engine builds) is planned without per-pair lists: both orientations are
costed from one sort of each side's lengths.  The same pairs in target-major
order take the general path, whose per-column row order is the same, so the
two plans must agree on every routing figure (pw_plan_info), and the results
must agree pair for pair: maxima, starting points and walks.
"""
import zlib
from collections import defaultdict

import numpy as np
import pytest

from tests._parity import AMINO, DNA, library_rows, planted_set, rand_seq
from tests.test_gpu_parity import blosum_params, dna_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


def pairs(nq, nt, query_major):
    if query_major:
        return np.repeat(np.arange(nq, dtype=np.int32), nt), np.tile(np.arange(nt, dtype=np.int32), nq)
    return np.tile(np.arange(nq, dtype=np.int32), nt), np.repeat(np.arange(nt, dtype=np.int32), nq)


def run(dev, O, qs, ts, p, query_major):
    from pypaswas_amd._native import PackedSeqs
    pq, pt = pairs(len(qs), len(ts), query_major)
    mss = np.array([O.min_start_score(p, len(qs[i]), len(ts[j])) for i, j in zip(pq, pt)], dtype=np.float32)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    Q, T = PackedSeqs.pack(qs), PackedSeqs.pack(ts)
    plan = dev.prepare(Q, T, pq, pt, mss, None, mode=1)
    try:
        info = plan.info()
        res = dev.execute(plan, p.minimum)
    finally:
        plan.release()
    return pq, pt, info, res


def by_pair(pq, pt, res):
    """Maxima, starting points and walk records keyed by (query, target)."""
    mx = {(int(q), int(t)): int(v) for q, t, v in zip(pq, pt, res.maxima.view(np.uint32))}
    sps = defaultdict(list)
    for s in res.starting_points:
        sps[(int(s["sequence"]), int(s["target"]))].append(tuple(s.tolist()))
    hits = defaultdict(list)
    for h in res.hits:
        rows = library_rows(res, h) if h["shown"] else ("", "", "")
        hits[(int(h["query"]), int(h["target"]))].append(
            (int(h["shown"]), int(h["s_start"]), int(h["s_end"]), int(h["t_start"]), int(h["t_end"]),
             int(h["matches"]), int(h["mismatches"]), float(h["score"])) + rows)
    return mx, dict(sps), dict(hits)


CASES = [
    # name, alphabet, queries, targets, query lengths, target lengths, affine
    ("reads-x-adapters-affine", DNA, 400, 24, (60, 700), (20, 80), True),
    ("reads-x-adapters-linear", DNA, 300, 16, (60, 700), (20, 80), False),
    ("adapters-x-reads-affine", DNA, 24, 300, (20, 80), (60, 700), True),
    ("proteins", AMINO, 40, 60, (30, 900), (30, 900), True),
    ("long-linear", DNA, 3, 4, (900, 2600), (1100, 3000), False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_all_vs_all_plan_equals_general_plan(dev, oracle, case):
    name, alphabet, nq, nt, ql, tl, affine = case
    print("planner", name, flush=True)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    qs, ts = planted_set(rng, nq, nt, ql, tl, alphabet, plant=0.7)
    p = blosum_params(oracle, minimum=20.0) if alphabet is AMINO else dna_params(oracle, affine, ff=0.05)
    pq1, pt1, info1, r1 = run(dev, oracle, qs, ts, p, True)
    pq2, pt2, info2, r2 = run(dev, oracle, qs, ts, p, False)
    assert info1 == info2
    a, b = by_pair(pq1, pt1, r1), by_pair(pq2, pt2, r2)
    assert a[0] == b[0]
    assert a[1] == b[1]
    assert a[2] == b[2]
    assert r1.n_traceback_pairs == r2.n_traceback_pairs and r1.n_traceback_pairs > 0
