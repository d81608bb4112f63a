"""Results laid out on the device against results assembled on the host.

When every pair of an execution finishes on the device, its records, op
characters and the three alignment rows of each shown hit are put into pair
order there (exclusive sums over the read-back slots, one wave per pair) and
copied once into the result; pw_set_assembly(0) reads them back per pair and
assembles them on the host instead.  Both must give the same result byte for
byte: maxima, starting points, hits, ops and every shown hit's rows.
"""
import zlib

import numpy as np
import pytest

from tests._parity import AMINO, DNA, planted_set, rand_seq
from tests.test_gpu_configs import all_pairs, oracle_params, run, same_result
from tests.test_gpu_parity import dna_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.set_assembly(True)
    d.close()


def shown_rows(res):
    """The alignment rows of the shown hits, in hit order."""
    al = bytes(res.aligned)
    h = res.hits
    out = []
    for off, n in zip(h["ops_offset"][h["shown"] == 1], h["length"][h["shown"] == 1]):
        out.append(al[3 * int(off):3 * int(off) + 3 * int(n)])
    return out


def both(dev, fn):
    dev.set_assembly(True)
    a = fn()
    dev.set_assembly(False)
    try:
        b = fn()
    finally:
        dev.set_assembly(True)
    return a, b


def test_config3_sample_device_layout_equals_host(dev, oracle):
    """200 x 2,000 proteins: 46 k gated pairs, several device rounds."""
    from pypaswas_amd import workloads as W
    qr, tr, st = W.config3(n_queries=200, n_targets=2000)
    q, t = W.reference_order(qr, tr, False)
    print("config3 200 x 2000, layout on the device vs on the host", flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    a, b = both(dev, lambda: run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=1))
    same_result(a, b)
    assert len(a.hits) > 40000
    assert shown_rows(a) == shown_rows(b)


@pytest.mark.parametrize("affine", [False, True])
def test_self_comparison_hidden_hits(dev, oracle, affine):
    """Reads against themselves: the same-id pairs' hits are hidden (no rows)."""
    rng = np.random.default_rng(zlib.crc32(b"assembly-self") + int(affine))
    seqs, _ = planted_set(rng, 60, 1, (80, 900), (20, 40), DNA, plant=0.0)
    ids = ["r%d" % i for i in range(len(seqs))]
    p = dna_params(oracle, affine, ff=0.05)
    pq, pt = all_pairs(len(seqs), len(seqs))
    print("self comparison, affine", affine, flush=True)
    a, b = both(dev, lambda: run(dev, oracle, p, seqs, seqs, pq, pt, ids, ids, mode=1))
    same_result(a, b)
    assert (a.hits["shown"] == 0).sum() >= len(seqs)
    assert shown_rows(a) == shown_rows(b)


def test_nothing_gated_and_graph_replays(dev, oracle):
    """No pair passes the gate (empty result), then graph replays of a plan."""
    from pypaswas_amd._native import PackedSeqs
    rng = np.random.default_rng(7)
    qs = [rand_seq(rng, 50, DNA) for _ in range(4)]
    ts = [rand_seq(rng, 60, DNA) for _ in range(5)]
    p = dna_params(oracle, False, ff=0.05)
    pq, pt = all_pairs(len(qs), len(ts))
    mss = np.full(len(pq), 1e9, dtype=np.float32)  # no start can reach it
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, 0.0, p.lower_limit, p.minimum)
    r = dev.align_pairs(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mss, None, p.minimum)
    assert r.n_traceback_pairs == 0 and len(r.starting_points) == 0 and len(r.hits) == 0
    qs, ts = planted_set(rng, 30, 20, (60, 500), (60, 500), AMINO, plant=0.6)
    from tests.test_gpu_parity import blosum_params
    p = blosum_params(oracle, minimum=20.0)
    pq, pt = all_pairs(len(qs), len(ts))
    mss = np.array([oracle.min_start_score(p, len(qs[i]), len(ts[j])) for i, j in zip(pq, pt)], dtype=np.float32)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    plan = dev.prepare(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mss, None, mode=1)
    try:
        dev.set_assembly(False)
        host = dev.execute(plan, p.minimum)
        dev.set_assembly(True)
        for k in range(3):
            g = dev.execute(plan, p.minimum, graph=True)
            same_result(host, g)
            assert shown_rows(host) == shown_rows(g)
            e = dev.execute(plan, p.minimum)
            same_result(host, e)
        assert host.n_traceback_pairs > 50
    finally:
        dev.set_assembly(True)
        plan.release()


def test_areas_grow_between_device_rounds(dev, oracle):
    """A context whose first areas are tiny (one candidate slot per pair, a
    64-byte ops area): the candidate, record and ops areas grow between device
    rounds, the records written before each growth carried over."""
    from pypaswas_amd import workloads as W
    from pypaswas_amd._native import Device
    qr, tr, st = W.config3(n_queries=60, n_targets=1000)
    q, t = W.reference_order(qr, tr, False)
    print("config3 60 x 1000, areas growing between device rounds", flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    roomy = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=1)
    d = Device(0)
    try:
        d.set_limits(candidates_per_pair=1, ops_bytes=64, grow=True)
        small = run(d, oracle, p, qs, ts, pq, pt, qids, tids, mode=1)
        same_result(roomy, small)
        assert shown_rows(roomy) == shown_rows(small)
        assert len(roomy.hits) > 5000
        del small
    finally:
        d.close()
