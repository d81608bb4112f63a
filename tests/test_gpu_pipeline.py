"""The device-resident execution: graph replays and host-driven tail rounds give the same results.

* a plan executed through its captured HIP graph, several times with eager
  executions and host work in between, returns exactly the eager result;
* work areas far too small for the job (pw_set_limits) push every gated pair
  through the tail rounds; results still match the oracle bit for bit.
"""
import zlib

import numpy as np
import pytest

from tests._parity import DNA, AMINO, compare, hip_all_vs_all, planted_set
from tests.test_gpu_parity import blosum_params, dna_params, overwrite_set

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


def same_result(a, b):
    assert np.array_equal(a.maxima.view(np.uint32), b.maxima.view(np.uint32))
    assert np.array_equal(a.starting_points, b.starting_points)
    assert np.array_equal(a.hits, b.hits)
    assert a.ops == b.ops
    assert a.n_traceback_pairs == b.n_traceback_pairs


def plan_for(dev, O, qs, ts, p, mode):
    from pypaswas_amd._native import PackedSeqs
    N, T = len(qs), len(ts)
    pq = np.repeat(np.arange(N, dtype=np.int32), T)
    pt = np.tile(np.arange(T, dtype=np.int32), N)
    mss = np.array([O.min_start_score(p, len(qs[i]), len(ts[j])) for i, j in zip(pq, pt)], dtype=np.float32)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    return dev.prepare(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mss, None, mode=mode)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("affine", [False, True])
def test_graph_replays_match_eager(dev, oracle, affine, mode):
    print("graph replay affine", affine, "mode", mode, flush=True)
    rng = np.random.default_rng(zlib.crc32(b"graph") + int(affine) + 2 * mode)
    # lane-per-pair jobs plus multi-stripe tile pairs, so both score streams are captured
    qs, ts = overwrite_set(rng, 70, 3, (200, 260), (20, 300), DNA)
    qs2, ts2 = planted_set(rng, 3, 2, (600, 900), (200, 500), DNA, plant=1.0)
    qs, ts = qs + qs2, ts + ts2
    p = dna_params(oracle, affine, minimum=8.0)
    plan = plan_for(dev, oracle, qs, ts, p, mode)
    try:
        eager = dev.execute(plan, p.minimum)
        compare(oracle, qs, ts, p, eager)
        for k in range(3):
            g = dev.execute(plan, p.minimum, graph=True)
            same_result(g, eager)
            e2 = dev.execute(plan, p.minimum)
            same_result(e2, eager)
    finally:
        plan.release()


def test_graph_protein(dev, oracle):
    rng = np.random.default_rng(5)
    qs, ts = planted_set(rng, 6, 5, (30, 500), (30, 500), AMINO, plant=0.8, sub=0.5, indel=0.05)
    p = blosum_params(oracle)
    plan = plan_for(dev, oracle, qs, ts, p, 1)
    try:
        eager = dev.execute(plan, p.minimum)
        compare(oracle, qs, ts, p, eager)
        for _ in range(2):
            same_result(dev.execute(plan, p.minimum, graph=True), eager)
    finally:
        plan.release()


@pytest.fixture()
def tight_dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    # a few KiB of direction area, one candidate slot per pair, 64 bytes of ops, no growth
    d.set_limits(max_direction_bytes=1 << 16, candidates_per_pair=1, ops_bytes=64, grow=False)
    yield d
    d.close()


@pytest.mark.parametrize("affine", [False, True])
def test_tail_rounds_match_oracle(tight_dev, dev, oracle, affine):
    print("tail rounds affine", affine, flush=True)
    rng = np.random.default_rng(21 + int(affine))
    qs, ts = planted_set(rng, 5, 4, (60, 700), (40, 300), DNA, plant=1.0, sub=0.03)
    p = dna_params(oracle, affine, minimum=5.0, ll=0.7, ff=0.0)
    tight = hip_all_vs_all(tight_dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, tight)
    roomy = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    same_result(tight, roomy)
    assert tight.n_traceback_pairs > 0


def test_tail_rounds_graph(tight_dev, oracle):
    """A graph execution whose pairs overflow finishes them in tail rounds, every replay."""
    rng = np.random.default_rng(33)
    qs, ts = planted_set(rng, 4, 3, (100, 400), (50, 200), DNA, plant=1.0, sub=0.03)
    p = dna_params(oracle, True, minimum=5.0, ll=0.8, ff=0.0)
    plan = plan_for(tight_dev, oracle, qs, ts, p, 1)
    try:
        eager = tight_dev.execute(plan, p.minimum)
        compare(oracle, qs, ts, p, eager)
        for _ in range(2):
            same_result(tight_dev.execute(plan, p.minimum, graph=True), eager)
    finally:
        plan.release()
