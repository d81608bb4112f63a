"""Checkpointed long pairs: the block-wise direction pass against the oracle.

Linear-gap pairs whose query and target both hold at least 1,024 residues are
scored with the query along the lanes at 8 residues per lane, one wave per
512-row stripe, and every 512 steps a stripe saves the wave's state and the
largest value of the 512 steps before.  Their direction pass recomputes only
the blocks of 512 steps up and left of the candidate starts (a path only moves
up and left), each block from its checkpoint and the boundary row the scoring
pass left above its stripe.  These tests compare such pairs with the oracle bit
for bit (maxima, starting points, walks):

* lengths at and around the block and stripe edges, with the shared region at
  the start, in the middle and at the end of both sequences (candidates in the
  first and in the last block and stripe);
* a low lower limit, so that many cells in many blocks are candidate starts
  and paths merge;
* the batched plan (checkpointed) against the plain launch (no checkpoints),
  against graph replays, and against tail rounds (every pair deferred: the
  whole matrix recomputed in one unit).
"""
import numpy as np
import pytest

from tests._parity import DNA, compare, hip_all_vs_all, mutate, plan_info, rand_seq
from tests.test_gpu_parity import dna_params

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


def edge_set(rng, qlens, tlens, region=300, where=("start", "middle", "end")):
    """Queries and targets of the given lengths; target k carries a mutated
    copy of a region of query k (k < min(len)), placed at the start, the middle
    or the end of both sequences in turn."""
    qs = [rand_seq(rng, n, DNA) for n in qlens]
    ts = [list(rand_seq(rng, n, DNA)) for n in tlens]
    for k in range(min(len(qs), len(ts))):
        w = where[k % len(where)]
        q = qs[k]
        if w == "start":
            a = 0
        elif w == "end":
            a = len(q) - region
        else:
            a = (len(q) - region) // 2
        piece = mutate(rng, q[a:a + region], 0.08, 0.02, DNA)[:region]
        t = ts[k]
        off = 0 if w == "start" else (len(t) - len(piece) if w == "end" else (len(t) - len(piece)) // 2)
        t[off:off + len(piece)] = list(piece)
    return qs, ["".join(t) for t in ts]


def test_block_and_stripe_edges_vs_oracle(dev, oracle):
    rng = np.random.default_rng(1701)
    qlens = [1024, 1535, 1537, 2048, 1100]
    tlens = [1024, 1089, 1536, 2049]
    qs, ts = edge_set(rng, qlens, tlens)
    p = dna_params(oracle, False, ff=0.02)
    info = plan_info(dev, oracle, qs, ts, p)
    assert info["checkpointed_pairs"] == len(qs) * len(ts)
    print("checkpoint edges", qlens, tlens, flush=True)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs >= 4


def test_many_candidates_vs_oracle(dev, oracle):
    """Lower limit 0.4: every diagonal cell of at least 0.4 x the pair maximum
    starts a path, in many blocks of several stripes; two copies of the shared
    region per target make merging and suppressed paths."""
    rng = np.random.default_rng(1702)
    qs = [rand_seq(rng, n, DNA) for n in (2600, 3100)]
    ts = []
    for k, n in enumerate((2900, 2300)):
        t = list(rand_seq(rng, n, DNA))
        q = qs[k]
        for off, a in ((150, 400), (n - 700, 1500)):
            piece = mutate(rng, q[a:a + 500], 0.06, 0.02, DNA)[:500]
            t[off:off + len(piece)] = list(piece)
        ts.append("".join(t))
    p = dna_params(oracle, False, ll=0.4, ff=0.02)
    print("checkpoint many candidates", flush=True)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert len(res.starting_points) > 20


def test_batched_equals_plain_graph_and_tail_rounds(dev, oracle):
    from pypaswas_amd._native import Device, PackedSeqs
    rng = np.random.default_rng(1703)
    qs, ts = edge_set(rng, [1800, 2500, 1030], [1300, 2700, 1500], region=400)
    qs.append(rand_seq(rng, 700, DNA))  # a pair below the checkpoint lengths in the same plan
    p = dna_params(oracle, False, ff=0.02)
    print("checkpoint batched vs plain, graph, tail rounds", flush=True)
    batched = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, batched)
    plain = hip_all_vs_all(dev, oracle, qs, ts, p, mode=0)
    same_result(batched, plain)
    # graph replays of the checkpointed plan
    N, T = len(qs), len(ts)
    pq = np.repeat(np.arange(N, dtype=np.int32), T)
    pt = np.tile(np.arange(T, dtype=np.int32), N)
    mss = np.array([oracle.min_start_score(p, len(qs[i]), len(ts[j])) for i, j in zip(pq, pt)], dtype=np.float32)
    plan = dev.prepare(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mss, None, mode=1)
    try:
        assert plan.info()["checkpointed_pairs"] == 9
        for k in range(3):
            same_result(batched, dev.execute(plan, p.minimum, graph=True))
            same_result(batched, dev.execute(plan, p.minimum))
    finally:
        plan.release()
    # tail rounds: a direction area too small for any gated pair at first
    d = Device(0)
    try:
        d.set_limits(max_direction_bytes=1 << 16, candidates_per_pair=1, ops_bytes=64, grow=False)
        tight = hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        same_result(batched, tight)
    finally:
        d.close()
