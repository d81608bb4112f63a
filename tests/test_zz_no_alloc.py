"""pw_score_pairs_noalloc: the score pass with no device allocation, against pw_score_pairs.

The same job through both entry points (through the C ABI), bit for bit, two
calls in a row with the second on a smaller input (the static scratch keeps
the first call's contents), in both launch modes, for linear and affine gaps,
with wave-per-pair tiles of split long pairs, lane-per-pair jobs (also split
into chunk items) and short pairs; and a job larger than the static scratch
is rejected.
"""
import numpy as np
import pytest

from tests._parity import DNA, mutate, planted_set, rand_seq

pytestmark = pytest.mark.gpu


def jobs(seed):
    rng = np.random.default_rng(seed)
    # 130 reads of similar length: two well-filled 64-lane jobs per adapter
    qs, ts = planted_set(rng, 130, 3, (200, 260), (20, 90), DNA, plant=0.7)
    # long pairs: several stripes and >= 256 columns, split over waves
    for _ in range(3):
        q = rand_seq(rng, int(rng.integers(1500, 2600)), DNA)
        qs.append(q)
        ts.append(mutate(rng, q[200:1400], 0.1, 0.02))
    return qs, ts


def pairs(nq, nt):
    return np.repeat(np.arange(nq, dtype=np.int32), nt), np.tile(np.arange(nt, dtype=np.int32), nq)


@pytest.fixture(scope="module", params=["default", "lane-per-pair"])
def dev(request):
    from pypaswas_amd._native import Device
    d = Device(0)
    d.routing_name = request.param
    if request.param == "lane-per-pair":
        # keep every lane-per-pair job whose lanes are busy enough (small plans
        # otherwise go to the wave-per-pair kernel)
        d.set_routing(-1, 1e9)
    yield d
    d.close()


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("affine", [False, True])
def test_noalloc_equals_score_pairs(dev, oracle, affine, mode):
    from pypaswas_amd._native import PackedSeqs
    print("noalloc affine", affine, "mode", mode, flush=True)
    p = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0,
                      gap_ext=(-1.0 if affine else None), minimum=10.0, highest=1.0)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext or 0.0, p.lower_limit, p.minimum)
    qs, ts = jobs(31 + affine)
    small_q, small_t = qs[:40], ts[:2]
    if mode == 1:
        # the first job has split long pairs (more wave-per-pair items than
        # such pairs) and, under the lane-per-pair routing, lane-per-pair jobs
        pq, pt = pairs(len(qs), len(ts))
        plan = dev.prepare(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, None, None, mode=1)
        info = plan.info()
        plan.release()
        assert info["wave_per_pair_items"] > 0 and info["wave_per_pair_cells"] > 0
        if dev.routing_name == "lane-per-pair":
            assert info["lane_per_pair_items"] > 0
    for q, t in ((qs, ts), (small_q, small_t)):
        qp, tp = PackedSeqs.pack(q), PackedSeqs.pack(t)
        pq, pt = pairs(len(q), len(t))
        want, _ = dev.score_pairs(qp, tp, pq, pt, mode=mode)
        got, ms = dev.score_pairs_noalloc(qp, tp, pq, pt, mode=mode)
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert ms > 0.0 and got.max() > 0


def test_noalloc_rejects_a_job_larger_than_the_scratch(dev, oracle):
    from pypaswas_amd._native import NOALLOC_SCRATCH_BYTES, PackedSeqs
    from pypaswas_amd.Core.Exceptions import InvalidOptionException
    p = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, minimum=10.0, highest=1.0)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, 0.0, p.lower_limit, p.minimum)
    big = "ACGT" * (NOALLOC_SCRATCH_BYTES // 4 + 1024)
    qp, tp = PackedSeqs.pack(["ACGTACGTAC"]), PackedSeqs.pack([big])
    one = np.zeros(1, np.int32)
    with pytest.raises(InvalidOptionException, match="static scratch"):
        dev.score_pairs_noalloc(qp, tp, one, one)
    # the context stays usable
    small = PackedSeqs.pack(["ACGTACGTACGT"])
    got, _ = dev.score_pairs_noalloc(qp, small, one, one)
    want, _ = dev.score_pairs(qp, small, one, one)
    assert np.array_equal(got, want)
