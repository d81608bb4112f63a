"""Autotuner: measured shape tables are valid, round-trip through the ABI, and never change results."""
import zlib

import numpy as np
import pytest

from tests._parity import AMINO, DNA, compare, hip_all_vs_all, planted_set
from tests.test_gpu_parity import blosum_params, dna_params, overwrite_set

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tuned():
    from pypaswas_amd._native import Device
    from oracle import oracle as O
    d = Device(0)
    p = dna_params(O, True)
    d.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext, p.lower_limit, p.minimum)
    print("autotune affine", flush=True)
    sec = d.autotune()
    print("autotune took %.2fs" % sec, flush=True)
    yield d
    d.close()


def same_result(a, b):
    assert np.array_equal(a.maxima.view(np.uint32), b.maxima.view(np.uint32))
    assert np.array_equal(a.starting_points, b.starting_points)
    assert np.array_equal(a.hits, b.hits)
    assert a.ops == b.ops


def test_table_is_measured(tuned):
    t = tuned.tuning(True)
    assert t["valid"]
    assert np.all(t["rates"] > 0)
    assert np.all(t["group_rates"] > 0)
    assert t["group_nb_long"] in (32, 48, 64, 80, 96)
    assert not tuned.tuning(False)["valid"]  # only the affine model was tuned


def test_tuned_results_match_oracle_and_untuned(tuned, oracle):
    rng = np.random.default_rng(zlib.crc32(b"tuned"))
    qs, ts = overwrite_set(rng, 70, 3, (200, 260), (20, 400), DNA)
    qs2, ts2 = planted_set(rng, 4, 3, (100, 1500), (60, 900), DNA, plant=1.0)
    qs, ts = qs + qs2, ts + ts2
    p = dna_params(oracle, True, minimum=8.0, ff=0.05)
    res = hip_all_vs_all(tuned, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    saved = tuned.tuning(True)
    tuned.set_tuning(True, None)
    try:
        assert not tuned.tuning(True)["valid"]
        same_result(hip_all_vs_all(tuned, oracle, qs, ts, p, mode=1), res)
    finally:
        tuned.set_tuning(True, saved["rates"], saved["group_nb_long"])
    again = tuned.tuning(True)
    assert again["valid"] and np.array_equal(again["rates"], saved["rates"])
    assert again["group_nb_long"] == saved["group_nb_long"]


def test_tuned_protein(tuned, oracle):
    rng = np.random.default_rng(61)
    qs, ts = planted_set(rng, 6, 5, (30, 900), (30, 900), AMINO, plant=0.8, sub=0.5, indel=0.05)
    p = blosum_params(oracle)
    compare(oracle, qs, ts, p, hip_all_vs_all(tuned, oracle, qs, ts, p, mode=1))


def test_set_tuning_rejects_bad_table(tuned):
    from pypaswas_amd.Core.Exceptions import InvalidOptionException
    with pytest.raises(InvalidOptionException):
        tuned.set_tuning(False, np.ones((7, 7, 4), np.float32), group_nb_long=40)
    bad = np.ones((7, 7, 4), np.float32)
    bad[0, 0, 0] = np.nan
    with pytest.raises(InvalidOptionException):
        tuned.set_tuning(False, bad, group_nb_long=64)


def test_plan_tuning_keeps_results(tuned, oracle):
    from pypaswas_amd._native import PackedSeqs
    rng = np.random.default_rng(zlib.crc32(b"plan-tune"))
    qs, ts = overwrite_set(rng, 140, 4, (150, 400), (40, 900), AMINO, plant=0.7, sub=0.4)
    p = blosum_params(oracle, minimum=25.0)
    N, T = len(qs), len(ts)
    pq = np.repeat(np.arange(N, dtype=np.int32), T)
    pt = np.tile(np.arange(T, dtype=np.int32), N)
    mss = np.array([oracle.min_start_score(p, len(qs[i]), len(ts[j])) for i, j in zip(pq, pt)], dtype=np.float32)
    tuned.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext, p.lower_limit, p.minimum)
    plan = tuned.prepare(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mss, None, mode=1)
    try:
        before = tuned.execute(plan, p.minimum)
        info = tuned.tune_plan(plan)
        print("plan tuning", info, flush=True)
        # the widths plan tuning times: the default long-job width (80) and 64, 96
        assert info["score_ms"] > 0 and info["group_chunk_width"] in (64, 80, 96)
        after = tuned.execute(plan, p.minimum)
        same_result(after, before)
        same_result(tuned.execute(plan, p.minimum, graph=True), before)
        compare(oracle, qs, ts, p, after)
    finally:
        plan.release()
