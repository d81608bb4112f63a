"""Bounded device waits: a hand-over wait that gives up ends in an error, not a hang.

The waits of the stripe hand-over of split long pairs (scoring and direction
pass) and of the chunk hand-over of split lane-per-pair jobs poll a progress
counter at most pw_set_wait_limit times.  With the limit set to 0 from the
host, the first wait that is not already satisfied gives up: every wave of the
execution leaves its work item, all kernels end normally, and the call reports
PW_EDEVICE (HardwareException) with the results discarded.  The same context
then runs the same job with the default limit and returns the results it
returned before.
"""
import zlib

import numpy as np
import pytest

from tests._parity import DNA, hip_all_vs_all

pytestmark = pytest.mark.gpu


def same_result(a, b):
    assert np.array_equal(a.maxima.view(np.uint32), b.maxima.view(np.uint32))
    assert np.array_equal(a.starting_points, b.starting_points)
    assert np.array_equal(a.hits, b.hits)
    assert a.ops == b.ops


@pytest.mark.parametrize("affine", [False, True])
def test_stripe_wait_gives_up_and_context_recovers(oracle, affine):
    from pypaswas_amd._native import Device, PackedSeqs
    from pypaswas_amd.Core.Exceptions import HardwareException
    from tests.test_gpu_long import long_pairs
    from tests.test_gpu_parity import dna_params
    print("wait limit, split long pairs, affine", affine, flush=True)
    rng = np.random.default_rng(70 + affine)
    qs, ts = long_pairs(rng, 3, (2500, 4000), DNA)
    p = dna_params(oracle, affine, ff=0.05)
    d = Device(0)
    try:
        want = hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        assert want.n_traceback_pairs >= 3
        d.set_wait_limit(0)
        with pytest.raises(HardwareException, match="gave up"):
            hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        pq = np.repeat(np.arange(len(qs), dtype=np.int32), len(ts))
        pt = np.tile(np.arange(len(ts), dtype=np.int32), len(qs))
        with pytest.raises(HardwareException, match="gave up"):
            d.score_pairs(PackedSeqs.pack(qs), PackedSeqs.pack(ts), pq, pt, mode=1)
        d.set_wait_limit(-1)
        same_result(hip_all_vs_all(d, oracle, qs, ts, p, mode=1), want)
    finally:
        d.close()


def test_chunk_wait_gives_up_and_context_recovers(oracle):
    from pypaswas_amd._native import Device
    from pypaswas_amd.Core.Exceptions import HardwareException
    from tests.test_gpu_parity import dna_params, overwrite_set
    print("wait limit, split lane-per-pair jobs", flush=True)
    rng = np.random.default_rng(zlib.crc32(b"split-waits"))
    rows, cols = overwrite_set(rng, 70, 3, (200, 260), (280, 420), DNA, plant=0.8, sub=0.05)
    qs, ts = cols, rows
    p = dna_params(oracle, True, minimum=8.0)
    d = Device(0)
    try:
        # the routing of the split-job tests: jobs split into one item per chunk
        d.set_routing(1, 150.0)
        want = hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        d.set_wait_limit(0)
        with pytest.raises(HardwareException, match="gave up"):
            hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        d.set_wait_limit(-1)
        same_result(hip_all_vs_all(d, oracle, qs, ts, p, mode=1), want)
    finally:
        d.close()


def test_wait_limit_rejects_out_of_range(oracle):
    from pypaswas_amd._native import Device
    from pypaswas_amd.Core.Exceptions import InvalidOptionException
    d = Device(0)
    try:
        with pytest.raises(InvalidOptionException):
            d.set_wait_limit(1 << 40)
        d.set_wait_limit(-1)
    finally:
        d.close()
