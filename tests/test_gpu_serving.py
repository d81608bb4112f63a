"""Serving loop: every job's records equal those of the same job run alone, and the oracle's.

Jobs of the serving mix (short reads against adapter panels, protein pairs,
long DNA pairs) go through the continuous-batching server all at once (the
server merges them into large batches), as a Poisson stream, and through a
direction arena too small for the long jobs (which then run alone and finish
in tail rounds).  Each job's result is compared bit for bit with
pw_align_pairs on that job alone; small jobs are also compared with the oracle.
"""
import numpy as np
import pytest

from tests._parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def mix():
    from pypaswas_amd import workloads as W
    from pypaswas_amd.serving import Job
    schemes = W.serving_schemes()
    raw = W.serving_jobs(seed=15, n_jobs=120, long_range=(1500, 3000))
    jobs = [Job.all_vs_all(W.SERVING_KINDS.index(k), schemes[W.SERVING_KINDS.index(k)], q, t) for k, q, t in raw]
    return raw, jobs, schemes


def alone(dev, scheme, job):
    dev.set_scoring(scheme.subst, scheme.char_offset, scheme.gap, scheme.gap_ext, scheme.lower_limit, scheme.minimum)
    return dev.align_pairs(job.queries, job.targets, job.pq, job.pt, job.mss, job.same, job.minimum, job.sp_cap,
                           mode=1)


def same_result(a, b):
    assert np.array_equal(a.maxima.view(np.uint32), b.maxima.view(np.uint32))
    assert np.array_equal(a.starting_points, b.starting_points)
    assert np.array_equal(a.hits, b.hits)
    assert a.ops == b.ops
    assert a.n_traceback_pairs == b.n_traceback_pairs


def server_for(dev, schemes, lanes=4, direction_bytes=256 << 20):
    from pypaswas_amd.serving import Server
    s = Server(dev, lanes=lanes, direction_bytes=direction_bytes)
    for i, sc in enumerate(schemes):
        s.set_scheme(i, sc)
    return s


def check_all(dev, raw, jobs, schemes, results):
    for k, job in enumerate(jobs):
        assert results[k] is not None, k
        same_result(results[k], alone(dev, schemes[job.scheme], job))


def test_schemes_are_the_reference_matrices(oracle, mix):
    raw, jobs, schemes = mix
    dna = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0))
    assert np.array_equal(schemes[0].subst, dna.subst_f32())
    assert np.array_equal(schemes[2].subst, dna.subst_f32())
    assert np.array_equal(schemes[1].subst, oracle.Params(matrix=oracle.BLOSUM62).subst_f32())
    assert (schemes[0].gap, schemes[0].gap_ext, schemes[1].gap, schemes[1].gap_ext, schemes[2].gap_ext) == \
        (-5.0, -1.0, -11.0, -1.0, 0.0)


def test_all_at_once_equals_alone(dev, mix):
    from pypaswas_amd.serving import replay
    raw, jobs, schemes = mix
    print("serving all at once", len(jobs), "jobs", flush=True)
    srv = server_for(dev, schemes)
    try:
        results, lat, wall = replay(srv, jobs)
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(jobs)
    assert st["batches"] < len(jobs)  # jobs were merged
    check_all(dev, raw, jobs, schemes, results)


def test_poisson_stream_equals_alone(dev, mix):
    from pypaswas_amd.serving import poisson_arrivals, replay
    raw, jobs, schemes = mix
    print("serving poisson", flush=True)
    srv = server_for(dev, schemes, lanes=3)
    try:
        results, lat, wall = replay(srv, jobs, poisson_arrivals(len(jobs), 3000.0, seed=3))
    finally:
        srv.close()
    assert np.all(lat > 0)
    check_all(dev, raw, jobs, schemes, results)


def test_small_arena_equals_alone(dev, mix):
    """An arena smaller than a long job: it runs alone, its pairs finish in tail rounds."""
    from pypaswas_amd.serving import replay
    raw, jobs, schemes = mix
    print("serving small arena", flush=True)
    srv = server_for(dev, schemes, lanes=2, direction_bytes=1 << 20)
    try:
        results, lat, wall = replay(srv, jobs[:40])
    finally:
        srv.close()
    check_all(dev, raw, jobs[:40], schemes, results)


def test_small_jobs_match_oracle(dev, oracle, mix):
    from pypaswas_amd.serving import replay
    raw, jobs, schemes = mix
    pick = [k for k, (kind, q, t) in enumerate(raw) if kind != "long"][:12]
    srv = server_for(dev, schemes)
    try:
        results, lat, wall = replay(srv, [jobs[k] for k in pick])
    finally:
        srv.close()
    for r, k in zip(results, pick):
        kind, q, t = raw[k]
        sc = schemes[jobs[k].scheme]
        matrix = oracle.BLOSUM62 if kind == "protein" else oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0)
        p = oracle.Params(matrix=matrix, gap=sc.gap, gap_ext=(sc.gap_ext or None), lower_limit=sc.lower_limit,
                          minimum=sc.minimum, filter_factor=sc.filter_factor, highest=sc.highest,
                          dna=(kind != "protein"))
        compare(oracle, [s for _, s in q], [s for _, s in t], p, r, qids=[i for i, _ in q], tids=[i for i, _ in t])


def test_submit_rejects_bad_input(dev, mix):
    from pypaswas_amd.Core.Exceptions import InvalidOptionException
    from pypaswas_amd.serving import Job, Server
    raw, jobs, schemes = mix
    srv = Server(dev, lanes=1, direction_bytes=1 << 20)
    try:
        with pytest.raises(InvalidOptionException):
            srv.submit(jobs[0])  # scheme not set
        srv.set_scheme(jobs[0].scheme, schemes[jobs[0].scheme])
        bad = Job(jobs[0].scheme, jobs[0].queries, jobs[0].targets, jobs[0].pq + 100, jobs[0].pt, jobs[0].mss,
                  jobs[0].same, jobs[0].minimum, jobs[0].cells)
        with pytest.raises(InvalidOptionException):
            srv.submit(bad)
    finally:
        srv.close()
