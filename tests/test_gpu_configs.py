"""The benchmark workloads under test: the HIP path against the oracle, bit for bit.

* Configuration 2 (20,000 reads x 64 adapters + reverse complements, affine
  DNA): a stride subsample of 400 reads x all 128 targets against the serial
  oracle pair by pair, and all 2.56 M pairs batched == plain.
* Configuration 3 (proteins, BLOSUM62 -11/-1): 20 x 200 against the oracle,
  and the 200 x 2,000 sample batched == plain.
* Configuration 4 (10,000 x 10,000 DNA pairs, linear): two full pairs and
  their reverse complements against the oracle with the direction pass's
  residues per lane left to the planner and forced to 2 and to 8 (chains of
  40, 79 and 20 stripes), and all 64 jobs batched == plain.
* Configuration 5 (serving mix) with long jobs of 5-20 kb: every job served
  in the mix equals the job run alone.

The oracle's side is oracle_serial_pair plus the walk, pair by pair
(tests._parity.compare_pairs); per-pair results do not depend on batching.
"""
import numpy as np
import pytest

from tests._parity import compare_pairs

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


def oracle_params(O, st):
    """The oracle's Params of a workload's settings dict."""
    dna = st["matrix_name"] == "DNA-RNA"
    if dna:
        matrix = O.dna_rna_matrix(float(st["match_score"]), float(st["mismatch_score"]), -1.0, 0.0)
        highest = float(st["match_score"])
    else:
        matrix = O.BLOSUM62
        highest = float(max(max(O.BLOSUM62)))
    ext = float(st["gap_extension"])
    return O.Params(matrix=matrix, gap=float(st["gap_score"]), gap_ext=(ext if ext != 0.0 else None),
                    lower_limit=float(st["lower_limit_score"]), minimum=float(st["minimum_score"]),
                    filter_factor=float(st["filter_factor"]), highest=highest, dna=dna)


def all_pairs(nq, nt):
    return np.repeat(np.arange(nq, dtype=np.int32), nt), np.tile(np.arange(nt, dtype=np.int32), nq)


def run(dev, O, p, queries, targets, pq, pt, qids, tids, mode=1):
    from pypaswas_amd._native import PackedSeqs
    qlen = np.array([len(s) for s in queries], dtype=np.float64)
    tlen = np.array([len(s) for s in targets], dtype=np.float64)
    # O.min_start_score over all pairs: highest * min(len) * filter_factor in double, stored f32
    mss = ((float(p.highest) * np.minimum(qlen[pq], tlen[pt])) * float(p.filter_factor)).astype(np.float32)
    same = (np.array(qids, dtype=object)[pq] == np.array(tids, dtype=object)[pt]).astype(np.uint8)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    return dev.align_pairs(PackedSeqs.pack(queries), PackedSeqs.pack(targets), pq, pt, mss, same, p.minimum,
                           mode=mode)


@pytest.fixture(scope="module")
def config2():
    from pypaswas_amd import workloads as W
    reads, adapters, st = W.config2()
    q, t = W.reference_order(reads, adapters, True)
    return q, t, st


def test_config2_subsample_vs_oracle(dev, oracle, config2):
    q, t, st = config2
    sub = q[::50]
    print("config2 subsample", len(sub), "x", len(t), flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in sub], [s for _, s in t]
    qids, tids = [i for i, _ in sub], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    res = run(dev, oracle, p, qs, ts, pq, pt, qids, tids)
    n = compare_pairs(oracle, qs, ts, pq, pt, p, res, qids, tids)
    assert n >= 100 and len(res.hits) >= 100


def test_config2_full_batched_equals_plain(dev, oracle, config2):
    q, t, st = config2
    print("config2 full batched vs plain", flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    batched = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=1)
    plain = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=0)
    same_result(batched, plain)
    assert len(pq) == 2_560_000 and batched.n_traceback_pairs > 5000


def test_config3_subsample_vs_oracle(dev, oracle):
    from pypaswas_amd import workloads as W
    qr, tr, st = W.config3(n_queries=20, n_targets=200)
    q, t = W.reference_order(qr, tr, False)
    print("config3 20 x 200", flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    for mode in (1, 0):
        res = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=mode)
        n = compare_pairs(oracle, qs, ts, pq, pt, p, res, qids, tids)
        assert n >= 20


def test_config3_sample_batched_equals_plain(dev, oracle):
    from pypaswas_amd import workloads as W
    qr, tr, st = W.config3(n_queries=200, n_targets=2000)
    q, t = W.reference_order(qr, tr, False)
    print("config3 200 x 2000 batched vs plain", flush=True)
    p = oracle_params(oracle, st)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    batched = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=1)
    plain = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=0)
    same_result(batched, plain)
    assert batched.n_traceback_pairs > 1000


@pytest.fixture(scope="module")
def config4():
    from pypaswas_amd import workloads as W
    qr, tr, st = W.config4()
    q, t = W.reference_order(qr, tr, True)
    qs, ts = [s for _, s in q], [s for _, s in t]
    qids, tids = [i for i, _ in q], [i for i, _ in t]
    pq, pt = all_pairs(len(qs), len(ts))
    mask = W.config4_pair_mask(q, t, pq, pt)
    return qs, ts, qids, tids, pq[mask], pt[mask], st


@pytest.mark.parametrize("rmax", [0, 2, 8])
def test_config4_two_jobs_vs_oracle(oracle, config4, rmax):
    """Two 10 kb x 10 kb jobs (each query against its target and the target's
    reverse complement) with the direction pass's residues per lane chosen by
    the plan (0), or capped at 2 and 8."""
    from pypaswas_amd._native import Device
    qs, ts, qids, tids, pq, pt, st = config4
    jobs = sorted(set(pq.tolist()))[:2]
    keep = np.isin(pq, jobs)
    pq2, pt2 = pq[keep], pt[keep]
    assert len(pq2) == 4
    print("config4 two jobs, residues cap", rmax, flush=True)
    p = oracle_params(oracle, st)
    d = Device(0)
    try:
        if rmax:
            d.set_routing(-1, 0.0, rmax)
        res = run(d, oracle, p, qs, ts, pq2, pt2, qids, tids)
        n = compare_pairs(oracle, qs, ts, pq2, pt2, p, res, qids, tids, threads=4)
        assert n >= 2 and res.n_traceback_pairs >= 2
    finally:
        d.close()


def test_config4_all_jobs_batched_equals_plain(dev, oracle, config4):
    qs, ts, qids, tids, pq, pt, st = config4
    print("config4 all jobs batched vs plain", flush=True)
    p = oracle_params(oracle, st)
    batched = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=1)
    plain = run(dev, oracle, p, qs, ts, pq, pt, qids, tids, mode=0)
    same_result(batched, plain)
    assert len(pq) == 128 and batched.n_traceback_pairs >= 64


def test_config5_long_jobs_served_equal_alone(dev):
    """The serving mix with long DNA jobs of 5-20 kb (plus reverse complements):
    every job served in the mix equals the job run alone."""
    from pypaswas_amd import workloads as W
    from pypaswas_amd.serving import Job, Server, replay
    schemes = W.serving_schemes()
    raw = W.serving_jobs(seed=20, n_jobs=80, long_range=(5000, 20000))
    longs = [k for k, (kind, _, _) in enumerate(raw) if kind == "long"]
    assert len(longs) >= 3
    assert max(max(len(s) for _, s in q) for kind, q, _ in raw if kind == "long") > 10000
    jobs = [Job.all_vs_all(W.SERVING_KINDS.index(k), schemes[W.SERVING_KINDS.index(k)], q, t) for k, q, t in raw]
    print("serving with", len(longs), "long jobs", flush=True)
    srv = Server(dev, lanes=4, direction_bytes=4 << 30)
    try:
        for i, sc in enumerate(schemes):
            srv.set_scheme(i, sc)
        results, lat, wall = replay(srv, jobs)
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(jobs)
    for k, job in enumerate(jobs):
        sc = schemes[job.scheme]
        dev.set_scoring(sc.subst, sc.char_offset, sc.gap, sc.gap_ext, sc.lower_limit, sc.minimum)
        alone = dev.align_pairs(job.queries, job.targets, job.pq, job.pt, job.mss, job.same, job.minimum,
                                job.sp_cap, mode=1)
        same_result(results[k], alone)
    assert sum(results[k].n_traceback_pairs for k in longs) >= 3
