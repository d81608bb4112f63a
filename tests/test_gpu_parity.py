"""HIP path vs the oracle on seeded inputs: maxima, starting points and walk, bit for bit.

Every case runs in both launch modes (0 = plain per-batch, 1 = batched persistent).
"""
import zlib

import numpy as np
import pytest

from tests._parity import AMINO, DNA, compare, hip_all_vs_all, plan_info, planted_set

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def gdev():
    """A context that keeps every lane-per-pair job whose lanes are busy enough:
    the default length rule sends the pairs of plans this small to the
    wave-per-pair kernel."""
    from pypaswas_amd._native import Device
    d = Device(0)
    d.set_routing(-1, 1e9)
    yield d
    d.close()


def dna_params(O, affine, minimum=10.0, ll=1.0, ff=0.2):
    return O.Params(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, gap_ext=(-1.0 if affine else None),
                    lower_limit=ll, minimum=minimum, filter_factor=ff, highest=1.0)


def blosum_params(O, minimum=30.0, ll=1.0, ff=0.05):
    return O.Params(matrix=O.BLOSUM62, gap=-11.0, gap_ext=-1.0, lower_limit=ll, minimum=minimum, filter_factor=ff,
                    highest=float(max(max(O.BLOSUM62))), dna=False)


CASES = [
    # name, alphabet, n queries, n targets, query lengths, target lengths, affine
    ("dna-linear-small", DNA, 6, 5, (20, 150), (10, 60), False),
    ("dna-affine-small", DNA, 6, 5, (20, 150), (10, 60), True),
    ("dna-linear-multistripe", DNA, 3, 3, (600, 1300), (100, 700), False),
    ("dna-affine-multistripe", DNA, 3, 3, (600, 1300), (100, 700), True),
]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_dna_parity(dev, oracle, case, mode):
    name, alpha, nq, nt, ql, tl, affine = case
    print("case", name, "mode", mode, flush=True)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    qs, ts = planted_set(rng, nq, nt, ql, tl, alpha, plant=0.7)
    p = dna_params(oracle, affine)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
    compare(oracle, qs, ts, p, res)


@pytest.mark.parametrize("mode", [0, 1])
def test_protein_affine_parity(dev, oracle, mode):
    rng = np.random.default_rng(7)
    qs, ts = planted_set(rng, 5, 5, (30, 400), (30, 400), AMINO, plant=0.8, sub=0.5, indel=0.05)
    p = blosum_params(oracle)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
    compare(oracle, qs, ts, p, res)


@pytest.mark.parametrize("ll", [1.0, 0.8, 0.5])
@pytest.mark.parametrize("affine", [False, True])
def test_lower_limit_many_starts(dev, oracle, ll, affine):
    """--llimit < 1 produces many starting points, merging paths and suppressed walks."""
    rng = np.random.default_rng(11)
    qs, ts = planted_set(rng, 4, 4, (60, 200), (40, 120), DNA, plant=1.0, sub=0.03)
    p = dna_params(oracle, affine, minimum=5.0, ll=ll, ff=0.0)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)


def test_repeats_and_ties(dev, oracle):
    """Homopolymers and tandem repeats: many equal maxima and merging paths."""
    qs = ["A" * 40, "ACACACACACACACACACAC", "GGGGTTTTGGGGTTTT" * 3]
    ts = ["A" * 25, "CACACACACA", "TTTTGGGG" * 4]
    for affine in (False, True):
        p = dna_params(oracle, affine, minimum=3.0, ll=1.0, ff=0.0)
        res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
        compare(oracle, qs, ts, p, res)


GROUP_CASES = [
    # queries of similar length and enough of them per target to be scored
    # lane-per-pair; targets above 96 residues are processed in chunks of 64
    ("group-dna-linear", DNA, 70, 4, (240, 300), (12, 300), False),
    ("group-dna-affine", DNA, 70, 4, (240, 300), (12, 300), True),
    ("group-dna-affine-short", DNA, 130, 3, (200, 260), (20, 80), True),
]


def overwrite_set(rng, nq, nt, ql, tl, alphabet, plant=0.6, sub=0.05):
    """Queries of length in ql; some carry a mutated target piece written over a window."""
    from tests._parity import mutate, rand_seq
    targets = [rand_seq(rng, int(rng.integers(tl[0], tl[1] + 1)), alphabet) for _ in range(nt)]
    queries = []
    for _ in range(nq):
        q = list(rand_seq(rng, int(rng.integers(ql[0], ql[1] + 1)), alphabet))
        if rng.random() < plant:
            t = targets[int(rng.integers(0, nt))]
            piece = mutate(rng, t, sub, 0.0, alphabet)[: len(q)]
            off = int(rng.integers(0, len(q) - len(piece) + 1))
            q[off:off + len(piece)] = list(piece)
        queries.append("".join(q))
    return queries, targets


@pytest.mark.parametrize("case", GROUP_CASES, ids=[c[0] for c in GROUP_CASES])
def test_group_kernel_parity(gdev, oracle, case):
    name, alpha, nq, nt, ql, tl, affine = case
    print("case", name, flush=True)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    qs, ts = overwrite_set(rng, nq, nt, ql, tl, alpha)
    p = dna_params(oracle, affine, minimum=8.0)
    info = plan_info(gdev, oracle, qs, ts, p)
    assert info["orientation"] == 0 and info["lane_per_pair_cells"] > 0, info
    res = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs > 0


def test_group_kernel_protein(gdev, oracle):
    rng = np.random.default_rng(99)
    qs, ts = overwrite_set(rng, 70, 3, (200, 260), (50, 700), AMINO, plant=0.8, sub=0.4)
    p = blosum_params(oracle, minimum=25.0)
    info = plan_info(gdev, oracle, qs, ts, p)
    assert info["orientation"] == 0 and info["lane_per_pair_cells"] > 0, info
    res = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    res0 = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=0)
    assert np.array_equal(res0.maxima, res.maxima)
    assert np.array_equal(res0.hits, res.hits)


@pytest.mark.parametrize("affine", [False, True])
def test_fractional_scores(gdev, oracle, affine):
    """Scores that are not whole numbers (and not binary fractions, so sums
    round): the lane-per-pair kernels run their unshifted form (the row shift
    needs whole numbers), the wave-per-pair kernels the same recurrences as
    always; both match the oracle bit for bit."""
    print("fractional affine", affine, flush=True)
    rng = np.random.default_rng(7 + int(affine))
    qs, ts = overwrite_set(rng, 70, 4, (60, 200), (20, 150), DNA)
    p = oracle.Params(matrix=oracle.dna_rna_matrix(1.1, -2.7, -1.3, 0.2), gap=-4.3,
                      gap_ext=(-0.9 if affine else None), lower_limit=1.0, minimum=7.5, filter_factor=0.2,
                      highest=1.1)
    info = plan_info(gdev, oracle, qs, ts, p)
    assert info["lane_per_pair_cells"] > 0, info
    res = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs > 0
    res0 = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=0)
    assert np.array_equal(res0.maxima.view(np.uint32), res.maxima.view(np.uint32))
    assert np.array_equal(res0.hits, res.hits)


@pytest.mark.parametrize("kind", ["dna-linear", "dna-affine", "protein"])
def test_group_kernel_query_columns(gdev, oracle, kind):
    """Few queries against many targets of similar length: the lane-per-pair jobs
    take the query as their shared column sequence (single- and multi-chunk)."""
    print("case", kind, flush=True)
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    alpha = AMINO if kind == "protein" else DNA
    rows, cols = overwrite_set(rng, 70, 3, (200, 260), (40, 300), alpha, plant=0.8,
                               sub=0.4 if kind == "protein" else 0.05)
    qs, ts = cols, rows  # 3 queries x 70 targets
    p = blosum_params(oracle, minimum=25.0) if kind == "protein" else dna_params(oracle, kind == "dna-affine",
                                                                                  minimum=8.0)
    info = plan_info(gdev, oracle, qs, ts, p)
    assert info["orientation"] == 1 and info["lane_per_pair_cells"] > 0, info
    res = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs > 0
    res0 = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=0)
    assert np.array_equal(res0.maxima, res.maxima)
    assert np.array_equal(res0.hits, res.hits)


@pytest.fixture(scope="module")
def sdev():
    """A context whose length rule is tight enough for a small plan's jobs to be
    split into one item per chunk (pipelined over rows), and loose enough for
    one chunk per wave."""
    from pypaswas_amd._native import Device
    d = Device(0)
    d.set_routing(1, 150.0)
    yield d
    d.close()


@pytest.mark.parametrize("kind", ["dna-linear", "dna-affine", "protein"])
def test_group_kernel_split_jobs(sdev, oracle, kind):
    """Lane-per-pair jobs split into chunk items that hand their right edge to the
    next item row by row: same maxima, starting points and walks as the oracle and
    as the plain launch."""
    print("case", kind, flush=True)
    rng = np.random.default_rng(zlib.crc32(("split-" + kind).encode()))
    alpha = AMINO if kind == "protein" else DNA
    rows, cols = overwrite_set(rng, 70, 3, (200, 260), (280, 420), alpha, plant=0.8,
                               sub=0.4 if kind == "protein" else 0.05)
    qs, ts = cols, rows  # 3 queries (the split column sequences) x 70 targets
    p = blosum_params(oracle, minimum=25.0) if kind == "protein" else dna_params(oracle, kind == "dna-affine",
                                                                                  minimum=8.0)
    info = plan_info(sdev, oracle, qs, ts, p)
    assert info["orientation"] == 1 and info["split_job_items"] > 0, info
    res = hip_all_vs_all(sdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs > 0
    res0 = hip_all_vs_all(sdev, oracle, qs, ts, p, mode=0)
    assert np.array_equal(res0.maxima, res.maxima)
    assert np.array_equal(res0.hits, res.hits)
