"""Long pairs: stripes split over several waves (scoring and direction pass) match the oracle.

Pairs with several stripes and at least 256 swept columns run one stripe per
wave, each stripe waiting for the bottom row of the one above.  These tests
cover such pairs against the oracle (linear, affine, protein), and compare the
split runs with runs where the split is impossible (plain launch for the
scoring pass; no room for boundary rows in the direction pass).  Pairs with
|q| + |t| >= 8192 also take the wave-per-pair start selection and walk.
"""
import numpy as np
import pytest

from tests._parity import AMINO, DNA, compare, hip_all_vs_all, mutate, rand_seq
from tests.test_gpu_parity import blosum_params, dna_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


def long_pairs(rng, n, length, alphabet, region=800, ident=0.85, indel=0.02):
    """n queries and n targets of about `length`, each target carrying a
    mutated region of the query with the same index."""
    qs, ts = [], []
    for _ in range(n):
        q = rand_seq(rng, int(rng.integers(length[0], length[1] + 1)), alphabet)
        t = list(rand_seq(rng, int(rng.integers(length[0], length[1] + 1)), alphabet))
        a = int(rng.integers(0, len(q) - region))
        piece = mutate(rng, q[a:a + region], 1.0 - ident, indel, alphabet)[:region]
        off = int(rng.integers(0, len(t) - len(piece)))
        t[off:off + len(piece)] = list(piece)
        qs.append(q)
        ts.append("".join(t))
    return qs, ts


def same_result(a, b):
    assert np.array_equal(a.maxima.view(np.uint32), b.maxima.view(np.uint32))
    assert np.array_equal(a.starting_points, b.starting_points)
    assert np.array_equal(a.hits, b.hits)
    assert a.ops == b.ops


@pytest.mark.parametrize("affine", [False, True])
def test_long_dna_vs_oracle(dev, oracle, affine):
    print("long dna affine", affine, flush=True)
    rng = np.random.default_rng(40 + int(affine))
    qs, ts = long_pairs(rng, 2, (2500, 4000), DNA)
    # filter factor low enough that the planted pairs pass the traceback gate
    p = dna_params(oracle, affine, ff=0.05)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs >= 2
    plain = hip_all_vs_all(dev, oracle, qs, ts, p, mode=0)
    same_result(plain, res)


@pytest.mark.parametrize("affine", [False, True])
def test_long_path_pairs_vs_oracle(dev, oracle, affine):
    """Pairs with |q| + |t| >= 8192 take the wave-per-pair start selection and
    walk (LDS windows over the direction area); the shorter pairs of the same
    batch the thread-per-pair kernels."""
    print("long path pairs affine", affine, flush=True)
    rng = np.random.default_rng(60 + int(affine))
    qs, ts = long_pairs(rng, 2, (4300, 5200), DNA, region=1500)
    qs.append(rand_seq(rng, 900, DNA))
    p = dna_params(oracle, affine, ff=0.05)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs >= 2


def test_long_protein_vs_oracle(dev, oracle):
    rng = np.random.default_rng(44)
    qs, ts = long_pairs(rng, 2, (1500, 2500), AMINO, region=500, ident=0.6)
    p = blosum_params(oracle, ff=0.02)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs >= 2


@pytest.fixture()
def nosplit_dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    # a direction limit this small leaves no room for boundary rows: every pair runs unsplit
    d.set_limits(max_direction_bytes=1 << 16)
    yield d
    d.close()


@pytest.mark.parametrize("affine", [False, True])
def test_split_equals_unsplit(dev, nosplit_dev, oracle, affine):
    rng = np.random.default_rng(50 + int(affine))
    qs, ts = long_pairs(rng, 3, (1200, 2200), DNA, region=500)
    p = dna_params(oracle, affine, ff=0.05)
    split = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    unsplit = hip_all_vs_all(nosplit_dev, oracle, qs, ts, p, mode=1)
    same_result(split, unsplit)
    assert split.n_traceback_pairs >= 3


@pytest.mark.parametrize("affine", [False, True])
def test_long_diagonal_runs_reach_the_edges(dev, oracle, affine):
    """Wave-per-pair start selection and walk take diagonal runs one cell per
    lane.  Long exact and near-exact diagonals here end at the matrix corner
    (query = target prefix), at column 0 (the target's start inside the query)
    and at row 0 (the query's start inside the target), crossing many LDS
    windows and stripes; a low lower limit (0.5 linear, 0.97 affine: ~2,300 and
    ~9,800 starting points) gives many starting points per pair whose paths run
    into cells an earlier path marked or walked."""
    print("long diagonal runs affine", affine, flush=True)
    rng = np.random.default_rng(70 + int(affine))
    t0 = rand_seq(rng, 4300, DNA)
    q0 = t0[:4100] + rand_seq(rng, 200, DNA)                  # corner: path to (0, 0)
    body = mutate(rng, rand_seq(rng, 3900, DNA), 0.03, 0.0, DNA)
    q1 = rand_seq(rng, 300, DNA) + body                       # target from column 0
    t1 = body[:3900] + rand_seq(rng, 600, DNA)
    body2 = rand_seq(rng, 3800, DNA)
    q2 = mutate(rng, body2, 0.05, 0.01, DNA) + rand_seq(rng, 500, DNA)  # query from row 0
    t2 = rand_seq(rng, 500, DNA) + body2
    qs, ts = [q0, q1, q2], [t0, t1, t2]
    assert all(len(q) + len(t) >= 8192 for q, t in zip(qs, ts))
    p = dna_params(oracle, affine, ll=(0.97 if affine else 0.5), ff=0.05)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)
    assert res.n_traceback_pairs >= 3
    assert len(res.starting_points) > 1000
    shown = res.hits[res.hits["shown"] != 0]
    # the shown paths end at the corner, at column 0 and at row 0
    assert {(int(h["s_start"]) == 0, int(h["t_start"]) == 0) for h in shown} >= {(True, True), (False, True), (True, False)}
