"""Edge cases of the HIP path vs the oracle: lengths at the kernels' block
edges (1 residue, multiples of 8 and 64, one either side), empty sequences,
an empty pair list, IUPAC ambiguity codes and a gate nothing passes.

The reference pads every sequence to a multiple of 8 and scores the padding
with the FILL value (SmithWaterman.py:222-286, default_main_cpu.cl:108-222),
so the block edges are where a layout slip would show; all comparisons are
bit for bit.
"""
import zlib

import numpy as np
import pytest

from tests._parity import DNA, compare, hip_all_vs_all, plan_info, planted_set, rand_seq

pytestmark = pytest.mark.gpu

EDGE_Q = [1, 2, 7, 8, 9, 63, 64, 65, 127, 128, 255, 256, 257]
EDGE_T = [1, 3, 8, 31, 64, 65, 96, 97, 129, 300]


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def gdev():
    """Lane-per-pair jobs kept for every length (see test_gpu_parity.gdev)."""
    from pypaswas_amd._native import Device
    d = Device(0)
    d.set_routing(-1, 1e9)
    yield d
    d.close()


def dna_params(O, affine, minimum=3.0, ff=0.0):
    return O.Params(matrix=O.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0, gap_ext=(-1.0 if affine else None),
                    lower_limit=1.0, minimum=minimum, filter_factor=ff, highest=1.0)


def edge_set(seed, alphabet=DNA):
    rng = np.random.default_rng(seed)
    qs = [rand_seq(rng, n, alphabet) for n in EDGE_Q]
    ts = [rand_seq(rng, n, alphabet) for n in EDGE_T]
    # a target that is a piece of the longest query, so long paths exist
    ts.append(qs[-1][10:200])
    return qs, ts


@pytest.mark.parametrize("affine", [False, True], ids=["linear", "affine"])
@pytest.mark.parametrize("mode", [0, 1])
def test_block_edge_lengths(dev, oracle, mode, affine):
    print("block edges mode", mode, "affine", affine, flush=True)
    qs, ts = edge_set(zlib.crc32(b"edges"))
    p = dna_params(oracle, affine)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
    compare(oracle, qs, ts, p, res)


@pytest.mark.parametrize("affine", [False, True], ids=["linear", "affine"])
def test_block_edge_lengths_lane_per_pair(gdev, oracle, affine):
    """The same lengths through the lane-per-pair scoring kernels (every
    target width class, chunked targets above 96 residues)."""
    print("block edges lane-per-pair affine", affine, flush=True)
    rng = np.random.default_rng(zlib.crc32(b"edges-group"))
    # enough queries of similar length per target to fill 64-lane jobs
    qs = [rand_seq(rng, int(n), DNA) for n in rng.integers(120, 131, 70)]
    ts = [rand_seq(rng, n, DNA) for n in EDGE_T] + [qs[0][5:100]]
    p = dna_params(oracle, affine)
    info = plan_info(gdev, oracle, qs, ts, p)
    assert info["lane_per_pair_cells"] > 0, info
    res = hip_all_vs_all(gdev, oracle, qs, ts, p, mode=1)
    compare(oracle, qs, ts, p, res)


@pytest.mark.parametrize("mode", [0, 1])
def test_empty_sequences(dev, oracle, mode):
    """Zero-length queries and targets score 0 and yield nothing, beside
    ordinary pairs of the same call."""
    print("empty sequences mode", mode, flush=True)
    rng = np.random.default_rng(zlib.crc32(b"empty"))
    qs, ts = planted_set(rng, 4, 3, (40, 120), (30, 90), DNA, plant=1.0)
    qs = qs[:2] + [""] + qs[2:]
    ts = [""] + ts
    for affine in (False, True):
        p = dna_params(oracle, affine)
        res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
        compare(oracle, qs, ts, p, res)
        mx = res.maxima.reshape(len(qs), len(ts))
        assert not mx[2].any() and not mx[:, 0].any()


@pytest.mark.parametrize("mode", [0, 1])
def test_empty_pair_list(dev, oracle, mode):
    from pypaswas_amd._native import PackedSeqs
    p = dna_params(oracle, True)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext, p.lower_limit, p.minimum)
    none = np.zeros(0, np.int32)
    res = dev.align_pairs(PackedSeqs.pack(["ACGT"]), PackedSeqs.pack(["ACGA"]), none, none,
                          np.zeros(0, np.float32), None, p.minimum, mode=mode)
    assert len(res.maxima) == 0 and len(res.starting_points) == 0 and len(res.hits) == 0
    assert res.n_traceback_pairs == 0


@pytest.mark.parametrize("mode", [0, 1])
def test_iupac_codes(dev, oracle, mode):
    """Ambiguity codes (N scores the 'any' value, R/Y/K/M/S/W match their
    bases) on both sides of the pairs."""
    print("iupac mode", mode, flush=True)
    alpha = "ACGTACGTACGTNRYKMSWU"
    rng = np.random.default_rng(zlib.crc32(b"iupac"))
    qs, ts = planted_set(rng, 6, 5, (30, 300), (20, 200), alpha, plant=0.8, sub=0.1)
    for affine in (False, True):
        p = oracle.Params(matrix=oracle.dna_rna_matrix(2.0, -3.0, -1.0, 0.5), gap=-5.0,
                          gap_ext=(-2.0 if affine else None), lower_limit=1.0, minimum=8.0, filter_factor=0.1,
                          highest=2.0)
        res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
        compare(oracle, qs, ts, p, res)


@pytest.mark.parametrize("mode", [0, 1])
def test_gate_passes_nothing(dev, oracle, mode):
    """A minimum score above every pair maximum: exact maxima, no starting
    points, no hits and no traceback work."""
    print("closed gate mode", mode, flush=True)
    rng = np.random.default_rng(zlib.crc32(b"closed"))
    qs, ts = planted_set(rng, 5, 4, (50, 700), (40, 300), DNA, plant=1.0)
    p = dna_params(oracle, True, minimum=1e6)
    res = hip_all_vs_all(dev, oracle, qs, ts, p, mode=mode)
    compare(oracle, qs, ts, p, res)
    assert res.maxima.max() > 0
    assert len(res.starting_points) == 0 and len(res.hits) == 0 and res.n_traceback_pairs == 0


@pytest.mark.parametrize("affine", [False, True])
@pytest.mark.parametrize("rmax", [2, 4, 8])
def test_direction_residues_per_lane(oracle, rmax, affine):
    """The direction pass with its residues per lane capped at 2, 4 and 8
    (pw_set_routing), linear and affine, on queries of 40 to 700 residues:
    one- and multi-stripe pairs of every stripe height the cap allows.  Under
    the cap 8 a query of 257-384 residues runs with 6 residues per lane; the
    test checks that such a query has a pair passing the traceback gate."""
    from pypaswas_amd._native import Device
    from tests._parity import AMINO
    print("direction residues cap", rmax, "affine", affine, flush=True)
    rng = np.random.default_rng(zlib.crc32(b"dir-residues") + affine)
    if affine:
        qs, ts = planted_set(rng, 8, 4, (40, 700), (60, 400), AMINO, plant=0.9, sub=0.3, indel=0.03)
        p = oracle.Params(matrix=oracle.BLOSUM62, gap=-11.0, gap_ext=-1.0, lower_limit=0.8, minimum=25.0,
                          filter_factor=0.05, highest=float(max(max(oracle.BLOSUM62))), dna=False)
    else:
        qs, ts = planted_set(rng, 8, 4, (40, 700), (60, 400), DNA, plant=0.9, sub=0.05, indel=0.02)
        p = dna_params(oracle, False, minimum=10.0, ff=0.05)
        p.lower_limit = 0.8
    # a 300-residue query carrying an exact copy of target 0's start: its pairs
    # with target 0 pass the gate and run with 6 residues per lane under the cap 8
    alpha = AMINO if affine else DNA
    piece = ts[0][:100]
    qs[0] = rand_seq(rng, 300 - len(piece), alpha) + piece
    d = Device(0)
    try:
        d.set_routing(-1, 0.0, rmax)
        info = plan_info(d, oracle, qs, ts, p)
        assert info["direction_residues_per_lane"] == rmax, info
        res = hip_all_vs_all(d, oracle, qs, ts, p, mode=1)
        compare(oracle, qs, ts, p, res)
        assert res.n_traceback_pairs > 0
        mx = res.maxima.reshape(len(qs), len(ts))
        mss = np.array([[oracle.min_start_score(p, len(q), len(t)) for t in ts] for q in qs], dtype=np.float32)
        gated = (mx >= p.minimum) & (mx >= mss)
        six = [i for i, q in enumerate(qs) if 256 < len(q) <= 384 and gated[i].any()]
        assert six, "no gated pair with a 257-384-residue query"
    finally:
        d.close()
