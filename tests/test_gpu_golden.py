"""The hand-derived golden cases through the HIP library (both launch modes).

The cases of tests/test_oracle_golden.py were worked on paper from the
reference's rules (default_main_cpu.cl:154-604, SmithWaterman.py:689-841);
here the GPU's maxima, starting points and walks are compared with those
literals directly, not through the oracle.
"""
import numpy as np
import pytest

from tests._parity import decode_ops, hip_all_vs_all
from tests.test_oracle_golden import CASES, params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from pypaswas_amd._native import Device
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=[c["name"].split()[0] for c in CASES])
def test_golden_case_on_gpu(dev, oracle, case, mode):
    print("golden", case["name"], "mode", mode, flush=True)
    q, t = case["q"], case["t"]
    p = params(oracle, case)
    res = hip_all_vs_all(dev, oracle, [q], [t], p, mode=mode)
    assert float(res.maxima[0]) == case["mx"]
    sps = [(int(s["blockX"]) * 8 + int(s["valueX"]), int(s["blockY"]) * 8 + int(s["valueY"]), float(s["score"]))
           for s in res.starting_points]
    assert sps == case["sps"]
    assert all(float(s["maxScore"]) == case["mx"] and float(s["posScore"]) == 0.0 for s in res.starting_points)
    got = []
    for h in res.hits:
        if h["shown"]:
            sm, al, tm = decode_ops(res.ops, int(h["ops_offset"]), int(h["length"]), q, t, int(h["s_start"]),
                                    int(h["t_start"]))
            got.append((True, int(h["s_start"]), int(h["s_end"]), int(h["t_start"]), int(h["t_end"]),
                        int(h["matches"]), int(h["mismatches"]), sm, al, tm))
        else:
            got.append((False,))
    assert got == case["walks"]
