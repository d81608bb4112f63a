"""The scoring pass's affine recurrence without the clamp of M.

The reference (default_main_cpu.cl:224-387) computes, per cell,
    M = max(T[i-1][j-1] + s, 0)            T = max(M, I, J)
    I = max(I[i][j-1] + E, M[i][j-1] + G)  J = max(J[i-1][j] + E, M[i-1][j] + G)
The HIP scoring kernels (score_kernel and group_kernel) instead feed the
unclamped d = T[i-1][j-1] + s into the gap candidates and clamp J at 0:
    I' = max(I'[i][j-1] + E, d[i][j-1] + G)
    J' = max(J'[i-1][j] + E, d[i-1][j] + G, 0)
    T  = max(d, I', J')
which is exact when G, E <= 0 (pw_set_scoring rejects anything else).  These
tests run both forms in float32 on random small matrices, the FILL padding
score (-1e10) included, and require every T to be bit-identical.
"""
import zlib

import numpy as np
import pytest

GAP_INIT = np.float32(-1e10)
FILL = np.float32(-1e10)


def reference_T(S, G, E):
    n, m = S.shape
    T = np.zeros((n + 1, m + 1), np.float32)
    M = np.zeros((n + 1, m + 1), np.float32)
    I = np.full((n + 1, m + 1), GAP_INIT, np.float32)
    J = np.full((n + 1, m + 1), GAP_INIT, np.float32)
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            M[i, j] = max(np.float32(T[i - 1, j - 1] + S[i - 1, j - 1]), np.float32(0))
            I[i, j] = max(np.float32(I[i, j - 1] + E), np.float32(M[i, j - 1] + G))
            J[i, j] = max(np.float32(J[i - 1, j] + E), np.float32(M[i - 1, j] + G))
            T[i, j] = max(M[i, j], I[i, j], J[i, j])
    return T[1:, 1:]


def clamp_free_T(S, G, E):
    n, m = S.shape
    T = np.zeros((n + 1, m + 1), np.float32)
    D = np.full((n + 1, m + 1), np.float32(0), np.float32)  # d of the boundary: M = 0 there
    I = np.full((n + 1, m + 1), GAP_INIT, np.float32)
    J = np.zeros((n + 1, m + 1), np.float32)  # J' of the boundary row: max(G, 0) = 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            D[i, j] = np.float32(T[i - 1, j - 1] + S[i - 1, j - 1])
            I[i, j] = max(np.float32(I[i, j - 1] + E), np.float32(D[i, j - 1] + G))
            J[i, j] = max(np.float32(J[i - 1, j] + E), np.float32(D[i - 1, j] + G), np.float32(0))
            T[i, j] = max(D[i, j], I[i, j], J[i, j])
    return T[1:, 1:]


def shifted_max(S, G, E, nb):
    """The row-shifted form of the lane-per-pair kernel, chunk by chunk (nb
    columns each, the right edge T^ and I^ handed to the next chunk) with the
    running maximum kept as rmax + sigma_r inside a row.  Returns the matrix
    maximum and T^ - sigma_r for every cell."""
    n, m = S.shape
    GmE = np.float32(G - E)
    edge = None  # per row: (T^ of the chunk's last column, I^ entering the next chunk)
    rmax = np.float32(0)
    out = np.zeros((n, m), np.float32)
    for c0 in range(0, m, nb):
        w = min(nb, m - c0)
        Tp = np.full(w, E, np.float32)   # T^ of the row above row 0: T = 0, sigma_-1 = E
        Jc = np.zeros(w, np.float32)
        dl = np.float32(E)
        new_edge = []
        sig1 = np.float32(-E)
        for r in range(n):
            sig = np.float32(sig1 + E)
            if edge is None:
                lb = (sig, np.float32(GmE + sig))
            else:
                lb = edge[r]
            ic = lb[1]
            rmax = np.float32(rmax + sig)
            Tn = np.empty(w, np.float32)
            for c in range(w):
                d = np.float32((dl if c == 0 else Tp[c - 1]) + np.float32(S[r, c0 + c] - E))
                mm = np.float32(d + GmE)
                P = np.float32(ic + E)
                T = max(d, P, Jc[c])
                Jc[c] = max(Jc[c], mm, sig1)
                ic = max(P, mm)
                Tn[c] = T
                rmax = max(rmax, T)
                out[r, c0 + c] = np.float32(T - sig)
            dl = lb[0]
            Tp = Tn
            rmax = np.float32(rmax - sig)
            new_edge.append((Tn[-1], ic))
            sig1 = np.float32(sig1 - E)
        edge = new_edge
    return rmax, out


@pytest.mark.parametrize("kind", ["dna", "protein"])
@pytest.mark.parametrize("gaps", [(-5.0, -1.0), (-11.0, -1.0), (0.0, -2.0), (-3.0, -3.0)])
def test_row_shifted_affine_matches_reference(kind, gaps):
    """Integral scores: the shifted form gives every T and the maximum bit for
    bit, across chunk edges."""
    G, E = np.float32(gaps[0]), np.float32(gaps[1])
    rng = np.random.default_rng(zlib.crc32(repr(("shift", kind, gaps)).encode()))
    for _ in range(5):
        n, m = int(rng.integers(1, 20)), int(rng.integers(1, 30))
        S = _scores(rng, n, m, kind)
        ref = reference_T(S, G, E)
        for nb in (4, 8, 64):
            mx, T = shifted_max(S, G, E, nb)
            assert np.array_equal(ref.view(np.uint32), T.view(np.uint32)), (kind, gaps, n, m, nb)
            assert mx == ref.max(), (kind, gaps, n, m, nb)


def _scores(rng, n, m, kind):
    if kind == "dna":
        a = rng.integers(0, 4, n)
        b = rng.integers(0, 4, m)
        S = np.where(a[:, None] == b[None, :], 1.0, -3.0).astype(np.float32)
    elif kind == "protein":
        S = rng.integers(-4, 12, (n, m)).astype(np.float32)
    else:  # fractional scores
        S = (rng.standard_normal((n, m)) * 2.5).astype(np.float32)
    # FILL padding rows and columns, as the kernels pad tiles
    pr, pc = rng.integers(0, 3), rng.integers(0, 3)
    if pr:
        S[-pr:, :] = FILL
    if pc:
        S[:, -pc:] = FILL
    return S


@pytest.mark.parametrize("kind", ["dna", "protein", "fractional"])
@pytest.mark.parametrize("gaps", [(-5.0, -1.0), (-11.0, -1.0), (0.0, -1.0), (-1.0, -0.5), (-3.0, -3.0)])
def test_clamp_free_affine_matches_reference(kind, gaps):
    G, E = np.float32(gaps[0]), np.float32(gaps[1])
    rng = np.random.default_rng(zlib.crc32(repr((kind, gaps)).encode()))
    for _ in range(6):
        n, m = int(rng.integers(1, 24)), int(rng.integers(1, 24))
        S = _scores(rng, n, m, kind)
        ref = reference_T(S, G, E)
        alt = clamp_free_T(S, G, E)
        assert np.array_equal(ref.view(np.uint32), alt.view(np.uint32)), (kind, gaps, n, m)


def _direction_codes(S, G, E, thr, clamp_free):
    """Direction byte (main 1 | STOP 4, I 2, J 3, 0 when T <= 0) and candidate
    flag of every cell, as the direction pass forms them
    (default_main_cpu.cl:283-345): the reference form from M, or the
    clamp-free form from d, J clamped at 0."""
    n, m = S.shape
    T = np.zeros((n + 1, m + 1), np.float32)
    M = np.zeros((n + 1, m + 1), np.float32)   # M (reference) or d (clamp-free); 0 on the boundary
    I = np.full((n + 1, m + 1), GAP_INIT, np.float32)
    J = np.full((n + 1, m + 1), np.float32(0) if clamp_free else GAP_INIT, np.float32)
    code = np.zeros((n, m), np.uint8)
    cand = np.zeros((n, m), bool)
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            d = np.float32(T[i - 1, j - 1] + S[i - 1, j - 1])
            mm = d if clamp_free else max(d, np.float32(0))
            left = M[i, j - 1] if (clamp_free or j > 1) else np.float32(0)
            I[i, j] = max(np.float32(I[i, j - 1] + E), np.float32(left + G))
            up = M[i - 1, j]
            if clamp_free:
                J[i, j] = max(np.float32(J[i - 1, j] + E), np.float32(up + G), np.float32(0))
            else:
                J[i, j] = max(np.float32(J[i - 1, j] + E), np.float32(up + G))
            t = max(mm, I[i, j], J[i, j])
            T[i, j] = t
            M[i, j] = mm
            md = M[i - 1, j - 1]
            stop = (md <= 0) if clamp_free else (md == 0)
            main = 1 | (4 if stop else 0)
            other = 2 if t == I[i, j] else 3
            pos = t > 0
            code[i - 1, j - 1] = (main if t == mm else other) if pos else 0
            cand[i - 1, j - 1] = pos and t == mm and mm >= thr
    return code, cand


@pytest.mark.parametrize("kind", ["dna", "protein", "fractional"])
@pytest.mark.parametrize("gaps", [(-5.0, -1.0), (-11.0, -1.0), (0.0, -1.0), (-1.0, -0.5)])
def test_clamp_free_direction_codes_match_reference(kind, gaps):
    """The direction pass's clamp-free form writes the reference's direction
    bytes and candidate flags."""
    G, E = np.float32(gaps[0]), np.float32(gaps[1])
    rng = np.random.default_rng(zlib.crc32(repr(("dirs", kind, gaps)).encode()))
    for _ in range(5):
        n, m = int(rng.integers(1, 20)), int(rng.integers(1, 20))
        S = _scores(rng, n, m, kind)
        thr = np.float32(rng.integers(1, 6))
        ref = _direction_codes(S, G, E, thr, False)
        alt = _direction_codes(S, G, E, thr, True)
        assert np.array_equal(ref[0], alt[0]) and np.array_equal(ref[1], alt[1]), (kind, gaps, n, m)


def test_clamp_free_needs_nonpositive_gaps():
    """A positive gap-open score breaks the equivalence (so the library rejects
    it): the check is not vacuous."""
    S = np.array([[2.0, -3.0, -3.0], [-3.0, -3.0, -3.0]], np.float32)
    G, E = np.float32(1.0), np.float32(-1.0)
    assert not np.array_equal(reference_T(S, G, E), clamp_free_T(S, G, E))


def _linear_codes(S, G, thr, diagonal_stop_only):
    """Linear-gap direction bytes (default_main_cpu.cl:162-177): the value is
    max(0, left, up, diag) and the code that of the first of left, up, diag
    equal to it when positive, with STOP when the predecessor along the code
    is 0 -- or, in the direction pass's form, when the diagonal predecessor
    is 0 (gap codes never stop for G <= 0)."""
    n, m = S.shape
    H = np.zeros((n + 1, m + 1), np.float32)
    code = np.zeros((n, m), np.uint8)
    cand = np.zeros((n, m), bool)
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            hd, left, up = H[i - 1, j - 1], H[i, j - 1], H[i - 1, j]
            d = np.float32(hd + S[i - 1, j - 1])
            lv, uv = np.float32(left + G), np.float32(up + G)
            cur = max(max(lv, uv), d, np.float32(0))
            isl, isu = cur == lv, cur == uv
            if diagonal_stop_only:
                b = 3 if isl else (2 if isu else (1 | (4 if hd == 0 else 0)))
            else:
                pv = left if isl else (up if isu else hd)
                b = (3 if isl else (2 if isu else 1)) | (4 if pv == 0 else 0)
            code[i - 1, j - 1] = b if cur > 0 else 0
            cand[i - 1, j - 1] = cur > 0 and not isl and not isu and cur >= thr
            H[i, j] = cur
    return code, cand


@pytest.mark.parametrize("kind", ["dna", "protein", "fractional"])
@pytest.mark.parametrize("G", [-5.0, -1.0, -0.5, 0.0])
def test_linear_direction_stop_on_diagonal_only(kind, G):
    G = np.float32(G)
    rng = np.random.default_rng(zlib.crc32(repr(("lin", kind, float(G))).encode()))
    for _ in range(6):
        n, m = int(rng.integers(1, 20)), int(rng.integers(1, 20))
        S = _scores(rng, n, m, kind)
        thr = np.float32(rng.integers(1, 6))
        ref = _linear_codes(S, G, thr, False)
        alt = _linear_codes(S, G, thr, True)
        assert np.array_equal(ref[0], alt[0]) and np.array_equal(ref[1], alt[1]), (kind, G, n, m)


def test_linear_stop_rule_needs_nonpositive_gap():
    """With G > 0 a gap cell can follow a zero cell, and the two forms part."""
    S = np.array([[-3.0, -3.0], [-3.0, -3.0]], np.float32)
    ref = _linear_codes(S, np.float32(1.0), np.float32(1.0), False)
    alt = _linear_codes(S, np.float32(1.0), np.float32(1.0), True)
    assert not np.array_equal(ref[0], alt[0])
