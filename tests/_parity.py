"""Helpers shared by the parity tests: seeded inputs and the HIP-vs-oracle comparison."""
from __future__ import annotations

import numpy as np

DNA = "ACGT"
AMINO = "ARNDCQEGHILKMFPSTWYV"


def rand_seq(rng, n, alphabet=DNA):
    return "".join(rng.choice(list(alphabet), size=n))


def mutate(rng, s, sub_rate, indel_rate=0.0, alphabet=DNA):
    out = []
    for c in s:
        r = rng.random()
        if r < indel_rate / 2:
            continue  # deletion
        if r < indel_rate:
            out.append(rng.choice(list(alphabet)))  # insertion before c
        out.append(rng.choice(list(alphabet)) if rng.random() < sub_rate else c)
    return "".join(out)


def planted_set(rng, nq, nt, qlen, tlen, alphabet=DNA, plant=0.5, sub=0.05, indel=0.02):
    """Queries and targets with shared regions so that hits exist."""
    targets = [rand_seq(rng, int(rng.integers(tlen[0], tlen[1] + 1)), alphabet) for _ in range(nt)]
    queries = []
    for _ in range(nq):
        q = rand_seq(rng, int(rng.integers(qlen[0], qlen[1] + 1)), alphabet)
        if rng.random() < plant and targets:
            t = targets[int(rng.integers(0, nt))]
            a = int(rng.integers(0, max(1, len(t) // 2)))
            b = int(rng.integers(a + 1, len(t) + 1))
            piece = mutate(rng, t[a:b], sub, indel, alphabet)
            pos = int(rng.integers(0, len(q) + 1))
            q = q[:pos] + piece + q[pos:]
        queries.append(q)
    return queries, targets


def hip_all_vs_all(dev, O, queries, targets, p, mode=1, same_id=None):
    from pypaswas_amd._native import PackedSeqs
    N, T = len(queries), len(targets)
    pq = np.repeat(np.arange(N, dtype=np.int32), T)
    pt = np.tile(np.arange(T, dtype=np.int32), N)
    mss = np.array([O.min_start_score(p, len(queries[i]), len(targets[j])) for i, j in zip(pq, pt)],
                   dtype=np.float32)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    return dev.align_pairs(PackedSeqs.pack(queries), PackedSeqs.pack(targets), pq, pt, mss, same_id, p.minimum,
                           mode=mode)


def plan_info(dev, O, queries, targets, p, mode=1):
    """Routing of the all-vs-all plan hip_all_vs_all runs (pw_plan_info)."""
    from pypaswas_amd._native import PackedSeqs
    N, T = len(queries), len(targets)
    pq = np.repeat(np.arange(N, dtype=np.int32), T)
    pt = np.tile(np.arange(T, dtype=np.int32), N)
    mss = np.zeros(len(pq), dtype=np.float32)
    dev.set_scoring(p.subst_f32(), p.char_offset, p.gap, p.gap_ext if p.gap_ext is not None else 0.0,
                    p.lower_limit, p.minimum)
    plan = dev.prepare(PackedSeqs.pack(queries), PackedSeqs.pack(targets), pq, pt, mss, None, mode=mode)
    try:
        return plan.info()
    finally:
        plan.release()


def sp_tuple(s):
    return (int(s["sequence"]) if not hasattr(s, "sequence") else int(s.sequence),)


def compare(O, queries, targets, p, res, qids=None, tids=None):
    """Asserts bit-exact parity of maxima, starting points and walk results."""
    N, T = len(queries), len(targets)
    br = O.sched_batch(queries, targets, p)
    # maxima, bit for bit
    hip_max = res.maxima.reshape(N, T)
    assert np.array_equal(hip_max.view(np.uint32), br.maxima.view(np.uint32)), \
        "maxima differ at {}".format(np.argwhere(hip_max != br.maxima)[:5].tolist())
    # starting points per pair in canonical order
    o_sp = {}
    for s in br.starting_points:
        o_sp.setdefault((s.sequence, s.target), []).append(
            (s.sequence, s.target, s.blockX, s.blockY, s.valueX, s.valueY, np.float32(s.score).view(np.uint32),
             np.float32(s.maxScore).view(np.uint32), np.float32(s.posScore).view(np.uint32)))
    h_sp = {}
    for s in res.starting_points:
        h_sp.setdefault((int(s["sequence"]), int(s["target"])), []).append(
            (int(s["sequence"]), int(s["target"]), int(s["blockX"]), int(s["blockY"]), int(s["valueX"]),
             int(s["valueY"]), np.float32(s["score"]).view(np.uint32), np.float32(s["maxScore"]).view(np.uint32),
             np.float32(s["posScore"]).view(np.uint32)))
    assert set(o_sp) == set(h_sp), "pairs with starting points differ: {} vs {}".format(
        sorted(set(o_sp) - set(h_sp))[:5], sorted(set(h_sp) - set(o_sp))[:5])
    for k in o_sp:
        assert o_sp[k] == h_sp[k], "starting points differ for pair {}".format(k)
    # walk
    qids = qids or ["q%d" % i for i in range(N)]
    tids = tids or ["t%d" % i for i in range(T)]
    ws = O.walk_c(br.directions, br.starting_points, queries, targets, qids, tids, p.minimum)
    o_w = {}
    for w in ws:
        o_w.setdefault((w.qi, w.ti), []).append(
            (w.shown, w.s_start, w.s_end, w.t_start, w.t_end, w.matches, w.mismatches,
             np.float32(w.score).view(np.uint32), w.sequence_match if w.shown else "",
             w.alignment if w.shown else "", w.target_match if w.shown else ""))
    h_w = {}
    for h in res.hits:
        qi, ti = int(h["query"]), int(h["target"])
        shown = bool(h["shown"])
        sm = al = tm = ""
        if shown:
            sm, al, tm = decode_ops(res.ops, int(h["ops_offset"]), int(h["length"]), queries[qi], targets[ti],
                                    int(h["s_start"]), int(h["t_start"]))
            assert library_rows(res, h) == (sm, al, tm), "library alignment rows differ from the op codes"
        h_w.setdefault((qi, ti), []).append(
            (shown, int(h["s_start"]), int(h["s_end"]), int(h["t_start"]), int(h["t_end"]), int(h["matches"]),
             int(h["mismatches"]), np.float32(h["score"]).view(np.uint32), sm, al, tm))
    assert set(o_w) == set(h_w)
    for k in o_w:
        assert o_w[k] == h_w[k], "walk differs for pair {}:\n{}\n{}".format(k, o_w[k][:3], h_w[k][:3])
    return br


def library_rows(res, h):
    """The query, alignment and target rows the library built for hit h (pw_result.aligned)."""
    n, a0 = int(h["length"]), 3 * int(h["ops_offset"])
    rows = bytes(res.aligned[a0:a0 + 3 * n]).decode("latin-1")
    return rows[:n], rows[n:2 * n], rows[2 * n:]


def decode_ops(ops: bytes, off: int, n: int, q: str, t: str, s_start: int, t_start: int):
    x, y = s_start, t_start
    sm, al, tm = [], [], []
    for k in range(n - 1, -1, -1):  # recorded from the alignment end back to its start
        op = ops[off + k]
        if op == ord("D"):
            a, b = q[x], t[y]
            sm.append(a); tm.append(b); al.append("|" if a.lower() == b.lower() else ".")
            x += 1; y += 1
        elif op == ord("L"):
            sm.append(q[x]); tm.append("-"); al.append(" "); x += 1
        else:
            sm.append("-"); tm.append(t[y]); al.append(" "); y += 1
    return "".join(sm), "".join(al), "".join(tm)


def oracle_pair(O, q, t, qid, tid, p, mp=None, cparams=None):
    """One pair through the serial restatement and the walk on its own directions.

    Returns (maximum, starting points as tuples, walks as tuples) in the forms
    compare() uses.  Pairs are independent: the IN_ALIGNMENT marks of a walk
    stay inside its pair, and the pair's walks keep their relative order in the
    stable ascending sort of SmithWaterman.py:699, so per-pair results equal
    those of any batch holding the pair.
    """
    mx, d, sps = O.serial_pair(q, t, p, mp=mp, cparams=cparams)
    sp_t = [(int(s.blockX), int(s.blockY), int(s.valueX), int(s.valueY), np.float32(s.score).view(np.uint32),
             np.float32(s.maxScore).view(np.uint32), np.float32(s.posScore).view(np.uint32)) for s in sps]
    walks = []
    if sps:
        ws = O.walk_c(d.reshape(1, 1, len(q), len(t)).copy(), sps, [q], [t], [qid], [tid], p.minimum)
        walks = [(w.shown, w.s_start, w.s_end, w.t_start, w.t_end, w.matches, w.mismatches,
                  np.float32(w.score).view(np.uint32), w.sequence_match if w.shown else "",
                  w.alignment if w.shown else "", w.target_match if w.shown else "") for w in ws]
    return np.float32(mx), sp_t, walks


def compare_pairs(O, queries, targets, pq, pt, p, res, qids=None, tids=None, threads=16, mss=None):
    """Bit-exact parity of a HIP result over an explicit pair list against the
    serial oracle run pair by pair (in a thread pool; the C restatement
    releases the interpreter lock).  Returns the number of pairs with
    starting points."""
    import concurrent.futures as cf
    qids = qids or ["q%d" % i for i in range(len(queries))]
    tids = tids or ["t%d" % i for i in range(len(targets))]
    n = len(pq)
    hip_sp, hip_w = {}, {}
    for s in res.starting_points:
        hip_sp.setdefault((int(s["sequence"]), int(s["target"])), []).append(
            (int(s["blockX"]), int(s["blockY"]), int(s["valueX"]), int(s["valueY"]),
             np.float32(s["score"]).view(np.uint32), np.float32(s["maxScore"]).view(np.uint32),
             np.float32(s["posScore"]).view(np.uint32)))
    for h in res.hits:
        qi, ti = int(h["query"]), int(h["target"])
        shown = bool(h["shown"])
        sm = al = tm = ""
        if shown:
            sm, al, tm = decode_ops(res.ops, int(h["ops_offset"]), int(h["length"]), queries[qi], targets[ti],
                                    int(h["s_start"]), int(h["t_start"]))
            assert library_rows(res, h) == (sm, al, tm), "library alignment rows differ from the op codes"
        hip_w.setdefault((qi, ti), []).append(
            (shown, int(h["s_start"]), int(h["s_end"]), int(h["t_start"]), int(h["t_end"]), int(h["matches"]),
             int(h["mismatches"]), np.float32(h["score"]).view(np.uint32), sm, al, tm))

    cparams = O.CParams(p)

    def one(i):
        qi, ti = int(pq[i]), int(pt[i])
        mp = None if mss is None else float(mss[i])
        return i, oracle_pair(O, queries[qi], targets[ti], qids[qi], tids[ti], p, mp, cparams)

    with_sp = 0
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        for i, (mx, sp_t, walks) in ex.map(one, range(n)):
            key = (int(pq[i]), int(pt[i]))
            assert np.float32(res.maxima[i]).view(np.uint32) == mx.view(np.uint32), \
                "maximum differs for pair {}: {} vs {}".format(key, res.maxima[i], mx)
            assert hip_sp.get(key, []) == sp_t, "starting points differ for pair {}".format(key)
            assert hip_w.get(key, []) == walks, "walk differs for pair {}:\n{}\n{}".format(
                key, hip_w.get(key, [])[:3], walks[:3])
            with_sp += bool(sp_t)
    assert sum(len(v) for v in hip_sp.values()) == len(res.starting_points)
    return with_sp
