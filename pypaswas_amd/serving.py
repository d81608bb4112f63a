"""Serving: alignment requests through the library's continuous-batching loop.

pyPaSWAS itself runs one job per process (pyPaSWAS/pypaswasall.py:98-124 ->
SmithWaterman.align_sequences, pyPaSWAS/Core/SmithWaterman.py:441-560).  Here a
request is such a job -- query and target records aligned all-vs-all with one
scoring scheme -- submitted to a long-lived server (``pw_server_*`` in
include/pypaswas_hip.h).  The server merges queued jobs of one scheme into a
batch whenever one of its stream lanes is idle and the direction arena has
room, runs batches on several lanes concurrently, and returns each job's
records exactly as the same job run alone would produce them.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from pypaswas_amd._native import AlignResult, Device, PackedSeqs, ResultRec, _ptr

Records = Sequence[Tuple[str, str]]


@dataclass
class Scheme:
    """Scoring constants of one scheme, as pw_set_scoring takes them."""
    subst: np.ndarray        # dim x dim f32, row = query residue
    char_offset: int
    gap: float
    gap_ext: float           # 0 selects the linear-gap kernels
    lower_limit: float
    minimum: float
    highest: float           # Score.highest_score, for the minimum start scores
    filter_factor: float

    @classmethod
    def from_score(cls, score, filter_factor: float) -> "Scheme":
        """From a Core.Scores object (the reference's Score interface)."""
        return cls(score.to_float32(), ord(score.char_offset), float(score.gap_score),
                   float(score.gap_extension) if score.gap_extension is not None else 0.0,
                   float(score.lower_limit_score), float(score.minimum_score), float(score.highest_score),
                   float(filter_factor))


@dataclass
class Job:
    """One request: all-vs-all pairs of its queries and targets under one scheme."""
    scheme: int
    queries: PackedSeqs
    targets: PackedSeqs
    pq: np.ndarray
    pt: np.ndarray
    mss: np.ndarray
    same: np.ndarray
    minimum: float
    cells: float
    sp_cap: int = 100

    @classmethod
    def all_vs_all(cls, scheme_id: int, scheme: Scheme, queries: Records, targets: Records, sp_cap: int = 100):
        from pypaswas_amd.Core.SmithWatermanHIP import job_arrays
        qp, tp, pq, pt, mss, same = job_arrays(queries, targets, scheme.highest, scheme.filter_factor)
        cells = float(np.sum(qp.length[pq].astype(np.float64) * tp.length[pt].astype(np.float64)))
        return cls(scheme_id, qp, tp, pq, pt, mss, same, scheme.minimum, cells, sp_cap)


class Server:
    """Continuous-batching server on one device (pw_server_*)."""

    def __init__(self, dev: Device, lanes: int = 4, direction_bytes: int = 4 << 30):
        self.dev = dev
        self.lib = dev.lib
        self.handle = ctypes.c_void_p()
        dev._check(self.lib.pw_server_create(dev.ctx, int(lanes), int(direction_bytes), ctypes.byref(self.handle)),
                   "pw_server_create")
        self._ids = np.zeros(1024, dtype=np.int64)

    def set_scheme(self, scheme_id: int, s: Scheme):
        m = np.ascontiguousarray(s.subst, dtype=np.float32)
        self.dev._check(self.lib.pw_server_set_scoring(self.handle, int(scheme_id),
                                                       m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                       int(m.shape[0]), int(s.char_offset), float(s.gap),
                                                       float(s.gap_ext), float(s.lower_limit), float(s.minimum)),
                        "pw_server_set_scoring")

    def submit(self, job: Job) -> int:
        jid = ctypes.c_int64()
        q, t = job.queries, job.targets
        rc = self.lib.pw_server_submit(self.handle, int(job.scheme), _ptr(q.data), _ptr(q.off), _ptr(q.length),
                                       len(q.length), _ptr(t.data), _ptr(t.off), _ptr(t.length), len(t.length),
                                       _ptr(job.pq), _ptr(job.pt), len(job.pq), _ptr(job.mss), _ptr(job.same),
                                       float(job.minimum), int(job.sp_cap), ctypes.byref(jid))
        self.dev._check(rc, "pw_server_submit")
        return int(jid.value)

    def poll(self, block: bool = False) -> List[int]:
        n = ctypes.c_int32()
        self.dev._check(self.lib.pw_server_poll(self.handle, 1 if block else 0, self._ids.ctypes.data,
                                                len(self._ids), ctypes.byref(n)), "pw_server_poll")
        return [int(v) for v in self._ids[:n.value]]

    def take(self, job_id: int) -> AlignResult:
        out = ctypes.POINTER(ResultRec)()
        self.dev._check(self.lib.pw_server_take(self.handle, int(job_id), ctypes.byref(out)), "pw_server_take")
        return self.dev._take_result(out)

    def stats(self) -> Dict[str, int]:
        v = np.zeros(11, dtype=np.int64)
        self.dev._check(self.lib.pw_server_stats(self.handle, v.ctypes.data, 11), "pw_server_stats")
        keys = ("batches", "admitted", "completed", "peak_arena_bytes", "queued", "running", "arena_bytes",
                "host_us_launch", "host_us_finish", "device_us", "reallocs")
        return {k: int(x) for k, x in zip(keys, v)}

    def close(self):
        if self.handle:
            self.lib.pw_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(n: int, qps: float, seed: int = 5) -> np.ndarray:
    """Arrival times (seconds from 0) of a Poisson stream at rate qps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, size=n))


def replay(server: Server, jobs: Sequence[Job], arrivals: Optional[np.ndarray] = None,
           watch: Optional[Dict[str, int]] = None):
    """Submits each job at its arrival time (all at once when arrivals is None),
    polls until every job completed.  Returns (results, latencies_s, wall_s)
    with results and latencies in job order; a job's latency runs from its
    arrival time to the poll that returned it.  With a dict `watch`, the
    largest queued and running job counts the server reported while the jobs
    ran are kept in it ("peak_queued", "peak_running")."""
    n = len(jobs)
    arr = np.zeros(n) if arrivals is None else np.asarray(arrivals, dtype=np.float64)
    results: List[Optional[AlignResult]] = [None] * n
    lat = np.zeros(n)
    pending: Dict[int, int] = {}
    i = 0
    t0 = time.perf_counter()
    while i < n or pending:
        now = time.perf_counter() - t0
        while i < n and arr[i] <= now:
            pending[server.submit(jobs[i])] = i
            i += 1
        done = server.poll(block=(i >= n))
        t = time.perf_counter() - t0
        if watch is not None:
            st = server.stats()
            watch["peak_queued"] = max(watch.get("peak_queued", 0), st["queued"])
            watch["peak_running"] = max(watch.get("peak_running", 0), st["running"])
        for jid in done:
            k = pending.pop(jid)
            lat[k] = t - arr[k]
            results[k] = server.take(jid)
        if not done and i < n:
            # idle until the next arrival, but keep polling running batches
            gap = arr[i] - (time.perf_counter() - t0)
            if gap > 2e-4:
                time.sleep(min(gap, 1e-4))
    wall = time.perf_counter() - t0
    return results, lat, wall
