"""ctypes binding of libpypaswas_hip.so (C ABI declared in include/pypaswas_hip.h).

There is no CPU fallback: if the library is missing or no HIP device is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading
import weakref
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from pypaswas_amd.Core.Exceptions import HardwareException, InvalidOptionException

LIB_NAME = "libpypaswas_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PW_OK = 0
PW_EINVAL = -1
PW_ENOMEM = -2
PW_EDEVICE = -3

# device memory of pw_score_pairs_noalloc per GPU (PW_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

# exported entry points, in the order include/pypaswas_hip.h declares them
EXPORTS = ("pw_init", "pw_destroy", "pw_device_mem_bytes", "pw_set_scoring", "pw_align_pairs", "pw_prepare",
           "pw_execute", "pw_release_plan", "pw_plan_info", "pw_execute_graph", "pw_set_limits", "pw_set_routing", "pw_set_assembly",
           "pw_set_wait_limit", "pw_score_pairs",
           "pw_score_pairs_noalloc", "pw_autotune", "pw_get_tuning", "pw_set_tuning", "pw_tune_plan", "pw_server_create", "pw_server_set_scoring", "pw_server_submit", "pw_server_poll", "pw_server_take",
           "pw_server_stats", "pw_server_destroy", "pw_free_result", "pw_last_error")


class StartingPointRec(ctypes.Structure):
    """36-byte StartingPoint record (default_main_cpu.cl:66-76)."""
    _fields_ = [("sequence", ctypes.c_uint32), ("target", ctypes.c_uint32), ("blockX", ctypes.c_uint32),
                ("blockY", ctypes.c_uint32), ("valueX", ctypes.c_uint32), ("valueY", ctypes.c_uint32),
                ("score", ctypes.c_float), ("maxScore", ctypes.c_float), ("posScore", ctypes.c_float)]


class HitRec(ctypes.Structure):
    _fields_ = [("query", ctypes.c_int32), ("target", ctypes.c_int32), ("shown", ctypes.c_int32),
                ("s_start", ctypes.c_int32), ("s_end", ctypes.c_int32), ("t_start", ctypes.c_int32),
                ("t_end", ctypes.c_int32), ("matches", ctypes.c_int32), ("mismatches", ctypes.c_int32),
                ("length", ctypes.c_int32), ("score", ctypes.c_float), ("sp_index", ctypes.c_int32),
                ("ops_offset", ctypes.c_int64)]


class ResultRec(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int64), ("maxima", ctypes.POINTER(ctypes.c_float)),
                ("n_starting_points", ctypes.c_int64), ("starting_points", ctypes.POINTER(StartingPointRec)),
                ("n_hits", ctypes.c_int64), ("hits", ctypes.POINTER(HitRec)), ("n_ops", ctypes.c_int64),
                ("ops", ctypes.POINTER(ctypes.c_char)), ("n_traceback_pairs", ctypes.c_int64),
                ("sp_truncated", ctypes.c_int64), ("cells", ctypes.c_double), ("ms_score", ctypes.c_double),
                ("ms_directions", ctypes.c_double), ("ms_traceback", ctypes.c_double), ("ms_walk", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("score_launches", ctypes.c_int64),
                ("ms_readback", ctypes.c_double), ("ms_tail", ctypes.c_double), ("ms_assemble", ctypes.c_double),
                ("ms_plan", ctypes.c_double), ("ms_upload", ctypes.c_double),
                ("aligned", ctypes.POINTER(ctypes.c_char))]


_lib = None
_lock = threading.Lock()


def library_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Loads the HIP library; raises HardwareException when it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HardwareException(
                "{} is missing: build it with `python -m pypaswas_amd.csrc.build` "
                "(there is no CPU fallback)".format(LIB_PATH))
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.pw_init.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.pw_destroy.argtypes = [vp]
        L.pw_device_mem_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.pw_set_scoring.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.pw_align_pairs.argtypes = [vp, vp, vp, vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, vp,
                                     ctypes.c_int64, vp, vp, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(ctypes.POINTER(ResultRec))]
        L.pw_prepare.argtypes = [vp, vp, vp, vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int64,
                                 vp, vp, ctypes.c_int32, ctypes.POINTER(vp)]
        L.pw_execute.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.POINTER(ResultRec))]
        L.pw_release_plan.argtypes = [vp, vp]
        L.pw_plan_info.argtypes = [vp, vp, ctypes.c_int32]
        L.pw_execute_graph.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.POINTER(ResultRec))]
        L.pw_set_limits.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]
        L.pw_set_routing.argtypes = [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32]
        L.pw_set_assembly.argtypes = [vp, ctypes.c_int32]
        L.pw_set_wait_limit.argtypes = [vp, ctypes.c_int64]
        L.pw_score_pairs.argtypes = [vp, vp, vp, vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, vp,
                                     ctypes.c_int64, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_double)]
        L.pw_score_pairs_noalloc.argtypes = L.pw_score_pairs.argtypes
        L.pw_free_result.argtypes = [vp, ctypes.POINTER(ResultRec)]
        L.pw_autotune.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.pw_get_tuning.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32)]
        L.pw_set_tuning.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32]
        L.pw_tune_plan.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), vp]
        L.pw_server_create.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(vp)]
        L.pw_server_set_scoring.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_float]
        L.pw_server_submit.argtypes = [vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int32,
                                       vp, vp, ctypes.c_int64, vp, vp, ctypes.c_double, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int64)]
        L.pw_server_poll.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.pw_server_take.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.POINTER(ResultRec))]
        L.pw_server_stats.argtypes = [vp, vp, ctypes.c_int32]
        L.pw_server_destroy.argtypes = [vp]
        L.pw_last_error.argtypes = [vp]
        L.pw_last_error.restype = ctypes.c_char_p
        for name in EXPORTS:
            if name != "pw_last_error":
                getattr(L, name).restype = ctypes.c_int
        _lib = L
        return L


@dataclass
class PackedSeqs:
    data: np.ndarray    # uint8 residues
    off: np.ndarray     # int64
    length: np.ndarray  # int32

    @staticmethod
    def pack(seqs: Sequence[str]) -> "PackedSeqs":
        enc = [s.encode("latin-1") for s in seqs]
        length = np.array([len(e) for e in enc], dtype=np.int32)
        off = np.zeros(len(enc), dtype=np.int64)
        if len(enc) > 1:
            off[1:] = np.cumsum(length[:-1], dtype=np.int64)
        data = np.frombuffer(b"".join(enc) + b"\0", dtype=np.uint8).copy()
        return PackedSeqs(data, off, length)


@dataclass
class AlignResult:
    maxima: np.ndarray          # [n_pairs] f32
    starting_points: np.ndarray  # structured, 36-byte records
    hits: np.ndarray             # structured
    ops: memoryview             # one byte per alignment column ('D', 'L', 'U'), bytes-like
    n_traceback_pairs: int
    sp_truncated: int
    cells: float
    ms_score: float
    ms_directions: float
    ms_traceback: float
    ms_walk: float
    ms_total: float
    score_launches: int
    ms_readback: float = 0.0
    ms_tail: float = 0.0
    ms_assemble: float = 0.0
    ms_plan: float = 0.0
    ms_upload: float = 0.0
    aligned: memoryview = memoryview(b"")  # 3 bytes per op: query, alignment and target rows of each shown hit


SP_DTYPE = np.dtype([("sequence", "<u4"), ("target", "<u4"), ("blockX", "<u4"), ("blockY", "<u4"),
                     ("valueX", "<u4"), ("valueY", "<u4"), ("score", "<f4"), ("maxScore", "<f4"),
                     ("posScore", "<f4")])
HIT_DTYPE = np.dtype([("query", "<i4"), ("target", "<i4"), ("shown", "<i4"), ("s_start", "<i4"), ("s_end", "<i4"),
                      ("t_start", "<i4"), ("t_end", "<i4"), ("matches", "<i4"), ("mismatches", "<i4"),
                      ("length", "<i4"), ("score", "<f4"), ("sp_index", "<i4"), ("ops_offset", "<i8")])
assert SP_DTYPE.itemsize == ctypes.sizeof(StartingPointRec) == 36
assert HIT_DTYPE.itemsize == ctypes.sizeof(HitRec)


def _release_result(dev_ref, lib, out):
    """Hands a result back to the library (its pinned buffers are recycled by a live device)."""
    dev = dev_ref()
    ctx = dev.ctx if dev is not None and dev.ctx else None
    lib.pw_free_result(ctx, out)


class _ResultHold:
    """Keeps a library result alive while any view of its buffers lives."""
    __slots__ = ("__weakref__",)


def _view(ptr, count: int, dtype: np.dtype, hold: _ResultHold) -> np.ndarray:
    """count records of dtype at a library pointer, as an array that keeps hold alive."""
    if count <= 0:
        return np.zeros(0, dtype)
    buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(ctypes.cast(ptr, ctypes.c_void_p).value)
    buf._hold = hold
    return np.frombuffer(buf, dtype=dtype)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class Device:
    """One library context on one GPU (single host thread, like the reference)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        rc = self.lib.pw_init(int(device), ctypes.byref(self.ctx))
        if rc != PW_OK:
            msg = self.lib.pw_last_error(None).decode(errors="replace")
            if rc == PW_EINVAL:
                raise InvalidOptionException(msg)
            raise HardwareException("HIP device {} unavailable: {}".format(device, msg))

    def _check(self, rc: int, what: str):
        if rc == PW_OK:
            return
        msg = "{}: {}".format(what, self.lib.pw_last_error(self.ctx).decode(errors="replace"))
        if rc == PW_EINVAL:
            raise InvalidOptionException(msg)
        if rc == PW_ENOMEM:
            raise MemoryError(msg)
        raise HardwareException(msg)

    def close(self):
        if self.ctx:
            self.lib.pw_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def global_mem_size(self) -> int:
        v = ctypes.c_uint64()
        self._check(self.lib.pw_device_mem_bytes(self.ctx, ctypes.byref(v)), "pw_device_mem_bytes")
        return int(v.value)

    def set_scoring(self, subst: np.ndarray, char_offset: int, gap: float, gap_ext: float, lower_limit: float,
                    minimum_score: float):
        m = np.ascontiguousarray(subst, dtype=np.float32)
        self._keep_subst = m
        self._check(self.lib.pw_set_scoring(self.ctx, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            int(m.shape[0]), int(char_offset), float(np.float32(gap)),
                                            float(np.float32(gap_ext)), float(np.float32(lower_limit)),
                                            float(np.float32(minimum_score))), "pw_set_scoring")

    def align_pairs(self, queries: PackedSeqs, targets: PackedSeqs, pq: np.ndarray, pt: np.ndarray,
                    min_start_score: np.ndarray, same_id: Optional[np.ndarray], minimum: float,
                    sp_cap: int = 100, mode: int = 1) -> AlignResult:
        pq = np.ascontiguousarray(pq, dtype=np.int32)
        pt = np.ascontiguousarray(pt, dtype=np.int32)
        mss = np.ascontiguousarray(min_start_score, dtype=np.float32)
        sid = None if same_id is None else np.ascontiguousarray(same_id, dtype=np.uint8)
        out = ctypes.POINTER(ResultRec)()
        rc = self.lib.pw_align_pairs(self.ctx, _ptr(queries.data), _ptr(queries.off), _ptr(queries.length),
                                     len(queries.length), _ptr(targets.data), _ptr(targets.off),
                                     _ptr(targets.length), len(targets.length), _ptr(pq), _ptr(pt), len(pq),
                                     _ptr(mss), _ptr(sid), float(minimum), int(sp_cap), int(mode),
                                     ctypes.byref(out))
        self._check(rc, "pw_align_pairs")
        return self._take_result(out)

    def _take_result(self, out) -> AlignResult:
        """Wraps a library result without copying: maxima, starting points,
        hits and ops are views of the library's buffers, which go back to the
        library once the last of these views is collected."""
        r = out.contents
        hold = _ResultHold()
        fin = weakref.finalize(hold, _release_result, weakref.ref(self), self.lib, out)
        fin.atexit = False
        n = int(r.n_pairs)
        maxima = _view(r.maxima, n, np.dtype(np.float32), hold)
        sps = _view(r.starting_points, int(r.n_starting_points), SP_DTYPE, hold)
        hits = _view(r.hits, int(r.n_hits), HIT_DTYPE, hold)
        n_ops = int(r.n_ops)
        if n_ops:
            buf = (ctypes.c_char * n_ops).from_address(ctypes.cast(r.ops, ctypes.c_void_p).value)
            buf._hold = hold
            ops = memoryview(buf).cast("B")
            abuf = (ctypes.c_char * (3 * n_ops)).from_address(ctypes.cast(r.aligned, ctypes.c_void_p).value)
            abuf._hold = hold
            aligned = memoryview(abuf).cast("B")
        else:
            ops = memoryview(b"")
            aligned = memoryview(b"")
        res = AlignResult(maxima, sps, hits, ops, int(r.n_traceback_pairs), int(r.sp_truncated),
                          float(r.cells), float(r.ms_score), float(r.ms_directions), float(r.ms_traceback),
                          float(r.ms_walk), float(r.ms_total), int(r.score_launches), float(r.ms_readback),
                          float(r.ms_tail), float(r.ms_assemble), float(r.ms_plan), float(r.ms_upload), aligned)
        # with no view left holding it (an empty result), the record goes back here
        del hold
        return res

    def prepare(self, queries: PackedSeqs, targets: PackedSeqs, pq: np.ndarray, pt: np.ndarray,
                min_start_score: Optional[np.ndarray], same_id: Optional[np.ndarray], mode: int = 1) -> "Plan":
        pq = np.ascontiguousarray(pq, dtype=np.int32)
        pt = np.ascontiguousarray(pt, dtype=np.int32)
        mss = None if min_start_score is None else np.ascontiguousarray(min_start_score, dtype=np.float32)
        sid = None if same_id is None else np.ascontiguousarray(same_id, dtype=np.uint8)
        handle = ctypes.c_void_p()
        rc = self.lib.pw_prepare(self.ctx, _ptr(queries.data), _ptr(queries.off), _ptr(queries.length),
                                 len(queries.length), _ptr(targets.data), _ptr(targets.off), _ptr(targets.length),
                                 len(targets.length), _ptr(pq), _ptr(pt), len(pq), _ptr(mss), _ptr(sid), int(mode),
                                 ctypes.byref(handle))
        self._check(rc, "pw_prepare")
        return Plan(self, handle)

    def execute(self, plan: "Plan", minimum: float, sp_cap: int = 100, graph: bool = False) -> AlignResult:
        """Runs a prepared plan; graph=True replays a captured HIP graph of its device work."""
        out = ctypes.POINTER(ResultRec)()
        fn = self.lib.pw_execute_graph if graph else self.lib.pw_execute
        self._check(fn(self.ctx, plan.handle, float(minimum), int(sp_cap), ctypes.byref(out)),
                    "pw_execute_graph" if graph else "pw_execute")
        return self._take_result(out)

    def autotune(self) -> float:
        """Measures the kernel shapes for the current scheme's gap model (pw_autotune); returns seconds."""
        sec = ctypes.c_double()
        self._check(self.lib.pw_autotune(self.ctx, ctypes.byref(sec)), "pw_autotune")
        return float(sec.value)

    def tuning(self, affine: bool) -> dict:
        """The tuning table of a gap model (pw_get_tuning)."""
        rates = np.zeros((7, 7, 4), dtype=np.float32)
        grates = np.zeros(5, dtype=np.float32)
        nb = ctypes.c_int32()
        valid = ctypes.c_int32()
        self._check(self.lib.pw_get_tuning(self.ctx, 1 if affine else 0, rates.ctypes.data, grates.ctypes.data,
                                           ctypes.byref(nb), ctypes.byref(valid)), "pw_get_tuning")
        return {"valid": bool(valid.value), "rates": rates, "group_rates": grates, "group_nb_long": int(nb.value)}

    def set_tuning(self, affine: bool, rates: Optional[np.ndarray], group_nb_long: int = 80):
        """Installs a saved table (pw_set_tuning); rates=None restores the cost model."""
        r = None if rates is None else np.ascontiguousarray(rates, dtype=np.float32).reshape(7, 7, 4)
        self._check(self.lib.pw_set_tuning(self.ctx, 1 if affine else 0, None if r is None else r.ctypes.data,
                                           int(group_nb_long)), "pw_set_tuning")

    def tune_plan(self, plan: "Plan") -> dict:
        """Times the plan's scoring pass under the routing choices and keeps the fastest (pw_tune_plan)."""
        sec = ctypes.c_double()
        ms = ctypes.c_double()
        choice = np.zeros(3, dtype=np.int32)
        self._check(self.lib.pw_tune_plan(self.ctx, plan.handle, ctypes.byref(sec), ctypes.byref(ms),
                                          choice.ctypes.data), "pw_tune_plan")
        return {"seconds": round(float(sec.value), 3), "score_ms": round(float(ms.value), 3),
                "group_max_share": float(choice[0]) / 1000.0, "group_chunk_width": int(choice[1]),
                "measured_shapes": bool(choice[2])}

    def set_limits(self, max_direction_bytes: int = 0, candidates_per_pair: int = 0, ops_bytes: int = 0,
                   grow: bool = True):
        """Work-area sizing of the traceback stages (pw_set_limits); 0 keeps a setting."""
        self._check(self.lib.pw_set_limits(self.ctx, int(max_direction_bytes), int(candidates_per_pair),
                                           int(ops_bytes), 1 if grow else 0), "pw_set_limits")

    def set_assembly(self, on_device: bool = True):
        """Result layout in pair order on the device (default) or on the host (pw_set_assembly)."""
        self._check(self.lib.pw_set_assembly(self.ctx, 1 if on_device else 0), "pw_set_assembly")

    def set_routing(self, orient: int = -1, group_max_share: float = 0.0, dir_residues: int = 0):
        """Routing rules of plans prepared afterwards (pw_set_routing); defaults restore them."""
        self._check(self.lib.pw_set_routing(self.ctx, int(orient), float(group_max_share), int(dir_residues)),
                    "pw_set_routing")

    def set_wait_limit(self, polls: int = -1):
        """Polls before a device hand-over wait gives up (pw_set_wait_limit); -1 restores the default."""
        self._check(self.lib.pw_set_wait_limit(self.ctx, int(polls)), "pw_set_wait_limit")

    def score_pairs(self, queries: PackedSeqs, targets: PackedSeqs, pq: np.ndarray, pt: np.ndarray,
                    mode: int = 1):
        pq = np.ascontiguousarray(pq, dtype=np.int32)
        pt = np.ascontiguousarray(pt, dtype=np.int32)
        out = np.zeros(len(pq), dtype=np.float32)
        ms = ctypes.c_double()
        rc = self.lib.pw_score_pairs(self.ctx, _ptr(queries.data), _ptr(queries.off), _ptr(queries.length),
                                     len(queries.length), _ptr(targets.data), _ptr(targets.off),
                                     _ptr(targets.length), len(targets.length), _ptr(pq), _ptr(pt), len(pq),
                                     int(mode), _ptr(out), ctypes.byref(ms))
        self._check(rc, "pw_score_pairs")
        return out, float(ms.value)

    def score_pairs_noalloc(self, queries: PackedSeqs, targets: PackedSeqs, pq: np.ndarray, pt: np.ndarray,
                            mode: int = 1):
        """score_pairs through pw_score_pairs_noalloc: no device memory allocated, the
        library's static scratch of NOALLOC_SCRATCH_BYTES per GPU instead."""
        pq = np.ascontiguousarray(pq, dtype=np.int32)
        pt = np.ascontiguousarray(pt, dtype=np.int32)
        out = np.zeros(len(pq), dtype=np.float32)
        ms = ctypes.c_double()
        rc = self.lib.pw_score_pairs_noalloc(self.ctx, _ptr(queries.data), _ptr(queries.off), _ptr(queries.length),
                                             len(queries.length), _ptr(targets.data), _ptr(targets.off),
                                             _ptr(targets.length), len(targets.length), _ptr(pq), _ptr(pt), len(pq),
                                             int(mode), _ptr(out), ctypes.byref(ms))
        self._check(rc, "pw_score_pairs_noalloc")
        return out, float(ms.value)


class Plan:
    """Device-resident sequences and work list (pw_prepare); release() frees them."""

    def __init__(self, dev: Device, handle):
        self.dev = dev
        self.handle = handle

    def info(self) -> dict:
        """Routing of the plan (pw_plan_info)."""
        v = np.zeros(10, dtype=np.int64)
        self.dev._check(self.dev.lib.pw_plan_info(self.handle, v.ctypes.data, 10), "pw_plan_info")
        keys = ("pairs", "lane_per_pair_items", "wave_per_pair_items", "orientation", "lane_per_pair_cells",
                "wave_per_pair_cells", "max_chunk_width", "direction_residues_per_lane", "split_job_items",
                "checkpointed_pairs")
        return {k: int(x) for k, x in zip(keys, v)}

    def release(self):
        if self.handle:
            self.dev.lib.pw_release_plan(self.dev.ctx, self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
