"""ctypes binding of libpolymult.so (include/polymult.h).

The Python face of the drop-in: the same three calls as the reference's
fherma_init / fherma_run / fherma_free (fherma.h:46-48), here
Multiplier(N, W, q) / Multiplier.run(a, b) / Multiplier.close(), with the
same argument meaning (N x L little-endian u32 limbs) and the same error
behaviour (a bad point or operand raises with the reference's message).

There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpolymult.so"

# Every C symbol include/polymult.h declares.
C_SYMBOLS = (
    "pm_init", "pm_run", "pm_run_device", "pm_run_device_scratch", "pm_free", "pm_last_error", "pm_query", "pm_point_info",
    "pm_dev_alloc", "pm_dev_free", "pm_copy_to_device", "pm_copy_to_host", "pm_stream_create", "pm_stream_sync",
    "pm_stream_destroy", "pm_sync", "pm_time_device",
    "pm_shard_bounds", "pm_run_device_sharded", "pm_time_shard",
    "pm_stage_rns", "pm_stage_ntt_fwd", "pm_stage_ntt_inv", "pm_stage_crt", "pm_probe_transfers", "pm_probe_copy",
    "pm_probe_valu",
    "pm_batch_create", "pm_batch_free", "pm_batch_last_error", "pm_batch_product_info", "pm_batch_run",
    "pm_batch_sync", "pm_batch_time", "pm_batch_autotune", "pm_batch_configure",
    "pm_server_create", "pm_server_free", "pm_server_last_error", "pm_server_submit", "pm_server_wait",
    "pm_server_stats_get", "pm_server_drive", "pm_server_saturate",
)
# The C++ entry points of include/fherma.h, mangled as the reference's harness binds them.
CXX_SYMBOLS = (
    "_Z11fherma_initRKN6fherma5PointE",
    "_Z10fherma_runPvRKN6fherma6InputsE",
    "_Z11fherma_freePv",
)


class PolymultError(RuntimeError):
    pass


class PmInfo(ctypes.Structure):
    _fields_ = [("N", ctypes.c_uint32), ("W", ctypes.c_uint32), ("L", ctypes.c_uint32), ("K", ctypes.c_uint32),
                ("delta", ctypes.c_uint64), ("w1", ctypes.c_uint32), ("w2", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device_budget", ctypes.c_uint64), ("pinned_budget", ctypes.c_uint64),
                ("max_batch", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32),
                ("host_threads", ctypes.c_uint32), ("col_tile", ctypes.c_int32),
                ("max_batch_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "failed", "batches", "deferred")] + \
               [(n, ctypes.c_double) for n in ("mean_batch", "p50_ms", "p99_ms", "max_ms", "mean_ms", "stage_ms",
                                               "issue_ms", "gpu_ms", "scatter_ms", "sched_busy",
                                               "completer_busy")] + \
               [(n, ctypes.c_uint64) for n in ("device_budget", "device_peak", "pinned_budget", "pinned_peak")]


class ServeReport(ctypes.Structure):
    _fields_ = [("offered_rate", ctypes.c_double), ("seconds", ctypes.c_double)] + \
               [(n, ctypes.c_uint64) for n in ("requests", "completed", "failed")] + \
               [(n, ctypes.c_double) for n in ("achieved_rate", "p50_ms", "p99_ms", "max_ms", "mean_ms",
                                               "mean_batch")] + \
               [(n, ctypes.c_uint64) for n in ("batches", "deferred", "device_peak")] + \
               [(n, ctypes.c_double) for n in ("max_late_ms", "stage_ms", "issue_ms", "gpu_ms", "scatter_ms",
                                               "sched_busy", "completer_busy", "queue_p50_ms", "queue_p99_ms",
                                               "service_p50_ms", "service_p99_ms")] + \
               [("point_p50_ms", ctypes.c_double * 8), ("point_p99_ms", ctypes.c_double * 8),
                ("cpu_util", ctypes.c_double), ("ctx_invol", ctypes.c_uint64), ("driver_ctx_invol", ctypes.c_uint64),
                ("verified", ctypes.c_uint64), ("mismatches", ctypes.c_uint64)]


def _as_dict(st: ctypes.Structure) -> dict:
    out = {}
    for name, _ in st._fields_:
        v = getattr(st, name)
        out[name] = list(v) if isinstance(v, ctypes.Array) else v
    return out


_LIB = None
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_vpp = ctypes.POINTER(ctypes.c_void_p)


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise PolymultError("libpolymult.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH))
    sig = {
        "pm_init": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_size_t,
                     ctypes.POINTER(_vp)], ctypes.c_int),
        "pm_run": ([_vp, _u32p, _u32p, _u32p], ctypes.c_int),
        "pm_run_device": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        "pm_run_device_scratch": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        "pm_free": ([_vp], None),
        "pm_last_error": ([], ctypes.c_char_p),
        "pm_query": ([_vp, ctypes.POINTER(PmInfo)], ctypes.c_int),
        "pm_point_info": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p,
                           ctypes.POINTER(ctypes.c_uint64), _u32p, ctypes.c_size_t], ctypes.c_int),
        "pm_dev_alloc": ([ctypes.c_size_t, ctypes.POINTER(_vp)], ctypes.c_int),
        "pm_dev_free": ([_vp], ctypes.c_int),
        "pm_copy_to_device": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
        "pm_copy_to_host": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
        "pm_stream_create": ([ctypes.POINTER(_vp)], ctypes.c_int),
        "pm_stream_sync": ([_vp], ctypes.c_int),
        "pm_stream_destroy": ([_vp], ctypes.c_int),
        "pm_sync": ([_vp], ctypes.c_int),
        "pm_time_device": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "pm_stage_rns": ([_vp, _vp, _vp], ctypes.c_int),
        "pm_stage_ntt_fwd": ([_vp, _vp], ctypes.c_int),
        "pm_stage_ntt_inv": ([_vp, _vp], ctypes.c_int),
        "pm_stage_crt": ([_vp, _vp, _vp], ctypes.c_int),
        "pm_probe_transfers": ([ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                               ctypes.c_int),
        "pm_probe_copy": ([ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "pm_probe_valu": ([ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "pm_shard_bounds": ([_vp, ctypes.c_int, _u32p], ctypes.c_int),
        "pm_run_device_sharded": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
        "pm_time_shard": ([_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
                          ctypes.c_int),
        "pm_batch_create": ([ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.POINTER(_vp)], ctypes.c_int),
        "pm_batch_free": ([_vp], None),
        "pm_batch_last_error": ([], ctypes.c_char_p),
        "pm_batch_product_info": ([_vp, ctypes.c_uint32, ctypes.POINTER(PmInfo)], ctypes.c_int),
        "pm_batch_run": ([_vp, _vpp, _vpp, _vpp, _vp], ctypes.c_int),
        "pm_batch_sync": ([_vp], ctypes.c_int),
        "pm_batch_time": ([_vp, _vpp, _vpp, _vpp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "pm_batch_autotune": ([_vp, _vpp, _vpp, _vpp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "pm_batch_configure": ([_vp] + [ctypes.c_int] * 5, ctypes.c_int),
        "pm_server_create": ([ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.POINTER(ServerConfig),
                              ctypes.POINTER(_vp)], ctypes.c_int),
        "pm_server_free": ([_vp], None),
        "pm_server_last_error": ([], ctypes.c_char_p),
        "pm_server_submit": ([_vp, ctypes.c_uint32, _u32p, _u32p, _u32p, ctypes.POINTER(ctypes.c_uint64)],
                             ctypes.c_int),
        "pm_server_wait": ([_vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "pm_server_stats_get": ([_vp, ctypes.c_int, ctypes.POINTER(ServerStats)], ctypes.c_int),
        "pm_server_drive": ([_vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                             ctypes.POINTER(ServeReport)], ctypes.c_int),
        "pm_server_saturate": ([_vp, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.c_uint64, ctypes.POINTER(ServeReport)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _LIB = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise PolymultError(load().pm_last_error().decode())


def _u32(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def point_info(N: int, W: int, L: int, q: np.ndarray, cap: int = 1024):
    """Host-only validation: (K, delta, primes). Raises like pm_init."""
    lib = load()
    q = _u32(q)
    K = ctypes.c_uint32()
    delta = ctypes.c_uint64()
    primes = np.zeros(cap, dtype=np.uint32)
    _check(lib.pm_point_info(N, W, L, _ptr(q), q.size, ctypes.byref(K), ctypes.byref(delta), _ptr(primes), cap))
    return K.value, delta.value, primes[:K.value].copy()


def probe_transfers(nbytes: int, threads: int = 16, reps: int = 5) -> dict:
    """PCIe and host staging rates in GB/s (see pm_probe_transfers)."""
    out = (ctypes.c_double * 7)()
    _check(load().pm_probe_transfers(nbytes, threads, reps, out))
    keys = ["h2d_pinned", "d2h_pinned", "h2d_pageable", "d2h_pageable", "stage_1thread", "stage_threads",
            "duplex_pinned"]
    return {k: round(v, 2) for k, v in zip(keys, out)}


def probe_copy(nbytes: int, iters: int = 10) -> float:
    """GB/s (read + write) of a plain device copy kernel over nbytes (see pm_probe_copy)."""
    g = ctypes.c_double()
    _check(load().pm_probe_copy(nbytes, iters, ctypes.byref(g)))
    return g.value


def probe_valu(iters: int = 256) -> dict:
    """Issue rates (tera lane-operations/s) of v_add_u32, v_mul_lo_u32,
    v_mul_hi_u32 and v_mad_u64_u32 (see pm_probe_valu)."""
    out = (ctypes.c_double * 4)()
    _check(load().pm_probe_valu(iters, out))
    return dict(zip(("add_u32", "mul_lo_u32", "mul_hi_u32", "mad_u64_u32"), [round(v, 3) for v in out]))


class DeviceBuffer:
    """Device memory owned through libpolymult's own HIP runtime."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        p = _vp()
        _check(load().pm_dev_alloc(nbytes, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(load().pm_copy_to_device(self.ptr, a.ctypes.data_as(_vp), a.nbytes))

    def download(self, shape, dtype=np.uint32) -> np.ndarray:
        """Synchronous copy to host; sync() the work that writes this buffer first."""
        out = np.empty(shape, dtype=dtype)
        _check(load().pm_copy_to_host(out.ctypes.data_as(_vp), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            _check(load().pm_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A hipStream_t from libpolymult's own HIP runtime, to pass as `stream`."""

    def __init__(self):
        p = _vp()
        _check(load().pm_stream_create(ctypes.byref(p)))
        self.handle = p

    def sync(self):
        _check(load().pm_stream_sync(self.handle))

    def close(self):
        if self.handle:
            _check(load().pm_stream_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Multiplier:
    """One point (N, W, q): the fherma_init / fherma_run / fherma_free triple."""

    def __init__(self, N: int, W: int, q: np.ndarray, L: int | None = None):
        lib = load()
        L = (W + 31) // 32 if L is None else L
        q = _u32(q).reshape(-1)
        h = _vp()
        _check(lib.pm_init(N, W, L, _ptr(q), q.size, ctypes.byref(h)))
        self._h = h
        info = PmInfo()
        _check(lib.pm_query(h, ctypes.byref(info)))
        self.info = info
        self.N, self.W, self.L, self.K = info.N, info.W, info.L, info.K

    def close(self):
        if getattr(self, "_h", None):
            load().pm_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the multiply ---------------------------------------------------------
    def run(self, a: np.ndarray, b: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        a, b = _u32(a), _u32(b)
        words = self.N * self.L
        if a.size != words or b.size != words:
            raise PolymultError("operand size does not match N x L")
        c = np.empty((self.N, self.L), dtype=np.uint32) if out is None else out
        _check(load().pm_run(self._h, _ptr(a), _ptr(b), _ptr(c)))
        return c

    def run_device(self, da: DeviceBuffer, db: DeviceBuffer, dc: DeviceBuffer, stream=None):
        """Enqueues c = a b on device buffers and returns. Without a stream,
        consecutive calls overlap on internal lanes (results still land
        in call order); call sync() before reading dc or changing da, db."""
        _check(load().pm_run_device(self._h, da.ptr, db.ptr, dc.ptr, stream))

    def run_device_scratch(self, da: DeviceBuffer, db: DeviceBuffer, dc: DeviceBuffer, stream=None):
        """The allocation-free device call (pm_run_device_scratch)."""
        _check(load().pm_run_device_scratch(self._h, da.ptr, db.ptr, dc.ptr, stream))

    def sync(self):
        """Waits for every device call of this multiplier."""
        _check(load().pm_sync(self._h))

    def time_device(self, da: DeviceBuffer, db: DeviceBuffer, dc: DeviceBuffer, iters: int) -> list[float]:
        ms = (ctypes.c_float * 6)()
        _check(load().pm_time_device(self._h, da.ptr, db.ptr, dc.ptr, iters, ms))
        return list(ms)

    # -- prime shards ---------------------------------------------------------
    def shard_bounds(self, shards: int) -> list[int]:
        b = np.zeros(shards + 1, dtype=np.uint32)
        _check(load().pm_shard_bounds(self._h, shards, _ptr(b)))
        return [int(x) for x in b]

    def run_device_sharded(self, da: DeviceBuffer, db: DeviceBuffer, dc: DeviceBuffer, shards: int, stream=None):
        _check(load().pm_run_device_sharded(self._h, da.ptr, db.ptr, dc.ptr, shards, stream))

    def time_shard(self, da: DeviceBuffer, db: DeviceBuffer, dc: DeviceBuffer, shards: int, iters: int) -> dict:
        ms = (ctypes.c_float * 7)()
        _check(load().pm_time_shard(self._h, da.ptr, db.ptr, dc.ptr, shards, iters, ms))
        keys = ["sharded_multiply", "rns", "col_fwd", "row_fused", "col_inv", "crt_partial", "combine"]
        return dict(zip(keys, list(ms)))

    # -- stage entry points ---------------------------------------------------
    def stage_rns(self, a: np.ndarray) -> np.ndarray:
        da = DeviceBuffer.from_array(_u32(a))
        dR = DeviceBuffer(self.K * self.N * 4)
        _check(load().pm_stage_rns(self._h, da.ptr, dR.ptr))
        return dR.download((self.K, self.N))

    def stage_ntt_fwd(self, R: np.ndarray) -> np.ndarray:
        dR = DeviceBuffer.from_array(_u32(R))
        _check(load().pm_stage_ntt_fwd(self._h, dR.ptr))
        return dR.download((self.K, self.N))

    def stage_ntt_inv(self, R: np.ndarray) -> np.ndarray:
        dR = DeviceBuffer.from_array(_u32(R))
        _check(load().pm_stage_ntt_inv(self._h, dR.ptr))
        return dR.download((self.K, self.N))

    def stage_crt(self, X: np.ndarray) -> np.ndarray:
        dX = DeviceBuffer.from_array(_u32(X))
        dc = DeviceBuffer(self.N * self.L * 4)
        _check(load().pm_stage_crt(self._h, dX.ptr, dc.ptr))
        return dc.download((self.N, self.L))


def _check_batch(rc: int):
    if rc != 0:
        raise PolymultError(load().pm_batch_last_error().decode())


class Batch:
    """Many products of uneven N/W in one graph launch (pm_batch_*).

    products: list of (N, W, q) with q a Python int or a uint32 limb array."""

    def __init__(self, products):
        lib = load()
        Ns = np.array([p[0] for p in products], dtype=np.uint32)
        Ws = np.array([p[1] for p in products], dtype=np.uint32)
        qs = []
        for N, W, q in products:
            L = (W + 31) // 32
            if isinstance(q, (int, np.integer)):
                q = np.frombuffer(int(q).to_bytes(4 * L, "little"), dtype="<u4")
            qs.append(np.ascontiguousarray(q, dtype=np.uint32).reshape(L))
        qc = np.ascontiguousarray(np.concatenate(qs))
        h = _vp()
        _check_batch(lib.pm_batch_create(len(products), _ptr(Ns), _ptr(Ws), _ptr(qc), ctypes.byref(h)))
        self._h = h
        self.count = len(products)
        self.info = []
        for i in range(self.count):
            inf = PmInfo()
            _check_batch(lib.pm_batch_product_info(h, i, ctypes.byref(inf)))
            self.info.append(inf)

    def close(self):
        if getattr(self, "_h", None):
            load().pm_batch_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _ptrs(bufs):
        arr = (ctypes.c_void_p * len(bufs))(*[b.ptr.value if isinstance(b.ptr, ctypes.c_void_p) else b.ptr
                                              for b in bufs])
        return arr

    def run(self, da, db, dc, stream=None):
        _check_batch(load().pm_batch_run(self._h, self._ptrs(da), self._ptrs(db), self._ptrs(dc), stream))

    def sync(self):
        _check_batch(load().pm_batch_sync(self._h))

    def time(self, da, db, dc, iters: int) -> float:
        ms = ctypes.c_float()
        _check_batch(load().pm_batch_time(self._h, self._ptrs(da), self._ptrs(db), self._ptrs(dc), iters,
                                          ctypes.byref(ms)))
        return ms.value

    def configure(self, col_tile=32, ceil_split=0, blocks_per_cu=0, group_mb=192, branches=1):
        _check_batch(load().pm_batch_configure(self._h, col_tile, ceil_split, blocks_per_cu, group_mb, branches))

    def autotune(self, da, db, dc, iters: int = 3) -> dict:
        best = (ctypes.c_int * 6)()
        _check_batch(load().pm_batch_autotune(self._h, self._ptrs(da), self._ptrs(db), self._ptrs(dc), iters, best))
        return {"col_tile": best[0], "ceil_split": best[1], "blocks_per_cu": best[2], "us_per_launch": best[3],
                "group_mb": best[4], "branches": best[5]}


def _check_server(rc: int):
    if rc < 0:
        raise PolymultError(load().pm_server_last_error().decode())
    return rc


def _q_array(q, L: int) -> np.ndarray:
    if isinstance(q, (int, np.integer)):
        q = np.frombuffer(int(q).to_bytes(4 * L, "little"), dtype="<u4")
    return np.ascontiguousarray(q, dtype=np.uint32).reshape(L)


class Server:
    """Serving loop over a catalog of points (pm_server_*): submit host
    operands for point i, wait for the result; a scheduler thread batches
    whatever is waiting under the HBM and pinned-memory budgets.

    points: list of (N, W, q); config keys: device_budget, pinned_budget,
    max_batch, max_inflight, host_threads, col_tile, max_batch_bytes (0 or
    absent: default)."""

    def __init__(self, points, **config):
        lib = load()
        self.points = [(int(N), int(W)) for N, W, _ in points]
        Ns = np.array([p[0] for p in points], dtype=np.uint32)
        Ws = np.array([p[1] for p in points], dtype=np.uint32)
        qc = np.ascontiguousarray(np.concatenate([_q_array(q, (W + 31) // 32) for _, W, q in points]))
        cfg = ServerConfig(**{k: int(v) for k, v in config.items()})
        h = _vp()
        _check_server(lib.pm_server_create(len(points), _ptr(Ns), _ptr(Ws), _ptr(qc), ctypes.byref(cfg),
                                           ctypes.byref(h)))
        self._h = h
        self._live = {}  # ticket -> (a, b, c) kept alive until done

    def close(self):
        if getattr(self, "_h", None):
            load().pm_server_free(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def submit(self, point: int, a: np.ndarray, b: np.ndarray) -> int:
        N, W = self.points[point]
        L = (W + 31) // 32
        a = _u32(a).reshape(N, L)
        b = _u32(b).reshape(N, L)
        c = np.empty((N, L), dtype=np.uint32)
        t = ctypes.c_uint64()
        _check_server(load().pm_server_submit(self._h, point, _ptr(a), _ptr(b), _ptr(c), ctypes.byref(t)))
        self._live[t.value] = (a, b, c)
        return t.value

    def wait(self, ticket: int, timeout: float = 60.0):
        """(c, latency_ms) once the request is done; None on timeout."""
        ms = ctypes.c_double()
        try:
            rc = _check_server(load().pm_server_wait(self._h, ticket, timeout, ctypes.byref(ms)))
        except PolymultError:
            self._live.pop(ticket, None)
            raise
        if rc == 1:
            return None
        _, _, c = self._live.pop(ticket)
        return c, ms.value

    def stats(self, reset: bool = False) -> dict:
        st = ServerStats()
        _check_server(load().pm_server_stats_get(self._h, int(reset), ctypes.byref(st)))
        return _as_dict(st)

    def drive(self, rate: float, seconds: float, weights, seed: int = 1) -> dict:
        """Open-loop Poisson load from a native driver; see pm_server_drive."""
        w = (ctypes.c_double * len(self.points))(*[float(x) for x in weights])
        rep = ServeReport()
        _check_server(load().pm_server_drive(self._h, rate, seconds, w, seed, ctypes.byref(rep)))
        return _as_dict(rep)

    def saturate(self, depth: int, seconds: float, weights, seed: int = 1) -> dict:
        """Closed-loop load with `depth` requests in flight; see pm_server_saturate."""
        w = (ctypes.c_double * len(self.points))(*[float(x) for x in weights])
        rep = ServeReport()
        _check_server(load().pm_server_saturate(self._h, depth, seconds, w, seed, ctypes.byref(rep)))
        return _as_dict(rep)
