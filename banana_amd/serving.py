"""Python view of the batching and serving entry points of include/bnn_flip.h.

`Batch` steps several FLIPSolver scenes concurrently (a worker per hardware queue, each scene on its
own stream); `Server` is the request loop (HBM-footprint admission, ready requests start as soon as
a worker is free). Both call libbnn_flip.so; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BnnError, Params, RequestResult, SceneSpec, ServerConfig, ServerStats
from .solver import make_params, make_spec


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def footprint(params, n_particles):
    """Device bytes a scene with these parameters and particle count keeps resident."""
    L = _lib.load()
    p = params if isinstance(params, Params) else make_params(params)
    b = C.c_uint64()
    st = L.bnn_flip_footprint(C.byref(p), int(n_particles), C.byref(b))
    if st != _lib.BNN_OK:
        raise BnnError(st, (L.bnn_flip_last_error(None) or b"").decode())
    return int(b.value)


class Batch:
    """Steps any set of FLIPSolver handles together: one fused launch per stage over all scenes
    (bnn_flip_batch_*); each scene's results are those of stepping it alone."""

    def __init__(self, timing=-1):
        self.L = _lib.load()
        h = C.c_void_p()
        st = self.L.bnn_flip_batch_create(C.byref(h))
        if st != _lib.BNN_OK:
            raise BnnError(st, (self.L.bnn_flip_batch_last_error(None) or b"").decode())
        self.h = h
        if timing >= 0:
            self.L.bnn_flip_batch_set_timing(self.h, int(timing))

    def close(self):
        if getattr(self, "h", None):
            self.L.bnn_flip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _call(self, fn, solvers, dts=None):
        n = len(solvers)
        hs = (C.c_void_p * max(n, 1))(*[s.h for s in solvers])
        status = np.zeros(max(n, 1), np.int32)
        args = [self.h, C.cast(hs, C.POINTER(C.c_void_p))]
        if dts is not None:
            d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, np.float32), (n,)), np.float32)
            args.append(d.ctypes.data_as(C.POINTER(C.c_float)))
        args += [n, status.ctypes.data_as(C.POINTER(C.c_int32))]
        st = fn(*args)
        if st not in (_lib.BNN_OK, 2):  # per-scene NOT_CONVERGED is reported in the status array
            raise BnnError(st, (self.L.bnn_flip_batch_last_error(self.h) or b"").decode())
        return status[:n]

    def step(self, solvers, dts):
        return self._call(self.L.bnn_flip_batch_step, solvers, dts)

    def advance_frame(self, solvers):
        return self._call(self.L.bnn_flip_batch_advance_frame, solvers)

    def pressure_projection(self, solvers, dts):
        return self._call(self.L.bnn_flip_batch_pressure_projection, solvers, dts)

    def set_shape(self, p2g_segment=8, pcg_chunk_first=8, pcg_chunk_max=8, items_per_group=16, min_groups=0, graphs=True):
        """Pins the launch shape of later calls (none of its settings changes a result)."""
        sh = _lib.BatchShape(int(p2g_segment), int(pcg_chunk_first), int(pcg_chunk_max), int(items_per_group), int(min_groups),
                             1 if graphs else 0)
        st = self.L.bnn_flip_batch_set_shape(self.h, C.byref(sh))
        if st != _lib.BNN_OK:
            raise BnnError(st, "invalid batch shape")

    def autotune(self):
        """Returns later calls to the autotuner."""
        self.L.bnn_flip_batch_set_shape(self.h, None)

    def info(self):
        """Launch counters since creation and the shape of the last call."""
        i = _lib.BatchInfo()
        st = self.L.bnn_flip_batch_get_info(self.h, C.byref(i))
        if st != _lib.BNN_OK:
            raise BnnError(st, "bnn_flip_batch_get_info failed")
        out = {k: getattr(i, k) for k, _ in _lib.BatchInfo._fields_ if k != "shape"}
        out["shape"] = {k: getattr(i.shape, k) for k, _ in _lib.BatchShape._fields_}
        return out


class Server:
    """Request loop over resident scenes (bnn_server_*)."""

    def __init__(self, hbm_budget_bytes=0, max_batch=0, workers=0, batching=True, barrier_rounds=0, max_batch_mcells=0):
        self.L = _lib.load()
        cfg = ServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.max_batch = int(max_batch)
        cfg.workers = int(workers)
        cfg.batching = 1 if batching else 0
        cfg.barrier_rounds = int(barrier_rounds)
        cfg.max_batch_mcells = int(max_batch_mcells)
        h = C.c_void_p()
        st = self.L.bnn_server_create(C.byref(cfg), C.byref(h))
        if st != _lib.BNN_OK:
            raise BnnError(st, (self.L.bnn_server_last_error(None) or b"").decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.bnn_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, st):
        if st != _lib.BNN_OK:
            raise BnnError(st, (self.L.bnn_server_last_error(self.h) or b"").decode())

    def now(self):
        return float(self.L.bnn_server_now(self.h))

    def add_scene(self, params, spec, seed):
        p = params if isinstance(params, Params) else make_params(params)
        s = spec if isinstance(spec, SceneSpec) else make_spec(spec)
        sid = C.c_uint64()
        self._c(self.L.bnn_server_add_scene(self.h, C.byref(p), C.byref(s), int(seed), C.byref(sid)))
        return int(sid.value)

    def release_scene(self, sid):
        self._c(self.L.bnn_server_release_scene(self.h, int(sid)))

    def submit(self, sid, kind, dt):
        rid = C.c_uint64()
        self._c(self.L.bnn_server_submit(self.h, int(sid), int(kind), float(dt), C.byref(rid)))
        return int(rid.value)

    def wait(self, rid, timeout=-1.0):
        r = RequestResult()
        self._c(self.L.bnn_server_wait(self.h, int(rid), float(timeout), C.byref(r)))
        return {k: getattr(r, k) for k, _ in RequestResult._fields_ if k != "pad"}

    def drain(self, timeout=-1.0):
        self._c(self.L.bnn_server_drain(self.h, float(timeout)))

    def particles(self, sid):
        n = C.c_uint64()
        self._c(self.L.bnn_server_get_particles(self.h, int(sid), None, None, C.byref(n)))
        pos = np.zeros((n.value, 3), np.float32)
        vel = np.zeros((n.value, 3), np.float32)
        self._c(self.L.bnn_server_get_particles(self.h, int(sid), _ptr(pos), _ptr(vel), C.byref(n)))
        return pos, vel

    def stats(self):
        s = ServerStats()
        self._c(self.L.bnn_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        self._c(self.L.bnn_server_reset_stats(self.h))
