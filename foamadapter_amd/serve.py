"""Serving: a resident request loop for implicit solves (libfvhip fv_server_*).

    srv = Server(device=0, hbm_budget=200 << 30, precision="mixed", batched=True)
    mid = srv.add_mesh(PolyMesh.hex(64, jitter=0.2, jitter_seed=1), renumber=True)
    rid = srv.submit_heat(mid, kappa=1e-2, dt=1e-3, steps=5, seed=7)
    T, st = srv.wait(rid)

Requests queue and are admitted while their device memory fits the budget;
with batched=True they run by continuous batching (systems join and leave the
batched CG between chunks), with batched=False one at a time through the plain
FP64 drop-in solve.  Latency is submission -> result on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import call, lib
from .mesh import PolyMesh


class FvServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("mode", C.c_int32), ("batched", C.c_int32),
                ("max_active", C.c_int32), ("tile_rows", C.c_int32), ("chunk", C.c_int32),
                ("use_graph", C.c_int32), ("solo_rows", C.c_int32)]


class FvRequest(C.Structure):
    _fields_ = [("mesh_id", C.c_int32), ("kind", C.c_int32), ("steps", C.c_int32),
                ("max_iter", C.c_int32), ("dt", C.c_double), ("kappa", C.c_double),
                ("rel_tol", C.c_double), ("h_wall_values", C.c_void_p), ("h_T0", C.c_void_p),
                ("seed", C.c_uint64)]


class FvRequestStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iters", C.c_int32), ("final_res", C.c_double),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class FvServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("rejected", C.c_int64), ("cancelled", C.c_int64), ("wall_s", C.c_double),
                ("solves_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("peak_bytes", C.c_size_t),
                ("peak_active", C.c_int32), ("resident_bytes", C.c_size_t),
                ("admit_ms", C.c_double), ("gpu_ms", C.c_double), ("advance_ms", C.c_double),
                ("chunks", C.c_int64), ("table_builds", C.c_int64), ("graph_builds", C.c_int64),
                ("chunk_event_ms", C.c_double), ("start_event_ms", C.c_double),
                ("solo_runs", C.c_int64), ("solo_ms", C.c_double)]


HEAT, POISSON = 0, 1
_SIGS = {
    "fv_server_create": [C.c_int, C.POINTER(FvServerOpts), C.POINTER(C.c_void_p)],
    "fv_server_add_mesh": [C.c_void_p, C.POINTER(_lib.FvPolyMeshDesc), C.c_int, C.POINTER(C.c_int)],
    "fv_server_mesh_cells": [C.c_void_p, C.c_int, C.POINTER(C.c_int32)],
    "fv_server_submit": [C.c_void_p, C.POINTER(FvRequest), C.POINTER(C.c_int64)],
    "fv_server_poll": [C.c_void_p, C.c_int64, C.POINTER(C.c_int)],
    "fv_server_wait": [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(FvRequestStats)],
    "fv_server_cancel": [C.c_void_p, C.c_int64],
    "fv_server_get_stats": [C.c_void_p, C.POINTER(FvServerStats)],
    "fv_server_request_bytes": [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_size_t)],
    "fv_server_destroy": [C.c_void_p],
}
for _n, _a in _SIGS.items():
    getattr(lib, _n).argtypes = _a
    getattr(lib, _n).restype = C.c_int
_lib._SIGS.update(_SIGS)


@dataclass
class RequestStats:
    status: int
    iters: int
    final_res: float
    queue_ms: float
    latency_ms: float


class Server:
    def __init__(self, device: int = 0, hbm_budget: int = 0, precision: str = "mixed",
                 batched: bool = True, max_active: int = 0, tile_rows: int = 0, chunk: int = 0,
                 use_graph: bool = True, solo_rows: int = 0):
        """solo_rows (batched): requests on meshes of at least this many cells run
        alone through the single-system solver between chunks (0: 1 << 20,
        negative: never)."""
        o = FvServerOpts(hbm_budget, _lib.FV_SOLVE_F32_IR if precision == "mixed" else _lib.FV_SOLVE_F64,
                         1 if batched else 0, max_active, tile_rows, chunk, 1 if use_graph else 0,
                         solo_rows)
        h = C.c_void_p()
        call("fv_server_create", device, C.byref(o), C.byref(h))
        self.handle = h
        self.n_patches = {}

    def add_mesh(self, poly: PolyMesh, renumber: bool = True) -> int:
        d, keep = poly.desc()
        mid = C.c_int()
        call("fv_server_add_mesh", self.handle, C.byref(d), 1 if renumber else 0, C.byref(mid))
        del keep
        self.n_patches[mid.value] = len(poly.patches)
        return mid.value

    def mesh_cells(self, mesh_id: int) -> int:
        n = C.c_int32()
        call("fv_server_mesh_cells", self.handle, mesh_id, C.byref(n))
        return n.value

    def _submit(self, req: FvRequest, keep=()) -> int:
        rid = C.c_int64()
        call("fv_server_submit", self.handle, C.byref(req), C.byref(rid))
        return rid.value

    def submit_heat(self, mesh_id: int, kappa: float, dt: float = 1e-3, steps: int = 5,
                    walls=None, T0=None, seed: int = 0, rel_tol: float = 1e-8,
                    max_iter: int = 2000) -> int:
        w = None if walls is None else np.ascontiguousarray(walls, np.float64)
        t0 = None if T0 is None else np.ascontiguousarray(T0, np.float64)
        req = FvRequest(mesh_id, HEAT, steps, max_iter, dt, kappa, rel_tol,
                        None if w is None else w.ctypes.data, None if t0 is None else t0.ctypes.data,
                        seed)
        return self._submit(req, (w, t0))

    def submit_poisson(self, mesh_id: int, seed: int = 0, rel_tol: float = 1e-8,
                       max_iter: int = 20000) -> int:
        req = FvRequest(mesh_id, POISSON, 1, max_iter, 1.0, 1.0, rel_tol, None, None, seed)
        return self._submit(req)

    def poll(self, rid: int) -> int:
        s = C.c_int()
        call("fv_server_poll", self.handle, rid, C.byref(s))
        return s.value

    def wait(self, rid: int, mesh_id: int | None = None, result: bool = True):
        out = None
        if result and mesh_id is not None:
            out = np.empty(self.mesh_cells(mesh_id))
        st = FvRequestStats()
        rc = lib.fv_server_wait(self.handle, rid, None if out is None else out.ctypes.data,
                                C.byref(st))
        stats = RequestStats(st.status, st.iters, st.final_res, st.queue_ms, st.latency_ms)
        return (out if rc == _lib.FV_OK else None), stats

    def cancel(self, rid: int):
        call("fv_server_cancel", self.handle, rid)

    def request_bytes(self, mesh_id: int, kind: int = HEAT) -> int:
        """Device bytes a request of `kind` (HEAT or POISSON) is admitted with."""
        n = C.c_size_t()
        call("fv_server_request_bytes", self.handle, mesh_id, kind, C.byref(n))
        return n.value

    def stats(self) -> dict:
        s = FvServerStats()
        call("fv_server_get_stats", self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in FvServerStats._fields_}

    def close(self):
        if self.handle:
            lib.fv_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
