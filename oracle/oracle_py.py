"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg to
check and time against; the engine never calls it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "dig-appolo_amd"))
import pydsmc as D  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oracle_count.restype = C.c_int64
        L.oracle_log_get.restype = C.c_int64
        L.oracle_destroy.restype = None
        _lib = L
    return _lib


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


PLANT_VHS_MASS, PLANT_VHS_HS_POWER = 1, 2
PLANT_LB_ONE_PARTNER, PLANT_LB_EXPONENT = 3, 4
PLANT_INFLOW_NO_FLUX_WEIGHT, PLANT_INFLOW_TANGENT_SCALE = 5, 6


@contextlib.contextmanager
def planted(which: int):
    """Oracles created inside the block carry one known mistake (see
    oracle_plant in dsmc_oracle.h), so a test can show its check fails."""
    if lib().oracle_plant(int(which)) != 0:
        raise ValueError(f"unknown plant {which}")
    try:
        yield
    finally:
        lib().oracle_plant(0)


def split_tags(mesh: "D.Mesh") -> np.ndarray:
    out = np.zeros(mesh.nface, dtype=np.uint8)
    lib().oracle_split_tags(C.byref(mesh.view), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def host_math(u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros((len(u), 6))
    lib().oracle_math(u.ctypes.data_as(C.POINTER(C.c_double)), C.c_int64(len(u)),
                      out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
    lib().oracle_philox(P(c), P(k), P(out))
    return out


def stream_uniforms(key, a, b, purpose, sub, n):
    out = np.zeros(n)
    lib().oracle_stream_uniforms(C.c_uint64(key), C.c_uint32(a), C.c_uint32(b), C.c_uint32(purpose),
                                 C.c_uint32(sub), C.c_int64(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def tri_hit(a, e1, e2, x, v):
    arrs = [np.ascontiguousarray(q, dtype=np.float64) for q in (a, e1, e2, x, v)]
    t = C.c_double()
    hit = lib().oracle_tri_hit(*[q.ctypes.data_as(C.POINTER(C.c_double)) for q in arrs], C.byref(t))
    return bool(hit), t.value


class Oracle:
    def __init__(self, mesh: "D.Mesh", p: "D.Params", seed: int = 1):
        self.mesh = mesh
        self.h = C.c_void_p()
        rc = lib().oracle_create(C.byref(mesh.view), C.byref(p), C.c_uint64(seed), C.byref(self.h))
        if rc != 0:
            raise RuntimeError("oracle_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_particles(self, parts: "D.Particles"):
        s = parts.soa()
        if lib().oracle_set_particles(self.h, C.byref(s)) != 0:
            raise RuntimeError("oracle_set_particles failed")

    def count(self) -> int:
        return lib().oracle_count(self.h)

    def particles(self) -> "D.Particles":
        parts = D.Particles(self.count())
        s = parts.soa()
        if lib().oracle_get_particles(self.h, C.byref(s)) != 0:
            raise RuntimeError("oracle_get_particles failed")
        return parts

    def step(self, istep: int, mask: int = D.STAGE_ALL):
        if lib().oracle_step(self.h, istep, mask) != 0:
            raise RuntimeError("oracle_step failed")

    def log_enable(self, on: bool = True):
        lib().oracle_log_enable(self.h, int(on))

    def log_get(self) -> np.ndarray:
        n = lib().oracle_log_get(self.h, None, 0)
        out = (D.CrossEvent * max(1, n))()
        lib().oracle_log_get(self.h, out, n)
        a = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_int32)), shape=(max(1, n) * 6,))
        return a.reshape(-1, 6)[:n].copy()

    def macro(self, which: int = D.MACRO_LOCAL) -> np.ndarray:
        out = np.zeros((self.mesh.ncell, 18))
        lib().oracle_macro_get(self.h, which, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def ns_apply(self, istep: int, ns: np.ndarray, accepted: np.ndarray | None = None) -> int:
        a = np.ascontiguousarray(ns, dtype=np.float64).reshape(-1, 6)
        acc = None if accepted is None else np.ascontiguousarray(accepted, dtype=np.uint8)
        nan_cell = C.c_int32()
        rc = lib().oracle_ns_apply(self.h, int(istep), a.ctypes.data_as(C.POINTER(C.c_double)),
                                   None if acc is None else acc.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   C.byref(nan_cell))
        if rc:
            raise RuntimeError(f"oracle_ns_apply failed ({rc})")
        return nan_cell.value

    def statistic_pre(self):
        lib().oracle_statistic_pre(self.h)

    def accum(self, family: int) -> np.ndarray:
        out = np.zeros((self.mesh.ncell, 18))
        lib().oracle_accum_get(self.h, family, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def cell_scalars(self) -> dict:
        nc = self.mesh.ncell
        d = {"count": np.zeros(nc, np.int32), "crmax": np.zeros(nc), "rem_coll": np.zeros(nc),
             "rem_emit": np.zeros(nc), "sparse_state": np.zeros(nc, np.uint8), "sparse_steps": np.zeros(nc, np.int32)}
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        lib().oracle_cell_scalars(self.h, P(d["count"], C.c_int32), P(d["crmax"], C.c_double),
                                  P(d["rem_coll"], C.c_double), P(d["rem_emit"], C.c_double),
                                  P(d["sparse_state"], C.c_uint8), P(d["sparse_steps"], C.c_int32))
        return d

    def surface(self) -> np.ndarray:
        out = np.zeros((self.mesh.nface, 6))
        lib().oracle_surface_get(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def surface_fields(self, istep: int, tally: np.ndarray | None = None) -> np.ndarray:
        out = np.zeros((self.mesh.nface, 12))
        t = None if tally is None else np.ascontiguousarray(tally, dtype=np.float64).reshape(-1, 6)
        if lib().oracle_surface_fields(self.h, int(istep), None if t is None else t.ctypes.data_as(C.POINTER(C.c_double)),
                                       out.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise RuntimeError("oracle_surface_fields failed")
        return out

    def stats(self) -> dict:
        s = D.Stats()
        lib().oracle_stats(self.h, C.byref(s))
        return s.as_dict()

    def times(self) -> np.ndarray:
        t = np.zeros(4)
        lib().oracle_times(self.h, t.ctypes.data_as(C.POINTER(C.c_double)))
        return t
