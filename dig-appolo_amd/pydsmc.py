"""Python binding of the MI355X DSMC engine's C ABI (include/dsmc_engine.h).

Thin ctypes layer over ``libdsmc_engine.so`` (built in-tree next to this
file).  Importing it without the built library raises: there is no CPU
fallback for any engine call.
"""
from __future__ import annotations

import ctypes as C
import os
import types

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdsmc_engine.so")

DSMC_OK = 0
STAGE_EMIT, STAGE_MOVE, STAGE_SORT, STAGE_COLLIDE, STAGE_SAMPLE = 1, 2, 4, 8, 16
STAGE_ALL = 31
MACRO_LOCAL, MACRO_RECORD, MACRO_FINAL, MACRO_OUTPUT = 0, 1, 2, 3
WALL_DIFFUSE, WALL_SPECULAR = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
TAG_INTERFACE, TAG_INLET, TAG_OUTLET, TAG_WALL = 3, 4, 5, 6
ORDER_MORTON = 1
MACRO_NAMES = ["rho", "ux", "uy", "uz", "T", "sxx", "sxy", "syy", "syz", "szz", "szx",
               "qx", "qy", "qz", "Tr", "qrx", "qry", "qrz"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_bp = C.POINTER(C.c_uint8)


class MeshView(C.Structure):
    _fields_ = [("ncell", C.c_int32), ("nface", C.c_int32), ("nnode", C.c_int32),
                ("node_xyz", _dp), ("face_nodes", _ip), ("face_cells", _ip), ("face_type", _ip),
                ("face_tag", _ip), ("face_split", _bp), ("face_normal", _dp), ("face_area", _dp),
                ("cell_nface", _ip), ("cell_faces", _ip), ("cell_volume", _dp), ("cell_centroid", _dp)]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Ma", "Kn", "aoa_deg", "T_in", "T_wall", "T_ref", "L0", "gamma",
                                          "omega", "alpha", "d_ref", "mass", "P_relax", "dr")] + \
               [(n, C.c_int32) for n in ("Npinitial", "NSCHEME", "NSS", "Nevery", "Nrepeat", "wall_model",
                                         "precision", "surface_stats", "inflow")] + \
               [(n, C.c_double) for n in ("kB", "n_ref", "v_rms", "v_in", "mu_ref", "p_ref", "dt_ref",
                                          "Twall_ref", "eta", "p_mass_r", "Neff", "dtime", "volume_total",
                                          "min_edge")] + \
               [("u_inf", C.c_double * 3), ("T_inf", C.c_double), ("coll_prefactor", C.c_double)]

    def copy(self) -> "Params":
        q = Params()
        C.pointer(q)[0] = self
        return q


class ParticlesSoA(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", _dp), ("y", _dp), ("z", _dp), ("vx", _dp), ("vy", _dp), ("vz", _dp),
                ("ir", _dp), ("cell", _ip), ("id", _up)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("particles", "emitted", "removed", "dropped", "wall_hits",
                                         "crossings", "coll_pairs", "collisions", "particle_steps")] + \
               [("last_step", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class KernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("calls", C.c_int64 * 8)]


TUNE_KEEP = -2147483648


def _tuning_dict(t) -> dict:
    return {f: getattr(t, f) for f in ("wide_min", "wide_grid", "move_chunk", "graph", "sort_map", "coll_block")}


class Tuning(C.Structure):
    _fields_ = [("wide_min", C.c_int32), ("wide_grid", C.c_int32), ("move_chunk", C.c_int32),
                ("graph", C.c_int32), ("sort_map", C.c_int32), ("coll_block", C.c_int32),
                ("reserved", C.c_int32 * 2)]


class CaseInfo(C.Structure):
    _fields_ = [("ncase", C.c_int32), ("cell_begin", C.c_int32), ("cell_end", C.c_int32),
                ("face_begin", C.c_int32), ("face_end", C.c_int32), ("particles", C.c_int64),
                ("particle_steps", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("steps", C.c_int32), ("steps_done", C.c_int32), ("kase", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_finish", C.c_double),
                ("estimate", C.c_int64), ("footprint", C.c_int64), ("particles", C.c_int64),
                ("particle_steps", C.c_int64), ("error", C.c_char * 256)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_RELEASED = 0, 1, 2, 3, 4


class CrossEvent(C.Structure):
    _fields_ = [("id", C.c_uint32), ("step", C.c_int32), ("seq", C.c_int32), ("face", C.c_int32),
                ("tri", C.c_int32), ("kind", C.c_int32)]


KERNEL_GROUPS = ["emit", "move", "sort_count", "sort_scan", "sort_place", "sort_gather", "collide_sample", "other"]

EXPORTED = [
    "dsmc_mesh_build", "dsmc_mesh_box_tets", "dsmc_mesh_box_hexes", "dsmc_mesh_capsule", "dsmc_mesh_get_view",
    "dsmc_mesh_destroy", "dsmc_params_apollo", "dsmc_params_box", "dsmc_params_derive", "dsmc_fill_freestream",
    "dsmc_engine_create", "dsmc_engine_destroy", "dsmc_particles_upload", "dsmc_particles_download",
    "dsmc_particle_count", "dsmc_step", "dsmc_run", "dsmc_track", "dsmc_track_log_enable", "dsmc_track_log_get",
    "dsmc_macro_get", "dsmc_statistic_pre", "dsmc_accum_get", "dsmc_cell_scalars_get", "dsmc_surface_get",
    "dsmc_surface_fields", "dsmc_stats_get", "dsmc_profile_enable", "dsmc_profile_get", "dsmc_last_run_ms",
    "dsmc_batch_create", "dsmc_case_info", "dsmc_case_surface_fields", "dsmc_fill_device",
    "dsmc_run_async", "dsmc_run_query", "dsmc_run_on_stream", "dsmc_engine_estimate", "dsmc_server_create", "dsmc_server_destroy",
    "dsmc_server_submit", "dsmc_server_pump", "dsmc_server_set_max_running", "dsmc_server_set_batching",
    "dsmc_server_set_tuning", "dsmc_server_request",
    "dsmc_server_engine", "dsmc_server_release",
    "dsmc_engine_footprint", "dsmc_engine_set_tuning", "dsmc_engine_get_tuning", "dsmc_autotune",
    "dsmc_device_math_check", "dsmc_device_accept_check", "dsmc_ns_apply", "dsmc_ns_apply_noalloc", "dsmc_write_fluent_cells", "dsmc_write_fluent_walls",
    "dsmc_mesh_read_cas", "dsmc_mesh_write_cas", "dsmc_mesh_renumber",
    "dsmc_copy_probe", "dsmc_last_error",
]

_lib = None


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"DSMC engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.dsmc_last_error.restype = C.c_char_p
        for name in EXPORTED:
            if name not in ("dsmc_last_error", "dsmc_mesh_destroy", "dsmc_engine_destroy", "dsmc_server_destroy"):
                getattr(L, name).restype = C.c_int
        L.dsmc_mesh_destroy.restype = None
        L.dsmc_engine_destroy.restype = None
        L.dsmc_server_destroy.restype = None
        _lib = L
    return _lib


ERR_ARG, ERR_HIP, ERR_MESH, ERR_CAPACITY, ERR_STATE = -1, -2, -3, -4, -5


class DsmcError(RuntimeError):
    """A failed engine call; `code` is its DSMC_ERR_* status."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def check(rc: int, what: str = ""):
    if rc != DSMC_OK:
        msg = lib().dsmc_last_error().decode(errors="replace")
        raise DsmcError(f"{what}: error {rc}: {msg}", rc)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_math(u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros((len(u), 6))
    check(lib().dsmc_device_math_check(_ptr(u, C.c_double), C.c_int64(len(u)), _ptr(out, C.c_double)),
          "device_math_check")
    return out


def device_accept(x: np.ndarray, pref: float, eta: float, crm: float) -> np.ndarray:
    """The collision kernels' FP32 estimate of the NTC acceptance ratio (0
    where unused) and the deterministic ratio, for speeds x: (n, 2)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((len(x), 2))
    check(lib().dsmc_device_accept_check(_ptr(x, C.c_double), C.c_int64(len(x)), C.c_double(pref),
                                         C.c_double(eta), C.c_double(crm), _ptr(out, C.c_double)),
          "device_accept_check")
    return out


def copy_probe(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Achieved HBM GB/s (read + write) of a plain device copy of nbytes."""
    g = C.c_double()
    check(lib().dsmc_copy_probe(C.c_int64(nbytes), reps, C.byref(g)), "copy_probe")
    return g.value


class Mesh:
    """Host mesh owned by the library (dsmc_mesh)."""

    def __init__(self, handle):
        self.h = handle
        self.view = MeshView()
        check(lib().dsmc_mesh_get_view(self.h, C.byref(self.view)), "mesh view")

    def __del__(self):
        if getattr(self, "h", None):
            lib().dsmc_mesh_destroy(self.h)
            self.h = None

    @staticmethod
    def box_tets(n: int, wall_tag: int = TAG_WALL) -> "Mesh":
        h = C.c_void_p()
        check(lib().dsmc_mesh_box_tets(n, wall_tag, C.byref(h)), "box_tets")
        return Mesh(h)

    @staticmethod
    def box_hexes(n: int, wall_tag: int = TAG_WALL) -> "Mesh":
        h = C.c_void_p()
        check(lib().dsmc_mesh_box_hexes(n, wall_tag, C.byref(h)), "box_hexes")
        return Mesh(h)

    @staticmethod
    def capsule(a: int = 25, c: int = 62, layers: int = 50, diameter: float = 3.912, tets: bool = False,
                jitter: float = 0.0, seed: int = 0) -> "Mesh":
        h = C.c_void_p()
        check(lib().dsmc_mesh_capsule(a, c, layers, C.c_double(diameter), int(tets), C.c_double(jitter),
                                      C.c_uint64(seed), C.byref(h)), "capsule")
        return Mesh(h)

    @staticmethod
    def read_cas(path: str, scale: float = 1.0) -> "Mesh":
        """Fluent ASCII case mesh (readCAS)."""
        h = C.c_void_p()
        check(lib().dsmc_mesh_read_cas(path.encode(), C.c_double(scale), C.byref(h)), "mesh_read_cas")
        return Mesh(h)

    def renumber(self, order: int = 1) -> tuple["Mesh", np.ndarray]:
        """The mesh with its cells in Morton order of their centroids, and
        new_to_old: per-cell results r of the renumbered mesh map back to
        this mesh's order as r_mesh[new_to_old] = r."""
        perm = np.zeros(self.ncell, dtype=np.int32)
        h = C.c_void_p()
        check(lib().dsmc_mesh_renumber(self.h, int(order), C.byref(h), _ptr(perm, C.c_int32)), "mesh_renumber")
        return Mesh(h), perm

    def write_cas(self, path: str):
        check(lib().dsmc_mesh_write_cas(C.byref(self.view), path.encode()), "mesh_write_cas")

    @staticmethod
    def build(node_xyz, face_nodes, face_cells, face_type, face_tag, ncell: int) -> "Mesh":
        node_xyz = np.ascontiguousarray(node_xyz, dtype=np.float64)
        face_nodes = np.ascontiguousarray(face_nodes, dtype=np.int32)
        face_cells = np.ascontiguousarray(face_cells, dtype=np.int32)
        face_type = np.ascontiguousarray(face_type, dtype=np.int32)
        face_tag = np.ascontiguousarray(face_tag, dtype=np.int32)
        h = C.c_void_p()
        check(lib().dsmc_mesh_build(ncell, len(face_type), len(node_xyz), _ptr(node_xyz, C.c_double),
                                    _ptr(face_nodes, C.c_int32), _ptr(face_cells, C.c_int32),
                                    _ptr(face_type, C.c_int32), _ptr(face_tag, C.c_int32), C.byref(h)),
              "mesh_build")
        return Mesh(h)

    @property
    def ncell(self):
        return self.view.ncell

    @property
    def nface(self):
        return self.view.nface

    def arrays(self) -> dict:
        v = self.view

        def arr(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)
        return {
            "node_xyz": arr(v.node_xyz, v.nnode * 3, np.float64).reshape(-1, 3),
            "face_nodes": arr(v.face_nodes, v.nface * 4, np.int32).reshape(-1, 4),
            "face_cells": arr(v.face_cells, v.nface * 2, np.int32).reshape(-1, 2),
            "face_type": arr(v.face_type, v.nface, np.int32),
            "face_tag": arr(v.face_tag, v.nface, np.int32),
            "face_split": arr(v.face_split, v.nface, np.uint8),
            "face_normal": arr(v.face_normal, v.nface * 3, np.float64).reshape(-1, 3),
            "face_area": arr(v.face_area, v.nface, np.float64),
            "cell_nface": arr(v.cell_nface, v.ncell, np.int32),
            "cell_faces": arr(v.cell_faces, v.ncell * 6, np.int32).reshape(-1, 6),
            "cell_volume": arr(v.cell_volume, v.ncell, np.float64),
            "cell_centroid": arr(v.cell_centroid, v.ncell * 3, np.float64).reshape(-1, 3),
        }


def params_apollo(monatomic: bool = True) -> Params:
    p = Params()
    check(lib().dsmc_params_apollo(C.byref(p), int(monatomic)), "params_apollo")
    return p


def params_box(Kn: float = 0.05) -> Params:
    p = Params()
    check(lib().dsmc_params_box(C.byref(p), C.c_double(Kn)), "params_box")
    return p


def derive(mesh: Mesh, p: Params) -> Params:
    check(lib().dsmc_params_derive(C.byref(mesh.view), C.byref(p)), "params_derive")
    return p


class Particles:
    """Numpy-owned SoA particle set."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ir")

    def __init__(self, n: int):
        self.n = int(n)
        for f in self.FIELDS:
            setattr(self, f, np.zeros(self.n, dtype=np.float64))
        self.cell = np.zeros(self.n, dtype=np.int32)
        self.id = np.zeros(self.n, dtype=np.uint32)

    def soa(self) -> ParticlesSoA:
        s = ParticlesSoA()
        s.n = self.n
        for f in self.FIELDS:
            setattr(s, f, _ptr(getattr(self, f), C.c_double))
        s.cell = _ptr(self.cell, C.c_int32)
        s.id = _ptr(self.id, C.c_uint32)
        return s

    def take(self, idx) -> "Particles":
        q = Particles(len(idx))
        for f in self.FIELDS + ("cell", "id"):
            setattr(q, f, np.ascontiguousarray(getattr(self, f)[idx]))
        return q

    def copy(self) -> "Particles":
        return self.take(np.arange(self.n))


def fill_freestream(mesh: Mesh, p: Params, ppc: float, temperature: float = 0.0, seed: int = 1) -> Particles:
    cnt = C.c_int64()
    empty = ParticlesSoA()
    check(lib().dsmc_fill_freestream(C.byref(mesh.view), C.byref(p), C.c_double(ppc), C.c_double(temperature),
                                     C.c_uint64(seed), C.byref(empty), C.byref(cnt)), "fill count")
    parts = Particles(cnt.value)
    s = parts.soa()
    check(lib().dsmc_fill_freestream(C.byref(mesh.view), C.byref(p), C.c_double(ppc), C.c_double(temperature),
                                     C.c_uint64(seed), C.byref(s), C.byref(cnt)), "fill")
    return parts


def concat_particles(parts: list) -> Particles:
    out = Particles(sum(q.n for q in parts))
    for f in Particles.FIELDS + ("cell", "id"):
        setattr(out, f, np.ascontiguousarray(np.concatenate([getattr(q, f) for q in parts])))
    return out


class Engine:
    """One DSMC case resident on the current HIP device."""

    def __init__(self, mesh: Mesh, p: Params, seed: int = 1, capacity: int = 0):
        self.mesh = mesh
        self.params = p
        self.h = C.c_void_p()
        check(lib().dsmc_engine_create(C.byref(mesh.view), C.byref(p), C.c_uint64(seed), C.c_int64(capacity),
                                       C.byref(self.h)), "engine_create")

    @classmethod
    def batch(cls, meshes: list, params: list, seeds: list, capacity: int = 0) -> "Engine":
        """Several cases in one engine (see dsmc_batch_create)."""
        self = cls.__new__(cls)
        self.meshes = list(meshes)
        self.case_params = list(params)
        self.h = C.c_void_p()
        k = len(meshes)
        views = (MeshView * k)(*[m.view for m in meshes])
        ps = (Params * k)(*params)
        sd = (C.c_uint64 * k)(*[int(x) for x in seeds])
        check(lib().dsmc_batch_create(k, views, ps, sd, C.c_int64(capacity), C.byref(self.h)), "batch_create")
        self.params = params[0]
        last = self.case_info(k - 1)
        # batch-wide sizes for the array getters: every case's cell range is
        # padded to a multiple of 64, so the batch ends at the last case's end
        # rounded up to 64
        self.mesh = types.SimpleNamespace(ncell=(last["cell_end"] + 63) // 64 * 64, nface=last["face_end"])
        return self

    def ncase(self) -> int:
        i = CaseInfo()
        check(lib().dsmc_case_info(self.h, 0, C.byref(i)), "case_info")
        return i.ncase

    def case_info(self, k: int) -> dict:
        i = CaseInfo()
        check(lib().dsmc_case_info(self.h, int(k), C.byref(i)), "case_info")
        return {f: getattr(i, f) for f, _ in CaseInfo._fields_}

    def case_particles(self, parts: Particles, k: int) -> Particles:
        """The particles of case k from a batch download, cells renumbered to the case's own mesh."""
        i = self.case_info(k)
        sel = np.flatnonzero((parts.cell >= i["cell_begin"]) & (parts.cell < i["cell_end"]))
        q = parts.take(sel)
        q.cell = (q.cell - i["cell_begin"]).astype(np.int32)
        return q

    def batch_particles(self, per_case: list) -> Particles:
        """Concatenates per-case particle sets, cells renumbered to the batch's."""
        out = []
        for k, q in enumerate(per_case):
            r = q.copy()
            r.cell = (r.cell + self.case_info(k)["cell_begin"]).astype(np.int32)
            out.append(r)
        return concat_particles(out)

    def fill_device(self, ppc, seeds, temperature=None) -> int:
        """Device-side initial fill (per case: ppc, seed, temperature or None)."""
        ppc = [float(x) for x in (ppc if hasattr(ppc, "__len__") else [ppc])]
        seeds = [int(x) for x in (seeds if hasattr(seeds, "__len__") else [seeds])]
        k = len(ppc)
        pa = (C.c_double * k)(*ppc)
        sd = (C.c_uint64 * k)(*seeds)
        ta = None
        if temperature is not None:
            t = [float(x) for x in (temperature if hasattr(temperature, "__len__") else [temperature])]
            ta = (C.c_double * k)(*t)
        n = C.c_int64()
        check(lib().dsmc_fill_device(self.h, pa, ta, sd, C.byref(n)), "fill_device")
        return n.value

    def case_surface_fields(self, k: int, istep: int, diameter: float):
        f = np.zeros((self.mesh.nface, 12))
        cd = C.c_double()
        q = C.c_double()
        check(lib().dsmc_case_surface_fields(self.h, int(k), istep, C.c_double(diameter), _ptr(f, C.c_double),
                                             C.byref(cd), C.byref(q)), "case_surface_fields")
        return f, cd.value, q.value

    def close(self):
        if getattr(self, "h", None):
            lib().dsmc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, parts: Particles):
        s = parts.soa()
        check(lib().dsmc_particles_upload(self.h, C.byref(s)), "upload")

    def count(self) -> int:
        n = C.c_int64()
        check(lib().dsmc_particle_count(self.h, C.byref(n)), "count")
        return n.value

    def download(self) -> Particles:
        n = self.count()
        parts = Particles(n)
        s = parts.soa()
        cnt = C.c_int64()
        check(lib().dsmc_particles_download(self.h, C.byref(s), C.byref(cnt)), "download")
        return parts

    def step(self, istep: int, mask: int = STAGE_ALL):
        check(lib().dsmc_step(self.h, istep, mask), f"step {istep}")

    def run(self, first: int, last: int) -> float:
        check(lib().dsmc_run(self.h, first, last), f"run {first}..{last}")
        ms = C.c_double()
        check(lib().dsmc_last_run_ms(self.h, C.byref(ms)), "last_run_ms")
        return ms.value

    def run_async(self, first: int, last: int):
        """Enqueue steps first..last on the engine's stream and return."""
        check(lib().dsmc_run_async(self.h, first, last), f"run_async {first}..{last}")

    def run_on_stream(self, first: int, last: int, stream: int):
        """Enqueue steps first..last ordered on a caller's HIP stream (the
        hipStream_t handle as an integer)."""
        check(lib().dsmc_run_on_stream(self.h, first, last, C.c_void_p(stream)), f"run_on_stream {first}..{last}")

    def run_query(self) -> bool:
        """Whether the last asynchronous run has completed (raises its errors)."""
        done = C.c_int32()
        check(lib().dsmc_run_query(self.h, C.byref(done)), "run_query")
        return bool(done.value)

    def track(self, istep: int):
        check(lib().dsmc_track(self.h, istep), "track")

    def log_enable(self, max_events: int):
        check(lib().dsmc_track_log_enable(self.h, C.c_int64(max_events)), "log_enable")

    def log_get(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().dsmc_track_log_get(self.h, None, C.c_int64(0), C.byref(n)), "log_count")
        out = (CrossEvent * max(1, n.value))()
        check(lib().dsmc_track_log_get(self.h, out, C.c_int64(n.value), C.byref(n)), "log_get")
        a = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_int32)), shape=(max(1, n.value) * 6,))
        return a.reshape(-1, 6)[: n.value].copy()

    def macro(self, which: int = MACRO_LOCAL) -> np.ndarray:
        out = np.zeros((self.mesh.ncell, 18))
        check(lib().dsmc_macro_get(self.h, which, _ptr(out, C.c_double)), "macro_get")
        return out

    def statistic_pre(self):
        check(lib().dsmc_statistic_pre(self.h), "statistic_pre")

    def accum(self, family: int) -> np.ndarray:
        out = np.zeros((self.mesh.ncell, 18))
        check(lib().dsmc_accum_get(self.h, family, _ptr(out, C.c_double)), "accum_get")
        return out

    def cell_scalars(self) -> dict:
        nc = self.mesh.ncell
        d = {"count": np.zeros(nc, np.int32), "crmax": np.zeros(nc), "rem_coll": np.zeros(nc),
             "rem_emit": np.zeros(nc), "sparse_state": np.zeros(nc, np.uint8), "sparse_steps": np.zeros(nc, np.int32)}
        check(lib().dsmc_cell_scalars_get(self.h, _ptr(d["count"], C.c_int32), _ptr(d["crmax"], C.c_double),
                                          _ptr(d["rem_coll"], C.c_double), _ptr(d["rem_emit"], C.c_double),
                                          _ptr(d["sparse_state"], C.c_uint8), _ptr(d["sparse_steps"], C.c_int32)),
              "cell_scalars")
        return d

    def surface(self) -> np.ndarray:
        out = np.zeros((self.mesh.nface, 6))
        check(lib().dsmc_surface_get(self.h, _ptr(out, C.c_double)), "surface_get")
        return out

    def surface_fields(self, istep: int, diameter: float):
        f = np.zeros((self.mesh.nface, 12))
        cd = C.c_double()
        q = C.c_double()
        check(lib().dsmc_surface_fields(self.h, istep, C.c_double(diameter), _ptr(f, C.c_double), C.byref(cd),
                                        C.byref(q)), "surface_fields")
        return f, cd.value, q.value

    def stats(self) -> dict:
        s = Stats()
        check(lib().dsmc_stats_get(self.h, C.byref(s)), "stats")
        return s.as_dict()

    def profile(self, on: bool = True):
        check(lib().dsmc_profile_enable(self.h, int(on)), "profile_enable")

    def kernel_times(self) -> dict:
        t = KernelTimes()
        check(lib().dsmc_profile_get(self.h, C.byref(t)), "profile_get")
        return {g: (t.ms[i], t.calls[i]) for i, g in enumerate(KERNEL_GROUPS)}

    @staticmethod
    def estimate(mesh: Mesh, p: Params, capacity: int) -> int:
        """Upper bound of the device bytes an engine for this case allocates."""
        b = C.c_int64()
        check(lib().dsmc_engine_estimate(C.byref(mesh.view), C.byref(p), C.c_int64(capacity), C.byref(b)),
              "estimate")
        return b.value

    def footprint(self) -> int:
        b = C.c_int64()
        check(lib().dsmc_engine_footprint(self.h, C.byref(b)), "footprint")
        return b.value

    def set_tuning(self, wide_min: int | None = None, wide_grid: int | None = None,
                   move_chunk: int | None = None, graph: int | None = None, sort_map: int | None = None,
                   coll_block: int | None = None):
        """Launch tuning; results are bit-identical for every setting."""
        t = Tuning()
        for f, v in (("wide_min", wide_min), ("wide_grid", wide_grid), ("move_chunk", move_chunk),
                     ("graph", graph), ("sort_map", sort_map), ("coll_block", coll_block)):
            setattr(t, f, TUNE_KEEP if v is None else int(v))
        check(lib().dsmc_engine_set_tuning(self.h, C.byref(t)), "set_tuning")

    def ns_apply(self, istep: int, ns: np.ndarray, accepted: np.ndarray | None = None) -> int:
        """Continuum coupling (molecular_velocity_change): ns is ncell x 6
        (rho/n_ref, u (3), T, Tr); returns the first coupled cell with a NaN
        NS state (left unchanged with every later cell), or -1."""
        a = np.ascontiguousarray(ns, dtype=np.float64).reshape(-1, 6)
        acc = None if accepted is None else np.ascontiguousarray(accepted, dtype=np.uint8)
        nan_cell = C.c_int32()
        check(lib().dsmc_ns_apply(self.h, int(istep), a.ctypes.data_as(C.POINTER(C.c_double)),
                                  None if acc is None else acc.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  C.byref(nan_cell)), "ns_apply")
        return nan_cell.value

    def ns_apply_noalloc(self, istep: int, ns_device: int, accepted_device: int | None = None) -> int:
        """ns_apply on NS fields already in device memory (addresses of
        ncell x 6 doubles and ncell bytes), with no device allocation
        (dsmc_ns_apply_noalloc)."""
        nan_cell = C.c_int32()
        check(lib().dsmc_ns_apply_noalloc(self.h, int(istep), C.c_void_p(ns_device),
                                          C.c_void_p(accepted_device) if accepted_device else None,
                                          C.byref(nan_cell)), "ns_apply_noalloc")
        return nan_cell.value

    def autotune(self, first_step: int, steps_per_trial: int = 2) -> tuple[dict, int]:
        """Tunes the launch parameters on live steps first_step, ...; returns
        (tuning, steps used).  Results do not depend on the tuning."""
        t = Tuning()
        used = C.c_int32()
        check(lib().dsmc_autotune(self.h, int(first_step), int(steps_per_trial), C.byref(t), C.byref(used)),
              "autotune")
        return _tuning_dict(t), used.value

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().dsmc_engine_get_tuning(self.h, C.byref(t)), "get_tuning")
        return _tuning_dict(t)


class _Borrowed(Engine):
    """An engine owned by a server (not destroyed by this object)."""

    def close(self):
        self.h = None

    def __del__(self):
        pass


class Server:
    """Serving loop with continuous batching (see dsmc_server_create)."""

    def __init__(self, budget_bytes: int = 0, chunk_steps: int = 50, auto_release: bool = True,
                 max_running: int = 0, batching: int = 1, tuning: dict | None = None):
        self.h = C.c_void_p()
        check(lib().dsmc_server_create(C.c_int64(budget_bytes), int(chunk_steps), int(auto_release),
                                       C.byref(self.h)), "server_create")
        self.meshes = {}
        if max_running:
            check(lib().dsmc_server_set_max_running(self.h, int(max_running)), "server_set_max_running")
        if batching != 1:
            check(lib().dsmc_server_set_batching(self.h, int(batching)), "server_set_batching")
        if tuning is not None:
            self.set_tuning(tuning)

    def set_tuning(self, tuning: dict | None):
        """Launch settings (Engine.set_tuning's keys; missing keys keep the
        engine default) for every engine built from now on; None restores
        the defaults."""
        if tuning is None:
            check(lib().dsmc_server_set_tuning(self.h, None), "server_set_tuning")
            return
        t = Tuning()
        for f in ("wide_min", "wide_grid", "move_chunk", "graph", "sort_map", "coll_block"):
            v = tuning.get(f)
            setattr(t, f, TUNE_KEEP if v is None else int(v))
        check(lib().dsmc_server_set_tuning(self.h, C.byref(t)), "server_set_tuning")

    def close(self):
        if getattr(self, "h", None):
            lib().dsmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, mesh: Mesh, p: Params, seed: int, ppc: float, steps: int, capacity: int = 0) -> int:
        """Queues a request; the server takes over the mesh."""
        rid = C.c_int64()
        check(lib().dsmc_server_submit(self.h, mesh.h, C.byref(p), C.c_uint64(seed), C.c_double(ppc), int(steps),
                                       C.c_int64(capacity), C.byref(rid)), "server_submit")
        self.meshes[rid.value] = (mesh.ncell, mesh.nface)
        mesh.h = None  # owned by the server now (its arrays go with it)
        return rid.value

    def pump(self) -> tuple[int, int, int]:
        r, q, f = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().dsmc_server_pump(self.h, C.byref(r), C.byref(q), C.byref(f)), "server_pump")
        return r.value, q.value, f.value

    def request(self, rid: int) -> dict:
        i = RequestInfo()
        check(lib().dsmc_server_request(self.h, int(rid), C.byref(i)), "server_request")
        d = {f: getattr(i, f) for f, _ in RequestInfo._fields_ if f != "error"}
        d["error"] = i.error.decode(errors="replace")
        return d

    def engine(self, rid: int, params: Params) -> Engine:
        """The engine of a finished, unreleased request (borrowed).  For a
        batched request it is the batch engine: the request's results are its
        case (request(rid)["kase"]) of it."""
        h = C.c_void_p()
        check(lib().dsmc_server_engine(self.h, int(rid), C.byref(h)), "server_engine")
        e = _Borrowed.__new__(_Borrowed)
        e.h = h
        e.params = params
        nc, nf = self.meshes[rid]
        e.mesh = types.SimpleNamespace(ncell=nc, nface=nf)
        n = e.ncase()
        if n > 1:
            # batch-wide sizes: the last case's range rounded up to 64 cells
            last = e.case_info(n - 1)
            e.mesh = types.SimpleNamespace(ncell=(last["cell_end"] + 63) // 64 * 64, nface=last["face_end"])
        return e

    def release(self, rid: int):
        check(lib().dsmc_server_release(self.h, int(rid)), "server_release")


def write_fluent_cells(path: str, p: Params, w: np.ndarray, zone_id: int = 2, first_id: int = 1):
    """Fluent cell fields file (out2Fluent_dsmc) of an ncell x 18 macro array."""
    a = np.ascontiguousarray(w, dtype=np.float64).reshape(-1, 18)
    check(lib().dsmc_write_fluent_cells(path.encode(), C.byref(p), a.shape[0],
                                        a.ctypes.data_as(C.POINTER(C.c_double)), int(zone_id), int(first_id)),
          "write_fluent_cells")


def write_fluent_walls(path: str, mesh: Mesh, fields: np.ndarray, zone_id: int = 7, first_id: int = 1):
    """Fluent wall fields file (outBoundaryStressHeat) of dsmc_surface_fields' nface x 12 output."""
    a = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1, 12)
    check(lib().dsmc_write_fluent_walls(path.encode(), C.byref(mesh.view), a.ctypes.data_as(C.POINTER(C.c_double)),
                                        int(zone_id), int(first_id)), "write_fluent_walls")
