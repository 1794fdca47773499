"""The continuum-coupling entry point that allocates nothing
(dsmc_ns_apply_noalloc) against dsmc_ns_apply, through the C ABI, bit for
bit.  Two calls in a row, the second on a smaller engine, so that the fixed
per-device scratch still holds the first call's larger contents when the
second runs.  Device inputs are allocated with the HIP runtime the engine
library uses."""
import ctypes

import numpy as np
import pytest

import pydsmc as D
from test_gpu_parity import _case, _hip_runtime, _ns_fields, _same_particles

pytestmark = pytest.mark.gpu

H2D = 1


def _to_device(hip, a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0
    return p


def test_noalloc_coupling_matches_ns_apply():
    hip = _hip_runtime()
    cells = []
    for kind in ("capsule", "box_tets"):
        mesh, p, parts = _case(kind)
        cells.append(mesh.ncell)
        print(f"[ns_apply_noalloc] {kind} cells={mesh.ncell} particles={parts.n}", flush=True)
        a = D.Engine(mesh, p, seed=21)
        b = D.Engine(mesh, p, seed=21)
        for e in (a, b):
            e.upload(parts)
            e.step(1)
            e.step(2)
        counts = a.cell_scalars()["count"].astype(np.float64)
        ns, acc, stop = _ns_fields(mesh, p, counts)
        dns, dacc = _to_device(hip, ns), _to_device(hip, acc)
        assert a.ns_apply(3, ns, acc) == stop
        assert b.ns_apply_noalloc(3, dns.value, dacc.value) == stop
        _same_particles(a.download(), b.download())
        assert np.array_equal(a.cell_scalars()["count"], b.cell_scalars()["count"])
        for e in (a, b):
            e.step(3)
            e.step(4)
        _same_particles(a.download(), b.download())
        # every cell coupled (no mask), no NaN
        ns2, _, _ = _ns_fields(mesh, p, a.cell_scalars()["count"].astype(np.float64), seed=5)
        ns2[:, 0] = np.abs(np.nan_to_num(ns2[:, 0], nan=1.0))
        dns2 = _to_device(hip, ns2)
        assert a.ns_apply(5, ns2) == -1
        assert b.ns_apply_noalloc(5, dns2.value) == -1
        _same_particles(a.download(), b.download())
        for k in ("particles", "collisions", "crossings"):
            assert a.stats()[k] == b.stats()[k], k
        # host memory is refused before anything runs
        with pytest.raises(D.DsmcError) as ei:
            b.ns_apply_noalloc(6, ns2.ctypes.data)
        assert ei.value.code == D.ERR_ARG
        for q in (dns, dacc, dns2):
            assert hip.hipFree(q) == 0
        a.close()
        b.close()
    assert cells[1] < cells[0]
