"""GPU parity through the mesh ingest and the output files.

* A case mesh written in the Fluent ASCII format and read back
  (dsmc_mesh_write_cas / dsmc_mesh_read_cas; readCAS,
  src/meshImport.cpp:177-576) runs bit-identically to the original mesh on
  the engine, and bit-identically to the oracle on the read mesh.
* The cell fields file (out2Fluent_dsmc, src/meshImport.cpp:1107-1182, called
  from out2dat, src/ProcessDSMC.cpp:3421-3424) and the wall fields file
  (outBoundaryStressHeat, src/ProcessDSMC.cpp:3494-3652) written from the
  engine's fields are byte-identical to those written from the oracle's.
"""
import numpy as np
import pytest

import oracle_py as O
import pydsmc as D
from test_gpu_parity import _same_particles

pytestmark = pytest.mark.gpu

DIAMETER = 3.912


def _capsule_case(mesh, seed=3):
    p = D.params_apollo(monatomic=False)
    p.Npinitial = 8
    D.derive(mesh, p)
    p.NSS = 3
    p.surface_stats = 1
    return p, D.fill_freestream(mesh, p, ppc=8.0, seed=seed)


def _capsule_mesh():
    return D.Mesh.capsule(a=6, c=12, layers=8, diameter=DIAMETER)


def test_cas_round_trip_runs_like_the_original(tmp_path):
    orig = _capsule_mesh()
    path = str(tmp_path / "capsule.cas")
    orig.write_cas(path)
    read = D.Mesh.read_cas(path)
    a, b = orig.arrays(), read.arrays()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    p, parts = _capsule_case(orig)
    pr, parts_r = _capsule_case(read)
    print(f"[cas round trip] cells={read.ncell} faces={read.nface} particles={parts.n}", flush=True)
    for f in ("x", "y", "z", "vx", "vy", "vz", "ir", "cell", "id"):
        assert np.array_equal(getattr(parts, f), getattr(parts_r, f)), f
    e0 = D.Engine(orig, p, seed=11)
    e0.upload(parts)
    e1 = D.Engine(read, pr, seed=11)
    e1.upload(parts_r)
    orc = O.Oracle(read, pr, seed=11)
    orc.set_particles(parts_r)
    for s in range(1, 7):
        e0.step(s)
        e1.step(s)
        orc.step(s)
    g = e1.download()
    _same_particles(g, e0.download())
    _same_particles(g, orc.particles())
    for which in (D.MACRO_LOCAL, D.MACRO_OUTPUT):
        assert np.array_equal(e1.macro(which), orc.macro(which)), which
    assert np.allclose(e1.surface(), orc.surface(), rtol=1e-12, atol=0.0)


def test_output_files_match_the_oracle(tmp_path):
    mesh = _capsule_mesh()
    p, parts = _capsule_case(mesh)
    eng = D.Engine(mesh, p, seed=11)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=11)
    orc.set_particles(parts)
    steps = 12
    for s in range(1, steps + 1):
        eng.step(s)
        orc.step(s)
    print(f"[output files] cells={mesh.ncell} particles={eng.count()} steps={steps}", flush=True)
    # cell fields file: what out2dat writes at this step
    w_e, w_o = eng.macro(D.MACRO_OUTPUT), orc.macro(D.MACRO_OUTPUT)
    assert np.array_equal(w_e, w_o)
    ce, co = tmp_path / "cells_engine.dat", tmp_path / "cells_oracle.dat"
    D.write_fluent_cells(str(ce), p, w_e)
    D.write_fluent_cells(str(co), p, w_o)
    assert ce.read_bytes() == co.read_bytes()
    # wall fields file: the engine's fields against the oracle's restatement
    # of outBoundaryStressHeat on the same tallies (bit for bit), and on the
    # oracle's own tallies (which differ only in the order of the engine's
    # atomic sums)
    f_e, cd, qpeak = eng.surface_fields(steps, DIAMETER)
    f_o = orc.surface_fields(steps, eng.surface())
    assert np.array_equal(f_e, f_o)
    we, wo = tmp_path / "walls_engine.dat", tmp_path / "walls_oracle.dat"
    D.write_fluent_walls(str(we), mesh, f_e)
    D.write_fluent_walls(str(wo), mesh, f_o)
    assert we.read_bytes() == wo.read_bytes()
    f_own = orc.surface_fields(steps)
    wall = np.linalg.norm(f_own[:, 5:8], axis=1) > 0.5
    assert wall.sum() > 0 and np.abs(f_own[wall, 8]).max() > 0
    np.testing.assert_allclose(f_e, f_own, rtol=1e-10, atol=1e-12 * np.abs(f_own).max())
    # drag coefficient: force on the wall along the free stream over
    # q_inf pi (D/2)^2 (q_inf = |U|^2 in the output's unit), from the
    # oracle's fields
    U = np.array(list(p.u_inf))
    area = mesh.arrays()["face_area"]
    drag = float(np.sum((f_own[wall, :3] @ U) / np.linalg.norm(U) * area[wall]))
    cd_o = drag / (U @ U * np.pi * 0.25 * DIAMETER ** 2)
    print(f"[output files] C_D engine {cd:.6f} oracle {cd_o:.6f}, peak q {qpeak:.4f}", flush=True)
    assert abs(cd - cd_o) <= 1e-9 * abs(cd_o)
    assert qpeak == pytest.approx(float(np.max(f_own[wall, 3] + f_own[wall, 4])), rel=1e-10)
