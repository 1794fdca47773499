"""The C++ host driver (dig-appolo_amd/host/dsmc_run.cpp: the reference's time
loop, src/meshImportTest.cpp:194-313, without MPI) on the GPU.  A short
Apollo run with surface tallies writes the reference-format cell fields file
(out2dat, src/ProcessDSMC.cpp:3421-3424) and wall fields file
(outBoundaryStressHeat, :3618-3621) at the reference's output cadence."""
import glob
import os
import subprocess

import numpy as np
import pytest

from test_fluent_output import _records

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dig-appolo_amd", "dsmc_run")


def test_host_driver_writes_reference_outputs(tmp_path):
    assert os.path.exists(DRIVER), "host driver not built (run __graft_entry__.build())"
    print("[host_driver] apollo 100 steps, 2 particles per cell, surface tallies", flush=True)
    r = subprocess.run([DRIVER, "apollo", "100", "2", "--surface", "--out", str(tmp_path)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout
    assert "Finished 50 of 100 steps" in out and "Finished 100 of 100 steps" in out
    assert "Total running time" in out and "C_D = " in out
    # output cadence: step 100 only (every 100 steps up to 2000)
    cells = glob.glob(str(tmp_path / "DSMCKn*_iter*_CELL372500.dat"))
    assert [os.path.basename(c).split("_iter")[1] for c in cells] == ["100_CELL372500.dat"]
    rec = _records(open(cells[0]).read())
    assert sorted(rec) == [700, 701, 702, 710, 711, 712]
    (zone, size, first, last, rho), = rec[700]
    assert (zone, size, first, last) == (2, 1, 1, 372500)
    assert rho.shape == (372500, 1) and np.isfinite(rho).all() and (rho >= 0).all() and rho.mean() > 0.1
    walls = glob.glob(str(tmp_path / "BoundaryStressHeatFluentKn*_iter100.dat"))
    assert len(walls) == 1
    wrec = _records(open(walls[0]).read())
    assert sorted(wrec) == [700, 701, 702, 703, 704, 705, 706]
    nwall = sum(blk[4].shape[0] for blk in wrec[701])
    assert nwall == 7450  # the capsule's wall quads
    # 704: pressure, 706: total heat flux; both positive on the windward side
    for var in (704, 706):
        v = np.concatenate([blk[4][:, 0] for blk in wrec[var]])
        assert v.shape == (7450,) and np.isfinite(v).all() and v.max() > 0.0, var
