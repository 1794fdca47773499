"""Fluent-format output files (CPU): the cell fields file of out2dat
(out2Fluent_dsmc, src/meshImport.cpp:1107-1182) and the wall fields file of
outBoundaryStressHeat (src/ProcessDSMC.cpp:3590-3652), parsed back."""
import math
import re

import numpy as np

import pydsmc as D


def _records(text):
    """{var-id: (zone, size, first, last, rows)} of the 300-records."""
    out = {}
    for m in re.finditer(r"\(300 \((\d+) (\d+) (\d+) 0 0 (\d+) (\d+)\)\n\(\n(.*?)\)\)\n", text, re.S):
        var, zone, size, first, last = (int(m.group(i)) for i in range(1, 6))
        rows = [[float(x) for x in ln.split()] for ln in m.group(6).strip("\n").split("\n")]
        out.setdefault(var, []).append((zone, size, first, last, np.array(rows)))
    return out


def test_cell_fields_file(tmp_path):
    m = D.Mesh.capsule(a=3, c=6, layers=3, diameter=3.912)
    p = D.params_apollo(monatomic=False)
    D.derive(m, p)
    rng = np.random.default_rng(3)
    w = rng.normal(size=(m.ncell, 18)) * 10.0 ** rng.integers(-3, 4, size=(m.ncell, 18))
    path = str(tmp_path / "cells.dat")
    D.write_fluent_cells(path, p, w, zone_id=2, first_id=1)
    text = open(path).read()
    assert text.startswith('(0 " (300 (var-id zone-id var-size 0 0 first-id last-id)(......))" )\n')
    rec = _records(text)
    assert sorted(rec) == [700, 701, 702, 710, 711, 712]
    ts = p.T_in / p.T_ref
    vs = math.sqrt(p.gamma / 2.0) * p.v_in / p.v_rms
    want = {700: (1, w[:, [0]]), 701: (3, w[:, 1:4] / vs), 702: (2, w[:, [4, 14]] / ts),
            710: (6, w[:, 5:11] / ts), 711: (3, w[:, 11:14] / (ts * vs)), 712: (3, w[:, 15:18] / (ts * vs))}
    for var, (size, vals) in want.items():
        (zone, sz, first, last, rows), = rec[var]
        assert (zone, sz, first, last) == (2, size, 1, m.ncell)
        assert rows.shape == vals.shape
        # default stream precision: 6 significant digits
        assert np.allclose(rows, vals, rtol=6e-6, atol=0), var


def test_wall_fields_file(tmp_path):
    m = D.Mesh.capsule(a=3, c=6, layers=3, diameter=3.912)
    a = m.arrays()
    walls = np.flatnonzero(a["face_tag"] == D.TAG_WALL)
    rng = np.random.default_rng(4)
    f = np.zeros((m.nface, 12))
    f[walls] = rng.normal(size=(walls.size, 12))
    path = str(tmp_path / "walls.dat")
    D.write_fluent_walls(path, m, f, zone_id=7, first_id=1)
    rec = _records(open(path).read())
    runs = np.split(walls, np.flatnonzero(np.diff(walls) != 1) + 1)
    assert sorted(rec) == [700, 701, 702, 703, 704, 705, 706]
    for var, cols in ((700, [0, 1, 2]), (701, [3]), (702, [4]), (703, [5, 6, 7]), (704, [8]), (705, [9, 10, 11])):
        assert len(rec[var]) == len(runs)
        for k, (zone, size, first, last, rows) in enumerate(rec[var]):
            r = runs[k]
            assert (zone, size, first, last) == (7 + k, len(cols), r[0] + 1, r[-1] + 1)
            # 16 significant digits (the reference's setprecision(16))
            assert np.allclose(rows, f[r][:, cols], rtol=1e-15, atol=0)
    for k, (_, _, _, _, rows) in enumerate(rec[706]):
        assert np.allclose(rows[:, 0], f[runs[k], 3] + f[runs[k], 4], rtol=1e-15, atol=0)
