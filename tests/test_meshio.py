"""Fluent ASCII case meshes (CPU): the reader (readCAS, src/meshImport.cpp:
177-576) on a hand-written two-hex file, and write/read round trips."""
import os

import numpy as np
import pytest

import pydsmc as D

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_hand_written_two_hexes():
    """Hex ids (c nodes, b faces), a uniform-quad zone and a mixed zone with
    node counts, c1 = 0 on the boundary; volumes, centroids and the shared
    face follow from the geometry."""
    m = D.Mesh.read_cas(os.path.join(GOLDEN, "two_hexes.cas"))
    a = m.arrays()
    assert (m.ncell, m.nface, m.view.nnode) == (2, 11, 12)
    assert np.allclose(a["cell_volume"], [1.0, 1.0], rtol=1e-14)
    assert np.allclose(a["cell_centroid"], [[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]], rtol=1e-14)
    assert list(a["face_tag"]) == [3] + [6] * 10
    assert list(a["face_type"]) == [4] * 11
    assert list(a["face_cells"][0]) == [0, 1]
    assert np.all(a["face_cells"][1:, 1] < 0)
    assert list(a["face_nodes"][0]) == [1, 4, 10, 7]  # 2 5 b 8, 0-based
    assert sorted(a["cell_nface"]) == [6, 6]
    # outward unit normals on the boundary: their area-weighted sum closes
    n = a["face_normal"] * a["face_area"][:, None]
    for c in (0, 1):
        s = np.zeros(3)
        for f in range(m.nface):
            c0, c1 = a["face_cells"][f]
            if c0 == c:
                s += n[f]
            elif c1 == c:
                s -= n[f]
        assert np.allclose(s, 0.0, atol=1e-14)


@pytest.mark.parametrize("kind", ["capsule", "capsule_tets", "box"])
def test_write_read_round_trip(tmp_path, kind):
    if kind == "capsule":
        m = D.Mesh.capsule(a=4, c=8, layers=3, diameter=3.912)
    elif kind == "capsule_tets":
        m = D.Mesh.capsule(a=3, c=6, layers=3, diameter=3.912, tets=True, jitter=0.2, seed=5)
    else:
        m = D.Mesh.box_tets(3)
    path = str(tmp_path / "m.cas")
    m.write_cas(path)
    r = D.Mesh.read_cas(path)
    a, b = m.arrays(), r.arrays()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_read_rejects_a_truncated_file(tmp_path):
    src = open(os.path.join(GOLDEN, "two_hexes.cas")).read()
    path = tmp_path / "bad.cas"
    path.write_text(src[: src.index("4 8 9 c b 2 0")])
    with pytest.raises(RuntimeError):
        D.Mesh.read_cas(str(path))


@pytest.mark.parametrize("old,new", [("(10 (1 1 c 1 3)", "(10 (1 0 c 1 3)"),
                                      ("(13 (3 1 1 2 4)", "(13 (3 0 1 2 4)")])
def test_zero_based_section_is_refused(tmp_path, old, new):
    """A node or face section numbered from 0 would put its first record
    before the start of the arrays: the reader refuses it."""
    text = open(os.path.join(GOLDEN, "two_hexes.cas")).read()
    assert old in text, old
    bad = tmp_path / "zero_based.cas"
    bad.write_text(text.replace(old, new, 1))
    with pytest.raises(D.DsmcError) as ei:
        D.Mesh.read_cas(str(bad))
    assert ei.value.code == D.ERR_MESH and "range" in str(ei.value)
