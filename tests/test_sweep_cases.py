"""The seeded capsule cases of the batched sweep and the serving queue are
valid meshes (CPU): every face's outward normal points away from the
centroid of each cell it bounds, so no tet of the 12-tet hex split is
inverted (inverted tets trap walks, which are then dropped)."""
import os
import sys

import numpy as np
import pytest

import pydsmc as D

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dig-appolo_amd"))
import sweep as SW  # noqa: E402


def _inward_faces(m):
    a = m.arrays()
    xyz, fn, fc, ft, nrm = a["node_xyz"], a["face_nodes"], a["face_cells"], a["face_type"], a["face_normal"]
    fctr = np.where(ft[:, None] == 4, xyz[fn[:, :4]].mean(1), xyz[fn[:, :3]].mean(1))
    cc = a["cell_centroid"]
    c0, c1 = fc[:, 0], fc[:, 1]
    bad = int(np.sum(np.einsum("ij,ij->i", fctr - cc[c0], nrm) <= 0))
    inner = (c1 >= 0) & (c1 < m.ncell)
    bad += int(np.sum(np.einsum("ij,ij->i", fctr[inner] - cc[c1[inner]], -nrm[inner]) <= 0))
    return bad


@pytest.mark.parametrize("seed,scale", [(3, 1.0), (29, 1.0), (0, 0.1), (5, 0.1), (1, 0.02), (2, 0.01)])
def test_sweep_case_meshes_have_no_inverted_cells(seed, scale):
    sp = SW.case_spec(seed, scale)
    m, p = SW.build_case(sp, D.PREC_MIXED, steps=100)
    assert m.ncell == sp["cells"]
    assert _inward_faces(m) == 0
    # near-wall cells are not slivers: the shortest edge is a fair fraction
    # of the body's size (diameter 1)
    assert p.min_edge > 1e-3, p.min_edge


def test_apollo_surrogate_has_no_inverted_cells():
    assert _inward_faces(D.Mesh.capsule(25, 62, 50, 3.912)) == 0
