"""Cell renumbering along a space-filling curve (dsmc_mesh_renumber), CPU:
the renumbered mesh is the same mesh (per-cell geometry permuted bit for
bit, faces untouched), tracking on it is the same tracking, and neighbouring
cells end up close in memory."""
import numpy as np
import pytest

import oracle_py as O
import pydsmc as D


@pytest.mark.parametrize("kind", ["box_tets", "box_hexes", "capsule_tets"])
def test_renumbered_mesh_is_a_permutation(kind):
    m = {"box_tets": lambda: D.Mesh.box_tets(6), "box_hexes": lambda: D.Mesh.box_hexes(7),
         "capsule_tets": lambda: D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912, tets=True, jitter=0.2,
                                                seed=3)}[kind]()
    r, perm = m.renumber(D.ORDER_MORTON)
    assert np.array_equal(np.sort(perm), np.arange(m.ncell))
    a, b = m.arrays(), r.arrays()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.ncell)
    for k in ("cell_volume", "cell_centroid", "cell_nface", "cell_faces"):
        assert np.array_equal(b[k], a[k][perm]), k
    for k in ("node_xyz", "face_nodes", "face_type", "face_tag", "face_split", "face_normal", "face_area"):
        assert np.array_equal(b[k], a[k]), k
    fc = a["face_cells"].copy()
    inner = (fc >= 0) & (fc < m.ncell)
    fc[inner] = inv[fc[inner]]
    assert np.array_equal(b["face_cells"], fc)


def test_tracking_is_unchanged_by_renumbering():
    """Specular walls and no collisions: the move draws nothing at random, so
    every particle ends where it ends on the original mesh, in the cell the
    permutation maps its original cell to, after the same face crossings."""
    m = D.Mesh.box_tets(5)
    r, perm = m.renumber(D.ORDER_MORTON)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.ncell)
    p = D.params_box(0.05)
    p.Npinitial = 20
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=20.0, temperature=1.0, seed=4)
    pr = parts.copy()
    pr.cell[:] = inv[parts.cell]
    out = []
    for mesh, pp in ((m, parts), (r, pr)):
        o = O.Oracle(mesh, p, seed=1)
        o.set_particles(pp)
        o.log_enable(True)
        for s in range(1, 9):
            o.step(s, D.STAGE_MOVE | D.STAGE_SORT)
        q = o.particles()
        order = np.argsort(q.id)
        ev = o.log_get()
        out.append((q, order, ev[np.lexsort((ev[:, 2], ev[:, 1], ev[:, 0]))]))
    (q0, o0, e0), (q1, o1, e1) = out
    assert o0.size == o1.size
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(q0, f)[o0], getattr(q1, f)[o1]), f
    assert np.array_equal(inv[q0.cell[o0]], q1.cell[o1])
    assert np.array_equal(e0, e1)  # crossings: same faces, tags and kinds


def test_morton_order_brings_neighbours_together():
    """A box numbered at random: after renumbering, most face neighbours are
    within 64 cells of each other (the generator's own box is already in
    Morton order, and renumbering it changes nothing measurable)."""
    b = D.Mesh.box_tets(12)
    a = b.arrays()
    shuffle = np.random.default_rng(1).permutation(b.ncell).astype(np.int32)
    fc = a["face_cells"].copy()
    inner = (fc >= 0) & (fc < b.ncell)
    fc[inner] = shuffle[fc[inner]]
    m = D.Mesh.build(a["node_xyz"], a["face_nodes"], fc, a["face_type"], a["face_tag"], b.ncell)
    r, _ = m.renumber(D.ORDER_MORTON)

    def near(mesh, w):
        f = mesh.arrays()["face_cells"]
        ok = (f[:, 1] >= 0) & (f[:, 1] < mesh.ncell)
        return float(np.mean(np.abs(f[ok, 0] - f[ok, 1]) < w))
    assert near(m, 64) < 0.05 and near(r, 64) > 0.75, (near(m, 64), near(r, 64))
