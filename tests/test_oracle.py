"""Pins the CPU oracle (CPU only, no GPU).

The reference ships no tests, golden vectors or fixtures for this path, so
each oracle function is checked against what the mathematics fixes:
published known-answer vectors (Philox), libm / extended precision (the
deterministic elementary functions), exact rational arithmetic
(Moller-Trumbore), hand-derived cases (quad splits, macro formulas),
conservation laws and closed-form rates (collisions, wall reflection,
inflow flux, sampling windows).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O
import pydsmc as D


# ---- random numbers -----------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10 (Salmon et al., SC'11)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(O.philox(ctr, key)) == want


def test_stream_uniforms_open_interval():
    u = O.stream_uniforms(12345, 7, 3, 1, 0, 400000)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / len(u))
    assert abs((u ** 2).mean() - 1 / 3) < 5e-3
    # different counters give different streams
    v = O.stream_uniforms(12345, 8, 3, 1, 0, 1000)
    assert not np.array_equal(u[:1000], v)


# ---- deterministic elementary functions ---------------------------------------------
def _ulp(x):
    return np.spacing(np.abs(x))


def test_det_math_against_libm():
    rng = np.random.default_rng(1)
    u = np.concatenate([rng.random(100000), 10.0 ** rng.uniform(-300, 0, 20000), [1.0, 0.5, 2.0 ** -53]])
    u = u[u > 0]
    h = O.host_math(u)
    assert np.all(np.abs(h[:, 0] - np.log(u)) <= 2 * _ulp(np.log(u)) + 1e-300)
    assert h[np.flatnonzero(u == 1.0)[0], 0] == 0.0
    e = np.exp(-20.0 * u)
    assert np.all(np.abs(h[:, 1] - e) <= 3 * _ulp(e) + 1e-320)
    # sin/cos(2 pi u) against extended precision
    two_pi = np.longdouble("6.283185307179586476925286766559005768394")
    ang = two_pi * u.astype(np.longdouble)
    s_ref, c_ref = np.sin(ang), np.cos(ang)
    assert np.max(np.abs(h[:, 2] - s_ref.astype(np.float64))) < 4e-16
    assert np.max(np.abs(h[:, 3] - c_ref.astype(np.float64))) < 4e-16
    assert np.array_equal(h[:, 4], np.sqrt(u))
    # exp(y log x): the rounding of y log x is amplified by |y log x|
    pw = u ** 0.26
    assert np.all(np.abs(h[:, 5] - pw) / pw < 4e-16 * (2 + np.abs(0.26 * np.log(u))))


# ---- Moller-Trumbore (intersectCachedTri) --------------------------------------------
def _exact_hit(a, e1, e2, x, v):
    F = [Fraction(float(q)) for q in np.concatenate([a, e1, e2, x, v])]
    A, E1, E2, X, V = F[0:3], F[3:6], F[6:9], F[9:12], F[12:15]

    def cross(p, q):
        return [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]

    def dot(p, q):
        return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]

    pv = cross(V, E2)
    det = dot(E1, pv)
    if det == 0:
        return None
    T = [X[i] - A[i] for i in range(3)]
    uu = dot(T, pv) / det
    qv = cross(T, E1)
    vv = dot(V, qv) / det
    t = dot(E2, qv) / det
    return uu, vv, t, det


def test_moller_trumbore_against_exact_rationals():
    rng = np.random.default_rng(2)
    checked = hits = 0
    for _ in range(3000):
        a, e1, e2 = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        v = rng.normal(size=3)
        # aim at a point near the triangle, from either side
        b1, b2, t0 = rng.uniform(-0.3, 1.2), rng.uniform(-0.3, 1.2), rng.uniform(-1.0, 3.0)
        x = a + b1 * e1 + b2 * e2 - v * t0
        ex = _exact_hit(a, e1, e2, x, v)
        if ex is None:
            continue
        uu, vv, t, det = ex
        margins = [float(uu), 1 - float(uu), float(vv), 1 - float(uu + vv), float(t)]
        if abs(float(det)) < 1e-6 or min(abs(m) for m in margins) < 1e-9:
            continue  # exact answer too close to a decision boundary for FP64
        want = uu >= 0 and uu <= 1 and vv >= 0 and uu + vv <= 1 and t > 0
        hit, tt = O.tri_hit(a, e1, e2, x, v)
        assert hit == want
        if hit:
            hits += 1
            assert abs(tt - float(t)) <= 1e-12 * max(1.0, abs(float(t)))
        checked += 1
    assert checked > 2000 and hits > 100


def test_moller_trumbore_edge_cases():
    a, e1, e2 = np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    down = np.array([0, 0, -1.0])
    # ray through an interior point at height 2 -> t = 2
    assert O.tri_hit(a, e1, e2, np.array([0.25, 0.25, 2.0]), down) == (True, 2.0)
    # vertex u = 0, v = 0 counts as a hit; outside the hypotenuse misses
    assert O.tri_hit(a, e1, e2, np.array([0.0, 0.0, 1.0]), down)[0]
    assert not O.tri_hit(a, e1, e2, np.array([0.6, 0.6, 1.0]), down)[0]
    # parallel ray (det = 0) and a hit behind the origin (t < 0) miss
    assert not O.tri_hit(a, e1, e2, np.array([0.2, 0.2, 1.0]), np.array([1.0, 0, 0]))[0]
    assert not O.tri_hit(a, e1, e2, np.array([0.2, 0.2, -1.0]), down)[0]
    # t <= 1e-14 (particle on the face) is not a hit
    assert not O.tri_hit(a, e1, e2, np.array([0.2, 0.2, 1e-15]), down)[0]


# ---- quad split choice --------------------------------------------------------------------
def _quad_mesh(P):
    # a single wedge-free test: a tet-like closed cell is not needed; the
    # split choice only reads the face, so build a one-hex mesh whose bottom
    # face is P and read its tag
    top = [np.array(q) + np.array([0, 0, 5.0]) for q in P]
    nodes = np.array(list(P) + top, dtype=float)
    hexf = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]
    m = D.Mesh.build(nodes, hexf, [[0, -1]] * 6, [4] * 6, [6] * 6, ncell=1)
    return m


def test_split_tags_hand_checked():
    # planar square: both splits score 0 -> 02
    sq = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    m = _quad_mesh(sq)
    assert m.arrays()["face_split"][0] == 0
    assert O.split_tags(m)[0] == 0
    # lifted corner: 02 scores 16 (1 - 1/2) = 8, 13 scores
    # 16 (1 - 1/sqrt 3) + (sqrt 3 - 1)/(sqrt 3 + 1) = 7.03 -> 13
    tw = [(0, 0, 0), (1, 0, 0), (1, 1, 1), (0, 1, 0)]
    m = _quad_mesh(tw)
    assert m.arrays()["face_split"][0] == 1
    assert O.split_tags(m)[0] == 1
    # lifting node 3 instead: diagonal 0-2 gives normals (-1,1,-1), (0,0,-1)
    # -> 16 (1 - 1/sqrt 3) + 0.268 = 7.03; diagonal 1-3 gives (0,1,-1),
    # (-1,0,-1) -> 16 (1 - 1/2) = 8 -> 02
    m = _quad_mesh([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 1)])
    assert m.arrays()["face_split"][0] == 0


def test_split_tags_match_oracle_on_capsule():
    m = D.Mesh.capsule(a=6, c=12, layers=4, jitter=0.25, seed=3)
    tags = m.arrays()["face_split"]
    quads = m.arrays()["face_type"] == 4
    assert np.array_equal(tags[quads], O.split_tags(m)[quads])
    assert 0 < (tags[quads] == 1).sum() < quads.sum()


# ---- mesh geometry ----------------------------------------------------------------------------
def test_apollo_surrogate_topology():
    m = D.Mesh.capsule()  # a=25, c=62, 50 layers
    a = m.arrays()
    # decoded from mesh/tecplot-derived/1720/3dapollo372500.cas.tmd
    assert (m.ncell, m.nface, m.view.nnode) == (372500, 1124950, 380052)
    tags = a["face_tag"]
    assert [(tags == t).sum() for t in (3, 4, 5, 6)] == [1110050, 3725, 3725, 7450]
    assert np.all(a["cell_nface"] == 6) and np.all(a["cell_volume"] > 0)
    w = a["node_xyz"][a["face_nodes"][tags == 6].ravel()]
    assert abs(w[:, 1].max() - 1.956) < 5e-3 and w[:, 0].min() > -1e-9


@pytest.mark.parametrize("kind", ["tets", "hexes"])
def test_box_geometry(kind):
    n = 5
    m = D.Mesh.box_tets(n) if kind == "tets" else D.Mesh.box_hexes(n)
    a = m.arrays()
    per = 1 / (6 * n ** 3) if kind == "tets" else 1 / n ** 3
    np.testing.assert_allclose(a["cell_volume"], per, rtol=1e-12)
    # closed cells: sum of outward vector areas vanishes
    acc = np.zeros((m.ncell, 3))
    fc = a["face_cells"]
    va = a["face_normal"] * a["face_area"][:, None]
    np.add.at(acc, fc[:, 0], va)
    inner = fc[:, 1] >= 0
    np.add.at(acc, fc[inner, 1], -va[inner])
    assert np.abs(acc).max() < 1e-14
    # boundary normals point out of the unit cube
    b = ~inner
    fcen = a["node_xyz"][a["face_nodes"][b][:, :3]].mean(axis=1)
    out = np.sign(fcen - 0.5) * (np.abs(fcen - 0.5) > 0.49)
    assert np.all(np.sum(out * a["face_normal"][b], axis=1) > 0.99)


def test_params_apollo_normalisation():
    m = D.Mesh.capsule(a=6, c=12, layers=4)
    p = D.derive(m, D.params_apollo(monatomic=True))
    # closed forms: v_rms = sqrt(2 kB T_ref / m); dt_ref = L0 / v_rms
    assert abs(p.v_rms - math.sqrt(2 * 1.3806e-23 * 273.15 / 4.65e-26)) < 1e-9
    assert abs(p.dt_ref - 3.912 / p.v_rms) < 1e-15
    assert abs(p.n_ref - 3.4305e19) / 3.4305e19 < 1e-4
    U = math.hypot(p.u_inf[0], p.u_inf[1])
    assert abs(U - 5 * math.sqrt(5 / 6) * math.sqrt(142.2 / 273.15)) < 1e-12
    assert abs(p.u_inf[1] / p.u_inf[0] - math.tan(math.radians(30))) < 1e-12
    pd = D.derive(m, D.params_apollo(monatomic=False))
    assert abs(math.hypot(pd.u_inf[0], pd.u_inf[1]) - 3.0183) < 1e-4


def test_fill_counts_and_moments():
    m = D.Mesh.box_tets(6)
    p = D.derive(m, D.params_box(0.05))
    parts = D.fill_freestream(m, p, ppc=40.0, temperature=1.0, seed=9)
    assert abs(parts.n - 40 * m.ncell) < 5 * math.sqrt(m.ncell)
    for c in ("vx", "vy", "vz"):
        v = getattr(parts, c)
        assert abs(v.mean()) < 5 * math.sqrt(0.5 / parts.n)
        assert abs(v.var() - 0.5) < 0.02
    assert np.all(parts.ir == 0)
    # positions inside the unit cube and spread over every cell
    for c in ("x", "y", "z"):
        assert getattr(parts, c).min() > 0 and getattr(parts, c).max() < 1
    assert np.bincount(parts.cell, minlength=m.ncell).min() > 20
