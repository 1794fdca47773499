"""Oracle physics against conservation laws and closed forms (CPU only)."""
import math
import re
import ctypes
import os

import numpy as np
import pytest

import oracle_py as O
import pydsmc as D

KB = 1.3806e-23


def _box(n=4, ppc=30.0, Kn=0.05, wall=D.WALL_SPECULAR, T=1.0, seed=5, hex_=False):
    m = D.Mesh.box_hexes(n) if hex_ else D.Mesh.box_tets(n)
    p = D.params_box(Kn)
    p.Npinitial = int(ppc)
    p.wall_model = wall
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=ppc, temperature=T, seed=seed)
    return m, p, parts


def test_collisions_conserve_momentum_and_energy():
    m, p, parts = _box(n=3, ppc=40)
    p.P_relax, p.dr = 0.3, 2.0  # exercise Larsen-Borgnakke exchange
    parts.ir[:] = np.random.default_rng(0).exponential(0.5, parts.n)
    o = O.Oracle(m, p, seed=3)
    o.set_particles(parts)
    P0 = np.array([parts.vx.sum(), parts.vy.sum(), parts.vz.sum()])
    E0 = 0.5 * (parts.vx ** 2 + parts.vy ** 2 + parts.vz ** 2).sum() + parts.ir.sum()
    for s in range(1, 11):
        o.step(s, D.STAGE_COLLIDE)
    q = o.particles()
    assert o.stats()["collisions"] > 1000
    P1 = np.array([q.vx.sum(), q.vy.sum(), q.vz.sum()])
    E1 = 0.5 * (q.vx ** 2 + q.vy ** 2 + q.vz ** 2).sum() + q.ir.sum()
    assert np.abs(P1 - P0).max() < 1e-9
    assert abs(E1 - E0) / E0 < 1e-13
    assert not np.allclose(q.vx, parts.vx)  # something happened


def test_specular_box_conserves_energy_and_count():
    m, p, parts = _box(n=4, ppc=20)
    o = O.Oracle(m, p, seed=7)
    o.set_particles(parts)
    E0 = (parts.vx ** 2 + parts.vy ** 2 + parts.vz ** 2).sum()
    for s in range(1, 31):
        o.step(s)
    q = o.particles()
    st = o.stats()
    assert q.n == parts.n and st["dropped"] == 0 and st["wall_hits"] > 0
    assert abs((q.vx ** 2 + q.vy ** 2 + q.vz ** 2).sum() - E0) / E0 < 1e-12
    assert q.x.min() >= 0 and q.x.max() <= 1 and q.y.min() >= 0 and q.z.max() <= 1


def _point_in_tet(P, tet):
    a, b, c, d = tet
    M = np.stack([b - a, c - a, d - a], axis=1)
    lam = np.linalg.solve(M, P - a)
    return np.all(lam >= -1e-9) and lam.sum() <= 1 + 1e-9


def test_tracking_lands_in_the_right_cell():
    m, p, parts = _box(n=3, ppc=15)
    a = m.arrays()
    o = O.Oracle(m, p, seed=1)
    o.set_particles(parts)
    o.step(1, D.STAGE_MOVE | D.STAGE_SORT)
    q = o.particles()
    # speeds are unchanged by specular walls; every particle is inside its cell
    sp0 = np.sort(parts.vx ** 2 + parts.vy ** 2 + parts.vz ** 2)
    sp1 = np.sort(q.vx ** 2 + q.vy ** 2 + q.vz ** 2)
    np.testing.assert_allclose(sp0, sp1, rtol=1e-14)
    nodes = a["node_xyz"]
    for k in range(0, q.n, 37):
        c = q.cell[k]
        ids = np.unique(a["face_nodes"][a["cell_faces"][c][:4]][:, :3])
        assert len(ids) == 4
        assert _point_in_tet(np.array([q.x[k], q.y[k], q.z[k]]), nodes[ids])
    # free flight: particles that hit no wall moved by exactly v * dtime
    ev = o.log_get() if False else None  # log not enabled here
    byid0 = np.argsort(parts.id)
    byid1 = np.argsort(q.id)
    x0 = np.stack([parts.x, parts.y, parts.z], 1)[byid0]
    x1 = np.stack([q.x, q.y, q.z], 1)[byid1]
    v0 = np.stack([parts.vx, parts.vy, parts.vz], 1)[byid0]
    v1 = np.stack([q.vx, q.vy, q.vz], 1)[byid1]
    free = np.all(v0 == v1, axis=1)
    assert free.mean() > 0.5
    np.testing.assert_allclose(x1[free], x0[free] + v0[free] * p.dtime, atol=1e-13)


def test_hard_sphere_collision_rate():
    # Bird (1994) eq. 1.10 / 4.64: nu = n sigma <c_r>, <c_r> = 4 sqrt(kT/(pi m));
    # collisions per cell and step = 1/2 N (N-1)/V Neff sigma <c_r> dt.
    m, p, parts = _box(n=4, ppc=40, Kn=0.05)
    o = O.Oracle(m, p, seed=9)
    o.set_particles(parts)
    steps = 8
    counts = np.bincount(parts.cell, minlength=m.ncell).astype(float)
    vol = m.arrays()["cell_volume"]
    sigma = math.pi * p.d_ref ** 2
    cr = 4 * math.sqrt(KB * p.T_ref / (math.pi * p.mass))
    dt = p.dtime * p.dt_ref
    per_step = np.sum(0.5 * counts * (counts - 1) / vol) * p.Neff * sigma * cr * dt
    for s in range(1, steps + 1):
        o.step(s, D.STAGE_COLLIDE)
    got = o.stats()["collisions"]
    want = per_step * steps
    assert abs(got - want) < 4 * math.sqrt(want) + 0.02 * want, (got, want)


def test_diffuse_wall_half_range_maxwellian():
    # one hex cell, particles launched into the x=0 wall: the re-emitted
    # normal speed has mean sqrt(pi Tw)/2, tangential components variance
    # Tw/2, rotational energy mean Tw/2 (v_rms units, src/ProcessDSMC.cpp:1808-1818)
    m = D.Mesh.box_hexes(1)
    p = D.params_box(0.05)
    p.wall_model = D.WALL_DIFFUSE
    p.T_wall = 300.0
    p.dr = 2.0
    D.derive(m, p)
    p.dtime = 0.02
    n = 200000
    parts = D.Particles(n)
    parts.x[:] = 0.005
    parts.y[:] = 0.5
    parts.z[:] = 0.5
    parts.vx[:] = -1.0
    parts.id[:] = np.arange(n)
    o = O.Oracle(m, p, seed=4)
    o.set_particles(parts)
    o.step(1, D.STAGE_MOVE | D.STAGE_SORT)
    q = o.particles()
    assert o.stats()["wall_hits"] == n
    Tw = 300.0 / 273.15
    tol = 5 / math.sqrt(n)
    assert q.vx.min() > 0
    assert abs(q.vx.mean() - math.sqrt(math.pi * Tw) / 2) < tol
    assert abs(q.vy.mean()) < tol and abs(q.vz.mean()) < tol
    assert abs(q.vy.var() - Tw / 2) < 4 * tol and abs(q.vz.var() - Tw / 2) < 4 * tol
    assert abs(q.ir.mean() - Tw / 2) < 4 * tol


def test_inflow_flux_matches_one_way_maxwellian():
    # Bird (1994) eq. 4.22: inward number flux through a surface element with
    # inward unit normal at angle theta to U: n/(2 sqrt(pi) beta) [exp(-s^2 cos^2)
    # + sqrt(pi) s cos (1 + erf(s cos))], s = U beta.  Over S steps the emitted
    # count per inlet cell equals S * rate up to the carried remainder (< 1).
    m = D.Mesh.capsule(a=5, c=10, layers=4)
    p = D.params_apollo(monatomic=True)
    p.Npinitial = 5
    D.derive(m, p)
    a = m.arrays()
    o = O.Oracle(m, p, seed=2)
    o.set_particles(D.Particles(0))
    S = 12
    for s in range(1, S + 1):
        o.step(s, D.STAGE_EMIT | D.STAGE_SORT)
    got = o.stats()["emitted"]
    beta = 1 / math.sqrt(2 * KB / p.mass * p.T_in)
    U = np.array(p.u_inf) * p.v_rms
    s_ = np.linalg.norm(U) * beta
    inlet = np.flatnonzero(a["face_tag"] == D.TAG_INLET)
    rate = 0.0
    for f in inlet:
        nodes = a["node_xyz"][a["face_nodes"][f][: a["face_type"][f]]]
        # flux through each triangle of the quad with its own normal
        tris = [(0, 1, 2), (0, 2, 3)] if a["face_split"][f] == 0 else [(0, 1, 3), (1, 2, 3)]
        for t in tris:
            P0, P1, P2 = nodes[list(t)]
            nv = np.cross(P1 - P0, P2 - P0)
            area = 0.5 * np.linalg.norm(nv)
            nin = -nv / np.linalg.norm(nv) * np.sign(np.dot(nv, a["face_normal"][f]))
            c = np.dot(U, nin) / np.linalg.norm(U)
            sc = s_ * c
            flux = p.n_ref / (2 * math.sqrt(math.pi) * beta) * (math.exp(-sc * sc) + math.sqrt(math.pi) * sc *
                                                                 (1 + math.erf(sc)))
            rate += flux * area * p.dtime / p.v_rms / p.Neff
    want = rate * S
    ninlet_cells = len(np.unique(a["face_cells"][inlet, 0]))
    assert abs(got - want) <= ninlet_cells, (got, want)


def test_sampling_matches_peculiar_velocity_definitions():
    # writeLocalMacro via statistic_macroPre against the moment definitions in
    # peculiar velocity c = v - u: T = 2/3 <c^2>, sigma_ij = 2 rho (<c_i c_j> -
    # delta_ij <c^2>/3), q_i = rho <c_i c^2>, q_r,i = 2 rho <c_i e_r>, Tr = 4/dr <e_r>
    m = D.Mesh.box_hexes(2)
    p = D.params_box(0.05)
    p.dr = 2.0
    D.derive(m, p)
    rng = np.random.default_rng(3)
    n = 800
    parts = D.Particles(n)
    parts.cell[:] = rng.integers(0, m.ncell, n)
    parts.x[:], parts.y[:], parts.z[:] = 0.5, 0.5, 0.5
    parts.vx[:] = rng.normal(0.7, 1.0, n)
    parts.vy[:] = rng.gamma(2.0, 0.5, n)
    parts.vz[:] = rng.normal(-0.2, 0.6, n)
    parts.ir[:] = rng.exponential(0.8, n)
    parts.id[:] = np.arange(n)
    o = O.Oracle(m, p, seed=1)
    o.set_particles(parts)
    o.statistic_pre()
    loc = o.macro(D.MACRO_LOCAL)
    vol = m.arrays()["cell_volume"]
    for c in range(m.ncell):
        sel = parts.cell == c
        v = np.stack([parts.vx[sel], parts.vy[sel], parts.vz[sel]], 1)
        er = parts.ir[sel]
        N = sel.sum()
        rho = p.Neff / vol[c] * N / p.n_ref
        u = v.mean(0)
        cc = v - u
        c2 = (cc ** 2).sum(1)
        T = 2 / 3 * c2.mean()
        P = cc.T @ cc / N
        sig = 2 * rho * (P - np.eye(3) * c2.mean() / 3)
        q = rho * (cc * c2[:, None]).mean(0)
        qr = 2 * rho * (cc * er[:, None]).mean(0)
        want = [rho, *u, T, sig[0, 0], sig[0, 1], sig[1, 1], sig[1, 2], sig[2, 2], sig[2, 0], *q,
                4 / p.dr * er.mean(), *qr]
        np.testing.assert_allclose(loc[c], want, rtol=1e-10, atol=1e-12)


def test_sampling_windows_closed_form():
    # frozen particles (sample only): stepsum = s m up to NSCHEME, then the
    # window recurrence keeps NSCHEME m; steady after NSS averages to the
    # single-step fields; the interval record every NGSIS equals them too.
    m = D.Mesh.box_hexes(2)
    p = D.params_box(0.05)
    p.dr = 2.0
    p.NSCHEME, p.NSS, p.Nevery, p.Nrepeat = 5, 4, 1, 10
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=30.0, temperature=1.0, seed=8)
    parts.ir[:] = np.random.default_rng(1).exponential(0.5, parts.n)
    o = O.Oracle(m, p, seed=1)
    o.set_particles(parts)
    o.statistic_pre()
    one = o.macro(D.MACRO_LOCAL)
    for s in range(1, 4):
        o.step(s, D.STAGE_SAMPLE)
    loc = o.macro(D.MACRO_LOCAL)
    # rho and every field carrying a factor rho (stress, heat fluxes) see the
    # partly filled window: 3 of NSCHEME = 5 samples; u, T, Tr do not
    scaled = [0, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17]
    want = one.copy()
    want[:, scaled] *= 3 / 5
    np.testing.assert_allclose(loc, want, rtol=1e-12, atol=1e-14)
    for s in range(4, 31):
        o.step(s, D.STAGE_SAMPLE)
    np.testing.assert_allclose(o.macro(D.MACRO_LOCAL), one, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(o.macro(D.MACRO_FINAL), one, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(o.macro(D.MACRO_RECORD), one, rtol=1e-11, atol=1e-13)
    # the interval accumulator restarts after every exchange step (s = 30)
    assert np.all(o.accum(2) == 0.0)
    np.testing.assert_allclose(o.accum(1)[:, 0], 26 * np.bincount(parts.cell, minlength=m.ncell))


def test_sparse_cell_state_machine():
    # a cell with fewer than 3 samples in an interval becomes "accumulating";
    # its interval sums are no longer reset and its record averages over the
    # steps accumulated since (src/ProcessDSMC.cpp:2798-2858).  NGSIS = 1 here.
    m = D.Mesh.box_hexes(2)
    p = D.params_box(0.05)
    p.NSCHEME, p.NSS, p.Nevery, p.Nrepeat = 5, 100, 1, 1
    D.derive(m, p)
    parts = D.Particles(6)
    parts.cell[:] = [0, 0, 1, 1, 1, 1]
    parts.x[:] = 0.25
    parts.vx[:] = [0.3, -0.1, 1.0, -1.0, 0.5, 0.1]
    parts.id[:] = np.arange(6)
    o = O.Oracle(m, p, seed=1)
    o.set_particles(parts)
    unit = p.Neff / m.arrays()["cell_volume"] / p.n_ref
    o.step(1, D.STAGE_SAMPLE)
    sc = o.cell_scalars()
    # cell 0: 2 samples < 3 -> accumulating; cell 1: 4 samples; empty cells: 0 < 3
    assert list(sc["sparse_state"]) == [1, 0, 1, 1, 1, 1, 1, 1]
    rec = o.macro(D.MACRO_RECORD)
    # 2 samples over max(1, Nrepeat) = 1 step; u = 0.1; T = 2/3 <c^2> = 2/3 * 0.04
    assert rec[0, 0] == pytest.approx(2 * unit[0]) and rec[0, 1] == pytest.approx(0.1)
    assert rec[0, 4] == pytest.approx(2 / 3 * 0.04)
    assert rec[1, 0] == pytest.approx(4 * unit[1]) and rec[1, 1] == pytest.approx(0.6 / 4)
    assert np.all(rec[2:] == 0.0)  # empty cells are invalid -> zeros
    acc = o.accum(2)
    assert acc[0, 0] == 2 and acc[1, 0] == 0  # sparse sums kept, normal reset
    o.step(2, D.STAGE_SAMPLE)
    assert o.cell_scalars()["sparse_steps"][0] == 2 and o.accum(2)[0, 0] == 4
    rec = o.macro(D.MACRO_RECORD)
    assert rec[0, 0] == pytest.approx(4 / 2 * unit[0]) and rec[0, 1] == pytest.approx(0.1)
    # out2dat shows record for sparse cells and before NSS
    out = o.macro(D.MACRO_OUTPUT)
    assert np.array_equal(out, rec)


# ---- the C ABI ------------------------------------------------------------------------------
def test_library_exports_every_declared_symbol():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "dsmc_engine.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = sorted(set(re.findall(r"\b(dsmc_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 30
    L = ctypes.CDLL(D.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(D.EXPORTED)


def test_params_and_mesh_errors_are_reported():
    with pytest.raises(D.DsmcError):
        D.Mesh.box_tets(0)
    m = D.Mesh.box_tets(2)
    p = D.params_box(0.05)
    p.Kn = -1.0
    with pytest.raises(D.DsmcError):
        D.derive(m, p)


def _ns_case(kind):
    """A small case after two steps, and NS fields for it: density scaled by
    0.6-1.5 per cell, random drift, T and T_r in [0.5, 2], 80% of cells
    coupled, a NaN density in one coupled cell at 70% of the cells."""
    if kind == "capsule":
        m = D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912)
        p = D.params_apollo(monatomic=False)  # diatomic: T_r > 0, so the shift applies
        p.Npinitial = 12
        D.derive(m, p)
        parts = D.fill_freestream(m, p, ppc=12.0, seed=4)
    else:
        m, p, parts = _box(n=3, ppc=20.0, T=1.0, seed=6)
    o = O.Oracle(m, p, seed=13)
    o.set_particles(parts)
    o.step(1)
    o.step(2)
    rng = np.random.default_rng(17)
    nc = m.ncell
    vol = m.arrays()["cell_volume"]
    cnt = np.bincount(o.particles().cell, minlength=nc).astype(np.float64)
    ns = np.empty((nc, 6))
    ns[:, 0] = cnt * p.Neff / (vol * p.n_ref) * rng.uniform(0.6, 1.5, nc)
    ns[:, 1:4] = rng.normal(0.0, 0.5, (nc, 3))
    ns[:, 4] = rng.uniform(0.5, 2.0, nc)
    ns[:, 5] = rng.uniform(0.5, 2.0, nc)
    acc = (rng.uniform(size=nc) < 0.8).astype(np.uint8)
    stop = int(np.flatnonzero(acc)[int(0.7 * acc.sum())])
    ns[stop, 0] = np.nan
    return m, p, o, ns, acc, stop


def _by_cell(q):
    order = np.lexsort((np.arange(q.n), q.cell))
    return {f: getattr(q, f)[order] for f in ("cell", "id", "x", "y", "z", "vx", "vy", "vz", "ir")}


@pytest.mark.parametrize("kind", ["capsule", "box"])
def test_ns_apply_counts_moments_and_untouched_cells(kind):
    """Continuum coupling on the oracle (molecular_velocity_change,
    src/ProcessGSIS.cpp:2239-2345): a coupled cell before the NaN cell ends
    with Iround(rho V n_ref / Neff) particles and, where the shift applies,
    exactly the NS mean velocity, temperature and rotational temperature;
    every other cell keeps its particles bit for bit; new particles get fresh
    ids and points inside their cell's bounding box."""
    m, p, o, ns, acc, stop = _ns_case(kind)
    before = _by_cell(o.particles())
    maxid = int(before["id"].max())
    assert o.ns_apply(3, ns, acc) == stop
    after = _by_cell(o.particles())
    nc = m.ncell
    a = m.arrays()
    vol = a["cell_volume"]
    nb = np.bincount(before["cell"], minlength=nc)
    na = np.bincount(after["cell"], minlength=nc)
    sb = np.concatenate([[0], np.cumsum(nb)])
    sa = np.concatenate([[0], np.cumsum(na)])
    fn, fc = a["face_nodes"], a["face_cells"]
    checked = 0
    for c in range(nc):
        B = {f: v[sb[c]:sb[c + 1]] for f, v in before.items()}
        A = {f: v[sa[c]:sa[c + 1]] for f, v in after.items()}
        if c >= stop or not acc[c]:
            for f in B:
                assert np.array_equal(A[f], B[f]), (c, f)
            continue
        x = ns[c, 0] * vol[c] * p.n_ref / p.Neff
        assert na[c] in (math.floor(x), math.floor(x) + 1), (c, na[c], x)
        new = A["id"] > maxid
        assert new.sum() == max(0, na[c] - nb[c])
        old_ids = set(B["id"].tolist())
        assert all(i in old_ids for i in A["id"][~new])
        if new.any():
            faces = np.flatnonzero((fc[:, 0] == c) | (fc[:, 1] == c))
            nodes = a["node_xyz"][np.unique(fn[faces][fn[faces] >= 0])]
            lo, hi = nodes.min(0) - 1e-12, nodes.max(0) + 1e-12
            pts = np.stack([A["x"][new], A["y"][new], A["z"][new]], 1)
            assert np.all(pts >= lo) and np.all(pts <= hi), c
        if p.dr == 0.0:
            # monatomic: T_r of the cell is not positive, so no shift
            # (src/ProcessGSIS.cpp:2304-2320): kept particles keep velocities
            vb = dict(zip(B["id"].tolist(), B["vx"].tolist()))
            assert all(vb[i] == vx for i, vx in zip(A["id"][~new].tolist(), A["vx"][~new].tolist()))
            checked += 1
        elif na[c] >= 2:
            v = np.stack([A["vx"], A["vy"], A["vz"]], 1)
            u = v.mean(0)
            T = 2.0 / (3.0 * na[c]) * (np.sum(v * v) - na[c] * np.dot(u, u))
            Tr = 4.0 / p.dr / na[c] * A["ir"].sum()
            if T > 0 and Tr > 0:
                assert np.allclose(u, ns[c, 1:4], rtol=0, atol=1e-12), (c, u, ns[c, 1:4])
                assert abs(T - ns[c, 4]) < 1e-11 * ns[c, 4], (c, T, ns[c, 4])
                assert abs(Tr - ns[c, 5]) < 1e-11 * ns[c, 5], (c, Tr, ns[c, 5])
                checked += 1
    assert checked > 10
    ids = after["id"]
    assert len(np.unique(ids)) == len(ids)


def test_ns_apply_refills_an_empty_cell_with_the_ns_maxwellian():
    """A coupled cell emptied by a zero NS density is refilled from the NS
    state (MaxwellianSample, src/ProcessGSIS.cpp:2424-2446): with many
    particles its drift, temperature and rotational temperature follow the NS
    values statistically before the exact shift, and exactly after it."""
    m, p, parts = _box(n=2, ppc=20.0, T=1.0, seed=8)
    p2 = p.copy()
    p2.dr = 2.0  # rotational energy present, so the shift applies
    o = O.Oracle(m, p2, seed=3)
    o.set_particles(parts)
    nc = m.ncell
    vol = m.arrays()["cell_volume"]
    ns = np.zeros((nc, 6))
    ns[:, 4] = ns[:, 5] = 1.0
    acc = np.zeros(nc, np.uint8)
    acc[0] = 1
    o.ns_apply(1, ns, acc)  # cell 0 emptied
    assert np.sum(o.particles().cell == 0) == 0
    ns[0] = [4000.0 * p.Neff / (vol[0] * p.n_ref), 0.3, -0.2, 0.1, 1.5, 0.8]
    o.ns_apply(2, ns, acc)
    q = o.particles()
    sel = q.cell == 0
    assert sel.sum() in (3999, 4000, 4001)
    v = np.stack([q.vx[sel], q.vy[sel], q.vz[sel]], 1)
    n = sel.sum()
    u = v.mean(0)
    T = 2.0 / (3.0 * n) * (np.sum(v * v) - n * np.dot(u, u))
    assert np.allclose(u, [0.3, -0.2, 0.1], atol=1e-12) and abs(T - 1.5) < 1e-11
    assert abs(4.0 / p2.dr / n * q.ir[sel].sum() - 0.8) < 1e-11


def test_oracle_threads_do_not_change_results():
    """The oracle's OpenMP paths (advection over cell ranges, the parallel
    cache2chain, collisions over cells) give the single-thread state bit for
    bit: recv_cache order and every bucket's order are kept.  The box has
    enough movers per step (~700k) for the parallel cache2chain."""
    mesh = D.Mesh.box_tets(16)
    p = D.params_box(0.05)
    p.Npinitial = 50
    D.derive(mesh, p)
    parts = D.fill_freestream(mesh, p, ppc=50.0, temperature=1.0, seed=5)
    out = []
    try:
        for nt in (1, 4):
            O.set_threads(nt)
            o = O.Oracle(mesh, p, seed=5)
            o.set_particles(parts)
            for s in (1, 2, 3):
                o.step(s)
            out.append((o.particles(), [o.accum(f) for f in (0, 1, 2)], o.stats()))
    finally:
        O.set_threads(1)
    (a, aa, sa), (b, ab, sb) = out
    for f in ("x", "y", "z", "vx", "vy", "vz", "ir"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(a.cell, b.cell) and np.array_equal(a.id, b.id)
    for fa, fb in zip(aa, ab):
        assert np.array_equal(fa, fb)
    assert sa["crossings"] == sb["crossings"] > 500000


def test_staged_step_counts_particle_steps_once():
    """A step run stage by stage (one call per reference driver stage, as the
    CPU baseline times it) counts its particles once, at the step's start,
    like a full step; a repeated or earlier stage opens a new step."""
    m = D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912)
    p = D.params_apollo(monatomic=True)
    p.Npinitial = 6
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=6.0, seed=3)
    full, staged = O.Oracle(m, p, seed=4), O.Oracle(m, p, seed=4)
    full.set_particles(parts)
    staged.set_particles(parts)
    stages = (D.STAGE_EMIT, D.STAGE_MOVE, D.STAGE_SORT, D.STAGE_COLLIDE, D.STAGE_SAMPLE)
    expect = 0
    for s in (1, 2, 3):
        expect += staged.stats()["particles"]
        full.step(s)
        for mask in stages:
            staged.step(s, mask)
        a, b = staged.stats(), full.stats()
        assert a["particle_steps"] == b["particle_steps"] == expect, (s, a, b)
        assert a["emitted"] == b["emitted"] and a["particles"] == b["particles"]
    # the same stage again is a new step, counted again
    c1 = staged.stats()["particles"]
    staged.step(4, D.STAGE_EMIT)
    c2 = staged.stats()["particles"]
    staged.step(4, D.STAGE_EMIT)
    assert staged.stats()["particle_steps"] == expect + c1 + c2
