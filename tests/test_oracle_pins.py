"""Closed-form pins for the oracle functions that bit-exact parity alone
cannot vouch for (engine and oracle share them), on the CPU.

Each check compares the oracle (or the derived case parameters both sides
use) with a closed form taken from kinetic theory, not with the formula the
code evaluates, and each is shown to fail under a plausible mistake: the
mistake is planted in the oracle (oracle_plant) or in a copy of the
parameters, and the same check must then reject it.

* VHS collision rate (src/ProcessDSMC.cpp:2583-2607,2626-2630,2679-2690):
  <sigma c_r> = 4 d_ref^2 sqrt(pi k T_ref / m) (T/T_ref)^(1-omega)
  (Bird 1994, eq. 4.64), at T = T_ref and T != T_ref.
* Larsen-Borgnakke exchange (:2700-2712): relaxation of a diatomic gas
  started with no rotational energy, against the mean-field solution of the
  serial two-partner exchange, and equipartition T_r = T_t at the end.
* Inflow velocities (sampleBoundaryParticle, :2416-2458): moments of freshly
  emitted particles against the flux-weighted drifting Maxwellian.
* Normalisation (setMeshMessage / setParticleMessage,
  src/meshImport.cpp:1315-1402): n_ref from the hard-sphere mean free path,
  Neff, dtime and the VHS prefactor.
* Wall fields (outBoundaryStressHeat, :3494-3557): the wall pressure of an
  equilibrium gas is n k T, and specular walls pass no heat.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_py as O
import pydsmc as D

KB = 1.3806e-23


# ---------------------------------------------------------------------------
# VHS collision rate
# ---------------------------------------------------------------------------
def _vhs_rate_ratio(T, plant=0, omega=0.74, n=4, ppc=40, warm=2, steps=8):
    """Accepted collisions over `steps` collide-only steps of an equilibrium
    box at temperature T (units of T_ref), divided by the closed-form count
    sum_cells 1/2 N (N-1)/V Neff <sigma c_r> dt.  Returns (ratio, relative
    Poisson sigma).  Two warm-up steps let each cell's (sigma c_r)max climb
    past the initial cap of 3 v_rms."""
    m = D.Mesh.box_tets(n)
    p = D.params_box(0.05)
    p.omega = omega
    p.Npinitial = ppc
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=ppc, temperature=T, seed=5)
    with O.planted(plant):
        o = O.Oracle(m, p, seed=9)
    o.set_particles(parts)
    for s in range(1, warm + 1):
        o.step(s, D.STAGE_COLLIDE)
    c0 = o.stats()["collisions"]
    for s in range(warm + 1, warm + steps + 1):
        o.step(s, D.STAGE_COLLIDE)
    got = o.stats()["collisions"] - c0
    counts = np.bincount(parts.cell, minlength=m.ncell).astype(float)
    vol = m.arrays()["cell_volume"]
    sigma_c = 4.0 * p.d_ref ** 2 * math.sqrt(math.pi * KB * p.T_ref / p.mass) * T ** (1.0 - omega)
    want = steps * float(np.sum(0.5 * counts * (counts - 1.0) / vol)) * p.Neff * sigma_c * p.dtime * p.dt_ref
    return got / want, 1.0 / math.sqrt(want)


def _rate_ok(ratio, sig):
    return abs(ratio - 1.0) < 4.0 * sig + 0.005


@pytest.mark.parametrize("T", [1.0, 0.5, 2.0])
def test_vhs_collision_rate_matches_bird(T):
    r, sig = _vhs_rate_ratio(T)
    print(f"[vhs] T={T} ratio {r:.4f} (sigma {sig:.4f})")
    assert _rate_ok(r, sig), (T, r, sig)


def test_vhs_rate_check_catches_the_molecular_mass():
    # (2 k T_ref / m)^eta instead of (2 k T_ref / m_r)^eta: 2^-eta = 0.85
    for T in (1.0, 2.0):
        r, sig = _vhs_rate_ratio(T, plant=O.PLANT_VHS_MASS)
        assert not _rate_ok(r, sig), (T, r)


def test_vhs_rate_check_needs_a_temperature_off_the_reference():
    # a constant hard-sphere cross-section pi d_ref^2 has exactly the VHS
    # rate at T = T_ref (eq. 4.64 with T = T_ref), so only T != T_ref can
    # tell the speed power c_r^(1 - 2 eta) from c_r
    r, sig = _vhs_rate_ratio(1.0, plant=O.PLANT_VHS_HS_POWER)
    assert _rate_ok(r, sig), r
    for T in (0.5, 2.0):
        r, sig = _vhs_rate_ratio(T, plant=O.PLANT_VHS_HS_POWER)
        assert not _rate_ok(r, sig), (T, r)


# ---------------------------------------------------------------------------
# Larsen-Borgnakke rotational exchange
# ---------------------------------------------------------------------------
def _lb_relaxation(plant=0, omega=0.74, P=0.18, n=5, ppc=60, steps=60):
    """Adiabatic box (collisions only) of a diatomic VHS gas at T_t = 1 with
    every Ir = 0.  Returns per step: collisions per particle so far, mean Ir
    over the initial translational energy per particle E0, and T_r/T_t."""
    m = D.Mesh.box_tets(n)
    p = D.params_box(0.05)
    p.omega, p.P_relax, p.dr, p.Npinitial = omega, P, 2.0, ppc
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=ppc, temperature=1.0, seed=3)
    parts.ir[:] = 0.0
    with O.planted(plant):
        o = O.Oracle(m, p, seed=4)
    o.set_particles(parts)
    N = parts.n
    E0 = 0.5 * float(np.sum(parts.vx ** 2 + parts.vy ** 2 + parts.vz ** 2)) / N
    rows = []
    for s in range(1, steps + 1):
        o.step(s, D.STAGE_COLLIDE)
        q = o.particles()
        et = 0.5 * float(np.mean(q.vx ** 2 + q.vy ** 2 + q.vz ** 2))
        er = float(q.ir.mean())
        # kT_t = (2/3) e_t (3 translational DOF), kT_r = e_r (2 rotational DOF)
        rows.append((o.stats()["collisions"] / N, er / E0, er / (2.0 * et / 3.0)))
    return rows


def _lb_prediction(c, omega=0.74, P=0.18):
    """Mean-field solution of the serial exchange.  A partner offered the
    exchange (probability P) gets eps' = (E_t + eps) x with 1 - x ~
    Beta(a, 1), a = 5/2 - omega, so E[x] = 1/(a + 1); the pair's E_t loses
    what the first partner gained before the second is offered.  With
    <E_t> = a kT_t over accepted VHS collisions (Bird 1994, eq. 4.80) and
    energy conservation e_t + e_r = E0:
      de_r/dc = K (2/3 E0 - 5/3 e_r),  K = P a (2 - P/(a+1)) / (a+1),
    c = accepted collisions per particle, so e_r/E0 = 2/5 (1 - exp(-5 K c/3))
    and the end state is equipartition, e_r = kT_t."""
    a = 2.5 - omega
    K = P * a * (2.0 - P / (a + 1.0)) / (a + 1.0)
    return 0.4 * (1.0 - math.exp(-5.0 * K * c / 3.0))


def _lb_ok(rows):
    curve = max(abs(er / _lb_prediction(c) - 1.0) for c, er, _ in rows[:30])
    final = abs(rows[-1][2] - 1.0)
    return curve < 0.03 and final < 0.025, curve, final


def test_larsen_borgnakke_relaxes_to_equipartition():
    ok, curve, final = _lb_ok(_lb_relaxation())
    print(f"[lb] worst curve deviation {curve:.4f}, final |Tr/Tt - 1| {final:.4f}")
    assert ok, (curve, final)


@pytest.mark.parametrize("plant", [O.PLANT_LB_ONE_PARTNER, O.PLANT_LB_EXPONENT])
def test_larsen_borgnakke_check_catches_planted_mistakes(plant):
    ok, curve, final = _lb_ok(_lb_relaxation(plant=plant))
    assert not ok, (plant, curve, final)


# ---------------------------------------------------------------------------
# Inflow velocity distribution
# ---------------------------------------------------------------------------
def _inlet_box(n=4):
    """Unit cube of n^3 hexes: the x = 0 face is the inlet, the other five
    are outlets."""
    b = D.Mesh.box_hexes(n)
    a = b.arrays()
    fc, xyz, fn = a["face_cells"], a["node_xyz"], a["face_nodes"]
    bnd = (fc[:, 1] < 0) | (fc[:, 1] >= b.ncell)
    cx = xyz[fn[:, :4], 0].mean(axis=1)
    tag = np.where(bnd, np.where(cx < 1e-12, D.TAG_INLET, D.TAG_OUTLET), D.TAG_INTERFACE)
    return D.Mesh.build(xyz, fn, fc, a["face_type"], tag, b.ncell)


def _emitted(plant=0):
    """One emission step of the Apollo free stream (diatomic) into an empty
    inlet box: ~2e5 particles.  A short dtime keeps the inflow trace from
    reaching the side outlets (under 0.1% leave)."""
    m = _inlet_box()
    p = D.params_apollo(monatomic=False)
    p.Npinitial = 1_200_000
    D.derive(m, p)
    p.dtime = 1e-3
    with O.planted(plant):
        o = O.Oracle(m, p, seed=21)
    o.set_particles(D.Particles(0))
    o.step(1, D.STAGE_EMIT | D.STAGE_SORT)
    return o.particles(), p


def _flux_moments(un, T):
    """<v_n>, <v_n^2> of the flux-weighted drifting Maxwellian
    f(v_n) ~ v_n exp(-(v_n - u_n)^2 / T), v_n > 0 (v_rms units: each
    component has variance T/2), by quadrature."""
    def f(v, k):
        return v ** (k + 1) * math.exp(-(v - un) ** 2 / T)
    hi = max(un, 0.0) + 12.0 * math.sqrt(T)
    z = integrate.quad(f, 0.0, hi, args=(0,), epsabs=0, epsrel=1e-12)[0]
    return [integrate.quad(f, 0.0, hi, args=(k,), epsabs=0, epsrel=1e-12)[0] / z for k in (1, 2)]


def _inflow_z(q, p):
    """z-scores of the emitted moments against the closed forms.  The x = 0
    inlet's inward normal is +x, so v_n = vx and (vy, vz) are tangential:
    Gaussian with mean (u_y, 0) and variance T/2; Ir ~ Exp(T/2)."""
    T = p.T_inf
    ux, uy = p.u_inf[0], p.u_inf[1]
    m1, m2 = _flux_moments(ux, T)

    def z(x, mu):
        return (float(x.mean()) - mu) / (float(x.std()) / math.sqrt(len(x)))
    return {"vn": z(q.vx, m1), "vn2": z(q.vx ** 2, m2), "vy": z(q.vy, uy), "vy2": z((q.vy - uy) ** 2, T / 2),
            "vz": z(q.vz, 0.0), "vz2": z(q.vz ** 2, T / 2), "ir": z(q.ir, T / 2)}


def test_inflow_velocities_are_the_flux_weighted_maxwellian():
    q, p = _emitted()
    assert q.n > 150_000
    zs = _inflow_z(q, p)
    print("[inflow] z " + " ".join(f"{k}={v:+.2f}" for k, v in zs.items()))
    assert max(abs(v) for v in zs.values()) < 4.5, zs


@pytest.mark.parametrize("plant,moment", [(O.PLANT_INFLOW_NO_FLUX_WEIGHT, "vn"),
                                          (O.PLANT_INFLOW_TANGENT_SCALE, "vy2")])
def test_inflow_check_catches_planted_mistakes(plant, moment):
    q, p = _emitted(plant)
    zs = _inflow_z(q, p)
    assert abs(zs[moment]) > 20.0, zs


# ---------------------------------------------------------------------------
# Case normalisation
# ---------------------------------------------------------------------------
def _normalisation_errors(p, n):
    """Relative errors of the derived quantities of the hard-sphere box of
    n^3 hexes x 6 tets against closed forms that do not use the derivation's
    own formulas:
      n_ref  = (5 pi/16) / (sqrt(2) pi d^2 Kn L0): the viscosity-based mean
               free path the reference uses is 5 pi/16 of the hard-sphere
               one 1/(sqrt(2) pi d^2 n) (Bird 1994, eqs. 4.52, 4.65);
      Neff   = n_ref V / (Npinitial N_cell) with V = 1 and N_cell = 6 n^3:
               the initial fill then represents n_ref;
      dtime  = 0.2 / n: the shortest edge of the Freudenthal split is the hex
               edge 1/n;
      prefactor: <sigma c_r> = pref <c_r^(1 - 2 eta)> over the relative-speed
               Maxwellian at T_ref equals 4 d^2 sqrt(pi k T_ref / m)
               (Bird eq. 4.64), by quadrature."""
    n_ref = (5.0 * math.pi / 16.0) / (math.sqrt(2.0) * math.pi * p.d_ref ** 2 * p.Kn * p.L0)
    neff = n_ref * 1.0 / (p.Npinitial * 6 * n ** 3)
    mr = 0.5 * p.mass
    b2 = mr / (2.0 * KB * p.T_ref)  # relative speeds: f(c) ~ c^2 exp(-b2 c^2)
    scale = 1.0 / math.sqrt(b2)
    w = lambda x, k: x ** (2 + k) * math.exp(-x * x)  # noqa: E731  (c = x * scale)
    z = integrate.quad(w, 0, 40, args=(0,), epsabs=0, epsrel=1e-13)[0]
    pw = 1.0 - 2.0 * p.eta
    mean_pow = integrate.quad(w, 0, 40, args=(pw,), epsabs=0, epsrel=1e-13)[0] / z * scale ** pw
    sc = 4.0 * p.d_ref ** 2 * math.sqrt(math.pi * KB * p.T_ref / p.mass)
    return {"n_ref": p.n_ref / n_ref - 1.0, "Neff": p.Neff / neff - 1.0, "dtime": p.dtime / (0.2 / n) - 1.0,
            "prefactor": p.coll_prefactor * mean_pow / sc - 1.0}


def _box_params(n, omega=0.5, ppc=37):
    m = D.Mesh.box_tets(n)
    p = D.params_box(0.05)
    p.Npinitial = ppc
    p.omega = omega
    D.derive(m, p)
    return m, p


@pytest.mark.parametrize("n,omega", [(3, 0.5), (5, 0.5), (4, 0.74), (4, 0.9)])
def test_normalisation_closed_forms(n, omega):
    m, p = _box_params(n, omega)
    err = _normalisation_errors(p, n)
    print(f"[normalisation] n={n} omega={omega} " + " ".join(f"{k}={v:.2e}" for k, v in err.items()))
    tol = {"n_ref": 1e-12, "Neff": 1e-12, "dtime": 1e-12, "prefactor": 1e-9}
    if omega != 0.5:
        # n_ref's closed form is the hard-sphere one; for VHS the viscosity
        # law changes it, which the Apollo value pins (test_oracle.py)
        err.pop("n_ref")
        err.pop("Neff")
    for k, v in err.items():
        assert abs(v) < tol[k], (k, v)


def test_normalisation_min_edge_on_a_curved_mesh():
    """dtime = 0.2 x the shortest face edge (setParticleMessage,
    src/meshImport.cpp:1396-1402), against a brute-force scan of every edge
    of a jittered capsule tet mesh."""
    m = D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912, tets=True, jitter=0.2, seed=3)
    p = D.derive(m, D.params_apollo(monatomic=True))
    a = m.arrays()
    xyz, fn, ft = a["node_xyz"], a["face_nodes"], a["face_type"]
    best = math.inf
    for k in range(4):
        nxt = np.where(ft == 4, (k + 1) % 4, np.where(k + 1 < 3, k + 1, 0))
        valid = k < ft
        if valid.any():
            d = np.linalg.norm(xyz[fn[valid, k]] - xyz[fn[valid, nxt[valid]]], axis=1)
            best = min(best, float(d.min()))
    assert p.min_edge == best
    assert abs(p.dtime / (0.2 * best) - 1.0) < 1e-15
    vol = float(a["cell_volume"].sum())
    assert abs(p.Neff * p.Npinitial * m.ncell / (vol * p.n_ref) - 1.0) < 1e-12


@pytest.mark.parametrize("field,mistake", [
    ("Neff", lambda p, n: p.n_ref / p.Npinitial),                      # cell count left out
    ("dtime", lambda p, n: 0.2 * math.sqrt(2.0) / n),                   # a face diagonal, not the shortest edge
    ("coll_prefactor", lambda p, n: p.coll_prefactor * 2.0 ** -p.eta),  # molecular mass for the reduced mass
    ("n_ref", lambda p, n: p.n_ref * 16.0 / (5.0 * math.pi)),           # hard-sphere mean free path
])
def test_normalisation_check_catches_mistakes(field, mistake):
    n = 4
    m, p = _box_params(n, omega=0.5 if field != "coll_prefactor" else 0.74)
    q = p.copy()
    setattr(q, field, mistake(q, n))
    err = _normalisation_errors(q, n)
    key = {"coll_prefactor": "prefactor"}.get(field, field)
    assert abs(err[key]) > 1e-3, (field, err)


# ---------------------------------------------------------------------------
# Wall fields (outBoundaryStressHeat)
# ---------------------------------------------------------------------------
def _box_wall_run(model, T, seed, steps=60):
    """Equilibrium box with steady wall tallies from step 1; returns the
    oracle, its parameters, the gas's density over n_ref and its sample
    temperature (units of T_ref)."""
    m = D.Mesh.box_tets(4)
    p = D.params_box(0.05)
    p.Npinitial = 40
    p.wall_model = model
    p.T_wall = 273.15 * T
    D.derive(m, p)
    p.surface_stats, p.NSS = 1, 0
    parts = D.fill_freestream(m, p, ppc=40, temperature=T, seed=seed)
    c = [parts.vx - parts.vx.mean(), parts.vy - parts.vy.mean(), parts.vz - parts.vz.mean()]
    t_sample = 2.0 / 3.0 * float(np.mean(c[0] ** 2 + c[1] ** 2 + c[2] ** 2))
    o = O.Oracle(m, p, seed=5)
    o.set_particles(parts)
    for s in range(1, steps + 1):
        o.step(s)
    return m, p, o, parts.n / (p.Npinitial * m.ncell), t_sample


def _wall_means(m, f):
    w = np.linalg.norm(f[:, 5:8], axis=1) > 0.5
    area = m.arrays()["face_area"][w]
    return (float(np.sum(f[w, 8] * area) / area.sum()), float(np.sum((f[w, 3] + f[w, 4]) * area) / area.sum()))


@pytest.mark.parametrize("model,T,seed", [(D.WALL_SPECULAR, 1.0, 3), (D.WALL_SPECULAR, 0.6, 5),
                                          (D.WALL_DIFFUSE, 1.0, 3)])
def test_wall_pressure_is_the_gas_pressure(model, T, seed):
    """Kinetic pressure on the walls of a gas at rest in equilibrium is
    p = n k T: in the wall output's unit n_ref k T_ref, rho T.  With specular
    walls the gas keeps its sample temperature and no energy crosses the
    wall (q = 0 exactly); a diffuse wall at T_w = T brings the gas to T_w."""
    steps = 60
    m, p, o, rho, t_sample = _box_wall_run(model, T, seed, steps)
    pw, q = _wall_means(m, o.surface_fields(steps))
    t_gas = t_sample if model == D.WALL_SPECULAR else T
    print(f"[wall] model={model} T={T} p/(rho T)={pw / (rho * t_gas):.4f} q={q:.2e}")
    assert abs(pw / (rho * t_gas) - 1.0) < 0.01
    if model == D.WALL_SPECULAR:
        assert q == 0.0
    else:
        assert abs(q) < 0.01 * pw
    # planted: the same tallies turned into fields with Neff halved (or any
    # wrong time, area or density scale) miss the gas pressure
    q2 = p.copy()
    q2.Neff *= 0.5
    bad = O.Oracle(m, q2, seed=5)
    pw_bad, _ = _wall_means(m, bad.surface_fields(steps, o.surface()))
    assert abs(pw_bad / (rho * t_gas) - 1.0) > 0.2


def _tet_nodes(a, c):
    """The four nodes of tet cell c (from its faces), in a fixed order."""
    nodes = []
    for f in a["cell_faces"][c][: a["cell_nface"][c]]:
        for v in a["face_nodes"][f]:
            if v >= 0 and v not in nodes:
                nodes.append(int(v))
    assert len(nodes) == 4
    return a["node_xyz"][nodes]


def test_initial_fill_closed_forms():
    """The initial free-stream fill (src/MeshparticalInitial.cpp:392-721 with
    the per-cell count rounding of src/meshImport.cpp:1409-1421), product
    host code that the device fill reproduces bit for bit, against what
    the sampling law fixes:
    * each cell gets floor(N_c) or floor(N_c) + 1 particles, N_c = its volume
      share of ppc x ncell, and the round-ups add up to the fractional parts
      (a plain floor, or a share of the cell count instead of the volume,
      fails);
    * every particle lies in its tet, and its barycentric coordinates have
      the uniform (Dirichlet(1,1,1,1)) mean 1/4 and variance 3/80;
    * velocities are the drifting Maxwellian: mean u_inf, variance T/2 per
      component (sqrt(T) for sqrt(T/2) fails by a factor 2), rotational
      energy exponential with mean T/2 for the diatomic gas."""
    m = D.Mesh.box_tets(5)
    # jitter-free box tets have equal volumes; use the capsule for the counts
    cap = D.Mesh.capsule(a=6, c=12, layers=8, diameter=3.912)
    p = D.derive(cap, D.params_apollo(monatomic=False))
    ppc = 30.0
    parts = D.fill_freestream(cap, p, ppc=ppc, seed=11)
    ac = cap.arrays()
    vol = ac["cell_volume"]
    ns = vol * ppc * cap.ncell / vol.sum()
    cnt = np.bincount(parts.cell, minlength=cap.ncell)
    up = cnt - np.floor(ns)
    assert set(np.unique(up)) <= {0.0, 1.0}
    frac = ns - np.floor(ns)
    sd = math.sqrt(float(np.sum(frac * (1 - frac))))
    print(f"[fill] {parts.n} particles; round-ups {up.sum():.0f} against {frac.sum():.1f} +- {sd:.1f}")
    assert abs(up.sum() - frac.sum()) < 5 * sd + 1
    # velocities: drifting Maxwellian and exponential rotational energy
    T = p.T_inf
    n = parts.n
    for comp, u in zip(("vx", "vy", "vz"), p.u_inf):
        v = getattr(parts, comp) - u
        assert abs(v.mean()) < 5 * math.sqrt(T / 2 / n), comp
        assert abs(v.var() / (T / 2) - 1) < 5 * math.sqrt(2 / n), comp
    ir = parts.ir
    assert ir.min() >= 0.0
    assert abs(ir.mean() / (T / 2) - 1) < 5 / math.sqrt(n)
    assert abs(ir.var() / (T / 2) ** 2 - 1) < 10 / math.sqrt(n)
    # positions: inside the tet, uniformly
    pb = D.derive(m, D.params_box(0.05))
    box = D.fill_freestream(m, pb, ppc=200.0, temperature=1.0, seed=4)
    a = m.arrays()
    lam = np.empty((box.n, 4))
    xyz = np.stack([box.x, box.y, box.z], axis=1)
    for c in range(m.ncell):
        sel = np.flatnonzero(box.cell == c)
        v = _tet_nodes(a, c)
        mat = np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])
        l123 = np.linalg.solve(mat, (xyz[sel] - v[0]).T).T
        lam[sel, 1:] = l123
        lam[sel, 0] = 1.0 - l123.sum(axis=1)
    assert lam.min() > -1e-12
    nb = box.n
    assert np.all(np.abs(lam.mean(axis=0) - 0.25) < 5 * math.sqrt(3 / 80 / nb))
    assert np.all(np.abs(lam.var(axis=0) / (3 / 80) - 1) < 0.05)
