"""Physics of the HIP engine at larger sizes than the oracle parity tests:
the adiabatic equilibrium box (energy conservation, hard-sphere collision
rate, Maxwellian kept) in FP64 and in the mixed-precision form, and the mixed
form's fields against the FP64 engine's."""
import math

import numpy as np
import pytest

import pydsmc as D

pytestmark = pytest.mark.gpu

KB = 1.3806e-23


def _box(n, ppc, prec, seed=5):
    m = D.Mesh.box_tets(n)
    p = D.params_box(0.05)
    p.Npinitial = int(ppc)
    p.NSS = 10
    p.precision = prec
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=ppc, temperature=1.0, seed=seed)
    return m, p, parts


def _energy(q):
    return 0.5 * float(np.sum(q.vx ** 2 + q.vy ** 2 + q.vz ** 2)) + float(np.sum(q.ir))


@pytest.mark.parametrize("prec,etol", [(D.PREC_FP64, 1e-11), (D.PREC_MIXED, 1e-6)])
def test_equilibrium_box(prec, etol):
    """Specular box of 16^3 x 6 tets, 40 particles per cell, hard spheres at
    T_ref: particle count and total energy conserved, collisions at the
    hard-sphere rate nu = n sigma <c_r>, <c_r> = 4 sqrt(kT/(pi m)) (Bird 1994,
    eqs. 1.10 and 4.64; per cell and step 1/2 N (N-1)/V Neff sigma <c_r> dt),
    and the velocity distribution stays Maxwellian (kurtosis 3)."""
    m, p, parts = _box(16, 40.0, prec)
    print(f"[equilibrium_box] prec={prec} cells={m.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(m, p, seed=17)
    eng.upload(parts)
    E0 = _energy(eng.download())
    vol = m.arrays()["cell_volume"]
    sigma = math.pi * p.d_ref ** 2
    cr = 4.0 * math.sqrt(KB * p.T_ref / (math.pi * p.mass))
    dt = p.dtime * p.dt_ref
    want = 0.0
    steps = 60
    for s in range(1, steps + 1):
        eng.step(s)
        cnt = eng.cell_scalars()["count"].astype(np.float64)
        want += np.sum(0.5 * cnt * (cnt - 1.0) / vol) * p.Neff * sigma * cr * dt
    st = eng.stats()
    q = eng.download()
    assert st["particles"] == parts.n and st["dropped"] == 0 and st["wall_hits"] > 0
    E1 = _energy(q)
    assert abs(E1 - E0) / E0 < etol, (E0, E1)
    got = st["collisions"]
    assert abs(got - want) < 4.0 * math.sqrt(want) + 0.01 * want, (got, want)
    for v in (q.vx, q.vy, q.vz):
        k = np.mean(v ** 4) / np.mean(v ** 2) ** 2
        assert abs(k - 3.0) < 0.03, k
        assert abs(np.mean(v ** 2) - 0.5) < 0.01


def test_mixed_matches_fp64_in_equilibrium():
    """Domain means of the steady-averaged fields (density, temperature,
    velocity) of the mixed form agree with the FP64 engine's."""
    out = {}
    for prec in (D.PREC_FP64, D.PREC_MIXED):
        m, p, parts = _box(12, 40.0, prec)
        eng = D.Engine(m, p, seed=23)
        eng.upload(parts)
        eng.run(1, 60)
        out[prec] = eng.macro(D.MACRO_FINAL)
    a, b = out[D.PREC_FP64], out[D.PREC_MIXED]
    for f, tol in ((0, 2e-3), (4, 2e-3)):
        ma, mb = a[:, f].mean(), b[:, f].mean()
        assert abs(ma - 1.0) < 0.01 and abs(mb - ma) < tol, (f, ma, mb)
    for f in (1, 2, 3):
        assert abs(a[:, f].mean()) < 5e-3 and abs(b[:, f].mean()) < 5e-3


@pytest.mark.parametrize("tets", [False, True])
def test_mixed_tracking_keeps_particles(tets):
    """The mixed engine's exit search (FP32 plane pre-selection confirmed by
    the exact test) on a capsule grid with warped hex faces, and on jittered
    tets: no particle lost over 30 steps, and the particle, crossing and wall
    hit counts track the FP64 engine's statistically (same initial state)."""
    if tets:
        m = D.Mesh.capsule(a=4, c=8, layers=5, diameter=3.912, tets=True, jitter=0.2, seed=5)
    else:
        m = D.Mesh.capsule(a=6, c=12, layers=8, diameter=3.912)
    p = D.params_apollo(monatomic=True)
    p.Npinitial = 20
    D.derive(m, p)
    parts = D.fill_freestream(m, p, ppc=20.0, seed=9)
    print(f"[mixed_tracking] tets={tets} cells={m.ncell} particles={parts.n}", flush=True)
    out = {}
    for prec in (D.PREC_FP64, D.PREC_MIXED):
        q = p.copy()
        q.precision = prec
        eng = D.Engine(m, q, seed=41)
        eng.upload(parts)
        eng.run(1, 30)
        out[prec] = eng.stats()
        eng.close()
    a, b = out[D.PREC_FP64], out[D.PREC_MIXED]
    assert a["dropped"] == 0 and b["dropped"] == 0, (a["dropped"], b["dropped"])
    for k in ("particles", "crossings", "wall_hits"):
        assert abs(b[k] - a[k]) <= 0.02 * a[k] + 4.0 * math.sqrt(a[k]), (k, a[k], b[k])
