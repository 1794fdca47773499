"""GPU parity: the HIP engine against the CPU oracle, through the C ABI.

FP64 engine results must be bit-identical to the oracle (same Philox draws,
same deterministic elementary functions, same operation order, no FMA
contraction): particle state, crossing sequences, per-cell collision state,
accumulators and macro fields.  Wall tallies are summed with atomics on the
GPU, so they are compared to a relative 1e-12.
"""
import numpy as np
import pytest

import oracle_py as O
import pydsmc as D

pytestmark = pytest.mark.gpu


def _same_particles(g, c):
    assert g.n == c.n
    for f in ("x", "y", "z", "vx", "vy", "vz", "ir"):
        a, b = getattr(g, f), getattr(c, f)
        bad = np.flatnonzero(a.view(np.int64) != b.view(np.int64))
        assert bad.size == 0, f"{f}: {bad.size} mismatches, first at {bad[:5]}: {a[bad[:3]]} vs {b[bad[:3]]}"
    assert np.array_equal(g.cell, c.cell)
    assert np.array_equal(g.id, c.id)


def test_acceptance_estimate_within_its_band():
    """The collision kernels decide u < c(cr v_rms)/(sigma c)max from an FP32
    estimate when u is outside a relative band of 2^-11 around it.  That is
    exact only if the estimate's error stays inside the band: checked here
    (to a quarter of the band) over speeds spanning the guarded range, for
    the Apollo gas's eta and for eta up to 0.5."""
    rng = np.random.default_rng(7)
    x = np.concatenate([2.0 ** rng.uniform(-60, 60, 200000), rng.uniform(1, 5000, 200000), [1.0, 2.0 ** -60, 2.0 ** 60]])
    p = D.params_apollo(monatomic=True)
    D.derive(D.Mesh.capsule(a=4, c=8, layers=3, diameter=3.912), p)
    for eta in (p.eta, 0.05, 0.5, -0.2):
        crm = 0.7 * x[200000:].max() ** (1 - 2 * eta) * p.coll_prefactor
        out = D.device_accept(x, p.coll_prefactor, eta, crm)
        est, ratio = out[:, 0], out[:, 1]
        used = est > 0
        assert used[200000:].all(), eta
        rel = np.abs(est[used] / ratio[used] - 1.0)
        print(f"[accept_estimate] eta={eta:.3f} used={used.sum()} max rel err={rel.max():.3e}", flush=True)
        assert rel.max() < 2.0 ** -13, (eta, rel.max())
    # hard spheres never use the estimate
    assert not (D.device_accept(x[:1000], p.coll_prefactor, 0.0, 1.0)[:, 0] > 0).any()


def _sorted_log(ev):
    if len(ev) == 0:
        return ev
    order = np.lexsort((ev[:, 2], ev[:, 1], ev[:, 0]))
    return ev[order]


def test_device_math_bitwise():
    rng = np.random.default_rng(0)
    u = np.concatenate([rng.random(200000), [1.0, 0.5, 0.25, 1e-300, 2.0 ** -53, 0.125, 0.875, 1 - 2 ** -53]])
    u = u[u > 0]
    dev = D.device_math(u)
    host = O.host_math(u)
    for k in range(6):
        bad = np.flatnonzero(dev[:, k].view(np.int64) != host[:, k].view(np.int64))
        assert bad.size == 0, f"function {k}: {bad.size} mismatches, u={u[bad[:3]]}"


def _case(kind, seed=3):
    if kind == "capsule":
        mesh = D.Mesh.capsule(a=6, c=12, layers=8, diameter=3.912)
        p = D.params_apollo(monatomic=False)
        p.Npinitial = 8
        D.derive(mesh, p)
        p.NSS = 3
        p.surface_stats = 1
        parts = D.fill_freestream(mesh, p, ppc=8.0, seed=seed)
    elif kind == "capsule_morton":
        # the capsule with its cells renumbered in Morton order
        mesh, _ = D.Mesh.capsule(a=6, c=12, layers=8, diameter=3.912).renumber(D.ORDER_MORTON)
        p = D.params_apollo(monatomic=False)
        p.Npinitial = 8
        D.derive(mesh, p)
        p.NSS = 3
        p.surface_stats = 1
        parts = D.fill_freestream(mesh, p, ppc=8.0, seed=seed)
    elif kind == "capsule_tets":
        mesh = D.Mesh.capsule(a=4, c=8, layers=5, diameter=3.912, tets=True, jitter=0.2, seed=5)
        p = D.params_apollo(monatomic=True)
        p.Npinitial = 6
        D.derive(mesh, p)
        p.NSS = 2
        parts = D.fill_freestream(mesh, p, ppc=6.0, seed=seed)
    elif kind == "box_dense":
        # > 64 particles per cell: strided moment sums, many candidates per cell
        mesh = D.Mesh.box_tets(3)
        p = D.params_box(0.05)
        p.Npinitial = 100
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=100.0, temperature=1.0, seed=seed)
    elif kind == "box_crowded":
        # ~400 particles per cell: more than a wave's LDS stage holds, so the
        # one-wave-per-cell kernel works on the cell in HBM (with its
        # per-particle conflict words in LDS)
        mesh = D.Mesh.box_tets(2)
        p = D.params_box(0.05)
        p.Npinitial = 400
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=400.0, temperature=1.0, seed=seed)
    elif kind == "box_vhs":
        # variable hard spheres (omega 0.74) with many candidates per cell:
        # the acceptance test's FP32 estimate and its exact fallback
        mesh = D.Mesh.box_tets(3)
        p = D.params_box(0.05)
        p.omega = 0.74
        p.Npinitial = 100
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=100.0, temperature=1.0, seed=seed)
    elif kind == "box_packed":
        # ~2000 particles per cell: more than the wide kernel's per-particle
        # conflict words cover, so its batches compare candidates lane by lane
        mesh = D.Mesh.box_tets(1)
        p = D.params_box(0.05)
        p.Npinitial = 2000
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=2000.0, temperature=1.0, seed=seed)
    elif kind == "box_sparse":
        # 1.5 particles per cell: empty cells, singletons and pairs (no NTC
        # candidates, one candidate), and cells emptied and refilled between
        # steps, in the moment sums, the windows and the sparse-cell records
        mesh = D.Mesh.box_tets(6)
        p = D.params_box(0.05)
        p.Npinitial = 2
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=1.5, temperature=1.0, seed=seed)
    elif kind == "box_hex":
        mesh = D.Mesh.box_hexes(6)
        p = D.params_box(0.05)
        p.Npinitial = 10
        p.wall_model = D.WALL_DIFFUSE
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=10.0, temperature=1.0, seed=seed)
    else:
        mesh = D.Mesh.box_tets(6)
        p = D.params_box(0.05)
        p.Npinitial = 10
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=10.0, temperature=1.0, seed=seed)
    return mesh, p, parts


# move_chunk: particles per move workgroup (lanes refill from the chunk); the
# schedule must not change any walk
@pytest.mark.parametrize("kind,move_chunk", [("box_tets", None), ("box_hex", None), ("capsule", None),
                                             ("capsule_tets", None), ("box_tets", 64), ("capsule", 256),
                                             ("capsule_tets", 4096), ("box_hex", 100)])
def test_track_parity(kind, move_chunk):
    mesh, p, parts = _case(kind)
    print(f"[track_parity] {kind} move_chunk={move_chunk} cells={mesh.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(mesh, p, seed=11)
    eng.set_tuning(move_chunk=move_chunk)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=11)
    orc.set_particles(parts)
    eng.log_enable(4_000_000)
    orc.log_enable(True)
    for s in (1, 2, 3):
        eng.step(s, D.STAGE_MOVE | D.STAGE_SORT)
        orc.step(s, D.STAGE_MOVE | D.STAGE_SORT)
    ge, ce = _sorted_log(eng.log_get()), _sorted_log(orc.log_get())
    assert len(ge) == len(ce) and len(ge) > 0
    assert np.array_equal(ge, ce)
    _same_particles(eng.download(), orc.particles())
    gs, cs = eng.stats(), orc.stats()
    for k in ("particles", "removed", "dropped", "wall_hits", "crossings"):
        assert gs[k] == cs[k], (k, gs[k], cs[k])


def _degenerate_starts(mesh, seed=7):
    """Particles of every cell started where the face walk is degenerate: on
    each of the cell's nodes, face centroids and edge midpoints (on a face,
    so the first exit can be at t = 0, shared with neighbouring cells), and
    at its centroid at rest, moving parallel to a face and moving along the
    axes (the box's faces are axis planes).  Velocities elsewhere are
    Maxwellian at T = 1/2 per component."""
    a = mesh.arrays()
    rng = np.random.default_rng(seed)
    X = a["node_xyz"]
    pos, vel, cells = [], [], []
    for c in range(mesh.ncell):
        faces = a["cell_faces"][c][: a["cell_nface"][c]]
        nodes, edges = [], set()
        for f in faces:
            fn = [int(v) for v in a["face_nodes"][f] if v >= 0]
            for k, v in enumerate(fn):
                if v not in nodes:
                    nodes.append(v)
                w = fn[(k + 1) % len(fn)]
                edges.add((min(v, w), max(v, w)))
        pts = [X[v] for v in nodes]
        pts += [X[[int(v) for v in a["face_nodes"][f] if v >= 0]].mean(axis=0) for f in faces]
        pts += [0.5 * (X[i] + X[j]) for i, j in sorted(edges)]
        for q in pts:
            pos.append(q)
            vel.append(rng.normal(0.0, np.sqrt(0.5), 3))
            cells.append(c)
        # a direction in the plane of face 0: across it a face whose normal
        # is not parallel to face 0's
        n0 = a["face_normal"][faces[0]]
        t = max((np.cross(n0, a["face_normal"][f]) for f in faces[1:]), key=np.linalg.norm)
        assert np.linalg.norm(t) > 1e-3
        for v in (np.zeros(3), t / np.linalg.norm(t), np.array([1.0, 0.0, 0.0]), np.array([0.0, -0.7, 0.0])):
            pos.append(a["cell_centroid"][c])
            vel.append(v)
            cells.append(c)
    pos, vel = np.array(pos), np.array(vel)
    P = D.Particles(len(pos))
    P.x[:], P.y[:], P.z[:] = pos[:, 0], pos[:, 1], pos[:, 2]
    P.vx[:], P.vy[:], P.vz[:] = vel[:, 0], vel[:, 1], vel[:, 2]
    P.cell[:] = cells
    P.id[:] = np.arange(P.n, dtype=np.uint32)
    return P


@pytest.mark.parametrize("kind", ["box_tets", "box_hex", "capsule"])
def test_track_parity_degenerate_starts(kind):
    """Walks that start on nodes, edges and faces (exits at t = 0, rays in a
    face's plane, particles at rest) take the reference's decisions bit for
    bit: the same crossings, wall hits and drops as the oracle.  Most of
    these particles are dropped on their first visit, as the reference drops
    them: a walk whose every hit has t <= 1e-14 finds no exit, and neither
    does a particle at rest (det = 0 on every face)."""
    if kind == "capsule":
        mesh = D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912)
        p = D.params_apollo(monatomic=True)
        p.Npinitial = 4
    else:
        mesh = D.Mesh.box_tets(4) if kind == "box_tets" else D.Mesh.box_hexes(4)
        p = D.params_box(0.05)
        p.Npinitial = 10
        if kind == "box_hex":
            p.wall_model = D.WALL_DIFFUSE
    D.derive(mesh, p)
    parts = _degenerate_starts(mesh)
    rest = np.flatnonzero((parts.vx == 0.0) & (parts.vy == 0.0) & (parts.vz == 0.0))
    print(f"[track_degenerate] {kind} cells={mesh.ncell} particles={parts.n} at rest={rest.size}", flush=True)
    eng = D.Engine(mesh, p, seed=11, capacity=2 * parts.n)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=11)
    orc.set_particles(parts)
    eng.log_enable(1_000_000)
    orc.log_enable(True)
    for s in (1, 2, 3):
        eng.step(s, D.STAGE_MOVE | D.STAGE_SORT)
        orc.step(s, D.STAGE_MOVE | D.STAGE_SORT)
    ge, ce = _sorted_log(eng.log_get()), _sorted_log(orc.log_get())
    assert len(ge) == len(ce) and len(ge) > 0
    assert np.array_equal(ge, ce)
    g = eng.download()
    _same_particles(g, orc.particles())
    gs, cs = eng.stats(), orc.stats()
    for k in ("particles", "removed", "dropped", "wall_hits", "crossings"):
        assert gs[k] == cs[k], (k, gs[k], cs[k])
    assert gs["dropped"] > 0 and gs["crossings"] > 0
    assert not np.isin(parts.id[rest], g.id).any()


@pytest.mark.parametrize("kind", ["box_tets", "box_hex", "capsule"])
def test_track_parity_fast_particles(kind):
    """Particles at 40 times their thermal speed: walks of tens of cells and
    several wall hits per step (specular walls reflected inside the move on
    the tet box; diffuse walls on the hex box and the capsule, whose walks
    continue in the wall kernel after each hit), and most of the capsule's
    particles leaving through its outflow.  Crossings, walls, exits and the
    particles stay bit-identical to the oracle."""
    if kind == "capsule":
        mesh = D.Mesh.capsule(a=4, c=8, layers=4, diameter=3.912)
        p = D.params_apollo(monatomic=False)
        p.Npinitial = 4
        p.surface_stats = 1
        p.NSS = 0
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=4.0, seed=5)
    else:
        mesh = D.Mesh.box_tets(4) if kind == "box_tets" else D.Mesh.box_hexes(4)
        p = D.params_box(0.05)
        p.Npinitial = 10
        if kind == "box_hex":
            p.wall_model = D.WALL_DIFFUSE
        D.derive(mesh, p)
        parts = D.fill_freestream(mesh, p, ppc=10.0, temperature=1.0, seed=5)
    for f in ("vx", "vy", "vz"):
        getattr(parts, f)[:] *= 40.0
    print(f"[track_fast] {kind} cells={mesh.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(mesh, p, seed=11)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=11)
    orc.set_particles(parts)
    eng.log_enable(2_000_000)
    orc.log_enable(True)
    for s in (1, 2, 3):
        eng.step(s, D.STAGE_MOVE | D.STAGE_SORT)
        orc.step(s, D.STAGE_MOVE | D.STAGE_SORT)
    ge, ce = _sorted_log(eng.log_get()), _sorted_log(orc.log_get())
    assert len(ge) == len(ce) and len(ge) > 0
    assert np.array_equal(ge, ce)
    _same_particles(eng.download(), orc.particles())
    gs, cs = eng.stats(), orc.stats()
    for k in ("particles", "removed", "dropped", "wall_hits", "crossings"):
        assert gs[k] == cs[k], (k, gs[k], cs[k])
    assert gs["wall_hits"] > 0 and gs["crossings"] > parts.n


# wide_min: None = default routing; 0 = every cell one wave per cell; -1 =
# every cell one thread per cell
@pytest.mark.parametrize("kind,wide_min", [("box_tets", None), ("capsule", None), ("capsule_tets", None),
                                           ("box_dense", None), ("box_dense", -1), ("capsule", 0),
                                           ("box_dense", 0), ("box_dense", 20), ("box_crowded", None),
                                           ("box_crowded", -1), ("box_packed", 0),
                                           ("box_vhs", -1), ("box_vhs", 0), ("capsule_morton", None),
                                           ("box_sparse", None), ("box_sparse", 0)])
def test_full_step_parity(kind, wide_min):
    mesh, p, parts = _case(kind)
    print(f"[full_step_parity] {kind} wide_min={wide_min} cells={mesh.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(mesh, p, seed=21)
    if wide_min is not None:
        eng.set_tuning(wide_min=wide_min)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=21)
    orc.set_particles(parts)
    eng.statistic_pre()
    orc.statistic_pre()
    for which in (D.MACRO_LOCAL, D.MACRO_RECORD):
        assert np.array_equal(eng.macro(which), orc.macro(which), equal_nan=True)
    nsteps = 6
    for s in range(1, nsteps + 1):
        eng.step(s)
        orc.step(s)
    _same_particles(eng.download(), orc.particles())
    gs, cs = eng.cell_scalars(), orc.cell_scalars()
    for k in gs:
        assert np.array_equal(gs[k], cs[k]), k
    for fam in (0, 1, 2):
        assert np.array_equal(eng.accum(fam), orc.accum(fam)), fam
    for which in (D.MACRO_LOCAL, D.MACRO_FINAL, D.MACRO_RECORD, D.MACRO_OUTPUT):
        assert np.array_equal(eng.macro(which), orc.macro(which), equal_nan=True), which
    st_g, st_c = eng.stats(), orc.stats()
    for k in st_g:
        if k != "last_step":
            assert st_g[k] == st_c[k], (k, st_g[k], st_c[k])
    sg, sc = eng.surface(), orc.surface()
    np.testing.assert_allclose(sg, sc, rtol=1e-12, atol=1e-12)


def test_long_run_parity():
    """60 full steps of the diatomic capsule (inflow, diffuse walls with
    surface tallies, steady sampling from step 20, GSIS record every 10
    steps): the engine stays bit-identical to the oracle over the whole run,
    so every sampled field, surface pressure and heat flux agrees exactly,
    not only within sampling noise."""
    mesh, p, parts = _case("capsule")
    p.NSS = 20
    p.Nevery = 5
    p.Nrepeat = 2
    print(f"[long_run_parity] capsule cells={mesh.ncell} particles={parts.n} steps=60", flush=True)
    eng = D.Engine(mesh, p, seed=77)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=77)
    orc.set_particles(parts)
    for s in range(1, 61):
        eng.step(s)
        orc.step(s)
    _same_particles(eng.download(), orc.particles())
    for fam in (0, 1, 2):
        assert np.array_equal(eng.accum(fam), orc.accum(fam)), fam
    for which in (D.MACRO_LOCAL, D.MACRO_FINAL, D.MACRO_RECORD, D.MACRO_OUTPUT):
        assert np.array_equal(eng.macro(which), orc.macro(which), equal_nan=True), which
    np.testing.assert_allclose(eng.surface(), orc.surface(), rtol=1e-12, atol=1e-12)
    st = eng.stats()
    assert st["collisions"] > 0 and st["wall_hits"] > 0 and st["emitted"] > 0


def test_empty_start_fills_from_the_inlet():
    """An empty domain at the start: the inflow alone fills the capsule, and
    every particle, accumulator and count matches the oracle's."""
    mesh, p, parts = _case("capsule")
    empty = parts.take(np.arange(0))
    print(f"[empty_start] capsule cells={mesh.ncell}", flush=True)
    eng = D.Engine(mesh, p, seed=21)
    eng.upload(empty)
    orc = O.Oracle(mesh, p, seed=21)
    orc.set_particles(empty)
    for s in range(1, 6):
        eng.step(s)
        orc.step(s)
    g = eng.download()
    assert g.n > 0
    _same_particles(g, orc.particles())
    for fam in (0, 1, 2):
        assert np.array_equal(eng.accum(fam), orc.accum(fam)), fam
    st_g, st_c = eng.stats(), orc.stats()
    for k in ("emitted", "removed", "dropped", "coll_pairs", "collisions"):
        assert st_g[k] == st_c[k], (k, st_g[k], st_c[k])


def _hip_runtime():
    """The HIP runtime the engine library is linked against (from this
    process's mappings), for creating a caller's stream in a test."""
    import ctypes
    D.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    assert paths, "HIP runtime not mapped"
    return ctypes.CDLL(sorted(paths)[0])


def test_run_on_a_caller_stream_matches_run():
    """Steps ordered on a caller's HIP stream (dsmc_run_on_stream, with work
    enqueued on that stream before and after the run) end in the state
    dsmc_run reaches, and are complete once the caller's stream is; a stream
    under graph capture is refused before anything is enqueued on it."""
    import ctypes
    mesh, p, parts = _case("capsule")
    print(f"[run_on_stream] capsule cells={mesh.ncell} particles={parts.n}", flush=True)
    a = D.Engine(mesh, p, seed=9)
    a.upload(parts)
    a.run(1, 6)
    b = D.Engine(mesh, p, seed=9)
    b.upload(parts)
    hip = _hip_runtime()
    st, buf, g = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    assert hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(1 << 20)) == 0
    assert hip.hipMemsetAsync(buf, 1, ctypes.c_size_t(1 << 20), st) == 0
    b.run_on_stream(1, 6, st.value)
    assert hip.hipMemsetAsync(buf, 2, ctypes.c_size_t(1 << 20), st) == 0
    assert hip.hipStreamSynchronize(st) == 0
    assert b.run_query()  # the caller's stream waited for the run
    _same_particles(a.download(), b.download())
    for fam in (0, 1, 2):
        assert np.array_equal(a.accum(fam), b.accum(fam)), fam
    assert hip.hipStreamBeginCapture(st, 2) == 0  # relaxed capture mode
    with pytest.raises(D.DsmcError, match="captured"):
        b.run_on_stream(7, 8, st.value)
    assert hip.hipStreamEndCapture(st, ctypes.byref(g)) == 0
    if g.value:
        hip.hipGraphDestroy(g)
    hip.hipFree(buf)
    hip.hipStreamDestroy(st)
    b.run(7, 8)
    a.run(7, 8)
    _same_particles(a.download(), b.download())


def test_capacity_overflow_is_an_error():
    """Inflow beyond the engine's capacity ends the step with a capacity
    error: the emission is cut at the capacity, so no kernel writes past it,
    and the engine can still be read and destroyed."""
    mesh, p, parts = _case("capsule")
    print(f"[capacity_overflow] capsule cells={mesh.ncell} particles={parts.n} capacity={parts.n + 5}", flush=True)
    eng = D.Engine(mesh, p, seed=21, capacity=parts.n + 5)
    eng.upload(parts)
    with pytest.raises(D.DsmcError, match="capacity"):
        for s in range(1, 4):
            eng.step(s)
    assert eng.download().n <= parts.n + 5
    eng.close()


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_MIXED])
@pytest.mark.parametrize("kind", ["box_dense", "box_crowded", "box_packed"])
def test_wide_matches_thread_path(kind, prec):
    """Every cell one wave (staged in LDS for box_dense, worked on in HBM for
    box_crowded and box_packed, with per-particle conflict words in LDS up to
    box_crowded's size and lane-by-lane conflict tests for box_packed) against
    every cell one thread: the same particles, cell scalars and accumulators
    bit for bit, in FP64 and in the mixed form (which has no oracle)."""
    mesh, p, parts = _case(kind)
    p.precision = prec
    print(f"[wide_vs_thread] {kind} prec={prec} cells={mesh.ncell} particles={parts.n}", flush=True)
    out = []
    for wm in (0, -1):
        eng = D.Engine(mesh, p, seed=31)
        eng.set_tuning(wide_min=wm)
        eng.upload(parts)
        eng.run(1, 6)
        out.append((eng.download(), eng.cell_scalars(), [eng.accum(f) for f in (0, 1, 2)], eng.stats()))
        eng.close()
    (qa, ca, aa, sa), (qb, cb, ab, sb) = out
    _same_particles(qa, qb)
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k
    for fa, fb in zip(aa, ab):
        assert np.array_equal(fa, fb)
    assert sa["collisions"] == sb["collisions"] > 0


@pytest.mark.parametrize("kind", ["capsule", "box_crowded"])
def test_collide_workgroup_size_does_not_change_results(kind):
    """The thread-per-cell collide+sample kernel at 64, 128 and 256 threads
    per workgroup (the autotuner's coll_block setting): particles, cell
    scalars, accumulators and counters bit-identical, on the capsule (cells
    of every size, crowded ones left to the wide kernel) and on a box whose
    cells all go through the thread path."""
    mesh, p, parts = _case(kind)
    print(f"[coll_block] {kind} cells={mesh.ncell} particles={parts.n}", flush=True)
    out = []
    for wg in (64, 128, 256):
        eng = D.Engine(mesh, p, seed=37)
        eng.set_tuning(coll_block=wg, wide_min=-1 if kind == "box_crowded" else 96)
        assert eng.tuning()["coll_block"] == wg
        eng.upload(parts)
        eng.run(1, 5)
        out.append((eng.download(), eng.cell_scalars(), [eng.accum(f) for f in (0, 1, 2)], eng.stats()))
        eng.close()
    q0, c0, a0, s0 = out[0]
    for q, c, a, st in out[1:]:
        _same_particles(q0, q)
        for k in c0:
            assert np.array_equal(c0[k], c[k]), k
        for fa, fb in zip(a0, a):
            assert np.array_equal(fa, fb)
        assert st == s0
    assert s0["collisions"] > 0


def test_collide_workgroup_size_is_checked():
    mesh, p, parts = _case("capsule")
    eng = D.Engine(mesh, p, seed=1)
    with pytest.raises(RuntimeError, match="coll_block"):
        eng.set_tuning(coll_block=100)
    eng.close()


def test_stagewise_matches_full_step():
    """Driving a step stage by stage, the way the reference driver calls
    preprocesseffquad, advection, cache2chain, collisionDSMC and
    statistic_macro (src/meshImportTest.cpp:207-234), gives the same state as
    one DSMC_STAGE_ALL call, and counts each step once."""
    mesh, p, parts = _case("capsule")
    print(f"[stagewise] capsule cells={mesh.ncell} particles={parts.n}", flush=True)
    full = D.Engine(mesh, p, seed=5)
    full.upload(parts)
    split = D.Engine(mesh, p, seed=5)
    split.upload(parts)
    stages = (D.STAGE_EMIT, D.STAGE_MOVE, D.STAGE_SORT, D.STAGE_COLLIDE, D.STAGE_SAMPLE)
    for s in range(1, 5):
        full.step(s)
        for m in stages:
            split.step(s, m)
    _same_particles(split.download(), full.download())
    for fam in (0, 1, 2):
        assert np.array_equal(split.accum(fam), full.accum(fam)), fam
    a, b = split.stats(), full.stats()
    for k in a:
        assert a[k] == b[k], (k, a[k], b[k])
    assert np.array_equal(split.macro(D.MACRO_OUTPUT), full.macro(D.MACRO_OUTPUT), equal_nan=True)


def test_irregular_stage_sequences_match_the_oracle():
    """After a regroup the engine holds the particles' positions through the
    regroup's permutation until the next move writes them in cell order.
    Every path that needs them in cell order first must see the same
    particles as the oracle: a regroup with no move (cache2chain without
    advection), emission in a call of its own before the move, a download
    between the regroup and the move, graph-replayed runs after that, and
    the continuum coupling after a regroup."""
    mesh, p, parts = _case("capsule")
    print(f"[stage sequences] capsule cells={mesh.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(mesh, p, seed=21)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=21)
    orc.set_particles(parts)

    def both(s, mask=D.STAGE_ALL):
        eng.step(s, mask)
        orc.step(s, mask)

    both(1)
    both(2, D.STAGE_SORT)                       # regroup with no move
    both(2, D.STAGE_COLLIDE | D.STAGE_SAMPLE)
    _same_particles(eng.download(), orc.particles())
    both(3, D.STAGE_EMIT)                       # emission on its own
    both(3, D.STAGE_MOVE | D.STAGE_SORT)
    _same_particles(eng.download(), orc.particles())  # positions put in cell order here
    both(3, D.STAGE_COLLIDE | D.STAGE_SAMPLE)
    both(4)
    eng.run(5, 9)                               # graph replays after the download
    for s in range(5, 10):
        orc.step(s)
    _same_particles(eng.download(), orc.particles())
    for fam in (0, 1, 2):
        assert np.array_equal(eng.accum(fam), orc.accum(fam)), fam
    # the continuum coupling after a full step, whose regroup left the
    # positions held through the permutation
    both(10)
    counts = eng.cell_scalars()["count"].astype(np.float64)
    ns, acc, stop = _ns_fields(mesh, p, counts)
    assert eng.ns_apply(11, ns, acc) == stop
    assert orc.ns_apply(11, ns, acc) == stop
    _same_particles(eng.download(), orc.particles())
    both(11)
    _same_particles(eng.download(), orc.particles())
    for fam in (0, 1, 2):
        assert np.array_equal(eng.accum(fam), orc.accum(fam)), fam


def test_stage_order_guard():
    """A move needs the particles grouped by cell: a second move without the
    regroup in between, or a collide after a move, is refused with an error
    instead of walking stale cells."""
    mesh, p, parts = _case("box_tets")
    eng = D.Engine(mesh, p, seed=3)
    eng.upload(parts)
    eng.step(1, D.STAGE_MOVE)
    with pytest.raises(D.DsmcError, match="regroup"):
        eng.step(2, D.STAGE_MOVE)
    with pytest.raises(D.DsmcError, match="regroup"):
        eng.step(1, D.STAGE_COLLIDE)
    eng.step(1, D.STAGE_SORT)
    eng.step(2, D.STAGE_ALL)
    assert eng.stats()["particles"] > 0


def test_autotune_steps_are_ordinary_steps():
    """The autotuner times live steps under different launch settings; the
    run must end bit-identical to an untuned run of the same steps."""
    mesh, p, parts = _case("capsule")
    print(f"[autotune] capsule cells={mesh.ncell} particles={parts.n}", flush=True)
    tuned = D.Engine(mesh, p, seed=9)
    tuned.upload(parts)
    tuning, used = tuned.autotune(1, 1)
    assert used >= 10 and tuning["move_chunk"] in (512, 1024, 2048, 4096, 8192)
    assert tuning["coll_block"] in (64, 128, 256)
    last = used + 2
    tuned.run(used + 1, last)
    plain = D.Engine(mesh, p, seed=9)
    plain.upload(parts)
    plain.run(1, last)
    _same_particles(tuned.download(), plain.download())
    for fam in (0, 1, 2):
        assert np.array_equal(tuned.accum(fam), plain.accum(fam)), fam
    a, b = tuned.stats(), plain.stats()
    for k in a:
        assert a[k] == b[k], (k, a[k], b[k])


def test_batch_cases_match_alone():
    """Three different cases in one batch engine: every case's particles,
    accumulators and fields are bit-identical to the same case run alone
    with the same seed and to the oracle run on that case; its wall tallies
    agree to 1e-12."""
    kinds = ["capsule", "capsule_tets", "box_tets"]
    seeds = [31, 32, 33]
    cases = [_case(k, seed=4 + i) for i, k in enumerate(kinds)]
    steps = 5
    alone = []
    for (mesh, p, parts), sd in zip(cases, seeds):
        eng = D.Engine(mesh, p, seed=sd)
        eng.upload(parts)
        eng.run(1, steps)
        alone.append({"parts": eng.download(), "acc": [eng.accum(f) for f in (0, 1, 2)],
                      "macro": eng.macro(D.MACRO_OUTPUT), "stats": eng.stats(), "surf": eng.surface()})
        eng.close()
        orc = O.Oracle(mesh, p, seed=sd)
        orc.set_particles(parts)
        for s in range(1, steps + 1):
            orc.step(s)
        _same_particles(alone[-1]["parts"], orc.particles())
        for f in (0, 1, 2):
            assert np.array_equal(alone[-1]["acc"][f], orc.accum(f)), f
        assert np.array_equal(alone[-1]["macro"], orc.macro(D.MACRO_OUTPUT), equal_nan=True)
    bat = D.Engine.batch([c[0] for c in cases], [c[1] for c in cases], seeds)
    print(f"[batch] cases={len(cases)} cells={bat.mesh.ncell}", flush=True)
    bat.upload(bat.batch_particles([c[2] for c in cases]))
    bat.run(1, steps)
    allp = bat.download()
    acc = [bat.accum(f) for f in (0, 1, 2)]
    mac = bat.macro(D.MACRO_OUTPUT)
    surf = bat.surface()
    for k, ref in enumerate(alone):
        info = bat.case_info(k)
        cb, ce = info["cell_begin"], info["cell_end"]
        _same_particles(bat.case_particles(allp, k), ref["parts"])
        assert info["particles"] == ref["parts"].n
        for f in range(3):
            assert np.array_equal(acc[f][cb:ce], ref["acc"][f]), (k, f)
        assert np.array_equal(mac[cb:ce], ref["macro"], equal_nan=True), k
        np.testing.assert_allclose(surf[info["face_begin"]:info["face_end"]], ref["surf"], rtol=1e-12, atol=1e-12)
    st = bat.stats()
    for key in ("particles", "emitted", "removed", "dropped", "wall_hits", "collisions", "particle_steps"):
        assert st[key] == sum(r["stats"][key] for r in alone), key


def test_graph_replay_matches_eager():
    """dsmc_run replays pairs of steps captured as a HIP graph: runs of
    several graph replays, with eager steps and host synchronisation between
    them, end bit-identical to the same steps launched kernel by kernel."""
    mesh, p, parts = _case("capsule")
    g = D.Engine(mesh, p, seed=41)
    g.set_tuning(graph=1)
    g.upload(parts)
    e = D.Engine(mesh, p, seed=41)
    e.set_tuning(graph=0)
    e.upload(parts)
    for first, last in ((1, 3), (4, 6), (7, 11), (12, 12), (13, 17)):
        g.run(first, last)
        e.run(first, last)
        _same_particles(g.download(), e.download())
        for fam in (0, 1, 2):
            assert np.array_equal(g.accum(fam), e.accum(fam)), (last, fam)
        a, b = g.stats(), e.stats()
        for k in a:
            assert a[k] == b[k], (last, k, a[k], b[k])
    assert np.array_equal(g.macro(D.MACRO_OUTPUT), e.macro(D.MACRO_OUTPUT), equal_nan=True)


@pytest.mark.parametrize("prec", [D.PREC_FP64, D.PREC_MIXED])
def test_device_fill_matches_host_fill(prec):
    """The device-side fill places bit-identically the particles of the host
    fill plus upload, for a single engine and for every case of a batch."""
    cases = [_case("capsule"), _case("box_tets"), _case("capsule_tets")]
    ppcs, seeds = [8.0, 10.0, 6.0], [101, 102, 103]
    temps = [0.0, 1.0, 0.0]
    for mesh, p, _ in cases:
        p.precision = prec
    host = []
    for (mesh, p, _), ppc, sd, T in zip(cases, ppcs, seeds, temps):
        parts = D.fill_freestream(mesh, p, ppc=ppc, temperature=T, seed=sd)
        a = D.Engine(mesh, p, seed=1)
        a.upload(parts)
        b = D.Engine(mesh, p, seed=1)
        n = b.fill_device(ppc, sd, T)
        assert n == parts.n
        _same_particles(b.download(), a.download())
        host.append(a.download())
    bat = D.Engine.batch([c[0] for c in cases], [c[1] for c in cases], [1, 2, 3])
    bat.fill_device(ppcs, seeds, temps)
    allp = bat.download()
    for k, ref in enumerate(host):
        _same_particles(bat.case_particles(allp, k), ref)


def _ns_fields(m, p, counts, seed=17):
    """NS fields for a coupling test: density scaled by 0.6-1.5 per cell,
    random drift, T and T_r in [0.5, 2], 80% of cells coupled, a NaN density
    in one coupled cell at 70% of the cells."""
    rng = np.random.default_rng(seed)
    nc = m.ncell
    vol = m.arrays()["cell_volume"]
    ns = np.empty((nc, 6))
    ns[:, 0] = counts * p.Neff / (vol * p.n_ref) * rng.uniform(0.6, 1.5, nc)
    ns[:, 1:4] = rng.normal(0.0, 0.5, (nc, 3))
    ns[:, 4] = rng.uniform(0.5, 2.0, nc)
    ns[:, 5] = rng.uniform(0.5, 2.0, nc)
    acc = (rng.uniform(size=nc) < 0.8).astype(np.uint8)
    stop = int(np.flatnonzero(acc)[int(0.7 * acc.sum())])
    ns[stop, 0] = np.nan
    return ns, acc, stop


@pytest.mark.parametrize("kind", ["capsule", "capsule_tets", "box_tets"])
def test_ns_apply_parity(kind):
    """Continuum coupling on the device (dsmc_ns_apply: molecular_velocity_change
    with replication, deletion and Maxwellian refill) against the oracle: the
    same particles bit for bit after the update, and after two more steps."""
    mesh, p, parts = _case(kind)
    print(f"[ns_apply_parity] {kind} cells={mesh.ncell} particles={parts.n}", flush=True)
    eng = D.Engine(mesh, p, seed=21)
    eng.upload(parts)
    orc = O.Oracle(mesh, p, seed=21)
    orc.set_particles(parts)
    for s in (1, 2):
        eng.step(s)
        orc.step(s)
    counts = eng.cell_scalars()["count"].astype(np.float64)
    ns, acc, stop = _ns_fields(mesh, p, counts)
    assert eng.ns_apply(3, ns, acc) == stop
    assert orc.ns_apply(3, ns, acc) == stop
    _same_particles(eng.download(), orc.particles())
    assert not np.array_equal(eng.cell_scalars()["count"], counts)
    for s in (3, 4):
        eng.step(s)
        orc.step(s)
    _same_particles(eng.download(), orc.particles())
    for k in ("particles", "collisions", "crossings"):
        assert eng.stats()[k] == orc.stats()[k], k
    # every cell coupled, no NaN: the whole mesh is updated
    ns2, _, _ = _ns_fields(mesh, p, eng.cell_scalars()["count"].astype(np.float64), seed=5)
    ns2[:, 0] = np.abs(np.nan_to_num(ns2[:, 0], nan=1.0))
    assert eng.ns_apply(5, ns2) == -1 and orc.ns_apply(5, ns2) == -1
    _same_particles(eng.download(), orc.particles())


def test_ns_apply_mixed_counts_match_fp64():
    """The mixed engine's coupling makes the same count, deletion and
    replication choices as the FP64 engine (they do not depend on the stored
    velocity precision): same cells, ids and positions; velocities agree to
    FP32 rounding."""
    mesh, p, parts = _case("capsule")
    out = []
    for prec in (D.PREC_FP64, D.PREC_MIXED):
        q = p.copy()
        q.precision = prec
        eng = D.Engine(mesh, q, seed=21)
        eng.upload(parts)
        counts = eng.cell_scalars()["count"].astype(np.float64)
        ns, acc, stop = _ns_fields(mesh, q, counts)
        assert eng.ns_apply(1, ns, acc) == stop
        out.append(eng.download())
        eng.close()
    a, b = out
    assert a.n == b.n
    for f in ("cell", "id", "x", "y", "z"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for f in ("vx", "vy", "vz"):
        assert np.allclose(getattr(a, f), getattr(b, f), rtol=1e-5, atol=1e-5), f


def test_ns_apply_refuses_oversized_and_infinite_states():
    """A coupled density that would make the particle total exceed the
    capacity (every cell near it, so a 32-bit sum of the per-cell counts
    would wrap) is a capacity error, and an infinite NS field before the NaN
    stop an argument error; in both cases the particles stay as they were.
    A NaN density still stops the update at its cell, as in the reference."""
    mesh, p, parts = _case("box_tets")
    eng = D.Engine(mesh, p, seed=21)
    eng.upload(parts)
    before = eng.download()
    counts = eng.cell_scalars()["count"].astype(np.float64)
    ns, _, _ = _ns_fields(mesh, p, counts)
    ns[:, 0] = np.abs(np.nan_to_num(ns[:, 0], nan=1.0))
    huge = ns.copy()
    huge[:, 0] = 1e30
    with pytest.raises(D.DsmcError) as ei:
        eng.ns_apply(1, huge)
    print(f"[ns_apply guard] huge density: {ei.value}", flush=True)
    assert ei.value.code == D.ERR_CAPACITY
    inf = ns.copy()
    inf[7, 4] = np.inf
    with pytest.raises(D.DsmcError) as ei:
        eng.ns_apply(1, inf)
    assert ei.value.code == D.ERR_ARG and "cell 7" in str(ei.value)
    after = eng.download()
    _same_particles(after, before)
    # the same infinite field after a NaN stop is never read
    late = ns.copy()
    late[3, 0] = np.nan
    late[7, 4] = np.inf
    assert eng.ns_apply(1, late) == 3
