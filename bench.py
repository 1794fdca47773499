"""Benchmark: DSMC particle-steps/s on one MI355X.

Workload (default): the equilibrium box — unit cube of 64^3 hexes x 6 tets
(1,572,864 cells), hard-sphere gas at T = 273.15 K, Kn = 0.05, 50 particles
per cell (~78.6M particles), specular walls, FP64.  One step = emit + move +
regroup + collide + sample.  Inputs are resident in HBM before the timed
region; the timed region is exactly --steps steps on the engine's stream,
bracketed by device synchronisation.

Also reported on the same JSON line: the mixed-precision engine on the same
particles, the roofline of the dominant kernel (HIP-event timed over the same
steps) and the CPU baseline (the C++ oracle on a bounded sample of the same
workload, on the host's cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dig-appolo_amd"))

import numpy as np  # noqa: E402

import pydsmc as D  # noqa: E402

METRIC = "particle-steps/s (move+collide+sample), FP64 vs mixed; p99 solve latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle for one launch of each kernel group
# (state S = 64 B FP64 / 48 B mixed, velocity+Ir S_v = 32 / 16 B,
# triangle cache M_tri = 80 B x triangles per cell, 896 B per-cell sampling)


def alg_bytes(kernel: str, S: int, Sv: int, mtri: float, ppc: float) -> float:
    if kernel == "move":
        return 2 * S + mtri / ppc
    if kernel == "sort":
        return 2 * S + 4
    if kernel == "collide_sample":
        return 2 * Sv + 896.0 / ppc
    raise KeyError(kernel)


def trial_steps(n_particles: int) -> int:
    """Steps per autotuner trial: about 3e7 particle-steps, so that a trial on
    a small case lasts long enough to rank the settings (2 to 8 steps)."""
    return max(2, min(8, round(3e7 / max(n_particles, 1))))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_case(n: int, ppc: float, mixed: bool):
    mesh = D.Mesh.box_tets(n)
    p = D.params_box(0.05)
    p.Npinitial = int(round(ppc))
    p.NSS = 1000
    p.precision = D.PREC_MIXED if mixed else D.PREC_FP64
    D.derive(mesh, p)
    return mesh, p


def make_apollo(ppc: float, mixed: bool):
    """Apollo capsule, plain DSMC: surrogate 372,500-cell hex O-grid, monatomic
    VHS, Ma 5, Kn 0.01, AoA 30, T_in 142.2 K, diffuse wall at 300 K, inflow."""
    mesh = D.Mesh.capsule(25, 62, 50, 3.912)
    p = D.params_apollo(monatomic=True)
    p.Npinitial = int(round(ppc))
    p.NSS = 250
    p.precision = D.PREC_MIXED if mixed else D.PREC_FP64
    D.derive(mesh, p)
    return mesh, p


def run_engine(mesh, p, parts, steps, warmup, seed=20261203, autotune=True, tuning=None, profile_pass=True):
    """Times `steps` full steps after `warmup` untimed ones, the way the
    engine runs in production: no profiling, steps replayed as HIP graphs
    (both buffer parities captured during set-up).  Then, unless
    profile_pass is off, a separate profiled pass of `steps` more steps
    (eager launches with a HIP event after each kernel group) gives the
    per-kernel split; it is labelled as such and never enters `value`."""
    eng = D.Engine(mesh, p, seed=seed)
    eng.upload(parts)
    tune_steps = 0
    if tuning:
        # launch settings given by the caller (a profiler run replaying the
        # settings a timed run picked)
        eng.set_tuning(**tuning)
        autotune = False
    tuning = eng.tuning()
    if autotune:
        # set-up: the autotuner times live steps (results do not depend on the
        # tuning it picks), then the case restarts from the same particles
        tuning, tune_steps = eng.autotune(1, trial_steps(parts.n))
        log(f"[bench] autotune ({tune_steps} steps): {tuning}")
    # set-up: capture the step-pair graphs of both buffer parities (each run
    # starts with one eager step, which flips the parity), then restart from
    # the same particles; the graphs stay valid across an upload
    eng.upload(parts)
    eng.run(1, 3)
    eng.run(4, 6)
    eng.upload(parts)
    if warmup > 0:
        eng.run(1, warmup)
    st0 = eng.stats()
    t0 = time.perf_counter()
    ms_events = eng.run(warmup + 1, warmup + steps)
    wall = time.perf_counter() - t0
    st1 = eng.stats()
    psteps = st1["particle_steps"] - st0["particle_steps"]
    kt, prof_ms = None, None
    if profile_pass:
        eng.profile(True)
        prof_ms = eng.run(warmup + steps + 1, warmup + 2 * steps)
        kt = eng.kernel_times()
        eng.profile(False)
    res = {
        "particle_steps": psteps,
        "event_ms": ms_events,
        "wall_s": wall,
        "value": psteps / (ms_events / 1e3),
        "ms_per_step": ms_events / steps,
        "kernels": kt,
        "profiled_ms_per_step": None if prof_ms is None else prof_ms / steps,
        "stats": st1,
        "footprint_GB": eng.footprint() / 1e9,
        "tuning": tuning,
        "tune_steps": tune_steps,
    }
    eng.close()
    return res


def cpu_share():
    """Host threads for the CPU baseline: every core this process may run
    on (its affinity mask), capped by OMP_NUM_THREADS when that is set (the
    share of a shared host given to this GPU), with the numbers behind it."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = "unknown CPU"
    try:
        import subprocess
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    if model == "unknown CPU":
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return threads, {"affinity_cores": aff, "nproc": os.cpu_count(), "OMP_NUM_THREADS": omp or None,
                     "model": model}


def cpu_baseline(threads: int, mesh, p, parts, what: str, first_step: int = 1, seconds_target: float = 15.0,
                 max_steps: int = 50, warm: bool = True):
    """The C++ oracle (a strict restatement of the reference's per-step
    path) from the state `parts`, on `threads` host threads: full steps,
    each stage by stage and timed like the reference driver's timers
    (src/meshImportTest.cpp:205-236: Preprocess = preprocesseffquad +
    statistic_macro, Advection, cache2chain, Collision), until
    seconds_target or max_steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    O.set_threads(threads)
    orc = O.Oracle(mesh, p, seed=7)
    orc.set_particles(parts)
    s = first_step
    if warm:
        orc.step(s)  # warm caches
        s += 1
    st0 = orc.stats()["particle_steps"]
    stages = (("Preprocess", D.STAGE_EMIT), ("Advection", D.STAGE_MOVE), ("cache2chain", D.STAGE_SORT),
              ("Collision", D.STAGE_COLLIDE), ("Preprocess", D.STAGE_SAMPLE))
    split = {"Preprocess": 0.0, "Advection": 0.0, "cache2chain": 0.0, "Collision": 0.0}
    t0 = time.perf_counter()
    steps = 0
    while True:
        for name, mask in stages:
            ts = time.perf_counter()
            orc.step(s, mask)
            split[name] += time.perf_counter() - ts
        steps += 1
        s += 1
        if time.perf_counter() - t0 > seconds_target or steps >= max_steps:
            break
    dt = time.perf_counter() - t0
    ps = orc.stats()["particle_steps"] - st0
    del orc
    return {
        "value": ps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what}, {steps} full steps in {dt:.1f} s, {threads} threads",
        "stage_ms_per_step": {k: round(1e3 * v / steps, 2) for k, v in split.items()},
    }


def cpu_pair(mesh, p, parts, what, first_step=1, seconds=15.0, single_seconds=6.0, threads=0):
    """All-share and one-thread CPU baselines of the same state (`threads`
    > 0 overrides the share)."""
    share, host = cpu_share()
    threads = threads or share
    what = f"{what} on {host['model']}"
    cpu = cpu_baseline(threads, mesh, p, parts, what, first_step, seconds_target=seconds)
    cpu["host"] = host
    log(f"[bench] cpu baseline: {cpu['value']:.4g} particle-steps/s on {threads} threads ({cpu['sample']})")
    if threads > 1:
        one = cpu_baseline(1, mesh, p, parts, what, first_step, seconds_target=single_seconds,
                           warm=False)
        cpu["single_thread"] = {"value": one["value"], "sample": one["sample"]}
        log(f"[bench] cpu baseline: {one['value']:.4g} particle-steps/s on 1 thread")
    return cpu


def apollo_record(steps: int = 500, nss: int = 250, ppc: float = 20.0, seed: int = 20261203, cpu: bool = True):
    """The Apollo case as the reference runs it (monatomic VHS, Ma 5,
    Kn 0.01, AoA 30, diffuse wall at 300 K, inflow, Npinitial 20 on the
    372,500-cell surrogate O-grid, `steps` steps with sampling from `nss`):
    FP64 and mixed, each timed over the whole job and over the developed
    flow (steps nss+1..steps); the CPU oracle is timed from the FP64
    engine's state at step nss, on the same host."""
    t_set = time.perf_counter()
    mesh, p = make_apollo(ppc, False)
    p.NSS = nss
    parts = D.fill_freestream(mesh, p, ppc=ppc, seed=seed)
    log(f"[bench] apollo: {mesh.ncell} cells, {parts.n} particles (set-up {time.perf_counter() - t_set:.1f} s)")
    out = {"workload": f"Apollo capsule surrogate hex O-grid ({mesh.ncell} cells), monatomic VHS, Ma 5, Kn 0.01, "
                       f"AoA 30, diffuse wall 300 K, inflow, Npinitial {ppc:g} ({parts.n} particles at start), "
                       f"{steps} steps, sampling from step {nss}",
           "cells": mesh.ncell, "particles_start": parts.n, "steps": steps}
    developed = None
    for tag, prec in (("fp64", D.PREC_FP64), ("mixed", D.PREC_MIXED)):
        q = p.copy()
        q.precision = prec
        eng = D.Engine(mesh, q, seed=seed)
        eng.upload(parts)
        tuning, _ = eng.autotune(1, trial_steps(parts.n))
        eng.upload(parts)
        eng.run(1, 3)
        eng.run(4, 6)
        eng.upload(parts)
        st0 = eng.stats()["particle_steps"]
        ms1 = eng.run(1, nss)
        st1 = eng.stats()["particle_steps"]
        if tag == "fp64" and cpu:
            developed = eng.download()
        ms2 = eng.run(nss + 1, steps)
        st = eng.stats()
        rec = {"value_job": round((st["particle_steps"] - st0) / ((ms1 + ms2) / 1e3), 1),
               "value_developed": round((st["particle_steps"] - st1) / (ms2 / 1e3), 1),
               "ms_per_step_developed": round(ms2 / (steps - nss), 4),
               "particles_end": int(st["particles"]), "dropped": int(st["dropped"]), "tuning": tuning}
        out[tag] = rec
        log(f"[bench] apollo {tag}: job {rec['value_job']:.4g}, developed flow {rec['value_developed']:.4g} "
            f"particle-steps/s ({rec['ms_per_step_developed']:.3f} ms/step)")
        eng.close()
    out["mixed_speedup_developed"] = round(out["mixed"]["value_developed"] / out["fp64"]["value_developed"], 3)
    if cpu:
        c = cpu_pair(mesh, p, developed,
                     f"C++ oracle, Apollo surrogate from the FP64 engine's state at step {nss} "
                     f"({developed.n} particles)", first_step=nss + 1)
        out["cpu_baseline"] = c
        out["gpu_vs_cpu_developed"] = round(out["fp64"]["value_developed"] / c["value"], 1)
        out["target"] = "FP64 engine >= 20x the host CPU on the Apollo case"
    return out


def _energy(q):
    return 0.5 * float(np.sum(q.vx ** 2 + q.vy ** 2 + q.vz ** 2, dtype=np.float64)) + float(np.sum(q.ir))


def validate_box(steps: int, n: int = 64, ppc: float = 50.0):
    """Equilibrium box at full size: total energy drift and the NTC collision
    count against the hard-sphere rate 1/2 N (N-1)/V Neff sigma <c_r> dt per
    cell and step, <c_r> = 4 sqrt(kT/(pi m)) (Bird 1994, eqs. 1.10, 4.64),
    summed over the cell counts of every step."""
    kb = 1.3806e-23
    out = {}
    for prec in (D.PREC_FP64, D.PREC_MIXED):
        mesh, p = make_case(n, ppc, prec == D.PREC_MIXED)
        parts = D.fill_freestream(mesh, p, ppc=ppc, temperature=1.0, seed=20261203)
        eng = D.Engine(mesh, p, seed=20261203)
        eng.upload(parts)
        eng.set_tuning(wide_min=-1, move_chunk=8192)
        e0 = _energy(eng.download())
        vol = mesh.arrays()["cell_volume"]
        sigma = math.pi * p.d_ref ** 2
        cr = 4.0 * math.sqrt(kb * p.T_ref / (math.pi * p.mass))
        dt = p.dtime * p.dt_ref
        want = 0.0
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            eng.run(s, s)
            cnt = eng.cell_scalars()["count"].astype(np.float64)
            want += float(np.sum(0.5 * cnt * (cnt - 1.0) / vol)) * p.Neff * sigma * cr * dt
            if s % 200 == 0:
                log(f"[validate box] prec={prec} step {s}/{steps} ({time.perf_counter() - t0:.1f} s)")
        st = eng.stats()
        q = eng.download()
        out["fp64" if prec == D.PREC_FP64 else "mixed"] = {
            "particles": int(st["particles"]), "dropped": int(st["dropped"]),
            "energy_rel_drift": (_energy(q) - e0) / e0,
            "collisions": int(st["collisions"]), "collisions_expected": want,
            "collision_rate_ratio": st["collisions"] / want,
            "collision_rate_sigma": math.sqrt(want) / want,
            "vx_kurtosis": float(np.mean(q.vx ** 4) / np.mean(q.vx ** 2) ** 2),
        }
        eng.close()
    return {"validate": "equilibrium box", "cells": mesh.ncell, "particles": parts.n, "steps": steps, **out}


def validate_apollo(steps: int, ppc: float = 100.0, nss: int = 5000, nscheme: int = 2000):
    """Apollo capsule at full size in FP64, FP64 with another seed (the
    sampling-noise floor) and mixed: drag coefficient, peak heat flux and the
    per-field relative L2 differences of the steady-averaged cell fields."""
    runs = {}
    for tag, prec, seed in (("fp64", D.PREC_FP64, 20261203), ("fp64_seed2", D.PREC_FP64, 7),
                            ("mixed", D.PREC_MIXED, 20261203)):
        mesh, p = make_apollo(ppc, prec == D.PREC_MIXED)
        p.NSS, p.NSCHEME, p.surface_stats = nss, nscheme, 1
        parts = D.fill_freestream(mesh, p, ppc=ppc, seed=seed)
        eng = D.Engine(mesh, p, seed=seed)
        eng.upload(parts)
        tuning, s = eng.autotune(1, trial_steps(parts.n))
        s += 1
        t0 = time.perf_counter()
        while s <= steps:
            last = min(steps, s + 499)
            eng.run(s, last)
            log(f"[validate apollo] {tag} step {last}/{steps} ({time.perf_counter() - t0:.1f} s)")
            s = last + 1
        wall = time.perf_counter() - t0
        fields, cd, qpeak = eng.surface_fields(steps, 3.912)
        st = eng.stats()
        runs[tag] = {"macro": eng.macro(D.MACRO_FINAL), "q": fields[:, 3] + fields[:, 4], "normal": fields[:, 5:8],
                     "uhat": np.array(list(p.u_inf)) / np.linalg.norm(list(p.u_inf)), "area": mesh.arrays()["face_area"],
                     "cd": cd, "qpeak": qpeak,
                     "particles": int(st["particles"]), "seconds": wall, "tuning": tuning,
                     "ps_per_s": st["particle_steps"] / max(wall, 1e-9)}
        eng.close()
    a = runs["fp64"]
    # stagnation region, chosen by geometry only (so that no run's sampling
    # noise picks the faces): the wall faces whose normal is most aligned
    # with the free stream, 2% of the wall faces; the wall faces are those
    # with a normal in the surface output
    wall = np.linalg.norm(a["normal"], axis=1) > 0.5
    align = np.where(wall, a["normal"] @ a["uhat"], -np.inf)
    hot = np.argsort(align)[::-1][: max(1, int(0.02 * wall.sum()))]
    area = a["area"]

    def stag(r):
        return float(np.sum(r["q"][hot] * area[hot]) / np.sum(area[hot]))

    def load(r):
        return float(np.sum(r["q"][wall] * area[wall]))
    names = ["rho", "ux", "uy", "uz", "T", "sxx", "sxy", "syy", "syz", "szz", "szx", "qx", "qy", "qz", "Tr",
             "qrx", "qry", "qrz"]
    rep = {"validate": "apollo", "cells": int(a["macro"].shape[0]), "steps": steps, "ppc": ppc, "NSS": nss,
           "NSCHEME": nscheme}
    for tag in ("fp64_seed2", "mixed"):
        b = runs[tag]
        fe = {}
        for f, nm in enumerate(names):
            den = float(np.linalg.norm(a["macro"][:, f]))
            if den > 0.0:
                fe[nm] = float(np.linalg.norm(b["macro"][:, f] - a["macro"][:, f])) / den
        rep[tag] = {"cd": b["cd"], "cd_rel_diff": (b["cd"] - a["cd"]) / a["cd"],
                    "qpeak_max_face": b["qpeak"], "qpeak_max_face_rel_diff": (b["qpeak"] - a["qpeak"]) / a["qpeak"],
                    "q_stagnation": stag(b), "q_stagnation_rel_diff": stag(b) / stag(a) - 1.0,
                    "heat_load": load(b), "heat_load_rel_diff": load(b) / load(a) - 1.0,
                    "field_rel_l2_diff": fe, "seconds": b["seconds"], "particle_steps_per_s": b["ps_per_s"],
                    "tuning": b["tuning"]}
    rep["stagnation_faces"] = int(len(hot))
    rep["fp64"] = {"cd": a["cd"], "qpeak_max_face": a["qpeak"], "q_stagnation": stag(a), "heat_load": load(a),
                   "seconds": a["seconds"], "particle_steps_per_s": a["ps_per_s"], "particles": a["particles"],
                   "tuning": a["tuning"]}
    rep["mixed_speedup"] = runs["mixed"]["ps_per_s"] / a["ps_per_s"]
    return rep


def run_sweep(args):
    """Batched sweep: `--sweep-cases` seeded capsule cases (seeds 0..N-1) in
    one batch engine, mixed precision, device fill, autotuned, steps replayed
    as HIP graphs; then every case alone, one after another, for the same
    steps (the unbatched schedule), for comparison."""
    import sweep as SW
    t0 = time.perf_counter()
    specs = [SW.case_spec(i, args.sweep_scale) for i in range(args.sweep_cases)]
    built = [SW.build_case(sp, D.PREC_MIXED, steps=2000) for sp in specs]
    log(f"[sweep] {len(specs)} cases built in {time.perf_counter() - t0:.1f} s")
    seeds = [1000 + sp["seed"] for sp in specs]
    want = sum(sp["ppc"] * m.ncell for sp, (m, _) in zip(specs, built))
    cap = int(1.25 * want) + 2_000_000 * len(specs)
    eng = D.Engine.batch([m for m, _ in built], [p for _, p in built], seeds, capacity=cap)
    n = eng.fill_device([sp["ppc"] for sp in specs], seeds)
    log(f"[sweep] batch: {eng.mesh.ncell} cells, {n} particles, {eng.footprint() / 1e9:.1f} GB "
        f"({time.perf_counter() - t0:.1f} s)")
    tuning, used = eng.autotune(1, trial_steps(n))
    s = used + 1
    if args.warmup:
        eng.run(s, s + args.warmup - 1)
        s += args.warmup
    st0 = eng.stats()
    ms = eng.run(s, s + args.steps - 1)
    st1 = eng.stats()
    ps = st1["particle_steps"] - st0["particle_steps"]
    batched = ps / (ms / 1e3)
    log(f"[sweep] batched: {batched:.4g} particle-steps/s, {ms / args.steps:.2f} ms/step, tuning {tuning}")
    eng.close()
    seq_ps, seq_ms = 0, 0.0
    if not args.no_sequential:
        for k, ((m, p), sp) in enumerate(zip(built, specs)):
            e1 = D.Engine(m, p, seed=seeds[k], capacity=int(1.25 * sp["ppc"] * m.ncell) + 2_000_000)
            e1.fill_device(sp["ppc"], seeds[k])
            e1.set_tuning(**{kk: v for kk, v in tuning.items() if kk != "graph"})
            e1.run(1, 1 + args.warmup)
            a = e1.stats()["particle_steps"]
            seq_ms += e1.run(2 + args.warmup, 1 + args.warmup + args.steps)
            seq_ps += e1.stats()["particle_steps"] - a
            e1.close()
            log(f"[sweep] case {k} alone done ({time.perf_counter() - t0:.1f} s)")
        log(f"[sweep] one case after another: {seq_ps / (seq_ms / 1e3):.4g} particle-steps/s")
    return {
        "metric": METRIC, "value": round(batched, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 velocities / f64 positions",
        "data": "synthetic (seeded capsule cases, device fill)",
        "config": {"workload": f"batched sweep of {len(specs)} capsule tet cases (seeds 0..{len(specs) - 1}, "
                               f"cell scale {args.sweep_scale:g}), Ma U[3,10], Kn logU[0.005,0.1], AoA U[0,40], "
                               "20-100 ppc, mixed precision, one engine, graph replay",
                   "cells": int(sum(m.ncell for m, _ in built)), "particles": int(n), "parallelism": "single GPU"},
        "sequential_value": None if not seq_ms else round(seq_ps / (seq_ms / 1e3), 1),
        "batched_speedup": None if not seq_ms else round(batched / (seq_ps / (seq_ms / 1e3)), 3),
        "tuning": tuning, "tune_steps": used,
    }


def run_serve(args):
    """Serving queue: `--serve-requests` seeded capsule cases arrive as a
    Poisson stream (`--serve-rate` per second), each a `--serve-steps` solve
    in mixed precision; the server admits them by HBM footprint and runs the
    admitted ones concurrently.  Latency = finish - arrival."""
    import sweep as SW
    t_set = time.perf_counter()
    specs = [SW.case_spec(i, args.serve_scale) for i in range(args.serve_requests)]
    built = [SW.build_case(sp, D.PREC_MIXED, steps=args.serve_steps) for sp in specs]
    log(f"[serve] {len(specs)} requests prepared in {time.perf_counter() - t_set:.1f} s")
    tuning = None
    if args.serve_tuning == "auto":
        # set-up, untimed: the autotuner's pick for an engine like those the
        # server builds (the first requests' cases in one batch engine), given
        # to every engine the server builds; results do not depend on it
        k = max(1, min(args.serve_batch, len(specs)))
        tb = [SW.build_case(sp, D.PREC_MIXED, steps=args.serve_steps) for sp in specs[:k]]
        seeds = [1000 + i for i in range(k)]
        want = sum(sp["ppc"] * m.ncell for sp, (m, _) in zip(specs, tb))
        cap = int(1.25 * want) + 2_000_000 * k
        eng = (D.Engine.batch([m for m, _ in tb], [p for _, p in tb], seeds, capacity=cap) if k > 1 else
               D.Engine(tb[0][0], tb[0][1], seed=seeds[0], capacity=cap))
        n = eng.fill_device([sp["ppc"] for sp in specs[:k]], seeds)
        tuning, used = eng.autotune(1, trial_steps(n))
        eng.close()
        del tb
        log(f"[serve] autotune on the first {k} requests' cases ({n} particles, {used} steps): {tuning}")
    rng = np.random.default_rng(20261203)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, len(specs)))
    arrivals -= arrivals[0]
    srv = D.Server(budget_bytes=0, chunk_steps=args.serve_chunk, max_running=args.serve_max_running,
                   batching=args.serve_batch, tuning=tuning)
    t0 = time.perf_counter()
    nxt, ids, last_log = 0, [], t0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(specs) and arrivals[nxt] <= now:
            m, p = built[nxt]
            ids.append(srv.submit(m, p, seed=1000 + nxt, ppc=specs[nxt]["ppc"], steps=args.serve_steps))
            nxt += 1
        running, queued, finished = srv.pump()
        if nxt == len(specs) and finished == len(specs):
            break
        if time.perf_counter() - last_log > 15.0:
            last_log = time.perf_counter()
            log(f"[serve] t={now:.1f} s: {nxt} arrived, {running} running, {queued} queued, {finished} finished")
        time.sleep(0.0005)
    makespan = time.perf_counter() - t0
    infos = [srv.request(i) for i in ids]
    srv.close()
    failed = [i["error"] for i in infos if i["state"] == D.REQ_FAILED]
    lat = np.array([i["t_finish"] - i["t_submit"] for i in infos])
    wait = np.array([i["t_admit"] - i["t_submit"] for i in infos])
    total_ps = float(sum(i["particle_steps"] for i in infos))
    p50, p99 = (float(x) for x in np.percentile(lat, [50, 99]))
    log(f"[serve] {len(infos)} requests in {makespan:.1f} s: p50 {p50:.2f} s, p99 {p99:.2f} s, "
        f"{total_ps / makespan:.4g} particle-steps/s, {len(failed)} failed")
    return {
        "metric": METRIC, "value": round(total_ps / makespan, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.serve_steps, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 velocities / f64 positions", "data": "synthetic (seeded capsule cases)",
        "config": {"workload": f"serving queue: {len(specs)} capsule requests (cell scale {args.serve_scale:g}), "
                               f"Poisson arrivals {args.serve_rate:g}/s, {args.serve_steps}-step solves, "
                               f"chunks of {args.serve_chunk} steps, HBM-footprint admission"
                               + (f", at most {args.serve_max_running} running" if args.serve_max_running else "")
                               + (f", up to {args.serve_batch} requests per engine" if args.serve_batch > 1 else
                                  ", one engine per request")
                               + f", {os.environ.get('GPU_MAX_HW_QUEUES', '4')} hardware queues"
                               + (f", launch settings {tuning} (autotuned in set-up)" if tuning else
                                  ", engine default launch settings")
                               + ", mixed precision",
                   "particles_initial": int(sum(i["particles"] for i in infos)), "parallelism": "single GPU"},
        "latency_s": {"p50": round(p50, 3), "p99": round(p99, 3), "max": round(float(lat.max()), 3),
                      "mean_queue_wait": round(float(wait.mean()), 3)},
        "makespan_s": round(makespan, 2), "failed": failed,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="box resolution (hexes per side)")
    ap.add_argument("--workload", choices=["box", "apollo", "sweep", "serve"], default="box")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="mean arrivals per second")
    ap.add_argument("--serve-steps", type=int, default=1000)
    ap.add_argument("--serve-chunk", type=int, default=50)
    ap.add_argument("--serve-max-running", type=int, default=8, help="requests running at once (0: no limit)")
    ap.add_argument("--serve-batch", type=int, default=4,
                    help="requests admitted together share one batch engine, up to this many")
    ap.add_argument("--serve-hw-queues", type=int, default=16,
                    help="HIP hardware queues of the serving process (sets GPU_MAX_HW_QUEUES; 0 keeps the environment's)")
    ap.add_argument("--serve-tuning", choices=["auto", "default"], default="auto",
                    help="auto: autotune once in set-up and give the pick to every engine the server builds")
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="multiplies every request's cell count (1: the sweep's 100k-1M-cell cases)")
    ap.add_argument("--sweep-cases", type=int, default=32)
    ap.add_argument("--sweep-scale", type=float, default=1.0, help="multiplies every sweep case's cell count")
    ap.add_argument("--no-sequential", action="store_true", help="sweep: skip the one-case-at-a-time comparison")
    ap.add_argument("--ppc", type=float, default=50.0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-apollo", action="store_true", help="skip the Apollo sub-record of the box workload")
    ap.add_argument("--tuning", type=json.loads, default=None,
                    help='launch settings instead of the autotuner, e.g. \'{"move_chunk": 4096, "wide_min": 64}\'')
    ap.add_argument("--validate", choices=["box", "apollo"],
                    help="run a full-size physics validation instead of the benchmark and print its report")
    ap.add_argument("--validate-steps", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB plain copy kernel (counter calibration, achievable HBM GB/s)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU per case; multi-GPU runs are not part of this build "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve" and args.serve_hw_queues > 0:
        # concurrently running engines each need a hardware queue of their
        # own (HIP's default is 4 per process); read at HIP initialisation,
        # which has not happened yet in this process
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.serve_hw_queues, 32))
    D.lib()  # fail loudly if the engine library is missing
    if args.validate == "box":
        print(json.dumps(validate_box(args.validate_steps or 2000)), flush=True)
        return
    if args.validate == "apollo":
        print(json.dumps(validate_apollo(args.validate_steps or 10000)), flush=True)
        return
    if args.workload == "sweep":
        print(json.dumps(run_sweep(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    copy_gbs = None
    if args.copy_probe:
        copy_gbs = D.copy_probe(1 << 30, 5)
        log(f"[bench] copy probe: 1 GiB x 5, {copy_gbs:.0f} GB/s (read + write)")
    t0 = time.perf_counter()
    if args.workload == "apollo":
        if args.ppc == 50.0:
            args.ppc = 20.0
        mesh, p = make_apollo(args.ppc, False)
        parts = D.fill_freestream(mesh, p, ppc=args.ppc, seed=20261203)
        ntri, label = 12.0, "apollo"
    else:
        mesh, p = make_case(args.n, args.ppc, False)
        parts = D.fill_freestream(mesh, p, ppc=args.ppc, temperature=1.0, seed=20261203)
        ntri, label = 4.0, f"box {args.n}^3x6 tets"
    log(f"[bench] {label}: {mesh.ncell} cells, {parts.n} particles (set-up {time.perf_counter() - t0:.1f} s)")
    fp64 = run_engine(mesh, p, parts, args.steps, args.warmup, autotune=not args.no_autotune, tuning=args.tuning)
    log(f"[bench] fp64: {fp64['value']:.4g} particle-steps/s, {fp64['ms_per_step']:.2f} ms/step")

    mixed = None
    if not args.no_mixed:
        pm = p.copy()
        pm.precision = D.PREC_MIXED
        mixed = run_engine(mesh, pm, parts, args.steps, args.warmup, autotune=not args.no_autotune,
                           tuning=args.tuning)
        log(f"[bench] mixed: {mixed['value']:.4g} particle-steps/s, {mixed['ms_per_step']:.2f} ms/step")

    # roofline of the dominant kernel group, from the HIP events around each
    # launch over the profiled pass (the same number of steps, right after
    # the timed ones; eager launches with an event after every kernel group)
    kt = fp64["kernels"]
    groups = {"move": kt["move"][0], "sort": sum(kt[g][0] for g in ("sort_count", "sort_scan", "sort_place",
                                                                   "sort_gather")),
              "collide_sample": kt["collide_sample"][0]}
    dom = max(groups, key=groups.get)
    per_particle = alg_bytes(dom, 64, 32, 80.0 * ntri, nmean_ppc := fp64["particle_steps"] / args.steps / mesh.ncell)
    nmean = fp64["particle_steps"] / args.steps
    launch_ms = groups[dom] / args.steps
    achieved = per_particle * nmean / (launch_ms / 1e3) / 1e9
    # HBM bytes per launch of that group, measured with rocprofv3 FETCH_SIZE /
    # WRITE_SIZE counters on this workload (profiles/traffic.json, calibrated
    # against a plain copy); null for workloads it does not cover
    traffic = None
    tkey = "box" if (args.workload == "box" and args.n == 64 and args.ppc == 50.0) else \
        ("apollo" if (args.workload == "apollo" and args.ppc == 20.0) else None)
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if tkey and os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(tkey, {}).get(dom, {}).get("traffic_bytes")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "bytes_per_particle": per_particle, "launch_ms": round(launch_ms, 4),
                "alg_bytes_per_launch": round(per_particle * nmean),
                "timing": f"HIP events on the engine's stream around each launch, over a profiled pass of "
                          f"{args.steps} steps right after the timed ones"}
    step_bytes = sum(alg_bytes(k, 64, 32, 80.0 * ntri, nmean_ppc) for k in ("move", "sort", "collide_sample"))

    cpu = None
    if not args.no_cpu:
        if args.workload == "apollo":
            what = f"C++ oracle, Apollo surrogate ({mesh.ncell} cells), Npinitial {args.ppc:g} ({parts.n} particles)"
        else:
            what = (f"C++ oracle, the same {args.n}^3x6-tet box ({mesh.ncell} cells), {args.ppc:g} ppc "
                    f"({parts.n} particles)")
        cpu = cpu_pair(mesh, p, parts, what, seconds=12.0, threads=args.cpu_threads)
    apollo = None
    if args.workload == "box" and not args.no_apollo:
        apollo = apollo_record(cpu=not args.no_cpu)

    out = {
        "metric": METRIC,
        "value": round(fp64["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Maxwellian fill of a generated mesh)",
        "config": {"workload": (f"equilibrium box {args.n}^3 hexes x 6 tets, hard spheres, Kn 0.05, "
                                f"{args.ppc:g} particles/cell, specular walls, full step "
                                "(emit+move+sort+collide+sample)") if args.workload == "box" else
                               ("Apollo capsule surrogate hex O-grid (372,500 cells), monatomic VHS, Ma 5, "
                                f"Kn 0.01, AoA 30, diffuse wall 300 K, inflow, Npinitial {args.ppc:g}, full step"),
                   "cells": mesh.ncell, "particles": parts.n, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "step_alg_GBs": round(step_bytes * nmean / (fp64["ms_per_step"] / 1e3) / 1e9, 1),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1]},
        "kernel_ms_note": "profiled pass (eager launches, an event after each kernel group), not the timed steps",
        "profiled_ms_per_step": round(fp64["profiled_ms_per_step"], 4),
        "mixed": None if mixed is None else {"value": round(mixed["value"], 1), "dtype": "f32 velocities / f64 positions",
                                             "ms_per_step": round(mixed["ms_per_step"], 4),
                                             "speedup_vs_fp64": round(mixed["value"] / fp64["value"], 3)},
        "gpu_vs_cpu": None if cpu is None else round(fp64["value"] / cpu["value"], 1),
        "footprint_GB": round(fp64["footprint_GB"], 2),
        "tuning": {"fp64": fp64["tuning"], "mixed": None if mixed is None else mixed["tuning"]},
        "copy_probe_GBs": None if copy_gbs is None else round(copy_gbs, 1),
        "timed_path": "graph replay of step pairs, no profiling",
        "apollo": apollo,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
