"""Batched sweep and serving on the GPU: the generated cases, the engine
footprint estimate, and the server's continuous batching, with an engine per
request and with requests sharing batch engines (each request's result
bit-identical alone and in the mix)."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import pydsmc as D

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dig-appolo_amd"))
import sweep as SW  # noqa: E402

pytestmark = pytest.mark.gpu


def _same(g, c):
    assert g.n == c.n
    for f in ("x", "y", "z", "vx", "vy", "vz", "ir", "cell", "id"):
        assert np.array_equal(getattr(g, f), getattr(c, f)), f


def test_footprint_estimate_bounds_the_engine():
    for seed in (0, 3):
        m, p = SW.build_case(SW.case_spec(seed, 0.02), D.PREC_MIXED, steps=20)
        cap = int(1.25 * p.Npinitial * m.ncell) + 2_000_000
        e = D.Engine(m, p, seed=1, capacity=cap)
        est = e.estimate(m, p, cap)
        got = e.footprint()
        print(f"[estimate] seed={seed} cells={m.ncell} est={est} got={got}", flush=True)
        assert got <= est <= 1.15 * got + (80 << 20)
        e.close()


def test_served_requests_match_alone():
    """Three requests served concurrently (chunks of 6 steps, so their runs
    interleave on the device) end bit-identical to each case run alone."""
    steps, seeds = 20, [501, 502, 503]
    specs = [SW.case_spec(i, 0.02) for i in (1, 2, 3)]
    srv = D.Server(budget_bytes=0, chunk_steps=6, auto_release=False)
    ids = []
    for sp, sd in zip(specs, seeds):
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        ids.append(srv.submit(m, p, seed=sd, ppc=sp["ppc"], steps=steps))
    while True:
        running, queued, finished = srv.pump()
        if finished == len(ids):
            break
    print(f"[serve] {[srv.request(i)['state'] for i in ids]}", flush=True)
    for rid, sp, sd in zip(ids, specs, seeds):
        info = srv.request(rid)
        assert info["state"] == D.REQ_DONE, info["error"]
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        alone = D.Engine(m, p, seed=sd)
        alone.fill_device(sp["ppc"], sd)
        alone.run(1, steps)
        served = srv.engine(rid, p)
        _same(served.download(), alone.download())
        for fam in (0, 1, 2):
            assert np.array_equal(served.accum(fam), alone.accum(fam)), fam
        assert info["particle_steps"] == alone.stats()["particle_steps"]
        alone.close()
        srv.release(rid)
    srv.close()


def test_batched_requests_match_alone():
    """With batching, the requests queued together (same step count) share
    one batch engine, so their kernels are the same launches, and each ends
    bit-identical to its case run alone; a request with another step count
    gets an engine of its own."""
    steps, seeds = 20, [701, 702, 703, 704]
    specs = [SW.case_spec(i, 0.02) for i in (7, 8, 9, 10)]
    nsteps = [steps, steps, steps, 12]
    srv = D.Server(budget_bytes=0, chunk_steps=6, auto_release=False, batching=4)
    ids = []
    for sp, sd, st in zip(specs, seeds, nsteps):
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=st)
        ids.append(srv.submit(m, p, seed=sd, ppc=sp["ppc"], steps=st))
    while srv.pump()[2] < len(ids):
        pass
    infos = [srv.request(i) for i in ids]
    print(f"[serve batched] kase {[i['kase'] for i in infos]} admits {[round(i['t_admit'], 3) for i in infos]}",
          flush=True)
    assert [i["kase"] for i in infos] == [0, 1, 2, 0]
    assert infos[0]["t_admit"] == infos[1]["t_admit"] == infos[2]["t_admit"]
    for rid, info, sp, sd, st in zip(ids, infos, specs, seeds, nsteps):
        assert info["state"] == D.REQ_DONE, info["error"]
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=st)
        alone = D.Engine(m, p, seed=sd)
        alone.fill_device(sp["ppc"], sd)
        alone.run(1, st)
        served = srv.engine(rid, p)
        k = info["kase"]
        ci = served.case_info(k)
        _same(served.case_particles(served.download(), k), alone.download())
        for fam in (0, 1, 2):
            assert np.array_equal(served.accum(fam)[ci["cell_begin"]:ci["cell_end"]], alone.accum(fam)), fam
        assert info["particle_steps"] == alone.stats()["particle_steps"]
        alone.close()
    for rid in ids:
        srv.release(rid)
    srv.close()


def test_server_tuning_reaches_its_engines():
    """dsmc_server_set_tuning: every engine the server builds afterwards gets
    the launch settings (checked when they are set), and the results stay
    bit-identical to each case run alone with the engine defaults."""
    steps, seeds = 12, [801, 802]
    specs = [SW.case_spec(i, 0.02) for i in (11, 12)]
    want = {"wide_min": 32, "wide_grid": 512, "move_chunk": 512, "graph": 0, "sort_map": 1, "coll_block": 128}
    srv = D.Server(budget_bytes=0, chunk_steps=5, auto_release=False, batching=2)
    with pytest.raises(RuntimeError, match="coll_block"):
        srv.set_tuning(dict(want, coll_block=100))
    srv.set_tuning(want)
    ids = []
    for sp, sd in zip(specs, seeds):
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        ids.append(srv.submit(m, p, seed=sd, ppc=sp["ppc"], steps=steps))
    while srv.pump()[2] < len(ids):
        pass
    for rid, sp, sd in zip(ids, specs, seeds):
        info = srv.request(rid)
        assert info["state"] == D.REQ_DONE, info["error"]
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        served = srv.engine(rid, p)
        got = served.tuning()
        print(f"[serve tuning] request {rid}: {got}", flush=True)
        assert {k: got[k] for k in want} == want
        alone = D.Engine(m, p, seed=sd)
        alone.fill_device(sp["ppc"], sd)
        alone.run(1, steps)
        k = info["kase"]
        _same(served.case_particles(served.download(), k), alone.download())
        alone.close()
    for rid in ids:
        srv.release(rid)
    srv.set_tuning(None)
    srv.close()


def test_running_cap_holds_requests_in_the_queue():
    """With at most one request running, three submitted requests run one
    after another in arrival order (each admitted only after the previous
    one finished), and all finish."""
    steps = 8
    specs = [SW.case_spec(i, 0.02) for i in (4, 5, 6)]
    srv = D.Server(budget_bytes=0, chunk_steps=4, max_running=1)
    ids = []
    for k, sp in enumerate(specs):
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        ids.append(srv.submit(m, p, seed=600 + k, ppc=sp["ppc"], steps=steps))
    while True:
        running, queued, finished = srv.pump()
        assert running <= 1
        if finished == len(ids):
            break
    infos = [srv.request(i) for i in ids]
    print(f"[serve cap] admits {[round(i['t_admit'], 3) for i in infos]}", flush=True)
    assert all(i["state"] == D.REQ_RELEASED for i in infos), [i["error"] for i in infos]
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit"] >= a["t_finish"]
    srv.close()


def test_cancelled_request_leaves_its_batch_mates_running():
    """A queued request released before admission is cancelled; with
    batching on, the requests admitted around it share one engine and
    finish."""
    steps = 8
    specs = [SW.case_spec(i, 0.02) for i in (11, 12, 13, 14)]
    srv = D.Server(budget_bytes=0, chunk_steps=4, auto_release=False, batching=4)
    ids = []
    for k, sp in enumerate(specs):
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=steps)
        ids.append(srv.submit(m, p, seed=900 + k, ppc=sp["ppc"], steps=steps))
    srv.release(ids[1])
    while srv.pump()[2] < len(ids):
        pass
    infos = [srv.request(i) for i in ids]
    print(f"[serve cancel] states {[i['state'] for i in infos]} kase {[i['kase'] for i in infos]}", flush=True)
    assert infos[1]["state"] == D.REQ_FAILED and "cancelled" in infos[1]["error"]
    for k in (0, 2, 3):
        assert infos[k]["state"] == D.REQ_DONE, infos[k]["error"]
        assert infos[k]["steps_done"] == steps and infos[k]["particle_steps"] > 0
    assert [infos[k]["kase"] for k in (0, 2, 3)] == [0, 1, 2]
    for k in (0, 2, 3):
        srv.release(ids[k])
    srv.close()


def test_fp64_served_requests_match_the_oracle():
    """FP64 requests served concurrently (the first two sharing a batch
    engine, the third in its own) end bit-identical to the oracle run on
    each case: the server's device fill, batching and chunked asynchronous
    runs change nothing of the reference path."""
    steps, seeds = 8, [811, 812, 813]
    specs = [SW.case_spec(i, 0.01) for i in (21, 22, 23)]
    srv = D.Server(budget_bytes=0, chunk_steps=3, auto_release=False, batching=2)
    ids = []
    for sp, sd in zip(specs, seeds):
        m, p = SW.build_case(sp, D.PREC_FP64, steps=steps)
        ids.append(srv.submit(m, p, seed=sd, ppc=sp["ppc"], steps=steps))
    while srv.pump()[2] < len(ids):
        pass
    infos = [srv.request(i) for i in ids]
    print(f"[serve fp64 vs oracle] kase {[i['kase'] for i in infos]} particles {[i['particles'] for i in infos]}",
          flush=True)
    assert [i["kase"] for i in infos] == [0, 1, 0]
    for rid, info, sp, sd in zip(ids, infos, specs, seeds):
        assert info["state"] == D.REQ_DONE, info["error"]
        m, p = SW.build_case(sp, D.PREC_FP64, steps=steps)
        orc = O.Oracle(m, p, seed=sd)
        orc.set_particles(D.fill_freestream(m, p, ppc=sp["ppc"], seed=sd))
        for s in range(1, steps + 1):
            orc.step(s)
        served = srv.engine(rid, p)
        k = info["kase"]
        ci = served.case_info(k)
        got = served.case_particles(served.download(), k)
        _same(got, orc.particles())
        assert np.array_equal(served.accum(0)[ci["cell_begin"]:ci["cell_end"]], orc.accum(0))
        assert info["particle_steps"] == orc.stats()["particle_steps"]
    for rid in ids:
        srv.release(rid)
    srv.close()
