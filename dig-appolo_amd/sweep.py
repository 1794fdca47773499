"""Seeded capsule cases for batched sweeps and the serving queue.

A case is a spherical-cap heat shield + 33 degree cone after-body (diameter
1 m) in a tet O-grid of 100k-1M cells (20 hex O-grid layers split 12 tets per
hex, interior nodes jittered; the surface resolution sets the cell count, so
the near-wall cells stay about as thick as they are wide and no tet of the
split is inverted), monatomic VHS gas, Mach uniform in [3, 10],
Kn log-uniform in [0.005, 0.1], angle of attack uniform in [0, 40] degrees,
20-100 particles per cell, diffuse wall at T_w, free-stream inflow, mixed
precision.  Case i is drawn from seed i, so a sweep or a request stream is
reproducible.
"""
from __future__ import annotations

import math

import numpy as np

import pydsmc as D


def case_spec(seed: int, scale: float = 1.0, Tw: float | None = None) -> dict:
    """Parameters of case `seed`; `scale` multiplies the cell count."""
    r = np.random.default_rng(seed)
    cells = r.uniform(1.0e5, 1.0e6) * scale
    # 12 tets per hex, `layers` hex layers over a closed surface of box
    # topology a x a x c (2 a^2 + 4 a c quads, c = 2.48 a); with 20 layers
    # of growth ratio 1.08 the first layer is about as thick as a surface
    # quad is wide
    layers = 20
    a = max(2, int(round(math.sqrt(cells / (12 * layers * (2 + 4 * 2.48))))))
    c = int(round(2.48 * a))
    quads = 2 * a * a + 4 * a * c
    if a < 5:
        # small (scaled-down) cases: coarser surfaces keep fewer layers
        a, c = 5, 12
        quads = 2 * a * a + 4 * a * c
        layers = max(2, int(round(cells / (12 * quads))))
    return {
        "seed": int(seed),
        "a": a,
        "c": c,
        "layers": layers,
        "cells": 12 * quads * layers,
        "Ma": float(r.uniform(3.0, 10.0)),
        "Kn": float(math.exp(r.uniform(math.log(0.005), math.log(0.1)))),
        "aoa": float(r.uniform(0.0, 40.0)),
        "ppc": int(r.integers(20, 101)),
        "Tw": float(r.uniform(250.0, 500.0)) if Tw is None else float(Tw),
    }


def build_case(spec: dict, precision: int = D.PREC_MIXED, steps: int = 2000):
    """Mesh and derived parameters of a case; sampling windows sized to `steps`."""
    mesh = D.Mesh.capsule(spec["a"], spec["c"], spec["layers"], diameter=1.0, tets=True, jitter=0.1,
                          seed=spec["seed"])
    p = D.params_apollo(monatomic=True)
    p.Ma = spec["Ma"]
    p.Kn = spec["Kn"]
    p.aoa_deg = spec["aoa"]
    p.L0 = 1.0
    p.T_wall = spec["Tw"]
    p.Npinitial = int(spec["ppc"])
    p.precision = precision
    p.NSS = max(1, steps // 2)
    p.NSCHEME = max(1, min(2000, steps // 2))
    D.derive(mesh, p)
    return mesh, p
