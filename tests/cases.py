"""Shared problem builders for the parity tests: one description drives both the GPU
engine (through the C ABI) and the CPU oracle."""
from __future__ import annotations

import numpy as np

import oracle as O


def case1d(C=40, prec=64, seed=11, bc=(0, 1, 0, 0), scales=(1.0,), thick=False, dt=0.05, phys_c=1.0,
           phys_a=1.0, tsurf=(1.0, 0.3), rng_seed=5):
    rng = np.random.default_rng(rng_seed)
    dx = rng.uniform(0.02, 0.08, C)
    lo, hi = (1.0, 3.0) if thick else (-1.0, 1.2)
    sa = 10 ** rng.uniform(lo, hi, C)
    ss = rng.uniform(0.0, 1.0, C) * sa * 0.5
    temp = rng.uniform(0.2, 1.0, C)
    fleck = 1.0 / (1.0 + rng.uniform(0.05, 2.0, C))
    radsource = np.where(rng.uniform(size=C) < 0.3, rng.uniform(0.5, 2.0, C), 0.0)
    return dict(geometry=1, prec=prec, nx=C, ny=1, seed=seed, bc=bc, scales=scales, dt=dt, phys_c=phys_c,
                phys_a=phys_a, dx=dx, dy=None, sa=sa, ss=ss, temp=temp, fleck=fleck, radsource=radsource,
                tsurf=[np.array([tsurf[0]]), np.array([tsurf[1]])])


def case2d(nx=12, ny=9, prec=64, seed=13, bc=(0, 1, 1, 0), scales=(1.0,), dt=0.05, phys_c=1.0, phys_a=1.0,
           rng_seed=7):
    rng = np.random.default_rng(rng_seed)
    dx = rng.uniform(0.03, 0.09, nx)
    dy = rng.uniform(0.03, 0.09, ny)
    C = nx * ny
    sa = 10 ** rng.uniform(-1.0, 1.2, C)
    ss = rng.uniform(0.0, 1.0, C) * sa * 0.5
    temp = rng.uniform(0.2, 1.0, C)
    fleck = 1.0 / (1.0 + rng.uniform(0.05, 2.0, C))
    radsource = np.where(rng.uniform(size=C) < 0.2, rng.uniform(0.5, 2.0, C), 0.0)
    tsurf = [rng.uniform(0.0, 1.0, nx) * (rng.uniform(size=nx) < 0.5), np.zeros(nx),
             rng.uniform(0.3, 1.0, ny), np.zeros(ny)]
    return dict(geometry=2, prec=prec, nx=nx, ny=ny, seed=seed, bc=bc, scales=scales, dt=dt, phys_c=phys_c,
                phys_a=phys_a, dx=dx, dy=dy, sa=sa, ss=ss, temp=temp, fleck=fleck, radsource=radsource,
                tsurf=tsurf)


def engine(case, capacity=1 << 16, quirks=0, **tuning):
    from MixedPrecisionIMC._hip import Engine

    e = Engine(geometry=case["geometry"], precision=case["prec"], nx=case["nx"], ny=case["ny"],
               energyscales=case["scales"], bc=case["bc"], seed=case["seed"], phys_c=case["phys_c"],
               phys_a=case["phys_a"], capacity=capacity, quirks=quirks, **tuning)
    e.set_geometry(case["dx"], case["dy"])
    return e


def engine_step(e, case, step):
    e.set_step_fields(step, case["dt"], case["sa"], case["ss"], case["fleck"], case["temp"], case["radsource"],
                      case["tsurf"])


def oracle_problem(case, quirks=0):
    op = O.OracleProblem(geometry=case["geometry"], prec=case["prec"], nx=case["nx"], ny=case["ny"],
                         scales=case["scales"], bc=case["bc"], seed=case["seed"], phys_c=case["phys_c"],
                         phys_a=case["phys_a"], quirks=quirks)
    st = op.step(dt=case["dt"], dx=case["dx"], dy=case["dy"], temp=case["temp"], fleck=case["fleck"],
                 sigma_a=case["sa"], sigma_s=case["ss"], radsource=case["radsource"], tsurf=case["tsurf"])
    return op, st
