"""Shared helpers of the full-run physics tests."""
from __future__ import annotations

import os

import numpy as np

from MixedPrecisionIMC import Constants, Input, Mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def suolson_inputs(prec=64, n_input=50000, n_max=500000):
    """The reference's Su-Olson deck run to t = 1 with its own dt = 0.002: its benchmark
    table (SuOlson.txt:71-72) is the radiation energy density at that time."""
    inp = Input.readInputs(os.path.join(GOLD, "SuOlson.txt"))
    inp["PRECISION"] = prec
    inp["ENDTIME"] = "1.0"
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(n_max)
    return inp


def bench_table(inp):
    return np.array(inp["XBENCH"]), np.array(inp["YBENCH"])


def mesh_of(inp):
    Constants.set_constants(inp)
    return Mesh.mesh_generation(inp)


def small_marshak(prec=32, n_input=200000, n_max=1000000, steps=None):
    inp = Input.readInputs(os.path.join(ROOT, "decks", "c2_marshak_fp32.txt"))
    inp["PRECISION"] = prec
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(n_max)
    return inp


def infinite_medium(prec=64, n_input=20000, n_max=200000):
    inp = Input.readInputs(os.path.join(GOLD, "InfiniteMedium.txt"))
    inp["PRECISION"] = prec
    inp["ENDTIME"] = "0.01"
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(n_max)
    return inp


def channel_small(prec=64):
    """A 2D crooked-pipe layout at reduced size (the reference's CrookedPipe deck with a
    constant dt and few particles)."""
    inp = Input.readInputs(os.path.join(GOLD, "CrookedPipe.txt"))
    inp["PRECISION"] = prec
    inp["TIMESTEPPING"] = "CONSTANT"
    inp["DT"] = "0.01"
    inp["ENDTIME"] = "0.1"
    inp["NINPUT"] = "20000"
    inp["NMAX"] = "100000"
    inp["CELLMIN"] = "1"
    return inp


def z_scores(a, b, rel_floor=0.0):
    """Per-cell z of the difference of two independent sample means (rows = seeds).
    rel_floor adds a deterministic precision floor (rel_floor * |mean|) to the standard
    error: cells that no radiation reaches have ~zero seed-to-seed variance, so their
    difference is the rounding of the tracking precision, not noise."""
    a, b = np.asarray(a), np.asarray(b)
    va = a.var(axis=0, ddof=1) / a.shape[0]
    vb = b.var(axis=0, ddof=1) / b.shape[0]
    den = np.sqrt(va + vb + (rel_floor * np.abs(b.mean(0))) ** 2)
    ok = den > 0
    z = np.zeros(a.shape[1])
    z[ok] = (a.mean(0)[ok] - b.mean(0)[ok]) / den[ok]
    return z, ok


def smoothed_at(x_points, centers, values, half_window=3):
    """Mean of the cells within +-half_window cells of each point (the table is smooth,
    so a 7-cell window lowers the MC noise without visible bias)."""
    out = []
    for x in x_points:
        i = int(np.argmin(np.abs(centers - x)))
        lo, hi = max(0, i - half_window), min(len(values), i + half_window + 1)
        out.append(float(np.mean(values[lo:hi])))
    return np.array(out)
