"""Output.plotting -- src/imc_output.jl:8-80, reduced to the saved arrays and the benchmark
comparison.

Plots.jl figures are not reproduced.  With SAVEVARS = TRUE the arrays the reference writes
to JLD2 at t_end (timesteps, centers, temp_saved, radenergy_saved, matenergy_saved,
imc_output.jl:51-57) are written to ``<outdir>/<NAME>_<PRECISION>.npz``.  With
BENCHMARK = TRUE the curve the reference draws over its plot (the deck's XBENCH / YBENCH
table against the PLOTVARS field, imc_output.jl:19-21, 27-29, 35-37) is compared in numbers
at t_end: printed, and stored in the same file when SAVEVARS = TRUE.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from .Input import get_scalar

# PLOTVARS -> the mesh field the reference plots (imc_output.jl:18, 26, 34)
PLOT_FIELDS = {"TEMPERATURE": "temp", "RADENERGY": "radenergydens", "MATENERGY": "matenergydens"}


def benchmark_comparison(inputs: dict, mesh) -> Optional[dict]:
    """The deck's benchmark table against this run, in 1D: the PLOTVARS field at the cell
    centres, linearly interpolated at each XBENCH abscissa (as the reference's two curves
    are read against each other), with the differences.  None when BENCHMARK is not TRUE or
    the geometry is 2D (the reference draws no benchmark curve on its heat map)."""
    if str(inputs.get("BENCHMARK", "FALSE")).strip().upper() != "TRUE" or mesh.geometry != "1D":
        return None
    plotvar = str(get_scalar(inputs, "PLOTVARS")).strip().upper()
    if plotvar not in PLOT_FIELDS:
        return None
    xb = np.asarray(inputs["XBENCH"], dtype=np.float64).ravel()
    yb = np.asarray(inputs["YBENCH"], dtype=np.float64).ravel()
    if xb.size != yb.size or xb.size == 0:
        raise ValueError(f"XBENCH and YBENCH must have the same, non-zero length ({xb.size}, {yb.size})")
    values = np.asarray(getattr(mesh, PLOT_FIELDS[plotvar]), dtype=np.float64).ravel()
    centers = np.asarray(mesh.centers, dtype=np.float64).ravel()
    imc = np.interp(xb, centers, values)
    diff = imc - yb
    nz = np.abs(yb) > 0
    rel = np.full(yb.shape, np.nan)
    rel[nz] = diff[nz] / yb[nz]
    return {"field": PLOT_FIELDS[plotvar], "x": xb, "benchmark": yb, "imc": imc, "diff": diff, "rel_diff": rel,
            "max_abs_diff": float(np.max(np.abs(diff))),
            "rms_rel_diff": float(np.sqrt(np.mean(rel[nz] ** 2))) if nz.any() else float("nan")}


def plotting(inputs: dict, mesh, simvars, outdir: str = "outputs") -> None:
    if simvars.t != simvars.t_end:
        return
    cmp = benchmark_comparison(inputs, mesh)
    if cmp is not None:
        print(f"Benchmark comparison ({cmp['field']} at {cmp['x'].size} points): max |IMC - benchmark| = "
              f"{cmp['max_abs_diff']:.4g}, RMS relative difference = {cmp['rms_rel_diff']:.4g}")
    if str(inputs.get("SAVEVARS", "FALSE")).strip() != "TRUE":
        return
    os.makedirs(outdir, exist_ok=True)
    name = f"{get_scalar(inputs, 'NAME')}_{inputs['PRECISION']}.npz"
    centers = mesh.centers if mesh.geometry == "1D" else np.concatenate(mesh.centers)
    extra = {}
    if cmp is not None:
        extra = {"xbench": cmp["x"], "ybench": cmp["benchmark"], "imc_at_xbench": cmp["imc"]}
    np.savez(os.path.join(outdir, name), timesteps=np.asarray(simvars.timesteps), centers=centers,
             temp_saved=np.asarray(mesh.temp_saved), radenergy_saved=np.asarray(mesh.radenergy_saved),
             matenergy_saved=np.asarray(mesh.matenergy_saved), **extra)
