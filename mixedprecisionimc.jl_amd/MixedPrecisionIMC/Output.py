"""Output.plotting -- src/imc_output.jl:8-80, reduced to the saved arrays.

Plots.jl figures are not reproduced.  With SAVEVARS = TRUE the arrays the reference writes
to JLD2 at t_end (timesteps, centers, temp_saved, radenergy_saved, matenergy_saved,
imc_output.jl:51-57) are written to ``<outdir>/<NAME>_<PRECISION>.npz``.
"""
from __future__ import annotations

import os

import numpy as np

from .Input import get_scalar


def plotting(inputs: dict, mesh, simvars, outdir: str = "outputs") -> None:
    if simvars.t != simvars.t_end:
        return
    if str(inputs.get("SAVEVARS", "FALSE")).strip() != "TRUE":
        return
    os.makedirs(outdir, exist_ok=True)
    name = f"{get_scalar(inputs, 'NAME')}_{inputs['PRECISION']}.npz"
    centers = mesh.centers if mesh.geometry == "1D" else np.concatenate(mesh.centers)
    np.savez(os.path.join(outdir, name), timesteps=np.asarray(simvars.timesteps), centers=centers,
             temp_saved=np.asarray(mesh.temp_saved), radenergy_saved=np.asarray(mesh.radenergy_saved),
             matenergy_saved=np.asarray(mesh.matenergy_saved))
