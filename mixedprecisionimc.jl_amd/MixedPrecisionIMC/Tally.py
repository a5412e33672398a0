"""Tally.tally -- src/imc_tally.jl:11-143.

Host part in FP64: initial material energy density at t = 0, energy increase
nrg_inc = sum_k (energydep_k - emittedenergy_k)/scale_k, total deposition, material energy
and temperature update (LINEARIZED: T = U^(1/4); otherwise T += nrg_inc/bee).  The census
radiation energy density (imc_tally.jl:157-190) comes from the GPU.
"""
from __future__ import annotations

import numpy as np

from . import Constants
from .Input import get_scalar
from .Utilities import sorter
from ._hip import Engine


def tally(inputs: dict, mesh, simvars, particles: Engine) -> None:
    if simvars.t == 0.0:
        T = mesh.temp.ravel(order="F")
        n = T.size
        vals = np.stack([mesh.fleck.ravel(order="F"), mesh.sigma_a[..., 0].ravel(order="F"),
                         np.full(n, Constants.phys_a), np.full(n, Constants.phys_c), T, T, T, T,
                         np.full(n, simvars.dt), np.full(n, mesh.distancescale)], axis=1)
        prod, _, _ = sorter(vals, [1.0])
        mesh.matenergydens = prod.reshape(mesh.shape, order="F")
    mesh.totalenergydep = 0.0
    nrg_inc = np.zeros(mesh.shape)
    vol = mesh.dx if mesh.geometry == "1D" else np.outer(mesh.dx, mesh.dy)
    for k, scale in enumerate(mesh.energyscales):
        nrg_inc = nrg_inc + (mesh.energydep[..., k] - mesh.emittedenergy[..., k]) / scale
        mesh.totalenergydep += float(np.sum(mesh.energydep[..., k] * vol / scale))
    mesh.energyincrease_saved.append(nrg_inc)
    mesh.matenergydens = mesh.matenergydens + nrg_inc
    if get_scalar(inputs, "LINEARIZED") == "TRUE":
        with np.errstate(invalid="ignore"):
            mesh.temp = mesh.matenergydens ** 0.25
    else:
        mesh.temp = mesh.temp + nrg_inc / mesh.bee
    mesh.radenergydens = particles.tally_census().reshape(mesh.shape, order="F")
    mesh.temp_saved.append(mesh.temp.copy())
    mesh.matenergy_saved.append(mesh.matenergydens.copy())
    mesh.radenergy_saved.append(mesh.radenergydens.copy())
