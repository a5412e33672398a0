"""EnergyCheck.energychecker -- src/imc_energycheck.jl:10-38 (host FP64).

error = (E_tot - E_dep - (E_rad - E_rad_old) - E_lost) / E_tot; stores E_rad as the next
step's E_rad_old and resets the lost energy.
"""
from __future__ import annotations

import numpy as np


def energychecker(inputs: dict, mesh, simvars, particles=None, radenergy=None) -> float:
    """radenergy: the radiation energy when already summed (the device-resident step forms it)."""
    if radenergy is not None:
        radenergy = float(radenergy)
    elif mesh.geometry == "1D":
        radenergy = float(np.sum(mesh.radenergydens * mesh.dx))
    else:
        radenergy = float(np.sum(mesh.radenergydens * mesh.dx[:, None] * mesh.dy[None, :]))
    err = (mesh.totalenergy - mesh.totalenergydep - (radenergy - mesh.radenergyold) - mesh.lostenergy) / mesh.totalenergy
    mesh.radenergyold = radenergy
    mesh.lostenergy = 0.0
    simvars.energy_errors.append(err)
    return err
