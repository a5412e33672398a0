"""Update.update -- src/imc_update.jl:12-70 in FP64 on the host.

beta = 4aT^3/bee (LINEARIZED: bee = 4aT^3, beta = 1); sigma = sigma0 * T^p with the
"MARSHAK WAVE" special case sigma_a = sigma0/T/T/T in 1D (imc_update.jl:32-34);
Fleck factor f = 1/(1 + sorter([s, alpha, beta, c, dt, sigma_a], [1])).
"""
from __future__ import annotations

import numpy as np

from . import Constants
from .Input import get_scalar
from .Utilities import sorter


def update(inputs: dict, mesh, simvars) -> None:
    a = Constants.phys_a
    T = mesh.temp
    T3 = T * T * T
    if get_scalar(inputs, "LINEARIZED") == "TRUE":
        mesh.bee = 4 * a * T3
        mesh.beta = np.ones_like(T)
    else:
        mesh.beta = (4 * a * T3) / mesh.bee
    with np.errstate(divide="ignore", over="ignore"):
        mesh.sigma_a[..., 0] = mesh.sigma_a[..., 1] * T ** mesh.sigma_a[..., 2]
        if simvars.geometry == "1D" and get_scalar(inputs, "NAME").upper() == "MARSHAK WAVE":
            mesh.sigma_a[..., 0] = mesh.sigma_a[..., 1] / T / T / T
        mesh.sigma_s[..., 0] = mesh.sigma_s[..., 1] * T ** mesh.sigma_s[..., 2]
    sa = mesh.sigma_a[..., 0].ravel(order="F")
    beta = np.broadcast_to(mesh.beta, T.shape).ravel(order="F")
    n = sa.size
    vals = np.stack([np.full(n, mesh.distancescale), np.full(n, Constants.alpha), beta,
                     np.full(n, Constants.phys_c), np.full(n, simvars.dt), sa], axis=1)
    prod, _, _ = sorter(vals, [1.0])
    mesh.fleck = (1.0 / (1.0 + prod)).reshape(T.shape, order="F")
