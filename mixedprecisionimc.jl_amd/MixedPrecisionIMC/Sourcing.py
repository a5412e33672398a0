"""Sourcing.sourcing -- src/imc_sourcing.jl:12-370, executed by libimc_hip.

The host uploads the step's coefficients (outputs of Update.update) and the engine computes
the emission, radiation-source and surface energies, the particle budget, the per-cell counts
and the births on the GPU.  Mutates mesh.emittedenergy, mesh.totalenergy and the particle
population, as the reference does.
"""
from __future__ import annotations

import numpy as np

from ._hip import Engine


def upload_step_fields(mesh, simvars, particles: Engine) -> None:
    """imc_set_step_fields: sigma_a[:,1], sigma_s[:,1], fleck, temp, radsource, temp_surf."""
    if mesh.geometry == "1D":
        tsurf = [np.array([mesh.temp_surf[0]]), np.array([mesh.temp_surf[1]])]
    else:
        tsurf = [np.asarray(v, dtype=np.float64) for v in mesh.temp_surf]
    particles.set_step_fields(simvars.stepindex, simvars.dt, mesh.sigma_a[..., 0], mesh.sigma_s[..., 0],
                              mesh.fleck, mesh.temp, mesh.radsource, tsurf)


def sourcing(mesh, simvars, particles: Engine) -> None:
    upload_step_fields(mesh, simvars, particles)
    emitted, etot, n = particles.source(simvars.n_input, simvars.n_max, simvars.cellmin)
    K = len(mesh.energyscales)
    mesh.emittedenergy = emitted.reshape(mesh.shape + (K,), order="F")
    mesh.totalenergy = etot


_DTYPES = {16: np.float16, 32: np.float32, 64: np.float64}


def sample_planck(simvars, particles: Engine, temp=1.0, stream: int = 0) -> np.ndarray:
    """Sourcing.sample_planck (imc_sourcing.jl:372-399): frequencies from the Planck spectrum,
    returned in the run's PRECISION as the reference returns them.  `temp` (scalar or array)
    multiplies the dimensionless sample, as at the reference's commented-out call site
    (`mesh.temp[cellindex] * sample_planck()`, imc_sourcing.jl:209,232,342,362; surfaces use
    `T_surface * sample_planck()`, :171,186); one frequency per entry of
    `temp`.  Drawn on the GPU (imc_sample_planck), keyed by (seed, stream, index)."""
    t = np.atleast_1d(np.asarray(temp, dtype=np.float64))
    return particles.sample_planck(t, stream=stream).astype(_DTYPES[int(simvars.precision)]).reshape(t.shape)

