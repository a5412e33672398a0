"""Transport -- src/imc_transport.jl, executed by libimc_hip.

MC (1D, :13-210) and MC2D (2D, :483-732) run the same GPU event loop; the geometry was fixed
when the engine was created.  Both overwrite mesh.energydep, add this step's vacuum loss to
mesh.lostenergy and the event-loop passes to simvars.iterations (as :73 does).
"""
from __future__ import annotations

from ._hip import Engine


def _run(mesh, simvars, particles: Engine) -> None:
    dep, lost, segs = particles.transport()
    K = len(mesh.energyscales)
    mesh.energydep = dep.reshape(mesh.shape + (K,), order="F")
    mesh.lostenergy += lost
    simvars.iterations += segs


def MC(mesh, simvars, particles: Engine) -> None:
    _run(mesh, simvars, particles)


def MC2D(mesh, simvars, particles: Engine) -> None:
    _run(mesh, simvars, particles)


def MC_RW(mesh, simvars, rwvars, particles: Engine) -> None:
    """The RANDOMWALK slot (imc_transport.jl:212-479).  The reference's random walk does not
    conserve energy (quirk Q6); here the slot is filled by the engine's corrected analytic
    random walk (imc_random_walk_enable) unless the learned surrogate has been enabled
    (Surrogate.train / Surrogate.load), which then takes its place."""
    if not getattr(particles, "surrogate_on", False) and not getattr(particles, "random_walk_on", False):
        particles.random_walk_enable(True)
    _run(mesh, simvars, particles)
