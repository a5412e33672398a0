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
    """The RANDOMWALK slot (imc_transport.jl:212-479).  The reference's analytic random walk
    does not conserve energy; this build runs the plain event loop here, which is the
    method the random walk approximates."""
    _run(mesh, simvars, particles)
