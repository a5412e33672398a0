"""MixedPrecisionIMC -- host side of the MI355X particle engine.

Mirrors the module layout and main loop of MixedPrecisionIMC.jl (src/MixedPrecisionIMC.jl):
``main([deck])`` reads a deck with the reference's keywords, builds the mesh and runs

    Update.update -> Sourcing.sourcing -> Transport.MC|MC2D|MC_RW -> Clean.clean
    -> Tally.tally -> EnergyCheck.energychecker -> Output.plotting -> timestep

per step (MixedPrecisionIMC.jl:136-176).  Deck parsing, mesh, Fleck factor, temperature
update and energy check run on the host in FP64; sourcing, transport, clean and the census
tally run in libimc_hip on the GPU.  ``particles`` is the device-resident population
(an ``Engine``) instead of a Vector of records.

``step_device`` (``run(..., device=True)``) runs the same loop body as one ``imc_step``: the
FP64 update, temperature update and energy sums then run on the GPU as well, and the host
keeps the clock, the saved histories and the energy check's final arithmetic.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import Clean, Constants, EnergyCheck, Input, Mesh, Output, Sourcing, Tally, Transport, Update, Utilities
from ._hip import REFLECT, VACUUM, Engine, ImcError

__all__ = ["main", "run", "step", "step_device", "timestep", "SimVars", "Engine", "ImcError", "Input", "Constants", "Mesh",
           "Update", "Sourcing", "Transport", "Clean", "Tally", "EnergyCheck", "Output", "Utilities"]


@dataclass
class SimVars:
    """MixedPrecisionIMC.jl:35-51 plus ``stepindex`` (Philox key of the step) and the
    per-step energy-balance errors."""
    t: float
    dt: float
    dt0: float
    k: float
    dtmax: float
    t_end: float
    timesteps: List[float]
    iterations: int
    n_input: int
    n_max: int
    cellmin: float
    pairwise: str
    BC: tuple
    precision: int
    geometry: str
    stepindex: int = 0
    energy_errors: List[float] = field(default_factory=list)


def timestep(timestepping: str, simvars: SimVars) -> None:
    """MixedPrecisionIMC.jl:181-222 (t kept in FP64)."""
    simvars.stepindex += 1
    if simvars.t == simvars.t_end:
        simvars.t = simvars.t_end + simvars.dt
        return
    if timestepping == "RAMP" and simvars.dt < simvars.dtmax:
        simvars.dt = simvars.dt * simvars.k
        if simvars.dt > simvars.dtmax:
            simvars.dt = simvars.dtmax
    if simvars.t + simvars.dt > simvars.t_end:
        simvars.dt = simvars.t_end - simvars.t
        simvars.t = simvars.t_end
    else:
        simvars.t = simvars.t + simvars.dt
    simvars.timesteps.append(simvars.t)


def _bc(name: str) -> int:
    name = name.strip().upper()
    if name == "REFLECT":
        return REFLECT
    if name == "VACUUM":
        return VACUUM
    raise ValueError(f"unknown boundary condition {name!r} (expected REFLECT or VACUUM)")


def setup(inputs: dict, *, precision: Optional[int] = None, seed: Optional[int] = None,
          capacity: Optional[int] = None, quirks: int = 0, tuning: Optional[dict] = None):
    """Builds (mesh, simvars, particles) from a parsed deck (MixedPrecisionIMC.jl:59-134)."""
    Constants.set_constants(inputs)
    mesh = Mesh.mesh_generation(inputs)
    prec = int(precision if precision is not None else inputs["PRECISION"])
    sd = int(seed if seed is not None else int(Input.get_scalar(inputs, "SEED")))
    ts = Input.get_scalar(inputs, "TIMESTEPPING").upper()
    if ts == "CONSTANT":
        dt = float(Input.get_scalar(inputs, "DT"))
        dt0 = k = dtmax = 0.0
    elif ts == "RAMP":
        dt0 = float(Input.get_scalar(inputs, "DT0"))
        k = float(Input.get_scalar(inputs, "K"))
        dtmax = float(Input.get_scalar(inputs, "DTMAX"))
        dt = dt0
    else:
        raise ValueError(f"unknown TIMESTEPPING {ts!r}")
    t_end = float(Input.get_scalar(inputs, "ENDTIME"))
    n_input = int(float(Input.get_scalar(inputs, "NINPUT")))
    n_max = int(float(Input.get_scalar(inputs, "NMAX")))
    cellmin = float(Input.get_scalar(inputs, "CELLMIN"))
    geometry = Input.get_scalar(inputs, "GEOMETRY").upper()
    if geometry == "1D":
        BC = (_bc(Input.get_scalar(inputs, "LEFTBC")), _bc(Input.get_scalar(inputs, "RIGHTBC")), VACUUM, VACUUM)
        nx, ny = mesh.Ncells, 1
    else:
        BC = tuple(_bc(Input.get_scalar(inputs, key)) for key in ("LEFTBC", "RIGHTBC", "TOPBC", "BOTTOMBC"))
        nx, ny = mesh.Ncells
    simvars = SimVars(t=0.0, dt=dt, dt0=dt0, k=k, dtmax=dtmax, t_end=t_end, timesteps=[], iterations=0,
                      n_input=n_input, n_max=n_max, cellmin=cellmin,
                      pairwise=str(inputs.get("PAIRWISE", "FALSE")), BC=BC, precision=prec, geometry=geometry)
    C = nx * ny
    cap = capacity if capacity is not None else int(1.25 * (n_max + n_input) + 4 * C * (cellmin + 1) + 4096)
    particles = Engine(geometry=1 if geometry == "1D" else 2, precision=prec, nx=nx, ny=ny,
                       energyscales=mesh.energyscales, bc=BC, seed=sd, phys_c=Constants.phys_c,
                       phys_a=Constants.phys_a, distancescale=mesh.distancescale, capacity=cap, quirks=quirks,
                       **(tuning or {}))
    particles.set_geometry(mesh.dx, mesh.dy if geometry == "2D" else None)
    return mesh, simvars, particles


def step(inputs: dict, mesh, simvars: SimVars, particles: Engine, timestepping: str) -> float:
    """One pass of the main loop body (MixedPrecisionIMC.jl:137-149 / :166-174)."""
    Update.update(inputs, mesh, simvars)
    Sourcing.sourcing(mesh, simvars, particles)
    if simvars.geometry == "2D":
        Transport.MC2D(mesh, simvars, particles)
    elif str(inputs.get("RANDOMWALK", "FALSE")).strip().upper() == "TRUE":
        Transport.MC_RW(mesh, simvars, None, particles)
    else:
        Transport.MC(mesh, simvars, particles)
    Clean.clean(particles)
    Tally.tally(inputs, mesh, simvars, particles)
    err = EnergyCheck.energychecker(inputs, mesh, simvars, particles)
    Output.plotting(inputs, mesh, simvars)
    timestep(timestepping, simvars)
    return err


def _set_device_material(inputs: dict, mesh, simvars: SimVars, particles: Engine) -> None:
    """imc_set_material from the mesh: sigma = coef * T^power (mesh.sigma_*[..., 1:3]), bee,
    temperature, radsource and surface temperatures, with the Update.update switches."""
    if mesh.geometry == "1D":
        tsurf = [np.array([mesh.temp_surf[0]]), np.array([mesh.temp_surf[1]])]
    else:
        tsurf = [np.asarray(v, dtype=np.float64) for v in mesh.temp_surf]
    linearized = Input.get_scalar(inputs, "LINEARIZED") == "TRUE"
    marshak = simvars.geometry == "1D" and Input.get_scalar(inputs, "NAME").upper() == "MARSHAK WAVE"
    particles.set_material(mesh.sigma_a[..., 1], mesh.sigma_a[..., 2], mesh.sigma_s[..., 1], mesh.sigma_s[..., 2],
                           None if linearized else mesh.bee, mesh.temp, mesh.radsource, tsurf, Constants.alpha,
                           linearized, marshak)
    particles.material_set = True


def step_device(inputs: dict, mesh, simvars: SimVars, particles: Engine, timestepping: str) -> float:
    """The loop body of ``step`` as one device-resident imc_step (Update.update, sourcing,
    transport, clean, Tally.tally and the energy sums on the GPU); the host keeps the clock,
    the saved histories and the energy check (imc_energycheck.jl:34).  The per-step coefficient
    arrays (fleck, sigma_a[..., 0], energydep, emittedenergy) stay on the device."""
    if not getattr(particles, "material_set", False):
        _set_device_material(inputs, mesh, simvars, particles)
    if (str(inputs.get("RANDOMWALK", "FALSE")).strip().upper() == "TRUE" and simvars.geometry == "1D"
            and not getattr(particles, "surrogate_on", False) and not getattr(particles, "random_walk_on", False)):
        particles.random_walk_enable(True)          # the MC_RW slot, as Transport.MC_RW fills it
    temp, rad, mat, r = particles.step(simvars.stepindex, simvars.dt, simvars.t == 0.0, simvars.n_input,
                                       simvars.n_max, simvars.cellmin)
    mesh.temp = temp.reshape(mesh.shape, order="F")
    mesh.radenergydens = rad.reshape(mesh.shape, order="F")
    mesh.matenergydens = mat.reshape(mesh.shape, order="F")
    mesh.totalenergy = r["totalenergy"]
    mesh.totalenergydep = r["totalenergydep"]
    mesh.lostenergy += r["lostenergy"]
    simvars.iterations += r["segments"]
    mesh.energyincrease_saved.append(particles.step_energy_increase().reshape(mesh.shape, order="F"))
    # the three fields are new arrays every step and are rebound, never written in place, so
    # the saved histories keep them without a copy
    mesh.temp_saved.append(mesh.temp)
    mesh.matenergy_saved.append(mesh.matenergydens)
    mesh.radenergy_saved.append(mesh.radenergydens)
    err = EnergyCheck.energychecker(inputs, mesh, simvars, particles, radenergy=r["radenergy"])
    Output.plotting(inputs, mesh, simvars)
    timestep(timestepping, simvars)
    return err


def run(deck, *, precision: Optional[int] = None, seed: Optional[int] = None, max_steps: int = 0,
        quirks: int = 0, verbose: bool = False, overrides: Optional[dict] = None, device: bool = False):
    """Runs a deck (path or parsed dict) to ENDTIME (or ``max_steps``) and returns
    (mesh, simvars, particles, info).  device=True runs each step as one imc_step."""
    inputs = Input.readInputs(deck) if isinstance(deck, str) else dict(deck)
    if overrides:
        inputs.update(overrides)
    mesh, simvars, particles = setup(inputs, precision=precision, seed=seed, quirks=quirks)
    timestepping = Input.get_scalar(inputs, "TIMESTEPPING").upper()
    t0 = time.perf_counter()
    nsteps = 0
    while simvars.t <= simvars.t_end and (max_steps <= 0 or nsteps < max_steps):
        err = (step_device if device else step)(inputs, mesh, simvars, particles, timestepping)
        nsteps += 1
        if verbose:
            print(f"step {nsteps} t={simvars.t:.6g} n={len(particles)} energy error {err:.3e}", flush=True)
    info = {"steps": nsteps, "wall_s": time.perf_counter() - t0, "timing": particles.timing()}
    return mesh, simvars, particles, info


def main(args) -> None:
    """MixedPrecisionIMC.jl:59-179: ``main([deck])``."""
    if not args:
        print("No input file provided, exiting... ")
        return
    print("Input file: ", args[0])
    mesh, simvars, particles, info = run(args[0], verbose=True)
    print(f"The simulation took {info['wall_s']} seconds to run ")


if __name__ == "__main__":
    import sys
    main(sys.argv[1:])
