"""ctypes binding of the CPU oracle (liboracle_imc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product.  See imc_oracle.h for the specification.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_imc.so")

VACUUM, REFLECT = 0, 1
Q1, Q2 = 1, 2

_DP = ctypes.POINTER(ctypes.c_double)


class Problem(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int32), ("prec", ctypes.c_int32), ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32), ("K", ctypes.c_int32), ("quirks", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4), ("seed", ctypes.c_uint32), ("pad_", ctypes.c_int32),
        ("scales", ctypes.c_double * 16), ("phys_c", ctypes.c_double), ("phys_a", ctypes.c_double),
        ("distancescale", ctypes.c_double),
    ]


class Step(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("dx", _DP), ("dy", _DP), ("temp", _DP), ("fleck", _DP),
        ("sigma_a", _DP), ("sigma_s", _DP), ("radsource", _DP), ("tsurf", _DP * 4),
    ]


class RunCfg(ctypes.Structure):
    _fields_ = [
        ("pb", Problem), ("dx", _DP), ("dy", _DP), ("sa0", _DP), ("sa_pow", _DP), ("ss0", _DP),
        ("ss_pow", _DP), ("bee", _DP), ("radsource", _DP), ("tsurf", _DP * 4),
        ("t_init", ctypes.c_double), ("alpha", ctypes.c_double),
        ("linearized", ctypes.c_int32), ("marshak", ctypes.c_int32), ("timestepping", ctypes.c_int32),
        ("nthreads", ctypes.c_int32),
        ("dt", ctypes.c_double), ("t_end", ctypes.c_double), ("dt0", ctypes.c_double), ("k", ctypes.c_double),
        ("dtmax", ctypes.c_double), ("n_input", ctypes.c_int64), ("n_max", ctypes.c_int64),
        ("cellmin", ctypes.c_double), ("max_steps", ctypes.c_int64), ("time_from", ctypes.c_int64),
    ]


class RunOut(ctypes.Structure):
    _fields_ = [
        ("temp", _DP), ("radenergydens", _DP), ("matenergydens", _DP), ("energy_error", _DP),
        ("err_cap", ctypes.c_int64), ("steps_done", ctypes.c_int64), ("histories", ctypes.c_int64),
        ("segments", ctypes.c_int64), ("t_hot", ctypes.c_double), ("t_final", ctypes.c_double),
        ("n_alive", ctypes.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.orc_round_half.restype = ctypes.c_double
        L.orc_round_half.argtypes = [ctypes.c_double]
        L.orc_sorter.restype = ctypes.c_int
        L.orc_sorter.argtypes = [ctypes.c_int, _DP, ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int, _DP, _DP]
        L.orc_source.restype = ctypes.c_int64
        L.orc_source.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Step), ctypes.c_uint32, ctypes.c_int64,
                                 ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _DP, ctypes.c_int64, _DP, _DP]
        L.orc_transport.restype = ctypes.c_int
        L.orc_transport.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Step), ctypes.c_uint32, _DP,
                                    ctypes.c_int64, _DP, _DP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        L.orc_clean.restype = ctypes.c_int64
        L.orc_clean.argtypes = [ctypes.POINTER(Problem), _DP, ctypes.c_int64]
        L.orc_census_tally.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(Step), _DP, ctypes.c_int64, _DP]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.POINTER(RunCfg), ctypes.POINTER(RunOut)]
        L.orc_planck.restype = None
        L.orc_planck.argtypes = [ctypes.c_int64, _DP, _DP, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_planck_terms.restype = None
        L.orc_planck_terms.argtypes = [ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def planck(temp, seed: int, stream: int) -> np.ndarray:
    """Planck frequencies temp_i * x_i with the engine's Philox layout (imc_oracle.cpp:
    orc_planck; the method of imc_sourcing.jl:372-399)."""
    t = np.ascontiguousarray(temp, dtype=np.float64)
    out = np.empty_like(t)
    lib().orc_planck(t.size, _p(t), _p(out), seed, stream)
    return out


def planck_terms(n: int, seed: int, stream: int, u_force: float = -1.0) -> np.ndarray:
    """Mixture terms n of samples 0 .. n-1 of `planck` (u_force >= 0 replaces the term uniform)."""
    out = np.empty(n, dtype=np.int64)
    lib().orc_planck_terms(n, seed, stream, u_force, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


def sample_planck_reference(uniforms, precision=np.float64, max_terms=1 << 20):
    """Sourcing.sample_planck (imc_sourcing.jl:372-399) as written, for one draw: `uniforms`
    yields rand(precision) values in the order the reference takes them (rn1, then rn1..rn4
    after the term is chosen).  n and the running sum of n^-4 are kept in `precision`; the
    test rn1 <= 90 nsum / pi^4 is formed in Float64 (the literal 90.0 and pi^4 promote).
    Returns (x, n), or (None, n) when `max_terms` terms pass without the test succeeding --
    the reference's loop has no other exit."""
    one = precision(1.0)
    n = one
    nsum = one
    rn1 = float(next(uniforms))
    for _ in range(max_terms):
        if rn1 <= 90.0 * float(nsum) / np.pi ** 4:
            r = [precision(next(uniforms)) for _ in range(4)]
            prod = precision(precision(precision(r[0] * r[1]) * r[2]) * r[3])
            x = precision(-1.0 * float(np.log(prod)) / float(n))
            return x, float(n)
        n = precision(n + one)
        nsum = precision(nsum + _inv_n4(n, precision))
    return None, float(n)


def _inv_n4(n, precision):
    """precision(1.0 / n^4) with n^4 rounded to `precision` (Float16 n^4 overflows to Inf
    from n = 16, so the term is 0 there)."""
    with np.errstate(over="ignore"):
        n4 = precision(float(n) ** 4)
    return precision(1.0 / float(n4))


def planck_nsum_limit(precision) -> float:
    """The value at which the reference's running sum of n^-4 (kept in `precision`) stops
    growing; 90 * limit / pi^4 is the largest first uniform its loop can accept."""
    one = precision(1.0)
    n, nsum = one, one
    while True:
        nxt = precision(nsum + _inv_n4(precision(n + one), precision))
        if nxt == nsum:
            return float(nsum)
        n, nsum = precision(n + one), nxt


def philox(ctr: Sequence[int], key: Sequence[int]):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def round_half(x: float) -> float:
    return lib().orc_round_half(float(x))


def sorter(prec: int, vals, scales, promote64=False):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    s = np.ascontiguousarray(scales, dtype=np.float64)
    prod = ctypes.c_double()
    sc = ctypes.c_double()
    k = lib().orc_sorter(prec, _p(v), len(v), int(promote64), _p(s), len(s), ctypes.byref(prod), ctypes.byref(sc))
    return prod.value, sc.value, k


class OracleProblem:
    """Holds a problem and a step's fields; the arrays are kept alive by this object."""

    def __init__(self, *, geometry=1, prec=64, nx=1, ny=1, scales=(1.0,), bc=(VACUUM,) * 4, seed=1234,
                 phys_c=1.0, phys_a=1.0, distancescale=1.0, quirks=0):
        pb = Problem()
        pb.geometry = geometry; pb.prec = prec; pb.nx = nx; pb.ny = ny if geometry == 2 else 1
        pb.K = len(scales); pb.quirks = quirks
        for i in range(4):
            pb.bc[i] = bc[i]
        pb.seed = seed & 0xFFFFFFFF
        for i, s in enumerate(scales):
            pb.scales[i] = s
        pb.phys_c = phys_c; pb.phys_a = phys_a; pb.distancescale = distancescale
        self.pb = pb
        self.geometry = geometry
        self.C = nx * (ny if geometry == 2 else 1)
        self.K = len(scales)
        self.width = 10 if geometry == 2 else 9
        self._keep = []

    def step(self, *, dt, dx, dy=None, temp, fleck, sigma_a, sigma_s, radsource, tsurf) -> Step:
        st = Step()
        arrs = {}
        for name, a in (("dx", dx), ("dy", dy if dy is not None else np.ones(1)), ("temp", temp),
                        ("fleck", fleck), ("sigma_a", sigma_a), ("sigma_s", sigma_s), ("radsource", radsource)):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
            arrs[name] = arr
            setattr(st, name, _p(arr))
        ts = [np.ascontiguousarray(np.atleast_1d(np.asarray(t, dtype=np.float64))) for t in tsurf]
        while len(ts) < 4:
            ts.append(np.zeros(1))
        for i in range(4):
            st.tsurf[i] = _p(ts[i])
        st.dt = dt
        self._keep.append((arrs, ts, st))
        return st

    def source(self, st: Step, step: int, census: np.ndarray, n_input, n_max, cellmin, cap=None):
        n_c = census.shape[0] if census is not None else 0
        cap = cap or (n_c + 4 * n_input + 8 * self.C * (int(cellmin) + 2) + 4096)
        rec = np.zeros((cap, self.width))
        if n_c:
            rec[:n_c] = census
        emitted = np.zeros(self.C * self.K)
        etot = ctypes.c_double()
        nb = lib().orc_source(ctypes.byref(self.pb), ctypes.byref(st), step, n_c, int(n_input), int(n_max),
                              float(cellmin), _p(rec), cap, _p(emitted), ctypes.byref(etot))
        if nb < 0:
            raise RuntimeError("oracle source capacity exceeded")
        return rec[: n_c + nb].copy(), emitted, etot.value, nb

    def transport(self, st: Step, step: int, rec: np.ndarray, nthreads: int = 1):
        rec = np.ascontiguousarray(rec, dtype=np.float64).copy()
        dep = np.zeros(self.C * self.K)
        lost = ctypes.c_double(0.0)
        segs = ctypes.c_int64(0)
        lib().orc_transport(ctypes.byref(self.pb), ctypes.byref(st), step, _p(rec), rec.shape[0], _p(dep),
                            ctypes.byref(lost), ctypes.byref(segs), nthreads)
        return rec, dep, lost.value, segs.value

    def clean(self, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, dtype=np.float64).copy()
        n = lib().orc_clean(ctypes.byref(self.pb), _p(rec), rec.shape[0])
        return rec[:n]

    def census(self, st: Step, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        rad = np.zeros(self.C)
        lib().orc_census_tally(ctypes.byref(self.pb), ctypes.byref(st), _p(rec), rec.shape[0], _p(rad))
        return rad


def run(*, geometry, prec, nx, ny=1, dx, dy=None, sa0, sa_pow, ss0, ss_pow, bee, radsource, tsurf, t_init,
        alpha, phys_c, phys_a, distancescale=1.0, scales=(1.0,), bc=(VACUUM,) * 4, seed=1234, linearized=False,
        marshak=False, timestepping="CONSTANT", dt=0.0, t_end=1.0, dt0=0.0, k=1.0, dtmax=0.0, n_input=1000,
        n_max=10000, cellmin=1.0, max_steps=0, time_from=0, nthreads=1, quirks=0):
    """Full run of the restated main loop; returns a dict of final fields and counters.
    histories, segments and t_hot count the steps from index time_from on."""
    C = nx * (ny if geometry == 2 else 1)
    cfg = RunCfg()
    pb = cfg.pb
    pb.geometry = geometry; pb.prec = prec; pb.nx = nx; pb.ny = ny if geometry == 2 else 1
    pb.K = len(scales); pb.quirks = quirks
    for i in range(4):
        pb.bc[i] = bc[i]
    pb.seed = seed & 0xFFFFFFFF
    for i, s in enumerate(scales):
        pb.scales[i] = s
    pb.phys_c = phys_c; pb.phys_a = phys_a; pb.distancescale = distancescale
    keep = {}

    def arr(name, a, n):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64).ravel(order="F"), (n,)).copy())
        keep[name] = v
        return _p(v)

    cfg.dx = arr("dx", dx, nx)
    cfg.dy = arr("dy", dy if dy is not None else np.ones(1), pb.ny)
    cfg.sa0 = arr("sa0", sa0, C); cfg.sa_pow = arr("sa_pow", sa_pow, C)
    cfg.ss0 = arr("ss0", ss0, C); cfg.ss_pow = arr("ss_pow", ss_pow, C)
    cfg.bee = arr("bee", bee, C); cfg.radsource = arr("radsource", radsource, C)
    ts = [np.ascontiguousarray(np.atleast_1d(np.asarray(t, dtype=np.float64))) for t in tsurf]
    while len(ts) < 4:
        ts.append(np.zeros(1))
    keep["ts"] = ts
    for i in range(4):
        cfg.tsurf[i] = _p(ts[i])
    cfg.t_init = t_init; cfg.alpha = alpha; cfg.linearized = int(linearized); cfg.marshak = int(marshak)
    cfg.timestepping = 0 if timestepping.upper() == "CONSTANT" else 1
    cfg.nthreads = nthreads
    cfg.dt = dt; cfg.t_end = t_end; cfg.dt0 = dt0; cfg.k = k; cfg.dtmax = dtmax
    cfg.n_input = int(n_input); cfg.n_max = int(n_max); cfg.cellmin = cellmin; cfg.max_steps = max_steps
    cfg.time_from = time_from
    out = RunOut()
    temp = np.zeros(C); rad = np.zeros(C); mat = np.zeros(C)
    cap = 100000
    err = np.zeros(cap)
    out.temp = _p(temp); out.radenergydens = _p(rad); out.matenergydens = _p(mat)
    out.energy_error = _p(err); out.err_cap = cap
    lib().orc_run(ctypes.byref(cfg), ctypes.byref(out))
    return {"temp": temp, "radenergydens": rad, "matenergydens": mat,
            "energy_error": err[: min(out.steps_done, cap)].copy(), "steps": out.steps_done,
            "histories": out.histories, "segments": out.segments, "t_hot": out.t_hot,
            "t_final": out.t_final, "n_alive": out.n_alive}


def run_deck(inputs: dict, mesh, *, prec=None, seed=None, max_steps=0, time_from=0, nthreads=1, quirks=0):
    """Runs the oracle's restated main loop on a deck parsed by the host package
    (Input.readInputs + Mesh.mesh_generation); used by tests and the CPU baseline."""
    def g(key):
        v = inputs[key]
        return str(v[0]) if isinstance(v, list) else str(v)

    geometry = 1 if g("GEOMETRY").upper() == "1D" else 2
    bcmap = {"VACUUM": VACUUM, "REFLECT": REFLECT}
    if geometry == 1:
        nx, ny = mesh.Ncells, 1
        bc = (bcmap[g("LEFTBC").upper()], bcmap[g("RIGHTBC").upper()], VACUUM, VACUUM)
        tsurf = [np.array([mesh.temp_surf[0]]), np.array([mesh.temp_surf[1]])]
    else:
        nx, ny = mesh.Ncells
        bc = tuple(bcmap[g(k).upper()] for k in ("LEFTBC", "RIGHTBC", "TOPBC", "BOTTOMBC"))
        tsurf = [np.asarray(t, dtype=np.float64) for t in mesh.temp_surf]
    ts = g("TIMESTEPPING").upper()
    kw = {}
    if ts == "CONSTANT":
        kw["dt"] = float(g("DT"))
    else:
        kw.update(dt0=float(g("DT0")), k=float(g("K")), dtmax=float(g("DTMAX")))
    return run(geometry=geometry, prec=int(prec if prec is not None else inputs["PRECISION"]), nx=nx, ny=ny,
               dx=mesh.dx, dy=mesh.dy if geometry == 2 else None,
               sa0=mesh.sigma_a[..., 1], sa_pow=mesh.sigma_a[..., 2], ss0=mesh.sigma_s[..., 1],
               ss_pow=mesh.sigma_s[..., 2], bee=mesh.bee, radsource=mesh.radsource, tsurf=tsurf,
               t_init=float(g("T_INIT")), alpha=float(g("ALPHA")), phys_c=float(g("PHYS_C")),
               phys_a=float(g("PHYS_A")), distancescale=mesh.distancescale, scales=tuple(mesh.energyscales),
               bc=bc, seed=int(seed if seed is not None else int(g("SEED"))),
               linearized=g("LINEARIZED") == "TRUE",
               marshak=(geometry == 1 and g("NAME").upper() == "MARSHAK WAVE"), timestepping=ts,
               t_end=float(g("ENDTIME")), n_input=int(float(g("NINPUT"))), n_max=int(float(g("NMAX"))),
               cellmin=float(g("CELLMIN")), max_steps=max_steps, time_from=time_from, nthreads=nthreads, quirks=quirks, **kw)
