"""ctypes binding of libimc_hip (include/imc_hip.h).

The shared library is built in-tree (mixedprecisionimc.jl_amd/csrc/libimc_hip.so).  There is
no fallback: if the library is missing or cannot be loaded every call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "csrc", "libimc_hip.so"))

IMC_OK = 0
IMC_EINVAL = -1
IMC_EHIP = -2
IMC_ENOMEM = -3
IMC_ENUMERIC = -4
VACUUM = 0
REFLECT = 1
QUIRK_Q1_RADSRC_NBODY = 1
QUIRK_Q2_LR_FACE_DX = 2

_STATUS = {IMC_EINVAL: "EINVAL", IMC_EHIP: "EHIP", IMC_ENOMEM: "ENOMEM", IMC_ENUMERIC: "ENUMERIC"}


class ImcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libimc_hip {_STATUS.get(code, code)}: {msg}")
        self.code = code


class ImcConfig(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("n_scales", ctypes.c_int32),
        ("quirks", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4),
        ("seed", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("energyscales", ctypes.c_double * 16),
        ("phys_c", ctypes.c_double),
        ("phys_a", ctypes.c_double),
        ("distancescale", ctypes.c_double),
        ("capacity", ctypes.c_int64),
        ("reserved", ctypes.c_int32 * 8),
    ]


class ImcTiming(ctypes.Structure):
    _fields_ = [
        ("ms_source", ctypes.c_double),
        ("ms_transport", ctypes.c_double),
        ("ms_clean", ctypes.c_double),
        ("ms_tally", ctypes.c_double),
        ("ms_transport_kernel", ctypes.c_double),
        ("n_transport_launches", ctypes.c_int64),
        ("histories", ctypes.c_int64),
        ("census_out", ctypes.c_int64),
        ("segments", ctypes.c_int64),
        ("transport_tally_mode", ctypes.c_int32),
        ("transport_grid", ctypes.c_int32),
        ("surrogate_events", ctypes.c_int64),
        ("truncated", ctypes.c_int64),
    ]


class SurrogateBox(ctypes.Structure):
    _fields_ = [("tau_cell_min", ctypes.c_float), ("tau_cell_max", ctypes.c_float),
                ("tau_b_min", ctypes.c_float), ("tau_b_max", ctypes.c_float),
                ("kappa_min", ctypes.c_float), ("kappa_max", ctypes.c_float),
                ("depth_min", ctypes.c_float)]


class ImcMaterial(ctypes.Structure):
    _fields_ = [("sigma_a_coef", ctypes.POINTER(ctypes.c_double)), ("sigma_a_power", ctypes.POINTER(ctypes.c_double)),
                ("sigma_s_coef", ctypes.POINTER(ctypes.c_double)), ("sigma_s_power", ctypes.POINTER(ctypes.c_double)),
                ("bee", ctypes.POINTER(ctypes.c_double)), ("temp", ctypes.POINTER(ctypes.c_double)),
                ("radsource", ctypes.POINTER(ctypes.c_double)), ("t_surf", ctypes.POINTER(ctypes.c_double) * 4),
                ("alpha", ctypes.c_double), ("linearized", ctypes.c_int32), ("marshak", ctypes.c_int32)]


class ImcStepResult(ctypes.Structure):
    _fields_ = [("totalenergy", ctypes.c_double), ("totalenergydep", ctypes.c_double),
                ("radenergy", ctypes.c_double), ("lostenergy", ctypes.c_double),
                ("histories", ctypes.c_int64), ("n_alive", ctypes.c_int64), ("segments", ctypes.c_int64)]


SURROGATE_NPARAM = 4868


_lib: Optional[ctypes.CDLL] = None

_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p

SYMBOLS = {
    "imc_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.POINTER(ImcConfig)]),
    "imc_destroy": (ctypes.c_int, [_VP]),
    "imc_last_error": (ctypes.c_char_p, [_VP]),
    "imc_set_geometry": (ctypes.c_int, [_VP, _DP, _DP]),
    "imc_set_step_fields": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_double, _DP, _DP, _DP, _DP, _DP,
                                           ctypes.POINTER(_DP)]),
    "imc_source": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, _DP, _DP, _I64P]),
    "imc_transport": (ctypes.c_int, [_VP, _DP, _DP, _I64P]),
    "imc_clean": (ctypes.c_int, [_VP, _I64P]),
    "imc_tally_census": (ctypes.c_int, [_VP, _DP]),
    "imc_particle_count": (ctypes.c_int, [_VP, _I64P]),
    "imc_export_particles": (ctypes.c_int, [_VP, _DP, ctypes.c_int64, _I64P]),
    "imc_import_particles": (ctypes.c_int, [_VP, _DP, ctypes.c_int64]),
    "imc_get_timing": (ctypes.c_int, [_VP, ctypes.POINTER(ImcTiming)]),
    "imc_reset_timing": (ctypes.c_int, [_VP]),
    "imc_synchronize": (ctypes.c_int, [_VP]),
    "imc_state_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int32)]),
    "imc_set_material": (ctypes.c_int, [_VP, ctypes.POINTER(ImcMaterial)]),
    "imc_step": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                ctypes.c_double, _DP, _DP, _DP, ctypes.POINTER(ImcStepResult)]),
    "imc_step_energy_increase": (ctypes.c_int, [_VP, _DP]),
    "imc_surrogate_configure": (ctypes.c_int, [_VP, ctypes.POINTER(SurrogateBox)]),
    "imc_surrogate_init": (ctypes.c_int, [_VP, ctypes.c_uint32]),
    "imc_surrogate_generate": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_uint32, _I64P]),
    "imc_surrogate_dataset": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_int64]),
    "imc_surrogate_train": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                           ctypes.POINTER(ctypes.c_float)]),
    "imc_surrogate_set_params": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "imc_surrogate_get_params": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "imc_surrogate_forward": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                                             ctypes.POINTER(ctypes.c_float)]),
    "imc_surrogate_gradient": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                              ctypes.c_int64, ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_float)]),
    "imc_surrogate_enable": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "imc_random_walk_enable": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_double]),
    "imc_sample_planck": (ctypes.c_int, [_VP, ctypes.c_int64, _DP, ctypes.c_uint32, _DP]),
}

_FP = ctypes.POINTER(ctypes.c_float)


def _fp(a):
    return a.ctypes.data_as(_FP)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def lib() -> ctypes.CDLL:
    """Loads libimc_hip.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImcError(IMC_EINVAL, f"{LIB_PATH} not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


class Engine:
    """One libimc_hip context: the device-resident particle population of a run."""

    def __init__(self, *, geometry: int, precision: int, nx: int, ny: int = 1,
                 energyscales: Sequence[float] = (1.0,), bc=(VACUUM, VACUUM, VACUUM, VACUUM),
                 seed: int = 1234, phys_c: float = 1.0, phys_a: float = 1.0,
                 distancescale: float = 1.0, capacity: int = 1 << 16, quirks: int = 0,
                 device: int = 0, tally_mode: int = 0, blocks_per_cu: int = 0, hash_bits: int = 0):
        L = lib()
        cfg = ImcConfig()
        cfg.geometry = geometry
        cfg.precision = precision
        cfg.nx = nx
        cfg.ny = ny if geometry == 2 else 1
        scales = list(energyscales)
        cfg.n_scales = len(scales)
        for i, s in enumerate(scales):
            cfg.energyscales[i] = float(s)
        for i in range(4):
            cfg.bc[i] = int(bc[i]) if i < len(bc) else VACUUM
        cfg.seed = int(seed) & 0xFFFFFFFF
        cfg.device = device
        cfg.phys_c = phys_c
        cfg.phys_a = phys_a
        cfg.distancescale = distancescale
        cfg.capacity = int(capacity)
        cfg.quirks = quirks
        cfg.reserved[0] = int(tally_mode)        # 0 auto, 1 wave rows, 2 block row, 3 global, 4 hashed
        cfg.reserved[1] = int(blocks_per_cu)     # 0 auto: occupancy-derived persistent grid
        cfg.reserved[2] = int(hash_bits)         # 0 auto: log2 of the hashed tally's slots
        self.geometry = geometry
        self.precision = precision
        self.nx = nx
        self.ny = cfg.ny
        self.C = nx * cfg.ny
        self.K = len(scales)
        self.scales = np.array(scales, dtype=np.float64)
        h = _VP()
        rc = L.imc_create(ctypes.byref(h), ctypes.byref(cfg))
        self._h = h
        if rc != IMC_OK:
            msg = L.imc_last_error(h).decode() if h.value else "imc_create failed"
            L.imc_destroy(h)
            self._h = None
            raise ImcError(rc, msg)
        self.n = 0

    # -- helpers ---------------------------------------------------------------
    def _check(self, rc: int):
        if rc != IMC_OK:
            raise ImcError(rc, lib().imc_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().imc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return int(self.n)

    # -- ABI calls --------------------------------------------------------------
    def set_geometry(self, dx, dy=None):
        self._dx = _f64(dx)
        self._dy = _f64(dy) if dy is not None else None
        self._check(lib().imc_set_geometry(self._h, _dp(self._dx), _dp(self._dy)))

    def set_step_fields(self, step: int, dt: float, sigma_a, sigma_s, fleck, temp, radsource, t_surf):
        arrs = [_f64(a) for a in (sigma_a, sigma_s, fleck, temp, radsource)]
        ts = [_f64(np.atleast_1d(a)) for a in t_surf]
        while len(ts) < 4:
            ts.append(np.zeros(1))
        tsp = (_DP * 4)(*[_dp(a) for a in ts])
        self._check(lib().imc_set_step_fields(self._h, int(step), float(dt), *[_dp(a) for a in arrs], tsp))

    def source(self, n_input: int, n_max: int, cellmin: float):
        emitted = np.zeros(self.C * self.K)
        etot = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().imc_source(self._h, int(n_input), int(n_max), float(cellmin), _dp(emitted),
                                     ctypes.byref(etot), ctypes.byref(n)))
        self.n = n.value
        return emitted, etot.value, n.value

    def transport(self):
        dep = np.zeros(self.C * self.K)
        lost = ctypes.c_double()
        segs = ctypes.c_int64()
        self._check(lib().imc_transport(self._h, _dp(dep), ctypes.byref(lost), ctypes.byref(segs)))
        return dep, lost.value, segs.value

    def clean(self) -> int:
        n = ctypes.c_int64()
        self._check(lib().imc_clean(self._h, ctypes.byref(n)))
        self.n = n.value
        return n.value

    def tally_census(self) -> np.ndarray:
        rad = np.zeros(self.C)
        self._check(lib().imc_tally_census(self._h, _dp(rad)))
        return rad

    def export_particles(self) -> np.ndarray:
        n = ctypes.c_int64()
        self._check(lib().imc_particle_count(self._h, ctypes.byref(n)))
        width = 10 if self.geometry == 2 else 9
        rec = np.zeros((max(n.value, 1), width))
        got = ctypes.c_int64()
        self._check(lib().imc_export_particles(self._h, _dp(rec), n.value, ctypes.byref(got)))
        return rec[: got.value]

    def import_particles(self, records: np.ndarray):
        width = 10 if self.geometry == 2 else 9
        rec = np.ascontiguousarray(np.asarray(records, dtype=np.float64).reshape(-1, width))
        self._check(lib().imc_import_particles(self._h, _dp(rec), rec.shape[0]))
        self.n = rec.shape[0]

    def timing(self) -> dict:
        t = ImcTiming()
        self._check(lib().imc_get_timing(self._h, ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in ImcTiming._fields_}

    def reset_timing(self):
        self._check(lib().imc_reset_timing(self._h))

    def synchronize(self):
        self._check(lib().imc_synchronize(self._h))

    def state_bytes(self) -> int:
        b = ctypes.c_int32()
        self._check(lib().imc_state_bytes(self._h, ctypes.byref(b)))
        return b.value

    # -- device-resident step ---------------------------------------------------
    def set_material(self, sigma_a_coef, sigma_a_power, sigma_s_coef, sigma_s_power, bee, temp, radsource,
                     t_surf, alpha: float, linearized: bool, marshak: bool):
        """imc_set_material: per-cell material description for imc_step (column-major)."""
        keep = [_f64(a) for a in (sigma_a_coef, sigma_a_power, sigma_s_coef, sigma_s_power, temp, radsource)]
        bee_a = _f64(bee) if bee is not None else None
        ts = [_f64(np.atleast_1d(a)) for a in t_surf]
        while len(ts) < 4:
            ts.append(np.zeros(1))
        m = ImcMaterial()
        m.sigma_a_coef, m.sigma_a_power, m.sigma_s_coef, m.sigma_s_power = (_dp(a) for a in keep[:4])
        m.temp, m.radsource = _dp(keep[4]), _dp(keep[5])
        m.bee = _dp(bee_a) if bee_a is not None else _DP()
        for q in range(4):
            m.t_surf[q] = _dp(ts[q])
        m.alpha = float(alpha)
        m.linearized = int(bool(linearized))
        m.marshak = int(bool(marshak))
        self._check(lib().imc_set_material(self._h, ctypes.byref(m)))

    def step(self, step: int, dt: float, first: bool, n_input: int, n_max: int, cellmin: float):
        """imc_step: one device-resident step.  Returns (temp, radenergydens, matenergydens, result)."""
        temp, rad, mat = np.empty(self.C), np.empty(self.C), np.empty(self.C)    # filled by imc_step
        r = ImcStepResult()
        self._check(lib().imc_step(self._h, int(step), float(dt), int(bool(first)), int(n_input), int(n_max),
                                   float(cellmin), _dp(temp), _dp(rad), _dp(mat), ctypes.byref(r)))
        self.n = r.n_alive
        return temp, rad, mat, {name: getattr(r, name) for name, _ in ImcStepResult._fields_}

    def step_energy_increase(self) -> np.ndarray:
        inc = np.empty(self.C)
        self._check(lib().imc_step_energy_increase(self._h, _dp(inc)))
        return inc

    # -- learned thick-cell surrogate (slot of Transport.MC_RW) -----------------
    def surrogate_configure(self, tau_cell=(2.5, 200.0), tau_b=(2.0, 1e4), kappa=(1e-5, 4.0), depth_min=0.7):
        box = SurrogateBox(tau_cell[0], tau_cell[1], tau_b[0], tau_b[1], kappa[0], kappa[1], depth_min)
        self._check(lib().imc_surrogate_configure(self._h, ctypes.byref(box)))

    def surrogate_init(self, seed: int = 0):
        self._check(lib().imc_surrogate_init(self._h, int(seed) & 0xFFFFFFFF))

    def surrogate_generate(self, n: int, seed: int = 0) -> int:
        ev = ctypes.c_int64()
        self._check(lib().imc_surrogate_generate(self._h, int(n), int(seed) & 0xFFFFFFFF, ctypes.byref(ev)))
        return ev.value

    def surrogate_dataset(self, n: int):
        f = np.zeros((n, 6), np.float32)
        t = np.zeros((n, 4), np.float32)
        self._check(lib().imc_surrogate_dataset(self._h, _fp(f), _fp(t), int(n)))
        return f, t

    def surrogate_train(self, epochs: int, batch: int = 8192, lr: float = 3e-3) -> np.ndarray:
        losses = np.zeros(epochs, np.float32)
        self._check(lib().imc_surrogate_train(self._h, int(epochs), int(batch), float(lr), _fp(losses)))
        return losses

    def surrogate_params(self) -> np.ndarray:
        p = np.zeros(SURROGATE_NPARAM, np.float32)
        self._check(lib().imc_surrogate_get_params(self._h, _fp(p)))
        return p

    def surrogate_set_params(self, params):
        p = _f32(params)
        assert p.size == SURROGATE_NPARAM
        self._check(lib().imc_surrogate_set_params(self._h, _fp(p)))

    def surrogate_forward(self, features) -> np.ndarray:
        f = _f32(features).reshape(-1, 6)
        out = np.zeros((f.shape[0], 4), np.float32)
        self._check(lib().imc_surrogate_forward(self._h, _fp(f), f.shape[0], _fp(out)))
        return out

    def surrogate_gradient(self, features, targets):
        f = _f32(features).reshape(-1, 6)
        t = _f32(targets).reshape(-1, 4)
        g = np.zeros(SURROGATE_NPARAM, np.float32)
        loss = ctypes.c_float()
        self._check(lib().imc_surrogate_gradient(self._h, _fp(f), _fp(t), f.shape[0], _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def surrogate_enable(self, on: bool = True):
        self._check(lib().imc_surrogate_enable(self._h, int(bool(on))))
        self.surrogate_on = bool(on)
        if on:
            self.random_walk_on = False

    def sample_planck(self, temp, stream: int = 0) -> np.ndarray:
        """imc_sample_planck: frequencies temp_i * x_i, x_i from the normalised Planck
        spectrum, keyed by (seed, stream, i)."""
        t = np.ascontiguousarray(temp, dtype=np.float64).ravel()
        out = np.empty_like(t)
        self._check(lib().imc_sample_planck(self._h, t.size, _dp(t), int(stream) & 0xFFFFFFFF, _dp(out)))
        return out

    def random_walk_enable(self, on: bool = True, r_min: float = 8.0):
        """imc_random_walk_enable: the analytic random walk in cells at least r_min collision
        mean free paths deep (the Transport.MC_RW slot); disables the surrogate."""
        self._check(lib().imc_random_walk_enable(self._h, int(bool(on)), float(r_min)))
        self.random_walk_on = bool(on)
        if on:
            self.surrogate_on = False
