"""Mesh.mesh_generation -- src/imc_mesh.jl:42-259, in FP64 on the host.

Differences from the reference, each deliberate: 2D UNIFORM meshes use YSIZE/DY for Ny
(the reference divides by DX, imc_mesh.jl:86 -- quirk Q3); a single 2D region rectangle
given without an enclosing tuple is accepted.
"""
from __future__ import annotations

import numpy as np

from .Input import get_scalar
from .Utilities import tointeger


class MeshStruct:
    """Fields of imc_mesh.jl:10-40 (arrays are numpy FP64; 2D arrays are (nx, ny))."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _linrange(a: float, b: float, n: int) -> np.ndarray:
    # Julia LinRange element j: fma(t, b, fma(-t, a, a)) with t = j/(n-1)
    if n == 1:
        return np.array([a])
    t = np.arange(n, dtype=np.float64) / (n - 1)
    return (a - t * a) + t * b


def _num(v) -> float:
    return float(v)


def region_joiner(geometry: str, regions, values, nodes, Ncells):
    """imc_mesh.jl:175-220."""
    if geometry == "1D":
        N = int(Ncells)
        if len(regions) <= 1:
            return np.full(N, _num(values[0]))
        out = np.zeros(N)
        ri = 0
        for i in range(N):
            if _num(regions[ri]) >= nodes[i + 1]:
                out[i] = _num(values[ri])
            else:
                ri += 1
                out[i] = _num(values[ri])
        return out
    nx, ny = Ncells
    rects = regions[0] if len(regions) else ""
    if isinstance(rects, str) or len(rects) <= 1:
        return np.full((nx, ny), _num(values[0]))
    # one rectangle written without the outer tuple: ((x0,x1),(y0,y1))
    if len(rects) == 2 and all(isinstance(r, tuple) and len(r) == 2 and
                               all(isinstance(q, (int, float)) for q in r) for r in rects):
        rects = (rects,)
    out = np.zeros((nx, ny))
    xn, yn = np.asarray(nodes[0]), np.asarray(nodes[1])
    for ri, rect in enumerate(rects):
        (x0, x1), (y0, y1) = rect
        xs = int(np.flatnonzero(xn <= x0)[-1])
        xe = int(np.flatnonzero(xn <= x1)[-1])
        ys = int(np.flatnonzero(yn <= y0)[-1])
        ye = int(np.flatnonzero(yn <= y1)[-1])
        out[xs:xe, ys:ye] = _num(values[ri])
    return out


def surface_definer(geometry: str, regions, values, nodes, Ncells):
    """imc_mesh.jl:222-259: 1D (left, right); 2D [bottom(nx), top(nx), left(ny), right(ny)]."""
    if geometry == "1D":
        return (_num(values[0]), _num(values[1]))
    nx, ny = Ncells
    vals = values[0]
    regs = regions[0]
    out = []
    for i in range(4):
        v = vals[i]
        sv = [str(v)] if not isinstance(v, tuple) else list(v)
        r = regs[i]
        sr = [r] if not isinstance(r, tuple) else list(r)
        axis_nodes = nodes[0] if i < 2 else nodes[1]
        n = nx if i < 2 else ny
        out.append(region_joiner("1D", sr, sv, axis_nodes, n))
    return out


def mesh_generation(inputs: dict) -> MeshStruct:
    geometry = get_scalar(inputs, "GEOMETRY").upper()
    meshtype = get_scalar(inputs, "MESHTYPE").upper()
    if geometry == "1D":
        if meshtype == "UNIFORM":
            xsize = float(get_scalar(inputs, "XSIZE"))
            dxv = float(get_scalar(inputs, "DX"))
            N = tointeger(xsize / dxv)
            dx = np.full(N, dxv)
            nodes = np.arange(N + 1, dtype=np.float64) / N * xsize
            centers = _linrange(dxv / 2, xsize - dxv / 2, N)
        else:
            nodes = np.asarray(inputs["MESHNODES"], dtype=np.float64)
            centers = (nodes[:-1] + nodes[1:]) / 2
            dx = nodes[1:] - nodes[:-1]
            N = len(dx)
        dy = np.ones(1)
        Ncells = N
        shape = (N,)
    else:
        if meshtype == "UNIFORM":
            xsize = float(get_scalar(inputs, "XSIZE"))
            ysize = float(get_scalar(inputs, "YSIZE"))
            dxv = float(get_scalar(inputs, "DX"))
            dyv = float(get_scalar(inputs, "DY"))
            nx, ny = tointeger(xsize / dxv), tointeger(ysize / dyv)
            dx, dy = np.full(nx, dxv), np.full(ny, dyv)
            xn = np.arange(nx + 1, dtype=np.float64) / nx * xsize
            yn = np.arange(ny + 1, dtype=np.float64) / ny * ysize
            centers = (_linrange(dxv / 2, xsize - dxv / 2, nx), _linrange(dyv / 2, ysize - dyv / 2, ny))
        else:
            xn = np.asarray(inputs["XMESHNODES"], dtype=np.float64)
            yn = np.asarray(inputs["YMESHNODES"], dtype=np.float64)
            centers = ((xn[:-1] + xn[1:]) / 2, (yn[:-1] + yn[1:]) / 2)
            dx, dy = xn[1:] - xn[:-1], yn[1:] - yn[:-1]
            nx, ny = len(dx), len(dy)
        nodes = (xn, yn)
        Ncells = (nx, ny)
        shape = (nx, ny)

    T_init = float(get_scalar(inputs, "T_INIT"))
    temp_surf = surface_definer(geometry, inputs["T_SURFACE_REGS"], inputs["T_SURFACE_VALS"], nodes, Ncells)
    temp = np.full(shape, T_init)
    sa = region_joiner(geometry, inputs["SIGMA_A_REGS"], inputs["SIGMA_A_VALS"], nodes, Ncells)
    sa_p = region_joiner(geometry, inputs["SIGMA_A_REGS"], inputs["SIGMA_A_POWERS"], nodes, Ncells)
    ss = region_joiner(geometry, inputs["SIGMA_S_REGS"], inputs["SIGMA_S_VALS"], nodes, Ncells)
    ss_p = region_joiner(geometry, inputs["SIGMA_S_REGS"], inputs["SIGMA_S_POWERS"], nodes, Ncells)
    es = inputs["ENERGYSCALES"]
    scales = sorted([float(v) for v in (es if isinstance(es, list) else [es])], reverse=True)
    distancescale = float(get_scalar(inputs, "DISTANCESCALE"))
    sa = sa / distancescale
    ss = ss / distancescale
    sigma_a = np.stack([sa, sa.copy(), sa_p], axis=-1)
    sigma_s = np.stack([ss, ss.copy(), ss_p], axis=-1)
    radsource = region_joiner(geometry, inputs["RADSOURCE_REGS"], inputs["RADSOURCE_VALS"], nodes, Ncells)
    bee = region_joiner(geometry, inputs["BEE_REGS"], inputs["BEE_VALS"], nodes, Ncells)
    K = len(scales)
    return MeshStruct(
        temp=temp, temp_surf=temp_surf, fleck=np.zeros(shape), beta=np.ones(shape),
        sigma=sigma_a + sigma_s, sigma_a=sigma_a, sigma_s=sigma_s,
        energydep=np.zeros(shape + (K,)), emittedenergy=np.zeros(shape + (K,)),
        totalenergy=0.0, totalenergydep=0.0, lostenergy=0.0, radenergyold=0.0,
        bee=bee, radsource=radsource, radenergydens=np.zeros(shape), matenergydens=np.zeros(shape),
        temp_saved=[temp.copy()], energyincrease_saved=[], radenergy_saved=[], matenergy_saved=[],
        energyscales=np.array(scales), distancescale=distancescale, centers=centers, nodes=nodes,
        dx=dx, dy=dy, perimeter=None, Ncells=Ncells, geometry=geometry, shape=shape,
    )
