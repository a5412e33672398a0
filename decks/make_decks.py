"""Writes the five workload decks (reference deck keywords, src/imc_input.jl format).

  c1_slab_fp64.txt        plain linearized slab, FP64 (400 cells)
  c2_marshak_fp32.txt     nonlinear Marshak wave, FP32 tracking (200 cells, ramped dt)
  c3_channel_fp16.txt     2D two-material channel, 280x80 graded mesh, FP16 tracking
  c4_thickthin_fp32.txt   1D alternating thick/thin 50-cell blocks, 2000 cells (seeded)
  c5_blocks2d_fp16.txt    2D 1024^2 mesh with 64x64 random material blocks (seed 7)

Keywords the workload descriptions leave open are pinned here and repeated in the deck
headers.  Run: python decks/make_decks.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

COMMON_TAIL = """
#### OUTPUTS ####
PLOTVARS = TEMPERATURE
BENCHMARK = FALSE
SAVEFIG = FALSE
SAVEANIMATION = FALSE
SAVEVARS = FALSE
"""


def fmt_list(vals, per_line=8):
    items = [f"{v:.10g}" for v in vals]
    lines = [", ".join(items[i:i + per_line]) for i in range(0, len(items), per_line)]
    return "[" + ",\n    ".join(lines) + "]"


def graded_nodes(breaks, counts, fine_at):
    """Piecewise nodes: segment i of [breaks[i], breaks[i+1]] gets counts[i] cells, graded
    geometrically toward the ends listed in fine_at (ratio 1.15)."""
    nodes = [breaks[0]]
    for i, n in enumerate(counts):
        a, b = breaks[i], breaks[i + 1]
        left, right = a in fine_at, b in fine_at
        if left and right:
            h = np.concatenate([1.15 ** np.arange(n // 2), 1.15 ** np.arange(n - n // 2)[::-1]])
        elif left:
            h = 1.15 ** np.arange(n)
        elif right:
            h = 1.15 ** np.arange(n)[::-1]
        else:
            h = np.ones(n)
        w = (b - a) * np.cumsum(h) / h.sum()
        nodes.extend((a + w).tolist())
    nodes[-1] = breaks[-1]
    return np.array(nodes)


def c1():
    return """NAME = Plain Slab
# Workload 1: 1D uniform slab x in [0,4], dx = 0.01 (400 cells), linearized, FP64.
PRECISION = FLOAT64
SEED = 1234
GEOMETRY = 1D
MESHTYPE = UNIFORM
XSIZE = 4.0
DX = 0.01
SIGMA_A_REGS = []
SIGMA_A_VALS = [1.0]
SIGMA_A_POWERS = [0.0]
SIGMA_S_REGS = []
SIGMA_S_VALS = [0.0]
SIGMA_S_POWERS = [0.0]
RADSOURCE_REGS = [0.5, 4.0]
RADSOURCE_VALS = [1.0, 0.0]
BEE_REGS = []
BEE_VALS = [1.0]
LEFTBC = VACUUM
RIGHTBC = REFLECT
TIMESTEPPING = CONSTANT
DT = 0.01
ENDTIME = 1.0
NINPUT = 100000
NMAX = 1000000
CELLMIN = 10
T_INIT = 0.0001
T_SURFACE_VALS = [0.0,0.0]
T_SURFACE_REGS = []
PHYS_C = 1.0
PHYS_A = 1.0
ALPHA = 1.0
LINEARIZED = TRUE
PAIRWISE = FALSE
RANDOMWALK = FALSE
ENERGYSCALES = [1.0]
DISTANCESCALE = 1.0
""" + COMMON_TAIL


def c2():
    return """NAME = Marshak Wave
# Workload 2: nonlinear Marshak wave, x in [0,0.2], 200 cells, sigma_a = 10 T^-3, bee = 0.3.
# FP32 tracking with compensated FP32 tallies.  Pinned: SEED 12345, CELLMIN 10.
PRECISION = FLOAT32
SEED = 12345
GEOMETRY = 1D
MESHTYPE = UNIFORM
XSIZE = 0.2
DX = 0.001
SIGMA_A_REGS = []
SIGMA_A_VALS = [10.0]
SIGMA_A_POWERS = [-3.0]
SIGMA_S_REGS = []
SIGMA_S_VALS = [0.0]
SIGMA_S_POWERS = [0.0]
RADSOURCE_REGS = []
RADSOURCE_VALS = [0.0]
BEE_REGS = []
BEE_VALS = [0.3]
LEFTBC = VACUUM
RIGHTBC = VACUUM
TIMESTEPPING = RAMP
DT0 = 0.0001
K = 1.1
DTMAX = 0.001
ENDTIME = 1.0
NINPUT = 2000000
NMAX = 10000000
CELLMIN = 10
T_INIT = 0.01
T_SURFACE_VALS = [1.0,0.0]
T_SURFACE_REGS = []
PHYS_C = 299.8
PHYS_A = 0.01372
ALPHA = 1.0
LINEARIZED = FALSE
PAIRWISE = FALSE
RANDOMWALK = FALSE
ENERGYSCALES = [1.0]
DISTANCESCALE = 1.0
""" + COMMON_TAIL


def c3():
    # 280 x 80 graded mesh on [0,7]x[0,2]; grading toward the wall/channel interfaces
    xb = [0.0, 2.5, 3.0, 4.0, 4.5, 7.0]
    xc = [90, 30, 40, 30, 90]
    yb = [0.0, 0.5, 1.0, 1.5, 2.0]
    yc = [20, 20, 20, 20]
    fine_x = {2.5, 3.0, 4.0, 4.5}
    fine_y = {0.5, 1.0, 1.5}
    xn = graded_nodes(xb, xc, fine_x)
    yn = graded_nodes(yb, yc, fine_y)
    assert len(xn) == 281 and len(yn) == 81
    regs = ("[((0.0, 7.0),(0.0,2.0)) , ((0.0, 3.0),(0.0,0.5)), ((2.5, 3.0),(0.0,1.5)), "
            "((2.5, 4.5),(1.0,1.5)), ((4.0, 4.5),(0.0,1.5)), ((4.0, 7.0),(0.0,0.5))]")
    return f"""NAME = Channel 2D
# Workload 3: two-material channel (crooked-pipe layout) on a 280x80 graded mesh over
# [0,7]x[0,2].  Walls sigma_a = 200, channel 0.2, bee = 0.1.  Left face y in [0,0.5] at
# T = 0.5.  FP16 position/direction tracking with FP32 weights.
# Pinned: SEED 12345, NINPUT 1000000, CELLMIN 1, PHYS_C 299.8, PHYS_A 0.01372.
PRECISION = FLOAT16
SEED = 12345
GEOMETRY = 2D
MESHTYPE = NONUNIFORM
XMESHNODES = {fmt_list(xn)}
YMESHNODES = {fmt_list(yn)}
SIGMA_A_REGS = {regs}
SIGMA_A_VALS = [200.0, 0.2, 0.2, 0.2, 0.2, 0.2]
SIGMA_A_POWERS = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
SIGMA_S_REGS = []
SIGMA_S_VALS = [0.0]
SIGMA_S_POWERS = [0.0]
BEE_REGS = []
BEE_VALS = [0.1]
RADSOURCE_REGS = []
RADSOURCE_VALS = [0]
LEFTBC = VACUUM
RIGHTBC = VACUUM
TOPBC = VACUUM
BOTTOMBC = REFLECT
TIMESTEPPING = CONSTANT
DT = 0.01
ENDTIME = 5.0
NINPUT = 1000000
NMAX = 50000000
CELLMIN = 1
T_INIT = 0.05
T_SURFACE_VALS = [(0.0),(0.0),(0.5,0.0),(0.0)]
T_SURFACE_REGS = [(7.0),(7.0),(0.5,2.0),(2.0)]
PHYS_C = 299.8
PHYS_A = 0.01372
ALPHA = 1.0
LINEARIZED = FALSE
PAIRWISE = FALSE
RANDOMWALK = FALSE
ENERGYSCALES = [1.0]
DISTANCESCALE = 1.0
""" + COMMON_TAIL


def c4(seed=4):
    rng = np.random.default_rng(seed)
    nb = 40
    ends = [round(0.25 * (i + 1), 6) for i in range(nb)]        # 50 cells of dx 0.005 per block
    thick = np.arange(nb) % 2 == 1                               # block 1 thin (source side), then alternate
    sa = np.where(thick, 10 ** rng.uniform(2.0, 4.0, nb), 10 ** rng.uniform(-1.0, 0.0, nb))
    ss = sa / 2
    return f"""NAME = Thick Thin Blocks
# Workload 4: 2000 cells over [0,10]; 40 alternating 50-cell blocks, thin sigma_a
# log-uniform in [0.1,1], thick log-uniform in [1e2,1e4], sigma_s = sigma_a/2 (seed {seed}).
# Left surface T = 1.  Pinned: bee 0.1, T_INIT 1.0 (hot, so the thick blocks are diffusive:
# collision optical thickness up to ~75 per cell), dt 1e-3 to t = 0.05, CELLMIN 1.
PRECISION = FLOAT32
SEED = 4444
GEOMETRY = 1D
MESHTYPE = UNIFORM
XSIZE = 10.0
DX = 0.005
SIGMA_A_REGS = {fmt_list(ends)}
SIGMA_A_VALS = {fmt_list(sa)}
SIGMA_A_POWERS = {fmt_list(np.zeros(nb))}
SIGMA_S_REGS = {fmt_list(ends)}
SIGMA_S_VALS = {fmt_list(ss)}
SIGMA_S_POWERS = {fmt_list(np.zeros(nb))}
RADSOURCE_REGS = []
RADSOURCE_VALS = [0.0]
BEE_REGS = []
BEE_VALS = [0.1]
LEFTBC = VACUUM
RIGHTBC = VACUUM
TIMESTEPPING = CONSTANT
DT = 0.001
ENDTIME = 0.05
NINPUT = 2000000
NMAX = 20000000
CELLMIN = 1
T_INIT = 1.0
T_SURFACE_VALS = [1.0,0.0]
T_SURFACE_REGS = []
PHYS_C = 299.8
PHYS_A = 0.01372
ALPHA = 1.0
LINEARIZED = FALSE
PAIRWISE = FALSE
RANDOMWALK = FALSE
ENERGYSCALES = [1.0]
DISTANCESCALE = 1.0
""" + COMMON_TAIL


def c5(seed=7):
    rng = np.random.default_rng(seed)
    nbx = 64
    w = 1.0 / nbx
    rects, sav, sap, ssv = [], [], [], []
    for j in range(nbx):
        for i in range(nbx):
            rects.append(f"(({i * w:.10g}, {(i + 1) * w:.10g}),({j * w:.10g},{(j + 1) * w:.10g}))")
    n = nbx * nbx
    sa = 10 ** rng.uniform(-1.0, 4.0, n)
    ss = rng.uniform(0.0, 1.0, n) * sa
    powr = np.where(rng.permutation(n) < n // 2, -3.0, 0.0)
    return f"""NAME = Blocks 2D
# Workload 5: 1024x1024 uniform mesh over [0,1]^2 with 64x64 random material blocks of
# 16x16 cells (seed {seed}): sigma_a log-uniform in [0.1,1e4], sigma_s = U[0,1] sigma_a,
# T^-3 power law in half of the blocks, bee = 0.1.  FP16 tracking.
# Pinned: ramp dt0 1e-4, K 1.1, dtmax 1e-3; NINPUT 20000000; CELLMIN 1; left face T = 1.
PRECISION = FLOAT16
SEED = 7
GEOMETRY = 2D
MESHTYPE = UNIFORM
XSIZE = 1.0
YSIZE = 1.0
DX = 0.0009765625
DY = 0.0009765625
SIGMA_A_REGS = [{", ".join(rects)}]
SIGMA_A_VALS = {fmt_list(sa)}
SIGMA_A_POWERS = {fmt_list(powr)}
SIGMA_S_REGS = [{", ".join(rects)}]
SIGMA_S_VALS = {fmt_list(ss)}
SIGMA_S_POWERS = {fmt_list(powr)}
BEE_REGS = []
BEE_VALS = [0.1]
RADSOURCE_REGS = []
RADSOURCE_VALS = [0]
LEFTBC = VACUUM
RIGHTBC = VACUUM
TOPBC = VACUUM
BOTTOMBC = VACUUM
TIMESTEPPING = RAMP
DT0 = 0.0001
K = 1.1
DTMAX = 0.001
ENDTIME = 1.0
NINPUT = 20000000
NMAX = 200000000
CELLMIN = 1
T_INIT = 0.01
T_SURFACE_VALS = [(0.0),(0.0),(1.0),(0.0)]
T_SURFACE_REGS = [(1.0),(1.0),(1.0),(1.0)]
PHYS_C = 299.8
PHYS_A = 0.01372
ALPHA = 1.0
LINEARIZED = FALSE
PAIRWISE = FALSE
RANDOMWALK = FALSE
ENERGYSCALES = [1.0]
DISTANCESCALE = 1.0
""" + COMMON_TAIL


def main():
    for name, fn in (("c1_slab_fp64.txt", c1), ("c2_marshak_fp32.txt", c2), ("c3_channel_fp16.txt", c3),
                     ("c4_thickthin_fp32.txt", c4), ("c5_blocks2d_fp16.txt", c5)):
        with open(os.path.join(HERE, name), "w") as f:
            f.write(fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
