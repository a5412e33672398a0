"""Host side (deck parser, mesh, update, timestep) and the C-ABI library surface.  CPU only.

test_reference_* restate the reference's own test set (test/runtests.jl:37-76) on its
fixture deck (tests/golden/test_input.txt).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import MixedPrecisionIMC as M
from MixedPrecisionIMC import Constants, Input, Mesh, Update, _hip

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _simvars(dt=0.01, geometry="2D"):
    return M.SimVars(t=0.0, dt=dt, dt0=0.0, k=0.0, dtmax=0.0, t_end=1.0, timesteps=[], iterations=0,
                     n_input=500, n_max=100000, cellmin=1.0, pairwise="FALSE", BC=(1, 0, 1, 1), precision=64,
                     geometry=geometry)


def test_reference_inputs_and_mesh():          # test/runtests.jl:37-52
    inputs = Input.readInputs(os.path.join(GOLD, "test_input.txt"))
    assert inputs and isinstance(inputs, dict) and "PRECISION" in inputs
    Constants.set_constants(inputs)
    assert isinstance(Constants.phys_c, float)
    mesh = Mesh.mesh_generation(inputs)
    assert isinstance(mesh.nodes, tuple) and len(mesh.nodes) == 2
    assert all(n.dtype == np.float64 for n in mesh.nodes)
    np.testing.assert_allclose(mesh.dx, np.full(10, 0.1))
    np.testing.assert_allclose(mesh.dy, np.full(4, 0.5))


def test_reference_update():                    # test/runtests.jl:59-76
    inputs = Input.readInputs(os.path.join(GOLD, "test_input.txt"))
    mesh = Mesh.mesh_generation(inputs)
    Constants.set_constants(inputs)
    Update.update(inputs, mesh, _simvars())
    assert np.any(mesh.beta != 0)
    assert np.any(mesh.sigma_a[..., 0] != 0)
    assert np.all(mesh.sigma_s[..., 0] == 0)
    # closed form of the Fleck factor for this deck: f = 1/(1 + alpha*beta*c*dt*sigma_a)
    beta = 4 * Constants.phys_a * 1.0 / 1.0
    np.testing.assert_allclose(mesh.fleck, 1 / (1 + 1.0 * beta * Constants.phys_c * 0.01 * 1.0), rtol=1e-15)


def test_parser_multiline_arrays_and_tables():
    inputs = Input.readInputs(os.path.join(GOLD, "MarshakWave.txt"))
    assert inputs["PRECISION"] == 64
    assert len(inputs["XBENCH"]) == 300 and len(inputs["YBENCH"]) == 300
    assert inputs["XBENCH"][0] == 0.00025 and inputs["YBENCH"][-1] == 0.010376228
    assert inputs["ENERGYSCALES"] == ["1.0"]              # single element stays a string vector
    su = Input.readInputs(os.path.join(GOLD, "SuOlson.txt"))
    assert su["PRECISION"] == 16 and len(su["XBENCH"]) == 12


def test_crooked_pipe_regions_and_surfaces():
    inputs = Input.readInputs(os.path.join(GOLD, "CrookedPipe.txt"))
    mesh = Mesh.mesh_generation(inputs)
    nx, ny = mesh.Ncells
    assert (nx, ny) == (106, 47)
    xc, yc = mesh.centers
    sa = mesh.sigma_a[..., 1]
    # rectangles are painted in order: walls 2000, channel 0.2 (CrookedPipe.txt SIGMA_A_REGS)
    def at(x, y):
        return sa[np.searchsorted(mesh.nodes[0], x) - 1, np.searchsorted(mesh.nodes[1], y) - 1]
    assert at(1.0, 0.25) == 0.2 and at(1.0, 1.0) == 2000.0 and at(2.7, 1.2) == 0.2
    assert at(3.5, 1.25) == 0.2 and at(3.5, 0.25) == 2000.0 and at(6.0, 0.25) == 0.2
    left = mesh.temp_surf[2]
    np.testing.assert_array_equal(left, np.where(mesh.nodes[1][1:] <= 0.5, 0.5, 0.0))
    assert np.all(mesh.temp_surf[0] == 0) and np.all(mesh.temp_surf[3] == 0)


def test_tuple_literals_cannot_execute_code(tmp_path):
    deck = tmp_path / "evil.txt"
    deck.write_text("GEOMETRY = 2D\nSIGMA_A_REGS = [(__import__('os').getcwd(), 1)]\n")
    with pytest.raises(ValueError):
        Input.readInputs(str(deck))


@pytest.mark.parametrize("deck", ["c1_slab_fp64", "c2_marshak_fp32", "c3_channel_fp16", "c4_thickthin_fp32"])
def test_workload_decks_parse(deck):
    inputs = Input.readInputs(os.path.join(ROOT, "decks", deck + ".txt"))
    Constants.set_constants(inputs)
    mesh = Mesh.mesh_generation(inputs)
    assert np.all(mesh.dx > 0)
    sv = _simvars(geometry=mesh.geometry)
    Update.update(inputs, mesh, sv)
    assert np.all((mesh.fleck > 0) & (mesh.fleck <= 1))


def test_marshak_fleck_is_temperature_independent():
    inputs = Input.readInputs(os.path.join(ROOT, "decks", "c2_marshak_fp32.txt"))
    Constants.set_constants(inputs)
    mesh = Mesh.mesh_generation(inputs)
    mesh.temp = np.linspace(0.01, 1.0, mesh.Ncells)
    sv = _simvars(dt=1e-4, geometry="1D")
    Update.update(inputs, mesh, sv)
    expect = 1 / (1 + 299.8 * 1e-4 * 4 * 0.01372 * 10 / 0.3)
    np.testing.assert_allclose(mesh.fleck, expect, rtol=1e-12)


def test_timestep_constant_and_ramp():
    sv = _simvars(dt=0.3)
    sv.t_end = 1.0
    ts = []
    while sv.t <= sv.t_end:
        ts.append(sv.t)
        M.timestep("CONSTANT", sv)
    assert ts == pytest.approx([0.0, 0.3, 0.6, 0.9, 1.0])
    sv = _simvars(dt=0.1)
    sv.dtmax, sv.k, sv.t_end = 0.25, 2.0, 1.0
    ts = []
    while sv.t <= sv.t_end:
        ts.append(sv.t)
        M.timestep("RAMP", sv)
    assert ts == pytest.approx([0.0, 0.2, 0.45, 0.7, 0.95, 1.0])


def test_abi_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "imc_hip.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(imc_\w+)\s*\(", header, re.M)))
    assert len(declared) >= 16
    L = _hip.lib()                       # loads the built library (no GPU call is made)
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) <= set(_hip.SYMBOLS)


def test_abi_struct_layouts_match_the_c_header(tmp_path):
    """sizeof and field offsets of every struct in include/imc_hip.h, as gcc lays them out,
    against the ctypes mirrors."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"imc_config": _hip.ImcConfig, "imc_timing": _hip.ImcTiming,
               "imc_surrogate_box": _hip.SurrogateBox, "imc_material": _hip.ImcMaterial,
               "imc_step_result": _hip.ImcStepResult}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "imc_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cname, fname, val = line.split()
            got[(cname, fname)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_abi_config_struct_layout():
    # imc_config as declared in include/imc_hip.h: 10 int32 + 2 int32 + 16 doubles + 3 doubles + int64 + 8 int32
    assert ctypes.sizeof(_hip.ImcConfig) == 4 * 12 + 8 * 19 + 8 + 4 * 8
    assert _hip.ImcConfig.energyscales.offset == 48


def test_benchmark_comparison_and_saved_arrays(tmp_path):
    """Output.plotting at t_end with BENCHMARK = TRUE and SAVEVARS = TRUE (the reference's Su-Olson
    deck): the run's PLOTVARS field (radenergydens) is read at each XBENCH abscissa by linear
    interpolation between cell centres, and the arrays of imc_output.jl:51-57 plus the table and
    the interpolated values are saved.  A field that is linear in x is interpolated exactly."""
    from MixedPrecisionIMC import Constants, Input, Mesh, Output
    inp = Input.readInputs(os.path.join(ROOT, "tests", "golden", "SuOlson.txt"))
    Constants.set_constants(inp)
    mesh = Mesh.mesh_generation(inp)
    mesh.radenergydens = 2.0 - 0.25 * mesh.centers
    mesh.temp_saved = [mesh.temp.copy()]
    mesh.radenergy_saved = [mesh.radenergydens.copy()]
    mesh.matenergy_saved = [mesh.matenergydens.copy()]
    cmp = Output.benchmark_comparison(inp, mesh)
    xb, yb = np.array(inp["XBENCH"]), np.array(inp["YBENCH"])
    inside = (xb >= mesh.centers[0]) & (xb <= mesh.centers[-1])
    np.testing.assert_allclose(cmp["imc"][inside], 2.0 - 0.25 * xb[inside], rtol=1e-12)
    np.testing.assert_allclose(cmp["diff"], cmp["imc"] - yb)
    assert cmp["field"] == "radenergydens" and cmp["max_abs_diff"] == pytest.approx(np.max(np.abs(cmp["diff"])))
    sv = M.SimVars(t=1.0, dt=0.01, dt0=0.0, k=0.0, dtmax=0.0, t_end=1.0, timesteps=[0.01, 1.0], iterations=0,
                   n_input=1, n_max=1, cellmin=1.0, pairwise="FALSE", BC=(0, 0, 0, 0), precision=64, geometry="1D")
    Output.plotting(inp, mesh, sv, outdir=str(tmp_path))
    files = list(tmp_path.glob("*.npz"))
    assert len(files) == 1
    with np.load(files[0]) as z:
        for key in ("timesteps", "centers", "temp_saved", "radenergy_saved", "matenergy_saved", "xbench", "ybench",
                    "imc_at_xbench"):
            assert key in z.files, key
        np.testing.assert_array_equal(z["imc_at_xbench"], cmp["imc"])
    inp["BENCHMARK"] = "FALSE"
    assert Output.benchmark_comparison(inp, mesh) is None
