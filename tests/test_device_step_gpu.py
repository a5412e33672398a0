"""The device-resident step (imc_set_material + imc_step) against the host-update step.

Both run the same deck with the same seed.  The host path does Update.update, Tally.tally and
the energy check in numpy FP64 and uploads the fields every step; the device path does the
same arithmetic in FP64 kernels, operation for operation.  Births and histories are then
identical, but transport tallies are added with unordered atomics (LDS or global), so two runs
agree only to the rounding of the tally type: the cases run with FP64 tracking and tallies,
where that is ~1e-16 and cannot flip a birth count, and compare every step to 1e-12.  pow()
(T^p with p != 0, and matE^(1/4) when linearized) may differ by an ulp between the device
library and glibc, which the same tolerance covers.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _deck(name, **over):
    from MixedPrecisionIMC import Input
    inp = Input.readInputs(os.path.join(ROOT, "decks", name))
    inp.update({k: (v if isinstance(v, list) else str(v)) for k, v in over.items()})
    return inp


CASES = {
    "marshak": lambda: _deck("c2_marshak_fp32.txt", NINPUT=200000, NMAX=1000000),
    "marshak_pow": lambda: _deck("c2_marshak_fp32.txt", NINPUT=200000, NMAX=1000000, NAME="Marshak Pow"),
    "slab_linearized_3scales": lambda: _deck("c1_slab_fp64.txt", NINPUT=20000, NMAX=200000,
                                             ENERGYSCALES=[1024.0, 8.0, 1.0]),
    "channel2d": lambda: _deck("c3_channel_fp16.txt", NINPUT=200000, NMAX=1000000),
    "channel2d_2scales": lambda: _deck("c3_channel_fp16.txt", NINPUT=100000, NMAX=500000,
                                       ENERGYSCALES=[64.0, 1.0]),
}
RTOL = 1e-12
STEPS = 4


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_step_matches_host_step(case):
    import MixedPrecisionIMC as M
    make, rtol = CASES[case], RTOL
    print(f"[device step] {case}", flush=True)
    inp_h, inp_d = make(), make()
    ts = str(inp_h["TIMESTEPPING"]).upper()
    mesh_h, sv_h, eng_h = M.setup(inp_h, precision=64)
    mesh_d, sv_d, eng_d = M.setup(inp_d, precision=64)
    for k in range(STEPS):
        err_h = M.step(inp_h, mesh_h, sv_h, eng_h, ts)
        err_d = M.step_device(inp_d, mesh_d, sv_d, eng_d, ts)
        assert eng_d.n == eng_h.n, (k, eng_d.n, eng_h.n)
        assert sv_d.iterations == sv_h.iterations
        for name in ("temp", "radenergydens", "matenergydens"):
            a, b = getattr(mesh_d, name), getattr(mesh_h, name)
            np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * np.abs(b).max(), err_msg=f"{name} step {k}")
        inc_d, inc_h = mesh_d.energyincrease_saved[-1], mesh_h.energyincrease_saved[-1]
        np.testing.assert_allclose(inc_d, inc_h, rtol=rtol, atol=rtol * np.abs(inc_h).max())
        assert mesh_d.totalenergy == pytest.approx(mesh_h.totalenergy, rel=rtol)
        assert mesh_d.totalenergydep == pytest.approx(mesh_h.totalenergydep, rel=rtol)
        assert abs(err_d - err_h) < 1e-12, (k, err_d, err_h)
        assert abs(err_d) < 1e-5
    td, th = eng_d.timing(), eng_h.timing()
    assert td["histories"] == th["histories"] and td["segments"] == th["segments"]
    eng_h.close()
    eng_d.close()


def test_device_step_requires_material_and_rejects_bad_input():
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC._hip import ImcError
    inp = _deck("c2_marshak_fp32.txt", NINPUT=20000, NMAX=100000)
    mesh, sv, eng = M.setup(inp)
    with pytest.raises(ImcError):
        eng.step(0, 1e-4, True, 1000, 10000, 1.0)
    bad = mesh.temp.copy()
    bad[3] = np.nan
    with pytest.raises(ImcError):
        eng.set_material(mesh.sigma_a[..., 1], mesh.sigma_a[..., 2], mesh.sigma_s[..., 1], mesh.sigma_s[..., 2],
                         mesh.bee, bad, mesh.radsource, [np.array([1.0]), np.array([0.0])], 1.0, False, True)
    eng.close()
