"""The analytic random walk (imc_random_walk_enable) on the GPU.

* With a threshold no cell reaches, every lane takes the plain event loop: the random-walk
  kernel then reproduces plain transport (FP64: identical populations and segment counts,
  temperatures to 1e-12, the rounding floor of unordered tally atomics).
* On the thick/thin block workload (12 steps, 4 seeds each) the random walk reproduces the
  plain-transport temperature profile within 5 standard errors + 2 % per cell (the diffusion
  limit is an approximation near the walk's threshold), with fewer event-loop passes, and
  every step balances energy (the walk conserves energy exactly).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c4_inputs(n_input=200000, n_max=2000000):
    from MixedPrecisionIMC import Input
    inp = Input.readInputs(os.path.join(ROOT, "decks", "c4_thickthin_fp32.txt"))
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(n_max)
    return inp


def test_random_walk_without_eligible_cells_is_plain_transport():
    import MixedPrecisionIMC as M
    print("[random walk] r_min 1e9 vs plain, FP64", flush=True)
    runs = []
    for rw in (False, True):
        inp = _c4_inputs(n_input=50000, n_max=500000)
        mesh, sv, parts = M.setup(inp, precision=64, seed=31)
        if rw:
            parts.random_walk_enable(True, r_min=1e9)
        for _ in range(3):
            M.step(inp, mesh, sv, parts, "CONSTANT")
        tm = parts.timing()
        runs.append((mesh.temp.copy(), parts.n, tm["segments"], tm["surrogate_events"]))
        parts.close()
    (tp, n_p, sp, _), (tr, n_r, sr, er) = runs
    assert n_p == n_r and sp == sr and er == 0
    np.testing.assert_allclose(tr, tp, rtol=1e-12, atol=0)


def test_random_walk_reproduces_plain_tallies():
    import MixedPrecisionIMC as M
    steps, seeds = 12, 4
    plain, walk, seg_p, seg_w, ev_w = [], [], 0, 0, 0
    for sd in range(seeds):
        print(f"[random walk] thick/thin seed {sd}", flush=True)
        for rw in (False, True):
            inp = _c4_inputs()
            mesh, sv, parts = M.setup(inp, seed=(900 if rw else 500) + sd)
            if rw:
                parts.random_walk_enable(True, r_min=5.0)
            for _ in range(steps):
                M.step(inp, mesh, sv, parts, "CONSTANT")
            tm = parts.timing()
            assert max(abs(e) for e in sv.energy_errors) < 2e-4
            if rw:
                seg_w += tm["segments"]; ev_w += tm["surrogate_events"]
                walk.append(mesh.temp.copy())
            else:
                seg_p += tm["segments"]
                plain.append(mesh.temp.copy())
            parts.close()
    plain, walk = np.array(plain), np.array(walk)
    mp, mw = plain.mean(0), walk.mean(0)
    se = np.sqrt(plain.var(0, ddof=1) / seeds + walk.var(0, ddof=1) / seeds)
    ratio = np.abs(mp - mw) / (5 * se + 0.02 * np.abs(mp))
    worst = int(np.argmax(ratio))
    msg = (f"segments plain {seg_p} random walk {seg_w} (walk events {ev_w}); worst cell {worst}: "
           f"plain {mp[worst]:.5f} walk {mw[worst]:.5f} se {se[worst]:.2e} ratio {ratio[worst]:.3f}; "
           f"max rel diff {np.max(np.abs(mp - mw) / mp):.4f}")
    print(msg)
    assert ev_w > 0, msg
    assert ratio[worst] < 1.0, msg
    assert seg_w < seg_p, msg


def _blocks2d_inputs(nx=128, n_input=200000):
    """The large 2D block deck (64 x 64 random material blocks over [0,1]^2) on an nx x nx mesh."""
    from MixedPrecisionIMC import Input
    inp = Input.readInputs(os.path.join(ROOT, "decks", "c5_blocks2d_fp16.txt"))
    inp["DX"] = inp["DY"] = str(1.0 / nx)
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(10 * n_input)
    return inp


def test_random_walk_2d_without_eligible_cells_is_plain_transport():
    import MixedPrecisionIMC as M
    print("[random walk 2D] r_min 1e9 vs plain, FP64", flush=True)
    runs = []
    for rw in (False, True):
        inp = _blocks2d_inputs(nx=64, n_input=50000)
        mesh, sv, parts = M.setup(inp, precision=64, seed=41)
        if rw:
            parts.random_walk_enable(True, r_min=1e9)
        for _ in range(3):
            M.step_device(inp, mesh, sv, parts, "RAMP")
        tm = parts.timing()
        runs.append((mesh.temp.copy(), parts.n, tm["segments"], tm["surrogate_events"]))
        parts.close()
    (tp, n_p, sp, _), (tr, n_r, sr, er) = runs
    assert n_p == n_r and sp == sr and er == 0
    np.testing.assert_allclose(tr, tp, rtol=1e-12, atol=0)


def test_random_walk_2d_reproduces_plain_heating():
    """128 x 128 cells, FP32, 8 steps, 4 seeds each: the temperature rise over T_INIT averaged
    over 16 x 16-cell regions, in the regions that heat (rise above 5% of the largest), agrees
    within 5 standard errors + 3%, with a tenth of the event-loop passes or fewer."""
    import MixedPrecisionIMC as M
    nx, steps, seeds = 128, 8, 4
    rise = {False: [], True: []}
    segs = {False: 0, True: 0}
    walk_events = 0
    for sd in range(seeds):
        print(f"[random walk 2D] blocks 128^2 seed {sd}", flush=True)
        for rw in (False, True):
            inp = _blocks2d_inputs(nx=nx)
            mesh, sv, parts = M.setup(inp, precision=32, seed=(700 if rw else 300) + sd)
            if rw:
                parts.random_walk_enable(True, r_min=5.0)
            for _ in range(steps):
                M.step_device(inp, mesh, sv, parts, "RAMP")
            tm = parts.timing()
            assert max(abs(e) for e in sv.energy_errors) < 2e-4
            segs[rw] += tm["segments"]
            if rw:
                walk_events += tm["surrogate_events"]
            t = np.asarray(mesh.temp).reshape(nx, nx, order="F") - 0.01
            rise[rw].append(t.reshape(8, nx // 8, 8, nx // 8).mean(axis=(1, 3)).ravel())
            parts.close()
    p, w = np.array(rise[False]), np.array(rise[True])
    mp, mw = p.mean(0), w.mean(0)
    se = np.sqrt(p.var(0, ddof=1) / seeds + w.var(0, ddof=1) / seeds)
    hot = mp > 0.05 * mp.max()
    ratio = np.abs(mp - mw)[hot] / (5 * se[hot] + 0.03 * mp[hot])
    msg = (f"segments plain {segs[False]} walk {segs[True]} (walk events {walk_events}); {hot.sum()} heated regions, "
           f"worst ratio {ratio.max():.3f}, mean relative difference {np.mean((mw - mp)[hot] / mp[hot]):.4f}")
    print(msg)
    assert walk_events > 0 and hot.sum() >= 4, msg
    assert ratio.max() < 1.0, msg
    assert segs[True] * 10 < segs[False], msg


@pytest.mark.parametrize("prec", [64, 32, 16])
@pytest.mark.parametrize("geom", ["1D", "2D"])
def test_random_walk_every_precision_balances_energy(prec, geom):
    """The walk kernels of every tracking precision: walk events happen and every step balances
    energy (the walk conserves it exactly; FP16/FP32 leave the tally rounding)."""
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Input
    print(f"[random walk] {geom} FP{prec} energy balance", flush=True)
    if geom == "1D":
        inp = _c4_inputs(n_input=100000, n_max=1000000)
        ts = "CONSTANT"
    else:
        inp = Input.readInputs(os.path.join(ROOT, "decks", "c3_channel_fp16.txt"))
        inp["SIGMA_A_VALS"] = [5000.0] + list(inp["SIGMA_A_VALS"][1:])   # walls thick enough to walk in
        inp["NINPUT"] = "100000"
        inp["NMAX"] = "1000000"
        ts = "RAMP" if str(inp["TIMESTEPPING"]).upper() == "RAMP" else "CONSTANT"
    mesh, sv, parts = M.setup(inp, precision=prec, seed=77)
    parts.random_walk_enable(True, r_min=3.0)
    for _ in range(4):
        M.step_device(inp, mesh, sv, parts, ts)
    tm = parts.timing()
    parts.close()
    tol = {64: 1e-9, 32: 2e-4, 16: 2e-3}[prec]
    assert tm["surrogate_events"] > 0
    assert max(abs(e) for e in sv.energy_errors) < tol, sv.energy_errors
