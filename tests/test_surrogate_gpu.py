"""The learned thick-cell surrogate on the GPU.

* forward / gradient of the BF16-MFMA kernels against the FP64 reference network
  (tests/mlp_oracle.py) -- with the same BF16 operand rounding: outputs within 2e-3,
  gradients within 2e-2 relative per layer; against pure FP64: cosine > 0.99;
* the GPU training-data generator against the oracle's plain transport of the same
  single-cell problem (energy-weighted outcomes within 5 standard errors);
* training lowers the loss and the network matches held-out Monte Carlo frequencies;
* transport with the surrogate conserves energy and reproduces the plain tallies.
"""
import numpy as np
import pytest

import mlp_oracle as R
import oracle as O
from cases import engine
from MixedPrecisionIMC import Surrogate

pytestmark = pytest.mark.gpu


def _dump(name, obj):
    """Keeps the raw numbers of a statistical test for inspection, in the folder named by the
    IMC_TEST_OUTPUT_DIR environment variable when it is set."""
    import json
    import os
    d = os.environ.get("IMC_TEST_OUTPUT_DIR", "")
    if d and os.path.isdir(d):
        with open(os.path.join(d, name + ".json"), "w") as f:
            json.dump(obj, f)


def _eng():
    case = dict(geometry=1, prec=32, nx=4, ny=1, seed=3, bc=(0, 0, 0, 0), scales=(1.0,), dt=1.0, phys_c=1.0,
                phys_a=1.0, dx=np.ones(4), dy=None)
    return engine(case, capacity=1 << 12)


def test_forward_matches_reference():
    e = _eng()
    p = R.init_params(5)
    e.surrogate_configure()
    e.surrogate_set_params(p.astype(np.float32))
    rng = np.random.default_rng(2)
    for n in (1000, 37, 64, 130):
        X = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
        got = e.surrogate_forward(X)
        ref, _ = R.forward(p.astype(np.float32).astype(np.float64), X, bf=True)
        np.testing.assert_allclose(got, ref, atol=2e-3)
    e.close()


def test_gradient_matches_reference():
    e = _eng()
    p = R.init_params(6).astype(np.float32)
    e.surrogate_configure()
    e.surrogate_set_params(p)
    rng = np.random.default_rng(4)
    n = 640
    X = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    stay = rng.integers(0, 2, n)
    Y = np.stack([stay, rng.integers(0, 2, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)], 1).astype(np.float32)
    g, loss = e.surrogate_gradient(X, Y)
    rl, rg = R.loss_and_grad(p.astype(np.float64), X, Y, bf=True)
    fl, fg = R.loss_and_grad(p.astype(np.float64), X, Y, bf=False)
    assert loss == pytest.approx(rl, rel=1e-3)
    W = R.unpack
    for name, a, b in zip(("W1", "b1", "W2", "b2", "W3", "b3"), W(g), W(rg)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-2, (name, err)
    cos = float(g @ fg / (np.linalg.norm(g) * np.linalg.norm(fg)))
    assert cos > 0.99
    e.close()


def _oracle_detector_run(tau_cell, tau_b, kappa, n, seed=5):
    """Plain transport (oracle) of particles started uniformly in x and mu inside a cell of
    optical thickness tau_cell (sigma_c = 1, f sigma_a = kappa, sigma_s = 0), flanked by
    black 'detector' cells; census after path tau_B.  Returns energy per particle that
    left through the left / right face, stayed to census, and was deposited in the cell."""
    dx = np.array([1.0, tau_cell, 1.0])
    sa = np.array([1e6, 1.0 + kappa, 1e6])
    fl = np.array([1.0, kappa / (1.0 + kappa), 1.0])
    op = O.OracleProblem(geometry=1, prec=64, nx=3, bc=(0, 0, 0, 0), seed=seed, phys_c=1.0)
    st = op.step(dt=tau_b, dx=dx, temp=np.ones(3), fleck=fl, sigma_a=sa, sigma_s=np.zeros(3),
                 radsource=np.zeros(3), tsurf=[np.zeros(1), np.zeros(1)])
    rng = np.random.default_rng(seed)
    mu = rng.uniform(-1, 1, n)
    rec = np.stack([np.ones(n), np.zeros(n), np.full(n, 2.0), rng.uniform(0, 1, n) * tau_cell, mu, np.ones(n),
                    np.ones(n), np.ones(n), np.ones(n)], 1)
    out, dep, lost, segs = op.transport(st, 0, rec, nthreads=0)
    alive = out[:, 7] != -1.0
    e_census = out[alive, 6].sum() / n
    return dep[0] * dx[0] / n, dep[2] * dx[2] / n, e_census, dep[1] * dx[1] / n


# kappa * tau_B < ln(100) so the oracle's 1% weight cut-off never ends a history first
@pytest.mark.parametrize("tau_cell,tau_b,kappa", [(30.0, 500.0, 1e-3), (60.0, 100.0, 3e-2), (25.0, 2000.0, 2e-4)])
def test_generator_matches_oracle_plain_transport(tau_cell, tau_b, kappa):
    e = _eng()
    eps = 1e-4
    e.surrogate_configure((tau_cell, tau_cell * (1 + eps)), (tau_b, tau_b * (1 + eps)), (kappa, kappa * (1 + eps)),
                          depth_min=0.0)
    n = 200000
    e.surrogate_generate(n, seed=11)
    _, T = e.surrogate_dataset(n)
    stay, right, frac, dep = T[:, 0], T[:, 1], T[:, 2], T[:, 3]
    w_exit = (1.0 - dep) * (1.0 - stay)
    gen = {"left": w_exit * (1 - right), "right": w_exit * right, "census": stay * np.exp(-kappa * tau_b),
           "deposit": (1 - stay) * dep + stay * (1 - np.exp(-kappa * tau_b))}
    no = 40000
    ol, orr, oc, od = _oracle_detector_run(tau_cell, tau_b, kappa, no)
    ref = {"left": ol, "right": orr, "census": oc, "deposit": od}
    for key, vals in gen.items():
        se = np.sqrt(vals.var() / n + vals.var() / no)
        assert abs(vals.mean() - ref[key]) < 5 * se + 1e-3 * abs(ref[key]), (key, vals.mean(), ref[key], se)
    e.close()


def test_training_learns_and_is_calibrated():
    e = _eng()
    info = Surrogate.train(e, n_samples=4_000_000, epochs=8, batch=8192, lr=3e-3, seed=1)
    losses = info["loss_per_epoch"]
    print("losses", losses, "generate", info["generate_s"], "train", info["train_s"])
    assert losses[-1] < losses[0]
    # calibration on fresh samples: binned predicted P(stay) against observed frequency
    e.surrogate_generate(200000, seed=99)
    X, T = e.surrogate_dataset(200000)
    P = e.surrogate_forward(X)
    bins = np.linspace(0, 1, 11)
    idx = np.digitize(P[:, 0], bins) - 1
    for b in range(10):
        sel = idx == b
        if sel.sum() > 2000:
            assert abs(P[sel, 0].mean() - T[sel, 0].mean()) < 0.03, (b, P[sel, 0].mean(), T[sel, 0].mean())
    ex = T[:, 0] == 0
    for o in (1, 2, 3):
        assert abs(P[ex, o].mean() - T[ex, o].mean()) < 0.02, (o, P[ex, o].mean(), T[ex, o].mean())
    e.close()


def _c4_inputs(n_input=200000, n_max=2000000):
    import os
    from MixedPrecisionIMC import Input
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inp = Input.readInputs(os.path.join(root, "decks", "c4_thickthin_fp32.txt"))
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(n_max)
    return inp


def test_surrogate_transport_reproduces_plain_tallies():
    """Thick/thin block workload, 12 steps: temperature profiles of plain and surrogate
    transport over 4 seeds each agree within 5 standard errors + 2 % per cell (the
    surrogate's census-position and exit-time models are approximations), and every
    step balances energy to FP32 rounding."""
    import MixedPrecisionIMC as M
    steps, seeds = 12, 4
    params = None
    plain, surr, seg_p, seg_s, t_p, t_s = [], [], 0, 0, 0.0, 0.0
    for sd in range(seeds):
        inp = _c4_inputs()
        mesh, sv, parts = M.setup(inp, seed=500 + sd)
        ts = "CONSTANT"
        for _ in range(steps):
            M.step(inp, mesh, sv, parts, ts)
        tm = parts.timing()
        seg_p += tm["segments"]; t_p += tm["ms_transport"]
        plain.append(mesh.temp.copy())
        assert max(abs(e) for e in sv.energy_errors) < 2e-4
        parts.close()
        inp = _c4_inputs()
        mesh, sv, parts = M.setup(inp, seed=700 + sd)
        if params is None:
            info = Surrogate.train(parts, seed=2)
            params = parts.surrogate_params()
            print("trained", info["loss_per_epoch"], info["generate_s"], info["train_s"])
        else:
            Surrogate.load(parts, params)
        for _ in range(steps):
            M.step(inp, mesh, sv, parts, ts)
        tm = parts.timing()
        seg_s += tm["segments"]; t_s += tm["ms_transport"]
        print("surrogate events", tm["surrogate_events"], "segments", tm["segments"])
        surr.append(mesh.temp.copy())
        assert max(abs(e) for e in sv.energy_errors) < 2e-4
        parts.close()
    plain, surr = np.array(plain), np.array(surr)
    mp, ms = plain.mean(0), surr.mean(0)
    se = np.sqrt(plain.var(0, ddof=1) / seeds + surr.var(0, ddof=1) / seeds)
    ratio = np.abs(mp - ms) / (5 * se + 0.02 * np.abs(mp))
    worst = int(np.argmax(ratio))
    msg = (f"segments plain {seg_p} surrogate {seg_s}; transport ms plain {t_p:.1f} surrogate {t_s:.1f}; "
           f"worst cell {worst}: plain {mp[worst]:.5f} surrogate {ms[worst]:.5f} se {se[worst]:.2e} "
           f"ratio {ratio[worst]:.3f}; max rel diff {np.max(np.abs(mp - ms) / mp):.4f}")
    print(msg)
    _dump("surrogate_e2e", {"plain": plain.tolist(), "surrogate": surr.tolist(), "seg_p": seg_p, "seg_s": seg_s,
                            "t_p": t_p, "t_s": t_s})
    assert ratio[worst] < 1.0, msg
    assert seg_s < seg_p, msg
