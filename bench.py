"""Throughput benchmark of the GPU particle engine on the Marshak-wave workload.

Workload (the metric's configuration): nonlinear Marshak wave in 1D, x in [0,0.2] with 200
uniform cells, sigma_a = 10 T^-3, bee = 0.3, left surface T = 1, ramped dt (1e-4 x 1.1 up to
1e-3), NINPUT 2e6, NMAX 1e7; FP32 tracking with compensated FP32 tallies (decks/c2_marshak_fp32.txt).

A step is one IMC timestep, run as one device-resident imc_step: FP64 update, sourcing,
transport, clean, census tally, FP64 temperature update and the energy-check sums on the GPU,
then the host's energy check and clock.  W warmup steps (the census builds to ~NMAX), then K
timed steps (`host_update_step` repeats the window with the update, temperature update and
energy check in numpy on the host).  value = histories transported in the K steps / wall time of
the K steps.  The roofline line is for the transport kernel (the dominant kernel), timed
with HIP events on the engine's stream: algorithmic bytes per launch =
S_state x (histories + census survivors).  The CPU baseline is the oracle (a C++ port of the
reference algorithm) run on the same deck for a bounded number of steps on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mixedprecisionimc.jl_amd")
sys.path.insert(0, PKG)

DECK = os.path.join(ROOT, "decks", "c2_marshak_fp32.txt")
DECK_SURR = os.path.join(ROOT, "decks", "c4_thickthin_fp32.txt")
DECK_SLAB = os.path.join(ROOT, "decks", "c1_slab_fp64.txt")
HBM_PEAK_GBS = 8000.0


TALLY_MODES = {0: "WAVEROW", 1: "BLOCKROW", 2: "GLOBAL", 3: "HASH"}


def profiled_valu_busy():
    """Fraction of the SIMDs' issue cycles the transport kernel keeps busy with VALU work on this
    workload, from the committed SQ counter summary (the kernel is issue-bound, not HBM-bound)."""
    path = os.path.join(ROOT, "profiles", "transport_marshak_valu.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("valu_busy_fraction:"):
                    return float(ln.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def profiled_traffic():
    """HBM bytes per transport launch on this workload, from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes, scaled by a plain copy's readings)."""
    path = os.path.join(ROOT, "profiles", "transport_marshak_hbm_traffic.md")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def _hot_s(tm) -> float:
    """GPU time of the particle loop (sourcing + transport + clean + census tally), seconds."""
    return (tm["ms_source"] + tm["ms_transport"] + tm["ms_clean"] + tm["ms_tally"]) / 1e3


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here as the CPU baseline only
    return O


def _threads() -> int:
    return min(16, os.cpu_count() or 1)


def oracle_window(inputs, mesh, warmup: int, steps: int, prec=None):
    """CPU rate of the oracle on timesteps warmup .. warmup+steps-1 of a deck (the earlier
    steps run untimed, so the sample has the same particle population as the GPU window)."""
    r = _oracle().run_deck(inputs, mesh, prec=prec, max_steps=warmup + steps, time_from=warmup,
                           nthreads=_threads())
    return r["histories"] / r["t_hot"], r


def gpu_run(inputs, warmup: int, steps: int, precision=None):
    """GPU hot-path rate on timesteps warmup .. warmup+steps-1 of a deck."""
    import MixedPrecisionIMC as M

    mesh, simvars, particles = M.setup(inputs, precision=precision)
    ts = str(inputs["TIMESTEPPING"]).upper()
    for _ in range(warmup):
        M.step_device(inputs, mesh, simvars, particles, ts)
    particles.synchronize()
    particles.reset_timing()
    done = 0
    while done < steps and simvars.t <= simvars.t_end:
        M.step_device(inputs, mesh, simvars, particles, ts)
        done += 1
    particles.synchronize()
    tm = particles.timing()
    particles.close()
    return tm["histories"] / _hot_s(tm), tm, done


def host_update_step(inputs, warmup: int, steps: int):
    """The same window with Update.update, Tally.tally and the energy check in numpy FP64 on the
    host and one field upload per step (the reference's module split), beside `value`, which
    runs each step as one device-resident imc_step."""
    import MixedPrecisionIMC as M

    mesh, simvars, particles = M.setup(inputs)
    ts = str(inputs["TIMESTEPPING"]).upper()
    for _ in range(warmup):
        M.step(inputs, mesh, simvars, particles, ts)
    particles.synchronize()
    particles.reset_timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        M.step(inputs, mesh, simvars, particles, ts)
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    particles.close()
    return {"histories_per_s": tm["histories"] / wall, "ms_per_step": 1e3 * wall / max(steps, 1)}


def cpu_baseline(inputs, mesh, warmup: int, steps: int, gpu_same):
    cpu, r = oracle_window(inputs, mesh, warmup, steps)
    if gpu_same is None:                       # the timed window was shorter than the sample
        rate, tm, _ = gpu_run(inputs, warmup, steps)
        gpu_same = (tm["histories"], tm["histories"] / rate)
    g_hist, g_hot = gpu_same
    return {"value": cpu, "unit": "histories/s", "cores": _threads(), "kind": "port",
            "sample": f"timesteps {warmup}..{warmup + steps - 1} of the same deck, i.e. the first {steps} "
                      f"steps of the GPU's timed window "
                      f"({r['histories']} histories, {r['segments']} segments; {warmup} earlier steps run untimed); "
                      f"C++ FP32 oracle, OpenMP over particles; sourcing+transport+clean+census time",
            "gpu_same_sample": {"histories_per_s_hot_path": g_hist / g_hot,
                                "speedup_hot_path": (g_hist / g_hot) / cpu}}


def fp64_leg(warmup: int, steps: int):
    """Plain FP64 tracking on the GPU against the FP64 oracle on the host cores: the Marshak
    deck at FLOAT64 on the same step window, and the reference's CPU slab case (400 cells,
    linearized, NINPUT 1e5, 100 steps) run to the end."""
    from MixedPrecisionIMC import Input, Mesh, Constants

    out = {}
    for name, deck, w, k in (("marshak_fp64", DECK, warmup, steps), ("slab_fp64", DECK_SLAB, 0, 1000)):
        inp = Input.readInputs(deck)
        inp["PRECISION"] = "64"
        g, tm, done = gpu_run(inp, w, k, precision=64)
        Constants.set_constants(inp)
        mesh = Mesh.mesh_generation(inp)
        c, r = oracle_window(inp, mesh, w, done, prec=64)
        out[name] = {"gpu_histories_per_s": g, "cpu_histories_per_s": c, "speedup": g / c, "cores": _threads(),
                     "steps": f"{w}..{w + done - 1}", "gpu_histories": tm["histories"],
                     "cpu_histories": r["histories"]}
    return out


def _c4(prec, n_input):
    from MixedPrecisionIMC import Input

    inp = Input.readInputs(DECK_SURR)
    inp["PRECISION"] = prec
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(10 * n_input)
    return inp


def _run_c4(prec, n_input, seed, steps, params=None, train=False, random_walk=False, rw_min=8.0):
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Surrogate

    inp = _c4(prec, n_input)
    mesh, simvars, particles = M.setup(inp, seed=seed)
    info = None
    if random_walk:
        particles.random_walk_enable(True, r_min=rw_min)
    elif train:
        info = Surrogate.train(particles, seed=seed)
        params = particles.surrogate_params()
    elif params is not None:
        Surrogate.load(particles, params)
    particles.reset_timing()
    for _ in range(steps):
        M.step_device(inp, mesh, simvars, particles, "CONSTANT")
    tm = particles.timing()
    particles.close()
    hot_s = (tm["ms_source"] + tm["ms_transport"] + tm["ms_clean"] + tm["ms_tally"]) / 1e3
    return mesh.temp.copy(), hot_s, tm, params, info


def surrogate_leg(steps=20, n_input=400000, seeds=12, ref_seeds=4, params=None, save_params=None):
    """Thick/thin blocks workload: FP32 plain vs FP32 + learned surrogate (and vs the analytic
    random walk, the non-learned filling of the same slot, r_min 8) at equal tally error
    against an FP64 GPU oracle (plain FP64, 4x the particles, ref_seeds seeds); `seeds` runs per
    method (each run's error is one draw of a noisy estimate: 12 seeds cut the variance of the
    measured FOM ratio to a third of 4 seeds').
    eps = relative RMS over cells of the final material temperature; FOM = 1/(eps^2 t).
    params: trained network parameters to use instead of training one (a .npy path or an array);
    save_params: a .npy path for the parameters the leg trained."""
    import numpy as np

    if isinstance(params, str):
        params = np.load(params)

    ref = np.mean([_run_c4(64, 4 * n_input, 9000 + s, steps)[0] for s in range(ref_seeds)], axis=0)
    out = {}
    train_info = None
    for label in ("plain", "surrogate", "random_walk"):
        eps2, t, hist, segs, nsurr = [], 0.0, 0, 0, 0
        for s in range(seeds):
            seed = {"plain": 0, "surrogate": 100, "random_walk": 200}[label] + s
            temp, hot, tm, params_out, info = _run_c4(32, n_input, seed, steps,
                                                      params=params if label == "surrogate" else None,
                                                      train=(label == "surrogate" and params is None),
                                                      random_walk=(label == "random_walk"))
            if label == "surrogate" and params is None:
                params, train_info = params_out, info
            eps2.append(float(np.mean(((temp - ref) / ref) ** 2)))
            t += hot
            hist += tm["histories"]
            segs += tm["segments"]
            nsurr += tm["surrogate_events"]
        e2 = float(np.mean(eps2))
        out[label] = {"eps_rms": e2 ** 0.5, "gpu_s_per_run": t / seeds, "histories_per_s": hist / t,
                      "segments_per_run": segs / seeds, "surrogate_events_per_run": nsurr / seeds,
                      "fom": 1.0 / (e2 * t / seeds)}
    out["fom_speedup"] = out["surrogate"]["fom"] / out["plain"]["fom"]
    out["time_speedup"] = out["plain"]["gpu_s_per_run"] / out["surrogate"]["gpu_s_per_run"]
    out["random_walk_fom_speedup"] = out["random_walk"]["fom"] / out["plain"]["fom"]
    out["workload"] = (f"thick/thin blocks 1D (decks/c4_thickthin_fp32.txt), {steps} steps, NINPUT {n_input}, "
                       f"{seeds} seeds each; oracle: FP64 plain, 4x particles, {ref_seeds} seeds")
    if train_info:
        out["training"] = {k: train_info[k] for k in ("samples", "events", "generate_s", "train_s")}
        out["training"]["final_loss"] = train_info["loss_per_epoch"][-1]
        if save_params:
            np.save(save_params, params)
    return out


DECK_CHANNEL = os.path.join(ROOT, "decks", "c3_channel_fp16.txt")
DECK_LARGE2D = os.path.join(ROOT, "decks", "c5_blocks2d_fp16.txt")


def large_2d_leg(warmup: int = 4, steps: int = 8, walk: bool = True):
    """The large 2D workload (1024 x 1024 cells, 64 x 64 random material blocks, NINPUT 2e7,
    NMAX 2e8; decks/c5_blocks2d_fp16.txt) on a window of its first steps.  Its deck asks for
    FP16 tracking, which the engine refuses (the cold T^-3 blocks have 1/sigma_c below binary16's
    range); the rate is measured with FP32 tracking and compensated FP32 tallies, each step one
    device-resident imc_step."""
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Input
    from MixedPrecisionIMC._hip import ImcError

    inp = Input.readInputs(DECK_LARGE2D)
    ts = str(inp["TIMESTEPPING"]).upper()
    out = {"workload": f"large 2D blocks (decks/c5_blocks2d_fp16.txt), steps {warmup}..{warmup + steps - 1}"}
    mesh, simvars, particles = M.setup(inp, precision=16)
    try:
        M.step_device(inp, mesh, simvars, particles, ts)
        out["fp16"] = "ran"
    except ImcError as exc:
        out["fp16"] = f"refused: {exc}"
    particles.close()
    mesh, simvars, particles = M.setup(inp, precision=32)
    for _ in range(warmup):
        M.step_device(inp, mesh, simvars, particles, ts)
    particles.synchronize()
    particles.reset_timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        M.step_device(inp, mesh, simvars, particles, ts)
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    out["fp32"] = {"histories_per_s": tm["histories"] / wall, "hot_path_histories_per_s": tm["histories"] / _hot_s(tm),
                   "segments_per_s_transport_kernel": tm["segments"] / (tm["ms_transport_kernel"] / 1e3),
                   "ms_per_step": 1e3 * wall / steps, "histories_per_step": tm["histories"] / steps,
                   "tally_mode": TALLY_MODES.get(tm["transport_tally_mode"]),
                   "max_abs_energy_error": max(abs(e) for e in simvars.energy_errors)}
    particles.close()
    if not walk:
        return out
    # the same window with the analytic random walk (at least 8 mean free paths from every face)
    mesh, simvars, particles = M.setup(inp, precision=32)
    particles.random_walk_enable(True, r_min=8.0)
    for _ in range(warmup):
        M.step_device(inp, mesh, simvars, particles, ts)
    particles.synchronize()
    particles.reset_timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        M.step_device(inp, mesh, simvars, particles, ts)
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    out["fp32_random_walk"] = {"histories_per_s": tm["histories"] / wall,
                               "hot_path_histories_per_s": tm["histories"] / _hot_s(tm),
                               "ms_per_step": 1e3 * wall / steps, "segments": tm["segments"],
                               "walk_events": tm["surrogate_events"],
                               "ms_transport_kernel": tm["ms_transport_kernel"],
                               "max_abs_energy_error": max(abs(e) for e in simvars.energy_errors)}
    particles.close()
    out["fom"] = large_2d_fom(inp)
    return out


def large_2d_fom(inp, steps: int = 8, seeds: int = 3, r_min: float = 8.0):
    """Random walk against plain transport on the large 2D workload at equal error: after `steps`
    steps, eps = relative RMS of the material temperature over the heated cells (rise over T_INIT
    above 5% of the largest) against an FP64 plain reference (2x the particles, `seeds` seeds);
    FOM = 1 / (eps^2 t), t = the run's GPU hot-path time (sourcing, transport, clean, census)."""
    import numpy as np
    import MixedPrecisionIMC as M

    ts = str(inp["TIMESTEPPING"]).upper()
    n_input = int(float(inp["NINPUT"]))

    def run(prec, seed, walk, scale=1):
        d = dict(inp)
        d["NINPUT"] = str(scale * n_input)
        d["NMAX"] = str(10 * scale * n_input)
        mesh, simvars, particles = M.setup(d, precision=prec, seed=seed, capacity=3 * scale * n_input)
        if walk:
            particles.random_walk_enable(True, r_min=r_min)
        for _ in range(steps):
            M.step_device(d, mesh, simvars, particles, ts)
        tm = particles.timing()
        particles.close()
        return np.asarray(mesh.temp, dtype=np.float64).ravel(), _hot_s(tm)

    from MixedPrecisionIMC import Input
    t_init = float(Input.get_scalar(inp, "T_INIT"))
    ref = np.mean([run(64, 5000 + s, False, 2)[0] for s in range(seeds)], axis=0)
    hot = (ref - t_init) > 0.05 * (ref - t_init).max()
    out = {"workload": f"steps 0..{steps - 1}, FP32, {seeds} seeds each; reference FP64 plain, 2x particles, "
                       f"{seeds} seeds; eps over the {int(hot.sum())} heated cells"}
    for label, walk in (("plain", False), ("random_walk", True)):
        eps2, t = [], 0.0
        for s in range(seeds):
            temp, hs = run(32, (6000 if walk else 7000) + s, walk)
            eps2.append(float(np.mean(((temp[hot] - ref[hot]) / ref[hot]) ** 2)))
            t += hs
        e2 = float(np.mean(eps2))
        out[label] = {"eps_rms": e2 ** 0.5, "gpu_s_per_run": t / seeds, "fom": 1.0 / (e2 * t / seeds)}
    out["random_walk_fom_speedup"] = out["random_walk"]["fom"] / out["plain"]["fom"]
    return out


def precision_study(steps=50, seeds=4, n_input=None):
    """2D channel workload (280x80 graded mesh): FP16 / FP32 / FP64 tracking against an FP64
    reference with 4x the particles.  eps = relative RMS over cells of the material
    temperature after `steps` steps (mean over seeds), with the reference's own standard
    error reported beside it."""
    import numpy as np
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Input

    def run(prec, scale, seed):
        inp = Input.readInputs(DECK_CHANNEL)
        n_in = int(float(Input.get_scalar(inp, "NINPUT"))) if n_input is None else n_input
        inp["NINPUT"] = str(int(n_in * scale))
        inp["NMAX"] = str(int(float(Input.get_scalar(inp, "NMAX")) * scale))
        mesh, sv, parts = M.setup(inp, precision=prec, seed=seed)
        err = None
        done = 0
        for _ in range(steps):
            try:
                M.step(inp, mesh, sv, parts, "CONSTANT")
            except Exception as exc:       # a tracking precision that breaks the run is a result
                err = f"step {done}: {exc}"
                break
            done += 1
        tm = parts.timing()
        parts.close()
        errs = [abs(e) for e in sv.energy_errors] or [float("nan")]
        return mesh.temp.ravel(order="F").copy(), tm, max(errs), err, done

    ref_runs = [run(64, 4, 7000 + s) for s in range(seeds)]
    assert all(r[3] is None for r in ref_runs), [r[3] for r in ref_runs]
    ref = np.mean([r[0] for r in ref_runs], axis=0)
    ref_se = np.std([r[0] for r in ref_runs], axis=0, ddof=1) / np.sqrt(seeds)
    out = {"workload": f"2D channel (decks/c3_channel_fp16.txt), {steps} steps, {seeds} seeds per mode; "
                       f"reference: FP64 with 4x NINPUT/NMAX, {seeds} seeds",
           "reference_rel_se_rms": float(np.sqrt(np.mean((ref_se / ref) ** 2)))}
    for prec in (16, 32, 64):
        rs = [run(prec, 1, 100 * prec + s) for s in range(seeds)]
        ok = [r for r in rs if r[3] is None]
        eps = [float(np.sqrt(np.mean(((r[0] - ref) / ref) ** 2))) for r in ok]
        hot = sum((r[1]["ms_source"] + r[1]["ms_transport"] + r[1]["ms_clean"] + r[1]["ms_tally"]) / 1e3 for r in rs)
        out[f"fp{prec}"] = {"eps_rms": float(np.mean(eps)) if eps else None, "eps_per_seed": eps,
                            "histories_per_s": sum(r[1]["histories"] for r in rs) / hot if hot > 0 else None,
                            "max_energy_error": max(r[2] for r in rs),
                            "failed_runs": [r[3] for r in rs if r[3] is not None],
                            "steps_done": [r[4] for r in rs]}
        print(f"[bench] precision study fp{prec}: eps {out[f'fp{prec}']['eps_rms']}, failures "
              f"{len(rs) - len(ok)}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--deck", default=DECK)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-host-update", action="store_true", help="skip the host-update comparison loop")
    ap.add_argument("--no-large-2d", action="store_true", help="skip the large 2D (1024 x 1024) leg")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="run only the surrogate leg (thick/thin blocks) and print it as the JSON line")
    ap.add_argument("--surrogate-params", default=None, help="use these trained parameters (.npy), no training")
    ap.add_argument("--save-surrogate-params", default=None, help="save the parameters the surrogate leg trains (.npy)")
    ap.add_argument("--large-2d-only", action="store_true",
                    help="run only the large 2D leg's plain FP32 window and print it as the JSON line")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="transport grid bound (0: automatic)")
    ap.add_argument("--precision-study", action="store_true",
                    help="also run the FP16/FP32/FP64 tally-error study on the 2D channel deck")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs on a "
              f"single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.surrogate_only:
        print(json.dumps({"surrogate": surrogate_leg(params=args.surrogate_params,
                                                     save_params=args.save_surrogate_params)}), flush=True)
        return
    if args.large_2d_only:
        print(json.dumps({"large_2d": large_2d_leg(walk=False)}), flush=True)
        return

    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Input

    inputs = Input.readInputs(args.deck)
    mesh, simvars, particles = M.setup(inputs, tuning={"blocks_per_cu": args.blocks_per_cu})
    timestepping = Input.get_scalar(inputs, "TIMESTEPPING").upper()
    for _ in range(args.warmup):
        M.step_device(inputs, mesh, simvars, particles, timestepping)
    particles.synchronize()
    particles.reset_timing()
    t0 = time.perf_counter()
    done = 0
    snap = None
    for _ in range(args.steps):
        if simvars.t > simvars.t_end:
            break
        M.step_device(inputs, mesh, simvars, particles, timestepping)
        done += 1
        if done == args.cpu_steps and not args.no_cpu:
            t_snap = particles.timing()      # GPU figures on the CPU baseline's sample
            snap = (t_snap["histories"], _hot_s(t_snap))
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    hist = tm["histories"]
    sbytes = particles.state_bytes()
    kern_s = tm["ms_transport_kernel"] / 1e3
    launches = max(tm["n_transport_launches"], 1)
    alg_bytes = sbytes * (hist + tm["census_out"])
    achieved = alg_bytes / kern_s / 1e9 if kern_s > 0 else 0.0
    traffic, traffic_src = profiled_traffic()
    valu_busy, valu_src = profiled_valu_busy()
    line = {
        "metric": "histories/sec at equal tally error vs FP64 oracle; surrogate vs plain speedup",
        "value": hist / wall,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Marshak wave 1D, 200 cells, sigma_a=10 T^-3, ramped dt, NINPUT 2e6, "
                               "NMAX 1e7, FP32 tracking + compensated FP32 tallies",
                   "deck": os.path.relpath(args.deck, ROOT), "precision": "FLOAT32",
                   "histories": hist, "segments": tm["segments"], "census_out": tm["census_out"],
                   "state_bytes": sbytes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "valu_busy": valu_busy, "valu_busy_source": valu_src,
                     "kernel": f"k_transport<32,{TALLY_MODES.get(tm['transport_tally_mode'])},1D>",
                     "grid_blocks": tm["transport_grid"],
                     "avg_launch_ms": tm["ms_transport_kernel"] / launches,
                     "bytes_per_launch": alg_bytes / launches,
                     "segments_per_s": tm["segments"] / kern_s if kern_s > 0 else 0.0},
        "gpu_time_ms": {k: tm[k] for k in ("ms_source", "ms_transport", "ms_clean", "ms_tally",
                                           "ms_transport_kernel")},
        "hot_path_histories_per_s": hist / _hot_s(tm),
    }
    print(f"[bench] {done} steps, {hist} histories, {tm['segments']} segments, wall {wall:.3f}s, "
          f"transport kernel {tm['ms_transport_kernel']:.1f} ms over {launches} launches", file=sys.stderr, flush=True)
    particles.close()
    if not args.no_host_update:
        line["host_update_step"] = host_update_step(inputs, args.warmup, done)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inputs, mesh, args.warmup, args.cpu_steps, snap)
        print(f"[bench] cpu baseline {line['cpu_baseline']['value']:.4g} histories/s", file=sys.stderr, flush=True)
        line["fp64"] = fp64_leg(args.warmup, args.cpu_steps)
        print(f"[bench] fp64 leg done", file=sys.stderr, flush=True)
    if not args.no_surrogate:
        line["surrogate"] = surrogate_leg()
        print(f"[bench] surrogate leg: FOM speedup {line['surrogate']['fom_speedup']:.2f}", file=sys.stderr,
              flush=True)
    if not args.no_large_2d:
        line["large_2d"] = large_2d_leg()
        print(f"[bench] large 2D leg: {line['large_2d']['fp32']['histories_per_s']:.4g} histories/s at FP32, "
              f"random walk FOM speedup {line['large_2d']['fom']['random_walk_fom_speedup']:.3g}",
              file=sys.stderr, flush=True)
    if args.precision_study:
        line["precision_study"] = precision_study()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
