"""Throughput benchmark of the GPU particle engine on the Marshak-wave workload.

Workload (the metric's configuration): nonlinear Marshak wave in 1D, x in [0,0.2] with 200
uniform cells, sigma_a = 10 T^-3, bee = 0.3, left surface T = 1, ramped dt (1e-4 x 1.1 up to
1e-3), NINPUT 2e6, NMAX 1e7; FP32 tracking with compensated FP32 tallies (decks/c2_marshak_fp32.txt).

A step is one IMC timestep: host FP64 update, GPU sourcing, transport, clean and census
tally, host temperature update and energy check.  W warmup steps (the census builds to
~NMAX), then K timed steps.  value = histories transported in the K steps / wall time of
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
HBM_PEAK_GBS = 8000.0


TALLY_MODES = {0: "WAVEROW", 1: "BLOCKROW", 2: "GLOBAL"}


def gpu_window(inputs, steps: int):
    """GPU rate on the same sample as the CPU baseline: the first `steps` timesteps."""
    import MixedPrecisionIMC as M

    mesh, simvars, particles = M.setup(inputs)
    ts = inputs["TIMESTEPPING"].upper() if isinstance(inputs["TIMESTEPPING"], str) else "RAMP"
    particles.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        M.step(inputs, mesh, simvars, particles, ts)
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    particles.close()
    hot = (tm["ms_source"] + tm["ms_transport"] + tm["ms_clean"] + tm["ms_tally"]) / 1e3
    return tm["histories"] / wall, tm["histories"] / hot


def cpu_baseline(inputs, mesh, steps: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here as the CPU baseline only

    threads = min(16, os.cpu_count() or 1)
    r = O.run_deck(inputs, mesh, max_steps=steps, nthreads=threads)
    g_wall, g_hot = gpu_window(inputs, steps)
    cpu = r["histories"] / r["t_hot"]
    return {"value": cpu, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"first {r['steps']} timesteps of the same deck from t=0 ({r['histories']} histories, "
                      f"C++ FP32 oracle, OpenMP over particles), source+transport+clean+census time",
            "gpu_same_sample": {"value_wall": g_wall, "value_hot_path": g_hot,
                                "speedup_hot_path": g_hot / cpu}}


def _c4(prec, n_input):
    from MixedPrecisionIMC import Input

    inp = Input.readInputs(DECK_SURR)
    inp["PRECISION"] = prec
    inp["NINPUT"] = str(n_input)
    inp["NMAX"] = str(10 * n_input)
    return inp


def _run_c4(prec, n_input, seed, steps, params=None, train=False):
    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Surrogate

    inp = _c4(prec, n_input)
    mesh, simvars, particles = M.setup(inp, seed=seed)
    info = None
    if train:
        info = Surrogate.train(particles, seed=seed)
        params = particles.surrogate_params()
    elif params is not None:
        Surrogate.load(particles, params)
    particles.reset_timing()
    for _ in range(steps):
        M.step(inp, mesh, simvars, particles, "CONSTANT")
    tm = particles.timing()
    particles.close()
    hot_s = (tm["ms_source"] + tm["ms_transport"] + tm["ms_clean"] + tm["ms_tally"]) / 1e3
    return mesh.temp.copy(), hot_s, tm, params, info


def surrogate_leg(steps=20, n_input=400000, seeds=4):
    """Thick/thin blocks workload: FP32 plain vs FP32 + learned surrogate at equal tally
    error against an FP64 GPU oracle (plain FP64, 4x the particles, 4 seeds).
    eps = relative RMS over cells of the final material temperature; FOM = 1/(eps^2 t)."""
    import numpy as np

    ref = np.mean([_run_c4(64, 4 * n_input, 9000 + s, steps)[0] for s in range(seeds)], axis=0)
    out = {}
    params = None
    train_info = None
    for label in ("plain", "surrogate"):
        eps2, t, hist, segs, nsurr = [], 0.0, 0, 0, 0
        for s in range(seeds):
            temp, hot, tm, params_out, info = _run_c4(32, n_input, 100 * (label == "surrogate") + s, steps,
                                                      params=params if label == "surrogate" else None,
                                                      train=(label == "surrogate" and params is None))
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
    out["workload"] = (f"thick/thin blocks 1D (decks/c4_thickthin_fp32.txt), {steps} steps, NINPUT {n_input}, "
                       f"{seeds} seeds each; oracle: FP64 plain, 4x particles, {seeds} seeds")
    if train_info:
        out["training"] = {k: train_info[k] for k in ("samples", "events", "generate_s", "train_s")}
        out["training"]["final_loss"] = train_info["loss_per_epoch"][-1]
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs on a "
              f"single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import MixedPrecisionIMC as M
    from MixedPrecisionIMC import Input

    inputs = Input.readInputs(args.deck)
    mesh, simvars, particles = M.setup(inputs)
    timestepping = Input.get_scalar(inputs, "TIMESTEPPING").upper()
    for _ in range(args.warmup):
        M.step(inputs, mesh, simvars, particles, timestepping)
    particles.synchronize()
    particles.reset_timing()
    t0 = time.perf_counter()
    done = 0
    for _ in range(args.steps):
        if simvars.t > simvars.t_end:
            break
        M.step(inputs, mesh, simvars, particles, timestepping)
        done += 1
    particles.synchronize()
    wall = time.perf_counter() - t0
    tm = particles.timing()
    hist = tm["histories"]
    sbytes = particles.state_bytes()
    kern_s = tm["ms_transport_kernel"] / 1e3
    launches = max(tm["n_transport_launches"], 1)
    alg_bytes = sbytes * (hist + tm["census_out"])
    achieved = alg_bytes / kern_s / 1e9 if kern_s > 0 else 0.0
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"k_transport<32,{TALLY_MODES.get(tm['transport_tally_mode'])},1D>",
                     "grid_blocks": tm["transport_grid"],
                     "avg_launch_ms": tm["ms_transport_kernel"] / launches,
                     "bytes_per_launch": alg_bytes / launches},
        "gpu_time_ms": {k: tm[k] for k in ("ms_source", "ms_transport", "ms_clean", "ms_tally",
                                           "ms_transport_kernel")},
    }
    print(f"[bench] {done} steps, {hist} histories, {tm['segments']} segments, wall {wall:.3f}s, "
          f"transport kernel {tm['ms_transport_kernel']:.1f} ms over {launches} launches", file=sys.stderr, flush=True)
    particles.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inputs, mesh, args.cpu_steps)
    if not args.no_surrogate:
        line["surrogate"] = surrogate_leg()
        print(f"[bench] surrogate leg: FOM speedup {line['surrogate']['fom_speedup']:.2f}", file=sys.stderr,
              flush=True)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
