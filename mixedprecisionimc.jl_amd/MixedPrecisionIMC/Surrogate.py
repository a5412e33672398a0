"""Learned thick-cell surrogate -- fills the RANDOMWALK slot (Transport.MC_RW,
src/imc_transport.jl:212-479) of the reference.

A 6 -> 64 -> 64 -> 4 network, trained on the GPU by libimc_hip's own BF16 matrix-core
kernels from plain single-cell histories that the GPU also generates, replaces the random
walk of a particle inside a 1D cell whose collision optical thickness sigma_c*dx lies in
the training box.  Outside the box the plain event loop runs.  The method is judged by the
plain method's tallies (tests/test_surrogate_gpu.py).
"""
from __future__ import annotations

import time

import numpy as np

DEFAULT_BOX = {"tau_cell": (2.5, 200.0), "tau_b": (2.0, 1.0e4), "kappa": (1.0e-5, 4.0), "depth_min": 0.7}


def train(particles, *, n_samples: int = 40_000_000, epochs: int = 60, batch: int = 8192, lr: float = 1.5e-3,
          seed: int = 0, box: dict | None = None) -> dict:
    """Generates the training set on the GPU, trains the network and enables it in transport."""
    b = dict(DEFAULT_BOX)
    if box:
        b.update(box)
    particles.surrogate_configure(b["tau_cell"], b["tau_b"], b["kappa"], b["depth_min"])
    t0 = time.perf_counter()
    events = particles.surrogate_generate(n_samples, seed=seed)
    t1 = time.perf_counter()
    particles.surrogate_init(seed)
    losses = particles.surrogate_train(epochs, batch=batch, lr=lr)
    t2 = time.perf_counter()
    particles.surrogate_enable(True)
    return {"samples": n_samples, "events": events, "generate_s": t1 - t0, "train_s": t2 - t1,
            "loss_per_epoch": np.asarray(losses).tolist(), "box": b}


def load(particles, params, box: dict | None = None) -> None:
    """Installs previously trained parameters (4868 floats) and enables the surrogate."""
    b = dict(DEFAULT_BOX)
    if box:
        b.update(box)
    particles.surrogate_configure(b["tau_cell"], b["tau_b"], b["kappa"], b["depth_min"])
    particles.surrogate_set_params(params)
    particles.surrogate_enable(True)
