"""Pins of the FP64 reference network against central finite differences.  CPU only."""
import numpy as np

import mlp_oracle as R


def test_parameter_count():
    assert R.NPARAM == 4868


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = R.init_params(3)
    X = rng.uniform(-1, 1, (37, 6))
    Y = np.stack([rng.integers(0, 2, 37), rng.integers(0, 2, 37), rng.uniform(0, 1, 37), rng.uniform(0, 1, 37)], 1)
    loss, g = R.loss_and_grad(p, X, Y)
    idx = rng.choice(R.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p); e[i] = h
        fd = (R.loss_and_grad(p + e, X, Y)[0] - R.loss_and_grad(p - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.14159], np.float32)
    got = R.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                     # tie -> even
    assert got[2] == 1.0 + 4 * 2 ** -8       # tie -> even (upwards)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.140625) < 1e-12
