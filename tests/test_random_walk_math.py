"""The analytic random walk's laws (imc_api.cpp: rw_exit_cdf and its inverse table;
imc_kernels.hip: rw_exit_a, rw_census_y), restated in numpy and checked against what the
mathematics fixes for one-dimensional diffusion on (-1, 1) started at 0, in the dimensionless
time a = D t / R0^2:

* the two series for the exit CDF (images for small a, eigenfunctions for large a) agree
  where both converge;
* the exit time drawn through the inverse table has mean 1/2 and second moment 5/12
  (E[T] = (1 - x^2) / 2 and E[T^2] from u'' = -2 E[T], u(+-1) = 0, at x = 0);
* the table's closed-form tail matches the series;
* the census position drawn for a walker not yet out follows the conditional density
  sum_k cos((2k+1) pi y / 2) exp(-(2k+1)^2 pi^2 a / 4) / S(a) (Kolmogorov distance < 0.02).
"""
import numpy as np
import pytest
from scipy.special import erfc

N = 4096


def cdf_small(a):
    r = 1.0 / (2.0 * np.sqrt(a))
    return sum((-2.0 if n & 1 else 2.0) * erfc((2 * n + 1) * r) for n in range(8))


def cdf_large(a):
    s = sum((-1.0 if k & 1 else 1.0) / (2 * k + 1) * np.exp(-((2 * k + 1) ** 2) * np.pi ** 2 * a / 4.0)
            for k in range(40))
    return 1.0 - 4.0 / np.pi * s


def exit_cdf(a):
    a = np.asarray(a, dtype=np.float64)
    out = np.where(a < 0.25, cdf_small(np.maximum(a, 1e-300)), cdf_large(a))
    return np.clip(np.where(a <= 0, 0.0, out), 0.0, 1.0)


def table():
    u = np.arange(N) / N
    lo, hi = np.zeros(N), np.full(N, 60.0)
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        below = exit_cdf(mid) < u
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return (0.5 * (lo + hi)).astype(np.float32)


def exit_a(tab, u):
    f = u * N
    i = f.astype(np.int64)
    tail = -(4.0 / np.pi ** 2) * np.log(np.pi / 4.0 * np.maximum(1.0 - u, 2.0 ** -30))
    j = np.minimum(i, N - 2)
    w = f - j
    interp = tab[j] + w * (tab[j + 1] - tab[j])
    return np.where(i >= N - 1, tail, interp)


def census_y(a, rng, n):
    if a >= 0.2:
        return (2.0 / np.pi) * np.arcsin(2.0 * rng.random(n) - 1.0)
    y = np.zeros(n)
    todo = np.ones(n, dtype=bool)
    sg = np.sqrt(2.0 * a)
    for _ in range(16):
        k = int(todo.sum())
        if k == 0:
            break
        yc = sg * rng.standard_normal(k)
        inside = np.abs(yc) < 1.0
        acc = inside & (rng.random(k) < 1.0 - np.exp((yc - 1.0) / a) - np.exp(-(yc + 1.0) / a))
        idx = np.flatnonzero(todo)
        y[idx[inside]] = yc[inside]
        todo[idx[acc]] = False
    return y


def test_exit_cdf_series_agree_where_both_converge():
    a = np.linspace(0.12, 0.6, 41)
    np.testing.assert_allclose(cdf_small(a), cdf_large(a), rtol=0, atol=1e-12)


def test_exit_time_moments_through_the_table():
    tab = table()
    rng = np.random.default_rng(5)
    a = exit_a(tab, rng.random(4_000_000).astype(np.float32).astype(np.float64))
    se1 = a.std() / np.sqrt(a.size)
    assert abs(a.mean() - 0.5) < 5 * se1 + 2e-4
    se2 = (a * a).std() / np.sqrt(a.size)
    assert abs((a * a).mean() - 5.0 / 12.0) < 5 * se2 + 5e-4


def test_exit_table_tail_matches_the_series():
    for u in (1.0 - 1.0 / N, 1.0 - 1e-5, 1.0 - 1e-7):
        a = -(4.0 / np.pi ** 2) * np.log(np.pi / 4.0 * (1.0 - u))
        assert exit_cdf(a) == pytest.approx(u, abs=1e-9)


@pytest.mark.parametrize("a", [0.03, 0.1, 0.19, 0.2, 0.5, 1.5])
def test_census_position_follows_the_conditional_density(a):
    y = np.linspace(-1.0, 1.0, 4001)
    dens = sum(np.cos((2 * k + 1) * np.pi * y / 2) * np.exp(-((2 * k + 1) ** 2) * np.pi ** 2 * a / 4)
               for k in range(200))
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(y))])
    cdf /= cdf[-1]
    rng = np.random.default_rng(11)
    s = np.sort(census_y(a, rng, 200_000))
    emp = np.arange(1, s.size + 1) / s.size
    model = np.interp(s, y, cdf)
    assert np.max(np.abs(emp - model)) < 0.02
