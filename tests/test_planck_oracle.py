"""Planck frequency sampling in the oracle (imc_oracle.cpp: orc_planck, orc_planck_terms; the
method of Sourcing.sample_planck, imc_sourcing.jl:372-399), pinned by what the mathematics fixes
for the normalised Planck spectrum p(x) = (15 / pi^4) x^3 / (e^x - 1):

* the moments E[x] = 4 zeta(5) / zeta(4), E[x^2] = 20 zeta(6) / zeta(4), E[1/x] = zeta(3) / (3 zeta(4))
  (Gamma-function integrals of x^k / (e^x - 1));
* the CDF 1 - (15 / pi^4) sum_n e^{-n x} (x^3 / n + 3 x^2 / n^2 + 6 x / n^3 + 6 / n^4)
  (Kolmogorov distance);
* the mixture term n: P(n) = 90 / (pi^4 n^4) for the inversion, and n^-4 / zeta(4, 65) for the
  rejection tail (u forced above the cumulative of the first 64 terms);
* temperature scaling nu = T x and determinism in (seed, stream, i).

The reference's own loop (restated as written in oracle.sample_planck_reference) is checked for
the property the engine's version fixes: its running sum of n^-4, rounded to Float32, stops below
zeta(4), so a first uniform in (90 sum / pi^4, 1) never ends the loop.  No reference test covers
sample_planck (the reference never calls it), so these closed forms are its pins.
"""
import os
import sys

import numpy as np
import pytest
from scipy.special import zeta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def planck_cdf(x):
    x = np.asarray(x, dtype=np.float64)
    n = np.arange(1, 200, dtype=np.float64)[:, None]
    s = np.exp(-n * x) * (x ** 3 / n + 3 * x ** 2 / n ** 2 + 6 * x / n ** 3 + 6 / n ** 4)
    return 1.0 - 15.0 / np.pi ** 4 * s.sum(axis=0)


@pytest.fixture(scope="module")
def draws():
    return O.planck(np.ones(2_000_000), 11, 5)


def test_moments_match_closed_forms(draws):
    x = draws
    for k, exact in ((1, 4 * zeta(5) / zeta(4)), (2, 20 * zeta(6) / zeta(4)), (-1, zeta(3) / (3 * zeta(4)))):
        v = x ** k
        se = v.std() / np.sqrt(v.size)
        assert abs(v.mean() - exact) < 5 * se, (k, v.mean(), exact, se)


def test_cdf_matches_planck_law(draws):
    s = np.sort(draws[:400_000])
    emp = np.arange(1, s.size + 1) / s.size
    assert np.max(np.abs(emp - planck_cdf(s))) < 3e-3


def test_cdf_series_is_a_distribution():
    # 199 terms: the truncated tail is (90 / pi^4) sum_{n >= 200} n^-4 ~ 4e-8 at x -> 0
    assert planck_cdf(np.array([1e-9]))[0] == pytest.approx(0.0, abs=1e-7)
    assert planck_cdf(np.array([60.0]))[0] == pytest.approx(1.0, abs=1e-12)
    # mean from the CDF: integral of 1 - F equals E[x]
    x = np.linspace(0, 60, 600001)
    assert np.trapezoid(1 - planck_cdf(x), x) == pytest.approx(4 * zeta(5) / zeta(4), rel=1e-6)


def test_term_law_of_the_inversion():
    t = O.planck_terms(1_000_000, 3, 9)
    p = 90 / np.pi ** 4 / np.arange(1, 7) ** 4.0
    cnt = np.bincount(t, minlength=8)[1:7]
    chi2 = np.sum((cnt - 1e6 * p) ** 2 / (1e6 * p))
    assert chi2 < 25, (cnt, 1e6 * p)        # 6 bins: P(chi2 > 25) < 5e-4


def test_term_law_of_the_tail():
    t = O.planck_terms(400_000, 3, 9, u_force=1 - 1e-9)
    assert t.min() >= 65
    ns = np.arange(65, 75)
    p = ns ** -4.0 / zeta(4, 65)
    cnt = np.array([(t == n).sum() for n in ns])
    rest = t.size - cnt.sum()
    exp_rest = t.size * (1 - p.sum())
    chi2 = np.sum((cnt - t.size * p) ** 2 / (t.size * p)) + (rest - exp_rest) ** 2 / exp_rest
    assert chi2 < 35, chi2                  # 11 bins: P(chi2 > 35) < 3e-4


def test_inversion_boundaries():
    w = 90 / np.pi ** 4
    assert O.planck_terms(1, 0, 0, u_force=0.0)[0] == 1
    assert O.planck_terms(1, 0, 0, u_force=w * (1 - 1e-12))[0] == 1
    assert O.planck_terms(1, 0, 0, u_force=w * (1 + 1e-12))[0] == 2
    c64 = w * sum(k ** -4.0 for k in range(1, 65))
    assert O.planck_terms(1, 0, 0, u_force=c64 * (1 - 1e-14))[0] == 64
    assert O.planck_terms(1, 0, 0, u_force=min(c64 * (1 + 1e-14), 1 - 1e-16))[0] >= 65


def test_temperature_scaling_and_keying():
    t = np.linspace(0.0, 3.0, 1000)
    a = O.planck(t, 5, 1)
    b = O.planck(np.ones_like(t), 5, 1)
    np.testing.assert_array_equal(a, t * b)
    assert a[0] == 0.0
    np.testing.assert_array_equal(O.planck(t, 5, 1), a)
    assert not np.array_equal(O.planck(t, 5, 2), a)
    assert not np.array_equal(O.planck(t, 6, 1), a)


def test_reference_loop_stops_short_of_one_in_float32():
    lim32 = O.planck_nsum_limit(np.float32)
    lim64 = O.planck_nsum_limit(np.float64)
    assert lim32 < zeta(4)
    bound = 90 * lim32 / np.pi ** 4
    # Julia's rand(Float32) takes values k 2^-24 up to 1 - 2^-24, above the bound
    assert 1 - 2.0 ** -24 > bound and 1 - bound == pytest.approx(1.03e-6, rel=0.02)
    assert 90 * lim64 / np.pi ** 4 < 1 - 2.0 ** -52

    def seq(vals):
        return iter(vals)
    # a first uniform above the bound keeps the reference's loop going through every term
    x, n = O.sample_planck_reference(seq([1 - 2.0 ** -24] + [0.5] * 4), np.float32, max_terms=20000)
    assert x is None and n > 19000
    # below it the loop ends with the Gamma(4, 1/n) draw -log(r1 r2 r3 r4) / n
    x, n = O.sample_planck_reference(seq([0.5, 0.5, 0.5, 0.5, 0.5]), np.float32)
    assert n == 1.0 and x == pytest.approx(-np.log(0.0625), rel=1e-6)


def test_reference_loop_ends_for_every_float16_uniform():
    # rand(Float16) stays at or below 1 - 2^-11, under the bound the Float16 sum reaches
    bound = 90 * O.planck_nsum_limit(np.float16) / np.pi ** 4
    assert 1 - 2.0 ** -11 <= bound
    x, n = O.sample_planck_reference(iter([1 - 2.0 ** -11] + [0.5] * 4), np.float16)
    assert x is not None and n >= 2
