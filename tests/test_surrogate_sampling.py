"""The two closed-form samplers of a surrogate exit (imc_kernels.hip: surr_exit_path,
surr_exit_mu), restated in numpy and checked against what the mathematics fixes:

* the exit path law IG(m, phi m) with phi = L^2 / (2 (sd m - L)), L = -ln(1 - d), has mean m
  and E[1 - exp(-sd s)] = d -- checked by integrating scipy's inverse-Gaussian density (an
  independent parameterisation), and the Michael-Schucany-Haas sampler is checked against
  that distribution (moments and a Kolmogorov-Smirnov test);
* the direction cosine from two Newton steps on mu^3 + mu^2 = 2u follows p(mu) = mu (1 + 1.5 mu):
  CDF residual below 1e-4 everywhere, and the sample mean matches 17/24.
"""
import numpy as np
import pytest
from scipy import integrate, stats

CASES = [(2.0, 0.3, 0.4), (1.0, 0.01, 0.0095), (5.0, 1.0, 0.9), (1.0, 2.0, 0.5), (0.2, 0.5, 0.09)]


def _shape(m, sd, d):
    lam = sd * m
    L = -np.log1p(-d)
    return L * L / (2.0 * (lam - L))          # phi


def _sample(m, sd, d, rng, n):
    lam = sd * m
    L = -np.log1p(-d)
    half_cv2 = min((lam - L) / (L * L), 50.0)
    z = rng.standard_normal(n)
    w = z * z * half_cv2
    xr = m / (1.0 + w + np.sqrt(w * (w + 2.0)))
    return np.where(rng.random(n) * (m + xr) <= m, xr, m * m / xr)


@pytest.mark.parametrize("m,sd,d", CASES)
def test_exit_path_law_matches_mean_and_deposit(m, sd, d):
    phi = _shape(m, sd, d)
    assert phi > 0
    lam_shape = phi * m                          # IG(mean m, shape phi m)
    dist = stats.invgauss(mu=m / lam_shape, scale=lam_shape)
    assert dist.mean() == pytest.approx(m, rel=1e-12)
    lt, _ = integrate.quad(lambda s: np.exp(-sd * s) * dist.pdf(s), 0, np.inf, limit=400)
    assert 1.0 - lt == pytest.approx(d, rel=1e-6)


@pytest.mark.parametrize("m,sd,d", CASES)
def test_exit_path_sampler_draws_that_law(m, sd, d):
    rng = np.random.default_rng(7)
    n = 400_000
    s = _sample(m, sd, d, rng, n)
    phi = min(_shape(m, sd, d), 1e300)
    lam_shape = phi * m
    dist = stats.invgauss(mu=m / lam_shape, scale=lam_shape)
    sd_mean = dist.std() / np.sqrt(n)
    assert abs(s.mean() - m) < 5 * sd_mean
    dep = 1.0 - np.exp(-sd * s)
    assert abs(dep.mean() - d) < 5 * dep.std() / np.sqrt(n)
    assert stats.kstest(s[:50_000], dist.cdf).pvalue > 1e-3


def _exit_mu(u):
    u = np.maximum(u, 2.0 ** -24)
    m = np.sqrt(u)
    for _ in range(2):
        m = m - (m * m * (m + 1.0) - 2.0 * u) / (m * (3.0 * m + 2.0))
    return np.minimum(m, 1.0)


def test_exit_direction_inverse_cdf():
    u = np.linspace(0.0, 1.0, 200_001)
    mu = _exit_mu(u)
    cdf = mu * mu * (1.0 + mu) / 2.0
    assert np.abs(cdf[1:] - u[1:]).max() < 1e-4
    assert np.all(np.diff(mu) >= 0)
    rng = np.random.default_rng(3)
    x = _exit_mu(rng.random(1_000_000))
    # E[mu] = int mu^2 (1 + 1.5 mu) = 1/3 + 3/8
    assert x.mean() == pytest.approx(17.0 / 24.0, abs=5 * x.std() / 1000.0)
