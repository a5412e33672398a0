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


# ---- the plane (imc_kernels.hip: rw2_exit_a, rw2_census, rw2_exit_dir; imc_api.cpp: DiskSeries) ----
# First exit of 2D diffusion from the centre of the unit disk, a = D t / R0^2:
# P(T <= a) = 1 - sum_k 2 e^{-j_k^2 a} / (j_k J1(j_k)); E[a] = 1/4 and E[a^2] = 3/32 (from
# Delta u = -1, Delta v = -2u with u = v = 0 on the rim: u = (1 - r^2)/4, v(0) = 3/32).
from scipy.special import j0, j1, jn_zeros  # noqa: E402

JZ = jn_zeros(0, 400)
CK = 2.0 / (JZ * j1(JZ))
NR = 1024


def disk_cdf(a):
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    s = np.zeros_like(a)
    for j, c in zip(JZ, CK):
        s += c * np.exp(-j * j * a)
    return np.where(a < 3.2e-5, 0.0, np.clip(1.0 - s, 0.0, 1.0))


def disk_tables():
    u = np.arange(N) / N
    lo, hi = np.zeros(N), np.full(N, 40.0)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        below = disk_cdf(mid) < u
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    exit_tab = (0.5 * (lo + hi)).astype(np.float32)
    ur = np.arange(NR) / (NR - 1)
    lo, hi = np.zeros(NR), np.ones(NR)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        below = mid * j1(JZ[0] * mid) / j1(JZ[0]) < ur
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return exit_tab, (0.5 * (lo + hi)).astype(np.float32)


def disk_exit_a(tab, u):
    f = u * N
    i = f.astype(np.int64)
    tail = np.log(1.6019747 / np.maximum(1.0 - u, 2.0 ** -30)) * 0.17291507
    j = np.minimum(i, N - 2)
    w = f - j
    return np.where(i >= N - 1, tail, tab[j] + w * (tab[j + 1] - tab[j]))


def disk_census(a, rad, rng, n):
    if a >= 0.2:
        f = rng.random(n) * (NR - 1)
        i = np.minimum(f.astype(np.int64), NR - 2)
        return rad[i] + (f - i) * (rad[i + 1] - rad[i])
    out = np.zeros(n)
    todo = np.ones(n, dtype=bool)
    sg = np.sqrt(2.0 * a)
    for _ in range(16):
        k = int(todo.sum())
        if k == 0:
            break
        rho = sg * np.sqrt(-2.0 * np.log(1.0 - rng.random(k)))
        inside = rho < 1.0
        acc = inside & (rng.random(k) < 1.0 - np.exp(-(1.0 - rho) / a))
        idx = np.flatnonzero(todo)
        out[idx[inside]] = rho[inside]
        todo[idx[acc]] = False
    return out


def exit_psi(rng, n):
    """rw2_exit_dir: cosine-law proposals sin psi = 2u - 1, accepted with probability
    (1 + (4/pi) cos psi) / (1 + 4/pi), at most 8 redraws."""
    sn = 2.0 * rng.random(n) - 1.0
    v = rng.random(n)
    for _ in range(8):
        cs = np.sqrt(np.maximum(1.0 - sn * sn, 0.0))
        rej = v * (1 + 4 / np.pi) >= 1.0 + (4 / np.pi) * cs
        if not rej.any():
            break
        sn[rej] = 2.0 * rng.random(int(rej.sum())) - 1.0
        v[rej] = rng.random(int(rej.sum()))
    return np.arcsin(sn)


def test_disk_constants_match_the_series():
    assert 0.17291507 == pytest.approx(1.0 / JZ[0] ** 2, rel=1e-7)
    assert 1.6019747 == pytest.approx(CK[0], rel=1e-7)


def test_disk_exit_time_moments_through_the_table():
    tab, _ = disk_tables()
    rng = np.random.default_rng(7)
    a = disk_exit_a(tab, rng.random(4_000_000).astype(np.float32).astype(np.float64))
    se1 = a.std() / np.sqrt(a.size)
    assert abs(a.mean() - 0.25) < 5 * se1 + 2e-4
    se2 = (a * a).std() / np.sqrt(a.size)
    assert abs((a * a).mean() - 3.0 / 32.0) < 5 * se2 + 3e-4


def test_disk_series_moments_exactly():
    # E[a] = sum c_k / j_k^2, E[a^2] = sum 2 c_k / j_k^4 from the CDF's Laplace form
    assert np.sum(CK / JZ ** 2) == pytest.approx(0.25, rel=1e-6)
    assert np.sum(2 * CK / JZ ** 4) == pytest.approx(3.0 / 32.0, rel=1e-9)


def test_disk_tail_matches_the_series():
    for u in (1.0 - 1.0 / N, 1.0 - 1e-5):
        a = np.log(CK[0] / (1.0 - u)) / JZ[0] ** 2
        assert disk_cdf(a)[0] == pytest.approx(u, abs=1e-9)


@pytest.mark.parametrize("a", [0.03, 0.1, 0.19, 0.2, 0.5, 1.5])
def test_disk_census_radius_follows_the_conditional_density(a):
    # heat kernel of the disk from its centre in the normalised eigenfunctions
    # J0(j_k r) / (sqrt(pi) J1(j_k)), times the ring's circumference
    r = np.linspace(0.0, 1.0, 4001)
    dens = r * sum(j0(j * r) / j1(j) ** 2 * np.exp(-j * j * a) for j in JZ[:200])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(r))])
    cdf /= cdf[-1]
    _, rad = disk_tables()
    s = np.sort(disk_census(a, rad, np.random.default_rng(13), 200_000))
    emp = np.arange(1, s.size + 1) / s.size
    assert np.max(np.abs(emp - np.interp(s, r, cdf))) < 0.02


def test_disk_exit_direction_law():
    psi = np.sort(exit_psi(np.random.default_rng(3), 400_000))
    F = 0.25 * (1.0 + np.sin(psi) + (2 / np.pi) * (psi + np.pi / 2) + np.sin(2 * psi) / np.pi)
    emp = np.arange(1, psi.size + 1) / psi.size
    assert np.max(np.abs(emp - F)) < 5e-3
    # E[cos psi] = (pi/2 + 16 / (3 pi)) / 4
    assert np.cos(psi).mean() == pytest.approx((np.pi / 2 + 16 / (3 * np.pi)) / 4, abs=2e-3)
