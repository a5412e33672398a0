"""Pins of the CPU oracle (oracle/) against what fixes each function independently of
its own formulas: published known-answer vectors, numpy's IEEE conversions, exact rational
arithmetic, closed-form transport solutions and conservation laws.  CPU only.

The reference's own tests hold no numeric vectors for this path (test/runtests.jl pins
parsing, types, update non-zero results and clean), so those are restated in test_host.py
and test_clean_reference_semantics below.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy import stats

import oracle as O
from cases import case1d, case2d, oracle_problem


# --- Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32_10) -------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert O.philox(ctr, key) == expect


# --- binary16 emulation vs numpy's IEEE float32 -> float16 conversion ------------------
def test_round_half_matches_numpy():
    rng = np.random.default_rng(3)
    vals = np.concatenate([
        rng.uniform(0, 0.05, 20000),
        rng.uniform(-1, 1, 20000) * 10.0 ** rng.uniform(-9, 5, 20000),
        [(1024 + m + 0.5) * 2.0 ** -18 for m in range(0, 1024, 3)],   # exact ties
        [(m + 0.5) * 2.0 ** -24 for m in range(0, 1024, 7)],          # subnormal ties
        [65504.0, 65519.0, 65520.0, 1e6, 2.0 ** -25, 2.0 ** -26, 0.0, -0.0],
    ]).astype(np.float32)
    with np.errstate(over="ignore"):
        expect = vals.astype(np.float16).astype(np.float64)
    got = np.array([O.round_half(float(v)) for v in vals])
    np.testing.assert_array_equal(got, expect)


# --- sorter (imc_utilities.jl:23-54) against exact rational products ---------------------
def test_sorter_fp64_matches_exact_product():
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = int(rng.integers(2, 12))
        vals = 10.0 ** rng.uniform(-6, 6, n) * rng.choice([1.0, 1.0, 1.0, -1.0], n)
        prod, scale, k = O.sorter(64, vals, [1.0])
        exact = Fraction(1)
        for v in vals:
            exact *= Fraction(float(v))
        assert k == 1 and scale == 1.0
        rel = abs(Fraction(prod) - exact) / abs(exact)
        assert float(rel) <= (n + 1) * 2.0 ** -53


def test_sorter_fp32_matches_exact_product_and_falls_back_on_overflow():
    # FP32: a product that overflows float with the first scale must use the next one
    vals = [3e20, 2e20, 1.0]
    prod, scale, k = O.sorter(32, vals, [1024.0, 1.0 / 1024.0])
    assert k == 2 and scale == 1.0 / 1024.0
    exact = 3e20 * 2e20 / 1024.0
    assert prod == pytest.approx(exact, rel=4 * 2.0 ** -24)
    # no scale works -> (0, 0, 0) as the reference's fallback
    prod, scale, k = O.sorter(32, [1e30, 1e30, 1e30], [1.0])
    assert (prod, scale, k) == (0.0, 0.0, 0)


def test_sorter_pairing_beats_sequential_in_fp32():
    # the reorder keeps 1e-30 * 1e30 products representable where left-to-right would underflow
    vals = [1e-30, 1e-30, 1e30, 1e30]
    prod, _, k = O.sorter(32, vals, [1.0])
    assert k == 1 and prod == pytest.approx(1.0, rel=1e-6)


# --- clean (imc_clean.jl:6-19; reference test test/runtests.jl:78-86) -------------------
def test_clean_reference_semantics():
    op = O.OracleProblem(geometry=1, prec=64, nx=4)
    rec = np.array([[1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, -1.0, 1.0]])
    assert op.clean(rec).shape[0] == 0            # the reference test: one flagged record removed
    recs = np.arange(45, dtype=float).reshape(5, 9)
    recs[[1, 3], 7] = -1.0
    out = op.clean(recs)
    np.testing.assert_array_equal(out, recs[[0, 2, 4]])   # order preserved


# --- sourcing closed forms and distributions ---------------------------------------------
def test_source_energy_and_counts_closed_form():
    case = case1d(prec=64, C=30)
    op, st = oracle_problem(case)
    n_input, cellmin = 20000, 3.0
    rec, emitted, etot, nb = op.source(st, 0, None, n_input, 10 * n_input, cellmin)
    C, dt = case["nx"], case["dt"]
    T, f, sa, dx = case["temp"], case["fleck"], case["sa"], case["dx"]
    a, c = case["phys_a"], case["phys_c"]
    e_body = f * sa * a * c * T ** 4 * dx * dt
    e_src = case["radsource"] * dx * dt
    e_face = a * c * np.array([case["tsurf"][0][0], case["tsurf"][1][0]]) ** 4 * dt / 4
    assert etot == pytest.approx(e_body.sum() + e_src.sum() + e_face.sum(), rel=1e-13)
    np.testing.assert_allclose(emitted, f * sa * a * c * T ** 4 * dt, rtol=1e-14)
    cells = rec[:, 2].astype(int) - 1
    n_face_l = int(np.rint(e_face[0] * n_input / etot))
    n_face_r = int(np.rint(e_face[1] * n_input / etot))
    n_body = np.maximum(np.rint(e_body * n_input / etot), cellmin).astype(int)
    n_src = np.where(e_src > 0, np.maximum(np.rint(e_src * n_input / etot), cellmin), 0).astype(int)
    assert nb == n_face_l + n_face_r + n_body.sum() + n_src.sum() == len(rec)
    # face births sit at 0.01 dx (left) / 0.99 dx (right) of the end cells (imc_sourcing.jl:167,182)
    s_ = case.get("s", 1.0)
    fl = (cells == 0) & (rec[:, 3] == 0.01 * dx[0] * s_)
    fr = (cells == C - 1) & (rec[:, 3] == 0.99 * dx[-1] * s_)
    assert fl.sum() == n_face_l and fr.sum() == n_face_r
    np.testing.assert_allclose([rec[fl, 6].sum(), rec[fr, 6].sum()], e_face, rtol=1e-12)
    vol = ~(fl | fr)
    # volume births: count and energy per cell are the body plus radiation-source closed forms
    np.testing.assert_array_equal(np.bincount(cells[vol], minlength=C), n_body + n_src)
    np.testing.assert_allclose(np.bincount(cells[vol], weights=rec[vol, 6], minlength=C), e_body + e_src,
                               rtol=1e-12)
    # births are stored in a stable order by start-time bucket floor(16 t/dt)
    bucket = np.clip(np.floor(rec[:, 1] / dt * 16.0), 0, 15)
    assert np.all(np.diff(bucket) >= 0)


def test_source_sampling_distributions():
    case = case1d(prec=64, C=4, tsurf=(1.0, 1.0))
    case["phys_c"] = 1.0
    op, st = oracle_problem(case)
    rec, _, _, nb = op.source(st, 1, None, 200000, 2000000, 1.0)
    left = rec[rec[:, 0] == 1.0]
    left = left[(left[:, 3] < 0.02 * case["dx"][0]) & (left[:, 4] > 0)]          # surface births at 0.01 dx
    # Lambertian surface directions mu = sqrt(xi): CDF mu^2
    assert stats.kstest(left[:, 4], lambda m: m ** 2).pvalue > 1e-3
    body = rec[(rec[:, 3] > 0.02 * case["dx"][0]) & (rec[:, 3] < 0.98 * case["dx"][0]) & (rec[:, 2] == 1)]
    assert stats.kstest(body[:, 4], stats.uniform(loc=-1, scale=2).cdf).pvalue > 1e-3
    assert stats.kstest(body[:, 1] / case["dt"], "uniform").pvalue > 1e-3


# --- transport closed forms -----------------------------------------------------------------
def _one_cell(prec=64, L=1.0, sa=0.7, bc=(0, 0, 0, 0), dt=0.5, c=1.0):
    case = dict(geometry=1, prec=prec, nx=1, ny=1, seed=1, bc=bc, scales=(1.0,), dt=dt, phys_c=c, phys_a=1.0,
                dx=np.array([L]), dy=None, sa=np.array([sa]), ss=np.zeros(1), temp=np.ones(1), fleck=np.ones(1),
                radsource=np.zeros(1), tsurf=[np.zeros(1), np.zeros(1)])
    return oracle_problem(case)


def test_transport_pure_absorber_closed_form():
    """f = 1 and sigma_s = 0: no collisions; deposit = E(1 - exp(-sa d)) with
    d = min((L - x)/mu, c (dt - t)), or all of E below the 1% cut-off."""
    L, sa, dt = 1.0, 0.7, 0.5
    op, st = _one_cell(L=L, sa=sa, dt=dt)
    rng = np.random.default_rng(9)
    n = 400
    x = rng.uniform(0, L, n); mu = rng.uniform(0.05, 1, n); t = rng.uniform(0, dt, n); E = rng.uniform(0.5, 2, n)
    rec = np.stack([np.ones(n), t, np.ones(n), x, mu, np.ones(n), E, E, np.ones(n)], axis=1)
    out, dep, lost, segs = op.transport(st, 0, rec)
    d = np.minimum((L - x) / mu, 1.0 * (dt - t))
    w = np.exp(-sa * d)
    dead = w <= 0.01
    dep_expect = np.where(dead, E, E * (1 - w)).sum() / L
    assert dep[0] == pytest.approx(dep_expect, rel=1e-12)
    escaped = (~dead) & ((L - x) / mu < dt - t)
    assert lost == pytest.approx((E * w)[escaped].sum(), rel=1e-12)
    census = (~dead) & ~escaped
    np.testing.assert_allclose(out[census, 6], (E * w)[census], rtol=1e-12)
    np.testing.assert_allclose(out[census, 3], (x + mu * (dt - t))[census], rtol=1e-12)
    assert segs == n


def test_transport_streaming_census_position():
    """sigma = 0: no deposit; a particle streams c(dt - t) across cells of a reflecting box."""
    C = 10
    case = dict(geometry=1, prec=64, nx=C, ny=1, seed=1, bc=(1, 1, 0, 0), scales=(1.0,), dt=0.37, phys_c=2.0,
                phys_a=1.0, dx=np.full(C, 0.1), dy=None, sa=np.zeros(C), ss=np.zeros(C), temp=np.ones(C),
                fleck=np.full(C, 0.5), radsource=np.zeros(C), tsurf=[np.zeros(1), np.zeros(1)])
    op, st = oracle_problem(case)
    rng = np.random.default_rng(2)
    n = 300
    cell = rng.integers(0, C, n); xl = rng.uniform(0, 0.1, n); mu = rng.uniform(-1, 1, n); t = np.zeros(n)
    E = np.ones(n)
    rec = np.stack([np.ones(n), t, cell + 1.0, xl, mu, np.ones(n), E, E, np.ones(n)], axis=1)
    out, dep, lost, segs = op.transport(st, 0, rec)
    assert np.all(dep == 0) and lost == 0
    # unfolded position in a reflecting box of length 1
    X = cell * 0.1 + xl + mu * 2.0 * 0.37
    Xf = np.mod(X, 2.0)
    Xf = np.where(Xf > 1.0, 2.0 - Xf, Xf)
    got = (out[:, 2] - 1) * 0.1 + out[:, 3]
    np.testing.assert_allclose(got, Xf, atol=1e-12)
    np.testing.assert_array_equal(out[:, 6], 1.0)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("geom", [1, 2])
def test_transport_conserves_energy(prec, geom):
    """Energy in = deposited * volume + lost + census energy (each history separately)."""
    case = case1d(prec=prec, bc=(0, 1, 0, 0)) if geom == 1 else case2d(prec=prec)
    op, st = oracle_problem(case)
    rec, _, _, _ = op.source(st, 0, None, 5000, 50000, 1.0)
    e_in = (rec[:, 6] if geom == 1 else rec[:, 7]).sum()
    out, dep, lost, segs = op.transport(st, 0, rec)
    alive = out[:, 7] != -1.0
    e_census = (out[alive, 6] if geom == 1 else out[alive, 7]).sum()
    vol = case["dx"] if geom == 1 else np.outer(case["dx"], case["dy"]).ravel(order="F")
    e_dep = (dep * vol).sum()
    rel = 1e-12 if prec == 64 else 2e-6
    assert e_dep + lost + e_census == pytest.approx(e_in, rel=rel)
    assert segs >= rec.shape[0]


def test_transport_reflecting_box_loses_nothing():
    case = case2d(prec=64, bc=(1, 1, 1, 1))
    op, st = oracle_problem(case)
    rec, _, _, _ = op.source(st, 0, None, 3000, 30000, 1.0)
    _, _, lost, _ = op.transport(st, 0, rec)
    assert lost == 0.0


def test_tree_sum_is_a_sum():
    """etot equals the exact sum of its terms to FP64 rounding (the tree order only)."""
    case = case1d(prec=64, C=700)
    op, st = oracle_problem(case)
    _, _, etot, _ = op.source(st, 0, None, 1000, 10000, 1.0)
    T, f, sa, dx, dt = case["temp"], case["fleck"], case["sa"], case["dx"], case["dt"]
    terms = list(f * sa * T ** 4 * dx * dt) + list(case["radsource"] * dx * dt) + \
        [case["tsurf"][0][0] ** 4 * dt / 4, case["tsurf"][1][0] ** 4 * dt / 4]
    exact = float(sum(Fraction(x) for x in terms))
    assert etot == pytest.approx(exact, rel=1e-13)


# Seed whose Philox block (slot 0, event 0, TAG_EVENT, 0) keyed (seed, step 0) has word 0 =
# 0xFFFFFF23: its 24-bit uniform is exactly 1, so the exponential variate -log(u) is 0.
ZERO_VARIATE_SEED = 56716214


def transparent_case(prec):
    """Three-cell reflecting box whose 1/sigma_c does not fit the tracking type: sigma_a = 1e-7
    (1/sigma_c = 2e7 > 65504) at FP16, sigma_a = 0 at FP32."""
    sa = 1e-7 if prec == 16 else 0.0
    return dict(geometry=1, prec=prec, nx=3, ny=1, seed=ZERO_VARIATE_SEED, bc=(1, 1, 0, 0), scales=(1.0,), dt=1.0,
                phys_c=1.0, phys_a=1.0, dx=np.full(3, 0.1), dy=None, sa=np.full(3, sa), ss=np.zeros(3),
                temp=np.ones(3), fleck=np.full(3, 0.5), radsource=np.zeros(3), tsurf=[np.zeros(1), np.zeros(1)])


def transparent_record():
    return np.array([[1.0, 0.0, 2.0, 0.05, 0.5, 1.0, 1.0, 1.0, 1.0]])


@pytest.mark.parametrize("prec", [32, 16])
def test_zero_variate_in_a_transparent_cell_stays_finite(prec):
    """The collision distance is variate * (1/sigma_c) with the reciprocal clamped to the
    largest finite value of the tracking type; unclamped, 0 * inf = NaN would make the
    history run forever.  Here the first event draws variate 0 in such a cell."""
    assert O.philox([0, 0, 2, 0], [ZERO_VARIATE_SEED, 0])[0] >> 8 == 0xFFFFFF
    case = transparent_case(prec)
    op, st = oracle_problem(case)
    out, dep, lost, segs = op.transport(st, 0, transparent_record())
    assert np.all(np.isfinite(out)) and np.all(np.isfinite(dep)) and np.isfinite(lost)
    assert out[0, 7] != -1.0 and segs < 100          # streams to census (path 1 = 10 cells)
    assert out[0, 6] == pytest.approx(1.0, rel=1e-3)
