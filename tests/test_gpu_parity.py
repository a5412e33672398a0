"""GPU parity against the CPU oracle, through the C ABI (libimc_hip).

Every case runs the same step on both sides with the same Philox keying, so births are
compared bit for bit and histories are replayed event by event.  Transcendentals differ
(ROCm device library on the GPU, glibc in the oracle), so floating-point fields are compared
with tolerances written in each test:
  FP64 : census fields rtol 1e-9, tallies rtol 1e-9 of the largest bin, alive set identical
  FP32 : alive-set mismatch <= 0.2 % of histories, tallies rtol 2e-3 of the total
  FP16 : alive-set mismatch <= 5 %, tallies within 5 % of the total (FP16 tracking)
"""
import numpy as np
import pytest

from cases import case1d, case2d, engine, engine_step, oracle_problem

pytestmark = pytest.mark.gpu

PRECS = [64, 32, 16]


def _fields(geometry):
    # (label, index) of record fields compared for census particles
    if geometry == 1:
        return [("t", 1), ("x", 3), ("mu", 4), ("E", 6), ("E0", 7)]
    return [("t", 0), ("x", 3), ("y", 4), ("theta", 5), ("E", 7), ("E0", 8)]


def _cells(rec, geometry):
    if geometry == 1:
        return rec[:, 2].astype(np.int64)
    return (rec[:, 1] + 1000 * rec[:, 2]).astype(np.int64)


def _alive(rec):
    return rec[:, 7] != -1.0


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("geom", [1, 2])
def test_source_matches_oracle(prec, geom):
    case = case1d(prec=prec) if geom == 1 else case2d(prec=prec)
    print(f"[source] geometry={geom} prec={prec}", flush=True)
    e = engine(case)
    engine_step(e, case, 3)
    emitted, etot, n = e.source(3000, 30000, 2.0)
    births = e.export_particles()
    op, st = oracle_problem(case)
    rec, o_emit, o_etot, nb = op.source(st, 3, None, 3000, 30000, 2.0)
    assert n == nb, (n, nb)
    assert etot == pytest.approx(o_etot, rel=1e-14)
    np.testing.assert_allclose(emitted, o_emit, rtol=1e-14, atol=0)
    np.testing.assert_array_equal(_cells(births, geom), _cells(rec, geom))
    for name, idx in _fields(geom):
        np.testing.assert_array_equal(births[:, idx], rec[:, idx], err_msg=name)
    e.close()


@pytest.mark.parametrize("geom", [1, 2])
def test_many_births_match_oracle(geom):
    """Enough births (6e5: 586 birth tiles x 16 start-time buckets = 9,376 histogram entries)
    that the birth-order scan spans several scan tiles of 4,096; the 2D mesh of 100 x 90 cells
    has 18,380 count entries, so the count scan does too.  Births equal the oracle's bit for
    bit."""
    case = case1d(prec=32) if geom == 1 else case2d(nx=100, ny=90, prec=32)
    print(f"[many births] geometry={geom}", flush=True)
    e = engine(case, capacity=1_500_000)
    engine_step(e, case, 3)
    emitted, etot, n = e.source(600000, 6000000, 2.0)
    births = e.export_particles()
    op, st = oracle_problem(case)
    rec, o_emit, o_etot, nb = op.source(st, 3, None, 600000, 6000000, 2.0)
    assert n == nb and n > 4096 * 1024 // 16, (n, nb)
    np.testing.assert_array_equal(_cells(births, geom), _cells(rec, geom))
    for name, idx in _fields(geom):
        np.testing.assert_array_equal(births[:, idx], rec[:, idx], err_msg=name)
    e.close()


def _replay(case, n_input=3000, step=2, **tuning):
    e = engine(case, **tuning)
    op, st = oracle_problem(case)
    rec, _, _, _ = op.source(st, step, None, n_input, 10 * n_input, 2.0)
    engine_step(e, case, step)
    e.import_particles(rec)
    dep, lost, segs = e.transport()
    g = e.export_particles()
    o_rec, o_dep, o_lost, o_segs = op.transport(st, step, rec)
    return e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("geom", [1, 2])
def test_transport_replay(prec, geom):
    case = case1d(prec=prec) if geom == 1 else case2d(prec=prec)
    print(f"[transport] geometry={geom} prec={prec}", flush=True)
    e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec) = _replay(case)
    n = rec.shape[0]
    a_g, a_o = _alive(g), _alive(o_rec)
    mismatch = np.count_nonzero(a_g != a_o)
    scale = np.abs(o_dep).max()
    total = np.abs(o_dep).sum()
    if prec == 64:
        assert mismatch == 0
        both = a_g & a_o
        np.testing.assert_array_equal(_cells(g[both], case["geometry"]), _cells(o_rec[both], case["geometry"]))
        for name, idx in _fields(case["geometry"]):
            np.testing.assert_allclose(g[both, idx], o_rec[both, idx], rtol=1e-9, atol=1e-12, err_msg=name)
        np.testing.assert_allclose(dep, o_dep, rtol=1e-9, atol=1e-9 * scale)
        assert lost == pytest.approx(o_lost, rel=1e-9, abs=1e-15)
        assert abs(segs - o_segs) <= max(1, o_segs // 100000)
    elif prec == 32:
        assert mismatch <= max(2, n // 500), (mismatch, n)
        assert np.abs(dep - o_dep).sum() <= 2e-3 * total
        assert lost == pytest.approx(o_lost, rel=2e-3, abs=1e-6 * total)
        assert abs(segs - o_segs) <= max(5, o_segs // 200)
    else:
        assert mismatch <= max(5, n // 20), (mismatch, n)
        assert np.abs(dep.sum() - o_dep.sum()) <= 5e-2 * total
    e.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("geom", [1, 2])
def test_clean_and_census_match_oracle(prec, geom):
    case = case1d(prec=prec) if geom == 1 else case2d(prec=prec)
    print(f"[clean] geometry={geom} prec={prec}", flush=True)
    e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec) = _replay(case)
    # clean the GPU population and compare with the oracle's clean of the GPU's own records:
    # order preserving removal (imc_clean.jl:13-18) must give exactly the same sequence
    n_alive = e.clean()
    after = e.export_particles()
    expect = op.clean(g)
    assert n_alive == expect.shape[0]
    np.testing.assert_array_equal(_cells(after, case["geometry"]), _cells(expect, case["geometry"]))
    for name, idx in _fields(case["geometry"]):
        np.testing.assert_array_equal(after[:, idx], expect[:, idx], err_msg=name)
    rad = e.tally_census()
    o_rad = op.census(st, expect)
    rtol = 1e-12 if prec == 64 else 1e-5
    np.testing.assert_allclose(rad, o_rad, rtol=rtol, atol=rtol * np.abs(o_rad).max())
    e.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_multi_scale_and_reflect(prec):
    """K = 3 energy scales and reflecting walls on both sides (1D)."""
    case = case1d(prec=prec, scales=(1024.0, 8.0, 1.0), bc=(1, 1, 0, 0), C=24, tsurf=(0.8, 0.8))
    e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec) = _replay(case)
    assert lost == 0.0 and o_lost == 0.0
    total = np.abs(o_dep).sum()
    tol = 1e-9 if prec == 64 else 2e-3
    assert np.abs(dep - o_dep).sum() <= tol * total
    e.close()


@pytest.mark.parametrize("prec", [32, 16])
def test_zero_variate_in_a_transparent_cell_gpu(prec):
    """GPU counterpart of test_oracle_pins.test_zero_variate_in_a_transparent_cell_stays_finite:
    a first event with exponential variate 0 in a cell whose 1/sigma_c overflows the tracking
    type ends finite, at census, like the oracle's history."""
    from test_oracle_pins import transparent_case, transparent_record
    case = transparent_case(prec)
    print(f"[zero variate] prec={prec}", flush=True)
    e = engine(case)
    engine_step(e, case, 0)
    rec = transparent_record()
    e.import_particles(rec)
    dep, lost, segs = e.transport()
    g = e.export_particles()
    op, st = oracle_problem(case)
    o_rec, o_dep, o_lost, o_segs = op.transport(st, 0, rec)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(dep)) and np.isfinite(lost)
    assert e.timing()["truncated"] == 0
    assert g[0, 7] != -1.0 and segs == o_segs
    assert g[0, 2] == o_rec[0, 2]
    assert g[0, 6] == pytest.approx(o_rec[0, 6], rel=1e-3)
    e.close()


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
@pytest.mark.parametrize("geom", [1, 2])
def test_transport_replay_every_tally_mode(mode, geom):
    """Per-wave LDS rows (1), one LDS row per block (2), replicated global FP64 atomics (3) and
    the hashed LDS table (4) give the oracle's tallies at FP64; blocks_per_cu=2 also exercises a
    small persistent grid."""
    case = case1d(prec=64) if geom == 1 else case2d(prec=64)
    print(f"[tally mode] mode={mode} geometry={geom}", flush=True)
    e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec) = _replay(case, tally_mode=mode,
                                                                                     blocks_per_cu=2)
    tm = e.timing()["transport_tally_mode"]
    assert tm == {1: 0, 2: 1, 3: 2, 4: 3}[mode]
    np.testing.assert_allclose(dep, o_dep, rtol=1e-9, atol=1e-9 * np.abs(o_dep).max())
    assert lost == pytest.approx(o_lost, rel=1e-9, abs=1e-15)
    assert np.count_nonzero(_alive(g) != _alive(o_rec)) == 0
    e.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_large_mesh_uses_the_hashed_tally(prec):
    """100 x 90 cells: a block row does not fit 64 KiB of LDS at FP64 (9000 bins), so the automatic
    choice is the hashed table (FP32 rows fit, so it is requested there); with more bins than its
    4096 slots, colliding bins take the global path.  Tallies match the oracle (FP64 to 1e-9;
    FP32 to the tolerances of the FP32 replay)."""
    case = case2d(nx=100, ny=90, prec=prec)
    print(f"[hashed tally] prec={prec}", flush=True)
    tuning = {} if prec == 64 else {"tally_mode": 4}
    e, op, st, rec, (dep, lost, segs, g), (o_dep, o_lost, o_segs, o_rec) = _replay(case, n_input=40000, **tuning)
    assert e.timing()["transport_tally_mode"] == 3
    if prec == 64:
        np.testing.assert_allclose(dep, o_dep, rtol=1e-9, atol=1e-9 * np.abs(o_dep).max())
        assert lost == pytest.approx(o_lost, rel=1e-9, abs=1e-15)
        assert np.count_nonzero(_alive(g) != _alive(o_rec)) == 0
    else:
        assert np.abs(dep - o_dep).sum() <= 2e-3 * np.abs(o_dep).sum()
    e.close()


def test_fp16_rejects_a_mean_free_path_it_cannot_represent():
    """sigma_c = 1e9: 1/sigma_c rounds to 0 in binary16, so every collision distance would be 0;
    the step is refused with IMC_ENUMERIC instead.  FP32 represents it and accepts the step."""
    from MixedPrecisionIMC._hip import ImcError
    for prec, ok in ((16, False), (32, True)):
        case = case1d(prec=prec, C=8)
        case["sa"] = case["sa"].copy()
        case["sa"][3] = 1e9
        e = engine(case)
        if ok:
            engine_step(e, case, 0)
        else:
            with pytest.raises(ImcError, match="mean free path"):
                engine_step(e, case, 0)
        e.close()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("geom", [1, 2])
def test_second_step_births_match_oracle(prec, geom):
    """The first sourcing sizes the buffers; later ones launch the births over an upper bound
    of their count without a host round trip.  The second step's population (census
    survivors of step 0, then the step-1 births) equals the oracle's sourcing from that
    census, bit for bit."""
    case = case1d(prec=prec) if geom == 1 else case2d(prec=prec)
    print(f"[second step births] geometry={geom} prec={prec}", flush=True)
    e = engine(case)
    engine_step(e, case, 0)
    e.source(3000, 30000, 2.0)
    e.transport()
    e.clean()
    census = e.export_particles()
    engine_step(e, case, 1)
    emitted, etot, n = e.source(3000, 30000, 2.0)
    g = e.export_particles()
    op, st = oracle_problem(case)
    rec, o_emit, o_etot, nb = op.source(st, 1, census, 3000, 30000, 2.0)
    assert n == rec.shape[0] == g.shape[0], (n, rec.shape, g.shape)
    assert etot == pytest.approx(o_etot, rel=1e-14)
    np.testing.assert_array_equal(_cells(g, geom), _cells(rec, geom))
    for name, idx in _fields(geom):
        np.testing.assert_array_equal(g[:, idx], rec[:, idx], err_msg=name)
    e.close()
