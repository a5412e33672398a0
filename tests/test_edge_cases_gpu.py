"""Edge cases through the C ABI, against the oracle where there is something to compare:

* an empty population: transport, clean, census tally and export do nothing and report zeros;
* ragged populations (1, 63, 65 and 2049 particles: partial waves, a partial work chunk, more
  than one clean tile) transported at FP64 equal the oracle's replay;
* a step in which every particle starts at t = 0 in cells so absorbing that none reaches census:
  clean leaves nothing, and the energy taken in equals the deposits plus the escapes;
* a source with no energy (zero temperatures, surfaces and radiation source) is refused with
  IMC_ENUMERIC, both on the first sourcing and on a later one (which launches its births
  without a host round trip and checks afterwards).
"""
import numpy as np
import pytest

from cases import case1d, case2d, engine, engine_step, oracle_problem

pytestmark = pytest.mark.gpu


def _width(geom):
    return 9 if geom == 1 else 10


def _fields(geom):
    if geom == 1:
        return [("t", 1), ("cell", 2), ("x", 3), ("mu", 4), ("E", 6), ("E0", 7)]
    return [("t", 0), ("ix", 1), ("iy", 2), ("x", 3), ("y", 4), ("theta", 5), ("E", 7), ("E0", 8)]


def _alive(rec, geom):
    return rec[:, 7] != -1.0 if geom == 1 else rec[:, 7] != -1.0


@pytest.mark.parametrize("geom", [1, 2])
def test_empty_population(geom):
    case = case1d(prec=64) if geom == 1 else case2d(prec=64)
    print(f"[edge] empty population, geometry={geom}", flush=True)
    e = engine(case)
    engine_step(e, case, 0)
    e.import_particles(np.zeros((0, _width(geom))))
    dep, lost, segs = e.transport()
    assert segs == 0 and lost == 0.0 and not dep.any()
    assert e.clean() == 0
    assert e.export_particles().shape == (0, _width(geom))
    assert not e.tally_census().any()
    e.close()


@pytest.mark.parametrize("n", [1, 63, 65, 2049])
@pytest.mark.parametrize("geom", [1, 2])
def test_ragged_population_replay(n, geom):
    case = case1d(prec=64) if geom == 1 else case2d(prec=64)
    print(f"[edge] ragged population n={n}, geometry={geom}", flush=True)
    op, st = oracle_problem(case)
    rec, _, _, _ = op.source(st, 2, None, 3000, 30000, 2.0)
    assert rec.shape[0] >= n
    rec = np.ascontiguousarray(rec[:n])
    e = engine(case)
    engine_step(e, case, 2)
    e.import_particles(rec)
    back = e.export_particles()
    for name, idx in _fields(geom):
        np.testing.assert_array_equal(back[:, idx], rec[:, idx], err_msg=name)
    dep, lost, segs = e.transport()
    g = e.export_particles()
    o_rec, o_dep, o_lost, o_segs = op.transport(st, 2, rec)
    assert segs == o_segs
    np.testing.assert_array_equal(_alive(g, geom), _alive(o_rec, geom))
    both = _alive(g, geom)
    for name, idx in _fields(geom):
        np.testing.assert_allclose(g[both, idx], o_rec[both, idx], rtol=1e-9, atol=1e-12, err_msg=name)
    np.testing.assert_allclose(dep, o_dep, rtol=1e-9, atol=1e-9 * max(np.abs(o_dep).max(), 1e-300))
    assert lost == pytest.approx(o_lost, rel=1e-9, abs=1e-15)
    e.close()


def test_everything_absorbed_before_census():
    case = case1d(prec=64)
    case["sa"] = case["sa"] * 1e5          # f sigma_a c dt >= 1e3 everywhere
    print("[edge] every particle absorbed or escaped", flush=True)
    op, st = oracle_problem(case)
    rec, _, _, _ = op.source(st, 1, None, 4000, 40000, 2.0)
    rec = rec.copy()
    rec[:, 1] = 0.0                        # all start at the beginning of the step
    e = engine(case)
    engine_step(e, case, 1)
    e.import_particles(rec)
    dep, lost, segs = e.transport()
    assert e.clean() == 0
    e_in = float(np.sum(rec[:, 6]))
    # deposits are energy per unit length (E / dx); scale index 0, scale 1
    e_dep = float(np.sum(dep[: case["nx"]] * case["dx"]))
    assert e_dep + lost == pytest.approx(e_in, rel=1e-12)
    o_rec, o_dep, o_lost, o_segs = op.transport(st, 1, rec)
    assert segs == o_segs
    np.testing.assert_allclose(dep, o_dep, rtol=1e-9, atol=1e-9 * np.abs(o_dep).max())
    e.close()


@pytest.mark.parametrize("geom", [1, 2])
def test_source_without_energy_is_refused(geom):
    from MixedPrecisionIMC._hip import ImcError
    case = case1d(prec=32) if geom == 1 else case2d(prec=32)
    print(f"[edge] zero source energy, geometry={geom}", flush=True)
    cold = dict(case)
    cold["temp"] = np.zeros_like(case["temp"])
    cold["radsource"] = np.zeros_like(case["radsource"])
    cold["tsurf"] = [np.zeros_like(t) for t in case["tsurf"]]
    e = engine(case)
    engine_step(e, cold, 0)                # first sourcing: refused before any birth
    with pytest.raises(ImcError):
        e.source(3000, 30000, 2.0)
    e.close()
    e = engine(case)
    engine_step(e, case, 0)                # a normal step sizes the buffers ...
    e.source(3000, 30000, 2.0)
    e.transport()
    e.clean()
    engine_step(e, cold, 1)                # ... so this one takes the direct path
    with pytest.raises(ImcError):
        e.source(3000, 30000, 2.0)
    e.close()
