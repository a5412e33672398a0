"""bench.py reads two figures of its roofline object from the committed rocprofv3 summaries:
HBM bytes per transport launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) and the VALU busy
fraction (SQ counters).  Both lines must stay parsable and plausible, and the summaries must give
the full commands that made them."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_traffic_and_valu_busy_are_parsable():
    import bench
    traffic, src = bench.profiled_traffic()
    assert src and traffic is not None
    busy, vsrc = bench.profiled_valu_busy()
    assert vsrc and busy is not None
    assert 1e7 < traffic < 1e10            # bytes per launch of the Marshak transport kernel
    assert 0.0 < busy <= 1.0


def test_profiles_state_their_commands():
    d = os.path.join(ROOT, "profiles")
    names = sorted(f for f in os.listdir(d) if f.endswith(".md"))
    assert names
    for f in names:
        text = open(os.path.join(d, f)).read()
        assert "rocprofv3 " in text, f
        assert "-- python3 bench.py" in text, f      # each runs committed files only
