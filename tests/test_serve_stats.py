"""The serving loop's latency summary (pointvortex_amd/serve.py, host logic only): runs that a
wall-clock cap stops are counted as unfinished, with the time they had waited as a lower bound
of their latency, rather than dropped from the percentiles."""
import numpy as np

from pointvortex_amd.serve import Result, ServerStats


def _result(rid, state, arrival, finished):
    return Result(rid, state, arrival, arrival, finished, 5, 1, "", "fp64")


def test_summary_without_unfinished_runs():
    st = ServerStats(results=[_result(k, "finished", 0.1 * k, 0.1 * k + 0.5 + 0.01 * k)
                              for k in range(10)], wall=2.0, attempts=60)
    s = st.summary()
    lat = [0.5 + 0.01 * k for k in range(10)]
    assert s["completed"] == 10 and s["unfinished"] == 0 and not s["latency_lower_bound"]
    assert np.isclose(s["p50_s"], np.percentile(lat, 50))
    assert np.isclose(s["p99_s"], np.percentile(lat, 99))
    assert np.isclose(s["solves_per_s"], 5.0)


def test_unfinished_runs_enter_the_percentiles_as_lower_bounds():
    res = [_result(k, "finished", 0.0, 1.0) for k in range(9)]
    res.append(_result(9, "unfinished", 2.0, 90.0))     # stopped by the cap after 88 s
    res.append(_result(10, "failed", 0.0, 0.3))          # failures stay out of the latencies
    res.append(_result(11, "not_arrived", 95.0, 90.0))   # never arrived: no latency
    st = ServerStats(results=res, wall=90.0, attempts=100)
    s = st.summary()
    assert s["completed"] == 9 and s["unfinished"] == 1 and s["failed"] == 1
    assert s["not_arrived"] == 1
    assert s["latency_lower_bound"]
    assert np.isclose(s["p99_s"], np.percentile([1.0] * 9 + [88.0], 99))
    assert s["p99_s"] > 80.0
    assert np.isclose(s["solves_per_s"], 9 / 90.0)
