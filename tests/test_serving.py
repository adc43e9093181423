"""Serving loop: continuous batching of mixed-size XAS / RIXS requests.

A request's spectra must be the same alone and in a mix where other requests
join and leave while it runs; the stated tolerance is exact equality, since
a grouped member's numbers do not depend on the other members.
"""
import numpy as np
import pytest

from edrixs_amd import serving

pytestmark = pytest.mark.gpu


def _engine(budget_gb=64.0):
    return serving.Engine(budget_gb=budget_gb, check_every=20, chunk=10, cf_tol=1e-6,
                          min_steps=40, max_steps=3000, n_in=2, npol=1, nw=300,
                          rixs_tol=1e-8, rixs_maxit=3000)


def _requests():
    reqs = serving.make_requests(6, rate=50.0, sizes=(12, 14), rixs_fraction=0.5, seed=3)
    kinds = [r.kind for r in reqs]
    assert "xas" in kinds and "rixs" in kinds
    return reqs


def test_serving_results_alone_equal_results_in_mix():
    import copy
    reqs = _requests()
    eng = _engine()
    # a budget that holds about two requests at a time, so later requests are
    # admitted while earlier ones are still running
    budget = 2.2 * max(eng.estimate_bytes(r) for r in reqs) / 1e9
    mix = _engine(budget).run(copy.deepcopy(reqs), realtime=False)
    assert len(mix) == len(reqs)
    by_id = {r.rid: r for r in mix}
    # some request was admitted after another had started its recurrences
    assert any(r.admitted > min(x.admitted for x in mix) for r in mix)
    for r in reqs:
        alone = _engine().run([copy.deepcopy(r)], realtime=False)[0]
        m = by_id[r.rid]
        assert m.nspectra == (1 if r.kind == "xas" else 2)
        assert m.steps == alone.steps
        for a, b in zip(m.spectra, alone.spectra):
            assert np.array_equal(a, b)
            assert np.all(np.isfinite(a)) and a.max() > 0


def test_serving_report_and_budget_admission():
    reqs = _requests()
    eng = _engine()
    done = eng.run(reqs, realtime=True)
    rep = serving.report(done)
    assert rep["requests"] == len(reqs) and rep["spectra"] == sum(r.nspectra for r in done)
    assert rep["latency_p50_s"] <= rep["latency_p99_s"] <= rep["latency_max_s"]
    assert all(r.done >= r.admitted >= r.arrival for r in done)
    assert abs(eng.used) < 1.0 and not eng.members


def test_serving_overlapped_setups_equal_serial_setups():
    # setups on the worker thread's own stream against setups between rounds
    # on the main thread: the same spectra, bit for bit, and the same steps
    import copy
    reqs = _requests()
    serial = _engine()
    serial.overlap = False
    a = {r.rid: r for r in serial.run(copy.deepcopy(reqs), realtime=False)}
    eng = _engine()
    b = {r.rid: r for r in eng.run(copy.deepcopy(reqs), realtime=False)}
    assert eng._stream is not None          # the worker's stream was used
    assert sorted(a) == sorted(b) == sorted(r.rid for r in reqs)
    for rid in a:
        assert a[rid].steps == b[rid].steps
        for x, y in zip(a[rid].spectra, b[rid].spectra):
            assert np.array_equal(x, y)
