"""Host-side logic that needs no device: serving request generation and
reports, the bench's reading of measured traffic and its one-GPU guard,
precision names."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_make_requests_poisson_and_parameter_ranges():
    from edrixs_amd import serving
    reqs = serving.make_requests(400, rate=20.0, sizes=(16, 18, 20, 22, 24), seed=5)
    t = np.array([r.arrival for r in reqs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 20.0) < 0.25 / 20.0          # mean inter-arrival 1/rate
    assert {r.norb for r in reqs} == {16, 18, 20, 22, 24}
    assert all(8 <= r.F2 <= 12 and 0.5 <= r.ten_dq <= 2 and 1 <= r.V <= 2 for r in reqs)
    assert {r.kind for r in reqs} == {"xas", "rixs"}
    again = serving.make_requests(400, rate=20.0, seed=5)
    assert [(r.kind, r.norb, r.seed) for r in again] == [(r.kind, r.norb, r.seed) for r in reqs]


def test_serving_report_percentiles():
    from edrixs_amd import serving
    done = []
    for i, lat in enumerate([1.0, 2.0, 3.0, 10.0]):
        r = serving.Request(rid=i, kind="xas", norb=16, seed=i, arrival=float(i))
        r.done = r.arrival + lat
        r.spectra = [np.zeros(3)]
        done.append(r)
    rep = serving.report(done)
    assert rep["requests"] == 4 and rep["spectra"] == 4
    assert rep["latency_p50_s"] == pytest.approx(2.5)
    assert rep["latency_max_s"] == 10.0 and rep["latency_p99_s"] <= 10.0
    assert rep["makespan_s"] == pytest.approx(13.0)


def test_serving_memory_estimate_grows_with_size():
    from edrixs_amd import serving
    eng = serving.Engine.__new__(serving.Engine)     # no device: only the estimator
    eng.precision, eng.n_in, eng.npol = "c128i", 4, 1
    est = [eng.estimate_bytes(serving.Request(0, k, n, 0, 0.0))
           for k in ("xas", "rixs") for n in (16, 20, 24)]
    assert est[0] < est[1] < est[2] and est[3] < est[4] < est[5]
    n_n = 6 * math.comb(24, 13)
    assert est[2] > n_n * 26 * 26                    # at least the H_n CSR bytes


def test_bench_reads_measured_traffic():
    sys.path.insert(0, ROOT)
    import bench
    for kernel, form in (("k_spmv_fac", "c128i"), ("k_spmv", "c128i"), ("k_spmv", "c64d"),
                         ("k_spmm_rl_k16", "c128i"), ("k_spmm_fac_k3", "c128i"),
                         ("k_spmm_fac_k2", "c128i")):
        v = bench.measured_traffic(kernel, form, 24)
        assert v is not None and 1e9 < v < 1e11
    assert bench.measured_traffic("k_spmv", "c128i", 20) is None
    assert bench.measured_traffic("k_nope", "c128i", 24) is None
    assert bench.traffic_key("k_spmm_fac", 3) == "k_spmm_fac_k3"
    assert bench.traffic_key("k_spmv_fac", 1) == "k_spmv_fac"


def test_precision_names():
    from edrixs_amd import core
    assert core._prec("c128") == core._prec("fp64") == 0
    assert core._prec("c64") == 1 and core._prec("mixed") == 2 and core._prec("indexed") == 3
    with pytest.raises(ValueError):
        core._prec("fp16")


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr and not r.stdout.strip()


def test_serving_admission_by_measured_bytes():
    # the admission decision alone (no device): a request is admitted when the
    # bytes the device holds now plus its scaled estimate fit the budget, or
    # when nothing runs; the oldest request past max_wait_s goes first
    from edrixs_amd import serving
    eng = serving.Engine(budget_gb=10.0, max_wait_s=5.0)
    reqs = serving.make_requests(3, rate=1.0, sizes=(16,), seed=1)
    for r, t, n in zip(reqs, (0.0, 1.0, 2.0), (20, 16, 18)):
        r.arrival, r.norb = t, n
    need16 = eng.need_bytes(reqs[1])
    assert 0 < need16 < 10e9
    # smallest first while nobody waited long
    assert eng.pick(reqs, running=1, now=3.0, in_use=0.0) is reqs[1]
    # over budget with others running: wait; alone: admitted anyway
    full = 10e9 - 0.5 * need16
    assert eng.pick(reqs, running=1, now=3.0, in_use=full) is None
    assert eng.pick(reqs, running=0, now=3.0, in_use=full) is reqs[1]
    # a measured size class scales the estimate
    eng.scale[16] = 3.0
    assert eng.need_bytes(reqs[1]) == 3.0 * eng.estimate_bytes(reqs[1])
    assert eng.pick(reqs, running=1, now=3.0, in_use=10e9 - 2.0 * need16) is None
    # the oldest past max_wait_s goes first, even if larger
    assert eng.pick(reqs, running=0, now=5.5, in_use=0.0) is reqs[0]
    assert eng.pick([], running=0, now=0.0, in_use=0.0) is None


def test_shifted_fom_rhs_block_rule(monkeypatch):
    # which multi-shift form shifted_fom_rhs picks, without a device: the
    # batched recurrence for 2-4 rows over a matrix with the slot-factored form
    # in the indexed precision, for 8 or more rows, or for a block within
    # SPLIT_BLOCK_BYTES; one recurrence per row otherwise
    import torch
    from edrixs_amd import core

    class FakeH:
        def __init__(self, n, fac):
            self.n, self._fac = n, fac

        def factored(self):
            return (6, 1, 0, 1, 1) if self._fac else None

    calls = []

    def batch(H, B, sigma, projections, **kw):
        calls.append(("batch", B.shape[0]))
        R, K = B.shape[0], len(sigma)
        return [None] * R, np.zeros((R, K), np.int32), np.zeros((R, K))

    def single(H, b, sigma, projections=(None,), **kw):
        calls.append(("single", 1))
        return [None], np.zeros(len(sigma), np.int32), np.zeros(len(sigma))

    monkeypatch.setattr(core, "shifted_fom_batch", batch)
    monkeypatch.setattr(core, "shifted_fom", single)
    big = 15_000_000
    cases = [(FakeH(big, True), 2, "c128i", ["batch"]),
             (FakeH(big, True), 2, "c128", ["single"] * 2),      # no factored product in c128
             (FakeH(big, False), 2, "c128i", ["single"] * 2),
             (FakeH(big, True), 5, "c128i", ["single"] * 5),     # the factored block stops at 4
             (FakeH(big, False), 8, "c64d", ["batch"]),
             (FakeH(1000, False), 3, "c128", ["batch"]),          # small block: one recurrence
             (FakeH(big, True), 1, "c128i", ["single"])]
    for H, R, prec, want in cases:
        calls.clear()
        B = torch.zeros((R, 4), dtype=torch.complex128)
        outs, it, res = core.shifted_fom_rhs(H, B, [1 + 1j, 2 + 1j], [None] * R, precision=prec)
        assert [c for c, _ in calls] == want, (H.n, R, prec, calls)
        assert it.shape == (R, 2) and res.shape == (R, 2) and len(outs) == R


def test_converged_spectra_compacts_the_block(monkeypatch):
    # converged_spectra without a device: a fake resumable batch whose vector j
    # reaches a final spectrum at step stop[j]; the block keeps only the
    # vectors still running after each chunk, in their order, and every
    # spectrum is taken at its own stopping chunk
    import torch
    from edrixs_amd import core
    stop = [40, 120, 40, 80, 120]
    keeps = []

    class FakeBatch:
        def __init__(self, H, V0, precision="c128", capacity=256):
            self.ids = [int(v) for v in V0[:, 0].real]
            self.K = len(self.ids)
            self.done = 0

        def advance(self, nsteps):
            self.done += nsteps
            return self.done

        def coefficients(self, count=None):
            n = self.done if count is None else count
            al = [np.full(n, float(i)) for i in self.ids]
            return al, [np.ones(n) for _ in self.ids], np.ones(self.K)

        def keep(self, idx):
            keeps.append(list(idx))
            self.ids = [self.ids[j] for j in idx]
            self.K = len(self.ids)

        def close(self):
            pass

    def fake_cf(al, be, nr, e_gs, grid, gamma):
        j, n = int(al[0]), len(al)
        return np.array([1.0 if n >= stop[j] else 1.0 + 1.0 / n])

    class FakeH:
        n = 100

        def factored(self):
            return None

    monkeypatch.setattr(core, "KrylovBatch", FakeBatch)
    monkeypatch.setattr(core, "cf_spectrum", fake_cf)
    V0 = torch.zeros((5, 100), dtype=torch.complex128)
    V0[:, 0] = torch.arange(5, dtype=torch.float64)
    spec, steps, change = core.converged_spectra(FakeH(), V0, 0.0, None, 0.1, chunk=20,
                                                 min_steps=20, max_steps=400)
    # a spectrum is final once unchanged over a chunk: vector j stops one
    # chunk after stop[j]
    assert steps == [60, 140, 60, 100, 140]
    # after 60 steps vectors 0 and 2 are final: keep 1, 3, 4; after 100 the
    # second of those (vector 3): keep positions 0 and 2 (vectors 1 and 4)
    assert keeps == [[1, 3, 4], [0, 2]]
    assert all(c <= 1e-8 for c in change)
