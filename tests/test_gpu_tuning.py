"""Autotuned schedules: the choice persists across processes, so a matrix's
products round the same way in every process that uses the same cache file,
and the tuning key tells apart same-shape matrices whose storage forms differ.

The matrices are H_n of a norb=18 impurity model (262,548 rows, above the
131,072-row tuning threshold) in the interleaved core-hole layout.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np, torch
from edrixs_amd import core, siam
m = siam.SiamModel(nbath=8, seed=7)
bn = core.Basis.combinations(m.norb, m.nocc + 1)
H = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved")
H.prepare("c128i")
i = np.arange(H.n)
x = torch.from_numpy(np.cos(0.37 * i) + 1j * np.sin(0.11 * i)).cuda()
a, b, nrm = core.build_krylov(H, x, 60, precision="c128i")
p = torch.cuda.get_device_properties(0)
print(json.dumps({"lanes": H.spmv_lanes("c128i"), "m": H.m, "n": H.n, "nnz": H.nnz,
                  "forms": 3 if H.factored() else 0,
                  "device": "%s:%d" % (p.gcnArchName.replace(" ", "_"), p.multi_processor_count),
                  "alpha": [float(v).hex() for v in a], "beta": [float(v).hex() for v in b],
                  "norm": float(nrm).hex()}))
"""


def _run(cache):
    env = dict(os.environ, EDX_TUNE_CACHE=str(cache))
    out = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_tuned_schedule_persists_across_processes(tmp_path):
    cache = tmp_path / "tune.txt"
    first = _run(cache)
    assert first["m"] >= 131072 and first["forms"] == 3
    key = "%s %d %d %d 3 1 3 " % (first["device"], first["m"], first["n"], first["nnz"])
    lines = cache.read_text().splitlines()
    assert any(ln.startswith(key) for ln in lines), lines
    # a second process reads the choice instead of timing again: same schedule,
    # the same (alpha, beta) bit for bit
    second = _run(cache)
    assert second == first
    # the file decides: a process given another recorded choice runs it
    other = 144 if first["lanes"] != 144 else 145
    forced = tmp_path / "forced.txt"
    forced.write_text(key + "%d\n" % other)
    third = _run(forced)
    assert third["lanes"] == other


def test_tuning_key_separates_storage_forms():
    from edrixs_amd import core, siam
    import torch
    m = siam.SiamModel(nbath=8, seed=7)
    fn = o.fock_sorted(m.norb, m.nocc + 1)
    bn = core.Basis.combinations(m.norb, m.nocc + 1)
    # the same sparsity pattern with every Coulomb coefficient scaled by its own
    # factor: 399 distinct off-diagonal values, too many for the packed words
    rng = np.random.default_rng(1)
    coul_b = m.coul_n.copy()
    f = 1 + 1e-3 * rng.random(len(coul_b))
    coul_b["re"] *= f
    coul_b["im"] *= f
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    B = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, coul_b)
    assert (A.m, A.n, A.nnz) == (B.m, B.n, B.nnz)
    RA = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    RB = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, coul_b)
    A.prepare("c128i")
    B.prepare("c128i")
    ga, gb = A.spmv_lanes("c128i"), B.spmv_lanes("c128i")
    assert ga and gb and gb not in (145, 206)      # B has no packed (or factored) form
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    xd = torch.from_numpy(x).cuda()
    for H, R in ((A, RA), (B, RB)):
        y = H.matvec(xd, precision="c128i").cpu().numpy()
        ref = o.matvec(R, x)
        scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
        assert np.abs(y - ref).max() < 1e-14 * scale
    # rebuilt matrices of each kind reuse their own kind's choice
    A2 = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    B2 = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, coul_b)
    A2.prepare("c128i")
    B2.prepare("c128i")
    assert A2.spmv_lanes("c128i") == ga and B2.spmv_lanes("c128i") == gb
