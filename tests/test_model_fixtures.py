"""The synthetic models' coefficient tensors against the reference's own
builders (fixtures written by tests/golden/make_model_fixtures.py from
get_umat_slater, get_trans_oper, atom_hsoc and cf_cubic_d).  The closed-form
checks (Racah energies, SOC eigenvalues, dipole sum rules) leave the signs and
phases open; these fix them: every entry to 1e-13."""
import os

import numpy as np

from edrixs_amd import siam

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model_fixtures.npz")


def _dense4(idx, val, n):
    u = np.zeros((n,) * 4, np.complex128)
    u[tuple(idx.astype(np.int64).T)] = val
    return u


def test_model_tensors_match_the_reference_builders():
    f = np.load(FIX)
    for name in ("bench", "full"):
        p = f["params_%s" % name]
        d, dp, zd, zp, dq = p[0:3], p[3:12], p[12], p[13], p[14]
        ud = _dense4(f["umat_d_%s_idx" % name], f["umat_d_%s_val" % name], 10)
        udp = _dense4(f["umat_dp_%s_idx" % name], f["umat_dp_%s_val" % name], 16)
        pairs = [(siam.umat_slater_d(*d), ud), (siam.umat_slater_dp(*dp), udp),
                 (siam.atom_hsoc(2, zd), f["hsoc_d_%s" % name]),
                 (siam.atom_hsoc(1, zp), f["hsoc_p_%s" % name]),
                 (siam.cf_cubic_d(dq), f["cf_cubic_d_%s" % name])]
        for ours, ref in pairs:
            assert ours.shape == ref.shape and np.abs(ref).max() > 0.05
            np.testing.assert_allclose(ours, ref, rtol=0, atol=1e-13)
    T = f["trans_dp"]
    np.testing.assert_allclose(siam.dipole_dp(), T, rtol=0, atol=1e-13)
    # the exchange part and the x / y phases are what the fixture fixes: the
    # complex conjugate dipole or the transposed exchange would not match
    assert np.abs(np.conj(T) - T).max() > 0.1


def test_benchmark_model_uses_the_fixture_parameters():
    f = np.load(FIX)
    p = f["params_bench"]
    m = siam.SiamModel(nbath=2, seed=0)
    assert m.params["F"] == tuple(p[0:3]) and m.params["zeta_d"] == p[12]
    assert m.params["zeta_p"] == p[13] and m.params["ten_dq"] == p[14]
    assert m.params["dp"] == tuple(p[6:10]) and m.params["pp"] == tuple(p[10:12])
