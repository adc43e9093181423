"""Writes tests/golden/model_fixtures.npz from the reference's own model
builders, run in the build container (where the reference checkout is at
/root/reference); the fixture, not the reference, travels with the tests.

The synthetic impurity models take their coefficient tensors from these
functions, whose conventions no closed form fixes (the sign of the exchange
part of U, the phases of the x and y dipoles, the phase convention of l.s):

  get_umat_slater('d', F0, F2, F4)                edrixs/coulomb_utensor.py:373-633
  get_umat_slater('dp', F0_dd, F2_dd, F4_dd, F0_dp, F2_dp, G1_dp, G3_dp, F0_pp, F2_pp)
  get_trans_oper('dp')                            edrixs/photon_transition.py:61-183
  atom_hsoc('d', zeta), atom_hsoc('p', zeta)      edrixs/soc.py:6-107
  cf_cubic_d(ten_dq)                              edrixs/angular_momentum.py:597-625

at the parameters of the benchmark models (Slater d (8, 10, 6.2), dp
exchange (6.0, 4.5, 2.6), zeta_d 0.08, zeta_p 11.0, 10Dq 1.2) and at a second
set with every parameter nonzero.  Rank-4 tensors are stored sparse (indices
and values of the entries above 1e-14).

Usage (build container only):  python tests/golden/make_model_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def sparse4(u):
    idx = np.argwhere(np.abs(u) > 1e-14)
    return idx.astype(np.int8), u[tuple(idx.T)].astype(np.complex128)


def main():
    sys.path.insert(0, REF)
    import edrixs
    out = {}
    sets = {"bench": dict(d=(8.0, 10.0, 6.2), dp=(8.0, 10.0, 6.2, 0.0, 6.0, 4.5, 2.6, 0.0, 0.0),
                          zd=0.08, zp=11.0, dq=1.2),
            "full": dict(d=(4.0, 2.0, 0.5), dp=(3.0, 7.0, 5.0, 1.3, 2.2, 1.7, 0.9, 0.6, 0.8),
                         zd=0.3, zp=2.5, dq=-0.7)}
    for name, p in sets.items():
        i, v = sparse4(np.asarray(edrixs.get_umat_slater("d", *p["d"]), dtype=np.complex128))
        out["umat_d_%s_idx" % name], out["umat_d_%s_val" % name] = i, v
        i, v = sparse4(np.asarray(edrixs.get_umat_slater("dp", *p["dp"]), dtype=np.complex128))
        out["umat_dp_%s_idx" % name], out["umat_dp_%s_val" % name] = i, v
        out["hsoc_d_%s" % name] = np.asarray(edrixs.atom_hsoc("d", p["zd"]), dtype=np.complex128)
        out["hsoc_p_%s" % name] = np.asarray(edrixs.atom_hsoc("p", p["zp"]), dtype=np.complex128)
        out["cf_cubic_d_%s" % name] = np.asarray(edrixs.cf_cubic_d(p["dq"]), dtype=np.complex128)
        out["params_%s" % name] = np.array(p["d"] + p["dp"] + (p["zd"], p["zp"], p["dq"]))
    out["trans_dp"] = np.asarray(edrixs.get_trans_oper("dp"), dtype=np.complex128)
    np.savez_compressed(os.path.join(HERE, "model_fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "model_fixtures.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
