"""CPU oracle for the EDRIXS exact-diagonalization core (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product package never does.

Two independent checkers live here:

* ``liboracle.so`` (``edx_oracle.c``): a step-by-step C restatement of the
  reference Fortran (``src/ham.f90``, ``src/spmv.f90``, ``src/lanczos.f90``,
  ``src/linsys.f90``, ``src/ed_driver.f90``); wrapped below with numpy arrays.
* ``dense_*`` helpers: Jordan-Wigner operator matrices built from first
  principles (products of single-fermion matrices), used to pin the C
  restatement on small Fock spaces (operator order, signs, index convention).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ROW_PLAIN, ROW_COREHOLE, ROW_FULLCORE = 0, 1, 2
COL_PLAIN, COL_COREHOLE, COL_VALONLY = 0, 1, 2

T2 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("re", "<f8"), ("im", "<f8")])
T4 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("ind3", "<i4"), ("ind4", "<i4"),
               ("re", "<f8"), ("im", "<f8")])


def compile_lib():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            compile_lib()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.orc_build.argtypes = [ctypes.c_int, i64, P, ctypes.c_int, i64, P, ctypes.c_int,
                                ctypes.c_int, i64, P, i64, P, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int, i64, i64,
                                P, P, P, P, P]
        L.orc_free.argtypes = [P]
        L.orc_matvec_csr.argtypes = [i64, P, P, P, P, P]
        L.orc_build_krylov.argtypes = [i64, P, P, P, P, ctypes.c_int, P, P, P, P]
        L.orc_lanczos.argtypes = [i64, P, P, P, P, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, P, P, P]
        L.orc_tridiag_eigs.argtypes = [ctypes.c_int, P, P, P, P]
        L.orc_pminres.argtypes = [i64, P, P, P, P, P, ctypes.c_double, ctypes.c_int, P, P]
        L.orc_denmat.argtypes = [i64, P, ctypes.c_int, P, P]
        L.orc_opavg.argtypes = [i64, P, P, P, P, P]
        L.orc_cf_spectrum.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, P, P, P]
        L.orc_make_newfock.argtypes = [ctypes.c_char, ctypes.c_int, i64, P, P]
        L.orc_binary_search.argtypes = [i64, P, i64]
        L.orc_binary_search.restype = i64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def as_t2(terms):
    """Accept an (n,4) [i, j, re, im] array or a T2 record array."""
    if isinstance(terms, np.ndarray) and terms.dtype == T2:
        return np.ascontiguousarray(terms)
    a = np.asarray(terms, dtype=np.float64).reshape(-1, 4)
    out = np.zeros(len(a), dtype=T2)
    out["ind1"], out["ind2"] = a[:, 0], a[:, 1]
    out["re"], out["im"] = a[:, 2], a[:, 3]
    return out


def as_t4(terms):
    if isinstance(terms, np.ndarray) and terms.dtype == T4:
        return np.ascontiguousarray(terms)
    a = np.asarray(terms, dtype=np.float64).reshape(-1, 6)
    out = np.zeros(len(a), dtype=T4)
    for k, name in enumerate(("ind1", "ind2", "ind3", "ind4")):
        out[name] = a[:, k]
    out["re"], out["im"] = a[:, 4], a[:, 5]
    return out


class CSR:
    """0-based CSR: rowptr int64[m+1], cols int32[nnz], vals complex128[nnz]."""

    def __init__(self, rowptr, cols, vals, ncols):
        self.rowptr, self.cols, self.vals, self.ncols = rowptr, cols, vals, ncols

    @property
    def m(self):
        return len(self.rowptr) - 1

    @property
    def nnz(self):
        return len(self.cols)

    def todense(self):
        d = np.zeros((self.m, self.ncols), dtype=np.complex128)
        for r in range(self.m):
            b, e = self.rowptr[r], self.rowptr[r + 1]
            d[r, self.cols[b:e]] = self.vals[b:e]
        return d


def build(rowkind, fock_left, colkind, fock_right, num_val_orbs, num_core_orbs, hop, coul,
          omega=0j, ham_sign=1.0, with_omega=True, row_begin=0, row_end=-1):
    L = lib()
    fl = np.ascontiguousarray(fock_left, dtype=np.int64)
    fr = np.ascontiguousarray(fock_right, dtype=np.int64)
    h = as_t2(hop) if hop is not None and len(hop) else np.zeros(0, dtype=T2)
    c = as_t4(coul) if coul is not None and len(coul) else np.zeros(0, dtype=T4)
    m = ctypes.c_int64()
    nnz = ctypes.c_int64()
    rp, cp, vp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.orc_build(rowkind, len(fl), _p(fl), colkind, len(fr), _p(fr), num_val_orbs,
                num_core_orbs, len(h), _p(h), len(c), _p(c), float(np.real(omega)),
                float(np.imag(omega)), float(ham_sign), int(with_omega), int(row_begin),
                int(row_end), ctypes.byref(m), ctypes.byref(nnz), ctypes.byref(rp),
                ctypes.byref(cp), ctypes.byref(vp))
    M, NNZ = m.value, nnz.value
    rowptr = np.ctypeslib.as_array((ctypes.c_int64 * (M + 1)).from_address(rp.value)).copy()
    if NNZ:
        cols = np.ctypeslib.as_array((ctypes.c_int32 * NNZ).from_address(cp.value)).copy()
        vals = np.ctypeslib.as_array((ctypes.c_double * (2 * NNZ)).from_address(vp.value))
        vals = vals.copy().view(np.complex128)
    else:
        cols = np.zeros(0, np.int32)
        vals = np.zeros(0, np.complex128)
    for q in (rp, cp, vp):
        if q.value:
            L.orc_free(q)
    if colkind == COL_COREHOLE:
        ncols = len(fr) * num_core_orbs
    else:
        ncols = len(fr)
    return CSR(rowptr, cols, vals, ncols)


# Reference builder names (src/ham.f90) on top of the generic row loop.
def build_ham_i(fock, hop, coul, omega=0j, ham_sign=1.0, **kw):
    return build(ROW_PLAIN, fock, COL_PLAIN, fock, 0, 0, hop, coul, omega, ham_sign, **kw)


def build_ham_n(fock_n, num_val_orbs, num_core_orbs, hop, coul, omega=0j, ham_sign=1.0, **kw):
    return build(ROW_COREHOLE, fock_n, COL_COREHOLE, fock_n, num_val_orbs, num_core_orbs,
                 hop, coul, omega, ham_sign, **kw)


def build_transop_i(fock_n, fock_i, num_val_orbs, num_core_orbs, trans, **kw):
    return build(ROW_COREHOLE, fock_n, COL_VALONLY, fock_i, num_val_orbs, num_core_orbs,
                 trans, None, 0j, 1.0, False, **kw)


def build_transop_f(fock_f, fock_n, num_val_orbs, num_core_orbs, trans, **kw):
    return build(ROW_FULLCORE, fock_f, COL_COREHOLE, fock_n, num_val_orbs, num_core_orbs,
                 trans, None, 0j, 1.0, False, **kw)


def matvec(A, x, y=None):
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.zeros(A.m, np.complex128) if y is None else np.array(y, dtype=np.complex128)
    L.orc_matvec_csr(A.m, _p(A.rowptr), _p(A.cols), _p(A.vals.view(np.float64)), _p(x), _p(y))
    return y


def build_krylov(A, v0, maxn):
    L = lib()
    v0 = np.ascontiguousarray(v0, dtype=np.complex128)
    alpha = np.zeros(maxn)
    beta = np.zeros(maxn + 1)
    neff = ctypes.c_int()
    norm = ctypes.c_double()
    L.orc_build_krylov(A.m, _p(A.rowptr), _p(A.cols), _p(A.vals.view(np.float64)), _p(v0),
                       maxn, ctypes.byref(neff), _p(alpha), _p(beta), ctypes.byref(norm))
    n = neff.value
    # pole-file convention (src/io.f90:515-541): alpha_1..n, beta_1..n
    return alpha[:n].copy(), beta[1:n + 1].copy(), norm.value


def lanczos(A, v0, nev, maxiter, tol, nvec):
    L = lib()
    v0 = np.ascontiguousarray(v0, dtype=np.complex128)
    ev = np.zeros(nev)
    vecs = np.zeros((nvec, A.m), np.complex128)
    nk = ctypes.c_int()
    rc = L.orc_lanczos(A.m, _p(A.rowptr), _p(A.cols), _p(A.vals.view(np.float64)), _p(v0),
                       nev, maxiter, tol, nvec, _p(ev), _p(vecs), ctypes.byref(nk))
    if rc:
        raise RuntimeError("orc_lanczos failed %d" % rc)
    return ev, vecs, nk.value


def tridiag_eigs(d, e, vectors=False):
    L = lib()
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = len(d)
    w = np.zeros(n)
    z = np.zeros((n, n)) if vectors else None
    rc = L.orc_tridiag_eigs(n, _p(d), _p(e), _p(w), _p(z) if vectors else None)
    if rc:
        raise RuntimeError("tridiagonal QL did not converge")
    return (w, z) if vectors else w


def pminres(A, b, tol, maxit):
    L = lib()
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros(A.m, np.complex128)
    it = ctypes.c_int()
    res = ctypes.c_double()
    L.orc_pminres(A.m, _p(A.rowptr), _p(A.cols), _p(A.vals.view(np.float64)), _p(b), _p(x),
                  tol, maxit, ctypes.byref(it), ctypes.byref(res))
    return x, it.value, res.value


def denmat(fock, num_val_orbs, vec):
    L = lib()
    f = np.ascontiguousarray(fock, dtype=np.int64)
    v = np.ascontiguousarray(vec, dtype=np.complex128)
    dm = np.zeros((num_val_orbs, num_val_orbs), np.complex128)
    L.orc_denmat(len(f), _p(f), num_val_orbs, _p(v), _p(dm))
    return dm


def opavg(A, v):
    """<v|A|v> as the opavg driver computes it (src/opavg_driver.f90:93-108)."""
    L = lib()
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros(2)
    L.orc_opavg(A.m, _p(A.rowptr), _p(A.cols), _p(A.vals.view(np.float64)), _p(v), _p(out))
    return complex(out[0], out[1])


def opavg_lines(evals, avgs):
    """The lines of opavg.dat: format (1i5, 4f20.10) with three values per
    record (src/io.f90:638-659)."""
    return ["%5d%20.10f%20.10f%20.10f" % (i + 1, e, a.real, a.imag)
            for i, (e, a) in enumerate(zip(evals, avgs))]


def cf_spectrum(alpha, beta, norm, e_gs, omega, gamma):
    """alpha, beta as in a pole file (alpha_1..n, beta_1..n)."""
    L = lib()
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    ga = np.ascontiguousarray(np.broadcast_to(gamma, om.shape), dtype=np.float64)
    out = np.zeros(len(om))
    L.orc_cf_spectrum(len(a), _p(a), _p(b), norm, e_gs, len(om), _p(om), _p(ga), _p(out))
    return out


def make_newfock(kind, pos, old):
    L = lib()
    new = np.zeros(1, np.int64)
    sgn = np.zeros(1, np.int32)
    L.orc_make_newfock(kind.encode(), pos, old, _p(new), _p(sgn))
    return int(new[0]), int(sgn[0])


def binary_search(vec, m):
    v = np.ascontiguousarray(vec, dtype=np.int64)
    return int(lib().orc_binary_search(len(v), _p(v), int(m)))


# ----------------------------------------------------------------- dense checks

def _popcount64(w):
    """Population count of each element of a uint64 array (SWAR bit sums)."""
    w = w - ((w >> np.uint64(1)) & np.uint64(0x5555555555555555))
    w = (w & np.uint64(0x3333333333333333)) + ((w >> np.uint64(2)) & np.uint64(0x3333333333333333))
    w = (w + (w >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    return (w * np.uint64(0x0101010101010101)) >> np.uint64(56)


def fock_sorted(norb, nocc):
    """All norb-bit words with popcount nocc in ascending order, by brute force:
    every word below 2**norb is tested, in ascending chunks, and the ones with
    nocc set bits are kept (the definition of the basis, edrixs/fock_basis.py:
    398-451, with no combinatorial shortcut)."""
    chunk = 1 << 22
    out = []
    for lo in range(0, 1 << norb, chunk):
        w = np.arange(lo, min(lo + chunk, 1 << norb), dtype=np.uint64)
        out.append(w[_popcount64(w) == nocc].astype(np.int64))
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def jw_annihilators(norb):
    """Dense c_p (p = 0..norb-1) on the 2^norb Fock space, occupation of p = bit p.

    Built as Kronecker products a_p = Z x ... x Z x a x I x ... x I with the
    string over the lower orbitals (so c_p|w> = (-1)^{#occ below p}|w - 2^p>).
    """
    a = np.array([[0.0, 1.0], [0.0, 0.0]])   # |0><1| on one mode (index = occupation)
    z = np.diag([1.0, -1.0])
    eye = np.eye(2)
    ops = []
    for p in range(norb):
        m = np.array([[1.0]])
        # bit q of the index <-> kron factor at position norb-1-q (big-endian kron)
        for q in reversed(range(norb)):
            f = z if q < p else (a if q == p else eye)
            m = np.kron(m, f)
        ops.append(m)
    return ops


def dense_operator(norb, emat=None, umat=None):
    """Dense sum_ij E_ij c+_i c_j + sum_ijkl U_ijkl c+_i c+_j c_k c_l on 2^norb states."""
    cs = jw_annihilators(norb)
    cd = [c.T.copy() for c in cs]
    dim = 2 ** norb
    H = np.zeros((dim, dim), dtype=np.complex128)
    if emat is not None:
        for i, j in zip(*np.nonzero(np.abs(emat) > 1e-12)):
            H += emat[i, j] * (cd[i] @ cs[j])
    if umat is not None:
        for i, j, k, l in zip(*np.nonzero(np.abs(umat) > 1e-12)):
            H += umat[i, j, k, l] * (cd[i] @ cd[j] @ cs[k] @ cs[l])
    return H


def n_choose_k(n, k):
    return math.comb(n, k)
