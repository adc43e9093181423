"""In-memory interface to the device ED core (libedx through ctypes).

Objects hold device-resident data: :class:`Basis` (sorted Fock words) and
:class:`CSRMatrix` (row pointers int64, columns int32, complex128 values plus an
optional complex64 copy).  Vectors are torch tensors on the HIP device
(complex128); torch only provides the device memory.

The functions mirror the Fortran routines they replace:

==========================  =====================================================
``build_ham_i``             ``build_ham_i``      src/ham.f90:2-349
``build_ham_n``             ``build_ham_n``      src/ham.f90:352-711
``build_transop_i``         ``build_transop_i``  src/ham.f90:714-912
``build_transop_f``         ``build_transop_f``  src/ham.f90:915-1115
``CSRMatrix.spmv``          ``matvec_csr``       src/spmv.f90:73-103 (y += A x)
``lanczos_eigs``            ``diag_ham_lanczos`` src/lanczos.f90:6-167
``lanczos_ground``          ``diag_ham_lanczos`` src/lanczos.f90:6-167 (two passes, no basis)
``block_lanczos_eigs``      ``diag_ham_lanczos`` src/lanczos.f90:6-167 (8-vector blocks)
``trlan_eigs``              ``diag_ham_arpack``  src/arpack.f90:2-112 (thick restart)
``build_krylov``            ``build_krylov_mp``  src/lanczos.f90:210-303
``minres``                  ``pminres_csr``      src/linsys.f90:3-128
``denmat``                  density matrix       src/ed_driver.f90:219-248
``cf_spectrum``             ``get_spectra_from_poles`` edrixs/plot_spectrum.py:9-61
==========================  =====================================================
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import EDX_C128, EDX_C64, EDX_C64D, EDX_C128I

T2 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("re", "<f8"), ("im", "<f8")])
T4 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("ind3", "<i4"), ("ind4", "<i4"),
               ("re", "<f8"), ("im", "<f8")])

PRECISIONS = {"c128": EDX_C128, "fp64": EDX_C128, "complex128": EDX_C128,
              "c64": EDX_C64, "complex64": EDX_C64,
              "c64d": EDX_C64D, "mixed": EDX_C64D,
              "c128i": EDX_C128I, "indexed": EDX_C128I,
              EDX_C128: EDX_C128, EDX_C64: EDX_C64, EDX_C64D: EDX_C64D, EDX_C128I: EDX_C128I}


def _prec(p):
    try:
        return PRECISIONS[p]
    except KeyError:
        raise ValueError("unknown precision %r" % (p,))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def torch():
    import torch as _t
    return _t


def device():
    return torch().device("cuda", 0)


def as_terms2(t):
    if t is None or len(t) == 0:
        return np.zeros(0, dtype=T2)
    t = np.asarray(t)
    if t.dtype.names:
        out = np.zeros(len(t), dtype=T2)
        for f in T2.names:
            out[f] = t[f]
        return out
    a = np.asarray(t, dtype=np.float64).reshape(-1, 4)
    out = np.zeros(len(a), dtype=T2)
    out["ind1"], out["ind2"], out["re"], out["im"] = a[:, 0], a[:, 1], a[:, 2], a[:, 3]
    return out


def as_terms4(t):
    if t is None or len(t) == 0:
        return np.zeros(0, dtype=T4)
    t = np.asarray(t)
    if t.dtype.names:
        out = np.zeros(len(t), dtype=T4)
        for f in T4.names:
            out[f] = t[f]
        return out
    a = np.asarray(t, dtype=np.float64).reshape(-1, 6)
    out = np.zeros(len(a), dtype=T4)
    for k, f in enumerate(("ind1", "ind2", "ind3", "ind4")):
        out[f] = a[:, k]
    out["re"], out["im"] = a[:, 4], a[:, 5]
    return out


class Basis:
    """Sorted Fock words resident on the device."""

    def __init__(self, handle):
        self._h = handle
        n = ctypes.c_int64()
        c = ctypes.c_int()
        _lib.call("edx_basis_size", handle, ctypes.byref(n), ctypes.byref(c))
        self.n = n.value
        self.complete = bool(c.value)

    @classmethod
    def combinations(cls, nbits, nocc):
        h = ctypes.c_void_p()
        _lib.call("edx_basis_combinations", nbits, nocc, ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_words(cls, words, nbits):
        w = np.ascontiguousarray(words, dtype=np.int64)
        h = ctypes.c_void_p()
        _lib.call("edx_basis_from_words", _ptr(w), len(w), nbits, ctypes.byref(h))
        return cls(h)

    def words(self):
        out = np.zeros(self.n, dtype=np.int64)
        _lib.call("edx_basis_download", self._h, _ptr(out))
        return out

    def lookup(self, words):
        w = np.ascontiguousarray(words, dtype=np.int64)
        out = np.zeros(len(w), dtype=np.int64)
        _lib.call("edx_basis_lookup", self._h, _ptr(w), len(w), _ptr(out))
        return out

    def __len__(self):
        return self.n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None and _lib._lib is not None:
            _lib._lib.edx_basis_free(h)
            self._h = None


class CSRMatrix:
    """CSR matrix resident on the device (0-based, int64 row pointers)."""

    def __init__(self, handle):
        self._h = handle
        m, n, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        p = ctypes.c_int()
        _lib.call("edx_csr_info", handle, ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                  ctypes.byref(p))
        self.m, self.n, self.nnz = m.value, n.value, nnz.value
        self.has_c64 = p.value == EDX_C64

    @property
    def shape(self):
        return (self.m, self.n)

    def download(self):
        rowptr = np.zeros(self.m + 1, dtype=np.int64)
        cols = np.zeros(self.nnz, dtype=np.int32)
        vals = np.zeros(self.nnz, dtype=np.complex128)
        _lib.call("edx_csr_download", self._h, _ptr(rowptr), _ptr(cols), _ptr(vals))
        return rowptr, cols, vals

    def values(self, precision="c128"):
        p = _prec(precision)
        out = np.zeros(self.nnz, dtype=np.complex128 if p == EDX_C128 else np.complex64)
        _lib.call("edx_csr_download_values", self._h, p, _ptr(out))
        return out

    def prepare(self, precision):
        """Make the value storage a precision needs resident on the device."""
        _lib.call("edx_csr_prepare", self._h, _prec(precision))
        return self

    def spmv_lanes(self, precision="c128", set=0):
        """Lanes per row of the single-vector SpMV (set > 0 fixes it; 0 = rule)."""
        out = ctypes.c_int()
        _lib.call("edx_csr_spmv_lanes", self._h, _prec(precision), int(set), ctypes.byref(out))
        return out.value

    def factored(self):
        """The slot-factored form made by prepare("c128i") on a core-hole matrix
        in the interleaved layout: (core slots, entries stored once per valence
        row, entries stored per row, typed block words, block types), or None
        when absent."""
        nc, hv, r = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("edx_csr_factored_info", self._h, ctypes.byref(nc), ctypes.byref(hv),
                  ctypes.byref(r))
        if not nc.value:
            return None
        nb, nt = ctypes.c_int64(), ctypes.c_int()
        _lib.call("edx_csr_factored_blocks", self._h, ctypes.byref(nb), ctypes.byref(nt))
        return nc.value, hv.value, r.value, nb.value, nt.value

    def add_c64(self):
        _lib.call("edx_csr_add_c64", self._h)
        self.has_c64 = True
        return self

    def spmv(self, x, y, precision="c128"):
        """y += A x for complex128 device tensors (src/spmv.f90:73-103)."""
        _check_vec(x, self.n)
        _check_vec(y, self.m)
        _lib.call("edx_spmv", self._h, _prec(precision), _dptr(x), _dptr(y))
        return y

    def matvec(self, x, precision="c128"):
        y = torch().zeros(self.m, dtype=torch().complex128, device=x.device)
        return self.spmv(x, y, precision)

    def spmm(self, X, Y, precision="c128"):
        """Y[:, k] += A X[:, k]; X, Y column-major complex128 given as (nrhs, n) tensors."""
        assert X.is_contiguous() and Y.is_contiguous()
        assert X.shape[1] == self.n and Y.shape[1] == self.m and X.shape[0] == Y.shape[0]
        _lib.call("edx_spmm", self._h, _prec(precision), X.shape[0], _dptr(X), _dptr(Y))
        return Y

    def todense(self):
        rp, c, v = self.download()
        d = np.zeros((self.m, self.n), np.complex128)
        for r in range(self.m):
            d[r, c[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
        return d

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None and _lib._lib is not None:
            _lib._lib.edx_csr_free(h)
            self._h = None


def _check_vec(t, n):
    T = torch()
    if t.dtype != T.complex128 or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
        raise ValueError("expected a contiguous complex128 device tensor of length %d" % n)


def _handle_out(fn, *args):
    h = ctypes.c_void_p()
    _lib.call(fn, *args, ctypes.byref(h))
    return CSRMatrix(h)


def build_ham_i(basis, hop, coul, omega=0j, ham_sign=1.0, precision="c128"):
    h, c = as_terms2(hop), as_terms4(coul)
    return _handle_out("edx_build_ham_i", basis._h, _ptr(h), len(h), _ptr(c), len(c),
                       float(np.real(omega)), float(np.imag(omega)), float(ham_sign),
                       _prec(precision))


# index layout of the core-hole space (edx_set_corehole_layout): "reference" is
# core_slot * dim + valence (src/ham.f90:419-446), "interleaved" is
# valence * num_core_orbs + core_slot (the same matrix, permuted)
COREHOLE_LAYOUTS = {"reference": 0, "interleaved": 1}


def _build_in_layout(layout, fn, *args):
    prev = ctypes.c_int()
    _lib.call("edx_set_corehole_layout", COREHOLE_LAYOUTS[layout], ctypes.byref(prev))
    try:
        return _handle_out(fn, *args)
    finally:
        _lib.call("edx_set_corehole_layout", prev.value, None)


def corehole_permutation(dim_val, num_core_orbs):
    """p with x_interleaved = x_reference[p] for core-hole-space vectors."""
    i = np.arange(dim_val * num_core_orbs, dtype=np.int64)
    return (i % num_core_orbs) * dim_val + i // num_core_orbs


def build_ham_n(basis_n, num_val_orbs, num_core_orbs, hop, coul, omega=0j, ham_sign=1.0,
                precision="c128", layout="reference"):
    h, c = as_terms2(hop), as_terms4(coul)
    return _build_in_layout(layout, "edx_build_ham_n", basis_n._h, num_val_orbs, num_core_orbs,
                            _ptr(h), len(h), _ptr(c), len(c), float(np.real(omega)),
                            float(np.imag(omega)), float(ham_sign), _prec(precision))


def build_transop_i(basis_n, basis_i, num_val_orbs, num_core_orbs, trans, precision="c128",
                    layout="reference"):
    t = as_terms2(trans)
    return _build_in_layout(layout, "edx_build_transop_i", basis_n._h, basis_i._h, num_val_orbs,
                            num_core_orbs, _ptr(t), len(t), _prec(precision))


def build_transop_f(basis_f, basis_n, num_val_orbs, num_core_orbs, trans, precision="c128",
                    layout="reference"):
    t = as_terms2(trans)
    return _build_in_layout(layout, "edx_build_transop_f", basis_f._h, basis_n._h, num_val_orbs,
                            num_core_orbs, _ptr(t), len(t), _prec(precision))


def lanczos_eigs(H, nev, maxiter=500, tol=1e-10, seed=1, deflate=True, nvec=None,
                 precision="c128"):
    """Lowest nev eigenvalues (numpy) and nvec eigenvectors (device tensor, nvec x n)."""
    T = torch()
    nvec = nev if nvec is None else nvec
    ev = np.zeros(nev)
    vecs = T.zeros((max(nvec, 1), H.n), dtype=T.complex128, device=device())
    it = ctypes.c_int()
    _lib.call("edx_lanczos_eigs", H._h, _prec(precision), nev, maxiter, float(tol),
              int(seed), int(bool(deflate)), nvec, _ptr(ev), _dptr(vecs), ctypes.byref(it))
    return ev, vecs[:nvec], it.value


def lanczos_ground(H, nev=1, maxiter=500, tol=1e-10, seed=1, nvec=None, precision="c128"):
    """Lowest nev eigenpairs by the reference's two-pass Lanczos
    (``diag_ham_lanczos`` src/lanczos.f90:6-167): no stored basis, no
    reorthogonalization, Ritz vectors from a second run of the recurrence.
    One copy of a degenerate level.  Returns (evals, evecs (nvec x n device
    tensor), steps)."""
    T = torch()
    nvec = nev if nvec is None else nvec
    ev = np.zeros(nev)
    vecs = T.zeros((max(nvec, 1), H.n), dtype=T.complex128, device=device())
    it = ctypes.c_int()
    _lib.call("edx_lanczos_ground", H._h, _prec(precision), nev, maxiter, float(tol), int(seed),
              nvec, _ptr(ev), _dptr(vecs), ctypes.byref(it))
    return ev, vecs[:nvec], it.value


def trlan_eigs(H, nev, ncv=None, maxiter=500, tol=1e-10, seed=1, deflate=True, nvec=None,
               precision="c128"):
    """Lowest nev eigenpairs by thick-restart Lanczos (the ARPACK path,
    ``diag_ham_arpack`` src/arpack.f90:2-112).  ncv: basis size (at least
    2 nev + 16 is used); maxiter: restarts.  Returns (evals,
    evecs (nvec x n device tensor), products with H, converged values)."""
    T = torch()
    nvec = nev if nvec is None else nvec
    ncv = nev + 2 if ncv is None else ncv
    ev = np.zeros(nev)
    vecs = T.zeros((max(nvec, 1), H.n), dtype=T.complex128, device=device())
    it, nc = ctypes.c_int(), ctypes.c_int()
    _lib.call("edx_trlan_eigs", H._h, _prec(precision), nev, ncv, maxiter, float(tol), int(seed),
              int(bool(deflate)), nvec, _ptr(ev), _dptr(vecs), ctypes.byref(it), ctypes.byref(nc))
    return ev, vecs[:nvec], it.value, nc.value


def lanczos_refine(H, start, nev, maxiter=200, tol=1e-10, nvec=None, precision="c128"):
    """Lanczos started from the sum of the rows of ``start`` (k x n device tensor)."""
    T = torch()
    nvec = nev if nvec is None else nvec
    st = start.contiguous()
    ev = np.zeros(nev)
    vecs = T.zeros((max(nvec, 1), H.n), dtype=T.complex128, device=device())
    it = ctypes.c_int()
    _lib.call("edx_lanczos_refine", H._h, _prec(precision), nev, maxiter, float(tol),
              st.shape[0], _dptr(st), nvec, _ptr(ev), _dptr(vecs), ctypes.byref(it))
    return ev, vecs[:nvec], it.value


def block_lanczos_eigs(H, nev, maxblocks=60, tol=1e-10, seed=1, nvec=None, precision="c128"):
    """Lowest nev eigenpairs by block Lanczos (blocks of 8, reorthogonalization
    on the FP64 matrix cores).  Returns (evals, evecs (nvec x n device tensor),
    block steps, residual estimates)."""
    T = torch()
    nvec = nev if nvec is None else nvec
    ev = np.zeros(nev)
    res = np.zeros(nev)
    vecs = T.zeros((max(nvec, 1), H.n), dtype=T.complex128, device=device())
    nb = ctypes.c_int()
    _lib.call("edx_block_lanczos_eigs", H._h, _prec(precision), nev, maxblocks, float(tol),
              int(seed), nvec, _ptr(ev), _dptr(vecs), ctypes.byref(nb), _ptr(res))
    return ev, vecs[:nvec], nb.value, res


def ground_state(H, nev=1, precision="c128", maxiter=500, tol=1e-10, seed=1, deflate=False,
                 refine_maxiter=200, solver="lanczos", ncv=None):
    """Lowest eigenpairs.  With precision 'c64' or 'c64d' the products round
    values to complex64 and the result is finished by an FP64 Lanczos started
    from those eigenvectors; 'c128' and 'c128i' are exact and need no
    refinement.  solver 'lanczos': one state without deflation takes the
    reference's two-pass Lanczos (no stored basis), several states or deflated
    restarts the fully reorthogonalized solver; 'trlan': thick-restart Lanczos
    over a basis of ncv (default 2 nev + 16) vectors, maxiter restarts, which
    finds several states with the fewest reorthogonalization flops.  Returns
    (evals, evecs, products with H or Lanczos steps)."""
    def lowest(p, t):
        if solver == "trlan":
            ev, v, nmv, _ = trlan_eigs(H, nev, ncv=ncv if ncv else 2 * nev + 16, maxiter=maxiter,
                                       tol=t, seed=seed, deflate=deflate, nvec=nev, precision=p)
            return ev, v, nmv
        return lanczos_eigs(H, nev, maxiter, t, seed, deflate, nev, p)
    if _prec(precision) in (EDX_C128, EDX_C128I):   # exact products: no refinement
        if nev == 1 and not deflate and solver == "lanczos":
            return lanczos_ground(H, 1, maxiter, tol, seed, 1, precision)
        return lowest(precision, tol)
    ev, vecs, it = lowest(precision, max(tol, 1e-6))
    ev2, vecs2, it2 = lanczos_refine(H, vecs, nev, refine_maxiter, tol, nev, "c128")
    return ev2, vecs2, it + it2


def build_krylov(H, v0, maxn, precision="c128"):
    """alpha_1..neff, beta_1..neff (pole-file convention) and norm = <v0|v0>."""
    _check_vec(v0, H.n)
    alpha = np.zeros(maxn)
    beta = np.zeros(maxn)
    neff = ctypes.c_int()
    norm = ctypes.c_double()
    _lib.call("edx_build_krylov", H._h, _prec(precision), _dptr(v0), maxn, ctypes.byref(neff),
              _ptr(alpha), _ptr(beta), ctypes.byref(norm))
    n = neff.value
    return alpha[:n].copy(), beta[:n].copy(), norm.value


def build_krylov_reorth(H, v0, maxn, precision="c128"):
    """Krylov tridiagonalization with full reorthogonalization (exact projection)."""
    _check_vec(v0, H.n)
    alpha = np.zeros(maxn)
    beta = np.zeros(maxn)
    neff = ctypes.c_int()
    norm = ctypes.c_double()
    _lib.call("edx_build_krylov_reorth", H._h, _prec(precision), _dptr(v0), maxn,
              ctypes.byref(neff), _ptr(alpha), _ptr(beta), ctypes.byref(norm))
    n = neff.value
    return alpha[:n].copy(), beta[:n].copy(), norm.value


def build_krylov_batch(H, V0, maxn, precision="c128"):
    """K independent Krylov recurrences in lockstep; V0 is a (K, n) device tensor.

    Returns lists (per RHS) of alpha_1..neff, beta_1..neff and norms."""
    V0 = V0.contiguous()
    K = V0.shape[0]
    assert V0.shape[1] == H.n and V0.dtype == torch().complex128
    alpha = np.zeros((K, maxn))
    beta = np.zeros((K, maxn))
    neff = np.zeros(K, dtype=np.int32)
    norm = np.zeros(K)
    _lib.call("edx_build_krylov_batch", H._h, _prec(precision), K, _dptr(V0), maxn, _ptr(neff),
              _ptr(alpha), _ptr(beta), _ptr(norm))
    return ([alpha[k, :neff[k]].copy() for k in range(K)],
            [beta[k, :neff[k]].copy() for k in range(K)], norm)


class KrylovBatch:
    """Resumable batched Krylov recurrence (edx_krylov_batch_*) over H from the
    rows of V0 (K, n); the device keeps vectors and the (alpha, beta) history."""

    def __init__(self, H, V0, precision="c128", capacity=256):
        V0 = V0.contiguous()
        assert V0.dim() == 2 and V0.shape[1] == H.n and V0.dtype == torch().complex128
        self.K = V0.shape[0]
        self.H = H                      # the state reads H: keep it alive
        self.done = 0
        self._h = None
        h = ctypes.c_void_p()
        _lib.call("edx_krylov_batch_start", H._h, _prec(precision), self.K, _dptr(V0),
                  int(capacity), ctypes.byref(h))
        self._h = h

    def advance(self, nsteps):
        """Up to nsteps more steps; returns the total steps taken (fewer than asked
        once every recurrence has broken down)."""
        d = ctypes.c_int()
        _lib.call("edx_krylov_batch_advance", self._h, int(nsteps), ctypes.byref(d))
        self.done = d.value
        return self.done

    def coefficients(self, count=None):
        """Per RHS: alpha_1..neff, beta_1..neff (pole-file convention) of the first
        ``count`` steps, and the norms <v0|v0>."""
        count = self.done if count is None else int(count)
        K = self.K
        alpha = np.zeros((K, max(count, 1)))
        beta = np.zeros((K, max(count, 1)))
        neff = np.zeros(K, dtype=np.int32)
        norm = np.zeros(K)
        _lib.call("edx_krylov_batch_read", self._h, count, _ptr(neff), _ptr(alpha), _ptr(beta),
                  _ptr(norm))
        return ([alpha[k, :neff[k]].copy() for k in range(K)],
                [beta[k, :neff[k]].copy() for k in range(K)], norm)

    def keep(self, indices):
        """Drop every recurrence but the rows ``indices`` (increasing) of the
        current batch; the kept ones continue unchanged as rows 0..len-1."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1))
        _lib.call("edx_krylov_batch_keep", self._h, len(idx), _ptr(idx))
        self.K = len(idx)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None and _lib._lib is not None:
            _lib._lib.edx_krylov_batch_free(h)
        self._h = None

    def __del__(self):
        self.close()


class MatrixFreeOp:
    """A Hamiltonian applied on the fly (edx_mf_*): the CSR build's terms, signs
    and columns, without storing the matrix.  Keeps its basis alive."""

    def __init__(self, handle, basis):
        self._h = handle
        self._basis = basis
        n = ctypes.c_int64()
        _lib.call("edx_mf_size", handle, ctypes.byref(n))
        self.m = self.n = n.value

    def apply(self, x, y):
        """y += H x for complex128 device tensors."""
        _check_vec(x, self.n)
        _check_vec(y, self.m)
        _lib.call("edx_mf_apply", self._h, _dptr(x), _dptr(y))
        return y

    def matvec(self, x):
        y = torch().zeros(self.m, dtype=torch().complex128, device=x.device)
        return self.apply(x, y)

    def build_krylov(self, v0, maxn):
        """build_krylov_mp over the operator: alpha, beta (pole-file convention), norm."""
        _check_vec(v0, self.n)
        alpha = np.zeros(maxn)
        beta = np.zeros(maxn)
        neff = ctypes.c_int()
        norm = ctypes.c_double()
        _lib.call("edx_build_krylov_mf", self._h, _dptr(v0), int(maxn), ctypes.byref(neff),
                  _ptr(alpha), _ptr(beta), ctypes.byref(norm))
        k = neff.value
        return alpha[:k].copy(), beta[:k].copy(), norm.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None and _lib._lib is not None:
            _lib._lib.edx_mf_free(h)
        self._h = None


def mf_ham_i(basis, hop, coul, omega=0j, ham_sign=1.0):
    h2, h4 = as_terms2(hop), as_terms4(coul)
    h = ctypes.c_void_p()
    _lib.call("edx_mf_ham_i", basis._h, _ptr(h2), len(h2), _ptr(h4), len(h4),
              float(np.real(omega)), float(np.imag(omega)), float(ham_sign), ctypes.byref(h))
    return MatrixFreeOp(h, basis)


def mf_ham_n(basis_n, num_val_orbs, num_core_orbs, hop, coul, omega=0j, ham_sign=1.0):
    h2, h4 = as_terms2(hop), as_terms4(coul)
    h = ctypes.c_void_p()
    _lib.call("edx_mf_ham_n", basis_n._h, int(num_val_orbs), int(num_core_orbs), _ptr(h2),
              len(h2), _ptr(h4), len(h4), float(np.real(omega)), float(np.imag(omega)),
              float(ham_sign), ctypes.byref(h))
    return MatrixFreeOp(h, basis_n)


class KrylovGroup:
    """Independent single-vector Krylov recurrences over different matrices,
    advanced together (edx_group_*): one SpMV launch per step serves every
    member.  Members are added and removed between advances; a member's
    coefficients do not depend on the other members."""

    def __init__(self, precision="c64d", max_steps=6000):
        self.max_steps = int(max_steps)
        self._h = None
        self._mats = {}                 # member -> matrix (kept alive while a member)
        h = ctypes.c_void_p()
        _lib.call("edx_group_create", _prec(precision), self.max_steps, ctypes.byref(h))
        self._h = h

    def add(self, H, v0):
        _check_vec(v0, H.n)
        mid = ctypes.c_int()
        _lib.call("edx_group_add", self._h, H._h, _dptr(v0), ctypes.byref(mid))
        self._mats[mid.value] = H
        return mid.value

    def remove(self, member):
        _lib.call("edx_group_remove", self._h, int(member))
        self._mats.pop(member, None)

    def advance(self, nsteps):
        """nsteps steps of every running member; returns how many still run."""
        live = ctypes.c_int()
        _lib.call("edx_group_advance", self._h, int(nsteps), ctypes.byref(live))
        return live.value

    def coefficients(self, member):
        """(alpha, beta, norm, running) of a member so far."""
        steps = ctypes.c_int()
        norm = ctypes.c_double()
        run = ctypes.c_int()
        _lib.call("edx_group_read", self._h, int(member), ctypes.byref(steps), None, None,
                  ctypes.byref(norm), ctypes.byref(run))
        a = np.zeros(max(steps.value, 1))
        b = np.zeros(max(steps.value, 1))
        _lib.call("edx_group_read", self._h, int(member), ctypes.byref(steps), _ptr(a), _ptr(b),
                  ctypes.byref(norm), ctypes.byref(run))
        n = steps.value
        return a[:n].copy(), b[:n].copy(), norm.value, bool(run.value)

    @property
    def members(self):
        return list(self._mats)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None and _lib._lib is not None:
            _lib._lib.edx_group_free(h)
        self._h = None
        self._mats = {}

    def __del__(self):
        self.close()


# A block of K interleaved vectors larger than this runs as K single-vector
# recurrences instead (each vector then stays in the 256 MB last-level cache).
SPLIT_BLOCK_BYTES = 192 << 20


def converged_spectra(H, V0, e_gs, grid, gamma, precision="c128", tol=1e-8, chunk=250,
                      min_steps=250, max_steps=4000, compact=True):
    """Continued-fraction spectra of the K start vectors (rows of V0), each taken at
    the first multiple of ``chunk`` steps (>= min_steps) where its relative L2
    change since the previous multiple is <= tol, or where its recurrence broke
    down.  Returns (spectra, steps, last_change); a spectrum that never met tol
    by max_steps keeps its last value and reports its last change.  Fewer than
    8 vectors whose interleaved block exceeds SPLIT_BLOCK_BYTES advance as
    separate single-vector recurrences, except 2 to 4 vectors over a matrix
    with the slot-factored form in the indexed precision, whose block product
    gathers contiguous segments of all vectors.  With ``compact`` a block
    drops the recurrences whose spectra are final (KrylovBatch.keep), so the
    later products cover only the vectors still running."""
    V0 = V0.contiguous()
    K = V0.shape[0]
    # a block of up to 4 vectors over a matrix with the slot-factored form stays
    # one block in the indexed precision (its products gather contiguous
    # segments of all K vectors); otherwise blocks of fewer than 8 vectors too
    # large for the last-level cache run as single-vector recurrences
    fac_block = 2 <= K <= 4 and _prec(precision) == EDX_C128I and H.factored() is not None
    if K < 8 and K * H.n * 16 > SPLIT_BLOCK_BYTES and not fac_block:
        groups = [[k] for k in range(K)]
    else:
        groups = [list(range(K))]
    cap = max(min_steps, chunk) + chunk
    kbs = [KrylovBatch(H, V0[g].contiguous() if len(g) < K else V0, precision, capacity=cap)
           for g in groups]
    prev = [None] * K
    out = [None] * K
    steps = [0] * K
    change = [np.inf] * K
    m = 0
    try:
        while True:
            target = min(max(m + chunk, min_steps), max_steps)
            for gi, kb in enumerate(kbs):
                g = groups[gi]
                if all(out[k] is not None for k in g):
                    continue
                got = kb.advance(target - kb.done)
                al, be, nr = kb.coefficients(got)
                for j, k in enumerate(g):
                    if out[k] is not None:
                        continue
                    s = cf_spectrum(al[j], be[j], nr[j], e_gs, grid, gamma)
                    ended = len(al[j]) < got or got < target
                    if prev[k] is not None:
                        den = np.linalg.norm(s)
                        change[k] = float(np.linalg.norm(s - prev[k]) / den) if den > 0 else 0.0
                    if ended or change[k] <= tol or target >= max_steps:
                        out[k] = s
                        steps[k] = len(al[j])
                        if ended:
                            change[k] = 0.0
                    prev[k] = s
                live = [j for j, k in enumerate(g) if out[k] is None]
                if compact and live and len(live) < len(g):
                    kb.keep(live)
                    groups[gi] = [g[j] for j in live]
            m = target
            if all(o is not None for o in out):
                break
    finally:
        for kb in kbs:
            kb.close()
    return out, steps, change


def minres_shifted_batch(H, sigma, B, tol=1e-10, maxit=500, precision="c128"):
    """Solve (sigma_k - H) x_k = b_k for the rows b_k of B (K, n) in one recurrence."""
    T = torch()
    B = B.contiguous()
    K = B.shape[0]
    sig = np.ascontiguousarray(np.asarray(sigma, dtype=np.complex128).reshape(K))
    X = T.zeros_like(B)
    it = np.zeros(K, dtype=np.int32)
    res = np.zeros(K)
    _lib.call("edx_minres_shifted_batch", H._h, _prec(precision), K, _ptr(sig.view(np.float64)),
              _dptr(B), _dptr(X), float(tol), int(maxit), _ptr(it), _ptr(res))
    return X, it, res


def shifted_fom(H, b, sigma, projections=(None,), tol=1e-8, maxit=2000, precision="c128"):
    """Solve (sigma_k - H) x_k = b for every shift from one Lanczos recurrence.

    ``projections`` lists CSR operators P (``None`` = identity); returns
    ``(outs, iters, res)`` where ``outs[p]`` is a (K, rows(P)) complex128 device
    tensor holding P x_k, and ``res`` the relative residual of each iterate."""
    T = torch()
    _check_vec(b, H.n)
    sig = np.ascontiguousarray(np.atleast_1d(np.asarray(sigma, dtype=np.complex128)))
    K = len(sig)
    projs = list(projections)
    outs = [T.zeros((K, P.m if P is not None else H.n), dtype=T.complex128, device=b.device)
            for P in projs]
    hp = (ctypes.c_void_p * len(projs))(*[P._h if P is not None else None for P in projs])
    op = (ctypes.c_void_p * len(projs))(*[o.data_ptr() for o in outs])
    it = np.zeros(K, dtype=np.int32)
    res = np.zeros(K)
    _lib.call("edx_shifted_fom_project", H._h, _prec(precision), _dptr(b), K,
              _ptr(sig.view(np.float64)), len(projs), hp, op, float(tol), int(maxit),
              _ptr(it), _ptr(res))
    return outs, it, res


def shifted_fom_batch(H, B, sigma, projections, tol=1e-8, maxit=2000, precision="c128"):
    """shifted_fom for the rows b_r of B (R, n) in one interleaved recurrence, the
    same shifts for every row: ``projections[r]`` (``None`` = identity) projects
    row r.  Returns ``(outs, iters, res)``: outs[r] is (K, rows(P_r)), iters and
    res are (R, K)."""
    T = torch()
    B = B.contiguous()
    R = B.shape[0]
    assert B.dim() == 2 and B.shape[1] == H.n and B.dtype == T.complex128
    assert len(projections) == R
    sig = np.ascontiguousarray(np.atleast_1d(np.asarray(sigma, dtype=np.complex128)))
    K = len(sig)
    outs = [T.zeros((K, P.m if P is not None else H.n), dtype=T.complex128, device=B.device)
            for P in projections]
    hp = (ctypes.c_void_p * R)(*[P._h if P is not None else None for P in projections])
    op = (ctypes.c_void_p * R)(*[o.data_ptr() for o in outs])
    it = np.zeros((R, K), dtype=np.int32)
    res = np.zeros((R, K))
    _lib.call("edx_shifted_fom_project_batch", H._h, _prec(precision), R, _dptr(B), K,
              _ptr(sig.view(np.float64)), hp, op, float(tol), int(maxit), _ptr(it), _ptr(res))
    return outs, it, res


def shifted_fom_rhs(H, B, sigma, projections, tol=1e-8, maxit=2000, precision="c128"):
    """shifted_fom for each row b_r of B (R, n) with its projection, in the
    results layout of shifted_fom_batch: one interleaved recurrence where the
    block product pays (the rule of converged_spectra: 2 to 4 rows over a
    matrix with the slot-factored form in the indexed precision, 8 or more
    rows, or a block within SPLIT_BLOCK_BYTES), else one recurrence per row
    (a larger interleaved block of a few rows overflows the last-level cache
    and its row-wise product loses to the single-vector ones)."""
    T = torch()
    B = B.contiguous()
    R = B.shape[0]
    fac_block = 2 <= R <= 4 and _prec(precision) == EDX_C128I and H.factored() is not None
    if R > 1 and (fac_block or R >= 8 or R * H.n * 16 <= SPLIT_BLOCK_BYTES):
        return shifted_fom_batch(H, B, sigma, projections, tol=tol, maxit=maxit,
                                 precision=precision)
    outs, its, rss = [], [], []
    for r in range(R):
        (Y,), it, res = shifted_fom(H, B[r], sigma, projections=(projections[r],), tol=tol,
                                    maxit=maxit, precision=precision)
        outs.append(Y)
        its.append(it)
        rss.append(res)
    return outs, np.stack(its), np.stack(rss)


def minres(A, b, tol=1e-10, maxit=500, precision="c128"):
    T = torch()
    _check_vec(b, A.n)
    x = T.zeros_like(b)
    it = ctypes.c_int()
    res = ctypes.c_double()
    _lib.call("edx_minres", A._h, _prec(precision), _dptr(b), _dptr(x), float(tol), int(maxit),
              ctypes.byref(it), ctypes.byref(res))
    return x, it.value, res.value


def denmat(basis, num_val_orbs, vec):
    _check_vec(vec, basis.n)
    dm = np.zeros((num_val_orbs, num_val_orbs), dtype=np.complex128)
    _lib.call("edx_denmat", basis._h, num_val_orbs, _dptr(vec), _ptr(dm))
    return dm


def cf_spectrum(alpha, beta, norm, e_gs, omega, gamma, weight=1.0, out=None):
    """Continued fraction of one pole set on the device, added into ``out``."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    ga = np.ascontiguousarray(np.broadcast_to(gamma, om.shape), dtype=np.float64)
    spec = np.zeros(len(om)) if out is None else out
    _lib.call("edx_cf_spectrum", len(a), _ptr(a), _ptr(b), float(norm), float(e_gs),
              float(weight), len(om), _ptr(om), _ptr(ga), _ptr(spec))
    return spec


def mem_in_use():
    v = ctypes.c_int64()
    _lib.call("edx_mem_in_use", ctypes.byref(v))
    return v.value


def set_stream(stream=None):
    """Queue this thread's library work on a torch stream (None: the legacy
    default stream).  Pair it with ``torch.cuda.stream(stream)`` so torch's
    work in the thread is ordered with the library's."""
    _lib.call("edx_set_stream", ctypes.c_void_p(stream.cuda_stream if stream is not None else 0))


def profile(enable=True):
    """Start (clearing) or stop recording SpMV/SpMM launch timings."""
    _lib.call("edx_profile", int(bool(enable)))


KERNELS = {"k_spmv": 0, "k_spmm_il": 1, "k_spmm_rl": 2, "k_spmm": 3, "k_mf": 4, "k_spmv_fac": 7,
           "k_spmm_fac": 8}
BLOCK_KERNELS = {"k_bproj": 5, "k_bupd": 6}   # matrix-core kernels (recorded as c128)


def profile_read(precision="c128", kernel=None, nrhs=None):
    """(launches, kernel milliseconds, algorithmic bytes) recorded for one
    precision and one SpMV/SpMM kernel (a name of KERNELS; None = all), and
    with ``nrhs`` the launches with that many right-hand sides only (above 16:
    every launch of more than 16)."""
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    by = ctypes.c_double()
    k = -1 if kernel is None else {**KERNELS, **BLOCK_KERNELS}[kernel]
    _lib.call("edx_profile_read_rhs", _prec(precision), k, -1 if nrhs is None else int(nrhs),
              ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by))
    return n.value, ms.value, by.value


def profile_read_alg(precision="c128", kernel=None, nrhs=None):
    """Summed algorithmic bytes of the same launches counted as the complex128
    CSR of the same operator (20 bytes per nonzero, row pointers, per vector x
    once and y written, and read where the epilogue accumulates), whatever
    form they read."""
    b = ctypes.c_double()
    k = -1 if kernel is None else {**KERNELS, **BLOCK_KERNELS}[kernel]
    _lib.call("edx_profile_read_alg", _prec(precision), k, -1 if nrhs is None else int(nrhs),
              ctypes.byref(b))
    return b.value


def profile_read_flops(kernel, precision="c128"):
    """Summed algorithmic flops of the recorded launches of a matrix-core kernel."""
    f = ctypes.c_double()
    _lib.call("edx_profile_read_flops", _prec(precision), BLOCK_KERNELS[kernel], ctypes.byref(f))
    return f.value
