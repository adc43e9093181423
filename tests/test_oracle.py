"""Pins the CPU oracle (oracle/edx_oracle.c) before anything is compared with it.

Each check is chosen so that a plausible slip in the restatement (dropped term,
wrong sign or index, swapped operator order, wrong accumulation order) fails:
brute-force Jordan-Wigner matrices, closed forms, textbook routines (scipy),
and the known answers in the reference's docstrings.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as o
from tests import helpers as H


# --------------------------------------------------------------- basis / signs

def test_fock_known_answers():
    # edrixs/fock_basis.py:161-162 and the 15-entry listing at 421-437 (C(6,2))
    assert list(H.fock_words(4, 2)) == [3, 5, 6, 9, 10, 12]
    assert list(H.fock_words(6, 2)) == [3, 5, 6, 9, 10, 12, 17, 18, 20, 24, 33, 34, 36, 40, 48]
    assert list(o.fock_sorted(6, 2)) == list(H.fock_words(6, 2))


def test_binary_search_matches_index():
    f = H.fock_words(8, 3)
    for k, w in enumerate(f):
        assert o.binary_search(f, w) == k
    assert o.binary_search(f, 1) == -1          # wrong popcount
    assert o.binary_search(f, 1 << 9 | 3) == -1  # out of range


def test_make_newfock_against_jordan_wigner():
    n = 6
    cs = o.jw_annihilators(n)
    for w in range(2 ** n):
        for p in range(n):
            e = np.zeros(2 ** n)
            e[w] = 1.0
            if (w >> p) & 1:
                out = cs[p] @ e
                new, sgn = o.make_newfock('-', p + 1, w)
            else:
                out = cs[p].T @ e
                new, sgn = o.make_newfock('+', p + 1, w)
            k = int(np.nonzero(out)[0][0])
            assert k == new and out[k] == sgn


def test_tridiag_eigs_against_scipy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 60):
        d = rng.normal(size=n)
        e = rng.normal(size=n - 1)
        w, z = o.tridiag_eigs(d, e, vectors=True)
        wr = sla.eigh_tridiagonal(d, e, eigvals_only=True)
        np.testing.assert_allclose(w, wr, atol=1e-12)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        for j in range(n):
            np.testing.assert_allclose(T @ z[j], w[j] * z[j], atol=1e-10)


# ------------------------------------------------------------------ CSR builds

def _check_csr_shape(A):
    assert A.rowptr[0] == 0 and np.all(np.diff(A.rowptr) >= 0)
    for r in range(A.m):
        c = A.cols[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0), "columns must be strictly ascending"


def test_build_ham_i_equals_dense_operator():
    rng = np.random.default_rng(7)
    n, N = 7, 3
    emat = H.random_hermitian_emat(n, rng)
    umat = H.random_hermitian_umat(n, rng, nnz=80)
    fock = H.fock_words(n, N)
    from edrixs_amd import siam
    A = o.build_ham_i(fock, siam.emat_terms(emat), siam.umat_terms(umat))
    _check_csr_shape(A)
    # dense from the text-rounded terms to compare exactly-built matrices
    Ht = o.dense_operator(n, _terms_to_emat(siam.emat_terms(emat), n),
                          _terms_to_umat(siam.umat_terms(umat), n))
    np.testing.assert_allclose(A.todense(), Ht[np.ix_(fock, fock)], atol=1e-12)
    assert np.allclose(A.todense(), A.todense().conj().T, atol=1e-12)


def _terms_to_emat(t, n):
    e = np.zeros((n, n), np.complex128)
    for r in t:
        e[r["ind1"] - 1, r["ind2"] - 1] += r["re"] + 1j * r["im"]
    return e


def _terms_to_umat(t, n):
    u = np.zeros((n,) * 4, np.complex128)
    for r in t:
        u[r["ind1"] - 1, r["ind2"] - 1, r["ind3"] - 1, r["ind4"] - 1] += r["re"] + 1j * r["im"]
    return u


def test_duplicates_sum_in_term_order():
    # three hopping terms on the same diagonal entry: (1e16 + 1) + (-1e16) = 0 in
    # term order, whereas any other order gives 1 -- pins src/m_types.f90:91-94.
    fock = H.fock_words(3, 1)
    hop = [[1, 1, 1e16, 0.0], [1, 1, 1.0, 0.0], [1, 1, -1e16, 0.0]]
    A = o.build_ham_i(fock, hop, None)
    assert A.vals[0] == 0.0
    # omega goes in last on the diagonal with ham_sign < 0 (src/ham.f90:156-176)
    hop2 = [[1, 1, 1.0, 0.0], [2, 1, 5.0, 0.0]]
    B = o.build_ham_i(fock, hop2, None, omega=3.0 + 0.5j, ham_sign=-1.0)
    d = B.todense()
    assert d[0, 0] == -1.0 + 3.0 + 0.5j and d[1, 1] == 3.0 + 0.5j and d[1, 0] == -5.0
    # retained exact zeros: a term pair cancelling to 0 stays stored
    hop3 = [[1, 2, 2.0, 0.0], [1, 2, -2.0, 0.0]]
    C = o.build_ham_i(fock, hop3, None)
    assert C.nnz == 1 and C.vals[0] == 0.0


def test_build_ham_n_and_transops_equal_dense():
    rng = np.random.default_rng(11)
    nv, nc, N = 5, 2, 2
    ntot = nv + nc
    emat = H.random_hermitian_emat(ntot, rng)
    umat = H.random_hermitian_umat(ntot, rng, nnz=120)
    from edrixs_amd import siam
    ht, ut = siam.emat_terms(emat), siam.umat_terms(umat)
    Hd = o.dense_operator(ntot, _terms_to_emat(ht, ntot), _terms_to_umat(ut, ntot))
    fock_i = H.fock_words(nv, N)
    fock_n = H.fock_words(nv, N + 1)
    rows_n = H.corehole_words(fock_n, nv, nc)
    A = o.build_ham_n(fock_n, nv, nc, ht, ut)
    _check_csr_shape(A)
    np.testing.assert_allclose(A.todense(), Hd[np.ix_(rows_n, rows_n)], atol=1e-12)
    # shifted form of the RIXS driver: omega - H_n (src/rixs_driver.f90:130-133)
    om = 1.25 + 0.3j
    S = o.build_ham_n(fock_n, nv, nc, ht, ut, omega=om, ham_sign=-1.0)
    np.testing.assert_allclose(S.todense(), om * np.eye(len(rows_n)) - Hd[np.ix_(rows_n, rows_n)],
                               atol=1e-12)
    # absorption operator: valence row index, core column index (edrixs/solvers.py:2414)
    T = np.zeros((ntot, ntot), np.complex128)
    T[0:nv, nv:ntot] = rng.normal(size=(nv, nc)) + 1j * rng.normal(size=(nv, nc))
    tt = siam.emat_terms(T)
    Td = o.dense_operator(ntot, _terms_to_emat(tt, ntot))
    rows_i = H.fullcore_words(fock_i, nv, nc)
    D = o.build_transop_i(fock_n, fock_i, nv, nc, tt)
    _check_csr_shape(D)
    np.testing.assert_allclose(D.todense(), Td[np.ix_(rows_n, rows_i)], atol=1e-12)
    # emission operator conj(T)^T: rows = final (full core), columns = intermediate
    tf = siam.emat_terms(np.conj(T).T)
    Tfd = o.dense_operator(ntot, _terms_to_emat(tf, ntot))
    F = o.build_transop_f(fock_i, fock_n, nv, nc, tf)
    _check_csr_shape(F)
    np.testing.assert_allclose(F.todense(), Tfd[np.ix_(rows_i, rows_n)], atol=1e-12)
    np.testing.assert_allclose(F.todense(), D.todense().conj().T, atol=1e-12)


def test_row_ranges_concatenate():
    rng = np.random.default_rng(5)
    n, N = 8, 4
    from edrixs_amd import siam
    ht = siam.emat_terms(H.random_hermitian_emat(n, rng))
    ut = siam.umat_terms(H.random_hermitian_umat(n, rng, nnz=40))
    fock = H.fock_words(n, N)
    A = o.build_ham_i(fock, ht, ut)
    B1 = o.build_ham_i(fock, ht, ut, row_begin=0, row_end=30)
    B2 = o.build_ham_i(fock, ht, ut, row_begin=30, row_end=-1)
    assert np.array_equal(np.concatenate([B1.cols, B2.cols]), A.cols)
    assert np.array_equal(np.concatenate([B1.vals, B2.vals]), A.vals)


def test_thread_count_changes_nothing():
    # the builds and products split rows over OpenMP threads; one thread and
    # several give the same arrays bit for bit (H_n of 38,610 rows: 18 chunks)
    import ctypes
    import ctypes.util
    gomp = ctypes.CDLL(ctypes.util.find_library("gomp") or "libgomp.so.1")
    m = H.small_siam(nbath=5, seed=2)
    fn = H.fock_words(m.norb, m.nocc + 1)
    x = np.random.default_rng(3).normal(size=len(fn) * m.ncore) + 0.5j
    orig = gomp.omp_get_max_threads()
    out = []
    try:
        for nt in (1, 4):
            gomp.omp_set_num_threads(nt)
            assert o.lib().orc_num_threads() == nt
            A = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
            out.append((A, o.matvec(A, x)))
    finally:
        gomp.omp_set_num_threads(orig)
    (A1, y1), (A4, y4) = out
    assert A1.m == 38610
    assert np.array_equal(A1.rowptr, A4.rowptr) and np.array_equal(A1.cols, A4.cols)
    assert np.array_equal(A1.vals.view(np.float64), A4.vals.view(np.float64))
    assert np.array_equal(y1.view(np.float64), y4.view(np.float64))


# ----------------------------------------------------------------- solvers

def _small_problem(seed=2, n=8, N=4):
    rng = np.random.default_rng(seed)
    from edrixs_amd import siam
    ht = siam.emat_terms(H.random_hermitian_emat(n, rng))
    ut = siam.umat_terms(H.random_hermitian_umat(n, rng, nnz=60))
    fock = H.fock_words(n, N)
    return o.build_ham_i(fock, ht, ut), fock, rng


def test_matvec_accumulates():
    A, _, rng = _small_problem()
    x = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    y0 = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    np.testing.assert_allclose(o.matvec(A, x, y0), y0 + A.todense() @ x, atol=1e-12)


def test_krylov_continued_fraction_equals_pole_sum():
    A, _, rng = _small_problem()
    Hd = A.todense()
    phi = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    # The reference recurrence has no reorthogonalization: once orthogonality is
    # lost it needs more than dim steps before the fraction equals the pole sum.
    a, b, norm = o.build_krylov(A, phi, 2 * A.m)
    assert abs(norm - np.vdot(phi, phi).real) < 1e-10
    om = np.linspace(-20, 20, 301)
    gam = 0.4
    egs = -1.5
    spec = o.cf_spectrum(a, b, norm, egs, om, gam)
    e, v = np.linalg.eigh(Hd)
    w = np.abs(v.conj().T @ phi) ** 2
    z = om + 1j * gam + egs
    exact = -1 / np.pi * np.imag((w[None, :] / (z[:, None] - e[None, :])).sum(1))
    np.testing.assert_allclose(spec, exact, rtol=1e-7, atol=1e-9)


def test_lanczos_eigenvalues_and_vectors():
    A, _, rng = _small_problem(seed=9)
    Hd = A.todense()
    e = np.linalg.eigvalsh(Hd)
    v0 = rng.normal(size=A.m) + 0j
    ev, vecs, nk = o.lanczos(A, v0, 2, 60, 1e-8, 2)
    assert 0 < nk <= 60
    np.testing.assert_allclose(ev, e[:2], atol=1e-9)
    for j in range(2):
        r = Hd @ vecs[j] - ev[j] * vecs[j]
        assert np.linalg.norm(r) < 1e-4
    # The reference recurrence keeps no orthogonality: run long enough and the
    # second Ritz value becomes a ghost copy of the first (edrixs/solvers.py:1928-1929),
    # and without convergence by max_niter the eigenvectors come back zero.
    ev2, vecs2, nk2 = o.lanczos(A, v0, 2, 100, 1e-14, 2)
    assert abs(ev2[0] - e[0]) < 1e-9 and abs(ev2[1] - e[0]) < 1e-9
    ev3, vecs3, nk3 = o.lanczos(A, v0, 2, 60, 1e-14, 2)
    assert nk3 == 0 and not np.any(vecs3)


def test_pminres_solves_shifted_system():
    A, fock, rng = _small_problem(seed=4)
    from edrixs_amd import siam
    Hd = A.todense()
    om = 0.7 + 0.5j
    S = Hd * 0
    np.fill_diagonal(S, om)
    S = S - Hd
    # the same matrix as built by the RIXS driver path
    b = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    Sc = o.CSR(*_dense_to_csr(S), A.m)
    x, it, res = o.pminres(Sc, b, 1e-10, 500)
    np.testing.assert_allclose(S @ x, b, atol=1e-8)
    assert res < 1e-8


def _dense_to_csr(D):
    rowptr = [0]
    cols, vals = [], []
    for r in range(D.shape[0]):
        nz = np.nonzero(D[r])[0]
        cols.extend(nz)
        vals.extend(D[r, nz])
        rowptr.append(len(cols))
    return (np.array(rowptr, np.int64), np.array(cols, np.int32), np.array(vals, np.complex128))


def test_denmat_equals_dense_expectation():
    n, N = 7, 3
    A, fock, rng = _small_problem(seed=6, n=n, N=N)
    v = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    v /= np.linalg.norm(v)
    dm = o.denmat(fock, n, v)
    cs = o.jw_annihilators(n)
    full = np.zeros(2 ** n, np.complex128)
    full[fock] = v
    for i in range(n):
        for j in range(n):
            ref = np.vdot(full, cs[i].T @ cs[j] @ full)   # <c+_i c_j>
            assert abs(dm[i, j] - ref) < 1e-12


def test_cf_single_pole_closed_form():
    # one level: I = -1/pi Im[norm/(w + i g + E - a)]
    om = np.linspace(-3, 3, 11)
    s = o.cf_spectrum([0.4], [0.0], 2.0, 0.1, om, 0.3)
    np.testing.assert_allclose(s, -1 / np.pi * np.imag(2.0 / (om + 0.3j + 0.1 - 0.4)), rtol=1e-14)


def test_opavg_equals_dense_expectation():
    # <v|O|v> for a non-Hermitian O (random one- and two-body terms without
    # the Hermitian partners) and a complex v: conjugating the wrong vector or
    # applying O^T instead of O changes the value
    rng = np.random.default_rng(12)
    n, N = 7, 3
    from edrixs_amd import siam
    emat = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * (rng.random((n, n)) < 0.5)
    umat = np.zeros((n,) * 4, np.complex128)
    for _ in range(40):
        i, j, k, l = rng.integers(0, n, size=4)
        if i != j and k != l:
            umat[i, j, k, l] += rng.normal() + 1j * rng.normal()
    et, ut = siam.emat_terms(emat), siam.umat_terms(umat)
    fock = H.fock_words(n, N)
    A = o.build_ham_i(fock, et, ut)
    Od = o.dense_operator(n, _terms_to_emat(et, n), _terms_to_umat(ut, n))[np.ix_(fock, fock)]
    assert not np.allclose(Od, Od.conj().T)
    v = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    ref = np.vdot(v, Od @ v)
    got = o.opavg(A, v)
    assert abs(got - ref) < 1e-12 * abs(ref) and abs(got.imag) > 1e-3
    assert abs(got - np.vdot(v, Od.T @ v)) > 1e-6 and abs(got - np.vdot(v.conj(), Od @ v)) > 1e-6
    # the file lines: (1i5, 4f20.10) with three values per record
    lines = o.opavg_lines([-1.5, 2.25], [got, 1.0 + 0.0j])
    assert lines[0] == "%5d%20.10f%20.10f%20.10f" % (1, -1.5, got.real, got.imag)
    assert len(lines[1]) == 65 and lines[1].startswith("    2")
