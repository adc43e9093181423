"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer work (basis words, row pointers, columns) and the complex128 CSR
values must be bit-exact (values compared with ==, which identifies +0/-0).
Floating-point solver results are held to the tolerances written next to
each assertion.
"""
import os

import numpy as np
import pytest

from oracle import oracle as o
from tests import helpers as H

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def core():
    from edrixs_amd import core as c
    import torch
    assert torch.cuda.is_available()
    c.torch().zeros(1, device="cuda")
    return c


def _assert_csr_equal(A_gpu, A_orc):
    rp, cols, vals = A_gpu.download()
    assert A_gpu.m == A_orc.m and A_gpu.n == A_orc.ncols
    assert np.array_equal(rp, A_orc.rowptr), "row pointers differ"
    assert np.array_equal(cols, A_orc.cols), "columns differ"
    assert np.array_equal(vals.real, A_orc.vals.real) and \
        np.array_equal(vals.imag, A_orc.vals.imag), "values differ"


def _to_dev(core, x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to("cuda")


# ----------------------------------------------------------------- basis

def test_basis_enumeration_and_rank(core):
    for nbits, nocc in ((1, 0), (1, 1), (6, 2), (12, 6), (20, 10), (24, 3)):
        b = core.Basis.combinations(nbits, nocc)
        ref = H.fock_words(nbits, nocc)
        assert b.complete and np.array_equal(b.words(), ref)
        idx = b.lookup(ref)
        assert np.array_equal(idx, np.arange(len(ref)))
    b = core.Basis.combinations(10, 4)
    bad = np.array([0, 1, 1 << 10 | 7, 0b1111, 0b11111, (1 << 63) - 1], dtype=np.int64)
    assert list(b.lookup(bad)) == [-1, -1, -1, 0, -1, -1]


def test_basis_from_words_modes(core):
    full = H.fock_words(10, 5)
    b = core.Basis.from_words(full, 10)
    assert b.complete
    sub = full[::3].copy()            # not the complete set: binary-search mode
    s = core.Basis.from_words(sub, 10)
    assert not s.complete
    probe = np.concatenate([sub, full[1::3], [0, 3]])
    ref = np.array([o.binary_search(sub, w) for w in probe])
    assert np.array_equal(s.lookup(probe), ref)
    with pytest.raises(Exception):
        core.Basis.from_words(sub[::-1].copy(), 10)


# ------------------------------------------------------------- CSR builds

def _rand_terms(n, seed, nnz_u=150):
    from edrixs_amd import siam
    rng = np.random.default_rng(seed)
    return (siam.emat_terms(H.random_hermitian_emat(n, rng)),
            siam.umat_terms(H.random_hermitian_umat(n, rng, nnz=nnz_u)))


@pytest.mark.parametrize("complete", [True, False])
def test_build_ham_i_bit_exact(core, complete):
    n, N = 10, 5
    ht, ut = _rand_terms(n, 1)
    fock = H.fock_words(n, N)
    if not complete:
        fock = fock[np.random.default_rng(0).random(len(fock)) < 0.7]
    b = core.Basis.from_words(fock, n)
    assert b.complete == complete
    for omega, sign in ((0j, 1.0), (2.5 + 0.3j, -1.0)):
        A = core.build_ham_i(b, ht, ut, omega=omega, ham_sign=sign)
        _assert_csr_equal(A, o.build_ham_i(fock, ht, ut, omega=omega, ham_sign=sign))


def test_build_ham_i_duplicate_order_and_zeros(core):
    fock = H.fock_words(3, 1)
    b = core.Basis.from_words(fock, 3)
    hop = [[1, 1, 1e16, 0.0], [1, 1, 1.0, 0.0], [1, 1, -1e16, 0.0], [1, 2, 2.0, 0.5],
           [1, 2, -2.0, -0.5]]
    A = core.build_ham_i(b, hop, None)
    _assert_csr_equal(A, o.build_ham_i(fock, hop, None))
    rp, c, v = A.download()
    # row 0: (1e16 + 1) - 1e16 == 0 on the diagonal, and the cancelled pair (2 - 2)
    # kept as an explicit zero in column 1; rows 1, 2 have no surviving terms
    assert A.nnz == 2 and list(rp) == [0, 2, 2, 2] and list(c) == [0, 1]
    assert v[0] == 0.0 and v[1] == 0.0


def test_build_siam_operators_bit_exact(core):
    m = H.small_siam(nbath=4, seed=7)          # 14 valence + 6 core spin-orbitals, N=7
    nv, nc, N = m.norb, m.ncore, m.nocc
    fi, fn = H.fock_words(nv, N), H.fock_words(nv, N + 1)
    bi, bn = core.Basis.from_words(fi, nv), core.Basis.from_words(fn, nv)
    _assert_csr_equal(core.build_ham_i(bi, m.hop_i, m.coul_i), o.build_ham_i(fi, m.hop_i, m.coul_i))
    Hn = core.build_ham_n(bn, nv, nc, m.hop_n, m.coul_n)
    _assert_csr_equal(Hn, o.build_ham_n(fn, nv, nc, m.hop_n, m.coul_n))
    om = -3.25 + 0.5j
    _assert_csr_equal(core.build_ham_n(bn, nv, nc, m.hop_n, m.coul_n, omega=om, ham_sign=-1.0),
                      o.build_ham_n(fn, nv, nc, m.hop_n, m.coul_n, omega=om, ham_sign=-1.0))
    for k in range(3):
        _assert_csr_equal(core.build_transop_i(bn, bi, nv, nc, m.trans_i[k]),
                          o.build_transop_i(fn, fi, nv, nc, m.trans_i[k]))
        _assert_csr_equal(core.build_transop_f(bi, bn, nv, nc, m.trans_f[k]),
                          o.build_transop_f(fi, fn, nv, nc, m.trans_f[k]))


def test_build_24orb_fixture_bit_exact(core):
    from edrixs_amd import iostream
    hop = iostream.read_terms(os.path.join(GOLD, "ed24orb_hopping_i.in"), 2)
    coul = iostream.read_terms(os.path.join(GOLD, "ed24orb_coulomb_i.in"), 4)
    for N in (3, 12):
        fock = H.fock_words(24, N) if N < 12 else core.Basis.combinations(24, 12).words()
        b = core.Basis.combinations(24, N)
        A = core.build_ham_i(b, hop, coul)
        rp, cols, vals = A.download()
        # oracle on windows of rows (full 2.7M-row oracle build is minutes)
        for lo, hi in ((0, 3000), (len(fock) // 2, len(fock) // 2 + 3000), (len(fock) - 3000, len(fock))):
            lo, hi = max(lo, 0), min(hi, len(fock))
            R = o.build_ham_i(fock, hop, coul, row_begin=lo, row_end=hi)
            seg = slice(rp[lo], rp[hi])
            assert np.array_equal(rp[lo:hi + 1] - rp[lo], R.rowptr)
            assert np.array_equal(cols[seg], R.cols)
            assert np.array_equal(vals[seg], R.vals)


def test_c64_copy_is_rounded_c128(core):
    ht, ut = _rand_terms(10, 3)
    b = core.Basis.combinations(10, 5)
    A = core.build_ham_i(b, ht, ut, precision="c64")
    assert A.has_c64
    v64 = A.values("c128")
    v32 = A.values("c64")
    rp, cols, _ = A.download()
    assert np.array_equal(v32, H.herm_c64_values(rp, cols, v64, A.n))
    # the copy is exactly Hermitian and each value is within one float ulp of FP64
    d32 = sp_dense(rp, cols, v32, A.m, A.n)
    assert np.array_equal(d32, d32.conj().T)
    assert np.all(np.abs(v32 - v64) <= 1.2e-7 * np.abs(v64) + 1e-300)
    # a rectangular operator keeps plain per-entry rounding
    m = H.small_siam(nbath=2, seed=3)
    bi = core.Basis.from_words(H.fock_words(m.norb, m.nocc), m.norb)
    bn = core.Basis.from_words(H.fock_words(m.norb, m.nocc + 1), m.norb)
    D = core.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[0])
    D.add_c64()
    assert np.array_equal(D.values("c64"), D.values("c128").astype(np.complex64))


def sp_dense(rp, cols, vals, m, n):
    out = np.zeros((m, n), vals.dtype)
    rows = np.repeat(np.arange(m), np.diff(rp))
    out[rows, cols] = vals
    return out


# ------------------------------------------------------------------- SpMV

def test_spmv_and_spmm_match_oracle(core):
    import torch
    m = H.small_siam(nbath=4, seed=2)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    rng = np.random.default_rng(5)
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    y0 = rng.normal(size=A.m) + 1j * rng.normal(size=A.m)
    ref = o.matvec(R, x, y0)
    y = A.spmv(_to_dev(core, x), _to_dev(core, y0)).cpu().numpy()
    # FP64: same products, different summation order -> a few ulps of the row norm
    scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
    assert np.abs(y - ref).max() < 1e-14 * scale
    A.add_c64()
    y32 = A.spmv(_to_dev(core, x), _to_dev(core, y0), precision="c64").cpu().numpy()
    # complex64 values and x with FP32 accumulation: ~1e-6 of the row scale
    assert np.abs(y32 - ref).max() < 2e-6 * scale
    X = rng.normal(size=(5, A.n)) + 1j * rng.normal(size=(5, A.n))
    Y = np.zeros((5, A.m), np.complex128)
    Yd = A.spmm(_to_dev(core, X), _to_dev(core, Y)).cpu().numpy()
    for k in range(5):
        assert np.abs(Yd[k] - o.matvec(R, X[k])).max() < 1e-14 * scale


def test_rectangular_transop_apply(core):
    m = H.small_siam(nbath=4, seed=2)
    fi, fn = H.fock_words(m.norb, m.nocc), H.fock_words(m.norb, m.nocc + 1)
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    D = core.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[2])
    R = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[2])
    x = np.random.default_rng(1).normal(size=D.n) + 0j
    y = D.matvec(_to_dev(core, x)).cpu().numpy()
    np.testing.assert_allclose(y, o.matvec(R, x), rtol=0, atol=1e-13)


# --------------------------------------------------------------- solvers

def _siam_hn(core, nbath=4, seed=2):
    m = H.small_siam(nbath=nbath, seed=seed)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    return m, fn, bn, core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n), \
        o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)


def test_krylov_matches_oracle(core):
    m, fn, bn, A, R = _siam_hn(core)
    phi = np.random.default_rng(8).normal(size=A.m) + 0.5j
    a, b, nrm = core.build_krylov(A, _to_dev(core, phi), 120)
    ra, rb, rn = o.build_krylov(R, phi, 120)
    assert len(a) == len(ra) == 120
    assert abs(nrm - rn) <= 1e-12 * rn
    # same recurrence, different summation order: first steps agree to ~1e-10
    np.testing.assert_allclose(a[:20], ra[:20], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b[:20], rb[:20], rtol=1e-9, atol=1e-9)
    om = np.linspace(-40, 40, 2000)
    s = o.cf_spectrum(a, b, nrm, 0.0, om, 0.5)
    sr = o.cf_spectrum(ra, rb, rn, 0.0, om, 0.5)
    assert np.linalg.norm(s - sr) / np.linalg.norm(sr) < 1e-8


def test_krylov_breakdown_stops(core):
    # an invariant 3-dimensional start space: beta_3 < 1e-10 ends the recurrence
    n, N = 6, 1
    fock = H.fock_words(n, N)
    hop = [[1, 1, 1.0, 0], [2, 2, 2.0, 0], [3, 3, 3.0, 0], [4, 4, 4.0, 0], [5, 5, 5.0, 0],
           [6, 6, 6.0, 0]]
    b = core.Basis.from_words(fock, n)
    A = core.build_ham_i(b, hop, None)
    phi = np.array([1, 1, 1, 0, 0, 0], dtype=np.complex128)
    a, bb, nrm = core.build_krylov(A, _to_dev(core, phi), 10)
    ra, rb, rn = o.build_krylov(o.build_ham_i(fock, hop, None), phi, 10)
    assert len(a) == len(ra) == 3
    np.testing.assert_allclose(a, ra, atol=1e-12)


def test_minres_matches_oracle_and_dense(core):
    m = H.small_siam(nbath=2, seed=4)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    om = -20.0 + 0.5j
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=om, ham_sign=-1.0)
    R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=om, ham_sign=-1.0)
    bvec = np.random.default_rng(3).normal(size=A.m) + 0j
    x, it, res = core.minres(A, _to_dev(core, bvec), tol=1e-10, maxit=2000)
    rx, rit, rres = o.pminres(R, bvec, 1e-10, 2000)
    assert abs(it - rit) <= 2
    xd = x.cpu().numpy()
    np.testing.assert_allclose(xd, rx, rtol=0, atol=1e-8 * np.abs(rx).max())
    assert res < 1e-8 * np.linalg.norm(bvec) + 1e-9
    np.testing.assert_allclose(R.todense() @ xd, bvec, atol=1e-7)


def test_lanczos_eigs_match_dense(core):
    m = H.small_siam(nbath=2, seed=5)          # 12 valence orbitals, N=6: dim 924
    fi = H.fock_words(m.norb, m.nocc)
    bi = core.Basis.from_words(fi, m.norb)
    A = core.build_ham_i(bi, m.hop_i, m.coul_i)
    e = np.linalg.eigvalsh(o.build_ham_i(fi, m.hop_i, m.coul_i).todense())
    ev, vecs, it = core.lanczos_eigs(A, 3, maxiter=300, tol=1e-10, seed=1)
    np.testing.assert_allclose(ev, e[:3], atol=1e-10)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    V = vecs.cpu().numpy()
    for k in range(3):
        assert np.linalg.norm(Hd @ V[k] - ev[k] * V[k]) < 1e-8


def test_lanczos_finds_degenerate_levels(core):
    # d^2 with Slater F's and cubic field: multiplets (degeneracies up to 6)
    from edrixs_amd import siam
    u = siam.umat_slater_d(4.0, 2.0, 0.5)
    e_cf = siam.cf_cubic_d(1.5)
    ht, ut = siam.emat_terms(e_cf), siam.umat_terms(u)
    fock = H.fock_words(10, 2)
    b = core.Basis.from_words(fock, 10)
    A = core.build_ham_i(b, ht, ut)
    e = np.linalg.eigvalsh(o.build_ham_i(fock, ht, ut).todense())
    nev = 8
    ev, _, _ = core.lanczos_eigs(A, nev, maxiter=45, tol=1e-10, seed=3, deflate=True)
    np.testing.assert_allclose(ev, e[:nev], atol=1e-9)


def test_denmat_matches_oracle(core):
    m = H.small_siam(nbath=2, seed=6)
    fi = H.fock_words(m.norb, m.nocc)
    bi = core.Basis.from_words(fi, m.norb)
    v = np.random.default_rng(2).normal(size=len(fi)) + 1j * np.random.default_rng(3).normal(size=len(fi))
    v /= np.linalg.norm(v)
    dm = core.denmat(bi, m.norb, _to_dev(core, v))
    np.testing.assert_allclose(dm, o.denmat(fi, m.norb, v), atol=1e-13)


def test_cf_kernel_matches_oracle(core):
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=300), np.abs(rng.normal(size=300))
    om = np.linspace(-10, 10, 2000)
    s = core.cf_spectrum(a, b, 1.7, 0.25, om, 0.5, weight=0.5)
    np.testing.assert_allclose(s, 0.5 * o.cf_spectrum(a, b, 1.7, 0.25, om, 0.5), rtol=1e-12,
                               atol=1e-14)


# ------------------------------------------------------- drop-in end to end

def test_fsolver_pipeline_matches_oracle(core, tmp_path, monkeypatch):
    """ed -> xas -> rixs -> opavg through the files, each output checked."""
    from edrixs_amd import iostream, solvers, spectra
    m = H.small_siam(nbath=2, seed=11)            # norb 12, N 6: dims 924 / 6*792
    wd = str(tmp_path)
    # the drivers' products in complex128 CSR instead of the exact indexed form
    monkeypatch.setenv("EDX_FSOLVER_C128", "1")
    evals128, _ = solvers.ed_siam(m, wd, neval=2, nvector=2, eigval_tol=1e-12, min_ndim=10)
    monkeypatch.delenv("EDX_FSOLVER_C128")
    evals, dm = solvers.ed_siam(m, wd, neval=2, nvector=2, eigval_tol=1e-12, min_ndim=10)
    fi = H.fock_words(m.norb, m.nocc)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e, V = np.linalg.eigh(Hd)
    np.testing.assert_allclose(evals, e[:2], atol=2e-10)        # f20.10 text + solver
    np.testing.assert_allclose(evals128, e[:2], atol=2e-10)
    for k in range(2):
        np.testing.assert_allclose(np.trace(dm[k]).real, m.nocc, atol=1e-8)
    # XAS with the ground state written by the ED step
    om = np.linspace(-10, 20, 500)
    xas, poles = solvers.xas_siam(m, om, wd, gamma_c=0.5, num_gs=1, nkryl=80)
    # oracle: same eigvec.1 file, same operators
    import struct
    raw = open(os.path.join(wd, "eigvec.1"), "rb").read()
    gs = np.frombuffer(raw[16 + 4:16 + 4 + 16 * len(fi)], dtype=np.complex128)
    egs = struct.unpack("<d", raw[4:12])[0]
    fn = H.fock_words(m.norb, m.nocc + 1)
    Rn = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    ref = np.zeros(len(om))
    for k in range(3):
        D = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[k])
        a, b, nrm = o.build_krylov(Rn, o.matvec(D, gs), 80)
        ref += o.cf_spectrum(a, b, nrm, egs, om, 0.5)
    ref /= 3
    assert spectra.relative_l2(xas, ref) < 1e-8
    # RIXS at one incident energy, one polarization pair
    pols = [([1, 0, 0], [0, 1, 0])]
    el = np.linspace(-1, 10, 300)
    rixs, _ = solvers.rixs_siam(m, [2.0], el, pols, wd, gamma_c=0.5, gamma_f=0.1, nkryl=60,
                                linsys_tol=1e-10, linsys_max=2000)
    ti, _ = m.pol_trans([1, 0, 0])
    _, tf = m.pol_trans([0, 1, 0])
    Di = o.build_transop_i(fn, fi, m.norb, m.ncore, ti)
    Af = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=2.0 + egs + 0.5j,
                       ham_sign=-1.0)
    x, _, _ = o.pminres(Af, o.matvec(Di, gs), 1e-10, 2000)
    Df = o.build_transop_f(fi, fn, m.norb, m.ncore, tf)
    Hf = o.build_ham_i(fi, m.hop_i, m.coul_i)
    a, b, nrm = o.build_krylov(Hf, o.matvec(Df, x), 60)
    rref = o.cf_spectrum(a, b, nrm, egs, el, 0.1)
    assert spectra.relative_l2(rixs[0, :, 0], rref) < 1e-7
    # opavg: <v|H|v> = E
    from edrixs_amd import fedrixs
    cwd = os.getcwd()
    os.chdir(wd)
    try:
        iostream.write_config(num_val_orbs=m.norb, num_gs=2)
        fedrixs.opavg_fsolver(0, 0, 1)
        d = np.loadtxt("opavg.dat", ndmin=2)
    finally:
        os.chdir(cwd)
    np.testing.assert_allclose(d[:, 2], d[:, 1], atol=1e-9)


def test_full_diag_path_matches_reference_test(core, tmp_path):
    """edrixs/tests/test_fortran.py:8-42: d^2, ed_solver=0, all 45 eigenvalues."""
    from edrixs_amd import fedrixs, iostream, siam
    emat = siam.cf_cubic_d(1.5)
    umat = siam.umat_slater_d(4.0, 2.0, 0.5)
    fock = H.fock_words(10, 2)
    Hd = o.dense_operator(10, emat, umat)[np.ix_(fock, fock)]
    e = np.linalg.eigvalsh(Hd)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with open("config.in", "w") as f:
            f.write("&control\ned_solver     =   0\nnum_val_orbs  =   20\nncv           =   150\n"
                    "neval         =   45\nnvector       =   45\nnum_gs        =   45\n"
                    "maxiter       =   1000\neigval_tol    =   1E-10\nidump         =   .true.\n"
                    "&end\n")
        iostream.write_fock_dec_by_N(10, 2)
        iostream.write_emat(emat, "hopping_i.in")
        iostream.write_umat(umat, "coulomb_i.in")
        fedrixs.ed_fsolver(0, 0, 1)
        _, e_gpu = np.loadtxt("eigvals.dat", unpack=True)
    finally:
        os.chdir(cwd)
    assert np.all(np.isclose(e, e_gpu))
    np.testing.assert_allclose(e_gpu, e, atol=1e-9)


def test_arpack_path_matches_reference_test(core, tmp_path):
    """The case of edrixs/tests/test_fortran.py:8-42 through ed_solver=2 (the
    ARPACK path, here thick-restart Lanczos with ncv = 150 and deflated
    restarts): all 45 eigenvalues of d^2, whose levels are up to 9-fold."""
    from edrixs_amd import fedrixs, iostream, siam
    emat = siam.cf_cubic_d(1.5)
    umat = siam.umat_slater_d(4.0, 2.0, 0.5)
    fock = H.fock_words(10, 2)
    e = np.linalg.eigvalsh(o.dense_operator(10, emat, umat)[np.ix_(fock, fock)])
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with open("config.in", "w") as f:
            f.write("&control\ned_solver     =   2\nnum_val_orbs  =   20\nncv           =   150\n"
                    "neval         =   45\nnvector       =   45\nnum_gs        =   45\n"
                    "maxiter       =   1000\neigval_tol    =   1E-10\nmin_ndim      =   10\n"
                    "idump         =   .true.\n&end\n")
        iostream.write_fock_dec_by_N(10, 2)
        iostream.write_emat(emat, "hopping_i.in")
        iostream.write_umat(umat, "coulomb_i.in")
        fedrixs.ed_fsolver(0, 0, 1)
        _, e_gpu = np.loadtxt("eigvals.dat", unpack=True)
    finally:
        os.chdir(cwd)
    np.testing.assert_allclose(e_gpu, e, atol=1e-9)


def test_missing_input_raises(core, tmp_path):
    from edrixs_amd import EdxError, fedrixs
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(EdxError):
            fedrixs.ed_fsolver(0, 0, 1)
    finally:
        os.chdir(cwd)


# ------------------------------------------------------ batched forms

def test_krylov_batch_matches_single_and_is_order_independent(core):
    import torch
    m, fn, bn, A, R = _siam_hn(core)
    rng = np.random.default_rng(12)
    V0 = rng.normal(size=(3, A.n)) + 1j * rng.normal(size=(3, A.n))
    Vd = _to_dev(core, V0)
    al, be, nr = core.build_krylov_batch(A, Vd, 80)
    for k in range(3):
        a1, b1, n1 = core.build_krylov(A, Vd[k].contiguous(), 80)
        ra, rb, rn = o.build_krylov(R, V0[k], 80)
        np.testing.assert_allclose(al[k][:20], ra[:20], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(al[k][:20], a1[:20], rtol=1e-11, atol=1e-11)
        assert abs(nr[k] - rn) <= 1e-12 * rn
        # alone (batch of one) vs in the batch: identical numbers
        ak, bk, nk = core.build_krylov_batch(A, Vd[k:k + 1].contiguous(), 80)
        assert np.array_equal(ak[0], al[k]) and np.array_equal(bk[0], be[k]) and nk[0] == nr[k]
    # reversed order gives the same per-RHS numbers
    alr, ber, nrr = core.build_krylov_batch(A, torch.flip(Vd, [0]).contiguous(), 80)
    for k in range(3):
        assert np.array_equal(alr[2 - k], al[k]) and np.array_equal(ber[2 - k], be[k])


def test_minres_shifted_batch_matches_oracle(core):
    m = H.small_siam(nbath=2, seed=4)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    rng = np.random.default_rng(9)
    sig = np.array([-20.0 + 0.5j, -15.0 + 0.5j, -5.0 + 0.3j, 10.0 + 0.5j])
    B = rng.normal(size=(4, Hn.n)) + 1j * rng.normal(size=(4, Hn.n))
    X, it, res = core.minres_shifted_batch(Hn, sig, _to_dev(core, B), tol=1e-10, maxit=3000)
    Xd = X.cpu().numpy()
    for k in range(4):
        S = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=sig[k], ham_sign=-1.0)
        rx, rit, rres = o.pminres(S, B[k], 1e-10, 3000)
        assert abs(int(it[k]) - rit) <= 3
        np.testing.assert_allclose(Xd[k], rx, rtol=0, atol=1e-7 * np.abs(rx).max())
        assert res[k] < 1e-8 * np.linalg.norm(B[k])


def test_split_c64d_product_error_is_offdiagonal_rounding_only(core):
    m, fn, bn, A, R = _siam_hn(core)
    A.prepare("c64d")
    rng = np.random.default_rng(1)
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    y = A.matvec(_to_dev(core, x), precision="c64d").cpu().numpy()
    # reference: exact diagonal + complex64-rounded (Hermitian-consistent) off-diagonal
    # values, FP64 arithmetic
    D = R.todense()
    dg = np.diag(np.diag(D))
    v32 = H.herm_c64_values(R.rowptr, R.cols, R.vals, R.ncols)
    off = sp_dense(R.rowptr, R.cols, v32.astype(np.complex128), R.m, R.ncols)
    off[np.arange(R.m), np.arange(R.m)] = 0
    assert np.array_equal(off, off.conj().T)
    ref = (dg + off) @ x
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    exact = D @ x
    assert np.abs(y - exact).max() < 1e-6 * np.abs(exact).max()


def test_krylov_reorth_is_exact_projection(core):
    # with full reorthogonalization the fraction from dim steps equals the pole sum
    m = H.small_siam(nbath=2, seed=8)
    fi = H.fock_words(m.norb, m.nocc)
    bi = core.Basis.from_words(fi, m.norb)
    A = core.build_ham_i(bi, m.hop_i, m.coul_i)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    phi = np.random.default_rng(4).normal(size=A.n) + 0j
    a, b, nrm = core.build_krylov_reorth(A, _to_dev(core, phi), 400)
    assert len(a) <= A.n
    e, v = np.linalg.eigh(Hd)
    w = np.abs(v.conj().T @ phi) ** 2
    om = np.linspace(-30, 30, 601)
    z = om + 0.3j
    exact = -1 / np.pi * np.imag((w[None, :] / (z[:, None] - e[None, :])).sum(1))
    np.testing.assert_allclose(o.cf_spectrum(a, b, nrm, 0.0, om, 0.3), exact, rtol=1e-9, atol=1e-12)


def test_rhs_parallel_batches_match_single(core):
    # K >= 8 runs the RHS-parallel SpMM schedule
    m, fn, bn, A, R = _siam_hn(core)
    rng = np.random.default_rng(21)
    K = 10
    V0 = rng.normal(size=(K, A.n)) + 1j * rng.normal(size=(K, A.n))
    al, be, nr = core.build_krylov_batch(A, _to_dev(core, V0), 60)
    for k in (0, 5, 9):
        ra, rb, rn = o.build_krylov(R, V0[k], 60)
        np.testing.assert_allclose(al[k][:20], ra[:20], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(be[k][:20], rb[:20], rtol=1e-9, atol=1e-9)
    for prec in ("c128", "c64d"):
        sig = -12.0 + 0.5j + np.arange(9) * 1.5
        B = rng.normal(size=(9, A.n)) + 0j
        X, it, res = core.minres_shifted_batch(A, sig, _to_dev(core, B), tol=1e-9, maxit=3000,
                                               precision=prec)
        Xd = X.cpu().numpy()
        Hd = R.todense()
        for k in range(9):
            r = sig[k] * Xd[k] - Hd @ Xd[k] - B[k]
            assert np.linalg.norm(r) < 1e-7 * np.linalg.norm(B[k]), (prec, k)


# ------------------------------------------------------- multi-shift RIXS solve

def _rixs_case(seed=3):
    m = H.small_siam(nbath=2, seed=seed)
    fi, fn = H.fock_words(m.norb, m.nocc), H.fock_words(m.norb, m.nocc + 1)
    Ri = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e, v = np.linalg.eigh(Ri)
    Rn = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Din = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[0]).todense()
    Dout = o.build_transop_f(fi, fn, m.norb, m.ncore, m.trans_f[1])
    b = Din @ v[:, 0]
    lo = np.linalg.eigvalsh(Rn.todense())[0]
    sig = lo + np.linspace(-2.0, 10.0, 6) + 0.5j
    return m, fi, fn, Rn, Dout, b, sig


@pytest.mark.parametrize("precision", ["c128", "c64d"])
def test_shifted_fom_matches_dense_solve(core, precision):
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Df = core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1])
    tol = 1e-10 if precision == "c128" else 1e-8
    (X, Y), it, res = core.shifted_fom(Hn, _to_dev(core, b), sig, projections=(None, Df),
                                       tol=tol, maxit=4000, precision=precision)
    X, Y = X.cpu().numpy(), Y.cpu().numpy()
    A = Rn.todense()
    Dd = Dout.todense()
    nb = np.linalg.norm(b)
    for k in range(len(sig)):
        S = sig[k] * np.eye(A.shape[0]) - A
        x = np.linalg.solve(S, b)
        true_res = np.linalg.norm(b - S @ X[k]) / nb
        assert res[k] <= tol and 0 < it[k] < 4000
        if precision == "c128":
            # the recurrence's residual tracks the true one (no reorthogonalization)
            assert true_res < 10 * tol
            assert np.linalg.norm(X[k] - x) < 1e-7 * np.linalg.norm(x)
            assert np.linalg.norm(Y[k] - Dd @ x) < 1e-7 * np.linalg.norm(Dd @ x)
            # the reference's per-shift MINRES reaches the same solution
            R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=sig[k], ham_sign=-1.0)
            rx, _, _ = o.pminres(R, b, 1e-12 * nb, 6000)
            assert np.linalg.norm(X[k] - rx) < 1e-7 * np.linalg.norm(rx)
        else:
            # complex64 off-diagonal values: operator rounding ~ 1e-7 relative
            assert true_res < 1e-5
            assert np.linalg.norm(Y[k] - Dd @ x) < 1e-5 * np.linalg.norm(Dd @ x)
        # the projection of the returned iterate is the projection of x
        assert np.linalg.norm(Y[k] - Dd @ X[k]) < 1e-12 * np.linalg.norm(Y[k]) + 1e-300


def test_shifted_fom_edge_cases(core):
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bn = core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    # b = 0: zero solution, no iterations
    (X,), it, res = core.shifted_fom(Hn, _to_dev(core, np.zeros(Hn.n)), sig[:2])
    assert np.all(X.cpu().numpy() == 0) and list(it) == [0, 0] and list(res) == [0, 0]
    # a real shift has no guaranteed nonzero pivots: rejected
    with pytest.raises(Exception, match="imaginary"):
        core.shifted_fom(Hn, _to_dev(core, b), [sig[0], 3.0 + 0j])
    # maxit caps the iteration; the reported residual is that of the last iterate
    (X,), it, res = core.shifted_fom(Hn, _to_dev(core, b), sig[:1], tol=0.0, maxit=7)
    A = sig[0] * np.eye(Hn.n) - Rn.todense()
    assert it[0] == 7
    tr = np.linalg.norm(b - A @ X.cpu().numpy()[0]) / np.linalg.norm(b)
    assert abs(tr - res[0]) < 1e-6 * tr
    # an eigenvector start converges at once (Lanczos breakdown after one step)
    ev = np.linalg.eigh(Rn.todense())[1][:, 5]
    (X,), it, res = core.shifted_fom(Hn, _to_dev(core, ev), sig[:3], tol=1e-12)
    w = np.linalg.eigvalsh(Rn.todense())[5]
    for k in range(3):
        assert it[k] == 1
        np.testing.assert_allclose(X.cpu().numpy()[k], ev / (sig[k] - w), atol=1e-12)


@pytest.mark.parametrize("precision", ["c128", "c64d", "c128i"])
def test_shifted_fom_batch_matches_dense_and_single_solves(core, precision):
    # three right-hand sides in one interleaved recurrence (the second a
    # different incident polarization, the third zero), each projected by its
    # own operator: the dense solutions, and the per-RHS recurrence to the
    # rounding of the two summation orders
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Df = core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1])
    Ri = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e, v = np.linalg.eigh(Ri)
    b2 = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[2]).todense() @ v[:, 0]
    Bm = np.stack([b, b2, np.zeros_like(b)])
    tol = 1e-10 if precision != "c64d" else 1e-8
    outs, it, res = core.shifted_fom_batch(Hn, _to_dev(core, Bm), sig, (Df, None, Df), tol=tol,
                                           maxit=4000, precision=precision)
    Y0, X1, Y2 = (t.cpu().numpy() for t in outs)
    assert it.shape == (3, len(sig)) and res.shape == (3, len(sig))
    assert np.all(it[2] == 0) and np.all(res[2] == 0) and not Y2.any()
    A = Rn.todense()
    Dd = Dout.todense()
    bound = 1e-7 if precision != "c64d" else 1e-5
    for k in range(len(sig)):
        S = sig[k] * np.eye(A.shape[0]) - A
        x0, x1 = np.linalg.solve(S, b), np.linalg.solve(S, b2)
        assert res[0, k] <= tol and res[1, k] <= tol and 0 < it[0, k] < 4000
        assert np.linalg.norm(Y0[k] - Dd @ x0) < bound * np.linalg.norm(Dd @ x0)
        assert np.linalg.norm(X1[k] - x1) < bound * np.linalg.norm(x1)
    # 40 steps of every shift: the per-RHS recurrence to the rounding of the
    # two summation orders (a converged run's step count is not comparable:
    # without reorthogonalization the rounding moves it by a few per cent)
    outs, it, res = core.shifted_fom_batch(Hn, _to_dev(core, Bm[:2]), sig, (Df, None), tol=0.0,
                                           maxit=40, precision=precision)
    for r, P in ((0, Df), (1, None)):
        (Ys,), its, rss = core.shifted_fom(Hn, _to_dev(core, Bm[r]), sig, projections=(P,),
                                           tol=0.0, maxit=40, precision=precision)
        Z, Ys = outs[r].cpu().numpy(), Ys.cpu().numpy()
        assert np.all(its == 40) and np.all(it[r] == 40)
        np.testing.assert_allclose(res[r], rss, rtol=1e-8)
        for k in range(len(sig)):
            assert np.linalg.norm(Z[k] - Ys[k]) < 1e-11 * np.linalg.norm(Ys[k]), (r, k)
    with pytest.raises(Exception, match="nrhs"):
        core.shifted_fom_batch(Hn, _to_dev(core, np.zeros((65, Hn.n), complex)), sig,
                               [None] * 65, precision=precision)


def test_shifted_fom_batch_over_the_factored_block_matches_single_solves(core, monkeypatch):
    # the path of the benchmark step: two right-hand sides over H_n in the
    # interleaved core-hole layout with its slot-factored form (indexed
    # precision), the product the factored 2-vector block (forced: a matrix
    # this size is below the autotuner's threshold), each projected by its own
    # emission operator, against one plain complex128 recurrence per right-hand
    # side in the reference layout (itself pinned to dense solves above)
    m, fn, bn, A, R = _siam_hn(core)
    fi = H.fock_words(m.norb, m.nocc)
    bi = core.Basis.from_words(fi, m.norb)
    lay = "interleaved"
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout=lay)
    Hn.prepare("c128i")
    assert Hn.factored() is not None
    Df = [core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[k], layout=lay) for k in (1, 2)]
    Dr = [core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[k]) for k in (1, 2)]
    p = core.corehole_permutation(len(fn), m.ncore)       # x_inter = x_ref[p]
    rng = np.random.default_rng(41)
    Bref = rng.normal(size=(2, A.n)) + 1j * rng.normal(size=(2, A.n))
    sig = np.linspace(-10.0, 20.0, 6) + 0.5j
    monkeypatch.setenv("EDX_FORCE_BSCHED", "3:4")
    outs, it, res = core.shifted_fom_rhs(Hn, _to_dev(core, Bref[:, p]), sig, Df, tol=1e-10,
                                         maxit=4000, precision="c128i")
    monkeypatch.delenv("EDX_FORCE_BSCHED")
    for r in range(2):
        (Ys,), its, rss = core.shifted_fom(A, _to_dev(core, Bref[r]), sig, projections=(Dr[r],),
                                           tol=1e-10, maxit=4000, precision="c128")
        Y, Ys = outs[r].cpu().numpy(), Ys.cpu().numpy()
        assert np.all(res[r] <= 1e-10) and np.all(it[r] > 0) and np.all(it[r] < 4000)
        for j in range(len(sig)):
            assert np.linalg.norm(Y[j] - Ys[j]) < 1e-7 * np.linalg.norm(Ys[j]), (r, j)


def test_shifted_fom_rhs_chooses_block_or_separate_recurrences(core, monkeypatch):
    # a small block runs as one interleaved recurrence (shifted_fom_batch's
    # results bit for bit); with the block-size limit at zero, as one
    # recurrence per row (shifted_fom's results bit for bit)
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Df = core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1])
    Bd = _to_dev(core, np.stack([b, np.roll(b, 3)]))
    kw = dict(tol=1e-10, maxit=4000, precision="c128")
    o1, i1, r1 = core.shifted_fom_rhs(Hn, Bd, sig, (Df, None), **kw)
    o2, i2, r2 = core.shifted_fom_batch(Hn, Bd, sig, (Df, None), **kw)
    assert all(np.array_equal(x.cpu().numpy(), y.cpu().numpy()) for x, y in zip(o1, o2))
    assert np.array_equal(i1, i2) and np.array_equal(r1, r2)
    monkeypatch.setattr(core, "SPLIT_BLOCK_BYTES", 0)
    o3, i3, r3 = core.shifted_fom_rhs(Hn, Bd, sig, (Df, None), **kw)
    for r, P in ((0, Df), (1, None)):
        (Y,), it, res = core.shifted_fom(Hn, Bd[r], sig, projections=(P,), **kw)
        assert np.array_equal(o3[r].cpu().numpy(), Y.cpu().numpy())
        assert np.array_equal(i3[r], it) and np.array_equal(r3[r], res)


def test_shifted_fom_packed_projection_equals_csr_projection(core, monkeypatch):
    # the projection read as packed words (column | value number, the table of
    # its few distinct values) gives the complex128 CSR's sums bit for bit: the
    # same values in the same order; one operator made before packing is
    # switched on, so it keeps the CSR
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    bd = _to_dev(core, b)
    outs = []
    for off in ("1", "0"):
        monkeypatch.setenv("EDX_NO_PACKED_PROJ", off)
        Df = core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1])
        (Y,), it, res = core.shifted_fom(Hn, bd, sig, projections=(Df,), tol=1e-10, maxit=4000)
        (Yb,), itb, resb = core.shifted_fom_batch(Hn, bd.reshape(1, -1), sig, (Df,), tol=1e-10,
                                                  maxit=4000)
        outs.append((Y.cpu().numpy(), it, res, Yb.cpu().numpy(), itb, resb))
    monkeypatch.delenv("EDX_NO_PACKED_PROJ")
    (Y0, it0, r0, Yb0, itb0, rb0), (Y1, it1, r1, Yb1, itb1, rb1) = outs
    assert np.array_equal(Y0, Y1) and np.array_equal(it0, it1) and np.array_equal(r0, r1)
    assert np.array_equal(Yb0, Yb1) and np.array_equal(itb0, itb1) and np.array_equal(rb0, rb1)
    Dd = Dout.todense()
    A = Rn.todense()
    for k in range(len(sig)):
        x = np.linalg.solve(sig[k] * np.eye(A.shape[0]) - A, b)
        assert np.linalg.norm(Y1[k] - Dd @ x) < 1e-7 * np.linalg.norm(Dd @ x)


def test_shifted_fom_windowed_updates_equal_per_step_updates(core, monkeypatch):
    # the projected updates made in windows of 2, 3 and 8 steps (one pass over
    # the shifts' vectors per window) against one pass per step, for a run
    # whose shifts stop at different steps and for runs capped at 7, 8 and 13
    # steps (the last window then partly filled or full).  The same scalars, so
    # the same iteration counts and residuals; the solutions to 1e-13 of their
    # scale (a window stores its earlier steps' scaled projections, a rounding
    # the per-step form may fuse into its next multiply-add)
    m, fi, fn, Rn, Dout, b, sig = _rixs_case()
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Df = core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1])
    bd = _to_dev(core, b)
    for tol, maxit in ((1e-10, 4000), (0.0, 7), (0.0, 8), (0.0, 13)):
        out = []
        for window in ("1", "2", "3", "8"):
            monkeypatch.setenv("EDX_MS_WINDOW", window)
            (X, Y), it, res = core.shifted_fom(Hn, bd, sig, projections=(None, Df), tol=tol,
                                               maxit=maxit, precision="c128")
            out.append((X.cpu().numpy(), Y.cpu().numpy(), it, res))
        X0, Y0, it0, r0 = out[0]
        for X1, Y1, it1, r1 in out[1:]:
            for k in range(len(sig)):
                assert np.abs(X1[k] - X0[k]).max() <= 1e-13 * np.abs(X0[k]).max(), (tol, maxit, k)
                assert np.abs(Y1[k] - Y0[k]).max() <= 1e-13 * np.abs(Y0[k]).max(), (tol, maxit, k)
            assert np.array_equal(it0, it1) and np.array_equal(r0, r1)
        if tol > 0:
            assert len(set(int(v) for v in it0)) > 1      # shifts stopped at different steps
    monkeypatch.delenv("EDX_MS_WINDOW")


# ------------------------------------------------- resumable Krylov, converged CF

def test_krylov_batch_resumes_bit_exact(core):
    m, fn, bn, A, R = _siam_hn(core)
    rng = np.random.default_rng(31)
    V0 = rng.normal(size=(3, A.n)) + 1j * rng.normal(size=(3, A.n))
    Vd = _to_dev(core, V0)
    al, be, nr = core.build_krylov_batch(A, Vd, 90)
    kb = core.KrylovBatch(A, Vd, capacity=16)      # grows its history twice
    assert kb.advance(20) == 20 and kb.advance(1) == 21 and kb.advance(69) == 90
    a2, b2, n2 = kb.coefficients()
    for k in range(3):
        assert np.array_equal(a2[k], al[k]) and np.array_equal(b2[k], be[k]) and n2[k] == nr[k]
    a3, b3, _ = kb.coefficients(50)
    for k in range(3):
        assert np.array_equal(a3[k], al[k][:50]) and np.array_equal(b3[k], be[k][:50])
    kb.close()


@pytest.mark.parametrize("precision", ["c128", "c64", "c128i"])
def test_krylov_batch_keep_continues_the_kept_recurrences(core, monkeypatch, precision):
    # a batch compacted to some of its vectors (6 -> 3 -> 1) continues each
    # kept recurrence exactly as the full batch does: with one product schedule
    # for every K (forced: the pipelined RHS-parallel one, whose per-vector sums
    # and dot partials do not depend on K), the coefficients are bit-identical
    m, fn, bn, A, R = _siam_hn(core)
    A.prepare(precision)
    rng = np.random.default_rng(33)
    V0 = rng.normal(size=(6, A.n)) + 1j * rng.normal(size=(6, A.n))
    Vd = _to_dev(core, V0)
    monkeypatch.setenv("EDX_FORCE_BSCHED", "2:4")
    full = core.KrylovBatch(A, Vd, precision, capacity=16)
    assert full.advance(80) == 80
    af, bf, nf = full.coefficients()
    full.close()
    kb = core.KrylovBatch(A, Vd, precision, capacity=16)
    assert kb.advance(30) == 30
    kb.keep([1, 3, 4])
    assert kb.K == 3
    a1, b1, n1 = kb.coefficients()
    for j, k in enumerate([1, 3, 4]):
        assert np.array_equal(a1[j], af[k][:30]) and np.array_equal(b1[j], bf[k][:30])
        assert n1[j] == nf[k]
    with pytest.raises(Exception, match="increase strictly"):
        kb.keep([1, 0])                  # refused before anything changes
    assert kb.K == 3
    assert kb.advance(30) == 60          # grows its history after the compaction
    kb.keep([2])
    assert kb.advance(20) == 80
    a2, b2, n2 = kb.coefficients()
    assert np.array_equal(a2[0], af[4]) and np.array_equal(b2[0], bf[4]) and n2[0] == nf[4]
    kb.close()
    monkeypatch.delenv("EDX_FORCE_BSCHED")


def test_converged_spectra_compaction_changes_nothing(core, monkeypatch):
    # converged_spectra drops the recurrences whose spectra are final; with one
    # product schedule for every K the spectra and step counts equal those of
    # the uncompacted block bit for bit
    m, fn, bn, A, R = _siam_hn(core)
    A.prepare("c128i")
    ev, vecs, _ = core.lanczos_eigs(A, 1, tol=1e-12)
    rng = np.random.default_rng(35)
    V0 = rng.normal(size=(8, A.n)) + 1j * rng.normal(size=(8, A.n))
    V0[5] = vecs[0].cpu().numpy()         # a near-eigenvector: its spectrum is final first
    V0[2] = V0[5] + 1e-3 * V0[2]
    Vd = _to_dev(core, V0)
    om = np.linspace(ev[0] - 5.0, ev[0] + 40.0, 301)
    monkeypatch.setenv("EDX_FORCE_BSCHED", "2:4")
    runs = [core.converged_spectra(A, Vd, ev[0], om, 0.3, precision="c128i", tol=1e-9, chunk=20,
                                   min_steps=20, max_steps=400, compact=c) for c in (False, True)]
    monkeypatch.delenv("EDX_FORCE_BSCHED")
    (s0, st0, c0), (s1, st1, c1) = runs
    assert st0 == st1 and c0 == c1
    assert len(set(st0)) > 1                                  # members stopped apart
    for k in range(8):
        assert np.array_equal(s0[k], s1[k]), k


def test_converged_spectra_equal_pole_sum(core):
    m = H.small_siam(nbath=2, seed=8)
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    rng = np.random.default_rng(5)
    Phi = rng.normal(size=(2, A.n)) + 1j * rng.normal(size=(2, A.n))
    e, v = np.linalg.eigh(Hd)
    egs = e[0]
    om = np.linspace(0.0, 10.0, 401)
    gam = 0.2
    for prec, tol, bound in (("c128", 1e-10, 1e-8), ("c64d", 1e-10, 1e-6)):
        spec, steps, change = core.converged_spectra(A, _to_dev(core, Phi), egs, om, gam,
                                                      precision=prec, tol=tol, chunk=40,
                                                      min_steps=40, max_steps=4000)
        for k in range(2):
            w = np.abs(v.conj().T @ Phi[k]) ** 2
            z = om + 1j * gam + egs
            exact = -1 / np.pi * np.imag((w[None, :] / (z[:, None] - e[None, :])).sum(1))
            assert change[k] <= tol and 0 < steps[k] < 4000
            assert np.linalg.norm(spec[k] - exact) < bound * np.linalg.norm(exact), (prec, k)


def test_converged_spectra_split_equals_block(core, monkeypatch):
    # forcing the one-recurrence-per-vector schedule changes no number
    m, fn, bn, A, R = _siam_hn(core)
    rng = np.random.default_rng(17)
    Phi = _to_dev(core, rng.normal(size=(3, A.n)) + 1j * rng.normal(size=(3, A.n)))
    om = np.linspace(-20.0, 20.0, 201)
    a = core.converged_spectra(A, Phi, 0.0, om, 0.5, tol=1e-9, chunk=30, min_steps=30)
    monkeypatch.setattr(core, "SPLIT_BLOCK_BYTES", 0)
    b = core.converged_spectra(A, Phi, 0.0, om, 0.5, tol=1e-9, chunk=30, min_steps=30)
    for k in range(3):
        assert np.array_equal(a[0][k], b[0][k]) and a[1][k] == b[1][k]


# ----------------------------------------------------------- grouped recurrences

@pytest.mark.parametrize("precision", ["c128", "c64d"])
def test_group_members_match_oracle_and_are_mix_independent(core, precision):
    m1, fn1, bn1, A1, R1 = _siam_hn(core, nbath=4, seed=2)
    m2 = H.small_siam(nbath=3, seed=5)
    f2 = H.fock_words(m2.norb, m2.nocc)
    A2 = core.build_ham_i(core.Basis.from_words(f2, m2.norb), m2.hop_i, m2.coul_i)
    R2 = o.build_ham_i(f2, m2.hop_i, m2.coul_i)
    rng = np.random.default_rng(41)
    va = rng.normal(size=A1.n) + 1j * rng.normal(size=A1.n)
    vb = rng.normal(size=A2.n) + 1j * rng.normal(size=A2.n)
    vc = rng.normal(size=A1.n) + 0j
    # member a alone
    g1 = core.KrylovGroup(precision, max_steps=200)
    a = g1.add(A1, _to_dev(core, va))
    g1.advance(60)
    al1, be1, n1, run1 = g1.coefficients(a)
    assert len(al1) == 60 and run1
    if precision == "c128":
        ra, rb, rn = o.build_krylov(R1, va, 60)
        np.testing.assert_allclose(al1[:20], ra[:20], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(be1[:20], rb[:20], rtol=1e-9, atol=1e-9)
        assert abs(n1 - rn) <= 1e-12 * rn
    # the same member in a changing mix: others join and leave mid-way
    g2 = core.KrylovGroup(precision, max_steps=200)
    b = g2.add(A2, _to_dev(core, vb))
    g2.advance(7)
    a2 = g2.add(A1, _to_dev(core, va))
    g2.advance(20)
    c = g2.add(A1, _to_dev(core, vc))
    g2.advance(15)
    g2.remove(b)
    g2.advance(25)
    al2, be2, n2, _ = g2.coefficients(a2)
    assert np.array_equal(al2, al1) and np.array_equal(be2, be1) and n2 == n1
    alb, beb, nb_, _ = g2.coefficients(c)
    assert len(alb) == 40
    if precision == "c128":
        rb2 = o.build_krylov(R1, vc, 40)
        np.testing.assert_allclose(alb[:20], rb2[0][:20], rtol=1e-9, atol=1e-9)
    g1.close()
    g2.close()


@pytest.mark.parametrize("precision", ["c128", "c64d"])
def test_group_pipelined_member_matches_oracle_and_is_mix_independent(core, precision, monkeypatch):
    # H_n of 68,640 rows (at least 65,536: the member's product runs on its own
    # pipelined sub-grid; rows of up to 35 nonzeros, beyond the 32 a lane group
    # prefetches) beside small members that keep the one-group-per-row product
    m, fn, bn, A, R = _siam_hn(core, nbath=6, seed=2)
    assert A.n >= 65536 and np.diff(R.rowptr).max() > 32
    m2 = H.small_siam(nbath=3, seed=5)
    f2 = H.fock_words(m2.norb, m2.nocc)
    A2 = core.build_ham_i(core.Basis.from_words(f2, m2.norb), m2.hop_i, m2.coul_i)
    rng = np.random.default_rng(43)
    va = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    vb = _to_dev(core, rng.normal(size=A2.n) + 1j * rng.normal(size=A2.n))
    g1 = core.KrylovGroup(precision, max_steps=200)
    a = g1.add(A, _to_dev(core, va))
    g1.advance(40)
    al1, be1, n1, run1 = g1.coefficients(a)
    g1.close()
    assert len(al1) == 40 and run1
    ra, rb, rn = o.build_krylov(R, va, 40)
    tol = 1e-9 if precision == "c128" else 1e-5
    np.testing.assert_allclose(al1[:20], ra[:20], rtol=tol, atol=tol)
    np.testing.assert_allclose(be1[:20], rb[:20], rtol=tol, atol=tol)
    assert abs(n1 - rn) <= 1e-12 * rn

    def mix():
        # a small member first, so the large one's sub-grid starts after padding blocks
        g = core.KrylovGroup(precision, max_steps=200)
        b = g.add(A2, vb)
        g.advance(5)
        a2 = g.add(A, _to_dev(core, va))
        c = g.add(A2, vb)
        g.advance(17)
        g.remove(b)
        g.advance(23)
        out = g.coefficients(a2), g.coefficients(c)
        g.close()
        return out

    (al2, be2, n2, _), cg = mix()
    assert np.array_equal(al2, al1) and np.array_equal(be2, be1) and n2 == n1
    monkeypatch.setenv("EDX_NO_GRAPH", "1")
    (al3, be3, n3, _), ce = mix()
    assert np.array_equal(al3, al1) and np.array_equal(be3, be1) and n3 == n1
    assert len(ce[0]) == 40 and np.array_equal(ce[0], cg[0]) and np.array_equal(ce[1], cg[1])


def test_group_breakdown_and_capacity(core):
    m, fn, bn, A, R = _siam_hn(core, nbath=2, seed=4)
    ev = np.linalg.eigh(R.todense())[1][:, 3]
    g = core.KrylovGroup("c128", max_steps=10)
    e = g.add(A, _to_dev(core, ev))                     # eigenvector: breaks down after one step
    r = g.add(A, _to_dev(core, np.ones(A.n) + 0j))      # runs into the 10-step capacity
    live = g.advance(30)
    ae, be, ne, rune = g.coefficients(e)
    ar, br, nr, runr = g.coefficients(r)
    assert live == 0 and not rune and not runr
    assert len(ae) == 1 and be[0] < 1e-10 and len(ar) == 10
    with pytest.raises(Exception, match="no member"):
        g.remove(12345)
    g.close()


# ------------------------------------------------------------ matrix-free apply

def test_matrix_free_apply_matches_csr(core):
    m = H.small_siam(nbath=4, seed=7)
    nv, nc, N = m.norb, m.ncore, m.nocc
    fi, fn = H.fock_words(nv, N), H.fock_words(nv, N + 1)
    bi, bn = core.Basis.from_words(fi, nv), core.Basis.from_words(fn, nv)
    rng = np.random.default_rng(13)
    cases = [(core.mf_ham_n(bn, nv, nc, m.hop_n, m.coul_n),
              o.build_ham_n(fn, nv, nc, m.hop_n, m.coul_n)),
             (core.mf_ham_n(bn, nv, nc, m.hop_n, m.coul_n, omega=-3.25 + 0.5j, ham_sign=-1.0),
              o.build_ham_n(fn, nv, nc, m.hop_n, m.coul_n, omega=-3.25 + 0.5j, ham_sign=-1.0)),
             (core.mf_ham_i(bi, m.hop_i, m.coul_i), o.build_ham_i(fi, m.hop_i, m.coul_i))]
    sub = fi[::2].copy()             # an incomplete basis: binary-search column lookup
    cases.append((core.mf_ham_i(core.Basis.from_words(sub, nv), m.hop_i, m.coul_i),
                  o.build_ham_i(sub, m.hop_i, m.coul_i)))
    for mf, R in cases:
        assert mf.n == R.m
        x = rng.normal(size=R.m) + 1j * rng.normal(size=R.m)
        y0 = rng.normal(size=R.m) + 1j * rng.normal(size=R.m)
        y = mf.apply(_to_dev(core, x), _to_dev(core, y0)).cpu().numpy()
        ref = o.matvec(R, x, y0)
        scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
        assert np.abs(y - ref).max() < 1e-14 * scale
        # a column of H: columns and signs exactly as the CSR build places them
        j = int(rng.integers(R.m))
        col = mf.matvec(_to_dev(core, np.eye(1, R.m, j)[0] + 0j)).cpu().numpy()
        dense_col = np.zeros(R.m, complex)
        for r in range(R.m):
            lo, hi = R.rowptr[r], R.rowptr[r + 1]
            hit = np.nonzero(R.cols[lo:hi] == j)[0]
            if len(hit):
                dense_col[r] = R.vals[lo + hit[0]]
        np.testing.assert_allclose(col, dense_col, rtol=1e-15, atol=1e-15 * np.abs(R.vals).max())


def test_matrix_free_krylov_matches_csr(core):
    m = H.small_siam(nbath=4, seed=2)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    mf = core.mf_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    phi = np.random.default_rng(8).normal(size=mf.n) + 0.5j
    a, b, nrm = mf.build_krylov(_to_dev(core, phi), 60)
    ra, rb, rn = o.build_krylov(R, phi, 60)
    assert len(a) == 60 and abs(nrm - rn) <= 1e-12 * rn
    np.testing.assert_allclose(a[:25], ra[:25], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b[:25], rb[:25], rtol=1e-9, atol=1e-9)


def test_matrix_free_rows_with_many_terms(core):
    # dense random terms: rows have more surviving terms than one compaction pass holds
    n, N = 14, 7
    rng = np.random.default_rng(23)
    from edrixs_amd import siam
    ht = siam.emat_terms(H.random_hermitian_emat(n, rng, density=1.0))
    ut = siam.umat_terms(H.random_hermitian_umat(n, rng, nnz=3000))
    fock = H.fock_words(n, N)
    R = o.build_ham_i(fock, ht, ut)
    assert np.diff(R.rowptr).max() > 128
    mf = core.mf_ham_i(core.Basis.from_words(fock, n), ht, ut)
    x = rng.normal(size=R.m) + 1j * rng.normal(size=R.m)
    y = mf.matvec(_to_dev(core, x)).cpu().numpy()
    ref = o.matvec(R, x)
    scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
    assert np.abs(y - ref).max() < 1e-14 * scale


def test_group_graph_replays_equal_eager_launches(core, monkeypatch):
    # steps replayed from a captured HIP graph give the numbers of steps launched
    # one by one, across several replays, re-captures and eager work between them
    m1, fn1, bn1, A1, R1 = _siam_hn(core, nbath=4, seed=2)
    m2 = H.small_siam(nbath=3, seed=5)
    f2 = H.fock_words(m2.norb, m2.nocc)
    A2 = core.build_ham_i(core.Basis.from_words(f2, m2.norb), m2.hop_i, m2.coul_i)
    rng = np.random.default_rng(3)
    va = _to_dev(core, rng.normal(size=A1.n) + 1j * rng.normal(size=A1.n))
    vb = _to_dev(core, rng.normal(size=A2.n) + 0j)

    def run():
        g = core.KrylovGroup("c64d", max_steps=400)
        a = g.add(A1, va)
        g.advance(37)                    # two replays + 5 single steps
        b = g.add(A2, vb)                # new membership: capture again
        _ = A1.matvec(va)                # eager work between replays
        g.advance(50)
        g.advance(16)
        out = (g.coefficients(a), g.coefficients(b))
        g.close()
        return out

    monkeypatch.setenv("EDX_NO_GRAPH", "1")
    eager = run()
    monkeypatch.delenv("EDX_NO_GRAPH")
    graph = run()
    for (ae, be, ne, _), (ag, bg, ng, _) in zip(eager, graph):
        assert len(ae) > 0 and np.array_equal(ae, ag) and np.array_equal(be, bg) and ne == ng


# ---------------------------------------------------------- indexed exact form

def test_indexed_form_products_are_exact(core):
    m, fn, bn, A, R = _siam_hn(core)
    A.prepare("c128i")
    rng = np.random.default_rng(19)
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    y0 = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    ref = o.matvec(R, x, y0)
    scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
    y = A.spmv(_to_dev(core, x), _to_dev(core, y0), precision="c128i").cpu().numpy()
    assert np.abs(y - ref).max() < 1e-14 * scale            # FP64 rounding only
    # Krylov coefficients: those of the exact operator
    a1, b1, n1 = core.build_krylov(A, _to_dev(core, x), 40, precision="c128i")
    ra, rb, rn = o.build_krylov(R, x, 40)
    np.testing.assert_allclose(a1[:25], ra[:25], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b1[:25], rb[:25], rtol=1e-9, atol=1e-9)
    # batched (K < 8 and K >= 8 schedules) and grouped recurrences accept it
    V = rng.normal(size=(10, A.n)) + 0j
    for K in (3, 10):
        al, be, nr = core.build_krylov_batch(A, _to_dev(core, V[:K]), 30, precision="c128i")
        rk = o.build_krylov(R, V[K - 1], 30)
        np.testing.assert_allclose(al[K - 1][:20], rk[0][:20], rtol=1e-9, atol=1e-9)
    g = core.KrylovGroup("c128i", max_steps=64)
    mid = g.add(A, _to_dev(core, V[0]))
    g.advance(30)
    ag = g.coefficients(mid)[0]
    np.testing.assert_allclose(ag[:20], o.build_krylov(R, V[0], 30)[0][:20], rtol=1e-9, atol=1e-9)
    g.close()
    # ground state in the indexed form equals the FP64 one
    ev, _, _ = core.ground_state(A, nev=1, precision="c128i", tol=1e-12)
    ev64, _, _ = core.ground_state(A, nev=1, precision="c128", tol=1e-12)
    assert abs(ev[0] - ev64[0]) < 1e-10


def test_indexed_form_many_values_and_capacity(core):
    from edrixs_amd import siam
    # dense random terms, 14 orbitals: 32,484 distinct off-diagonal values still fit
    n, N = 14, 7
    rng = np.random.default_rng(23)
    ht = siam.emat_terms(H.random_hermitian_emat(n, rng, density=1.0))
    ut = siam.umat_terms(H.random_hermitian_umat(n, rng, nnz=3000))
    fock = H.fock_words(n, N)
    A = core.build_ham_i(core.Basis.from_words(fock, n), ht, ut)
    A.prepare("c128i")
    R = o.build_ham_i(fock, ht, ut)
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    y = A.matvec(_to_dev(core, x), precision="c128i").cpu().numpy()
    scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
    assert np.abs(y - o.matvec(R, x)).max() < 1e-14 * scale
    # 16 orbitals: 149,766 distinct values do not fit the table
    rng = np.random.default_rng(23)
    ht = siam.emat_terms(H.random_hermitian_emat(16, rng, density=1.0))
    ut = siam.umat_terms(H.random_hermitian_umat(16, rng, nnz=6000))
    B = core.build_ham_i(core.Basis.combinations(16, 8), ht, ut)
    with pytest.raises(Exception, match="distinct off-diagonal values"):
        B.prepare("c128i")
    # the matrix stays usable in the other forms
    xb = _to_dev(core, np.ones(B.n) + 0j)
    assert np.isfinite(B.matvec(xb, precision="c64d").cpu().numpy()).all()


@pytest.mark.parametrize("precision", ["c128", "c128i"])
def test_block_lanczos_matches_dense(core, precision):
    m = H.small_siam(nbath=2, seed=5)          # dim 924
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e = np.linalg.eigvalsh(Hd)
    ev, vecs, nb, res = core.block_lanczos_eigs(A, 3, maxblocks=115, tol=1e-10, seed=1,
                                                precision=precision)
    np.testing.assert_allclose(ev, e[:3], atol=1e-10)
    assert (res < 1e-10).all(), (nb, res)
    V = vecs.cpu().numpy()
    np.testing.assert_allclose(V.conj() @ V.T, np.eye(3), atol=1e-12)
    for k in range(3):
        assert np.linalg.norm(Hd @ V[k] - ev[k] * V[k]) < 1e-9


def test_block_lanczos_degenerate_multiplets_and_invariant_subspaces(core):
    from edrixs_amd import siam
    # d^3 with Slater F's and a cubic field (dim 120): the lowest levels are
    # 4-, 4- and 6-fold; one block of 8 resolves each copy
    ht = siam.emat_terms(siam.cf_cubic_d(1.5))
    ut = siam.umat_terms(siam.umat_slater_d(4.0, 2.0, 0.5))
    fock = H.fock_words(10, 3)
    A = core.build_ham_i(core.Basis.from_words(fock, 10), ht, ut)
    Hd = o.build_ham_i(fock, ht, ut).todense()
    e = np.linalg.eigvalsh(Hd)
    ev, vecs, nb, res = core.block_lanczos_eigs(A, 14, maxblocks=15, tol=1e-10, seed=3)
    np.testing.assert_allclose(ev, e[:14], atol=1e-10)
    V = vecs.cpu().numpy()
    np.testing.assert_allclose(V.conj() @ V.T, np.eye(14), atol=1e-11)
    assert np.abs(Hd @ V.T - V.T * ev).max() < 1e-9
    # d^2 in a cubic field without Coulomb terms (dim 45): three levels, 15-,
    # 24- and 6-fold.  The block Krylov space is invariant after a few steps;
    # the deficient block is refilled with random directions, which bring the
    # copies of the 15-fold ground level beyond the first eight
    ut0 = siam.umat_terms(np.zeros((10, 10, 10, 10), complex))
    fock2 = H.fock_words(10, 2)
    A2 = core.build_ham_i(core.Basis.from_words(fock2, 10), ht, ut0)
    e2 = np.linalg.eigvalsh(o.build_ham_i(fock2, ht, ut0).todense())
    ev2, vecs2, nb2, res2 = core.block_lanczos_eigs(A2, 12, maxblocks=5, tol=1e-10, seed=4)
    np.testing.assert_allclose(ev2, e2[:12], atol=1e-10)
    W = vecs2.cpu().numpy()
    np.testing.assert_allclose(W.conj() @ W.T, np.eye(12), atol=1e-11)


def test_block_lanczos_matches_single_vector_lanczos(core):
    m = H.small_siam(nbath=6, seed=7)          # 16 orbitals, N=8: dim 12,870
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    ev1, v1, _ = core.lanczos_eigs(A, 3, maxiter=800, tol=1e-11, seed=2, deflate=True)
    ev, vecs, nb, res = core.block_lanczos_eigs(A, 3, maxblocks=150, tol=1e-10, seed=2)
    assert (res < 1e-10).all(), (nb, res)
    np.testing.assert_allclose(ev, ev1, atol=1e-10)
    # eigenvectors: |<v_block|v_single>| = 1 for the nondegenerate ground state
    ov = abs(complex((vecs[0].conj() * v1[0]).sum().item()))
    assert abs(ov - 1.0) < 1e-9
    # true residual |H x - e x| of every returned vector
    for k in range(3):
        y = A.matvec(vecs[k].contiguous())
        assert float((y - ev[k] * vecs[k]).norm()) < 1e-9
    # mixed-precision products: eigenvalues of the rounded matrix
    evd, _, _, _ = core.block_lanczos_eigs(A, 3, maxblocks=150, tol=1e-9, seed=2,
                                           precision="c64d", nvec=0)
    assert np.abs(evd - ev1).max() < 1e-5
    # one d electron: 10 states, below one block of 16 rows
    from edrixs_amd import siam
    ht = siam.emat_terms(siam.cf_cubic_d(1.5))
    ut = siam.umat_terms(siam.umat_slater_d(4.0, 2.0, 0.5))
    tiny = core.build_ham_i(core.Basis.from_words(H.fock_words(10, 1), 10), ht, ut)
    with pytest.raises(Exception, match="at least 16 rows"):
        core.block_lanczos_eigs(tiny, 1)


@pytest.mark.parametrize("precision", ["c128", "c128i"])
def test_two_pass_lanczos_ground_state(core, precision):
    # the reference's procedure (no basis, no reorthogonalization, Ritz vector
    # from a second run of the recurrence) against dense eigh and against the
    # fully reorthogonalized solver from the same start vector
    m = H.small_siam(nbath=2, seed=5)          # dim 924
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e, V = np.linalg.eigh(Hd)
    ev, vecs, steps = core.lanczos_ground(A, 1, maxiter=600, tol=1e-10, seed=1, precision=precision)
    assert abs(ev[0] - e[0]) < 1e-10 and 10 <= steps < 600
    x = vecs[0].cpu().numpy()
    assert abs(abs(np.vdot(V[:, 0], x)) - 1.0) < 1e-9
    assert np.linalg.norm(Hd @ x - ev[0] * x) < 1e-7
    evn, _, _ = core.lanczos_ground(A, 1, maxiter=600, tol=1e-10, seed=1, nvec=0, precision=precision)
    assert evn[0] == ev[0]
    # 12,870 rows: same energy and state as the reorthogonalized solver
    m2 = H.small_siam(nbath=6, seed=7)
    f2 = H.fock_words(m2.norb, m2.nocc)
    B = core.build_ham_i(core.Basis.from_words(f2, m2.norb), m2.hop_i, m2.coul_i)
    e1, v1, _ = core.lanczos_eigs(B, 1, maxiter=800, tol=1e-11, seed=2, deflate=False,
                                  precision=precision)
    e2, v2, _ = core.ground_state(B, 1, precision=precision, maxiter=800, tol=1e-11, seed=2)
    assert abs(e1[0] - e2[0]) < 1e-10
    assert abs(abs(complex((v1[0].conj() * v2[0]).sum().item())) - 1.0) < 1e-9
    y = B.matvec(v2[0].contiguous())
    assert float((y - e2[0] * v2[0]).norm()) < 1e-7


def test_trlan_matches_dense_many_states(core):
    m = H.small_siam(nbath=2, seed=5)          # dim 924
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    Hd = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e = np.linalg.eigvalsh(Hd)
    # ARPACK's default ncv = nev + 2 (raised to 2 nev + 16 = 96) and a larger basis; the
    # 40 lowest levels of this model lie within 2.7 of each other, 3.4 % of the spectrum
    for ncv in (42, 120):
        ev, vecs, nmv, nconv = core.trlan_eigs(A, 40, ncv=ncv, maxiter=400, tol=1e-12, seed=1)
        np.testing.assert_allclose(ev, e[:40], atol=1e-9)
        assert nconv == 40
        V = vecs.cpu().numpy()
        np.testing.assert_allclose(V.conj() @ V.T, np.eye(40), atol=1e-10)
        assert np.abs(Hd @ V.T - V.T * ev).max() < 1e-8 and nmv > 0
    # few states, no deflation, eigenvectors only for the first two
    ev, vecs, _, nconv = core.trlan_eigs(A, 3, ncv=5, maxiter=400, tol=1e-12, seed=2,
                                         deflate=False, nvec=2)
    np.testing.assert_allclose(ev, e[:3], atol=1e-10)
    assert vecs.shape[0] == 2 and nconv == 3
    with pytest.raises(Exception, match="ncv"):
        core.trlan_eigs(A, 3, ncv=0)


def test_trlan_degenerate_levels_and_invariant_subspaces(core):
    from edrixs_amd import siam
    ht = siam.emat_terms(siam.cf_cubic_d(1.5))
    ut = siam.umat_terms(siam.umat_slater_d(4.0, 2.0, 0.5))
    # d^3 (dim 120): 4-, 4- and 6-fold lowest levels; deflated restarts find every copy
    fock = H.fock_words(10, 3)
    A = core.build_ham_i(core.Basis.from_words(fock, 10), ht, ut)
    Hd = o.build_ham_i(fock, ht, ut).todense()
    e = np.linalg.eigvalsh(Hd)
    ev, vecs, _, _ = core.trlan_eigs(A, 14, ncv=30, maxiter=300, tol=1e-12, seed=3)
    np.testing.assert_allclose(ev, e[:14], atol=1e-10)
    V = vecs.cpu().numpy()
    np.testing.assert_allclose(V.conj() @ V.T, np.eye(14), atol=1e-10)
    assert np.abs(Hd @ V.T - V.T * ev).max() < 1e-8
    # d^2 without Coulomb terms (dim 45, levels 15-, 24- and 6-fold): every
    # Krylov space is invariant after three steps
    ut0 = siam.umat_terms(np.zeros((10, 10, 10, 10), complex))
    fock2 = H.fock_words(10, 2)
    A2 = core.build_ham_i(core.Basis.from_words(fock2, 10), ht, ut0)
    e2 = np.linalg.eigvalsh(o.build_ham_i(fock2, ht, ut0).todense())
    ev2, vecs2, _, _ = core.trlan_eigs(A2, 12, ncv=20, maxiter=100, tol=1e-12, seed=4)
    np.testing.assert_allclose(ev2, e2[:12], atol=1e-10)
    W = vecs2.cpu().numpy()
    np.testing.assert_allclose(W.conj() @ W.T, np.eye(12), atol=1e-10)


def test_trlan_matches_single_vector_lanczos(core):
    m = H.small_siam(nbath=6, seed=7)          # 16 orbitals, N=8: dim 12,870
    fi = H.fock_words(m.norb, m.nocc)
    A = core.build_ham_i(core.Basis.from_words(fi, m.norb), m.hop_i, m.coul_i)
    ev1, v1, _ = core.lanczos_eigs(A, 10, maxiter=1500, tol=1e-11, seed=2, deflate=True)
    ev, vecs, nmv, nconv = core.trlan_eigs(A, 10, ncv=30, maxiter=500, tol=1e-12, seed=2)
    np.testing.assert_allclose(ev, ev1, atol=1e-9)
    assert nconv == 10
    for k in range(10):
        y = A.matvec(vecs[k].contiguous())
        assert float((y - ev[k] * vecs[k]).norm()) < 1e-8


def test_interleaved_corehole_layout_is_the_permuted_matrix(core):
    m = H.small_siam(nbath=2, seed=11)
    fi, fn = H.fock_words(m.norb, m.nocc), H.fock_words(m.norb, m.nocc + 1)
    bi, bn = core.Basis.from_words(fi, m.norb), core.Basis.from_words(fn, m.norb)
    p = core.corehole_permutation(len(fn), m.ncore)      # x_inter = x_ref[p]
    pinv = np.empty_like(p)
    pinv[p] = np.arange(len(p))

    def triplets(A, rmap=None, cmap=None):
        rp, c, v = A.download()
        r = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        c = c.astype(np.int64)
        if rmap is not None:
            r = rmap[r]
        if cmap is not None:
            c = cmap[c]
        order = np.lexsort((c, r))
        return r[order], c[order], v[order]

    pairs = [
        (core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n),
         core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved"), True, True),
        (core.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[0]),
         core.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[0], layout="interleaved"), True, False),
        (core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1]),
         core.build_transop_f(bi, bn, m.norb, m.ncore, m.trans_f[1], layout="interleaved"), False, True),
    ]
    for ref, inter, rows_ch, cols_ch in pairs:
        a = triplets(ref, pinv if rows_ch else None, pinv if cols_ch else None)
        b = triplets(inter)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)          # same entries, bit for bit
    # the layout is per call: a plain build afterwards is the reference layout again
    again = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    for x, y in zip(triplets(again), triplets(pairs[0][0])):
        np.testing.assert_array_equal(x, y)
    # XAS Krylov on the permuted pair: same (alpha, beta) up to rounding
    Hn_r, Hn_i = pairs[0][0], pairs[0][1]
    D_r, D_i = pairs[1][0], pairs[1][1]
    gs = _to_dev(core, np.random.default_rng(4).normal(size=len(fi)) + 0j)
    a1, b1, n1 = core.build_krylov(Hn_r, D_r.matvec(gs), 60)
    a2, b2, n2 = core.build_krylov(Hn_i, D_i.matvec(gs), 60)
    # (alpha ~ 100 here: rounding of the reordered sums, relative 1e-12)
    np.testing.assert_allclose(a2, a1, rtol=1e-12)
    np.testing.assert_allclose(b2, b1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())
    assert abs(n1 - n2) < 1e-12 * n1
    with pytest.raises(Exception, match="unknown core-hole layout"):
        core._lib.call("edx_set_corehole_layout", 7, None)


@pytest.mark.parametrize("precision", ["c128", "c64d"])
def test_batched_minres_graph_replay_equals_eager(core, monkeypatch, precision):
    # the iterations after the first are one captured odd/even pair replayed;
    # results must equal the eager launches bit for bit, call after call, with
    # eager work between the calls, for K on both SpMM schedules
    m = H.small_siam(nbath=2, seed=4)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    rng = np.random.default_rng(10)
    for K in (3, 12):
        sig = -20.0 + 3.0 * np.arange(K) + 0.4j
        B = _to_dev(core, rng.normal(size=(K, Hn.n)) + 1j * rng.normal(size=(K, Hn.n)))
        monkeypatch.setenv("EDX_NO_GRAPH", "1")
        Xe, ite, rese = core.minres_shifted_batch(Hn, sig, B, tol=1e-10, maxit=777,
                                                  precision=precision)
        monkeypatch.delenv("EDX_NO_GRAPH")
        for rep in range(3):
            Xg, itg, resg = core.minres_shifted_batch(Hn, sig, B, tol=1e-10, maxit=777,
                                                      precision=precision)
            assert np.array_equal(itg, ite)
            assert np.array_equal(resg, rese)
            assert bool((Xg == Xe).all())
            _ = Hn.matvec(B[0].contiguous())     # eager work between the replayed calls
        # an odd iteration cap ends with one eager iteration after the replays
        Xo, ito, _ = core.minres_shifted_batch(Hn, sig, B, tol=1e-14, maxit=8,
                                               precision=precision)
        monkeypatch.setenv("EDX_NO_GRAPH", "1")
        Xo2, ito2, _ = core.minres_shifted_batch(Hn, sig, B, tol=1e-14, maxit=8,
                                                 precision=precision)
        monkeypatch.delenv("EDX_NO_GRAPH")
        assert np.array_equal(ito, ito2) and bool((Xo == Xo2).all())


@pytest.mark.parametrize("code", [143, 144, 184])
def test_pipelined_spmv_schedules_match_oracle(core, code):
    # the pipelined grid-stride schedules (chosen by the autotuner on large
    # matrices) forced on a small one, in every value form, with rows longer
    # than the U * G nonzeros a lane prefetches (dense random terms)
    from edrixs_amd import siam
    m = H.small_siam(nbath=4, seed=2)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    rng = np.random.default_rng(6)
    n12, N12 = 12, 6
    ht = siam.emat_terms(H.random_hermitian_emat(n12, rng, density=1.0))
    ut = siam.umat_terms(H.random_hermitian_umat(n12, rng, nnz=400))
    f12 = H.fock_words(n12, N12)
    B = core.build_ham_i(core.Basis.from_words(f12, n12), ht, ut)
    RB = o.build_ham_i(f12, ht, ut)
    assert np.diff(RB.rowptr).max() > 64          # beyond the prefetched window
    for Amat, Rmat in ((A, R), (B, RB)):
        x = rng.normal(size=Amat.n) + 1j * rng.normal(size=Amat.n)
        y0 = rng.normal(size=Amat.m) + 1j * rng.normal(size=Amat.m)
        ref = o.matvec(Rmat, x, y0)
        scale = np.abs(Rmat.vals).max() * np.abs(x).max() * np.diff(Rmat.rowptr).max()
        for form, tol in (("c128", 1e-14), ("c128i", 1e-14), ("c64d", 2e-6), ("c64", 2e-6)):
            Amat.prepare(form)
            Amat.spmv_lanes(form, set=code)
            assert Amat.spmv_lanes(form) == code
            y = Amat.spmv(_to_dev(core, x), _to_dev(core, y0), precision=form).cpu().numpy()
            assert np.abs(y - ref).max() < tol * scale, (form, code)
    # a Krylov recurrence (fused epilogue and dot partials) on the forced schedule
    A.spmv_lanes("c128", set=8)
    v0 = _to_dev(core, rng.normal(size=A.n) + 0j)
    a1, b1, _ = core.build_krylov(A, v0, 40)
    A.spmv_lanes("c128", set=code)
    a2, b2, _ = core.build_krylov(A, v0, 40)
    np.testing.assert_allclose(a2, a1, rtol=1e-12)
    np.testing.assert_allclose(b2, b1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())
    with pytest.raises(Exception, match="SpMV schedule"):
        A.spmv_lanes("c128", set=146)


def test_packed_indexed_schedule_145(core):
    # the pipelined product built for five workgroups per CU, which only the
    # packed indexed form (column | value number in one word) may take
    m = H.small_siam(nbath=4, seed=2)
    fn = H.fock_words(m.norb, m.nocc + 1)
    bn = core.Basis.from_words(fn, m.norb)
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    R = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
    rng = np.random.default_rng(31)
    x = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    y0 = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    ref = o.matvec(R, x, y0)
    scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
    A.prepare("c128i")
    A.spmv_lanes("c128i", set=145)
    y = A.spmv(_to_dev(core, x), _to_dev(core, y0), precision="c128i").cpu().numpy()
    assert np.abs(y - ref).max() < 1e-14 * scale
    a1, b1, _ = core.build_krylov(A, _to_dev(core, x), 40, precision="c128i")
    A.spmv_lanes("c128i", set=144)
    a2, b2, _ = core.build_krylov(A, _to_dev(core, x), 40, precision="c128i")
    np.testing.assert_allclose(a1, a2, rtol=1e-12)
    np.testing.assert_allclose(b1, b2, rtol=1e-12, atol=1e-12 * np.abs(a2).max())
    A.spmv_lanes("c128", set=145)
    with pytest.raises(Exception, match="packed indexed form"):
        A.spmv(_to_dev(core, x), _to_dev(core, y0), precision="c128")


@pytest.mark.parametrize("precision", ["c128", "c128i", "c64d"])
def test_pipelined_rhs_parallel_spmm_matches_other_schedules(core, monkeypatch, precision):
    # the pipelined RHS-parallel SpMM (chosen by the autotuner for large
    # matrices) forced on a small one: a 16-vector Krylov batch and a 12-vector
    # shifted MINRES against the default schedules, rows longer than the 32
    # nonzeros a lane group prefetches included
    m, fn, bn, A, R = _siam_hn(core)
    A.prepare(precision)
    assert np.diff(R.rowptr).max() > 32
    rng = np.random.default_rng(21)
    V = _to_dev(core, rng.normal(size=(16, A.n)) + 1j * rng.normal(size=(16, A.n)))
    monkeypatch.delenv("EDX_FORCE_BSCHED", raising=False)
    a0, b0, n0 = core.build_krylov_batch(A, V, 60, precision=precision)
    monkeypatch.setenv("EDX_FORCE_BSCHED", "2:4")
    a1, b1, n1 = core.build_krylov_batch(A, V, 60, precision=precision)
    if precision == "c128i":
        # schedule 4 (packed words, value table in LDS) makes the same sums in
        # the same order as schedule 2 on the same grid
        monkeypatch.setenv("EDX_FORCE_BSCHED", "4:4")
        a4, b4, n4 = core.build_krylov_batch(A, V, 60, precision=precision)
        monkeypatch.setenv("EDX_FORCE_BSCHED", "2:4")
        for k in range(16):
            assert np.array_equal(a4[k], a1[k]) and np.array_equal(b4[k], b1[k])
        assert np.array_equal(n4, n1)
    # 60 steps of the plain recurrence amplify the summation-order rounding
    for k in range(16):
        np.testing.assert_allclose(a1[k], a0[k], rtol=1e-10)
        np.testing.assert_allclose(b1[k], b0[k], rtol=1e-10, atol=1e-10 * np.abs(a0[k]).max())
    sig = -20.0 + 2.5 * np.arange(12) + 0.4j
    B = V[:12].contiguous()
    X1, it1, _ = core.minres_shifted_batch(A, sig, B, tol=1e-10, maxit=600, precision=precision)
    monkeypatch.delenv("EDX_FORCE_BSCHED")
    X0, it0, _ = core.minres_shifted_batch(A, sig, B, tol=1e-10, maxit=600, precision=precision)
    assert np.abs(it1.astype(int) - it0.astype(int)).max() <= 2
    err = (X1 - X0).abs().max().item() / X0.abs().max().item()
    assert err < 1e-8


def test_gs_sector_search_matches_dense_sectors(core):
    # do_ed=0 sector search on a 12-orbital impurity model with the bath
    # shifted down (the ground sector is not the half-filled one), against the
    # dense ground energies of every sector
    from edrixs_amd import solvers, siam
    m = siam.SiamModel(nbath=2, seed=3, bath_range=(-6.0, -4.0))
    hop, coul = m.hop_i, m.coul_i
    e_dense = {}
    for n in range(0, m.norb + 1):
        fock = H.fock_words(m.norb, n)
        e_dense[n] = float(np.linalg.eigvalsh(o.build_ham_i(fock, hop, coul).todense())[0])
    n_best = min(e_dense, key=e_dense.get)
    import io
    logs = {}
    for conc in (True, False):
        logs[conc] = io.StringIO()
        out = solvers.search_gs_sector(m.norb, hop, coul, imp_orbs=range(10), log=logs[conc],
                                       concurrent=conc)
        if conc:
            n_gs, e_gs, table = out
        else:
            # the sectors evaluated two at a time on their own streams give the
            # sequential walk's result and log exactly
            assert out == (n_gs, e_gs, table)
            assert logs[True].getvalue() == logs[False].getvalue()
    assert n_gs == n_best and n_gs != m.norb // 2
    assert abs(e_gs - e_dense[n_best]) < 1e-9
    for n, e, occ in table:
        assert abs(e - e_dense[n]) < 1e-9
        assert -1e-9 <= occ <= 10 + 1e-9
    assert table[0][0] == n_gs and [t[1] for t in table] == sorted(t[1] for t in table)


def test_slot_factored_form_matches_oracle(core, monkeypatch):
    # H_n in the interleaved core-hole layout: the off-diagonal entries that keep
    # the core slot come from valence-only terms and are the same for the six
    # slots of a valence row, so the factored form stores them once per valence
    # row (H_v); the entries that move the core hole go into typed blocks
    # (valence column, slot shift), and valence rows whose slot rows differ stay
    # whole per row (R).  The split is checked against counts taken from the
    # oracle's CSR (reference layout), and the factored product (schedule 206)
    # against the oracle's matvec: on the plain model, and on one with
    # core-valence terms that keep the slot, so some valence rows differ
    # between slots; each also built without typed blocks (slot-changing
    # entries per row).
    from edrixs_amd import siam
    base = H.small_siam(nbath=4, seed=2)
    nv, nc = base.norb, base.ncore
    u = base.umat_n.copy()
    for pc, d1, d2 in ((nv + 0, 0, 3), (nv + 4, 1, 6)):   # c+_p c+_d1 c_d2 c_p and partner
        u[pc, d1, d2, pc] += 0.3 + 0.2j
        u[pc, d2, d1, pc] += 0.3 - 0.2j
    variants = [(base.coul_n, True), (siam.umat_terms(u), False)]
    fn = H.fock_words(nv, base.nocc + 1)
    dv = len(fn)
    bn = core.Basis.from_words(fn, nv)
    p = core.corehole_permutation(dv, nc)            # x_inter = x_ref[p]
    pinv = np.empty_like(p)
    pinv[p] = np.arange(len(p))
    rng = np.random.default_rng(23)
    for coul, all_same in variants:
        R = o.build_ham_n(fn, nv, nc, base.hop_n, coul)
        # expected split from the oracle's entries: per valence row, the
        # slot-keeping off-diagonal (valence column, value) lists of the six
        # slots, and the slot-moving entries by (valence column, shift)
        lists = [[[] for _ in range(nc)] for _ in range(dv)]
        moves = [dict() for _ in range(dv)]
        n_off = 0
        for r in range(R.m):
            s, v = divmod(r, dv)
            for j in range(R.rowptr[r], R.rowptr[r + 1]):
                c = int(R.cols[j])
                if c == r:
                    continue
                n_off += 1
                s2, v2 = divmod(c, dv)
                if s2 == s:
                    lists[v][s].append((v2, R.vals[j].real, R.vals[j].imag))
                else:
                    bits = np.array([R.vals[j].real, R.vals[j].imag]).view(np.int64)
                    moves[v].setdefault((v2, s2 - s), {})[s] = (int(bits[0]), int(bits[1]))
        same = [all(lists[v][s] == lists[v][0] for s in range(1, nc)) for v in range(dv)]
        hv = sum(len(lists[v][0]) for v in range(dv) if same[v])
        assert all(same) == all_same and any(same)
        n_off_row = np.diff(R.rowptr) - 1                  # every row has its diagonal
        r_rows = sum(int(n_off_row[s_ * dv + v]) for s_ in range(nc) for v in range(dv)
                     if not same[v])
        blocks = sum(len(moves[v]) for v in range(dv) if same[v])
        types = {(d, tuple(sorted(ent.items()))) for v in range(dv) if same[v]
                 for (_, d), ent in moves[v].items()}
        x = rng.normal(size=dv * nc) + 1j * rng.normal(size=dv * nc)
        y0 = rng.normal(size=x.size) + 1j * rng.normal(size=x.size)
        ref = o.matvec(R, x[pinv], y0[pinv])[p]
        scale = np.abs(R.vals).max() * np.abs(x).max() * np.diff(R.rowptr).max()
        for with_blocks in (True, False):
            if with_blocks:
                monkeypatch.delenv("EDX_NO_FAC_BLOCKS", raising=False)
            else:
                monkeypatch.setenv("EDX_NO_FAC_BLOCKS", "1")
            A = core.build_ham_n(bn, nv, nc, base.hop_n, coul, layout="interleaved")
            A.prepare("c128i")
            f = A.factored()
            assert f is not None and f[0] == nc and f[1] == hv, (f, hv)
            if with_blocks:
                assert f[2] == r_rows and f[3] == blocks and f[4] == len(types), \
                    (f, r_rows, blocks, len(types))
                assert blocks > 0 and 0 < f[4] <= 1024
            else:
                assert f[2] == n_off - nc * hv and f[3] == 0 and f[4] == 0, f
                # rows of more than 8 remainder entries exist, so the loop past
                # the 8 words a lane prefetches runs
                r_len = [n_off_row[s_ * dv + v] - (len(lists[v][s_]) if same[v] else 0)
                         for s_ in range(nc) for v in range(dv)]
                assert max(r_len) > 8 and sum(r_len) == f[2]
            A.spmv_lanes("c128i", set=206)
            assert A.spmv_lanes("c128i") == 206
            y = A.spmv(_to_dev(core, x), _to_dev(core, y0), precision="c128i").cpu().numpy()
            assert np.abs(y - ref).max() < 1e-14 * scale, with_blocks
            # a Krylov recurrence (fused epilogue, dot partials) on the factored product
            a1, b1, _ = core.build_krylov(A, _to_dev(core, x), 40, precision="c128i")
            A.spmv_lanes("c128i", set=144)
            a2, b2, _ = core.build_krylov(A, _to_dev(core, x), 40, precision="c128i")
            np.testing.assert_allclose(a1, a2, rtol=1e-12)
            np.testing.assert_allclose(b1, b2, rtol=1e-12, atol=1e-12 * np.abs(a2).max())
    monkeypatch.delenv("EDX_NO_FAC_BLOCKS", raising=False)
    # matrices outside the form: none is made, and the schedule refuses them
    Aref = core.build_ham_n(bn, nv, nc, base.hop_n, base.coul_n)      # reference layout
    Aref.prepare("c128i")
    assert Aref.factored() is None
    Aref.spmv_lanes("c128i", set=206)
    with pytest.raises(Exception, match="slot-factored form"):
        Aref.matvec(_to_dev(core, np.ones(Aref.n, np.complex128)), precision="c128i")


def _read_eigvec(path, ndim):
    """eigvec.N: Fortran unformatted records (real(8) eigenvalue; complex(8) vector)."""
    import struct
    raw = open(path, "rb").read()
    e = struct.unpack("<d", raw[4:12])[0]
    v = np.frombuffer(raw[16 + 4:16 + 4 + 16 * ndim], dtype=np.complex128).copy()
    return e, v


def test_opavg_of_an_operator_other_than_h_matches_oracle(core, tmp_path):
    """opavg_fsolver (src/opavg_driver.f90) with O != H: a non-Hermitian set of
    one- and two-body terms in hopping_i.in / coulomb_i.in, averaged over the
    two eigenvectors an ED run wrote; opavg.dat against the oracle, line by line
    in the (1i5, 4f20.10) format of src/io.f90:638-659."""
    import re
    from edrixs_amd import fedrixs, iostream, siam, solvers
    m = H.small_siam(nbath=2, seed=11)
    wd = str(tmp_path)
    solvers.ed_siam(m, wd, neval=2, nvector=2, eigval_tol=1e-12, min_ndim=10)
    rng = np.random.default_rng(21)
    n = m.norb
    emat = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * (rng.random((n, n)) < 0.3)
    umat = np.zeros((n,) * 4, np.complex128)
    for _ in range(60):
        i, j, k, l = rng.integers(0, n, size=4)
        if i != j and k != l:
            umat[i, j, k, l] += rng.normal() + 1j * rng.normal()
    et, ut = siam.emat_terms(emat), siam.umat_terms(umat)
    fi = H.fock_words(n, m.nocc)
    O = o.build_ham_i(fi, et, ut)
    cwd = os.getcwd()
    os.chdir(wd)
    try:
        iostream.write_terms(et, "hopping_i.in")
        iostream.write_terms(ut, "coulomb_i.in")
        iostream.write_config(num_val_orbs=n, num_gs=2)
        fedrixs.opavg_fsolver(0, 0, 1)
        lines = open("opavg.dat").read().splitlines()
        vecs = [_read_eigvec("eigvec.%d" % k, len(fi)) for k in (1, 2)]
    finally:
        os.chdir(cwd)
    assert len(lines) == 2
    ref = [o.opavg(O, v) for _, v in vecs]
    assert all(abs(a.imag) > 1e-3 for a in ref)       # a genuinely complex average
    want = o.opavg_lines([e for e, _ in vecs], ref)
    for got, exp in zip(lines, want):
        assert re.fullmatch(r"[ \d]{4}\d( {0,19}-?\d+\.\d{10}){3}", got) and len(got) == 65, got
        assert got[:5] == exp[:5]
        g, x = np.array(got[5:].split(), float), np.array(exp[5:].split(), float)
        np.testing.assert_allclose(g, x, atol=2e-10)


def test_two_ground_states_through_the_xas_and_rixs_drivers(core, tmp_path):
    """num_gs = 2 over a degenerate ground level: an atomic d7 shell with
    spin-orbit coupling and a cubic field (no bath, time-reversal symmetric, an
    odd electron count) has a Kramers doublet lowest, 0.053 eV below a quartet.
    Each eigenvector of the doublet is arbitrary within it, but the
    Boltzmann-weighted sum of the two spectra (edrixs/plot_spectrum.py:43-60)
    is not: XAS and RIXS through the file drivers against dense Lorentzian sums
    over the oracle's operators."""
    from edrixs_amd import solvers
    m = H.small_siam(nbath=0, seed=11, nocc=7)         # norb 10, N 7: dims 120 / 6*45
    wd = str(tmp_path)
    evals, _ = solvers.ed_siam(m, wd, neval=2, nvector=2, eigval_tol=1e-12, min_ndim=10)
    fi, fn = H.fock_words(m.norb, m.nocc), H.fock_words(m.norb, m.nocc + 1)
    Hi = o.build_ham_i(fi, m.hop_i, m.coul_i).todense()
    e, V = np.linalg.eigh(Hi)
    assert e[1] - e[0] < 1e-9 and e[2] - e[1] > 1e-3
    np.testing.assert_allclose(evals, e[:2], atol=2e-10)
    Hn = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n).todense()
    en, Vn = np.linalg.eigh(Hn)
    T = 300.0
    om = np.linspace(-10, 20, 400)
    xas, _ = solvers.xas_siam(m, om, wd, gamma_c=0.5, num_gs=2, nkryl=3000, temperature=T)
    ref = np.zeros(len(om))
    for k in range(3):
        D = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[k]).todense()
        for g in range(2):                              # equal weights: a degenerate level
            amp = np.abs(Vn.conj().T @ (D @ V[:, g])) ** 2
            ref += 0.5 * (-1 / np.pi) * np.imag(
                (amp[None, :] / (om[:, None] + 0.5j + e[0] - en[None, :])).sum(1))
    ref /= 3
    assert np.linalg.norm(xas - ref) / np.linalg.norm(ref) < 1e-6
    # RIXS: one incident energy, one polarization pair, both ground states
    w_in, el = 2.0, np.linspace(-1, 10, 300)
    pols = [([1, 0, 0], [0, 1, 0])]
    rixs, _ = solvers.rixs_siam(m, [w_in], el, pols, wd, gamma_c=0.5, gamma_f=0.1, num_gs=2,
                                nkryl=2000, linsys_tol=1e-12, linsys_max=4000, temperature=T)
    ti, _ = m.pol_trans([1, 0, 0])
    _, tf = m.pol_trans([0, 1, 0])
    Di = o.build_transop_i(fn, fi, m.norb, m.ncore, ti).todense()
    Df = o.build_transop_f(fi, fn, m.norb, m.ncore, tf).todense()
    rref = np.zeros(len(el))
    for g in range(2):
        x = np.linalg.solve((w_in + e[0] + 0.5j) * np.eye(len(en)) - Hn, Di @ V[:, g])
        amp = np.abs(V.conj().T @ (Df @ x)) ** 2
        rref += 0.5 * (-1 / np.pi) * np.imag(
            (amp[None, :] / (el[:, None] + 0.1j + e[0] - e[None, :])).sum(1))
    got = rixs[0, :, 0]
    assert np.linalg.norm(got - rref) / np.linalg.norm(rref) < 1e-6


def test_group_factored_member_matches_oracle_and_is_mix_independent(core, monkeypatch):
    # an interleaved-layout H_n in an indexed (c128i) group takes the
    # slot-factored product; its (alpha, beta) against the oracle's recurrence
    # from the same start vector (reference layout: the vector permuted), and
    # the same bits alone, in a changing mix and without graph replay
    m, fn, bn, _, R = _siam_hn(core, nbath=4, seed=2)
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved")
    A.prepare("c128i")
    assert A.factored() is not None
    p = core.corehole_permutation(len(fn), m.ncore)      # x_inter = x_ref[p]
    m2 = H.small_siam(nbath=3, seed=5)
    f2 = H.fock_words(m2.norb, m2.nocc)
    A2 = core.build_ham_i(core.Basis.from_words(f2, m2.norb), m2.hop_i, m2.coul_i)
    rng = np.random.default_rng(47)
    vref = rng.normal(size=A.n) + 1j * rng.normal(size=A.n)
    va = _to_dev(core, vref[p])
    vb = _to_dev(core, rng.normal(size=A2.n) + 1j * rng.normal(size=A2.n))
    g1 = core.KrylovGroup("c128i", max_steps=200)
    a = g1.add(A, va)
    g1.advance(40)
    al1, be1, n1, run1 = g1.coefficients(a)
    g1.close()
    assert len(al1) == 40 and run1
    ra, rb, rn = o.build_krylov(R, vref, 40)
    np.testing.assert_allclose(al1[:20], ra[:20], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(be1[:20], rb[:20], rtol=1e-9, atol=1e-9)
    assert abs(n1 - rn) <= 1e-12 * rn

    def mix():
        g = core.KrylovGroup("c128i", max_steps=200)
        b = g.add(A2, vb)
        g.advance(5)
        a2 = g.add(A, va)
        c = g.add(A2, vb)
        g.advance(17)
        g.remove(b)
        g.advance(23)
        out = g.coefficients(a2), g.coefficients(c)
        g.close()
        return out

    (al2, be2, n2, _), cg = mix()
    assert np.array_equal(al2, al1) and np.array_equal(be2, be1) and n2 == n1
    monkeypatch.setenv("EDX_NO_GRAPH", "1")
    (al3, be3, n3, _), ce = mix()
    assert np.array_equal(al3, al1) and np.array_equal(be3, be1) and n3 == n1
    assert np.array_equal(ce[0], cg[0]) and np.array_equal(ce[1], cg[1])


@pytest.mark.parametrize("K", [2, 3, 4])
def test_factored_multi_rhs_product_matches_single_vectors(core, monkeypatch, K):
    # the factored product over K interleaved vectors (forced schedule 3:4 of
    # the multi-RHS products) in a batched Krylov recurrence: each vector's
    # (alpha, beta) against the single-vector recurrence and the oracle's
    m, fn, bn, _, R = _siam_hn(core, nbath=4, seed=2)
    A = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved")
    A.prepare("c128i")
    assert A.factored() is not None
    p = core.corehole_permutation(len(fn), m.ncore)
    rng = np.random.default_rng(53 + K)
    V = rng.normal(size=(K, A.n)) + 1j * rng.normal(size=(K, A.n))
    monkeypatch.setenv("EDX_FORCE_BSCHED", "3:4")
    al, be, nr = core.build_krylov_batch(A, _to_dev(core, V), 50, precision="c128i")
    monkeypatch.delenv("EDX_FORCE_BSCHED")
    for k in range(K):
        a1, b1, n1 = core.build_krylov(A, _to_dev(core, V[k]), 50, precision="c128i")
        np.testing.assert_allclose(al[k], a1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())
        np.testing.assert_allclose(be[k], b1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())
        assert abs(nr[k] - n1) <= 1e-13 * n1
        vref = np.empty_like(V[k])
        vref[p] = V[k]                                   # x_inter = x_ref[p]
        ra, rb, rn = o.build_krylov(R, vref, 50)
        np.testing.assert_allclose(al[k][:20], ra[:20], rtol=1e-9, atol=1e-9)
    # the schedule refuses a matrix without the form
    B = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    B.prepare("c128i")
    monkeypatch.setenv("EDX_FORCE_BSCHED", "3:4")
    with pytest.raises(Exception, match="slot-factored"):
        core.build_krylov_batch(B, _to_dev(core, V), 5, precision="c128i")


def test_profile_counts_every_launch_beyond_the_event_ring(core):
    # more products than the profiler's ring of 16,384 event pairs holds: every
    # launch is counted with its bytes, in each precision and kernel apart
    m, fn, bn, A, R = _siam_hn(core, nbath=2, seed=1)
    A.prepare("c128i")
    x = _to_dev(core, np.random.default_rng(3).normal(size=A.n) + 0.25j)
    y = core.torch().zeros(A.m, dtype=core.torch().complex128, device="cuda")
    nrep = 20000
    core.profile(True)
    for _ in range(nrep):
        A.spmv(x, y, precision="c128")
    for _ in range(7):
        A.spmv(x, y, precision="c128i")
    n128, ms128, by128 = core.profile_read("c128", "k_spmv")
    n128i, _, _ = core.profile_read("c128i", None)
    core.profile(False)
    assert n128 == nrep and n128i == 7
    assert ms128 > 0 and by128 > 0
    # bytes of one product: values and columns, row pointers, x, y read and written
    one = A.nnz * 20.0 + (A.m + 1) * 8.0 + A.n * 16.0 + A.m * 32.0
    assert abs(by128 / nrep - one) <= 1e-9 * one
    # a cleared profile starts from zero
    core.profile(True)
    assert core.profile_read("c128", "k_spmv")[0] == 0
    core.profile(False)
