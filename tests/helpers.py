"""Shared test utilities: small random Hermitian models and basis helpers."""
import itertools
import math

import numpy as np

from edrixs_amd import siam


def fock_words(norb, nocc):
    return np.array(sorted(sum(1 << p for p in c)
                           for c in itertools.combinations(range(norb), nocc)), dtype=np.int64)


def random_hermitian_emat(n, rng, density=0.5, complex_vals=True):
    a = rng.normal(size=(n, n))
    if complex_vals:
        a = a + 1j * rng.normal(size=(n, n))
    a = a * (rng.random((n, n)) < density)
    return (a + a.conj().T) / 2


def random_hermitian_umat(n, rng, nnz=60, complex_vals=True):
    """U with U_ijkl = conj(U_lkji) so sum U c+c+cc is Hermitian."""
    u = np.zeros((n,) * 4, dtype=np.complex128)
    for _ in range(nnz):
        i, j, k, l = rng.integers(0, n, size=4)
        if i == j or k == l:
            continue
        v = rng.normal() + (1j * rng.normal() if complex_vals else 0)
        u[i, j, k, l] += v
        u[l, k, j, i] += np.conj(v)
    return u


def core_words(num_core_orbs):
    """Core configurations with one hole, in the reference order (hole at bit nc-k)."""
    nc = num_core_orbs
    return [(1 << nc) - 1 - (1 << (nc - k)) for k in range(1, nc + 1)]


def corehole_words(fock_val, num_val_orbs, num_core_orbs):
    """Full words of the intermediate space, index (k-1)*dim_val + v."""
    return np.array([(c << num_val_orbs) | int(v) for c in core_words(num_core_orbs)
                     for v in fock_val], dtype=np.int64)


def fullcore_words(fock_val, num_val_orbs, num_core_orbs):
    full = (1 << num_core_orbs) - 1
    return np.array([(full << num_val_orbs) | int(v) for v in fock_val], dtype=np.int64)


def small_siam(nbath=2, seed=1, nocc=None):
    """A genuine d + bath + 2p-core model small enough for dense checks."""
    return siam.SiamModel(nbath=nbath, seed=seed, nocc=nocc)


def ncomb(n, k):
    return math.comb(n, k)


def herm_c64_values(rowptr, cols, vals, n):
    """Expected complex64 copy of a square CSR: an off-diagonal entry whose
    transposed partner is its conjugate to 1e-13 relative is rounded from the
    FP64 Hermitian average (a_rc + conj(a_cr)) / 2, so the stored copy stays
    exactly Hermitian; a diagonal entry whose imaginary part is rounding
    residue (<= 1e-13 of the real part) becomes real; all other entries round
    independently."""
    rowptr = np.asarray(rowptr, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.complex128)
    rows = np.repeat(np.arange(len(rowptr) - 1, dtype=np.int64), np.diff(rowptr))
    keys = rows * n + cols                     # ascending: rows, then sorted columns
    pk = cols * n + rows
    pos = np.clip(np.searchsorted(keys, pk), 0, max(len(keys) - 1, 0))
    found = (keys[pos] == pk) & (cols != rows) if len(keys) else np.zeros(0, bool)
    a, b = vals, vals[pos]
    close = (np.abs(a.real - b.real) + np.abs(a.imag + b.imag)
             <= 1e-13 * (np.abs(a.real) + np.abs(a.imag) + np.abs(b.real) + np.abs(b.imag)))
    use = found & close
    out = vals.copy()
    out[use] = (a.real[use] + b.real[use]) * 0.5 + 1j * ((a.imag[use] - b.imag[use]) * 0.5)
    diag = (cols == rows) & (np.abs(vals.imag) <= 1e-13 * np.abs(vals.real))
    out[diag] = vals.real[diag]
    return out.astype(np.complex64)
