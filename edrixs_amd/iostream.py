"""Text formats at the solver boundary.

The Fortran core and its Python callers exchange everything through files in
the working directory; these writers and readers produce and parse the same
formats (the doubles the Fortran side reads are defined by the text):

* ``write_emat`` / ``write_umat``  edrixs/iostream.py:165-235
* ``write_config``                 edrixs/iostream.py:238-276
* ``read_poles_from_file``         edrixs/iostream.py:279-323
* ``write_fock_dec_by_N``          edrixs/fock_basis.py:398-451
"""
import numpy as np

_SPACE = "    "


def _term_lines(idx, vals, nidx, fmt_int, fmt_float):
    fmt = (fmt_int + _SPACE) * nidx + (fmt_float + _SPACE) * 2 + "\n"
    return [fmt.format(*(int(i) + 1 for i in ii), v.real, v.imag) for ii, v in zip(idx, vals)]


def write_emat(emat, fname, tol=1e-12, fmt_int="{:10d}", fmt_float="{:.15f}"):
    """Nonzeros (|v| > tol) of a rank-2 matrix in C order, 1-based indices."""
    emat = np.asarray(emat)
    a1, a2 = np.nonzero(np.abs(emat) > tol)
    with open(fname, "w") as f:
        if len(a1) == 0:
            f.write("{:10d}\n".format(1))
            f.write(((fmt_int + _SPACE) * 2 + (fmt_float + _SPACE) * 2 + "\n")
                    .format(1, 1, 0.0, 0.0))
            return
        f.write("{:20d}\n".format(len(a1)))
        f.writelines(_term_lines(zip(a1, a2), emat[a1, a2], 2, fmt_int, fmt_float))


def write_umat(umat, fname, tol=1e-12, fmt_int="{:10d}", fmt_float="{:.15f}"):
    """Nonzeros of a rank-4 tensor in C order, 1-based indices."""
    umat = np.asarray(umat)
    idx = np.nonzero(np.abs(umat) > tol)
    with open(fname, "w") as f:
        if len(idx[0]) == 0:
            f.write("{:10d}\n".format(1))
            f.write(((fmt_int + _SPACE) * 4 + (fmt_float + _SPACE) * 2 + "\n")
                    .format(1, 1, 1, 1, 0.0, 0.0))
            return
        f.write("{:20d}\n".format(len(idx[0])))
        f.writelines(_term_lines(zip(*idx), umat[idx], 4, fmt_int, fmt_float))


def write_terms(terms, fname):
    """Write an already-extracted term record array (fields ind1.., re, im)."""
    names = [n for n in terms.dtype.names if n.startswith("ind")]
    nidx = len(names)
    with open(fname, "w") as f:
        f.write("{:20d}\n".format(len(terms)))
        fmt = ("{:10d}" + _SPACE) * nidx + ("{:.15f}" + _SPACE) * 2 + "\n"
        for t in terms:
            f.write(fmt.format(*(int(t[n]) for n in names), float(t["re"]), float(t["im"])))


def write_config(directory=".", ed_solver=1, num_val_orbs=2, num_core_orbs=2, neval=1,
                 nvector=1, ncv=1, idump=True, num_gs=1, maxiter=500, linsys_max=1000,
                 min_ndim=1000, nkryl=500, eigval_tol=1e-8, linsys_tol=1e-10, omega_in=0.0,
                 gamma_in=0.1):
    """The &control namelist read by the solvers (same keys and defaults)."""
    items = [("ed_solver", ed_solver), ("num_val_orbs", num_val_orbs),
             ("num_core_orbs", num_core_orbs), ("neval", neval), ("nvector", nvector),
             ("ncv", ncv), ("idump", ".true." if idump else ".false."), ("num_gs", num_gs),
             ("maxiter", maxiter), ("linsys_max", linsys_max), ("min_ndim", min_ndim),
             ("nkryl", nkryl), ("eigval_tol", eigval_tol), ("linsys_tol", linsys_tol),
             ("omega_in", omega_in), ("gamma_in", gamma_in)]
    with open(directory + "/config.in", "w") as f:
        f.write("&control\n")
        for k, v in items:
            f.write("%s=%s\n" % (k, v))
        f.write("&end\n")


def read_poles_from_file(file_list):
    """Pole files (#number_of_poles / #enegry / #normalization, then i alpha beta)."""
    poles = {"npoles": [], "eigval": [], "norm": [], "alpha": [], "beta": []}
    for fname in file_list:
        with open(fname) as f:
            neff = int(f.readline().split()[1])
            eigval = float(f.readline().split()[1])
            norm = float(f.readline().split()[1])
            alpha, beta = [], []
            for _ in range(neff):
                tok = f.readline().split()
                alpha.append(float(tok[1]))
                beta.append(float(tok[2]))
        poles["npoles"].append(neff)
        poles["eigval"].append(eigval)
        poles["norm"].append(norm)
        poles["alpha"].append(alpha)
        poles["beta"].append(beta)
    return poles


def fock_words(norb, nocc):
    """Ascending norb-bit words with popcount nocc (colex enumeration on the host)."""
    from math import comb
    n = comb(norb, nocc)
    out = np.zeros(n, dtype=np.int64)
    if n == 0:
        return out
    # iterate combinations in colex order with Gosper's hack (ascending words)
    if nocc == 0:
        return out
    w = (1 << nocc) - 1
    for i in range(n):
        out[i] = w
        c = w & -w
        r = w + c
        w = (((r ^ w) >> 2) // c) | r
    return out


def write_fock_dec_by_N(N, r, fname="fock_i.in"):
    """Count, then the ascending words of N orbitals with r electrons."""
    words = fock_words(N, r)
    with open(fname, "w") as f:
        f.write("%d\n" % len(words))
        np.savetxt(f, words, fmt="%d")
    return len(words)


def read_terms(fname, rank):
    """Read a hopping (rank 2) or Coulomb (rank 4) term file into a record array."""
    from .core import T2, T4
    with open(fname) as f:
        n = int(f.readline().split()[0])
        data = np.loadtxt(f, ndmin=2, max_rows=n) if n else np.zeros((0, rank + 2))
    dt = T2 if rank == 2 else T4
    out = np.zeros(n, dtype=dt)
    for k in range(rank):
        out["ind%d" % (k + 1)] = data[:, k].astype(np.int64)
    out["re"], out["im"] = data[:, rank], data[:, rank + 1]
    return out
