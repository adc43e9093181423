"""Synthetic single-impurity Anderson models with a 2p core shell.

This is the input side of the ED core: it produces the one-body matrix
``emat``, the two-body tensor ``umat`` and the dipole transition operators in
the conventions the reference's file writers and Fortran core consume:

* spin-orbital order inside a shell with angular momentum l is
  ``(m + l) * 2 + sigma`` for m = -l..l, sigma = 0 (up), 1 (down)
  (edrixs/coulomb_utensor.py:185-192);
* ``H = sum E_ij c+_i c_j + sum U_ijkl c+_i c+_j c_k c_l`` where U is the
  Slater-integral tensor with the last two indices swapped and the 1/2 factor
  of edrixs/coulomb_utensor.py:213-229;
* Gaunt coefficients ``c^k(l1 m1, l2 m2)`` as in edrixs/coulomb_utensor.py:75-78,
  evaluated here from the Racah formula for Wigner 3j symbols;
* spin-orbit matrix zeta * l.s in complex harmonics (edrixs/soc.py:6-107);
* cubic crystal field (edrixs/angular_momentum.py:597-625) with the real to
  complex harmonic transform of edrixs/basis_transform.py:109-200;
* dipole operators x, y, z from Clebsch-Gordan coefficients
  (edrixs/photon_transition.py:12-34).

Term values pass through the ``'{:.15f}'`` text rendering the reference's
``write_emat``/``write_umat`` apply (edrixs/iostream.py:189,225), so a model
built in memory gives the same doubles as one read back from files.
"""
import math
from fractions import Fraction

import numpy as np

# ------------------------------------------------------------- angular momentum


def _fact(n):
    return math.factorial(n)


def wigner3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol for integer arguments (Racah's closed formula)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    tri = Fraction(_fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3),
                   _fact(j1 + j2 + j3 + 1))
    pre = tri * (_fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2)
                 * _fact(j3 + m3) * _fact(j3 - m3))
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (_fact(k) * _fact(j3 - j2 + k + m1) * _fact(j3 - j1 + k - m2)
               * _fact(j1 + j2 - j3 - k) * _fact(j1 - k - m1) * _fact(j2 - k + m2))
        s += Fraction((-1) ** k, den)
    sign = -1 if (j1 - j2 - m3) % 2 else 1
    # sign * s * sqrt(pre), rounded once from the exact square
    sq = s * s * pre
    val = math.sqrt(sq.numerator) / math.sqrt(sq.denominator)
    return float(sign * (1 if s >= 0 else -1) * val)


def clebsch_gordan(j1, j2, J, m1, m2, M):
    """<j1 m1; j2 m2 | J M> for integer arguments."""
    if m1 + m2 != M:
        return 0.0
    sign = -1 if (j1 - j2 + M) % 2 else 1
    return sign * math.sqrt(2 * J + 1) * wigner3j(j1, j2, J, m1, m2, -M)


def gaunt_ck(l1, l2):
    """c^k(l1 m1, l2 m2) = sqrt(4pi/(2k+1)) int Y*_{l1m1} Y_{k,m1-m2} Y_{l2m2}.

    Indexed [k, m1 + l1, m2 + l2] like edrixs.get_gaunt.
    """
    res = np.zeros((l1 + l2 + 1, 2 * l1 + 1, 2 * l2 + 1))
    for k in range(l1 + l2 + 1):
        if (l1 + l2 + k) % 2 or not (abs(l1 - l2) <= k <= l1 + l2):
            continue
        w0 = wigner3j(l1, k, l2, 0, 0, 0)
        for i1, m1 in enumerate(range(-l1, l1 + 1)):
            for i2, m2 in enumerate(range(-l2, l2 + 1)):
                w = wigner3j(l1, k, l2, -m1, m1 - m2, m2)
                res[k, i1, i2] = ((-1) ** (m1 % 2)) * math.sqrt((2 * l1 + 1) * (2 * l2 + 1)) * w0 * w
    return res


def umat_slater(l_list, fk):
    """Coulomb tensor of edrixs/coulomb_utensor.py:80-229 (same keys for ``fk``).

    ``fk[(k, i, j, t, u)]`` are Slater integrals with 1-based shell indices.
    """
    labels = []
    for ish, l in enumerate(l_list):
        for m in range(-l, l + 1):
            for spin in range(2):
                labels.append((ish + 1, l, m, spin))
    kmax = 2 * max(l_list)
    ck = {}
    for la in set(l_list):
        for lb in set(l_list):
            g = gaunt_ck(la, lb)
            ck[(la, lb)] = g
    n = len(labels)
    umat = np.zeros((n, n, n, n), dtype=np.complex128)

    def c(k, la, ma, lb, mb):
        g = ck[(la, lb)]
        if k >= g.shape[0]:
            return 0.0
        return g[k, ma + la, mb + lb]

    for o1 in range(n):
        i1, l1, m1, s1 = labels[o1]
        for o2 in range(n):
            i2, l2, m2, s2 = labels[o2]
            for o3 in range(n):
                i3, l3, m3, s3 = labels[o3]
                if s1 != s3:
                    continue
                for o4 in range(n):
                    i4, l4, m4, s4 = labels[o4]
                    if s2 != s4 or (m1 + m2) != (m3 + m4):
                        continue
                    res = 0.0
                    for k in range(kmax + 1):
                        f = fk.get((k, i1, i2, i3, i4))
                        if f is not None:
                            res += c(k, l1, m1, l3, m3) * c(k, l4, m4, l2, m2) * f
                    umat[o1, o2, o4, o3] = res
    return umat / 2.0


def umat_slater_d(F0, F2, F4):
    return umat_slater([2], {(0, 1, 1, 1, 1): F0, (2, 1, 1, 1, 1): F2, (4, 1, 1, 1, 1): F4})


def umat_slater_dp(F0_dd, F2_dd, F4_dd, F0_dp, F2_dp, G1_dp, G3_dp, F0_pp, F2_pp):
    """d shell (orbitals 0-9) + p shell (orbitals 10-15), key layout of get_umat_slater('dp')."""
    fk = {(0, 1, 1, 1, 1): F0_dd, (2, 1, 1, 1, 1): F2_dd, (4, 1, 1, 1, 1): F4_dd,
          (0, 1, 2, 1, 2): F0_dp, (0, 2, 1, 2, 1): F0_dp,
          (2, 1, 2, 1, 2): F2_dp, (2, 2, 1, 2, 1): F2_dp,
          (1, 1, 2, 2, 1): G1_dp, (1, 2, 1, 1, 2): G1_dp,
          (3, 1, 2, 2, 1): G3_dp, (3, 2, 1, 1, 2): G3_dp,
          (0, 2, 2, 2, 2): F0_pp, (2, 2, 2, 2, 2): F2_pp}
    return umat_slater([2, 1], fk)


def atom_hsoc(l, zeta):
    """zeta * l.s on |m, sigma>, index (m+l)*2+sigma, sigma=0 up (+1/2), 1 down."""
    n = 2 * (2 * l + 1)
    h = np.zeros((n, n), dtype=np.complex128)
    for m in range(-l, l + 1):
        iu, idn = (m + l) * 2, (m + l) * 2 + 1
        h[iu, iu] = 0.5 * m
        h[idn, idn] = -0.5 * m
        if m < l:   # <m+1, dn| l+ s- |m, up> = sqrt(l(l+1) - m(m+1))
            a = math.sqrt(l * (l + 1) - m * (m + 1))
            jdn = (m + 1 + l) * 2 + 1
            h[jdn, iu] = 0.5 * a
            h[iu, jdn] = 0.5 * a
    return zeta * h


def tmat_c2r_d():
    """Columns: real d orbitals (dz2, dzx, dzy, dx2-y2, dxy) in |2,m> components."""
    s2 = math.sqrt(2.0)
    t = np.zeros((5, 5), dtype=np.complex128)
    t[2, 0] = 1.0
    t[1, 1], t[3, 1] = 1 / s2, -1 / s2
    t[1, 2], t[3, 2] = 1j / s2, 1j / s2
    t[0, 3], t[4, 3] = 1 / s2, 1 / s2
    t[0, 4], t[4, 4] = 1j / s2, -1j / s2
    return t


def cf_cubic_d(ten_dq):
    """Cubic crystal field on the 10 d spin-orbitals, complex harmonics."""
    t = tmat_c2r_d()
    ts = np.zeros((10, 10), dtype=np.complex128)
    ts[0::2, 0::2] = t
    ts[1::2, 1::2] = t
    dreal = np.zeros((10, 10), dtype=np.complex128)
    e = np.array([0.6, -0.4, -0.4, 0.6, -0.4]) * ten_dq
    dreal[0::2, 0::2] = np.diag(e)
    dreal[1::2, 1::2] = np.diag(e)
    return ts @ dreal @ ts.conj().T


def dipole_dp():
    """<d m1 s| r_pol |p m2 s> for pol = x, y, z, shape (3, 10, 6)."""
    op = np.zeros((3, 5, 3), dtype=np.complex128)
    for i1, m1 in enumerate(range(-2, 3)):
        for i2, m2 in enumerate(range(-1, 2)):
            t1 = clebsch_gordan(1, 1, 2, m2, -1, m1)
            t2 = clebsch_gordan(1, 1, 2, m2, 1, m1)
            t3 = clebsch_gordan(1, 1, 2, m2, 0, m1)
            op[0, i1, i2] = (t1 - t2) * math.sqrt(2.0) / 2.0
            op[1, i1, i2] = (t1 + t2) * 1j * math.sqrt(2.0) / 2.0
            op[2, i1, i2] = t3
    out = np.zeros((3, 10, 6), dtype=np.complex128)
    for i in range(3):
        out[i, 0::2, 0::2] = op[i]
        out[i, 1::2, 1::2] = op[i]
    return out


# ------------------------------------------------------------------- term lists

T2 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("re", "<f8"), ("im", "<f8")])
T4 = np.dtype([("ind1", "<i4"), ("ind2", "<i4"), ("ind3", "<i4"), ("ind4", "<i4"),
               ("re", "<f8"), ("im", "<f8")])


def _text_round(x):
    """The double that the reference's Fortran reads back from '{:.15f}'."""
    return float("{:.15f}".format(x))


def emat_terms(emat, tol=1e-12):
    """Nonzeros of a rank-2 matrix in numpy C order, 1-based, text-rounded.

    Mirrors edrixs/iostream.py:165-198 (write_emat), including the single
    placeholder term (1, 1, 0, 0) written for an all-zero matrix.
    """
    a1, a2 = np.nonzero(np.abs(emat) > tol)
    if len(a1) == 0:
        out = np.zeros(1, dtype=T2)
        out["ind1"], out["ind2"] = 1, 1
        return out
    out = np.zeros(len(a1), dtype=T2)
    out["ind1"], out["ind2"] = a1 + 1, a2 + 1
    v = emat[a1, a2]
    out["re"] = [_text_round(x) for x in v.real]
    out["im"] = [_text_round(x) for x in v.imag]
    return out


def umat_terms(umat, tol=1e-12):
    """Rank-4 counterpart (edrixs/iostream.py:201-235)."""
    idx = np.nonzero(np.abs(umat) > tol)
    if len(idx[0]) == 0:
        out = np.zeros(1, dtype=T4)
        for n in ("ind1", "ind2", "ind3", "ind4"):
            out[n] = 1
        return out
    out = np.zeros(len(idx[0]), dtype=T4)
    for k, n in enumerate(("ind1", "ind2", "ind3", "ind4")):
        out[n] = idx[k] + 1
    v = umat[idx]
    out["re"] = [_text_round(x) for x in v.real]
    out["im"] = [_text_round(x) for x in v.imag]
    return out


# ----------------------------------------------------------------- SIAM models

class SiamModel:
    """Terms of one impurity model: valence = 10 d + nbath bath spin-orbitals, 6 p core.

    Attributes (all term arrays are 1-based, reference term order):
      norb (valence spin-orbitals), ncore (=6), nocc (valence electrons, initial),
      hop_i, coul_i (initial / final Hamiltonian), hop_n, coul_n (core-hole
      Hamiltonian), trans[k] (k = x, y, z dipole, absorption direction:
      valence index row, core index column).
    """

    def __init__(self, nbath=10, nocc=None, seed=0, F=(8.0, 10.0, 6.2), zeta_d=0.08,
                 ten_dq=1.2, V=1.5, bath_range=(-3.0, 3.0), zeta_p=11.0,
                 dp=(0.0, 6.0, 4.5, 2.6), pp=(0.0, 0.0), bath_levels=None):
        self.nbath = nbath
        self.norb = 10 + nbath
        self.ncore = 6
        self.nocc = self.norb // 2 if nocc is None else nocc
        self.params = dict(nbath=nbath, seed=seed, F=tuple(F), zeta_d=zeta_d, ten_dq=ten_dq,
                           V=V, bath_range=tuple(bath_range), zeta_p=zeta_p, dp=tuple(dp),
                           pp=tuple(pp))
        nv, ntot = self.norb, self.norb + 6
        rng = np.random.default_rng(seed)
        if bath_levels is None:
            bath_levels = rng.uniform(bath_range[0], bath_range[1], size=nbath)
        self.bath_levels = np.asarray(bath_levels, dtype=float)
        emat = np.zeros((ntot, ntot), dtype=np.complex128)
        emat[0:10, 0:10] = atom_hsoc(2, zeta_d) + cf_cubic_d(ten_dq)
        for b in range(nbath):
            ib = 10 + b
            jd = b % 10          # d partner of bath spin-orbital b
            emat[ib, ib] += self.bath_levels[b]
            emat[jd, ib] += V
            emat[ib, jd] += np.conj(V)
        emat_n = emat.copy()
        emat_n[nv:ntot, nv:ntot] += atom_hsoc(1, zeta_p)
        self.emat_i, self.emat_n = emat, emat_n

        ud = umat_slater_d(*F)
        udp = umat_slater_dp(F[0], F[1], F[2], dp[0], dp[1], dp[2], dp[3], pp[0], pp[1])
        idx_d = list(range(10))
        idx_dp = list(range(10)) + [nv + i for i in range(6)]
        umat_i = np.zeros((ntot,) * 4, dtype=np.complex128)
        umat_i[np.ix_(idx_d, idx_d, idx_d, idx_d)] = ud
        umat_n = np.zeros((ntot,) * 4, dtype=np.complex128)
        umat_n[np.ix_(idx_dp, idx_dp, idx_dp, idx_dp)] = udp
        self.umat_i, self.umat_n = umat_i, umat_n

        tr = dipole_dp()
        self.trans_mat = np.zeros((3, ntot, ntot), dtype=np.complex128)
        self.trans_mat[:, 0:10, nv:ntot] = tr

        self.hop_i = emat_terms(self.emat_i)
        self.coul_i = umat_terms(self.umat_i)
        self.hop_n = emat_terms(self.emat_n)
        self.coul_n = umat_terms(self.umat_n)
        self.trans_i = [emat_terms(self.trans_mat[k]) for k in range(3)]
        # emission operator as the RIXS caller writes it: conj(T)^T (edrixs/solvers.py:2689)
        self.trans_f = [emat_terms(np.conj(self.trans_mat[k]).T) for k in range(3)]

    def pol_trans(self, polvec):
        """Absorption operator sum_k T_k * e_k and emission conj(sum_k T_k e_k)^T."""
        t = np.tensordot(np.asarray(polvec, dtype=np.complex128), self.trans_mat, axes=1)
        return emat_terms(t), emat_terms(np.conj(t).T)

    @property
    def dims(self):
        nv, n = self.norb, self.nocc
        return dict(ndim_i=math.comb(nv, n), ndim_n_nocore=math.comb(nv, n + 1),
                    ndim_n=6 * math.comb(nv, n + 1))
