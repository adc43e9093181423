"""File-driven orchestration around the drop-in solvers.

Counterparts of ``ed_siam_fort`` / ``xas_siam_fort`` / ``rixs_siam_fort``
(edrixs/solvers.py:1819-2704) for :class:`edrixs_amd.siam.SiamModel`: write the
term, fock and config files, call ``fedrixs`` (ours by default), read the
results back, and post-process poles into spectra (isotropic XAS averages the
three dipole components, edrixs/solvers.py:2453-2470).
"""
import contextlib
import os

import numpy as np

from . import iostream, spectra


@contextlib.contextmanager
def _cwd(path):
    old = os.getcwd()
    os.makedirs(path, exist_ok=True)
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


def _fsolver(fsolver):
    if fsolver is None:
        from . import fedrixs as fsolver
    return fsolver


def ed_siam(model, workdir=".", neval=3, nvector=3, maxiter=500, eigval_tol=1e-10,
            ed_solver=1, min_ndim=100, fsolver=None):
    """Ground state(s): eigvals.dat, denmat.dat and eigvec.N (edrixs/solvers.py:2123-2148)."""
    fs = _fsolver(fsolver)
    with _cwd(workdir):
        iostream.write_terms(model.hop_i, "hopping_i.in")
        iostream.write_terms(model.coul_i, "coulomb_i.in")
        iostream.write_terms(model.hop_n, "hopping_n.in")
        iostream.write_terms(model.coul_n, "coulomb_n.in")
        iostream.write_config(ed_solver=ed_solver, num_val_orbs=model.norb, neval=neval,
                              nvector=nvector, ncv=neval + 2, idump=True, maxiter=maxiter,
                              min_ndim=min_ndim, eigval_tol=eigval_tol)
        iostream.write_fock_dec_by_N(model.norb, model.nocc, "fock_i.in")
        fs.ed_fsolver(0, 0, 1)
        data = np.loadtxt("eigvals.dat", ndmin=2)
        evals = data[:neval, 1]
        d = np.loadtxt("denmat.dat", ndmin=2)
        shape = (nvector, model.norb, model.norb)
        denmat = d[:, 3].reshape(shape) + 1j * d[:, 4].reshape(shape)
    return evals, denmat


def xas_siam(model, ominc, workdir=".", gamma_c=0.5, num_gs=1, nkryl=300, temperature=1.0,
             fsolver=None, write_fock=True):
    """Isotropic XAS over the three dipole components (edrixs/solvers.py:2268-2475)."""
    fs = _fsolver(fsolver)
    om = np.asarray(ominc, dtype=float)
    gam = np.broadcast_to(np.asarray(gamma_c, dtype=float), om.shape)
    xas = np.zeros(len(om))
    poles = []
    with _cwd(workdir):
        iostream.write_config(num_val_orbs=model.norb, num_core_orbs=model.ncore,
                              num_gs=num_gs, nkryl=nkryl)
        if write_fock:
            iostream.write_fock_dec_by_N(model.norb, model.nocc, "fock_i.in")
            iostream.write_fock_dec_by_N(model.norb, model.nocc + 1, "fock_n.in")
        iostream.write_terms(model.hop_n, "hopping_n.in")
        iostream.write_terms(model.coul_n, "coulomb_n.in")
        for k in range(3):
            iostream.write_terms(model.trans_i[k], "transop_xas.in")
            fs.xas_fsolver(0, 0, 1)
            p = iostream.read_poles_from_file(["xas_poles.%d" % (i + 1) for i in range(num_gs)])
            xas += spectra.get_spectra_from_poles(p, om, gam, temperature)
            poles.append(p)
    return xas / 3.0, spectra.merge_pole_dicts(poles)


def rixs_siam(model, ominc, eloss, pol_vectors, workdir=".", gamma_c=0.5, gamma_f=0.1,
              num_gs=1, nkryl=300, linsys_max=1000, linsys_tol=1e-10, temperature=1.0,
              fsolver=None, write_fock=True):
    """RIXS for each incident energy and (e_in, e_out) polarization pair
    (edrixs/solvers.py:2478-2704).  Returns rixs[iom, ieloss, ipol] and poles."""
    fs = _fsolver(fsolver)
    om = np.asarray(ominc, dtype=float)
    el = np.asarray(eloss, dtype=float)
    gc = np.broadcast_to(np.asarray(gamma_c, dtype=float), om.shape)
    gf = np.broadcast_to(np.asarray(gamma_f, dtype=float), el.shape)
    rixs = np.zeros((len(om), len(el), len(pol_vectors)))
    poles = []
    with _cwd(workdir):
        if write_fock:
            iostream.write_fock_dec_by_N(model.norb, model.nocc, "fock_i.in")
            iostream.write_fock_dec_by_N(model.norb, model.nocc + 1, "fock_n.in")
            iostream.write_fock_dec_by_N(model.norb, model.nocc, "fock_f.in")
        iostream.write_terms(model.hop_i, "hopping_i.in")
        iostream.write_terms(model.coul_i, "coulomb_i.in")
        iostream.write_terms(model.hop_n, "hopping_n.in")
        iostream.write_terms(model.coul_n, "coulomb_n.in")
        for iom, omega in enumerate(om):
            iostream.write_config(num_val_orbs=model.norb, num_core_orbs=model.ncore,
                                  omega_in=omega, gamma_in=gc[iom], num_gs=num_gs, nkryl=nkryl,
                                  linsys_max=linsys_max, linsys_tol=linsys_tol)
            per_om = []
            for ip, (ei, ef) in enumerate(pol_vectors):
                ti, _ = model.pol_trans(ei)
                _, tf = model.pol_trans(ef)
                iostream.write_terms(ti, "transop_rixs_i.in")
                iostream.write_terms(tf, "transop_rixs_f.in")
                fs.rixs_fsolver(0, 0, 1)
                p = iostream.read_poles_from_file(["rixs_poles.%d" % (i + 1)
                                                   for i in range(num_gs)])
                rixs[iom, :, ip] = spectra.get_spectra_from_poles(p, el, gf, temperature)
                per_om.append(p)
            poles.append(per_om)
    return rixs, poles


def search_gs_sector(norb, hop, coul, imp_orbs, precision="c128i", tol=1e-10, maxiter=500,
                     log=None, concurrent=True):
    """Particle-number sector of the ground state: the ``do_ed=0`` search of
    ``ed_siam_fort`` (edrixs/solvers.py:2156-2263), in memory on the device.

    Starts at N = norb // 2 and walks the sectors N+1, N+2, ... and N-1, N-2,
    ... in the reference's order: a direction stops at the first sector whose
    ground energy is above the best so far, and a sector that improves on it
    stops the other direction.  Each sector's Hamiltonian is built on the
    device from the valence terms (hop, coul) and its lowest eigenvalue found
    with the Lanczos solver in ``precision``.  Returns (n_gs, e_gs, table) with
    table a list of (N, E0, impurity occupation) sorted by energy, as the
    reference writes to search_result.dat; the impurity occupation is the trace
    of the ground state's density matrix over the orbitals ``imp_orbs``.  With
    ``log`` (a writable file) each sector is printed as the walk takes it, like
    search_gs.log.

    With ``concurrent`` the first three sectors, and then the two sectors of
    each step while both directions are open, are evaluated at once on two
    worker threads with their own streams; the walk then takes them in the
    reference's order and drops a sector it would not have evaluated, so the
    result and the log are those of the sequential walk."""
    from . import core

    imp = np.asarray(list(imp_orbs), dtype=int)
    done = {}

    def sector(n):
        b = core.Basis.combinations(norb, n)
        H = core.build_ham_i(b, hop, coul)
        ev, vecs, _ = core.ground_state(H, nev=1, precision=precision, maxiter=maxiter, tol=tol)
        occ = float(np.real(np.trace(core.denmat(b, norb, vecs[0].contiguous())[np.ix_(imp, imp)])))
        return (n, float(ev[0]), occ)

    pool = None
    if concurrent:
        import threading
        from concurrent.futures import ThreadPoolExecutor
        import torch as _torch
        local = threading.local()

        def on_stream(n):
            if not hasattr(local, "stream"):
                local.stream = _torch.cuda.Stream()     # default priority
            s_ = local.stream
            core.set_stream(s_)
            try:
                with _torch.cuda.stream(s_):
                    r = sector(n)
                    s_.synchronize()
            finally:
                core.set_stream(None)
            return r

        pool = ThreadPoolExecutor(max_workers=2)

    def prefetch(ns):
        """Start the sectors of ns not yet evaluated (concurrent mode)."""
        if pool is None:
            return
        for n in ns:
            if 0 <= n <= norb and n not in done:
                done[n] = pool.submit(on_stream, n)

    def take(n):
        """The walk evaluates sector n (logging it in walk order)."""
        if pool is None:
            r = sector(n)
        else:
            prefetch([n])
            r = done[n].result()
        if log is not None:
            print("%d %.10f %.10f" % r, file=log, flush=True)
        return r

    n0 = norb // 2
    try:
        prefetch([n0, n0 + 1, n0 - 1])
        res = [take(n0)]
        n_gs, e_gs = n0, res[0][1]
        up = down = True
        for i in range(norb // 2):
            if up and down:
                prefetch([n0 + i + 1, n0 - i - 1])
            if up:
                r = take(n0 + i + 1)
                res.append(r)
                if r[1] > e_gs:
                    up = False
                else:
                    down = False
                    n_gs, e_gs = r[0], r[1]
            if down:
                r = take(n0 - i - 1)
                res.append(r)
                if r[1] > e_gs:
                    down = False
                else:
                    up = False
                    n_gs, e_gs = r[0], r[1]
    finally:
        if pool is not None:
            pool.shutdown(wait=True)
    res.sort(key=lambda t: t[1])
    return n_gs, e_gs, res
