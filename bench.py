#!/usr/bin/env python3
"""XAS+RIXS spectra/s of the device ED core on the norb=24, N=12 impurity model.

One step = one end-to-end pass of the hot path over one synthetic model
(10 d + 14 bath spin-orbitals, 2p core; ground-state space 2,704,156 rows,
core-hole space 14,976,864 rows), all on the device:

  1. CSR build of H_i, H_n, the three dipole absorption operators and the
     emission operators of the RIXS polarizations;
  2. ground state (Lanczos; the mixed forms finish with FP64 refinement);
  3. isotropic XAS: the three Krylov recurrences from D_k|gs> run as one
     batched recurrence over H_n for at least nkryl steps and until each
     continued fraction on nw points has converged (changes by less than
     cf_tol over cf_chunk further steps);
  4. RIXS at n_in incident energies x 2 polarizations: the shifted systems
     (w_in + E_gs + i Gamma_c - H_n) x = D_in|gs> of one polarization solved
     together by one multi-shift Galerkin recurrence over the unshifted H_n
     that keeps only D_out x (relative residual rixs_tol), then the
     final-state Krylov recurrences over H_i batched and run to converged
     continued fractions on the nw energy-loss points the same way.

Spectra per step = 1 (isotropic XAS) + 2 n_in (RIXS columns).  Three forms
are timed, the first giving ``value``: c128i (exact complex128 products: each
off-diagonal nonzero stored as the 16-bit number of its value in a table of
the matrix's distinct values, FP64 diagonal, vectors and accumulation), c64d
(complex64 off-diagonal values with the exact diagonal, FP64 vectors,
accumulation and scalars) and c128 (plain complex128 CSR).  The spectra of each
form are compared with the c128 form's.  A plain FP64 run of the reference's
own procedure on the GPU (one MINRES and one final-state recurrence per
(w_in, pol)), the serving loop and the CPU oracle baseline are reported beside.
"""
import argparse
import json
import os
import sys
import time

T_START = time.monotonic()   # the wall-clock budget counts from process start
DEADLINE = T_START + 480.0

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "XAS+RIXS spectra/s at norb=24,N=12: mixed vs FP64 GPU at 1e-6 err; p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix-core peak
PROFILE_KIND = {"c128": 0, "c64": 1, "c64d": 2, "c128i": 3}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--budget", type=float, default=480.0,
                   help="wall-clock seconds for the whole run, counted from process start: "
                        "warm-up is capped at one step (it holds the schedule autotuning), "
                        "as many of the requested steps are timed as leave room for the "
                        "comparison forms, the plain FP64 run, serving and the CPU baseline, "
                        "and each of those runs only if its estimated time still fits")
    p.add_argument("--nbath", type=int, default=14)
    p.add_argument("--nev", type=int, default=3,
                   help="lowest states of H_i found to tol 1e-10 (thick-restart Lanczos); the "
                        "spectra start from the lowest")
    p.add_argument("--nkryl", type=int, default=300)
    p.add_argument("--nw", type=int, default=2000)
    p.add_argument("--n-in", type=int, default=8, help="RIXS incident energies")
    p.add_argument("--rixs-tol", type=float, default=1e-8,
                   help="relative residual of the RIXS intermediate-state solves")
    p.add_argument("--rixs-maxit", type=int, default=6000)
    p.add_argument("--cf-tol", type=float, default=1e-8,
                   help="Krylov recurrences stop once a further cf-chunk steps change "
                        "their continued fraction by less than this (relative L2)")
    p.add_argument("--cf-chunk", type=int, default=250)
    p.add_argument("--nkryl-max", type=int, default=5000)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--forms", default="c128i,c64d,c128",
                   help="forms to time; the first is the reported value")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--exact", action="store_true",
                   help="also compare with the fully reorthogonalized FP64 projection")
    p.add_argument("--cpu-rows", type=int, default=200000)
    p.add_argument("--serve", type=int, default=12,
                   help="requests in the serving-loop measurement (0: skip it)")
    p.add_argument("--serve-budget-gb", type=float, default=96.0)
    p.add_argument("--plain-pairs", type=int, default=4,
                   help="RIXS (w_in, pol) pairs the plain FP64 reference procedure is timed on "
                        "(0: skip it; 2*n_in: all)")
    p.add_argument("--serve-only", action="store_true", help="run only the serving measurement")
    p.add_argument("--plain-only", action="store_true",
                   help="run only the plain FP64 procedure (the reference's, on the GPU), "
                        "--plain-pairs of its RIXS pairs, for a kernel trace of that form")
    p.add_argument("--config2", action="store_true",
                   help="run only the 128-RHS batched RIXS comparison on the norb=20 model")
    p.add_argument("--config4", action="store_true",
                   help="run only the norb=26 ground state + XAS + matrix-free measurement")
    p.add_argument("--eigs", action="store_true",
                   help="run only the lowest-three-states comparison (single-vector vs block "
                        "Lanczos) on the norb=20 and norb=24 models")
    p.add_argument("--eigs-nbath", default="10,14")
    p.add_argument("--cpu-check", action="store_true",
                   help="host only: the reference's whole procedure by the oracle at "
                        "--cpu-check-nbath against the cpu_baseline cost model")
    p.add_argument("--cpu-check-nbath", type=int, default=8)
    p.add_argument("--cpu-check-nin", type=int, default=2)
    p.add_argument("--probe", action="store_true",
                   help="run only a short fixed set of SpMV/SpMM launches (for counter runs)")
    p.add_argument("--probe-form", default="c128i")
    p.add_argument("--launches", type=int, default=10)
    return p.parse_args()


class Workload:
    GAMMA_C = 0.5      # core-hole broadening (XAS and RIXS intermediate), eV
    GAMMA_F = 0.1      # final-state broadening, eV
    POLS = (((1.0, 0.0, 0.0), (0.0, 1.0, 0.0)),   # (e_in, e_out) pairs
            ((0.0, 1.0, 0.0), (0.0, 0.0, 1.0)))

    def __init__(self, args):
        from edrixs_amd import core, siam
        self.core = core
        self.args = args
        self.model = siam.SiamModel(nbath=args.nbath, seed=args.seed)
        m = self.model
        self.bi = core.Basis.combinations(m.norb, m.nocc)
        self.bn = core.Basis.combinations(m.norb, m.nocc + 1)
        self.omega = np.linspace(-10.0, 30.0, args.nw)
        self.eloss = np.linspace(0.0, 10.0, args.nw)
        self.pol_terms = [(m.pol_trans(ei)[0], m.pol_trans(ef)[1]) for ei, ef in self.POLS]
        self.w_in = None   # set from the XAS edge of the first pass

    def step(self, precision):
        import torch
        c, m, a = self.core, self.model, self.args
        T = {}
        t0 = time.perf_counter()
        Hi = c.build_ham_i(self.bi, m.hop_i, m.coul_i)
        # core-hole space in the interleaved layout (valence * 6 + core slot): the
        # same operators permuted, spectra equal up to rounding
        lay = "interleaved"
        Hn = c.build_ham_n(self.bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout=lay)
        D = [c.build_transop_i(self.bn, self.bi, m.norb, m.ncore, m.trans_i[k], layout=lay)
             for k in range(3)]
        Din = [c.build_transop_i(self.bn, self.bi, m.norb, m.ncore, ti, layout=lay)
               for ti, _ in self.pol_terms]
        Dout = [c.build_transop_f(self.bi, self.bn, m.norb, m.ncore, tf, layout=lay)
                for _, tf in self.pol_terms]
        Hi.prepare(precision)
        Hn.prepare(precision)
        torch.cuda.synchronize()
        T["build"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        # the lowest nev states to tol 1e-10 (thick-restart Lanczos: a single-vector
        # two-pass run can return a ghost copy of the lowest level as the second);
        # the spectra start from the lowest, as num_gs = 1 does
        ev, vecs, it = c.ground_state(Hi, nev=a.nev, precision=precision, maxiter=400, tol=1e-10,
                                      solver="trlan")
        k0 = int(np.argmin(ev))
        egs = float(ev[k0])
        gs = vecs[k0].contiguous()
        torch.cuda.synchronize()
        T["ground_state"] = time.perf_counter() - t0
        # isotropic XAS: three components in one batched recurrence
        t0 = time.perf_counter()
        Phi = torch.stack([D[k].matvec(gs) for k in range(3)]).contiguous()
        # at least nkryl steps, continued until each fraction has converged
        xas_comp, xsteps, xchange = c.converged_spectra(
            Hn, Phi, egs, self.omega, self.GAMMA_C, precision=precision, tol=a.cf_tol,
            chunk=a.cf_chunk, min_steps=a.nkryl, max_steps=a.nkryl_max)
        xas = sum(xas_comp) / 3.0
        torch.cuda.synchronize()
        T["xas"] = time.perf_counter() - t0
        if self.w_in is None:   # incident energies: 10 eV window starting 2 eV below the edge
            edge = self.omega[int(np.argmax(xas))]
            self.w_in = np.linspace(edge - 2.0, edge + 8.0, a.n_in)
        # RIXS: one multi-shift recurrence over H_n serves every incident energy; the
        # incident polarizations run as one interleaved recurrence (a 2-vector block
        # product per step) where H_n has the slot-factored form, else one each,
        # each projected by its own D_out; only D_out x is kept
        t0 = time.perf_counter()
        npol = len(self.POLS)
        sigma = np.array([w + egs + 1j * self.GAMMA_C for w in self.w_in])
        B = torch.stack([Din[p].matvec(gs) for p in range(npol)]).contiguous()
        proj, it_b, res_b = c.shifted_fom_rhs(Hn, B, sigma, Dout, tol=a.rixs_tol,
                                              maxit=a.rixs_maxit, precision=precision)
        iters = [int(v) for v in it_b.reshape(-1)]
        res = [float(v) for v in res_b.reshape(-1)]
        torch.cuda.synchronize()
        T["rixs_solve"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        # (w_in, pol) order: q = i_w * npol + p
        Phf = torch.stack([proj[p][i] for i in range(len(self.w_in))
                           for p in range(npol)]).contiguous()
        # final-state recurrences run until each continued fraction has converged
        # on the energy-loss grid (the narrow final-state width needs far more
        # than the XAS step count)
        rixs, ksteps, kchange = c.converged_spectra(
            Hi, Phf, egs, self.eloss, self.GAMMA_F, precision=precision, tol=a.cf_tol,
            chunk=a.cf_chunk, min_steps=a.nkryl, max_steps=a.nkryl_max)
        torch.cuda.synchronize()
        T["rixs_krylov"] = time.perf_counter() - t0
        info = dict(nnz_i=Hi.nnz, nnz_n=Hn.nnz, m_n=Hn.m, fac_n=Hn.factored(), lanczos_iters=it,
                    rixs_iters=iters, rixs_res_rel_max=max(res),
                    rixs_krylov_steps=[int(v) for v in ksteps],
                    xas_krylov_steps=[int(v) for v in xsteps], xas_cf_change_max=float(max(xchange)),
                    rixs_cf_change_max=float(max(kchange)),
                    phases_s=T, egs=egs, lowest_states=[float(v) for v in np.sort(ev)])
        out = dict(egs=egs, xas=xas, xas_comp=xas_comp, rixs=rixs, info=info, gs=gs)
        del Hi, Hn, D, Din, Dout, Phi, proj, Phf, vecs
        return out


def plain_fp64(work, ref, pairs):
    """The reference's procedure on the GPU in complex128, one (w_in, pol) at a
    time as its driver runs (src/rixs_driver.f90:127-218): rebuild
    w_in + E_gs + i Gamma_c - H_n, MINRES to the same relative residual, apply
    D_out, then one final-state Krylov recurrence to a converged fraction; XAS
    as three single-vector recurrences.  ``pairs`` of the 2 n_in RIXS pairs
    (spread over the incident energies) are run and timed one by one; the RIXS
    time is scaled to all pairs.  Their spectra are compared with the batched
    multi-shift form's."""
    import torch
    c, m, a = work.core, work.model, work.args
    T = {}
    t0 = time.perf_counter()
    Hi = c.build_ham_i(work.bi, m.hop_i, m.coul_i)
    Hn = c.build_ham_n(work.bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    D = [c.build_transop_i(work.bn, work.bi, m.norb, m.ncore, m.trans_i[k]) for k in range(3)]
    torch.cuda.synchronize()
    T["build"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ev, vecs, it = c.ground_state(Hi, nev=a.nev, precision="c128", maxiter=400, tol=1e-10,
                                  solver="trlan")
    k0 = int(np.argmin(ev))
    egs = float(ev[k0])
    gs = vecs[k0].contiguous()
    torch.cuda.synchronize()
    T["ground_state"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    xas = 0.0
    for k in range(3):
        phi = D[k].matvec(gs).reshape(1, -1).contiguous()
        sp, _, _ = c.converged_spectra(Hn, phi, egs, work.omega, work.GAMMA_C, precision="c128",
                                       tol=a.cf_tol, chunk=a.cf_chunk, min_steps=a.nkryl,
                                       max_steps=a.nkryl_max)
        xas = xas + sp[0] / 3.0
    torch.cuda.synchronize()
    T["xas"] = time.perf_counter() - t0
    if work.w_in is None:   # run alone: the incident energies from its own XAS edge
        edge = work.omega[int(np.argmax(xas))]
        work.w_in = np.linspace(edge - 2.0, edge + 8.0, a.n_in)
    del Hn
    npol = len(work.POLS)
    allp = [(i, p) for i in range(len(work.w_in)) for p in range(npol)]
    pick = sorted(set(int(round(v)) for v in np.linspace(0, len(allp) - 1, pairs)))
    pair_s, iters, diffs = [], [], []
    for q in pick:
        i, p = allp[q]
        t0 = time.perf_counter()
        sig = work.w_in[i] + egs + 1j * work.GAMMA_C
        A = c.build_ham_n(work.bn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=sig, ham_sign=-1.0)
        Din = c.build_transop_i(work.bn, work.bi, m.norb, m.ncore, work.pol_terms[p][0])
        Dout = c.build_transop_f(work.bi, work.bn, m.norb, m.ncore, work.pol_terms[p][1])
        b = Din.matvec(gs)
        bn = float(torch.linalg.vector_norm(b))
        x, nit, res = c.minres(A, b, tol=a.rixs_tol * bn, maxit=a.rixs_maxit)
        del A
        phi = Dout.matvec(x).reshape(1, -1).contiguous()
        spec, _, _ = c.converged_spectra(Hi, phi, egs, work.eloss, work.GAMMA_F, precision="c128",
                                         tol=a.cf_tol, chunk=a.cf_chunk, min_steps=a.nkryl,
                                         max_steps=a.nkryl_max)
        torch.cuda.synchronize()
        pair_s.append(time.perf_counter() - t0)
        iters.append(int(nit))
        if ref is not None:
            diffs.append(rel(spec[0], ref["rixs"][q]))
        log("plain fp64 pair %d (w_in %.2f, pol %d): %.2f s, %d MINRES iterations" %
            (q, work.w_in[i], p, pair_s[-1], nit))
    t_rixs = float(np.mean(pair_s)) * len(allp)
    total = T["build"] + T["ground_state"] + T["xas"] + t_rixs
    nspec = 1 + len(allp)
    return {"spectra_per_s": nspec / total, "seconds_per_step": total, "phases_s": T,
            "rixs_pairs_timed": len(pick), "rixs_pairs_total": len(allp),
            "rixs_pair_seconds": pair_s, "minres_iters": iters,
            "rixs_rel_l2_vs_multishift_c128_max": max(diffs) if diffs else None,
            "note": ("builds, ground state and XAS timed in full; the RIXS time is the mean of "
                     "the timed pairs times the number of pairs" if len(pick) < len(allp)
                     else "every pair timed")}


def left():
    """Seconds left of the run's wall-clock budget."""
    return DEADLINE - time.monotonic()


def timed(work, precision, steps, warmup, reserve=lambda t: 0.0):
    """``warmup`` untimed steps, then up to ``steps`` timed ones.  The first
    timed step always runs; a further one only if the budget left, less
    ``reserve(t)`` for the work that follows, still holds one step at the
    slowest step time t seen so far (warm-up included).  Returns the timed
    seconds, the number of timed steps, the last step's output and the kernel
    profile of the timed region."""
    import torch
    c = work.core
    slowest = 0.0
    for i in range(warmup):
        t = time.perf_counter()
        work.step(precision)
        slowest = max(slowest, time.perf_counter() - t)
        log("%s warmup %d: %.2f s" % (precision, i, slowest))
    torch.cuda.synchronize()
    c.profile(True)
    t0 = time.perf_counter()
    out = None
    done = 0
    while done < steps:
        if done > 0 and left() - reserve(slowest) < slowest:
            log("%s: budget holds no further step (%.0f s left, step %.1f s, %.0f s reserved); "
                "%d of %d steps timed" % (precision, left(), slowest, reserve(slowest), done, steps))
            break
        ts = time.perf_counter()
        out = work.step(precision)
        slowest = max(slowest, time.perf_counter() - ts)
        done += 1
        log("%s step %d done at %.2f s  phases %s" % (precision, done - 1, time.perf_counter() - t0,
            {k: round(v, 3) for k, v in out["info"]["phases_s"].items()}))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = {k: c.profile_read(precision, k) + (c.profile_read_alg(precision, k),)
            for k in c.KERNELS}
    prof["all"] = c.profile_read(precision) + (c.profile_read_alg(precision),)
    # the same per count of right-hand sides (one template instance each)
    prof["by_rhs"] = {}
    for k in c.KERNELS:
        for r in range(1, 18):
            got = c.profile_read(precision, k, r)
            if got[0]:
                prof["by_rhs"][(k, r)] = got + (c.profile_read_alg(precision, k, r),)
    c.profile(False)
    return dt, done, out, prof


# Wall-clock cost of the work after the timed steps, in units of one step of
# the reported form (c128i) where it scales with it, else in seconds: one
# step of each comparison form, the plain FP64 run with its default pairs,
# the serving measurement and the CPU baseline sample.
COST_STEPS = {"c128": 1.6, "c64d": 1.3, "c64": 1.3, "c128i": 1.0, "plain": 2.2}
COST_SECONDS = {"serve": 45.0, "cpu": 30.0}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def exact_reference(work, gs, results):
    """Fully reorthogonalized FP64 Krylov projection for the z dipole component,
    taken to the step count the FP64 form converged at (bounded by the HBM that
    keeping every basis vector needs), and the same recurrence with the mixed
    (c64d) products; every form's converged z spectrum is compared with it."""
    import torch
    c, m, a = work.core, work.model, work.args
    Hn = c.build_ham_n(work.bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Dz = c.build_transop_i(work.bn, work.bi, m.norb, m.ncore, m.trans_i[2])
    phi = Dz.matvec(gs)
    egs = results["c128"]["egs"]
    free, _ = torch.cuda.mem_get_info()
    steps = min(results["c128"]["info"]["xas_krylov_steps"][2],
                int(0.5 * free / (16.0 * Hn.n)) - 2)
    t = time.perf_counter()
    al, be, nr = c.build_krylov_reorth(Hn, phi, steps, "c128")
    ref = c.cf_spectrum(al, be, nr, egs, work.omega, work.GAMMA_C)
    log("reorthogonalized FP64 reference (%d steps): %.1f s" % (steps, time.perf_counter() - t))
    Hn.prepare("c64d")
    al2, be2, nr2 = c.build_krylov_reorth(Hn, phi, steps, "c64d")
    mix = c.cf_spectrum(al2, be2, nr2, egs, work.omega, work.GAMMA_C)
    out = {"component": "z", "steps": steps, "c64d_reorth_vs_exact": rel(mix, ref)}
    for form, r in results.items():
        out["%s_converged_vs_exact" % form] = rel(r["xas_comp"][2], ref)
    return out


def cpu_costs(m, rows, reps=3):
    """Oracle costs measured on row windows of H_i and H_n (the middle third of
    each, ``rows`` rows): seconds per built row and per nonzero of a product,
    with the oracle's row loops over the OpenMP threads the host gives."""
    from oracle import oracle as o
    from edrixs_amd import iostream
    fi = iostream.fock_words(m.norb, m.nocc)
    fn = iostream.fock_words(m.norb, m.nocc + 1)
    mi, mn = len(fi), len(fn) * m.ncore
    lo_i, lo_n = mi // 3, mn // 3
    ri, rn = min(rows, mi - lo_i), min(rows, mn - lo_n)
    t = time.perf_counter()
    Ai = o.build_ham_i(fi, m.hop_i, m.coul_i, row_begin=lo_i, row_end=lo_i + ri)
    tbi = (time.perf_counter() - t) / ri
    t = time.perf_counter()
    An = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, row_begin=lo_n, row_end=lo_n + rn)
    tbn = (time.perf_counter() - t) / rn
    xi, xn = np.ones(mi, np.complex128), np.ones(mn, np.complex128)
    t = time.perf_counter()
    for _ in range(reps):
        o.matvec(Ai, xi)
    tmi = (time.perf_counter() - t) / reps / max(Ai.nnz, 1)
    t = time.perf_counter()
    for _ in range(reps):
        o.matvec(An, xn)
    tmn = (time.perf_counter() - t) / reps / max(An.nnz, 1)
    return dict(mi=mi, mn=mn, ri=ri, rn=rn, reps=reps, tbi=tbi, tbn=tbn, tmi=tmi, tmn=tmn,
                cores=int(o.lib().orc_num_threads()))


def cpu_model(c, nnz_i, nnz_n, lanczos_iters, xas_steps, rixs_iters, rixs_krylov_steps):
    """Seconds of the reference's procedure predicted from the window costs:
    builds by rows, products by nonzeros times the iteration counts (ED, one
    XAS component per Krylov run rebuilding H_n, one RIXS (w_in, pol) call per
    pair rebuilding shifted H_n and H_f).  Vector operations are not counted."""
    build_i, build_n = c["tbi"] * c["mi"], c["tbn"] * c["mn"]
    spmv_i, spmv_n = c["tmi"] * nnz_i, c["tmn"] * nnz_n
    t_ed = build_i + lanczos_iters * spmv_i
    t_xas = sum(build_n + k * spmv_n for k in xas_steps)
    t_rixs = sum(build_n + it * spmv_n + build_i + k * spmv_i
                 for it, k in zip(rixs_iters, rixs_krylov_steps))
    return t_ed, t_xas, t_rixs


def cpu_baseline(work, info, rows):
    """Oracle (C restatement, its row loops over the OpenMP threads the host
    gives, OMP_NUM_THREADS) on a bounded sample of the same step,
    following the reference's procedure: one fsolver call per RIXS (w_in, pol)
    that rebuilds H_n (shifted), the operators and H_f, runs MINRES and the
    final Krylov; XAS rebuilds H_n per component.  Per-row build cost and
    per-nonzero matvec cost are measured on row windows and scaled
    (cpu_model; ``--cpu-check`` compares that model with the whole procedure
    run at a smaller size)."""
    c = cpu_costs(work.model, rows)
    # the reference runs one MINRES per (w_in, pol); its iteration count is taken
    # as that of the Galerkin iterate the device solve stopped at for that shift
    xs = [float(np.mean(info["xas_krylov_steps"]))] * 3
    t_ed, t_xas, t_rixs = cpu_model(c, info["nnz_i"], info["nnz_n"], info["lanczos_iters"], xs,
                                    info["rixs_iters"], info["rixs_krylov_steps"])
    total = t_ed + t_xas + t_rixs
    n_rixs = len(info["rixs_iters"])
    nspec = 1 + n_rixs
    return {"value": nspec / total, "unit": "spectra/s", "cores": c["cores"], "kind": "port",
            "sample": ("oracle (%d threads, rows in parallel) built %d H_i and %d H_n rows and ran %d row-block "
                       "matvecs on each; per-row and per-nonzero costs scaled to the step "
                       "(ED + 3 XAS components + %d RIXS (w_in, pol) calls, each rebuilding "
                       "its matrices as the reference driver does): %.0f s for %d spectra"
                       % (c["cores"], c["ri"], c["rn"], c["reps"], n_rixs, total, nspec)),
            "seconds_per_step": total}


def cpu_check(args):
    """The reference's whole procedure run by the oracle on the host (no GPU):
    ED by Lanczos, XAS as three Krylov runs each rebuilding H_n, and RIXS as
    one call per (w_in, pol) that rebuilds sigma - H_n, solves it by MINRES,
    applies D_out and runs the final-state Krylov over a rebuilt H_f
    (src/ed_driver.f90, src/xas_driver.f90, src/rixs_driver.f90), at
    ``--cpu-check-nbath``; then the cost model cpu_baseline scales from row
    windows, fed with this run's own iteration counts."""
    from oracle import oracle as o
    from edrixs_amd import iostream, siam
    m = siam.SiamModel(nbath=args.cpu_check_nbath, seed=args.seed)
    fi = iostream.fock_words(m.norb, m.nocc)
    fn = iostream.fock_words(m.norb, m.nocc + 1)
    rng = np.random.default_rng(args.seed)
    omega = np.linspace(-10.0, 30.0, args.nw)
    T = {}
    t0 = time.perf_counter()
    Hi = o.build_ham_i(fi, m.hop_i, m.coul_i)
    v0 = rng.normal(size=Hi.m) + 1j * rng.normal(size=Hi.m)
    ev, vecs, it_ed = o.lanczos(Hi, v0, 1, 2000, 1e-10, 1)
    egs, gs = float(ev[0]), vecs[0]
    T["ed"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    xas, xsteps = 0.0, []
    for k in range(3):
        Hn = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n)
        D = o.build_transop_i(fn, fi, m.norb, m.ncore, m.trans_i[k])
        a, b, nrm = o.build_krylov(Hn, o.matvec(D, gs), args.nkryl)
        xsteps.append(len(a))
        xas = xas + o.cf_spectrum(a, b, nrm, egs, omega, Workload.GAMMA_C)
    T["xas"] = time.perf_counter() - t0
    nnz_n = Hn.nnz
    del Hn
    edge = omega[int(np.argmax(xas))]
    w_in = np.linspace(edge - 2.0, edge + 8.0, args.cpu_check_nin)
    t0 = time.perf_counter()
    iters, ksteps = [], []
    for w in w_in:
        for ei, ef in Workload.POLS:
            sig = w + egs + 1j * Workload.GAMMA_C
            A = o.build_ham_n(fn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=sig, ham_sign=-1.0)
            tin, tout = m.pol_trans(ei)[0], m.pol_trans(ef)[1]
            Din = o.build_transop_i(fn, fi, m.norb, m.ncore, tin)
            Dout = o.build_transop_f(fi, fn, m.norb, m.ncore, tout)
            bvec = o.matvec(Din, gs)
            x, it, _ = o.pminres(A, bvec, args.rixs_tol * np.linalg.norm(bvec), args.rixs_maxit)
            Hf = o.build_ham_i(fi, m.hop_i, m.coul_i)
            a, b, nrm = o.build_krylov(Hf, o.matvec(Dout, x), args.nkryl)
            iters.append(it)
            ksteps.append(len(a))
    T["rixs"] = time.perf_counter() - t0
    measured = T["ed"] + T["xas"] + T["rixs"]
    c = cpu_costs(m, args.cpu_rows)
    t_ed, t_xas, t_rixs = cpu_model(c, Hi.nnz, nnz_n, it_ed, xsteps, iters, ksteps)
    model = t_ed + t_xas + t_rixs
    return {"norb": m.norb, "rows_i": Hi.m, "rows_n": c["mn"], "nnz_i": Hi.nnz, "nnz_n": nnz_n,
            "cores": c["cores"], "window_rows": [c["ri"], c["rn"]],
            "measured_s": {k: round(v, 2) for k, v in T.items()}, "measured_total_s": measured,
            "model_s": {"ed": t_ed, "xas": t_xas, "rixs": t_rixs}, "model_total_s": model,
            "model_over_measured": model / measured,
            "counts": {"lanczos": it_ed, "xas_krylov": xsteps, "minres": iters,
                       "rixs_krylov": ksteps}}


def serve(args):
    """The serving form: ``--serve`` requests with Poisson arrivals (20 per
    second), norb drawn from {16, 18, 20, 22, 24}, N = norb / 2, per-request
    model parameters, half XAS (isotropic, 3 components) and half RIXS (4
    incident energies, one polarization pair), run by the continuous-batching
    engine in c128i with spectra converged to 1e-6.  Reports latency
    percentiles beside spectra/s."""
    from edrixs_amd import serving
    reqs = serving.make_requests(args.serve, rate=20.0, sizes=(16, 18, 20, 22, 24),
                                 rixs_fraction=0.5, seed=args.seed)
    eng = serving.Engine(budget_gb=args.serve_budget_gb, precision="c128i", check_every=100,
                         chunk=25, cf_tol=1e-6, min_steps=200, max_steps=5000, n_in=4, npol=1,
                         nw=1000, rixs_tol=1e-6, rixs_maxit=5000)
    done = eng.run(reqs, realtime=True, log=log)
    rep = serving.report(done)
    rep.update({"arrival_rate_per_s": 20.0, "sizes_norb": [16, 18, 20, 22, 24],
                "hbm_budget_gb": args.serve_budget_gb, "precision": "c128i",
                "kinds": {k: sum(r.kind == k for r in done) for k in ("xas", "rixs")},
                "norb_of_requests": [r.norb for r in sorted(done, key=lambda r: r.rid)]})
    return rep


def config4(args):
    """The largest single-GPU case: norb = 26 (10 d + 16 bath), N = 13.  The
    ground state of H_i (10,400,600 rows) in c64d with FP64 refinement, its
    energy error against an FP64 Lanczos run, isotropic XAS over H_n
    (57,946,200 rows) to converged continued fractions, and the matrix-free
    apply of the same H_n checked against the stored CSR and timed."""
    import torch
    c = __import__("edrixs_amd.core", fromlist=["core"])
    from edrixs_amd import siam
    out = {"workload": "norb=26 N=13 SIAM (10 d + 16 bath, 2p core)"}
    m = siam.SiamModel(nbath=16, seed=args.seed)
    t0 = time.perf_counter()
    bi = c.Basis.combinations(m.norb, m.nocc)
    bn = c.Basis.combinations(m.norb, m.nocc + 1)
    Hi = c.build_ham_i(bi, m.hop_i, m.coul_i)
    Hi.prepare("c64d")
    torch.cuda.synchronize()
    out["H_i"] = {"rows": Hi.m, "nnz": Hi.nnz, "build_s": time.perf_counter() - t0}
    t0 = time.perf_counter()
    ev64, _, it64 = c.ground_state(Hi, nev=1, precision="c128", maxiter=500, tol=1e-10)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    ev, vecs, it = c.ground_state(Hi, nev=1, precision="c64d", maxiter=500, tol=1e-10)
    tmix = time.perf_counter() - t0
    out["ground_state"] = {"E_fp64": float(ev64[0]), "E_c64d_refined": float(ev[0]),
                           "abs_err_eV": abs(float(ev[0]) - float(ev64[0])),
                           "fp64_s": t64, "c64d_s": tmix}
    log("config4 ground state: %.10f (err %.2e)" % (ev[0], out["ground_state"]["abs_err_eV"]))
    gs = vecs[0].contiguous()
    egs = float(ev[0])
    t0 = time.perf_counter()
    Hn = c.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    Hn.prepare("c64d")
    torch.cuda.synchronize()
    out["H_n"] = {"rows": Hn.m, "nnz": Hn.nnz, "build_s": time.perf_counter() - t0}
    D = [c.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[k]) for k in range(3)]
    Phi = torch.stack([D[k].matvec(gs) for k in range(3)]).contiguous()
    omega = np.linspace(-10.0, 30.0, args.nw)
    c.profile(True)
    t0 = time.perf_counter()
    comp, steps, change = c.converged_spectra(Hn, Phi, egs, omega, 0.5, precision="c64d",
                                              tol=args.cf_tol, chunk=args.cf_chunk,
                                              min_steps=args.nkryl, max_steps=args.nkryl_max)
    torch.cuda.synchronize()
    txas = time.perf_counter() - t0
    n, ms, by = c.profile_read("c64d", "k_spmv")
    c.profile(False)
    out["xas"] = {"seconds": txas, "krylov_steps": steps, "cf_change": change,
                  "k_spmv_launches": n, "k_spmv_avg_us": ms / max(n, 1) * 1e3,
                  "k_spmv_alg_gbs": by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0}
    log("config4 XAS: %.1f s, steps %s" % (txas, steps))
    # matrix-free apply of the same H_n
    mf = c.mf_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    x = gs.new_zeros(Hn.n)
    x.copy_(Phi[2])
    y_csr = Hn.matvec(x)
    y_mf = mf.matvec(x)
    diff = float((y_mf - y_csr).abs().max() / y_csr.abs().max())
    c.profile(True)
    for _ in range(3):
        mf.matvec(x)
    nm, msm, bym = c.profile_read("c128", "k_mf")
    c.profile(False)
    csr_bytes = Hn.nnz * 20.0 + (Hn.m + 1) * 8.0 + Hn.n * 16.0 + Hn.m * 32.0
    out["matrix_free"] = {"rows": mf.n, "avg_ms": msm / max(nm, 1),
                          "alg_bytes": bym / max(nm, 1), "alg_gbs": bym / (msm * 1e-3) / 1e9,
                          "stored_csr_equivalent_bytes": csr_bytes,
                          "csr_equivalent_gbs": csr_bytes / (msm / max(nm, 1) * 1e-3) / 1e9,
                          "rel_diff_vs_csr": diff}
    log("config4 matrix-free apply: %.1f ms, rel diff vs CSR %.1e" % (msm / max(nm, 1), diff))
    return out


def eigs(args):
    """Lowest three states of H_i to tol 1e-10 (the norb=20 N=10 model asks for
    them; norb=24 N=12 shown beside it): single-vector Lanczos with full
    reorthogonalization and deflated restarts against block Lanczos (blocks of
    8, reorthogonalization on the FP64 matrix cores), both in the exact forms.
    The block run's projection and update kernels are timed live and priced
    against the HBM and FP64 matrix-core peaks."""
    import torch
    c = __import__("edrixs_amd.core", fromlist=["core"])
    from edrixs_amd import siam
    out = {}
    for nbath in [int(x) for x in args.eigs_nbath.split(",")]:
        m = siam.SiamModel(nbath=nbath, seed=args.seed)
        bi = c.Basis.combinations(m.norb, m.nocc)
        Hi = c.build_ham_i(bi, m.hop_i, m.coul_i)
        Hi.prepare("c128i")
        torch.cuda.synchronize()
        r = {"rows": Hi.m, "nnz": Hi.nnz}
        for form in ("c128", "c128i"):
            t0 = time.perf_counter()
            ev1, _, it1 = c.lanczos_eigs(Hi, 3, maxiter=800, tol=1e-10, seed=1, deflate=True,
                                         precision=form)
            t1 = time.perf_counter() - t0
            c.profile(True)
            t0 = time.perf_counter()
            evb, vb, nb, res = c.block_lanczos_eigs(Hi, 3, maxblocks=100, tol=1e-10, seed=1,
                                                    precision=form)
            tb = time.perf_counter() - t0
            kern = {}
            for k in c.BLOCK_KERNELS:
                n_, ms, by = c.profile_read("c128", k)
                fl = c.profile_read_flops(k)
                if n_:
                    kern[k] = {"launches": n_, "total_s": ms * 1e-3, "avg_us": ms / n_ * 1e3,
                               "alg_gbs": by / (ms * 1e-3) / 1e9,
                               "hbm_frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "tflops": fl / (ms * 1e-3) / 1e12,
                               "fp64_mfma_frac": fl / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS}
            c.profile(False)
            x = vb[0].contiguous()
            hx = Hi.matvec(x)
            gres = float((hx - float(evb[0]) * x).abs().max())
            r[form] = {"single_vector": {"seconds": t1, "krylov_steps": it1,
                                         "evals": [float(v) for v in ev1]},
                       "block": {"seconds": tb, "block_steps": nb, "evals": [float(v) for v in evb],
                                 "residual_estimates": [float(v) for v in res],
                                 "gs_residual_max_abs": gres, "kernels": kern},
                       "speedup_block_vs_single": t1 / tb,
                       "max_abs_eval_diff": float(np.abs(evb - ev1).max())}
            log("eigs nbath=%d %s: single %.2f s (%d steps), block %.2f s (%d blocks), "
                "max |dE| %.1e" % (nbath, form, t1, it1, tb, nb, r[form]["max_abs_eval_diff"]))
            del vb
        out["norb%d" % m.norb] = r
        del Hi
        torch.cuda.empty_cache()
    return out


def config2(args):
    """Batched RIXS on the norb=20 model: 64 incident energies over 10 eV x 2
    polarizations = 128 shifted systems (w_in + E_gs + i Gamma_c - H_n) x =
    D_pol|gs> to relative residual 1e-8, then final-state continued fractions
    over H_i on 2000 loss points.  Three routes to the same 128 spectra:
    (1) c64d multi-shift recurrences (one per polarization, projected by D_out),
    (2) c64d batched MINRES over the interleaved SpMM (128 right-hand sides),
    (3) the reference's per-RHS FP64 MINRES loop (edx_minres, one system at a
    time).  Routes 1 and 2 are compared with route 3 spectrum by spectrum."""
    import torch
    c = __import__("edrixs_amd.core", fromlist=["core"])
    from edrixs_amd import siam
    m = siam.SiamModel(nbath=10, seed=args.seed)
    bi = c.Basis.combinations(m.norb, m.nocc)
    bn = c.Basis.combinations(m.norb, m.nocc + 1)
    Hi = c.build_ham_i(bi, m.hop_i, m.coul_i)
    Hn = c.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n)
    ev, vecs, _ = c.ground_state(Hi, nev=1, precision="c128", maxiter=400, tol=1e-10)
    egs = float(ev[0])
    gs = vecs[0].contiguous()
    e0n, _, _ = c.lanczos_eigs(Hn, 1, maxiter=300, tol=1e-8)
    edge = float(e0n[0]) - egs
    n_in = 64
    w_in = np.linspace(edge - 2.0, edge + 8.0, n_in)
    sigma = w_in + egs + 0.5j
    eloss = np.linspace(0.0, 10.0, 2000)
    pols = (((1.0, 0.0, 0.0), (0.0, 1.0, 0.0)), ((0.0, 1.0, 0.0), (0.0, 0.0, 1.0)))
    Din = [c.build_transop_i(bn, bi, m.norb, m.ncore, m.pol_trans(ei)[0]) for ei, _ in pols]
    Dout = [c.build_transop_f(bi, bn, m.norb, m.ncore, m.pol_trans(ef)[1]) for _, ef in pols]
    b = [Din[p].matvec(gs) for p in range(2)]
    Hi.prepare("c64d")
    Hn.prepare("c64d")
    kw = dict(tol=args.cf_tol, chunk=args.cf_chunk, min_steps=args.nkryl, max_steps=args.nkryl_max)

    def finals(phis, precision):
        P = torch.stack(phis).contiguous()
        spec, steps, _ = c.converged_spectra(Hi, P, egs, eloss, 0.1, precision=precision, **kw)
        return spec, steps

    out = {"workload": "norb=20 N=10 SIAM, 64 w_in x 2 pols = 128 RHS on H_n (%d rows)" % Hn.n}
    torch.cuda.synchronize()
    # (1) multi-shift, c64d
    t0 = time.perf_counter()
    phis, its = [], []
    for p in range(2):
        (Y,), it, res = c.shifted_fom(Hn, b[p], sigma, projections=(Dout[p],), tol=1e-8,
                                      maxit=8000, precision="c64d")
        phis += [Y[k] for k in range(n_in)]
        its.append(int(it.max()))
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    s1, st1 = finals(phis, "c64d")
    torch.cuda.synchronize()
    out["multishift_c64d"] = {"solve_s": t_solve, "total_s": time.perf_counter() - t0,
                              "recurrence_steps": its}
    log("config2 multi-shift: solve %.2f s, total %.2f s" % (t_solve, out["multishift_c64d"]["total_s"]))
    # (2) batched MINRES over the SpMM, c64d
    t0 = time.perf_counter()
    B = torch.stack([b[p] for p in range(2) for _ in range(n_in)]).contiguous()
    sig2 = np.concatenate([sigma, sigma])
    bn_min = float(torch.linalg.vector_norm(B, dim=1).min())
    X, it2, res2 = c.minres_shifted_batch(Hn, sig2, B, tol=1e-8 * bn_min, maxit=8000,
                                          precision="c64d")
    torch.cuda.synchronize()
    t_solve2 = time.perf_counter() - t0
    phis2 = [Dout[q // n_in].matvec(X[q].contiguous()) for q in range(2 * n_in)]
    s2, st2 = finals(phis2, "c64d")
    torch.cuda.synchronize()
    out["batched_minres_c64d"] = {"solve_s": t_solve2, "total_s": time.perf_counter() - t0,
                                  "iterations_max": int(it2.max())}
    log("config2 batched MINRES: solve %.2f s, total %.2f s" % (t_solve2, out["batched_minres_c64d"]["total_s"]))
    del X, B
    # (3) per-RHS FP64 MINRES loop (the reference's procedure, matrices resident)
    t0 = time.perf_counter()
    phis3, its3 = [], []
    for p in range(2):
        nb = float(torch.linalg.vector_norm(b[p]))
        for k in range(n_in):
            A = c.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, omega=sigma[k], ham_sign=-1.0)
            x, it3, _ = c.minres(A, b[p], tol=1e-8 * nb, maxit=8000)
            phis3.append(Dout[p].matvec(x))
            its3.append(int(it3))
            del A
    torch.cuda.synchronize()
    t_solve3 = time.perf_counter() - t0
    s3 = []
    for q in range(2 * n_in):
        sp, _ = finals([phis3[q]], "c128")
        s3.append(sp[0])
    torch.cuda.synchronize()
    out["per_rhs_minres_fp64"] = {"solve_s": t_solve3, "total_s": time.perf_counter() - t0,
                                  "iterations_max": max(its3)}
    log("config2 per-RHS FP64 MINRES: solve %.2f s, total %.2f s" % (t_solve3, out["per_rhs_minres_fp64"]["total_s"]))
    out["multishift_vs_fp64_rel_l2_max"] = max(rel(a, r) for a, r in zip(s1, s3))
    out["batched_minres_vs_fp64_rel_l2_max"] = max(rel(a, r) for a, r in zip(s2, s3))
    out["speedup_multishift_vs_per_rhs_fp64"] = (out["per_rhs_minres_fp64"]["total_s"] /
                                                 out["multishift_c64d"]["total_s"])
    out["speedup_batched_minres_vs_per_rhs_fp64"] = (out["per_rhs_minres_fp64"]["total_s"] /
                                                     out["batched_minres_c64d"]["total_s"])
    out["spectra"] = 2 * n_in
    return out


def traffic_key(kernel, nrhs):
    """Name of a kernel instance in the traffic summaries: the kernel, with
    _k<K> for a block of K right-hand sides."""
    return kernel if nrhs == 1 else "%s_k%d" % (kernel, nrhs)


def measured_traffic(kernel, form, norb):
    """HBM bytes per launch of a kernel as two rocprofv3 counter runs measured
    it (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel calibration),
    from the summaries under profiles/; None when none covers this workload."""
    import glob
    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        with open(fn) as f:
            for line in f:
                t = line.split()
                if len(t) == 5 and t[0] == "traffic:" and t[1:4] == [kernel, form, "norb%d" % norb]:
                    return float(t[4])
    return None


def probe(args):
    """A short fixed workload for hardware-counter runs: the SpMV/SpMM launches
    of the step at full size (single-vector SpMV over H_n with the Krylov
    epilogue, the K=2 and K=3 blocks over H_n, the K=16 final-state block over
    H_i),
    each ``--launches`` times, plus a plain device copy of 1 GiB as the
    calibration kernel of known traffic."""
    import torch
    c = __import__("edrixs_amd.core", fromlist=["core"])
    from edrixs_amd import siam
    m = siam.SiamModel(nbath=args.nbath, seed=args.seed)
    bi = c.Basis.combinations(m.norb, m.nocc)
    bn = c.Basis.combinations(m.norb, m.nocc + 1)
    Hi = c.build_ham_i(bi, m.hop_i, m.coul_i)
    Hn = c.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout="interleaved")   # as the step
    f = args.probe_form
    Hi.prepare(f)
    Hn.prepare(f)
    g = torch.Generator(device="cuda").manual_seed(1)
    def rnd(*shape):
        return torch.randn(*shape, dtype=torch.complex128, device="cuda", generator=g)
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("probe: copy done")
    kb = c.KrylovBatch(Hn, rnd(1, Hn.n), f, capacity=args.launches + 1)   # k_spmv, Krylov epilogue
    kb.advance(args.launches)
    kb.close()
    torch.cuda.synchronize()
    log("probe: %d k_spmv Krylov steps on H_n done" % args.launches)
    for K in (2, 3):   # the RIXS solves' and the XAS vectors' blocks
        kb = c.KrylovBatch(Hn, rnd(K, Hn.n), f, capacity=args.launches + 1)
        kb.advance(args.launches)
        kb.close()
        torch.cuda.synchronize()
        log("probe: %d K=%d steps on H_n done" % (args.launches, K))
    kb = c.KrylovBatch(Hi, rnd(16, Hi.n), f, capacity=args.launches + 1)
    kb.advance(args.launches)
    kb.close()
    torch.cuda.synchronize()
    log("probe: %d K=16 steps on H_i done" % args.launches)
    print(json.dumps({"probe": f, "launches": args.launches, "nnz_H_n": Hn.nnz, "nnz_H_i": Hi.nnz,
                      "rows_H_n": Hn.m, "rows_H_i": Hi.m}))


def main():
    global DEADLINE
    args = parse()
    DEADLINE = T_START + args.budget
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, and the "
              "multi-GPU layout is a design note, not a measured path (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.cpu_check:
        print(json.dumps({"cpu_check": cpu_check(args)}))
        return
    import torch
    torch.zeros(1, device="cuda")
    if args.probe:
        probe(args)
        return
    if args.serve_only:
        print(json.dumps({"serving": serve(args)}))
        return
    if args.plain_only:
        print(json.dumps({"plain_fp64": plain_fp64(Workload(args), None, max(1, args.plain_pairs))}))
        return
    if args.config4:
        print(json.dumps({"config4": config4(args)}))
        return
    if args.config2:
        print(json.dumps({"config2": config2(args)}))
        return
    if args.eigs:
        print(json.dumps({"eigs": eigs(args)}))
        return
    work = Workload(args)
    forms = args.forms.split(",")
    main_form = forms[0]
    want_plain = "c128" in forms and args.plain_pairs > 0

    def reserve(t):
        r = sum(COST_STEPS.get(f, 1.6) for f in forms[1:]) * t
        r += COST_STEPS["plain"] * t * (args.plain_pairs / 4.0) if want_plain else 0.0
        r += COST_SECONDS["serve"] * (args.serve / 12.0) if args.serve > 0 else 0.0
        r += 0.0 if args.no_cpu else COST_SECONDS["cpu"]
        return r

    results, times, nsteps, profs, skipped = {}, {}, {}, {}, {}
    # the reported form first: its warm-up (at most one step, which holds the
    # schedule autotuning) also warms the device and library for the others
    warmup = min(args.warmup, 1)
    dt, k, out, prof = timed(work, main_form, args.steps, warmup, reserve)
    results[main_form], times[main_form], nsteps[main_form], profs[main_form] = out, dt, k, prof
    t_step = dt / k
    for f in forms[1:]:
        need = COST_STEPS.get(f, 1.6) * t_step
        if left() < need:
            skipped[f] = "budget: %.0f s left, about %.0f s needed" % (left(), need)
            log("skipping form %s (%s)" % (f, skipped[f]))
            continue
        dt, k, out, prof = timed(work, f, 1, 0)
        results[f], times[f], nsteps[f], profs[f] = out, dt, k, prof
    forms = [f for f in forms if f in results]
    nspec = 1 + 2 * args.n_in
    t_main = times[main_form]
    # the dominant kernel: the SpMV/SpMM variant and count of right-hand sides
    # (one template instance) with the most kernel time
    by_rhs = profs[main_form]["by_rhs"]
    kname, krhs = max(by_rhs, key=lambda kr: by_rhs[kr][1])
    n, ms, by, aby = by_rhs[(kname, krhs)]
    gbs = aby / (ms * 1e-3) / 1e9 if ms > 0 else 0.0          # algorithmic (CSR) bytes
    sgbs = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0          # the stored form's bytes
    m = work.model
    forms_out = {}
    ref = results.get("c128")
    for f in forms:
        r = results[f]
        kern = {}
        for k, v in profs[f].items():
            if k == "by_rhs":
                continue
            pn, pms, pby, paby = v
            if pn:
                kern[k] = {"launches": pn, "total_s": pms * 1e-3, "avg_us": pms / pn * 1e3,
                           "alg_gbs": paby / (pms * 1e-3) / 1e9 if pms > 0 else 0.0,
                           "stored_gbs": pby / (pms * 1e-3) / 1e9 if pms > 0 else 0.0}
        inst = {}
        for (k, nr), (pn, pms, pby, paby) in sorted(profs[f]["by_rhs"].items()):
            inst[traffic_key(k, nr)] = {"launches": pn, "total_s": pms * 1e-3,
                                        "avg_us": pms / pn * 1e3,
                                        "alg_gbs": paby / (pms * 1e-3) / 1e9 if pms > 0 else 0.0,
                                        "stored_gbs": pby / (pms * 1e-3) / 1e9 if pms > 0 else 0.0}
        d = {"spectra_per_s": nspec * nsteps[f] / times[f],
             "ms_per_step": times[f] / nsteps[f] * 1e3, "steps": nsteps[f],
             "phases_s": r["info"]["phases_s"],
             "spmv_kernels": kern,
             "spmv_instances": inst,
             "lanczos_iters": r["info"]["lanczos_iters"],
             "lowest_states_eV": r["info"]["lowest_states"],
             "rixs_iters_max": max(r["info"]["rixs_iters"]),
             "rixs_krylov_steps_max": max(r["info"]["rixs_krylov_steps"]),
             "rixs_krylov_steps": r["info"]["rixs_krylov_steps"],
             "rixs_iters": r["info"]["rixs_iters"],
             "rixs_cf_change_max": r["info"]["rixs_cf_change_max"],
             "xas_krylov_steps_max": max(r["info"]["xas_krylov_steps"]),
             "xas_cf_change_max": r["info"]["xas_cf_change_max"],
             "rixs_rel_residual_max": r["info"]["rixs_res_rel_max"]}
        if ref is not None and f != "c128":
            d["ground_energy_abs_err_eV"] = abs(r["egs"] - ref["egs"])
            d["lowest_states_abs_err_eV"] = float(np.max(np.abs(
                np.array(r["info"]["lowest_states"]) - np.array(ref["info"]["lowest_states"]))))
            d["xas_rel_l2_vs_fp64"] = rel(r["xas"], ref["xas"])
            d["rixs_rel_l2_vs_fp64_max"] = max(rel(a, b) for a, b in zip(r["rixs"], ref["rixs"]))
        forms_out[f] = d
    line = {
        "metric": METRIC,
        "value": nspec * nsteps[main_form] / t_main,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": nsteps[main_form],
        "steps_requested": args.steps,
        "warmup": warmup,
        "warmup_requested": args.warmup,
        "budget_s": args.budget,
        "ms_per_step": t_main / nsteps[main_form] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"c128i": "f64 (complex128): exact values through a table of the distinct "
                           "off-diagonal values (value number per nonzero, packed with its "
                           "column in one 32-bit word), f64 diagonal",
                  "c64d": "c64 off-diagonal values, f64 diagonal/vectors/accumulation/scalars",
                  "c64": "c64 values and vectors, f32 accumulation, f64 scalars",
                  "c128": "f64 (complex128)"}[main_form],
        "data": "synthetic",
        "config": {
            "workload": ("norb=24 N=12 SIAM (10 d + 14 bath, 2p core), end to end: CSR builds of "
                         "H_i (%d rows) and H_n (%d rows), ground state, isotropic XAS (3 batched "
                         "Krylov recurrences, >= %d steps, to converged continued fractions) and RIXS at %d incident energies x 2 polarizations "
                         "(one multi-shift Galerkin recurrence over H_n for both polarizations, a 2-vector "
                         "block where H_n has the slot-factored form, projected by D_out, + batched "
                         "final-state Krylov over H_i to converged continued fractions), "
                         "continued fractions on %d points"
                         % (work.bi.n, work.bn.n * m.ncore, args.nkryl, args.n_in, args.nw)),
            "form": main_form, "spectra_per_step": nspec,
            "nnz_H_i": results[main_form]["info"]["nnz_i"],
            "nnz_H_n": results[main_form]["info"]["nnz_n"],
        },
        "roofline": {
            "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_key(kname, krhs), main_form, m.norb),
            "kernel": "%s, %d right-hand side%s (%s form), every such launch of the timed region"
                      % (kname, krhs, "" if krhs == 1 else "s", main_form),
            "nrhs": krhs,
            "launches": n, "avg_us": ms / max(n, 1) * 1e3, "bytes_per_launch": aby / max(n, 1),
            "bytes_counted": ("algorithmic bytes of the product as the complex128 CSR of the same "
                              "FP64 operator: nnz*(16+4) + 8*(m+1) + K*(16*n + 16*m*(1+a)), a = 1 "
                              "(the Krylov epilogue accumulates into y)"),
            "stored_form": {
                "achieved": sgbs, "frac": sgbs / HBM_PEAK_GBS, "unit": "GB/s",
                "bytes_per_launch": by / max(n, 1),
                "bytes_counted": ("the bytes the launch's storage form reads and writes: index "
                                  "words, row pointers, value and type tables, diagonal, and per "
                                  "right-hand side x once, y written, the epilogue's addend and "
                                  "dot vector")},
        },
        "forms": forms_out,
    }
    if "c128" in results:
        line["speedup_vs_fp64"] = (times["c128"] / nsteps["c128"]) / (t_main / nsteps[main_form])
    if not args.no_cpu:
        need = COST_SECONDS["cpu"]
        rows = args.cpu_rows if left() >= need else max(20000, int(args.cpu_rows * left() / need))
        cb = cpu_baseline(work, results[main_form]["info"], rows)
        line["cpu_baseline"] = cb
        line["gpu_vs_cpu"] = line["value"] / cb["value"]
    if args.serve > 0:
        need = COST_SECONDS["serve"] * args.serve / 12.0
        if left() >= need:
            line["serving"] = serve(args)
        else:
            skipped["serving"] = "budget: %.0f s left, about %.0f s needed" % (left(), need)
    if want_plain:
        need = COST_STEPS["plain"] * t_step * args.plain_pairs / 4.0
        if left() >= need:
            pl = plain_fp64(work, results["c128"], args.plain_pairs)
            line["plain_fp64"] = pl
            line["speedup_vs_plain_fp64"] = line["value"] / pl["spectra_per_s"]
        else:
            skipped["plain_fp64"] = "budget: %.0f s left, about %.0f s needed" % (left(), need)
    if args.exact and "c128" in results:
        line["accuracy_vs_exact_projection"] = exact_reference(work, results["c128"]["gs"], results)
    if skipped:
        line["skipped"] = skipped
    line["wall_s"] = time.monotonic() - T_START
    print(json.dumps(line))


if __name__ == "__main__":
    main()
