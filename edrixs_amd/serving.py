"""Serving loop: mixed XAS / RIXS requests on one GPU with continuous batching.

Each request is an impurity model of its own size and parameters.  Requests
arrive over time; a request is admitted as soon as the HBM it will hold fits
in the budget next to what the device holds now, and leaves as soon as its
spectra have converged, freeing its memory for the next one.  What the device
holds is measured, not estimated: the library's own allocations
(``edx_mem_in_use``) plus torch's.  What a new request will hold is estimated
from its dimensions, scaled by what earlier requests of the same size class
were measured to take (the bytes the device gained over their setup).  Every Krylov
recurrence of every running request (the three XAS dipole components, the
RIXS final states of each (incident energy, polarization)) is a member of
the :class:`core.KrylovGroup` of its request's size class (norb): a step of
all of them is one SpMV launch over all their rows, so requests that arrive
while others run share the device instead of queueing behind them.  Each
round gives every size class about ``slice_s`` of device time (its step
count is set from its measured step time), so small requests are not held to
the pace of the largest ones.

Per-request work that is not a Krylov step -- CSR builds, the ground state
(Lanczos), the edge estimate and the RIXS multi-shift solve -- runs on a
worker thread with its own stream (default priority), one request at a time,
while this thread keeps stepping the groups; the finished setup's recurrences
join their group between rounds.  The smallest queued request is set up first
unless one has waited longer than ``max_wait_s``.

Results do not depend on the mix: a member's coefficients are the same in any
group, and a spectrum is taken at the first multiple of ``check_every`` of
the member's OWN step count at which it changed by less than ``cf_tol``
(the recurrence's whole (alpha, beta) history stays available, so the check
does not depend on when the group's steps happened).
"""
import dataclasses
import math
import time

import numpy as np

from . import core, siam

GAMMA_C = 0.5
GAMMA_F = 0.1
LAYOUT = "interleaved"   # core-hole index valence * 6 + slot (permuted operators, same spectra)


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                  # "xas" or "rixs"
    norb: int
    seed: int
    arrival: float             # seconds after the start of the run
    F2: float = 10.0
    ten_dq: float = 1.2
    V: float = 1.5
    admitted: float = math.nan
    done: float = math.nan
    spectra: list = dataclasses.field(default_factory=list)
    steps: list = dataclasses.field(default_factory=list)
    egs: float = math.nan

    @property
    def latency(self):
        return self.done - self.arrival

    @property
    def nspectra(self):
        return len(self.spectra)


def make_requests(n, rate=20.0, sizes=(16, 18, 20, 22, 24), rixs_fraction=0.5, seed=0):
    """n requests with Poisson arrivals (mean ``rate`` per second), norb drawn
    from ``sizes``, N = norb / 2, and per-request model parameters:
    F2 in [8, 12] eV, 10Dq in [0.5, 2] eV, V in [1, 2] eV, bath energies in
    [-3, 3] eV (seeded per request)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    out = []
    for i in range(n):
        out.append(Request(rid=i, kind="rixs" if rng.random() < rixs_fraction else "xas",
                           norb=int(rng.choice(sizes)), seed=int(rng.integers(1 << 30)),
                           arrival=float(t[i]), F2=float(rng.uniform(8, 12)),
                           ten_dq=float(rng.uniform(0.5, 2.0)), V=float(rng.uniform(1.0, 2.0))))
    return out


def _dims(norb):
    n_i = math.comb(norb, norb // 2)
    n_n = 6 * math.comb(norb, norb // 2 + 1)
    return n_i, n_n


class Engine:
    """Continuous-batching XAS / RIXS server on one GPU."""

    def __init__(self, budget_gb=64.0, precision="c128i", check_every=100, chunk=25,
                 cf_tol=1e-6, min_steps=200, max_steps=4000, n_in=4, npol=1, nw=1000,
                 rixs_tol=1e-6, rixs_maxit=4000, slice_s=0.1, max_wait_s=10.0, overlap=True):
        self.budget = budget_gb * 1e9
        self.overlap = overlap
        self._stream = None     # the worker's stream for setups
        self.precision = precision
        self.check_every = check_every
        self.chunk = chunk
        self.cf_tol = cf_tol
        self.min_steps = min_steps
        self.max_steps = max_steps
        self.n_in = n_in
        self.npol = npol
        self.omega = np.linspace(-10.0, 30.0, nw)
        self.eloss = np.linspace(0.0, 10.0, nw)
        self.rixs_tol = rixs_tol
        self.rixs_maxit = rixs_maxit
        self.slice_s = slice_s
        self.max_wait_s = max_wait_s
        self.groups = {}        # norb -> KrylovGroup of that size class
        self.step_s = {}        # norb -> measured seconds per group step
        self.used = 0.0         # bytes the running requests were measured to take
        self.scale = {}         # norb -> measured / estimated bytes of past setups
        self.device_bytes = self._device_bytes

    def group_for(self, norb):
        if norb not in self.groups:
            self.groups[norb] = core.KrylovGroup(self.precision, max_steps=self.max_steps)
        return self.groups[norb]

    @property
    def members(self):
        return [m for g in self.groups.values() for m in g.members]

    # ------------------------------------------------------------ per request
    def estimate_bytes(self, req):
        """HBM a request holds while it runs: its CSR matrices (nonzeros per row
        from the model's term counts, bounded) and its Krylov vectors."""
        n_i, n_n = _dims(req.norb)
        per_nnz = {"c64d": 32.0, "c128i": 34.0}.get(self.precision, 20.0)   # CSR + its copies
        mats = (n_i * 18 + n_n * 26) * per_nnz + (n_i + n_n) * 32.0
        vecs = 16.0 * (2 * 3 * n_n if req.kind == "xas"
                       else 2 * self.n_in * self.npol * n_i + 6 * n_n + 2 * self.n_in * n_i)
        return mats + vecs

    def need_bytes(self, req):
        """Estimated bytes of a request, scaled by the measured/estimated ratio
        of earlier requests of its size class (1 until one was measured)."""
        return self.estimate_bytes(req) * self.scale.get(req.norb, 1.0)

    @staticmethod
    def _device_bytes():
        """Device memory in use now: the library's allocations plus torch's."""
        import torch
        return float(core.mem_in_use()) + float(torch.cuda.memory_allocated())

    def pick(self, queue, running, now, in_use):
        """The queued request to admit now, or None.  The oldest request that
        has waited longer than max_wait_s goes first, else the smallest; it is
        admitted if nothing runs or if in_use plus its need fits the budget
        (a request larger than the whole budget thus runs alone)."""
        if not queue:
            return None
        old = [r for r in queue if now - r.arrival > self.max_wait_s]
        req = min(old, key=lambda r: r.arrival) if old else \
            min(queue, key=lambda r: (r.norb, r.arrival))
        if running == 0 or in_use + self.need_bytes(req) <= self.budget:
            return req
        return None

    def _model(self, req):
        f2 = req.F2
        return siam.SiamModel(nbath=req.norb - 10, nocc=req.norb // 2, seed=req.seed,
                              F=(8.0, f2, 0.62 * f2), ten_dq=req.ten_dq, V=req.V)

    def prepare(self, req):
        """Builds, ground state and (RIXS) the intermediate-state solve: the
        request's matrices and the start vectors of its recurrences.  Touches
        no group, so it may run on a worker thread (its own stream)."""
        m = self._model(req)
        p = self.precision
        bi = core.Basis.combinations(m.norb, m.nocc)
        bn = core.Basis.combinations(m.norb, m.nocc + 1)
        Hi = core.build_ham_i(bi, m.hop_i, m.coul_i)
        ev, vecs, _ = core.ground_state(Hi, nev=1, precision=p, maxiter=400, tol=1e-10)
        egs = float(ev[0])
        gs = vecs[0].contiguous()
        req.egs = egs
        Hn = core.build_ham_n(bn, m.norb, m.ncore, m.hop_n, m.coul_n, layout=LAYOUT)
        st = {"req": req, "mats": [Hi, Hn], "starts": [], "grid": None, "gamma": None}
        if req.kind == "xas":
            for k in range(3):
                D = core.build_transop_i(bn, bi, m.norb, m.ncore, m.trans_i[k], layout=LAYOUT)
                st["starts"].append((Hn, D.matvec(gs)))
            st["grid"], st["gamma"] = self.omega, GAMMA_C
            st["combine"] = "average"
        else:
            # incident energies: a 10 eV window from 2 eV below the absorption
            # edge (lowest core-hole state), estimated by a short Lanczos run
            e0n, _, _ = core.lanczos_eigs(Hn, 1, maxiter=120, tol=1e-6, precision=p)
            edge = float(e0n[0]) - egs
            w_in = np.linspace(edge - 2.0, edge + 8.0, self.n_in)
            sigma = w_in + egs + 1j * GAMMA_C
            for pol in range(self.npol):
                ein = np.eye(3)[pol % 3]
                eout = np.eye(3)[(pol + 1) % 3]
                tin = m.pol_trans(ein)[0]
                tout = m.pol_trans(eout)[1]
                Din = core.build_transop_i(bn, bi, m.norb, m.ncore, tin, layout=LAYOUT)
                Dout = core.build_transop_f(bi, bn, m.norb, m.ncore, tout, layout=LAYOUT)
                (Y,), _, _ = core.shifted_fom(Hn, Din.matvec(gs), sigma, projections=(Dout,),
                                              tol=self.rixs_tol, maxit=self.rixs_maxit,
                                              precision=p)
                for k in range(len(sigma)):
                    st["starts"].append((Hi, Y[k].contiguous()))
            st["grid"], st["gamma"] = self.eloss, GAMMA_F
            st["combine"] = "each"
            st["mats"] = [Hi]          # H_n no longer needed
        return st

    def attach(self, st):
        """The request's recurrences join its size class's group (main thread)."""
        import torch
        group = self.group_for(st["req"].norb)
        st["group"] = group
        st["members"] = [group.add(H, v) for H, v in st["starts"]]
        # the group copied the start vectors on this thread's stream: let the
        # copies finish before the vectors' memory can go back to the worker
        torch.cuda.current_stream().synchronize()
        st["starts"] = []
        st["prev"] = {mid: None for mid in st["members"]}
        st["checked"] = {mid: 0 for mid in st["members"]}
        st["result"] = {}
        return st

    def setup(self, req):
        """prepare() and attach() on this thread."""
        return self.attach(self.prepare(req))

    def _prepare_on_stream(self, req):
        """prepare() on the calling (worker) thread's own stream, finished
        before it returns; also the device bytes the setup took."""
        import torch
        if self._stream is None:
            self._stream = torch.cuda.Stream()     # default priority
        s = self._stream
        core.set_stream(s)
        try:
            with torch.cuda.stream(s):
                before = self.device_bytes()
                st = self.prepare(req)
                s.synchronize()
                st["took"] = max(0.0, self.device_bytes() - before)
        finally:
            core.set_stream(None)
        return st

    def poll(self, st):
        """Check the request's members at their own step multiples; True once
        every spectrum has converged (or its recurrence ended)."""
        ce = self.check_every
        for mid in st["members"]:
            if mid in st["result"]:
                continue
            a, b, nrm, running = st["group"].coefficients(mid)
            steps = len(a)
            while st["checked"][mid] + ce <= steps or (not running and st["checked"][mid] < steps):
                c = min(st["checked"][mid] + ce, steps)
                s = core.cf_spectrum(a[:c], b[:c], nrm, st["req"].egs, st["grid"], st["gamma"])
                prev = st["prev"][mid]
                st["checked"][mid] = c
                st["prev"][mid] = s
                ended = c == steps and not running
                conv = False
                if prev is not None and c >= self.min_steps:
                    den = np.linalg.norm(s)
                    conv = den == 0 or np.linalg.norm(s - prev) <= self.cf_tol * den
                if conv or ended or c >= self.max_steps:
                    st["result"][mid] = (s, c)
                    break
        return len(st["result"]) == len(st["members"])

    def finish(self, st):
        req = st["req"]
        res = [st["result"][mid] for mid in st["members"]]
        if st["combine"] == "average":
            req.spectra = [sum(s for s, _ in res) / len(res)]
        else:
            req.spectra = [s for s, _ in res]
        req.steps = [c for _, c in res]
        for mid in st["members"]:
            st["group"].remove(mid)
        st["mats"].clear()

    # -------------------------------------------------------------- the loop
    def run(self, requests, realtime=True, log=None):
        """Serve ``requests`` (arrival times in seconds from now).  With
        realtime=False every request is present at the start (for tests).
        With ``overlap`` (the default) one request's setup runs on a worker
        thread and its own stream while the groups keep stepping; otherwise
        setups run between rounds on this thread."""
        import torch
        from concurrent.futures import ThreadPoolExecutor
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, active, done = [], [], []
        t0 = time.perf_counter()
        pool = ThreadPoolExecutor(max_workers=1) if self.overlap else None
        inflight = None           # the setup running on the worker

        def now():
            return time.perf_counter() - t0

        def admit(st, took):
            req = st["req"]
            est = self.estimate_bytes(req)
            if est > 0 and took > 0:
                self.scale[req.norb] = max(self.scale.get(req.norb, 0.0), took / est)
            st["bytes"] = took
            self.used += took
            active.append(self.attach(st))
            if log:
                log("admit r%d %s norb=%d at %.2f s, ready at %.2f s (%d running)" %
                    (req.rid, req.kind, req.norb, req.admitted, now(), len(active)))

        try:
            while pending or queue or active or inflight is not None:
                t = now()
                while pending and (not realtime or pending[0].arrival <= t):
                    queue.append(pending.pop(0))
                # a finished setup joins its group; with nothing else to step,
                # wait for it
                if inflight is not None and (inflight.done() or not active):
                    admit(inflight.result(), inflight.result()["took"])
                    inflight = None
                # admission against the HBM budget by the bytes the device
                # holds now, one setup at a time
                if inflight is None and queue:
                    req = self.pick(queue, len(active), now(), self.device_bytes())
                    if req is not None:
                        queue.remove(req)
                        req.admitted = now()
                        if pool is not None:
                            inflight = pool.submit(self._prepare_on_stream, req)
                        else:
                            before = self.device_bytes()
                            st = self.prepare(req)
                            torch.cuda.synchronize()
                            admit(st, max(0.0, self.device_bytes() - before))
                if active:
                    # every size class gets about slice_s of device time per
                    # round: small requests take many cheap steps while large
                    # ones take few
                    for norb in sorted({st["req"].norb for st in active}):
                        g = self.groups[norb]
                        per = self.step_s.get(norb)
                        n = self.chunk if per is None else int(min(1000, max(1, self.slice_s / per)))
                        t1 = time.perf_counter()
                        g.advance(n)
                        dt = (time.perf_counter() - t1) / n
                        self.step_s[norb] = dt if per is None else 0.5 * (per + dt)
                    still = []
                    for st in active:
                        if self.poll(st):
                            self.finish(st)
                            torch.cuda.current_stream().synchronize()
                            st["req"].done = now()
                            self.used -= st["bytes"]
                            done.append(st["req"])
                            if log:
                                r = st["req"]
                                log("done r%d %s norb=%d latency %.2f s" % (r.rid, r.kind, r.norb,
                                                                           r.latency))
                        else:
                            still.append(st)
                    active = still
                elif inflight is None and pending and not queue:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
        return done


def report(done):
    """Latency percentiles and throughput of a served set."""
    lat = np.array([r.latency for r in done])
    nspec = sum(r.nspectra for r in done)
    span = max(r.done for r in done) - min(r.arrival for r in done)
    return {"requests": len(done), "spectra": nspec, "makespan_s": span,
            "spectra_per_s": nspec / span if span > 0 else float("nan"),
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max())}
