"""Continuous-batching impurity-energy server on one GPU.

Requests (N atoms, Morse D/re/a, impurity position) arrive over time.  Each is
routed to a grid class n in {64, 96, 128} at spacing 0.75 A (box half-width
0.375 n A): the smallest class whose box holds a droplet of radius
R = 2.22 N^(1/3) A plus 10 A of vacuum on each side.  Every class owns one
context with a fixed number of slots, sized from an HBM budget with
effpot_footprint_bytes.  A request is admitted into a free slot as soon as it
has arrived; its state stays on the device until it converges (FP32 phase,
then FP64 polish); the loop of each class runs in its own host thread, so the
classes' kernels overlap on the GPU through their own streams.  Each slot's
trajectory depends only on its own request (deterministic per-slot kernels and
reductions), so a result is the same alone and in the mix.

A request whose mean field turns out too stiff for the class time step (its
energy stops being finite) is restarted from its initial state with half the
time step, in the same slot and without touching the other slots
(effpot_set_dt), up to `max_retries` times.

There is no reference counterpart: the reference runs one density file per
process (loop.sh:9-13).
"""
import heapq
import math
import threading
import time
from dataclasses import dataclass, field

from . import Context, SolveOpts, make_params, make_morse, default_dt, footprint_bytes
from . import SLOT_START, SLOT_DONE, SLOT_FAILED, CORE_OTC, PREC_MIXED

GRID_CLASSES = (64, 96, 128)
SPACING = 0.75


@dataclass
class Request:
    rid: int
    natoms: float
    D: float
    re: float
    a: float
    pos: tuple
    arrival: float = 0.0
    seed: int = 0
    # filled by the server
    grid: int = 0
    start: float = 0.0
    finish: float = 0.0
    eimp: float = float("nan")
    etot: float = float("nan")
    steps: int = 0
    converged: bool = False
    failed: bool = False
    dt: float = 0.0       # time step of the final attempt
    retries: int = 0


def radius(natoms):
    return 2.22 * natoms ** (1.0 / 3.0)


def grid_for(natoms):
    R = radius(natoms)
    for n in GRID_CLASSES:
        if 2 * R + 20.0 <= SPACING * n:
            return n
    return None


def class_params(n, slots, precision=PREC_MIXED, h_core=2.377):
    return make_params(n, xmax=SPACING * n / 2, n4=300.0, core=CORE_OTC, h_core=h_core, precision=precision,
                       batch=slots, psi_complex=0)


def plan_slots(hbm_budget_bytes, weights=None, precision=PREC_MIXED, max_slots=64):
    """Slots per grid class so that all contexts fit the HBM budget; the budget
    is split in proportion to `weights` (expected share of requests)."""
    weights = weights or {64: 1.0, 96: 1.0, 128: 1.0}
    tot = sum(weights.values())
    plan = {}
    for n in GRID_CLASSES:
        share = hbm_budget_bytes * weights.get(n, 0.0) / tot
        per = footprint_bytes(class_params(n, 1, precision))
        plan[n] = max(1, min(max_slots, int(share // per)))
    return plan


class _Lane:
    """One grid class: a context, its slots and the thread that ticks it."""

    def __init__(self, n, slots, opts_kw, precision, max_retries=2):
        self.n = n
        self.p = class_params(n, slots, precision)
        self.ctx = Context(self.p)
        self.B = slots
        self.slots = self.ctx.new_slots()
        self.req = [None] * slots
        for s in range(slots):
            self.ctx.set_active(s, False)
        self.opts = SolveOpts(dt=default_dt(self.p), **opts_kw)
        self.precision = precision
        self.max_retries = max_retries
        self.queue = []          # arrived, waiting for a slot
        self.lock = threading.Lock()
        self.done = []

    def free_slots(self):
        return [s for s in range(self.B) if self.req[s] is None]

    def admit(self, req, now):
        s = self.free_slots()[0]
        req.dt = self.opts.dt
        req.retries = 0
        req.start = now
        self._load(s, req)

    def _load(self, s, req):
        self.ctx.set_dt(req.dt, slot=s)
        self.ctx.set_active(s, True)
        self.ctx.set_natoms(req.natoms, slot=s)
        self.ctx.init_fermi(0.0218, radius(req.natoms), 0.6, 1e-6, req.seed, slot=s)
        self.ctx.set_impurity(make_morse(rimp=req.pos, D=req.D, re=req.re, a=req.a), slot=s)
        self.slots[s].phase = SLOT_START
        self.slots[s].fp32 = 1 if self.precision == PREC_MIXED else 0
        self.slots[s].steps = 0
        self.slots[s].converged = 0
        req.grid = self.n
        self.req[s] = req

    def step(self, t0, progress=None):
        busy = any(r is not None for r in self.req)
        if busy:
            self.ctx.tick(self.opts, self.slots)
        now = time.perf_counter() - t0
        for s in range(self.B):
            r = self.req[s]
            if r is None:
                continue
            if self.slots[s].phase == SLOT_FAILED and r.retries < self.max_retries:
                r.retries += 1
                r.dt *= 0.5
                self._load(s, r)
                continue
            if self.slots[s].phase in (SLOT_DONE, SLOT_FAILED):
                r.finish = now
                r.eimp = self.slots[s].e.eimpu
                r.etot = self.slots[s].e.etot
                r.steps = self.slots[s].steps
                r.converged = bool(self.slots[s].converged)
                r.failed = self.slots[s].phase == SLOT_FAILED
                self.done.append(r)
                if progress:
                    progress(r)
                self.req[s] = None
                self.ctx.set_active(s, False)
        return busy


class Server:
    def __init__(self, slots_per_class, tol=1e-9, tol_fp32=1e-7, max_steps=60000, pener=50,
                 precision=PREC_MIXED, max_retries=2):
        kw = dict(tol=tol, n_rk=3, pener=pener, max_steps=max_steps, polish_steps=max_steps,
                  tol_fp32=tol_fp32, check_finite=0)
        self.lanes = {n: _Lane(n, k, kw, precision, max_retries) for n, k in slots_per_class.items() if k > 0}

    def close(self):
        for ln in self.lanes.values():
            ln.ctx.close()

    def run(self, requests, progress=None):
        """Serve `requests` (with .arrival seconds from start) in real time.
        Returns them completed, with start/finish/latency fields set;
        `progress(request)` is called as each one completes."""
        pending = [(r.arrival, r.rid, r) for r in requests]
        heapq.heapify(pending)
        for r in requests:
            r.grid = grid_for(r.natoms) or 0
            if r.grid == 0:
                r.failed = True
        total = sum(1 for r in requests if not r.failed)
        plock = threading.Lock()
        t0 = time.perf_counter()

        def lane_loop(ln):
            while True:
                now = time.perf_counter() - t0
                with plock:
                    keep = []
                    while pending and pending[0][0] <= now:
                        it = heapq.heappop(pending)
                        if it[2].grid == ln.n:
                            ln.queue.append(it[2])
                        else:
                            keep.append(it)
                    for it in keep:
                        heapq.heappush(pending, it)
                    no_more = not any(it[2].grid == ln.n for it in pending)
                while ln.queue and ln.free_slots():
                    ln.admit(ln.queue.pop(0), time.perf_counter() - t0)
                busy = ln.step(t0, progress)
                if not busy:
                    if no_more and not ln.queue:
                        return
                    time.sleep(0.001)

        threads = [threading.Thread(target=lane_loop, args=(ln,)) for ln in self.lanes.values()]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        done = [r for ln in self.lanes.values() for r in ln.done]
        assert len(done) == total, (len(done), total)
        return requests


def make_stream(nreq, rate, seed=0, pos_frac=1.5):
    """Poisson arrivals at `rate` per second; N log-uniform in [100, 3000];
    D ~ U[10, 60] K, re ~ U[3.5, 5] A, a ~ U[0.8, 1.4]/A; position uniform in
    the ball of radius pos_frac R (clamped to 8 A inside the box)."""
    import random
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(nreq):
        t += rng.expovariate(rate)
        N = math.exp(rng.uniform(math.log(100.0), math.log(3000.0)))
        R = radius(N)
        while True:
            v = [rng.uniform(-1, 1) for _ in range(3)]
            if sum(x * x for x in v) <= 1.0:
                break
        n = grid_for(N)
        lim = SPACING * n / 2 - 8.0 if n else 0.0
        pos = tuple(max(-lim, min(lim, pos_frac * R * x)) for x in v)
        out.append(Request(rid=i, natoms=N, D=rng.uniform(10, 60), re=rng.uniform(3.5, 5.0),
                           a=rng.uniform(0.8, 1.4), pos=pos, arrival=t, seed=1000 + i))
    return out


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q
    f, c = math.floor(k), math.ceil(k)
    return xs[f] + (xs[c] - xs[f]) * (k - f)
