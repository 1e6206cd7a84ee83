"""Continuous-batching impurity-energy server on one GPU.

Requests (N atoms, Morse D/re/a, impurity position) arrive over time.  Each is
routed to a grid class n in {64, 96, 128} at spacing 0.75 A (box half-width
0.375 n A): the smallest class whose box holds a droplet of radius
R = 2.22 N^(1/3) A plus 10 A of vacuum on each side.

Admission is by HBM footprint.  The slots of a grid class live in blocks:
contexts of `block_slots[n]` slots whose device footprint
(effpot_footprint_bytes) is reserved from the server's HBM budget when the
block is created and returned when it is destroyed.  An arrived request takes
a free slot of its class; if there is none, a new block is created when the
budget still holds its footprint, else the request waits; a block whose slots
have all finished is kept for the next request of its class, unless another
class waits for budget: then it is destroyed and its memory goes there.  A
request's state stays on the device from admission to convergence (FP32
phase, then FP64 polish).  The blocks of each class are ticked by one host
thread, so the classes' kernels overlap on the GPU through their own streams.  Each slot's trajectory depends
only on its own request (deterministic per-slot kernels and reductions), so a
result is the same alone and in the mix, whatever block it lands in.

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
from dataclasses import dataclass

from . import Context, SolveOpts, make_params, make_morse, default_dt, footprint_bytes
from . import SLOT_START, SLOT_DONE, SLOT_FAILED, CORE_OTC, PREC_MIXED

GRID_CLASSES = (64, 96, 128)
SPACING = 0.75
# slots per block: one 128^3 mixed slot holds ~0.9 GB, a 64^3 one ~0.1 GB
BLOCK_SLOTS = {64: 8, 96: 4, 128: 2}


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
    is split in proportion to `weights` (expected share of requests).  For a
    fixed plan (Server(slots_per_class=...)) instead of admission by footprint."""
    weights = weights or {64: 1.0, 96: 1.0, 128: 1.0}
    tot = sum(weights.values())
    plan = {}
    for n in GRID_CLASSES:
        share = hbm_budget_bytes * weights.get(n, 0.0) / tot
        per = footprint_bytes(class_params(n, 1, precision))
        plan[n] = max(1, min(max_slots, int(share // per)))
    return plan


class Budget:
    """HBM bytes reserved by the blocks alive, against a fixed total, and the
    grid classes that wait for some of it."""

    def __init__(self, total):
        self.total = total
        self.used = 0
        self.peak = 0
        self.starved = set()
        self.lock = threading.Lock()

    def take(self, nbytes, who):
        with self.lock:
            if self.used + nbytes > self.total:
                self.starved.add(who)
                return False
            self.used += nbytes
            self.peak = max(self.peak, self.used)
            self.starved.discard(who)
            return True

    def give(self, nbytes):
        with self.lock:
            self.used -= nbytes

    def others_waiting(self, who):
        with self.lock:
            return bool(self.starved - {who})

    def not_waiting(self, who):
        with self.lock:
            self.starved.discard(who)


class _Block:
    """One context of a grid class: `B` slots stepped together."""

    def __init__(self, n, B, precision, nbytes):
        self.p = class_params(n, B, precision)
        self.ctx = Context(self.p)
        self.B = B
        self.nbytes = nbytes
        self.slots = self.ctx.new_slots()
        self.req = [None] * B
        for s in range(B):
            self.ctx.set_active(s, False)

    def free_slots(self):
        return [s for s in range(self.B) if self.req[s] is None]

    def idle(self):
        return all(r is None for r in self.req)


class _Lane:
    """One grid class: its blocks, its queue and the thread that ticks them."""

    def __init__(self, n, block_slots, opts_kw, precision, budget, fixed=False, max_retries=2):
        self.n = n
        self.block_slots = block_slots
        self.precision = precision
        self.budget = budget
        self.fixed = fixed       # one block created up front and kept (a fixed slot plan)
        self.opts = SolveOpts(dt=default_dt(class_params(n, 1, precision)), **opts_kw)
        self.block_bytes = footprint_bytes(class_params(n, block_slots, precision))
        self.max_retries = max_retries
        self.blocks = []
        self.queue = []          # arrived, waiting for a slot
        self.done = []
        self.created = 0
        if fixed:
            self._new_block()

    def _new_block(self):
        if not self.budget.take(self.block_bytes, self.n):
            return None
        try:
            b = _Block(self.n, self.block_slots, self.precision, self.block_bytes)
        except Exception:
            self.budget.give(self.block_bytes)
            raise
        self.blocks.append(b)
        self.created += 1
        return b

    def admit_waiting(self, now):
        """Admit queued requests into free slots, creating blocks while the
        budget holds them; returns False when a request cannot ever fit."""
        while self.queue:
            blk = next((b for b in self.blocks if b.free_slots()), None)
            if blk is None and not self.fixed:
                if self.block_bytes > self.budget.total:
                    return False
                blk = self._new_block()
            if blk is None:
                return True  # wait for a slot or for budget
            req = self.queue.pop(0)
            req.dt = self.opts.dt
            req.retries = 0
            req.start = now
            self._load(blk, blk.free_slots()[0], req)
        return True

    def _load(self, blk, s, req):
        c = blk.ctx
        c.set_dt(req.dt, slot=s)
        c.set_active(s, True)
        c.set_natoms(req.natoms, slot=s)
        c.init_fermi(0.0218, radius(req.natoms), 0.6, 1e-6, req.seed, slot=s)
        c.set_impurity(make_morse(rimp=req.pos, D=req.D, re=req.re, a=req.a), slot=s)
        blk.slots[s].phase = SLOT_START
        blk.slots[s].fp32 = 1 if self.precision == PREC_MIXED else 0
        blk.slots[s].steps = 0
        blk.slots[s].converged = 0
        req.grid = self.n
        blk.req[s] = req

    def step(self, t0, progress=None):
        busy = False
        for blk in self.blocks:
            if not blk.idle():
                busy = True
                blk.ctx.tick(self.opts, blk.slots)
        now = time.perf_counter() - t0
        for blk in self.blocks:
            for s in range(blk.B):
                r = blk.req[s]
                if r is None:
                    continue
                ph = blk.slots[s].phase
                if ph == SLOT_FAILED and r.retries < self.max_retries:
                    r.retries += 1
                    r.dt *= 0.5
                    self._load(blk, s, r)
                    continue
                if ph in (SLOT_DONE, SLOT_FAILED):
                    r.finish = now
                    r.eimp = blk.slots[s].e.eimpu
                    r.etot = blk.slots[s].e.etot
                    r.steps = blk.slots[s].steps
                    r.converged = bool(blk.slots[s].converged)
                    r.failed = ph == SLOT_FAILED
                    self.done.append(r)
                    if progress:
                        progress(r)
                    blk.req[s] = None
                    blk.ctx.set_active(s, False)
        if not self.queue:
            self.budget.not_waiting(self.n)
        # idle blocks stay (a later request of this class reuses them) unless
        # another class waits for budget: then their memory goes back
        if not self.fixed and not self.queue and self.budget.others_waiting(self.n):
            for blk in [b for b in self.blocks if b.idle()]:
                blk.ctx.close()
                self.budget.give(blk.nbytes)
                self.blocks.remove(blk)
        return busy

    def close(self):
        for blk in self.blocks:
            blk.ctx.close()
            self.budget.give(blk.nbytes)
        self.blocks = []


class Server:
    """Serve requests in real time on one GPU.

    budget_bytes: HBM the server may reserve for solver state; requests are
    admitted by the footprint of their grid class's blocks (block_slots per
    class, default BLOCK_SLOTS).  slots_per_class: instead, a fixed plan of one
    block per class with that many slots, created up front."""

    def __init__(self, slots_per_class=None, budget_bytes=None, block_slots=None, tol=1e-9, tol_fp32=1e-7,
                 max_steps=60000, pener=50, precision=PREC_MIXED, max_retries=2, graph_steps=24):
        kw = dict(tol=tol, n_rk=3, pener=pener, max_steps=max_steps, polish_steps=max_steps,
                  tol_fp32=tol_fp32, check_finite=0, graph_steps=graph_steps)
        if slots_per_class is not None:
            self.budget = Budget(float("inf"))
            self.lanes = {n: _Lane(n, k, kw, precision, self.budget, fixed=True, max_retries=max_retries)
                          for n, k in slots_per_class.items() if k > 0}
        else:
            if budget_bytes is None:
                raise ValueError("give budget_bytes (admission by footprint) or slots_per_class (fixed plan)")
            self.budget = Budget(budget_bytes)
            bs = dict(BLOCK_SLOTS)
            bs.update(block_slots or {})
            self.lanes = {n: _Lane(n, bs[n], kw, precision, self.budget, max_retries=max_retries)
                          for n in GRID_CLASSES}

    def close(self):
        for ln in self.lanes.values():
            ln.close()

    def stats(self):
        """Peak HBM reserved, the budget and the blocks created per class."""
        return {"peak_bytes": self.budget.peak, "budget_bytes": self.budget.total,
                "blocks_created": {n: ln.created for n, ln in self.lanes.items()}}

    def run(self, requests, progress=None):
        """Serve `requests` (with .arrival seconds from start) in real time.
        Returns them completed, with start/finish/latency fields set;
        `progress(request)` is called as each one completes."""
        pending = [(r.arrival, r.rid, r) for r in requests]
        heapq.heapify(pending)
        for r in requests:
            r.grid = grid_for(r.natoms) or 0
            if r.grid not in self.lanes:
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
                if not ln.admit_waiting(time.perf_counter() - t0):
                    for r in ln.queue:  # a block of this class exceeds the whole budget
                        r.failed = True
                        ln.done.append(r)
                    ln.queue = []
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
