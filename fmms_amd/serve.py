"""Serving loop: continuous batching of independent N-body systems on one GPU.

Requests (a new system plus a number of leapfrog steps) arrive over time.
The server admits them under an HBM budget and computes each newcomer's
initial forces on its own, on a second context and stream, in a worker
thread, while the batched steps of the active systems go on. From the first
step boundary after those forces are done it advances the newcomer
together with every other active system: all active systems form one
ensemble (fmmh_set_systems) whose state (positions, velocities, masses,
accelerations, potentials) stays packed in device memory, so one batched
kick-drift-kick step (fmmh_leapfrog_device) advances them all with one
launch per stage. Finished systems leave at the next boundary, their final
state copied out. Each system's trajectory is the one it would have alone:
the forest evaluates every system exactly as its standalone tree would.

The reference has no serving layer (crates/visual/src/sim.rs steps one
system per Sim); this is the batched form of its Sim::step loop
(sim.rs:539-594) for many systems.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import FMMH_E_OVER_BUDGET, Fmm, FmmError, hip, ic_disk, ic_hernquist, ic_plummer

# Device bytes per body of one evaluation context at p = 8 (bodies, keys,
# cells, expansions and lists), measured on the 1M-body Plummer workload and
# rounded up: the starting estimate for admission under the HBM budget.
BYTES_PER_BODY = 1600
# Packed state per body: pos, vel, acc (3 doubles each), mass, pot.
STATE_BYTES = 88
# Allowance for what the server's two contexts hold whatever the batch
# (size-independent buffers, growth slack of the arenas, the admission
# context's buffers for its largest newcomer so far).
OVERHEAD_BYTES = 256 << 20


def bytes_per_body(p):
    """Starting estimate of one context's device bytes per body at expansion
    order p: a fixed part (bodies, keys, cells, lists) and the expansions,
    which grow with (p + 1)^2 (BYTES_PER_BODY at p = 8)."""
    return 600.0 + 1000.0 * (p + 1) ** 2 / 81.0


@dataclass
class Request:
    """A system to evolve: initial state (host arrays) and a step count."""
    rid: int
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    steps: int
    arrival: float = 0.0  # seconds since the start of the run
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    steps_done: int = 0
    rejected: bool = False  # larger than the whole HBM budget: never admitted
    final: tuple = field(default=None, repr=False)

    @property
    def n(self):
        return self.mass.shape[0]


def make_request(rid, model, n, steps, seed, arrival=0.0):
    """A Plummer, Hernquist or exponential-disk system (total mass 1)."""
    if model == "plummer":
        s = ic_plummer(n, rmax=20.0, seed=seed)
    elif model == "hernquist":
        s = ic_hernquist(n, a=1.0, rmax=100.0, seed=seed)
    elif model == "disk":
        s = ic_disk(n, rd=1.0, zd=0.05, rmax=10.0, m_central=0.1, seed=seed)
    else:
        raise ValueError(model)
    return Request(rid, s.pos, s.vel, s.mass, steps, arrival)


class _Slot:
    """One admitted system: its device state lives either in its own arrays
    (`arrays`, while waiting to join) or in the packed ensemble at `offset`."""

    def __init__(self, req, arrays):
        self.req = req
        self.arrays = arrays  # dict name -> DeviceArray (pos, vel, mass, acc, pot)
        self.offset = None    # position in the packed arrays once joined


class Server:
    """Continuous-batching leapfrog server (one GPU, one ensemble context)."""

    FIELDS = (("pos", 3), ("vel", 3), ("mass", 1), ("acc", 3), ("pot", 1))

    def __init__(self, p=8, theta=0.5, softening=1e-3, n_leaf=32, dt=1e-4,
                 hbm_budget=200e9, precision="fp64", soft_mac=0.0):
        self.kw = dict(p=p, eps=math.inf, n_leaf=n_leaf, theta_max=theta, softening=softening,
                       precision=precision, soft_mac=soft_mac)
        self.dt = dt
        self.hbm_budget = float(hbm_budget)
        # each context also enforces the budget on its own allocations
        self.ctx = Fmm(None, None, hbm_budget_bytes=self.hbm_budget, **self.kw)
        self.admit_ctx = Fmm(None, None, hbm_budget_bytes=self.hbm_budget, **self.kw)
        # device bytes per body: the estimate for p, raised to what the
        # ensemble context actually holds once it has run
        self.bpb = bytes_per_body(p)
        self.active = []      # admitted systems, in packing order
        self.queue = []       # arrived, not yet admitted
        # admitted systems whose initial forces are being computed: (slot,
        # future), in admission order; one worker, as the admission context
        # takes one call at a time
        self.pending = []
        self._pool = ThreadPoolExecutor(max_workers=1)
        self.packed = None    # dict name -> DeviceArray over all active systems
        self.dirty = True     # membership changed since the last packing
        self.step_ms = []     # wall time of each batched step
        self.step_bodies = []

    # ---------------------------------------------------------- admission
    def _bytes(self, n):
        return (self.bpb + STATE_BYTES) * n

    def _in_use(self):
        return OVERHEAD_BYTES + sum(self._bytes(s.req.n) for s in self.active) + \
            sum(self._bytes(s.req.n) for s, _ in self.pending)

    def _fits_alone(self, n):
        return OVERHEAD_BYTES + self._bytes(n) <= self.hbm_budget

    def submit(self, req):
        """Queue a request; one that could not fit even alone is rejected
        (FmmError with code FMMH_E_OVER_BUDGET) instead of waiting forever."""
        if not self._fits_alone(req.n):
            req.rejected = True
            raise FmmError(FMMH_E_OVER_BUDGET,
                           f"request {req.rid}: {req.n} bodies need about "
                           f"{self._bytes(req.n):.3g} bytes, over the {self.hbm_budget:.3g}-byte budget")
        self.queue.append(req)

    def _admit(self, now):
        kept = []
        for req in self.queue:
            if not self._fits_alone(req.n):  # the measured footprint outgrew the budget
                req.rejected = True
                continue
            if self._in_use() + self._bytes(req.n) > self.hbm_budget:
                kept.append(req)
                continue
            n = req.n
            arr = {"pos": hip.DeviceArray.from_host(np.ascontiguousarray(req.pos)),
                   "vel": hip.DeviceArray.from_host(np.ascontiguousarray(req.vel)),
                   "mass": hip.DeviceArray.from_host(np.ascontiguousarray(req.mass)),
                   "acc": hip.DeviceArray((n, 3)), "pot": hip.DeviceArray((n,))}
            # a(x0) on its own (the same forces it gets inside an ensemble),
            # overlapped with the batched steps; it joins once they are done
            req.admitted = now
            self.pending.append((_Slot(req, arr), self._pool.submit(self._initial_forces, n, arr)))
        self.queue = kept

    def _initial_forces(self, n, arr):
        self.admit_ctx.set_systems([n])
        self.admit_ctx.accel_device(n, arr["pos"].ptr, arr["mass"].ptr, arr["acc"].ptr,
                                    arr["pot"].ptr)

    def _join(self):
        """Moves admitted systems whose initial forces are done (in admission
        order) into the active set; with nothing active, waits for the first."""
        while self.pending and (self.pending[0][1].done() or not self.active):
            slot, fut = self.pending.pop(0)
            fut.result()  # re-raises a failure of the admission call
            self.active.append(slot)
            self.dirty = True

    def close(self):
        """Waits for admissions in flight and stops the worker thread."""
        self._pool.shutdown(wait=True)

    # ------------------------------------------------------------ packing
    def _pack(self):
        """New packed arrays for the current active set: joined systems move
        from the old packed arrays, newcomers from their own arrays."""
        total = sum(s.req.n for s in self.active)
        new = {name: hip.DeviceArray((total, w) if w > 1 else (total,))
               for name, w in self.FIELDS}
        off = 0
        for s in self.active:
            for name, w in self.FIELDS:
                if s.offset is None:
                    src = s.arrays[name].ptr
                else:
                    src = self.packed[name].ptr + s.offset * w * 8
                hip.copy_d2d(new[name].ptr + off * w * 8, src, s.req.n * w * 8)
            off += s.req.n
        hip.synchronize_copies()
        off = 0
        for s in self.active:
            if s.offset is None:
                for a in s.arrays.values():
                    a.free()
                s.arrays = None
            s.offset = off
            off += s.req.n
        if self.packed is not None:
            for a in self.packed.values():
                a.free()
        self.packed = new
        self.ctx.set_systems([s.req.n for s in self.active])
        self.dirty = False

    def _retire(self, slot):
        """Final state of a finished system (host arrays) from the packed set."""
        out = []
        for name, w in self.FIELDS:
            if name == "mass":
                continue
            shape = (slot.req.n, w) if w > 1 else (slot.req.n,)
            a = hip.DeviceArray(shape)
            hip.copy_d2d(a.ptr, self.packed[name].ptr + slot.offset * w * 8, slot.req.n * w * 8)
            out.append(a)
        hip.synchronize_copies()
        res = tuple(a.download() for a in out)
        for a in out:
            a.free()
        return res

    # --------------------------------------------------------------- step
    def step(self, now):
        """One step boundary: admit arrivals (their initial forces start on
        the admission context), join the admitted systems whose initial
        forces are done (packing), one batched KDK step of every active
        system, retire finished systems. Returns the number of bodies
        stepped."""
        t0 = time.perf_counter()
        self._admit(now)
        self._join()
        if not self.active:
            return 0
        if self.dirty:
            self._pack()
        total = sum(s.req.n for s in self.active)
        p = self.packed
        self.ctx.leapfrog_device(total, p["pos"].ptr, p["vel"].ptr, p["mass"].ptr, p["acc"].ptr,
                                 p["pot"].ptr, self.dt, 1, refresh_every=1, acc_valid=True)
        if total >= 100_000:  # size-independent buffers no longer dominate the arena
            self.bpb = max(self.bpb, self.ctx.device_bytes() / total)
        done = []
        for s in self.active:
            s.req.steps_done += 1
            if s.req.steps_done >= s.req.steps:
                done.append(s)
        for s in done:
            s.req.final = self._retire(s)
            s.req.finished = now
        if done:
            self.active = [s for s in self.active if s not in done]
            self.dirty = True
        self.step_ms.append(1e3 * (time.perf_counter() - t0))
        self.step_bodies.append(total)
        return total

    def run(self, requests, max_seconds=None):
        """Serve requests (sorted by arrival, seconds since start) until all
        are finished; arrivals are admitted at the first step boundary after
        their arrival time."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t_start = time.perf_counter()
        i = 0
        while i < len(pending) or self.queue or self.active or self.pending:
            now = time.perf_counter() - t_start
            if max_seconds is not None and now > max_seconds:
                break
            while i < len(pending) and pending[i].arrival <= now:
                try:
                    self.submit(pending[i])
                except FmmError:
                    pass  # marked rejected; it never runs
                i += 1
            if not self.queue and not self.active and not self.pending:
                time.sleep(max(0.0, min(pending[i].arrival - now, 0.01)))
                continue
            self.step(now)
        return time.perf_counter() - t_start

    def stats(self):
        ms = np.array(self.step_ms)
        bodies = np.array(self.step_bodies, dtype=float)
        if ms.size == 0:
            return {}
        return {"steps": int(ms.size), "p50_ms": float(np.percentile(ms, 50)),
                "p99_ms": float(np.percentile(ms, 99)),
                "body_forces_per_s": float(bodies.sum() / (ms.sum() * 1e-3)),
                "mean_batch_bodies": float(bodies.mean())}


def run_alone(req, p=8, theta=0.5, softening=1e-3, n_leaf=32, dt=1e-4, soft_mac=0.0,
              precision="fp64"):
    """The same system stepped on its own (the reference behaviour: one Sim)."""
    from . import Integrator

    f = Fmm(None, None, p=p, eps=math.inf, n_leaf=n_leaf, theta_max=theta, softening=softening,
            soft_mac=soft_mac, precision=precision)
    it = Integrator(f, req.pos, req.vel, req.mass)
    it.step(dt, steps=0)
    t0 = time.perf_counter()
    for _ in range(req.steps):
        it.step(dt, steps=1, refresh_every=1)
    el = time.perf_counter() - t0
    return it.state(), el


__all__ = ["OVERHEAD_BYTES", "Request", "STATE_BYTES", "Server", "bytes_per_body", "make_request", "run_alone", "FmmError"]
