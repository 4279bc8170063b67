"""Continuous-batching request loop for BUG evolutions on one GPU.

A request is a whole evolution: a chain (XX/XXZ, domain-wall or Neel start) or a Heston QTT
pricing problem, with its own length, maxdim, grid size, step count and precision.  The engine
keeps a fixed number of slots; each slot is a host thread with its own library context and HIP
stream, on which its request's state and environments stay resident from the first step to the
last.  Requests wait in arrival order and are admitted as soon as a slot is free AND their
estimated HBM footprint fits in what the running requests leave of the budget, so a request that
arrives while others are on the GPU starts without waiting for them to finish, and the steps of
all running requests overlap on the device (different streams).  Each request's arithmetic is
the same alone and in the mix: the kernels are deterministic, a request's launches go to its own
stream in program order, and nothing numerical is shared between slots.

Precision policy: a request's ``precision`` ("mixed" or "fp64") applies to chains (Hermitian
generators); Heston requests (non-Hermitian, stiff) always run FP64.  In FP32 their stiff
generators turn contraction rounding into spurious rank growth (a 2^7 x 2^7 grid: bond dimension
2 -> 14 and 3x the time over three steps) or overflow (2^10 x 2^10); the form actually used is
reported on the result.

Reported per request: queueing delay (arrival to admission), run time, and latency (arrival to
completion); per run: p50/p99 latency and the aggregate step rate.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import bug_step as B
from . import cstep as C
from . import mps as M
from . import qtt as Q
from .blocks import device
from .mpo import xxz_mpo


@dataclass
class Request:
    """One evolution.  ``kind`` "chain": ``L``, ``delta``, ``start`` ("dw" or "neel"),
    ``maxdim``, ``steps``, ``dt`` (tau = -i dt); ``kind`` "heston": ``bits`` per axis,
    ``params`` (r, q, kappa, theta, sigma, rho), ``maxdim``, ``steps``, ``dt`` (tau = dt, real,
    hermitian = False).  ``order``, ``precision`` and ``seed`` apply to both.  ``driver`` "c"
    steps through the library's step-level entry point ``bug_step`` (the step orchestrated in
    C++), "python" through the package's ``cbe_bug_step``; both run the same kernels."""

    kind: str
    params: Dict
    order: int = 2
    precision: str = "mixed"
    seed: Optional[int] = None
    keep_state: bool = False
    driver: str = "c"
    rid: int = -1
    arrival: float = 0.0


@dataclass
class Result:
    rid: int
    kind: str
    steps: int
    arrival: float
    admitted: float
    finished: float
    footprint: int
    statuses: List[str]
    observable: np.ndarray          # chain: S^z profile; heston: bond dimensions
    state: Optional[object] = None  # the final device state when ``keep_state``
    error: Optional[str] = None
    precision: str = ""             # the arithmetic form the request ran in

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def queueing(self) -> float:
        return self.admitted - self.arrival


def footprint_bytes(req: Request) -> int:
    """Estimated HBM high-water mark of a request: the state, one environment stack and the
    step's working set (frames up to twice the bond, their enlarged environments and the root
    Krylov basis) at the request's maxdim, complex FP64, with a fixed allowance for plans and
    scratch.  An estimate for admission, not a measurement."""
    p = req.params
    if req.kind == "chain":
        L, d, w = p["L"], 2, 5
    elif req.kind == "heston":
        L, d, w = 2 * p["bits"], 2, 9
    else:
        raise ValueError(f"unknown request kind {req.kind!r}")
    D = min(p["maxdim"], 2 ** (L // 2))
    state = L * d * D * D * 16
    envs = (L + 1) * w * D * D * 16
    work = 4 * w * (2 * D) ** 2 * 16 + 32 * (2 * D) ** 2 * 16
    return int(2 * (state + envs + work) + (64 << 20))


def _build(req: Request):
    p = req.params
    if req.kind == "chain":
        start = {"dw": M.domain_wall_state, "neel": M.neel_state}[p["start"]]
        psi = start(p["L"])
        mpo = xxz_mpo(p["L"], delta=p["delta"])
        tau, herm = -1j * p["dt"], True
    else:
        n = p["bits"]
        g = Q.qtt_grid(n, n, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)
        init = Q.qtt_separable(g, lambda x: math.exp(0.5 * x), lambda v: math.exp(-v))
        psi = Q.qtt_mps(init)
        mpo = Q.heston_mpo(g, Q.HestonParams(*p["params"]))
        tau, herm = complex(p["dt"]), False
    return psi, mpo, tau, herm


def precision_used(req: Request) -> str:
    """The form a request runs in: its own for chains, FP64 for Heston (see the module notes)."""
    if req.precision not in ("mixed", "fp64"):
        raise ValueError(f"unknown precision {req.precision!r}")
    return req.precision if req.kind == "chain" else "fp64"


class Cancelled(RuntimeError):
    """A request stopped between two steps because its engine was cancelled."""


def run_request(req: Request, stop: Optional[threading.Event] = None):
    """Run one request to completion on the calling thread's current stream; with ``stop`` set
    (``Engine.cancel``) it ends between two steps with ``Cancelled``."""
    psi, mpo, tau, herm = _build(req)
    form = precision_used(req)
    if req.driver not in ("c", "python"):
        raise ValueError(f"unknown driver {req.driver!r}")
    kw = dict(order=req.order, maxdim=req.params["maxdim"], hermitian=herm,
              work_ceiling_rule="maxdim", precision=form, rng=req.seed)
    statuses = []
    cstate = cmpo = None
    if req.driver == "c":
        cstate, cmpo = C.CState(psi), C.CMPO(mpo)
    try:
        for _ in range(req.params["steps"]):
            if stop is not None and stop.is_set():
                raise Cancelled(f"cancelled after {len(statuses)} of {req.params['steps']} steps")
            if cstate is not None:
                statuses.append(C.bug_step(cstate, cmpo, tau, **kw)["status"])
            else:
                info = B.cbe_bug_step(psi, mpo, tau, **kw)
                statuses.append(B.budget_status(info.ledger))
        if cstate is not None:
            obs = (cstate.expect_diagonal() if req.kind == "chain"
                   else np.array(cstate.bond_dims()))
            psi = cstate.to_mps() if req.keep_state else None
        else:
            obs = M.magnetisation(psi) if req.kind == "chain" else np.array(psi.bond_dims())
        torch.cuda.current_stream().synchronize()
    finally:
        if cstate is not None:
            cstate.close()
            cmpo.close()
    return psi, statuses, obs


class Engine:
    """``slots`` concurrent requests at most, within ``hbm_budget`` bytes of estimated
    footprint (default: 80 % of the device memory free when the engine starts)."""

    def __init__(self, slots: int = 4, hbm_budget: Optional[int] = None):
        if slots < 1:
            raise ValueError("an engine needs at least one slot")
        torch.cuda.init()
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(device())
            hbm_budget = int(0.8 * free)
        self.slots = slots
        self.budget = int(hbm_budget)
        self.in_use = 0
        self._cv = threading.Condition()
        self._pending: List[Request] = []
        self._work: "queue.Queue[Optional[Request]]" = queue.Queue()
        self._running = 0
        self._next = 0
        self._closed = False
        self._stop = threading.Event()
        self.results: Dict[int, Result] = {}
        self._threads = [threading.Thread(target=self._worker, name=f"bug-slot-{i}", daemon=True)
                         for i in range(slots)]
        self._sched = threading.Thread(target=self._scheduler, name="bug-admission", daemon=True)
        for t in self._threads:
            t.start()
        self._sched.start()

    # -- client side ----------------------------------------------------------------------------
    def submit(self, req: Request) -> int:
        fp = footprint_bytes(req)
        if fp > self.budget:
            raise ValueError(f"request needs ~{fp / 2**30:.1f} GiB, over the engine's budget of "
                             f"{self.budget / 2**30:.1f} GiB")
        with self._cv:
            if self._closed:
                raise RuntimeError("engine is closed")
            req.rid = self._next
            self._next += 1
            req.arrival = time.perf_counter()
            self._pending.append(req)
            self._cv.notify_all()
        return req.rid

    def wait(self, rids: Optional[Sequence[int]] = None, timeout: Optional[float] = None):
        """Block until the given requests (default: all submitted) have finished."""
        deadline = None if timeout is None else time.perf_counter() + timeout
        with self._cv:
            while True:
                want = range(self._next) if rids is None else rids
                if all(r in self.results for r in want):
                    return [self.results[r] for r in want]
                left = None if deadline is None else deadline - time.perf_counter()
                if left is not None and left <= 0:
                    raise TimeoutError("requests still running")
                self._cv.wait(left)

    def cancel(self):
        """Drop the requests still waiting for admission (they get no result) and make the running
        ones stop after their current step (their results carry a ``Cancelled`` error)."""
        with self._cv:
            self._pending.clear()
            self._stop.set()
            self._cv.notify_all()

    def close(self):
        """Refuse new requests, let the admitted and waiting ones finish, stop the slots."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
            while self._pending or self._running:
                self._cv.wait()
        for _ in self._threads:
            self._work.put(None)
        for t in self._threads:
            t.join()
        self._sched.join()

    # -- engine side ----------------------------------------------------------------------------
    def _scheduler(self):
        """Admit in arrival order while a slot is free and the head request fits the budget."""
        with self._cv:
            while True:
                while self._pending and self._running < self.slots:
                    head = self._pending[0]
                    fp = footprint_bytes(head)
                    if self.in_use + fp > self.budget:
                        break
                    self._pending.pop(0)
                    self.in_use += fp
                    self._running += 1
                    head._admitted = time.perf_counter()
                    head._footprint = fp
                    self._work.put(head)
                if self._closed and not self._pending:
                    return
                self._cv.wait()

    def _worker(self):
        stream = torch.cuda.Stream(device=device())
        while True:
            req = self._work.get()
            if req is None:
                return
            err, psi, statuses, obs = None, None, [], np.zeros(0)
            try:
                with torch.cuda.stream(stream):
                    psi, statuses, obs = run_request(req, self._stop)
            except Exception as e:          # reported on the result, the slot stays alive
                err = f"{type(e).__name__}: {e}"
            try:
                form = precision_used(req)
            except ValueError:
                form = req.precision
            res = Result(req.rid, req.kind, len(statuses), req.arrival, req._admitted,
                         time.perf_counter(), req._footprint, statuses, obs,
                         psi if req.keep_state else None, err, form)
            with self._cv:
                self.results[req.rid] = res
                self.in_use -= req._footprint
                self._running -= 1
                self._cv.notify_all()


def summarize(results: Sequence[Result], makespan: float) -> Dict:
    """p50/p99 latency, queueing and the aggregate step rate of a finished run."""
    lat = np.array([r.latency for r in results])
    q = np.array([r.queueing for r in results])
    steps = sum(r.steps for r in results)
    return {
        "requests": len(results),
        "errors": sum(r.error is not None for r in results),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
        "p50_queueing_s": round(float(np.percentile(q, 50)), 3),
        "steps": steps,
        "makespan_s": round(makespan, 3),
        "steps_per_s": round(steps / makespan, 3) if makespan > 0 else None,
    }


def poisson_mix(n: int, rate: float, seed: int = 0, scale: str = "small") -> List[tuple]:
    """(arrival offset in s, Request) pairs of the serving mix: 60 % chains, 40 % Heston.

    ``scale = "full"`` draws the full mix (L in {32, 64, 96, 128}, maxdim in {64, 256, 512, 1024},
    20-200 steps; Heston grids 2^8-2^11 per axis, maxdim 32-128, 10-50 steps); ``"small"`` the
    same shapes cut to what a few minutes on one GPU hold (L in {16, 24, 32}, maxdim in
    {16, 32, 64}, 2-6 steps; Heston grids 2^5-2^7, maxdim 16-32, 2-4 steps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.random() < 0.6:
            if scale == "full":
                L, maxdim, steps = (int(rng.choice([32, 64, 96, 128])),
                                    int(rng.choice([64, 256, 512, 1024])), int(rng.integers(20, 201)))
            else:
                L, maxdim, steps = (int(rng.choice([16, 24, 32])), int(rng.choice([16, 32, 64])),
                                    int(rng.integers(2, 7)))
            req = Request("chain", {"L": L, "delta": float(rng.choice([0.0, 0.5, 1.0])),
                                    "start": str(rng.choice(["dw", "neel"])), "maxdim": maxdim,
                                    "steps": steps, "dt": 0.05})
        else:
            params = (rng.uniform(0, 0.05), rng.uniform(0, 0.05), rng.uniform(0.5, 4),
                      rng.uniform(0.02, 0.15), rng.uniform(0.1, 0.8), rng.uniform(-0.9, 0))
            if scale == "full":
                bits, maxdim, steps = (int(rng.integers(8, 12)), int(rng.integers(32, 129)),
                                       int(rng.integers(10, 51)))
            else:
                bits, maxdim, steps = (int(rng.integers(5, 8)), int(rng.integers(16, 33)),
                                       int(rng.integers(2, 5)))
            req = Request("heston", {"bits": bits, "params": params, "maxdim": maxdim,
                                     "steps": steps, "dt": 0.02})
        out.append((t, req))
    return out


def serve_mix(mix: Sequence[tuple], slots: int, precision: str, timeout: float = 900.0) -> Dict:
    """Replay a timed mix (``poisson_mix``) against a fresh engine: each request is submitted at
    its arrival offset; returns ``summarize`` of the requests that completed.  When ``timeout``
    seconds pass first, the engine is cancelled (waiting requests dropped, running ones stopped
    after their current step) and ``cancelled`` counts the requests left out of the summary."""
    eng = Engine(slots=slots)
    t0 = time.perf_counter()
    deadline = t0 + timeout
    rids = []
    try:
        for offset, req in mix:
            req.precision = precision
            delay = t0 + offset - time.perf_counter()
            if delay > 0:
                if t0 + offset > deadline:
                    break
                time.sleep(delay)
            rids.append(eng.submit(req))
        try:
            eng.wait(rids, timeout=max(deadline - time.perf_counter(), 0.0))
        except TimeoutError:
            eng.cancel()
    finally:
        eng.close()
    done = [eng.results[r] for r in rids if r in eng.results and eng.results[r].error is None]
    failed = [eng.results[r] for r in rids
              if r in eng.results and eng.results[r].error is not None
              and not eng.results[r].error.startswith("Cancelled")]
    makespan = (max(r.finished for r in done) - t0) if done else 0.0
    out = summarize(done, makespan) if done else {"requests": 0, "steps": 0}
    out["errors"] = len(failed)
    out["cancelled"] = len(mix) - len(done) - len(failed)
    out["statuses_met"] = sum(s == "met" for r in done for s in r.statuses)
    return out
