"""Normalised imaginary-time evolution on the device: ``ite_step`` / ``ite_evolve`` and the step
schedules.

One step is ``cbe_bug_step`` at tau = -dbeta on H - eshift (eshift the state's energy by default),
with the root report and its outside-Galerkin defect |tau| ||(I - P) H psi_1|| / ||psi_1||, then
renormalisation through the centre tensor.  A step whose ledger or defect fails is redone from
the saved state: with the measured closure loop when the closure did not close, then, while
retries remain and the failure is not a resource cap, as sub-steps whose number follows from
the defect's O(dbeta^(p+1)) scaling.  Every contraction runs in the library; the control flow
and the energies' final sums are host code.

Reference (madhavmen/BondUpdateGalerkin):
  - src/BUG/ITE/ite_step.jl:1-301     _ite_energy, _ite_normalise!, _ite_nsplit, _ite_capped,
                                      _ite_attempt!, ite_step!, _ite_result, ite_evolve!
  - src/BUG/ITE/schedules.jl:17-51    log_panel_steps, uniform_beta_steps, beta_grid,
                                      uniform_schedule, geometric_schedule
"""
from __future__ import annotations

import math
import time
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import environments as ev
from . import tensor_ops as T
from .bug_step import cbe_bug_step
from .mpo import MPO
from .mps import SymMPS
from .probes import ProbeStream

ITE_MAXSPLIT = 64
ROOT_METHODS = ("exp",)


def overlap(psi: SymMPS) -> float:
    """<psi|psi> by the left transfer chain (any gauge)."""
    C = None
    for A in psi.tensors:
        C = T.overlap_left(A, A if C is None else T.core_times_site(C, A))
    return float(T.norm(C))


def mpo_energy(psi: SymMPS, mpo: MPO) -> float:
    """<psi|H|psi> in any gauge: the one-site effective operator at the centre between the
    environments of the state's own tensors."""
    L = len(psi)
    c = psi.center if 0 <= psi.center < L else 0
    E = ev.boundary_left(mpo, psi.bond_leg(0))
    for i in range(c):
        E = ev.push_left(E, psi[i], mpo[i])
    R = ev.boundary_right(mpo, psi.bond_leg(L))
    for i in range(L - 1, c, -1):
        R = ev.push_right(R, psi[i], mpo[i])
    A = psi[c]
    return float(T.inner(A, ev.apply_one_site(E, mpo[c], R, A)).real)


def energy(psi: SymMPS, mpo: MPO) -> float:
    """``_ite_energy``: <psi|H|psi> / <psi|psi>."""
    return mpo_energy(psi, mpo) / max(overlap(psi), np.finfo(float).eps)


def normalise(psi: SymMPS) -> float:
    """``_ite_normalise!`` through the centre tensor; returns the old norm."""
    n = T.norm(psi[psi.center])
    if not (math.isfinite(n) and n > 0):
        raise ArithmeticError(f"imaginary-time step produced a state of norm {n}")
    psi[psi.center] = T.scale(1.0 / n, psi[psi.center])
    return n


def nsplit(attempt: dict, tol_basis: float, order: int) -> int:
    """``_ite_nsplit``: (resid / tol_basis)^(1/(p+1)) with a 20% margin for an outside-defect
    failure, clamped to [2, 64]; 2 for any other reason."""
    r = attempt["resid"]
    if not ("outside" in attempt["unmet"] and math.isfinite(r) and r > tol_basis):
        return 2
    return min(max(math.ceil(1.2 * (r / tol_basis) ** (1.0 / (order + 1))), 2), ITE_MAXSPLIT)


def capped(attempt: dict) -> bool:
    """``_ite_capped``: failed at the rank cap or the work ceiling, which a smaller step cannot
    widen."""
    lg = attempt["info"].ledger
    return bool(lg.maxdim_reached or lg.workcap_reached)


def _nanmax(xs):
    xs = list(xs)
    return float("nan") if any(math.isnan(x) for x in xs) else max(xs)


def attempt_step(psi: SymMPS, mpo: MPO, dbeta: float, *, order: int, tol_basis: float,
                 tol_trunc: float, tol_root: float, maxdim: int, eshift, outside_residual: bool,
                 closure_checks: bool, parallel: bool, rng, step_kw: dict) -> dict:
    """``_ite_attempt!``: energy, the step at tau = -dbeta with the shifted generators and the
    root report, renormalisation, the energy after, the verdicts."""
    t0 = time.perf_counter()
    e_before = energy(psi, mpo)
    es = e_before if eshift == "energy" else float(eshift)
    rep: dict = {}
    info = cbe_bug_step(psi, mpo, complex(-dbeta), order=order, tol_basis=tol_basis,
                        tol_trunc=tol_trunc, tol_expv=tol_root, maxdim=maxdim, eshift=es,
                        root_report=rep, outside_residual=outside_residual,
                        closure_checks=closure_checks, parallel=parallel, rng=rng, **step_kw)
    growth = normalise(psi)
    e_after = energy(psi, mpo)
    unmet = list(info.ledger.unmet)
    # `not resid <= tol`: a NaN defect is a failed measurement, not a pass
    if outside_residual and not (rep["resid"] <= tol_basis):
        unmet.append("outside")
    return dict(dbeta=dbeta, ok=not unmet, unmet=unmet, eshift=es, e_before=e_before,
                e_pre_root=rep["e_pre"], e_root=rep["e_root"], e_after=e_after,
                de_root=rep["e_root"] - rep["e_pre"], de_compression=e_after - rep["e_root"],
                outside=rep["outside"], resid=rep["resid"], resid_initial=rep["resid_initial"],
                outside_passes=rep["outside_passes"], outside_added=rep["outside_added"],
                root_err=info.ledger.expv_err, root_iters=info.ledger.expv_iters,
                discarded=info.ledger.discarded, growth=growth,
                chi=max(psi.bond_dims() + [0]), r_aug=info.ledger.r_aug,
                closure_checks=closure_checks, wall=time.perf_counter() - t0, info=info)


def ite_result(status: str, pieces: List[dict], attempts: List[dict], t0: float) -> dict:
    """``_ite_result``: the accepted pieces folded into one record (first energy before, last
    after, root and compression changes summed, worst defects)."""
    unmet: List[str] = []
    for p in pieces:
        for u in p["unmet"]:
            if u not in unmet:
                unmet.append(u)
    return dict(status=status, unmet=unmet, dbeta=sum(p["dbeta"] for p in pieces),
                e_before=pieces[0]["e_before"], e_pre_root=pieces[0]["e_pre_root"],
                e_root=pieces[-1]["e_root"], e_after=pieces[-1]["e_after"],
                de_root=sum(p["de_root"] for p in pieces),
                de_compression=sum(p["de_compression"] for p in pieces),
                resid=_nanmax(p["resid"] for p in pieces),
                outside=_nanmax(p["outside"] for p in pieces),
                root_err=max(p["root_err"] for p in pieces),
                discarded=max(p["discarded"] for p in pieces), chi=pieces[-1]["chi"],
                nsub=sum(p.get("nsub", 1) for p in pieces), attempts=attempts,
                wall=time.perf_counter() - t0)


def _snapshot(psi: SymMPS) -> SymMPS:
    return SymMPS([t.clone() for t in psi.tensors], psi.charges, psi.center)


def ite_step(psi: SymMPS, mpo: MPO, dbeta: float, *, order: int = 2, root_method: str = "exp",
             tol_basis: float = 1e-5, tol_trunc: float = 1e-6, tol_root: float = 1e-6,
             maxdim: int = 128, eshift="energy", outside_residual: bool = True, retries: int = 2,
             on_fail: str = "flag", parallel: bool = False, rng: Optional[ProbeStream] = None,
             **step_kw) -> dict:
    """``ite_step!``: one normalised step over dbeta, in place, accepted when the ledger is met
    and the outside defect is within tol_basis; otherwise redone with the measured closure (for
    an unclosed closure), then, while retries > 0 and the failure is not a resource cap, as
    ``nsplit`` sub-steps.  A step still failing has status "failed" (or raises with on_fail =
    "error").  Only the exponential root (``root_method = "exp"``) is built here."""
    if not dbeta > 0:
        raise ValueError(f"dbeta must be positive, got {dbeta}")
    if root_method not in ROOT_METHODS:
        raise ValueError(f"root_method must be one of {ROOT_METHODS}, got {root_method!r}")
    if on_fail not in ("flag", "error"):
        raise ValueError("on_fail must be 'flag' or 'error'")
    if retries < 0:
        raise ValueError(f"retries must be >= 0, got {retries}")
    t0 = time.perf_counter()
    rng = ProbeStream(0x5EED) if rng is None else rng
    normalise(psi)
    saved = _snapshot(psi)
    akw = dict(order=order, tol_basis=tol_basis, tol_trunc=tol_trunc, tol_root=tol_root,
               maxdim=maxdim, eshift=eshift, outside_residual=outside_residual,
               parallel=parallel, rng=rng)

    def restore():
        snap = _snapshot(saved)
        for i in range(len(psi)):
            psi[i] = snap[i]
        psi.center = snap.center

    attempts = []
    a = attempt_step(psi, mpo, float(dbeta), closure_checks=False, step_kw=step_kw, **akw)
    attempts.append(a)
    if not a["ok"] and order > 1 and "closure" in a["unmet"]:
        restore()
        a = attempt_step(psi, mpo, float(dbeta), closure_checks=True, step_kw=step_kw, **akw)
        attempts.append(a)
    if a["ok"]:
        return ite_result("met" if len(attempts) == 1 else "retried", [a], attempts, t0)
    if retries > 0 and not capped(a):
        restore()
        n = nsplit(a, tol_basis, order)
        halves = []
        for _ in range(n):
            sub = ite_step(psi, mpo, dbeta / n, root_method=root_method, retries=retries - 1,
                           on_fail="flag", **akw, **step_kw)
            attempts.extend(sub["attempts"])
            halves.append(sub)
        ok = all(h["status"] != "failed" for h in halves)
        res = ite_result("retried" if ok else "failed", halves, attempts, t0)
        if not ok and on_fail == "error":
            raise ArithmeticError(f"imaginary-time step dbeta={dbeta} failed after retries: "
                                  f"unmet {res['unmet']}")
        return res
    if on_fail == "error":
        raise ArithmeticError(f"imaginary-time step dbeta={dbeta} failed: unmet {a['unmet']}")
    return ite_result("failed", [a], attempts, t0)


def ite_evolve(psi: SymMPS, mpo: MPO, steps: Sequence[float], observe: Optional[Callable] = None,
               **kw) -> List[dict]:
    """``ite_evolve!``: one ``ite_step`` per entry of ``steps``; each record carries the beta
    reached and, with ``observe``, obs = observe(psi).  A failed step is recorded and the
    evolution continues (on_fail = "error" stops it)."""
    log, beta = [], 0.0
    for db in steps:
        rec = ite_step(psi, mpo, db, **kw)
        beta += db
        rec = dict(beta=beta, **rec)
        if observe is not None:
            rec["obs"] = observe(psi)
        log.append(rec)
    return log


def log_panel_steps(beta_max: float, panel_count: int, per_panel: int = 4) -> List[float]:
    """``log_panel_steps``: panel_count dyadic panels with endpoints 0, beta_max/2^(P-1), ...,
    beta_max/2, beta_max, each split into per_panel uniform steps."""
    if not beta_max > 0:
        raise ValueError(f"beta_max must be positive, got {beta_max}")
    if panel_count < 1:
        raise ValueError(f"log_panel_steps needs panel_count >= 1, got {panel_count}")
    if per_panel < 1:
        raise ValueError(f"per_panel must be >= 1, got {per_panel}")
    edges = [0.0] + [float(beta_max) / 2.0 ** (panel_count - p) for p in range(1, panel_count + 1)]
    steps: List[float] = []
    for p in range(panel_count):
        steps += [(edges[p + 1] - edges[p]) / per_panel] * per_panel
    return steps


def uniform_beta_steps(beta_max: float, step_count: int) -> List[float]:
    """``uniform_beta_steps``: step_count equal steps reaching beta_max."""
    if not beta_max > 0:
        raise ValueError(f"beta_max must be positive, got {beta_max}")
    if step_count < 1:
        raise ValueError(f"step_count must be >= 1, got {step_count}")
    return [float(beta_max) / step_count] * step_count


def beta_grid(steps: Sequence[float]) -> List[float]:
    """``beta_grid``: 0 and the cumulative sums of a step list."""
    return [0.0] + [float(x) for x in np.cumsum(np.asarray(steps, dtype=np.float64))]


def uniform_schedule(step_size: float, step_count: int) -> List[float]:
    """``uniform_schedule``: step_count steps of step_size."""
    return [float(step_size)] * step_count


def geometric_schedule(first_step: float, ratio: float, hmax: float,
                       step_count: int) -> List[float]:
    """``geometric_schedule``: dtau_k = min(hmax, first_step ratio^k), k = 0 .. step_count-1."""
    return [min(float(hmax), first_step * ratio ** k) for k in range(step_count)]
