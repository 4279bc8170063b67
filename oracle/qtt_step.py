"""ORACLE (test infrastructure): ``qtt_bug_step!`` for a linear flow, on the oracle's step.

* ``qtt_step_tolerance``  -- src/QTTS/Tooling/nonlinear_step.jl:25-45 (tolerances in the units of
                             the step's O(h^(p+1)) local error, floored at 1e-14)
* ``qtt_root_for_order``  -- src/BUG/TimeEvolution/stage_solvers.jl:161-165
* ``qtt_bug_step``        -- nonlinear_step.jl:102-337 on its ``stage_solver = :krylov_rational``
                             branch (the default for a ``LinearFlow``, flows.jl:20-44): the BUG
                             engine's step at tau = pref h with the rational root ``R(tau A_G)``
                             of the order's implicit Runge-Kutta method, unshifted so that the
                             flow's decay is kept, Arnoldi cap 200, then the frame audits
* ``frame_defect``        -- src/QTTS/Tooling/frames.jl:180-185

Pinned by tests/test_oracle_qtt_step.py: the stability functions against the Radau IIA and
trapezoid Butcher tableaux (R(z) = 1 + z b^T (I - z A)^{-1} 1) and against exp(z) to each
method's order; a step whose frames span the whole space against the dense R(hA) psi; the
tolerance scaling on hand values.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import blocks as ob
from . import bug as obug
from . import krylov as ok
from . import ite as oite
from . import linalg as la

QTT_KRYLOV_MARGIN = 0.5       # _QTT_KRYLOV_MARGIN, nonlinear_step.jl:8
QTT_RATIONAL_MAXITER = 200    # _QTT_RATIONAL_MAXITER, :17
QTT_TOL_FLOOR = 1e-14         # _QTT_TOL_FLOOR, :23


def qtt_step_tolerance(tol: float, h: float, order: int, scaling: str = "order",
                       h_ref: float = 0.1) -> float:
    if scaling not in ("order", "fixed"):
        raise ValueError(f"tol_scaling must be order or fixed, got {scaling!r}")
    if not h_ref > 0:
        raise ValueError("tol_h_ref must be positive")
    if scaling == "fixed" or tol == 0:
        return float(tol)
    factor = min(1.0, (abs(float(h)) / float(h_ref)) ** (order + 1))
    return max(float(tol) * factor, min(float(tol), QTT_TOL_FLOOR))


def qtt_root_for_order(order: int) -> str:
    roots = {2: "trapezoid", 3: "radau2", 4: "radau3"}
    if order not in roots:
        raise ValueError(f"nonlinear BUG orders are 2, 3, 4; got {order}")
    return roots[order]


class LinearFlow:
    """F(t, psi) = pref A psi for a fixed operator A (``mpo``: the oracle's MPO sites)."""

    def __init__(self, mpo, pref: complex = 1.0):
        self.mpo = mpo
        self.pref = complex(pref)


def frame_defect(left: Sequence[ob.Fam], right: Sequence[ob.Fam]) -> float:
    """max ||Q^H Q - I|| (spectral) over the left frames and ||Q Q^H - I|| over the right."""
    worst = 0.0
    for F in left:
        if F is not None:
            worst = max(worst, _gram_defect(la.overlap_left(F, F)))
    for F in right:
        if F is not None:
            worst = max(worst, _gram_defect(la.overlap_right(F, F)))
    return worst


def _gram_defect(G: ob.Fam) -> float:
    worst = 0.0
    for (k, q), blk in G.blocks.items():
        d = np.linalg.norm(blk - np.eye(blk.shape[0]), 2) if blk.shape[0] == blk.shape[1] \
            else np.linalg.norm(blk, 2)
        worst = max(worst, float(d))
    return worst


def qtt_bug_step(psi: obug.MPS, flow: LinearFlow, t: float, h: float, *, order: int = 2,
                 root: Optional[str] = None, tol_stage: float = 1e-10, tol_basis: float = 1e-6,
                 tol_trunc: float = 1e-8, maxdim: int = 128, keep: Sequence = (),
                 stage_solver: str = "auto", tol_scaling: str = "order", tol_h_ref: float = 0.1,
                 tol_krylov: Optional[float] = None, seed: int = 0x5EED, **step_kw) -> dict:
    """One step of the linear flow from t to t + h in place; returns the step's report."""
    solver = "krylov_rational" if stage_solver == "auto" else stage_solver
    if solver == "newton_gmres":
        raise NotImplementedError("the Newton-GMRES stage solver (nonlinear flows) is not restated")
    if solver != "krylov_rational":
        raise ValueError("stage_solver must be auto, newton_gmres or krylov_rational")
    if not isinstance(flow, LinearFlow):
        raise ValueError("stage_solver = krylov_rational needs a LinearFlow")
    root = qtt_root_for_order(order) if root is None else root
    ok.check_root_method(root)

    def step_tol(x):
        return qtt_step_tolerance(x, h, order, tol_scaling, tol_h_ref)

    tol_step = dict(basis=step_tol(tol_basis), trunc=step_tol(tol_trunc), stage=step_tol(tol_stage))
    tau = complex(flow.pref * h)
    # the exponential root runs on H - <H> (its scale restored below); a rational root unshifted
    eshift = oite.energy(psi, flow.mpo) if root == "exp" else 0.0
    tk = step_tol(QTT_KRYLOV_MARGIN * tol_basis if tol_krylov is None else tol_krylov)
    opts = obug.StepOptions(
        order=order, tol_basis=tol_step["basis"], tol_trunc=tol_step["trunc"], maxdim=maxdim,
        hermitian=False, keep_states=list(keep), truncate=True, eshift=eshift, tol_krylov=tk,
        root_method=root, tol_expv=tol_step["stage"], root_shift=False,
        root_maxiter=QTT_RATIONAL_MAXITER if root != "exp" else ok.ROOT_MAXITER, seed=seed,
        **step_kw)
    frames = {}
    info = obug.cbe_bug_step(psi, flow.mpo, tau, opts, frames_report=frames)
    if eshift != 0.0:                  # exp(tau (A - <A>)) psi times exp(tau <A>)
        c = psi.center
        psi.sites[c] = psi.sites[c].scale(np.exp(tau * eshift))
    lg = info.ledger
    return dict(converged=lg.expv_err <= tol_step["stage"], root_error=lg.expv_err,
                krylov=lg.expv_iters, discarded=lg.discarded, status=lg.status,
                rank_capped=max(psi.bond_dims()) >= maxdim, miss_input=lg.proj_err,
                frame_defect=frame_defect(frames.get("left", []), frames.get("right", [])),
                tol_step=tol_step, root=root, solver=solver, chi=max(psi.bond_dims()))
