"""``qtt_bug_step``: one rank-adaptive BUG step of a linear QTT flow on the device.

The linear-flow branch of ``qtt_bug_step!`` (src/QTTS/Tooling/nonlinear_step.jl:102-337,
``stage_solver = :krylov_rational``, the default for a ``LinearFlow``, flows.jl:20-44): the BUG
engine's step (``cbe_bug_step``) at tau = pref h on the flow's fixed operator A, with the root
solved as R(tau A_G) S_n -- the stability function of the order's implicit Runge-Kutta method
(trapezoid for order 2, Radau IIA with two and three stages for orders 3 and 4;
stage_solvers.jl:161-165) on the Arnoldi basis of the Galerkin operator, unshifted so that the
flow's decay is kept, with a basis of up to 200 vectors -- then the step's compression and the
frame audits (frames.jl:173-185).  The step's tolerances are applied in the units of its
O(h^(p+1)) local error (``qtt_step_tolerance``, nonlinear_step.jl:25-45).  ``root = "exp"`` runs
the exponential root on A - <A> and restores the scalar exp(tau <A>).

The nonlinear branch (the Newton-GMRES stage solver with densities rebuilt at every stage
iterate) is not built here: ``stage_solver = "newton_gmres"`` raises.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import bug_step as B
from . import krylov as K
from . import tensor_ops as T
from .mpo import MPO
from .mps import SymMPS

QTT_KRYLOV_MARGIN = 0.5       # share of tol_basis given to the local Krylov test
QTT_RATIONAL_MAXITER = 200    # Arnoldi basis cap of the one-stage rational root
QTT_TOL_FLOOR = 1e-14         # smallest per-step tolerance handed out


def qtt_step_tolerance(tol: float, h: float, order: int, scaling: str = "order",
                       h_ref: float = 0.1) -> float:
    """``tol`` times (|h| / h_ref)^(order + 1) (never looser than ``tol``), floored at 1e-14
    unless ``tol`` is already below; ``scaling = "fixed"`` returns ``tol``."""
    if scaling not in ("order", "fixed"):
        raise ValueError(f"tol_scaling must be 'order' or 'fixed', got {scaling!r}")
    if not h_ref > 0:
        raise ValueError("tol_h_ref must be positive")
    if scaling == "fixed" or tol == 0:
        return float(tol)
    factor = min(1.0, (abs(float(h)) / float(h_ref)) ** (order + 1))
    return max(float(tol) * factor, min(float(tol), QTT_TOL_FLOOR))


def qtt_root_for_order(order: int) -> str:
    """The root method of BUG order 2, 3, 4: trapezoid, Radau IIA(2), Radau IIA(3)."""
    roots = {2: "trapezoid", 3: "radau2", 4: "radau3"}
    if order not in roots:
        raise ValueError(f"nonlinear BUG orders are 2, 3, 4; got {order}")
    return roots[order]


class LinearFlow:
    """d psi / dt = pref A psi for a fixed operator ``mpo`` (e.g. the Heston generator)."""

    def __init__(self, mpo: MPO, pref: complex = 1.0):
        self.mpo = mpo
        self.pref = complex(pref)


def frame_defect(left: Sequence, right: Sequence) -> float:
    """Largest isometry defect ||Q^H Q - I|| (left frames) and ||Q Q^H - I|| (right frames),
    spectral norm per sector block, from the frames' Gram matrices on the device."""
    worst = 0.0
    for F, left_side in [(f, True) for f in left] + [(f, False) for f in right]:
        if F is None:
            continue
        G = T.overlap_left(F, F) if left_side else T.overlap_right(F, F)
        for (k, q), blk in G.to_blocks().items():
            d = np.linalg.norm(blk - np.eye(blk.shape[0]), 2) if blk.shape[0] == blk.shape[1] \
                else np.linalg.norm(blk, 2)
            worst = max(worst, float(d))
    return worst


def qtt_bug_step(psi: SymMPS, flow: LinearFlow, t: float, h: float, *, order: int = 2,
                 root: Optional[str] = None, tol_stage: float = 1e-10, tol_basis: float = 1e-6,
                 tol_trunc: float = 1e-8, maxdim: int = 128, keep: Sequence[SymMPS] = (),
                 stage_solver: str = "auto", tol_scaling: str = "order", tol_h_ref: float = 0.1,
                 tol_krylov: Optional[float] = None, **step_kw) -> dict:
    """One step of ``flow`` from t to t + h in place; returns the step's report (convergence of
    the root, its error and Krylov size, the discarded weight, the audits, the tolerances used).
    ``keep`` states are kept exactly in the frames (Heston mass); ``step_kw`` go to
    ``cbe_bug_step`` (e.g. ``precision``, ``parallel``, ``rng``)."""
    solver = "krylov_rational" if stage_solver == "auto" else stage_solver
    if solver == "newton_gmres":
        raise NotImplementedError("the Newton-GMRES stage solver (nonlinear flows) is not built")
    if solver != "krylov_rational":
        raise ValueError("stage_solver must be 'auto', 'newton_gmres' or 'krylov_rational'")
    if not isinstance(flow, LinearFlow):
        raise ValueError("stage_solver = 'krylov_rational' needs a LinearFlow")
    root = qtt_root_for_order(order) if root is None else root
    K.check_root_method(root)

    def step_tol(x):
        return qtt_step_tolerance(x, h, order, tol_scaling, tol_h_ref)

    tol_step = dict(basis=step_tol(tol_basis), trunc=step_tol(tol_trunc), stage=step_tol(tol_stage))
    tau = complex(flow.pref * h)
    eshift = 0.0
    if root == "exp":
        from .ite import energy
        eshift = energy(psi, flow.mpo)
    tk = step_tol(QTT_KRYLOV_MARGIN * tol_basis if tol_krylov is None else tol_krylov)
    frames: dict = {}
    info = B.cbe_bug_step(
        psi, flow.mpo, tau, order=order, tol_basis=tol_step["basis"],
        tol_trunc=tol_step["trunc"], maxdim=maxdim, hermitian=False, keep_states=list(keep),
        truncate=True, eshift=eshift, tol_krylov=tk, root_method=root, tol_root=tol_step["stage"],
        root_shift=False, root_maxiter=QTT_RATIONAL_MAXITER if root != "exp" else K.ROOT_MAXITER,
        frames_report=frames, **step_kw)
    if eshift != 0.0:                       # exp(tau (A - <A>)) psi times exp(tau <A>)
        T.scale_(complex(np.exp(tau * eshift)), psi[psi.center])
    lg = info.ledger
    return dict(converged=lg.expv_err <= tol_step["stage"], root_error=lg.expv_err,
                krylov=lg.expv_iters, discarded=lg.discarded, status=B.budget_status(lg),
                rank_capped=max(psi.bond_dims()) >= maxdim, miss_input=lg.proj_err,
                frame_defect=frame_defect(frames.get("left", []), frames.get("right", [])),
                tol_step=tol_step, root=root, solver=solver, chi=max(psi.bond_dims()),
                seconds=dict(info.seconds))
