"""ORACLE (test infrastructure): the holdout-certified adaptive DD controller.

Restates, for the U(1) / no-symmetry families of ``oracle.blocks``:

* ``pivoted_qr``     -- column-pivoted Householder QR (LAPACK ``geqp3``'s greedy rule: the next
                        pivot is the column of largest remaining norm, the first of equals), norms
                        recomputed at every step, R diagonal made real non-negative;
* ``thin_qr``        -- ``_thin_qr`` (src/BUG/Randomised/dd_cbe.jl:184-259): per charge sector of
                        the new leg, the pivoted QR of the stacked blocks, the columns whose pivot
                        exceeds ``thr`` kept (right side: of the conjugate transpose);
* ``dd_select``      -- ``_dd_select`` (dd_cbe.jl:287-321): output DD projection, then the
                        two-pass thin QR ("twice is enough", ``_REORTH_KEEP = 1/sqrt 2``, :273);
* ``dd_probe``       -- ``_dd_probe`` with a sector ``boost`` (``_boost!``, cbe.jl:206-216);
* ``sample_residual_full`` / ``target_sectors`` -- ``_sample_residual`` / ``_target_sectors!``
                        (src/BUG/Randomised/sketch.jl:141-201);
* ``dd_controller``  -- ``_dd_controller`` (src/BUG/Randomised/dd_cbe_controller.jl:26-212) with
                        ``batch_blocks = 1``, ``cap_selection = :storage``, ``audit = false``;
* ``initial_width``  -- the controller's first width (dd_cbe_expand.jl:89-93).

Pinned: ``pivoted_qr`` chooses scipy's (LAPACK geqp3) pivots and |R| diagonal and spans the same
columns (tests/test_oracle_adaptive.py); the controller's step replays the reference's frozen
runs, which were made with this mode (test_fixtures_b155550.jl:35-38), at their tolerances.
"""
from __future__ import annotations

import math
from typing import Dict, Optional, Tuple

import numpy as np

from . import linalg as la
from .blocks import Fam

REORTH_KEEP = 1.0 / math.sqrt(2.0)
DD_HOLDOUT = 4
ADAPTIVE_GROWTH = 1.2


def pivoted_qr(A: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """A[:, perm] = Q R with greedy column pivoting; returns (Q (m x k), diag R (k, real >= 0),
    perm), k = min(m, n)."""
    A = np.array(A, dtype=np.complex128)
    m, n = A.shape
    k = min(m, n)
    perm = np.arange(n)
    vs, taus, betas = [], [], []
    for j in range(k):
        norms = np.sum(np.abs(A[j:, j:]) ** 2, axis=0)
        p = j + int(np.argmax(norms))
        if p != j:
            A[:, [j, p]] = A[:, [p, j]]
            perm[[j, p]] = perm[[p, j]]
        x = A[j:, j].copy()
        alpha = x[0]
        x2 = float(np.sum(np.abs(x[1:]) ** 2))
        if x2 == 0.0 and alpha.imag == 0.0:
            tau, beta, v = 0.0, alpha.real, np.zeros(m - j, complex)
            v[0] = 1.0
        else:
            nrm = math.sqrt(abs(alpha) ** 2 + x2)
            beta = -nrm if alpha.real >= 0 else nrm
            tau = complex((beta - alpha.real) / beta, -alpha.imag / beta)
            v = x / (alpha - beta)
            v[0] = 1.0
        A[j:, j] = 0
        A[j, j] = beta
        if tau != 0:
            # A <- (I - conj(tau) v v^H) A on the trailing columns
            w = v.conj() @ A[j:, j + 1:]
            A[j:, j + 1:] -= np.conj(tau) * np.outer(v, w)
        vs.append(v)
        taus.append(tau)
        betas.append(beta)
    Q = np.zeros((m, k), dtype=np.complex128)
    Q[:k, :k] = np.eye(k)
    for j in range(k - 1, -1, -1):
        if taus[j] == 0:
            continue
        v = vs[j]
        w = v.conj() @ Q[j:, j:]
        Q[j:, j:] -= taus[j] * np.outer(v, w)
    d = np.array(betas, dtype=float)
    sign = np.where(d < 0, -1.0, 1.0)
    return Q * sign[None, :], np.abs(d), perm


def _width_sectors(T: Fam, side: str) -> Dict[int, int]:
    return dict(T.col) if side == "left" else dict(T.row)


def thin_qr(sketch: Fam, side: str, thr: float) -> Optional[Fam]:
    """Orthonormal basis of the span of ``sketch``'s new-leg columns (left: columns of the left
    matricisation; right: rows of the right one), sector by sector, keeping the columns whose
    pivot exceeds ``thr``."""
    if side == "left":
        secs = la.left_sectors(sketch)
        out_blocks, width = {}, {}
        for qc, parts in secs.items():
            M = la.left_matrix(sketch, qc, parts)
            if M.shape[0] == 0:
                continue
            Q, r, _ = pivoted_qr(M)
            kept = int(np.sum(r > thr))
            if kept == 0:
                continue
            width[qc] = kept
            r0 = 0
            for s, ql, d in parts:
                out_blocks[(s, ql)] = Q[r0:r0 + d, :kept].copy()
                r0 += d
        if not width:
            return None
        return Fam(sketch.row, width, sketch.shifts, out_blocks)
    secs = la.right_sectors(sketch)
    out_blocks, width = {}, {}
    for ql, parts in secs.items():
        M = la.right_matrix(sketch, ql, parts)
        if M.shape[1] == 0:
            continue
        Q, r, _ = pivoted_qr(M.conj().T)
        kept = int(np.sum(r > thr))
        if kept == 0:
            continue
        width[ql] = kept
        Qh = Q[:, :kept].conj().T
        c0 = 0
        for s, qr, d in parts:
            out_blocks[(s, ql)] = Qh[:, c0:c0 + d].copy()
            c0 += d
    if not width:
        return None
    return Fam(width, sketch.col, sketch.shifts, out_blocks)


def _perp(frame: Fam, T: Fam, side: str) -> Fam:
    return la.perp_left(frame, T) if side == "left" else la.perp_right(frame, T)


def dd_select(frame: Fam, sketch: Fam, side: str) -> Optional[Fam]:
    """``_dd_select``: projected = P^D sketch; thin QR at the floor 100 eps max(|sketch|,
    |projected|), then the thin QR of P^D(first pass) keeping pivots >= 1/sqrt 2."""
    from .bug import CBE_EPS

    projected = _perp(frame, sketch, side)
    nz, ny = sketch.norm(), projected.norm()
    if ny <= CBE_EPS * max(nz, 1.0):
        return None
    first = thin_qr(projected, side, CBE_EPS * max(nz, ny))
    if first is None:
        return None
    return thin_qr(_perp(frame, first, side), side, REORTH_KEEP)


def dd_probe(frame: Fam, side: str, width: int, rng, boost: Optional[Dict[int, int]] = None):
    """``_dd_probe`` with an optional sector boost (``_boost!``: extra columns per sector, capped
    by the sector's discarded dimension)."""
    from .bug import _graded_draw, alloc_columns, frame_widths, fused_dims

    dfull = fused_dims(frame, side)
    dT = frame_widths(frame, side)
    dC = {q: max(d - dT.get(q, 0), 0) for q, d in dfull.items()}
    if sum(dC.values()) == 0:
        return None
    cols = alloc_columns(dC, width)
    if boost:
        for q, extra in boost.items():
            if q not in dC:
                continue
            have = cols.get(q, 0)
            cols[q] = have + min(max(extra, 0), max(dC[q] - have, 0))
    Om = _graded_draw(frame, side, dfull, cols, rng)
    proj = la.perp_left(frame, Om) if side == "left" else la.perp_right(frame, Om)
    return proj if proj.norm() > 0 else None


def sample_residual_full(basis, project, Y: Fam, Om: Fam, side: str):
    """``_sample_residual`` -> (estimate, per-sector squared residual / columns, holdout
    columns per sector); Y is the raw sketch of the holdout Om."""
    from .bug import CBE_EPS

    cols = dict(Om.row) if side == "left" else dict(Om.col)
    cols = {q: n for q, n in cols.items() if n > 0}
    per: Dict[int, float] = {}
    if not cols:
        return 0.0, per, cols
    residual = project(Y)
    if residual.norm() <= CBE_EPS:
        return 0.0, per, cols
    if basis is not None:
        residual = project(_perp(basis, residual, side))
    sq: Dict[int, float] = {}
    for (s, q), b in residual.blocks.items():
        w = q + residual.shifts[s] if side == "left" else q
        sq[w] = sq.get(w, 0.0) + float(np.vdot(b, b).real)
    for w, v in sq.items():
        n = cols.get(w, 0)
        if n <= 0:
            raise RuntimeError("Test B: residual sector without holdout column")
        per[w] = v / n
    return math.sqrt(sum(per.values())), per, cols


def target_sectors(boost: Dict[int, int], per: Dict[int, float], Om, side: str, alloc) -> int:
    """``_target_sectors!``: every sector whose scaled residual exceeds its equal share of the
    sample allowance gets as many extra columns as the holdout gave it (at least one); ``side``
    is the expanding side (the holdout's width leg is its row leg for a left expansion)."""
    if not per:
        return 0
    have = {} if Om is None else (dict(Om.row) if side == "left" else dict(Om.col))
    share = alloc.sample ** 2 / len(per)
    n = 0
    for q, r2 in per.items():
        if alloc.scale ** 2 * r2 > share:
            boost[q] = boost.get(q, 0) + max(have.get(q, 0), 1)
            n += 1
    return n


def initial_width(U0: Fam, V0: Fam, old_rank: int, growth: float = ADAPTIVE_GROWTH) -> int:
    dmax = max(sum(U0.row.values()), old_rank, sum(V0.col.values()))
    return math.ceil(1.2 * max(math.ceil(growth * dmax) - old_rank, 1))


def dd_controller(frame: Fam, side: str, draw, sketch, project, width0: int, wcap: int, ncap: int,
                  alloc, ledger, holdout_width: int = DD_HOLDOUT,
                  acceptance_factor: float = 1.0):
    """``_dd_controller``: grow the DD basis block by block until an independent holdout finds
    the rest within the side's share.  ``draw(width, boost)`` returns a DD-projected probe on the
    opposite side, ``sketch`` the raw MPO action on it.  Returns (basis or None, status)."""
    from .bug import trim_total, oplus

    state = {"aug": frame, "dd": None, "count": 0, "discovered": 0}
    boost: Dict[int, int] = {}

    def width_of(T):
        return 0 if T is None else sum(_width_sectors(T, side).values())

    def probe_width(Om):
        return sum(Om.row.values()) if side == "left" else sum(Om.col.values())

    def absorb(sketched):
        new = dd_select(state["aug"], sketched, side)
        if new is None:
            return 0
        room = ncap - width_of(state["dd"])
        if room <= 0:
            return 0
        if width_of(new) > room:
            new = trim_total(new, side, room)
            if new is None:
                return 0
        state["dd"] = new if state["dd"] is None else oplus(state["dd"], new, side)
        state["aug"] = oplus(state["aug"], new, side)
        return width_of(new)

    def stream(Om, field):
        w = probe_width(Om)
        if field == "n_probes":
            state["discovered"] += w
            ledger.n_probes += w
        else:
            ledger.n_holdout += w
        return sketch(Om)

    def fresh(columns):
        Om = draw(columns, None)
        return None if Om is None else (Om, stream(Om, "n_holdout"))

    def resid(check):
        return sample_residual_full(state["dd"], project, check[1], check[0], side)

    def newsector():
        if state["dd"] is None:
            return False
        fw = _width_sectors(frame, side)
        return any(fw.get(q, 0) == 0 for q, n in _width_sectors(state["dd"], side).items() if n)

    width = min(max(width0, 1), max(min(wcap, ncap), 1))
    last = draw(width, None)
    if last is None:
        return None, "empty"
    state["count"] += absorb(stream(last, "n_probes"))
    check = fresh(holdout_width)
    escalated = False
    while True:
        if check is None:
            return state["dd"], "met"
        est, per, hcols = resid(check)
        ind = alloc.scale * est
        ok = ind <= acceptance_factor * alloc.sample
        if ok and not escalated and (ind > alloc.sample / 2 or newsector()):
            escalated = True
            wide = fresh(2 * holdout_width)
            if wide is not None:
                check = wide
                continue
        capped = state["discovered"] >= wcap or state["count"] >= ncap
        if ok or capped:
            ledger.max_sample = max(ledger.max_sample, ind)
            if ok:
                return state["dd"], "met"
            ledger.unmet_("sample")
            ledger.workcap_reached = True
            return state["dd"], "work_limited"
        nt = target_sectors(boost, per, check[0], side, alloc)
        ledger.n_targeted += nt
        next_check = fresh(holdout_width)
        added = absorb(check[1])
        if nt == 0:
            width = min(2 * width, wcap)
        remaining = ncap - width_of(state["dd"])
        if remaining > 0:
            blk = draw(min(width, max(remaining + 2, 1)), boost if boost else None)
            if blk is not None:
                last = blk
                added += absorb(stream(blk, "n_probes"))
        check = next_check
        state["count"] += added
        if added == 0:
            ledger.max_sample = max(ledger.max_sample, ind)
            ledger.unmet_("sample")
            return state["dd"], "basis_limited"
