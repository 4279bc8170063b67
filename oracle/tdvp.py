"""Oracle (test infrastructure only): the symmetric two-site TDVP step, restated in numpy.

Only the tests use this module, and only as the checker.  One step of tau is a forward sweep
of two-site updates over tau/2 (each followed by the backward one-site update over -tau/2 at the
next site) and the mirrored backward sweep; the two-site generator is the one-site operator of
the fused site pair (local space s1 (x) s2, MPO W1 W2 summed over the shared channel).

Reference files (madhavmen/BondUpdateGalerkin):
  - src/Baselines/TDVP/tdvp2.jl:1-301          TDVP2Info, _tdvp2_bond!, _tdvp2_backward,
                                               tdvp2_step!
  - src/BUG/Core/Environments/mpo_environments.jl:316-344   apply_h_two_site
  - src/BUG/Core/RankAdaptation/truncation.jl:41-70          _cut_divisor, _budget_svd
  - src/BUG/Core/LocalSolvers/root_methods.jl:359-402        expv (reorth off by default)
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import blocks as ob
from . import krylov as ok
from . import linalg as la
from .bug import MPS, canonical, left_boundary, right_boundary
from .blocks import Fam, MpoSite


@dataclass
class TDVP2Info:
    """``TDVP2Info``: the largest bond, the largest two-site block formed, the largest relative
    weight one cut threw away, the working set, the operator applications, and the cumulative
    relative discarded norm."""
    max_bond: int
    theta_elements: int
    discarded: float
    peak_elements: int
    krylov_dims: int
    cumulative: float
    backward_growth: float = float("-inf")


def fused_mpo(W1: MpoSite, W2: MpoSite, d: int) -> MpoSite:
    """The pair's MPO as one site of local dimension d^2: W12[a,b](s1' d + s2', s1 d + s2) =
    sum_c W1[a,c](s1', s1) W2[c,b](s2', s2)."""
    by_c = {}
    for c, b, so, si, v in W2.entries:
        by_c.setdefault(c, []).append((b, so, si, v))
    acc = {}
    for a, c, so1, si1, v1 in W1.entries:
        for b, so2, si2, v2 in by_c.get(c, ()):
            key = (a, b, so1 * d + so2, si1 * d + si2)
            acc[key] = acc.get(key, 0.0) + v1 * v2
    return MpoSite(W1.cl, W2.cr, [(a, b, so, si, v) for (a, b, so, si), v in acc.items()])


def two_site(A: Fam, B: Fam) -> Fam:
    """Theta[(s1, s2)] = A[s1] B[s2] with slice s1 d + s2, shift q(s1) + q(s2)."""
    d = len(A.shifts)
    shifts = [A.shifts[s1] + B.shifts[s2] for s1 in range(d) for s2 in range(d)]
    terms = [(s1 * d + s2, s1, s2, la.OP_N, la.OP_N, 1.0) for s1 in range(d) for s2 in range(d)]
    return la.product(A.row, B.col, shifts, A, B, terms)


def _pair_sectors(T: Fam, q1, q2):
    """Sector Q of the middle bond: row segments (s1, l-charge Q - q1[s1]) and column segments
    (s2, r-charge Q + q2[s2]) of the two-site matrix, in local-state order."""
    d1, d2 = len(q1), len(q2)
    Qs = sorted({ql + q1[s1] for s1 in range(d1) for ql in T.row})
    out = {}
    for Q in Qs:
        rows = [(s1, Q - q1[s1], T.row[Q - q1[s1]]) for s1 in range(d1) if Q - q1[s1] in T.row]
        cols = [(s2, Q + q2[s2], T.col[Q + q2[s2]]) for s2 in range(d2) if Q + q2[s2] in T.col]
        if rows and cols:
            out[Q] = (rows, cols)
    return out


def svd_two_site(T: Fam, q1, q2, cutoff: float = 0.0, nkeep: Optional[int] = None):
    """Theta = A S B over (l, s1) | (s2, r): A (l, s1, bond) left isometry, S {Q: sv},
    B (bond, s2, r) right isometry; Telum's global selection (cutoff relative to sv_max, nkeep).
    Returns (A, S, B, kept, trunc)."""
    d2 = len(q2)
    secs = _pair_sectors(T, q1, q2)
    facs, spectra = {}, {}
    for Q, (rows, cols) in secs.items():
        M = np.zeros((sum(r[2] for r in rows), sum(c[2] for c in cols)), dtype=np.complex128)
        r0 = 0
        for s1, ql, dl in rows:
            c0 = 0
            for s2, qr, dr in cols:
                b = T.blocks.get((s1 * d2 + s2, ql))
                if b is not None:
                    M[r0:r0 + dl, c0:c0 + dr] = b
                c0 += dr
            r0 += dl
        U, s, Vh = la._block_svd(M)
        facs[Q] = (U, s, Vh)
        spectra[Q] = s
    keep, kept, trunc = la._select(spectra, cutoff, nkeep)
    ab, bb, bond, S = {}, {}, {}, {}
    for Q, (U, s, Vh) in facs.items():
        k = keep[Q]
        if k == 0:
            continue
        bond[Q] = k
        S[Q] = s[:k]
        rows, cols = secs[Q]
        r0 = 0
        for s1, ql, dl in rows:
            ab[(s1, ql)] = U[r0:r0 + dl, :k]
            r0 += dl
        c0 = 0
        for s2, qr, dr in cols:
            bb[(s2, Q)] = Vh[:k, c0:c0 + dr]
            c0 += dr
    A = Fam(T.row, bond, list(q1), ab)
    B = Fam(bond, T.col, list(q2), bb)
    return A, S, B, kept, trunc


def budget_rank(kept, trunc, maxdim: int, share_of):
    """``_budget_svd``'s rank: the smallest rank whose discarded Schmidt norm fits
    ``share_of(total weight)``, never above maxdim nor below 1."""
    nk = min(len(kept), maxdim)
    total = la.sq_weight(kept) + la.sq_weight(trunc)
    share = share_of(total)
    base = la.sq_weight(trunc) + la.sq_weight(kept[nk:])
    keep, _ = ok.tail_rank(kept[:nk], 1.0, share, nk, base=base)
    return max(keep, 1)


def cut_divisor(L: int, root: int) -> float:
    """``_cut_divisor(L, root)`` with root 1-based."""
    return math.sqrt(root - 1) + math.sqrt(L - root)


def bug_root_link(L: int) -> int:
    """``_bug_root_link`` (1-based bond)."""
    return max(1, min(L - 1, L // 2))


_VEC = dict(inner=ob.inner, norm=lambda x: x.norm(), axpy=lambda a, x, y: y.axpy(a, x),
            scale=lambda a, x: x.scale(a))


def _expv(apply, tau, v, hermitian, maxiter, tol, conv_tol, substeps):
    if hermitian:
        return ok.lanczos_expv(apply, tau, v, tol=tol, maxiter=maxiter, reorth=False,
                               conv_tol=conv_tol, substeps=substeps, **_VEC)
    return ok.arnoldi_expv(apply, tau, v, tol=tol, maxiter=maxiter, conv_tol=conv_tol,
                           substeps=substeps, **_VEC)


def tdvp2_step(psi: MPS, mpo, tau: complex, *, maxdim: int = 200, trunc_thresh: float = 1e-12,
               maxiter: int = 30, tol: float = 1e-15, hermitian: bool = True,
               conv_tol: float = 0.0, substeps: int = 1, weight_tol: Optional[float] = None,
               eshift: float = 0.0, backward_report: bool = False) -> TDVP2Info:
    """``tdvp2_step!`` (tdvp2.jl:184-297): one symmetric two-site TDVP step in place, second
    order.  ``weight_tol`` replaces the relative ``trunc_thresh`` cut by the budgeted one (each
    cut's allowance weight_tol ||Theta|| / cut_divisor at the BUG's default root); ``eshift``
    integrates H - eshift; ``backward_report`` records the largest backward-substep growth
    beyond the mean-energy drift."""
    L = len(psi)
    if L < 2:
        raise ValueError("tdvp2_step needs at least two sites")
    if len(mpo) != L:
        raise ValueError(f"chain has {len(mpo)} sites, state has {L}")
    if weight_tol is not None and not weight_tol >= 0:
        raise ValueError(f"weight_tol must be non-negative, got {weight_tol}")
    tau = complex(tau)
    half = tau / 2
    d = len(psi.sites[0].shifts)
    q = list(psi.sites[0].shifts)
    wdiv = cut_divisor(L, bug_root_link(L) + 1)
    acc = dict(apps=0, theta=0, worst=0.0, peak=0, w_fwd=0.0, w_bwd=0.0, n2=0.0,
               growth=float("-inf"))

    def stored():
        return sum(sum(b.size for b in f.blocks.values()) for f in psi.sites)

    def backward(Lenv, W, Renv, M):
        if backward_report:
            n0 = M.norm()
            e1 = float(ob.inner(M, ob.apply_one_site(Lenv, W, Renv, M, shift=eshift)).real) / n0 ** 2

        def apply(x):
            acc["apps"] += 1
            return ob.apply_one_site(Lenv, W, Renv, x, shift=eshift)
        out, _ = _expv(apply, -half, M, hermitian, maxiter, tol, conv_tol, substeps)
        if backward_report:
            acc["growth"] = max(acc["growth"], math.log(out.norm() / n0) - (-half).real * e1)
        return out

    def bond(i, Lenv, Renv, forward: bool, sweep: int):
        Theta = two_site(psi.sites[i], psi.sites[i + 1])
        nth = sum(b.size for b in Theta.blocks.values())
        acc["theta"] = max(acc["theta"], nth)
        acc["peak"] = max(acc["peak"], stored() + nth)
        W12 = fused_mpo(mpo[i], mpo[i + 1], d)

        def apply(x):
            acc["apps"] += 1
            return ob.apply_one_site(Lenv, W12, Renv, x, shift=eshift)
        Theta, _ = _expv(apply, half, Theta, hermitian, maxiter, tol, conv_tol, substeps)
        if weight_tol is None:
            A, S, B, kept, trunc = svd_two_site(Theta, q, q, max(trunc_thresh, 1e-14), maxdim)
        else:
            A, S, B, kept, trunc = svd_two_site(Theta, q, q, 0.0, maxdim)
            k = budget_rank(kept, trunc, maxdim, lambda n2: weight_tol * math.sqrt(n2) / wdiv)
            if k < len(kept):
                A, S, B, kept, trunc = svd_two_site(Theta, q, q, 0.0, k)
        tot = la.sq_weight(kept) + la.sq_weight(trunc)
        acc["worst"] = max(acc["worst"], 0.0 if tot == 0 else math.sqrt(la.sq_weight(trunc) / tot))
        acc["w_fwd" if sweep == 1 else "w_bwd"] += la.sq_weight(trunc)
        acc["n2"] = max(acc["n2"], tot)
        if forward:
            psi.sites[i] = A
            lnext = ob.push_left(Lenv, A, mpo[i])
            M = la.diag_times(S, B, "left")
            if i < L - 2:
                M = backward(lnext, mpo[i + 1], Renv, M)
            psi.sites[i + 1] = M
            psi.center = i + 1
            return lnext
        psi.sites[i + 1] = B
        rnext = ob.push_right(Renv, B, mpo[i + 1])
        M = la.diag_times(S, A, "right")
        if i > 0:
            M = backward(Lenv, mpo[i], rnext, M)
        psi.sites[i] = M
        psi.center = i
        return rnext

    canonical(psi, 0)
    renv = [None] * (L + 1)
    renv[L] = right_boundary(mpo, psi)
    for i in range(L - 1, 1, -1):
        renv[i] = ob.push_right(renv[i + 1], psi.sites[i], mpo[i])
    lch = left_boundary(mpo, psi)
    for i in range(L - 1):
        lch = bond(i, lch, renv[i + 2], True, 1)
    lenv = [left_boundary(mpo, psi)]
    for i in range(L - 2):
        lenv.append(ob.push_left(lenv[-1], psi.sites[i], mpo[i]))
    rch = right_boundary(mpo, psi)
    for i in range(L - 2, -1, -1):
        canonical(psi, i)
        rch = bond(i, lenv[i], rch, False, 2)
    cum = ((math.sqrt(acc["w_fwd"]) + math.sqrt(acc["w_bwd"])) / math.sqrt(acc["n2"])
           if acc["n2"] > 0 else 0.0)
    return TDVP2Info(max(psi.bond_dims() + [0]), acc["theta"], acc["worst"], acc["peak"],
                     acc["apps"], cum, acc["growth"])
