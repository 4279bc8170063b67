"""Two-site TDVP on the device: the reference's memory and accuracy baseline for CBE-BUG, run on
the same kernels as the BUG step.

One symmetric step of tau sweeps the bonds left to right with two-site updates over tau/2 (each
followed by the backward one-site update over -tau/2 at the next site) and back.  The two-site
generator is the library's one-site H_eff apply on the fused site pair: Theta carries the d^2
slices (s1, s2) with shift q(s1) + q(s2), and the pair's MPO W1 W2 (summed over the shared
channel) is one MPO site of local dimension d^2, so the H_eff kernel runs unchanged.  The split
is one batched SVD per middle-bond charge sector of the (l, s1) | (s2, r) matricisation, gathered
and scattered by block-copy launches.

Reference (madhavmen/BondUpdateGalerkin):
  - src/Baselines/TDVP/tdvp2.jl:1-301          TDVP2Info, _tdvp2_bond!, _tdvp2_backward,
                                               tdvp2_step!
  - src/BUG/Core/Environments/mpo_environments.jl:316-344   apply_h_two_site
  - src/BUG/Core/RankAdaptation/truncation.jl:41-70          _cut_divisor, _budget_svd
  - src/BUG/Core/LocalSolvers/root_methods.jl:359-402        expv (reorth off by default)
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from . import environments as ev
from . import krylov as K
from . import tensor_ops as T
from .blocks import BlockTensor, Layout, Leg
from .bug_step import canonical
from .mpo import MPO, MPOSite
from .mps import SymMPS


@dataclass
class TDVP2Info:
    """``TDVP2Info``: the largest bond, the largest two-site block formed, the largest relative
    weight one cut threw away, the working set (elements), the operator applications, and the
    cumulative relative discarded norm; ``backward_growth`` the largest backward-substep growth
    beyond the mean-energy drift when asked for."""
    max_bond: int
    theta_elements: int
    discarded: float
    peak_elements: int
    krylov_dims: int
    cumulative: float
    backward_growth: float = float("-inf")


def fused_mpo(W1: MPOSite, W2: MPOSite, d: int) -> MPOSite:
    """The pair's MPO as one site of local dimension d^2: W12[a,b](s1' d + s2', s1 d + s2) =
    sum_c W1[a,c](s1', s1) W2[c,b](s2', s2)."""
    by_c: Dict[int, list] = {}
    for c, b, so, si, v in zip(W2.a, W2.b, W2.sout, W2.sin, W2.coeff):
        by_c.setdefault(int(c), []).append((int(b), int(so), int(si), complex(v)))
    acc: Dict[tuple, complex] = {}
    for a, c, so1, si1, v1 in zip(W1.a, W1.b, W1.sout, W1.sin, W1.coeff):
        for b, so2, si2, v2 in by_c.get(int(c), ()):
            key = (int(a), b, int(so1) * d + so2, int(si1) * d + si2)
            acc[key] = acc.get(key, 0.0) + complex(v1) * v2
    ents = list(acc.items())
    return MPOSite(
        cl=np.asarray(W1.cl, dtype=np.int32), cr=np.asarray(W2.cr, dtype=np.int32),
        a=np.array([k[0] for k, _ in ents], dtype=np.int32),
        b=np.array([k[1] for k, _ in ents], dtype=np.int32),
        sout=np.array([k[2] for k, _ in ents], dtype=np.int32),
        sin=np.array([k[3] for k, _ in ents], dtype=np.int32),
        coeff=np.array([v for _, v in ents], dtype=np.complex128))


def two_site(A: BlockTensor, B: BlockTensor) -> BlockTensor:
    """Theta[(s1, s2)] = A[s1] B[s2], slice s1 d + s2 with shift q(s1) + q(s2): one grouped
    product launch."""
    qa, qb = A.layout.shifts, B.layout.shifts
    shifts = [qa[s1] + qb[s2] for s1 in range(len(qa)) for s2 in range(len(qb))]
    terms = [(s1 * len(qb) + s2, s1, s2, T.OP_N, T.OP_N, 1.0)
             for s1 in range(len(qa)) for s2 in range(len(qb))]
    return T.product(A.layout.row, B.layout.col, shifts, A, B, terms)


class _TwoSiteSVD:
    """Per-sector SVDs of the (l, s1) | (s2, r) matricisation of a two-site tensor; selections
    read from one factorization."""

    def __init__(self, Th: BlockTensor, q1, q2):
        lay = Th.layout
        self.lay, self.q1, self.q2 = lay, list(q1), list(q2)
        rowd, cold = lay.row.as_dict(), lay.col.as_dict()
        d2 = len(q2)
        Qs = sorted({ql + q for q in q1 for ql in rowd})
        self.secs = {}
        for Q in Qs:
            rows = [(s1, Q - q1[s1], rowd[Q - q1[s1]]) for s1 in range(len(q1)) if Q - q1[s1] in rowd]
            cols = [(s2, Q + q2[s2], cold[Q + q2[s2]]) for s2 in range(d2) if Q + q2[s2] in cold]
            if rows and cols:
                self.secs[Q] = (rows, cols)
        self.qlist = list(self.secs)
        self.mats = [(sum(r[2] for r in self.secs[Q][0]), sum(c[2] for c in self.secs[Q][1]))
                     for Q in self.qlist]
        self.batch = T._SvdBatch(self.mats)
        tab = T._block_table(lay)
        g = T.CopyList()
        for i, Q in enumerate(self.qlist):
            rows, cols = self.secs[Q]
            r0 = 0
            for s1, ql, dl in rows:
                c0 = 0
                for s2, qr, dr in cols:
                    blk = tab.get((s1 * d2 + s2, ql))
                    if blk is not None:
                        T._gather_plain(g, self.batch, i, blk[0], blk[1], r0, c0, dl, dr)
                    c0 += dr
                r0 += dl
        # blocks absent from the layout stay zero in the gathered matrices
        self.batch.ws.zero_()
        g.run(Th.data, self.batch.ws)
        self.batch.run()
        self.spectra = {Q: self.batch.spectra[i] for i, Q in enumerate(self.qlist)}

    def select(self, cutoff: float = 0.0, nkeep: Optional[int] = None):
        """(A (l, s1, bond) left isometry, S {Q: sv}, B (bond, s2, r) right isometry, kept,
        trunc) by Telum's global selection."""
        keep, kept, trunc = T.select(self.spectra, cutoff, nkeep)
        bond = Leg.make((Q, keep[Q]) for Q in self.qlist if keep[Q] > 0)
        A = T.zeros(Layout(self.lay.row, bond, self.q1))
        B = T.zeros(Layout(bond, self.lay.col, self.q2))
        at, bt = T._block_table(A.layout), T._block_table(B.layout)
        cu_u, cu_v, cv_u, cv_v = T.CopyList(), T.CopyList(), T.CopyList(), T.CopyList()
        S = {}
        batch = self.batch
        for i, Q in enumerate(self.qlist):
            k = keep[Q]
            if k == 0:
                continue
            S[Q] = batch.spectra[i][:k]
            tall = batch.tall[i]
            rows, cols = self.secs[Q]
            r0 = 0
            for s1, ql, dl in rows:
                o = at[(s1, ql)]
                batch.add_u_rows(cu_u if tall else cu_v, i, r0, dl, k, o[0], o[1])
                r0 += dl
            c0 = 0
            for s2, qr, dr in cols:
                o = bt[(s2, Q)]
                batch.add_vh_cols(cv_v if tall else cv_u, i, c0, dr, k, o[0], o[1])
                c0 += dr
        cu_u.run(batch.U, A.data)
        cu_v.run(batch.Vh, A.data)
        cv_v.run(batch.Vh, B.data)
        cv_u.run(batch.U, B.data)
        return A, S, B, kept, trunc


def svd_two_site(Th: BlockTensor, q1, q2, cutoff: float = 0.0, nkeep: Optional[int] = None):
    """Theta = A S B over (l, s1) | (s2, r); returns (A, S, B, kept, trunc)."""
    return _TwoSiteSVD(Th, q1, q2).select(cutoff, nkeep)


def budget_rank(kept, trunc, maxdim: int, share_of) -> int:
    """``_budget_svd``'s rank: the smallest rank whose discarded Schmidt norm fits
    ``share_of(total weight)``, never above maxdim nor below 1."""
    nk = min(len(kept), maxdim)
    total = T.sq_weight(kept) + T.sq_weight(trunc)
    share = share_of(total)
    base = T.sq_weight(trunc) + T.sq_weight(kept[nk:])
    keep, _ = K.tail_rank(kept[:nk], 1.0, share, nk, base=base)
    return max(keep, 1)


def cut_divisor(L: int, root: int) -> float:
    """``_cut_divisor(L, root)`` with root 1-based: each budgeted cut's allowance is
    weight_tol ||psi|| / cut_divisor."""
    return math.sqrt(root - 1) + math.sqrt(L - root)


def bug_root_link(L: int) -> int:
    """``_bug_root_link`` (1-based bond)."""
    return max(1, min(L - 1, L // 2))


def _expv(apply, tau, v, hermitian, maxiter, tol, conv_tol, substeps):
    if hermitian:
        return K.lanczos_expv(apply, tau, v, tol=tol, maxiter=maxiter, reorth=False,
                              conv_tol=conv_tol, substeps=substeps)
    return K.arnoldi_expv(apply, tau, v, tol=tol, maxiter=maxiter, conv_tol=conv_tol,
                          substeps=substeps)


def tdvp2_step(psi: SymMPS, mpo: MPO, tau: complex, *, maxdim: int = 200,
               trunc_thresh: float = 1e-12, maxiter: int = 30, tol: float = 1e-15,
               hermitian: bool = True, conv_tol: float = 0.0, substeps: int = 1,
               weight_tol: Optional[float] = None, eshift: float = 0.0,
               backward_report: bool = False, backward_norms: Optional[List[float]] = None
               ) -> TDVP2Info:
    """``tdvp2_step!``: one symmetric two-site TDVP step of tau in place, second order.
    ``trunc_thresh`` is Telum's relative singular-value cut with at most ``maxdim`` kept;
    ``weight_tol`` replaces it by the budgeted cut (each allowance weight_tol ||Theta|| /
    cut_divisor at the BUG's default root); ``eshift`` integrates H - eshift;
    ``backward_report`` records the largest backward-substep growth beyond the mean-energy drift
    and ``backward_norms`` receives every backward substep's ||M_after|| / ||M_before||."""
    L = len(psi)
    if L < 2:
        raise ValueError("tdvp2_step needs at least two sites")
    if len(mpo) != L:
        raise ValueError(f"chain has {len(mpo)} sites, state has {L}")
    if weight_tol is not None and not weight_tol >= 0:
        raise ValueError(f"weight_tol must be non-negative, got {weight_tol}")
    tau = complex(tau)
    half = tau / 2
    eshift = float(eshift)
    d = psi.d
    q = list(psi[0].layout.shifts)
    wdiv = cut_divisor(L, bug_root_link(L) + 1)
    fused = [fused_mpo(mpo[i], mpo[i + 1], d) for i in range(L - 1)]
    acc = dict(apps=0, theta=0, worst=0.0, peak=0, w_fwd=0.0, w_bwd=0.0, n2=0.0,
               growth=float("-inf"))

    def stored():
        return sum(t.layout.nelem for t in psi.tensors)

    def backward(Lenv, W, Renv, M):
        n0 = T.norm(M)
        if backward_report:
            e1 = float(T.inner(M, ev.apply_one_site(Lenv, W, Renv, M, shift=eshift)).real) / n0 ** 2

        def apply(x):
            acc["apps"] += 1
            return ev.apply_one_site(Lenv, W, Renv, x, shift=eshift)
        out, _ = _expv(apply, -half, M, hermitian, maxiter, tol, conv_tol, substeps)
        if backward_report:
            acc["growth"] = max(acc["growth"], math.log(T.norm(out) / n0) - (-half).real * e1)
        if backward_norms is not None:
            backward_norms.append(T.norm(out) / n0)
        return out

    def bond(i, Lenv, Renv, forward: bool, sweep: int):
        Th = two_site(psi[i], psi[i + 1])
        acc["theta"] = max(acc["theta"], Th.layout.nelem)
        acc["peak"] = max(acc["peak"], stored() + Th.layout.nelem)
        W12 = fused[i]

        def apply(x):
            acc["apps"] += 1
            return ev.apply_one_site(Lenv, W12, Renv, x, shift=eshift)
        Th, _ = _expv(apply, half, Th, hermitian, maxiter, tol, conv_tol, substeps)
        fac = _TwoSiteSVD(Th, q, q)
        if weight_tol is None:
            A, S, B, kept, trunc = fac.select(max(trunc_thresh, 1e-14), maxdim)
        else:
            A, S, B, kept, trunc = fac.select(0.0, maxdim)
            k = budget_rank(kept, trunc, maxdim, lambda n2: weight_tol * math.sqrt(n2) / wdiv)
            if k < len(kept):
                A, S, B, kept, trunc = fac.select(0.0, k)
        tot = T.sq_weight(kept) + T.sq_weight(trunc)
        acc["worst"] = max(acc["worst"], 0.0 if tot == 0 else math.sqrt(T.sq_weight(trunc) / tot))
        acc["w_fwd" if sweep == 1 else "w_bwd"] += T.sq_weight(trunc)
        acc["n2"] = max(acc["n2"], tot)
        if forward:
            psi[i] = A
            lnext = ev.push_left(Lenv, A, mpo[i])
            M = T.diag_scale(B, S, "rows")
            if i < L - 2:
                M = backward(lnext, mpo[i + 1], Renv, M)
            psi[i + 1] = M
            psi.center = i + 1
            return lnext
        psi[i + 1] = B
        rnext = ev.push_right(Renv, B, mpo[i + 1])
        M = T.diag_scale(A, S, "cols")
        if i > 0:
            M = backward(Lenv, mpo[i], rnext, M)
        psi[i] = M
        psi.center = i
        return rnext

    canonical(psi, 0)
    renv: List[Optional[BlockTensor]] = [None] * (L + 1)
    renv[L] = ev.boundary_right(mpo, psi.bond_leg(L))
    for i in range(L - 1, 1, -1):
        renv[i] = ev.push_right(renv[i + 1], psi[i], mpo[i])
    lch = ev.boundary_left(mpo, psi.bond_leg(0))
    for i in range(L - 1):
        lch = bond(i, lch, renv[i + 2], True, 1)
    lenv = [ev.boundary_left(mpo, psi.bond_leg(0))]
    for i in range(L - 2):
        lenv.append(ev.push_left(lenv[-1], psi[i], mpo[i]))
    rch = ev.boundary_right(mpo, psi.bond_leg(L))
    for i in range(L - 2, -1, -1):
        canonical(psi, i)
        rch = bond(i, lenv[i], rch, False, 2)
    cum = ((math.sqrt(acc["w_fwd"]) + math.sqrt(acc["w_bwd"])) / math.sqrt(acc["n2"])
           if acc["n2"] > 0 else 0.0)
    return TDVP2Info(max(psi.bond_dims() + [0]), acc["theta"], acc["worst"], acc["peak"],
                     acc["apps"], cum, acc["growth"])
