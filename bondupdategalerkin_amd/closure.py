"""Order-p closure of the root frontiers on the device (the order > 1 stage of ``cbe_bug_step``).

After the two half-sweeps the frames hold the state's own subspace; for order p > 1 the step closes
the frontiers F_j = (tau^j / j!) B^j S0 (B = P H P at the root) for j = 1 .. p-1 against the global
operator: one enriching frontier sweep per degree measures, bond by bond, the frontier's
outside-space action with a folded sketch (the two-site action is never formed), grows the bonds
whose leak exceeds their share by the same DD-QR CBE and local Krylov stage as the half-sweeps,
and re-orthonormalises every grown frame by its polar factor.  Under ``closure_checks`` every
frontier is first measured on the incoming frames (non-enriching sweeps) and enriching passes
loop until one adds nothing; a measuring sweep with exact bond actions also gives the
outside-Galerkin action of a root vector (the imaginary-time driver's defect).  Every contraction is a grouped
block GEMM of the library; factorizations are the batched QR / SVD kernels; the probes are drawn
on the host from the step's counter-based stream and uploaded.

Reference (madhavmen/BondUpdateGalerkin), ``frontier_tol = 0``, ``fold_omega`` unset so folded
for order > 1, measurements run in order (the reference's path with fewer than four Julia
threads, ``_frontier_parallel`` false):
  - _single_pass_closure!      src/BUG/Core/RankAdaptation/krylov_closure_loop.jl:203-281
  - _krylov_closure! (checks)  src/BUG/Core/RankAdaptation/krylov_closure_loop.jl:18-194
  - _measure_frontiers, _closure_loop!, estimate_global_complement_action
                               src/BUG/Core/Sweeps/frontier_closure.jl:238-410
  - KLFrames, _projected_frontiers, _frontier_mps, _project_root, _root_generator,
    _FRONTIER_FLOOR            src/BUG/Core/RankAdaptation/krylov_closure_frames.jl
  - _frontier_sweep!           src/BUG/Core/Sweeps/frontier_closure.jl:42-236
  - _polar_frame, _frontier_left, _frontier_right
                               src/BUG/Core/Sweeps/frontier_sweeps.jl:20-33, 46-199, 201-354
  - _frame_onto_left/right, _bond_action
                               src/BUG/Core/RankAdaptation/krylov_frontier_bonds.jl:13-51, 80-106
  - _fold_sketch_closures, sampled_leak   src/BUG/Randomised/cbe.jl:62-132, 355-391

Sites are 0-based, bond b joins sites b and b+1, ``rl`` is the root bond.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from . import environments as ev
from . import tensor_ops as T
from .blocks import BlockTensor
from .mps import SymMPS

FRONTIER_FLOOR = 1e-14


@dataclass
class KLFrames:
    """The frames of one step: ``left[0..rl]``, ``right[rl..L-2]`` (``right[b]`` is site b+1),
    ``lenv`` the environment on link rl through ``left[0..rl-1]``, ``renv`` the one on link rl+2
    through ``right[rl+1..L-2]``; ``gen`` counts the enrichments."""
    left: List[Optional[BlockTensor]]
    right: List[Optional[BlockTensor]]
    rl: int
    lenv: BlockTensor
    renv: BlockTensor
    gen: int = 0
    eshift: float = 0.0          # the generators' energy shift: B - eshift, H_eff - eshift


@dataclass
class SweepResult:
    added: int
    leak: float
    met: bool
    leaks_left: List[float] = field(default_factory=list)
    leaks_right: List[float] = field(default_factory=list)
    capped: bool = False         # an enrichment of the sweep stopped at the work ceiling


def closure_maxpass(order: int) -> int:
    """``_closure_maxpass`` (krylov_closure_frames.jl:12)."""
    return 2 * order + 2


class _Ctx:
    """What every bond of a frontier sweep shares: the step's settings and the probe stream."""

    def __init__(self, mpo, tau, alloc, wcap, ex, stream, recurrence, prec, htol, fold=True,
                 exact=False, parallel=False, eshift=0.0):
        self.mpo, self.tau, self.alloc, self.wcap = mpo, tau, alloc, wcap
        self.ex, self.stream, self.recurrence = ex, stream, recurrence
        self.prec, self.htol = prec, htol
        self.fold, self.exact = fold, exact
        self.parallel = parallel
        self.eshift = eshift

    def with_stream(self, stream) -> "_Ctx":
        """The same settings drawing from another probe stream."""
        return _Ctx(self.mpo, self.tau, self.alloc, self.wcap, self.ex, stream, self.recurrence,
                    self.prec, self.htol, self.fold, self.exact, self.parallel, self.eshift)


def root_generator(mpo, fr: KLFrames, prec: int):
    """B = P H P - eshift at the root: two environment pushes, then zero-site applies."""
    Lr = ev.push_left(fr.lenv, fr.left[fr.rl], mpo[fr.rl], precision=prec)
    Rr = ev.push_right(fr.renv, fr.right[fr.rl], mpo[fr.rl + 1], precision=prec)
    es = fr.eshift
    return lambda C: ev.apply_zero_site(Lr, Rr, C, shift=es, precision=prec)


def project_root(psi: SymMPS, fr: KLFrames) -> BlockTensor:
    """S = (W (x) Z)^dag psi: the left and right overlap chains, then their product."""
    L = len(psi)
    AL = None
    for i in range(fr.rl + 1):
        c = psi[i] if AL is None else T.core_times_site(AL, psi[i])
        AL = T.overlap_left(fr.left[i], c)
    AR = None
    for b in range(L - 2, fr.rl - 1, -1):
        c = psi[b + 1] if AR is None else T.site_times_core(psi[b + 1], AR)
        AR = T.overlap_right(c, fr.right[b])
    return T.core_times_core(AL, AR)


def projected_frontiers(apply_B, root_start: BlockTensor, tau: complex, count: int):
    """F_j = (tau / j) B F_{j-1}, j = 1 .. count."""
    out, F = [], root_start
    for j in range(1, count + 1):
        F = apply_B(F)
        T.scale_(tau / j, F)
        out.append(F)
    return out


def frontier_mps(psi: SymMPS, fr: KLFrames, core: BlockTensor) -> SymMPS:
    """(W (x) Z) core as an MPS with its centre on site rl."""
    L = len(psi)
    sites = [fr.left[i] for i in range(fr.rl)]
    sites.append(T.site_times_core(fr.left[fr.rl], core))
    sites += [fr.right[b] for b in range(fr.rl, L - 1)]
    return SymMPS(sites, psi.charges, center=fr.rl)


POLAR_NS_DEFECT = 1e-4


def polar_frame(frame: BlockTensor, side: str) -> BlockTensor:
    """U V^dag of frame = U S V^dag: the orthonormal frame nearest ``frame``, same span and width.

    The frames of a frontier sweep are orthonormal up to roundoff, so the polar factor is taken by
    two Newton-Schulz steps X <- X (3 - X^dag X) / 2 (four grouped GEMMs; each squares the Gram
    defect e = ||X^dag X - 1||_F up to a factor 3/4, so from e <= 1e-4 it ends below 1e-16, at the
    same polar factor) when e, estimated as sqrt(||G||^2 - 2 ||X||^2 + n), is below 1e-4;
    otherwise by the batched SVD."""
    left = side == "left"
    n = frame.layout.col.total if left else frame.layout.row.total
    gram = (lambda X: T.overlap_left(X, X)) if left else (lambda X: T.overlap_right(X, X))
    G = gram(frame)
    d2 = T.norm(G) ** 2 - 2.0 * T.norm(frame) ** 2 + n
    if n > 0 and d2 < POLAR_NS_DEFECT ** 2:
        X = frame
        for it in range(2):
            Y = T.site_times_core(X, G) if left else T.core_times_site(G, X)
            T.scale_(-0.5, Y)
            T.axpy_(1.5, X, Y)
            X = Y
            if it == 0:
                G = gram(X)
        return X
    if left:
        r = T.svd_left(frame, cutoff=0.0)
        return T.site_times_core(r.U, r.Vh)
    r = T.svd_right(frame, cutoff=0.0)
    return T.core_times_site(r.U, r.Vh)


def frame_onto_left(left_iso: BlockTensor, carry: BlockTensor, right: BlockTensor):
    """The two-site frame with the left isometry prescribed; right = L Q by a unique QR."""
    from .bug_step import BondFrame
    Lr, V0 = T.qr_right(right)
    S0 = T.core_times_core(T.overlap_left(left_iso, carry), Lr)
    return BondFrame(left_iso, S0, V0, left_iso.layout.col.total)


def frame_onto_right(left: BlockTensor, carry: BlockTensor, right_iso: BlockTensor):
    """The two-site frame with the right isometry prescribed; left = Q R by a unique QR."""
    from .bug_step import BondFrame
    U0, Rl = T.qr_left(left)
    S0 = T.core_times_core(Rl, T.overlap_right(carry, right_iso))
    return BondFrame(U0, S0, right_iso, right_iso.layout.row.total)


def sampled_leak(fr, cx: _Ctx, bond: int, lenv, renv, side: str, count: int) -> float:
    """The folded-sketch estimate of one bond's outside-space action: graded Gaussian columns on
    the full opposite local space streamed through the two-site action (four grouped launches,
    the action never formed), projected off the frame, normalised per sector (Test B)."""
    from .bug_step import graded_probe, sample_residual, stream_left, stream_right
    W1, W2 = cx.mpo[bond], cx.mpo[bond + 1]
    if side == "left":
        Om = graded_probe(fr.V0, "right", count, cx.stream)
        if Om is None:
            return 0.0
        project = lambda t: T.perp_left(fr.U0, t)
        Y = project(stream_left(lenv, W1, W2, renv, T.site_times_core(fr.U0, fr.S0), fr.V0, Om,
                                cx.prec))
    else:
        Om = graded_probe(fr.U0, "left", count, cx.stream)
        if Om is None:
            return 0.0
        project = lambda t: T.perp_right(fr.V0, t)
        Y = project(stream_right(lenv, W1, W2, renv, fr.U0, T.core_times_site(fr.S0, fr.V0), Om,
                                 cx.prec))
    return sample_residual(None, project, Y, Om, side)


def exact_leak(fr, cx: _Ctx, bond: int, lenv, renv, side: str) -> float:
    """The unfolded leak (``_bond_action`` with fold = false, ``_left_leak``/``_right_leak``,
    krylov_frontier_bonds.jl:57-106): ||(I - U0 U0^dag) X|| on the left, ||X (I - V0^dag V0)|| on
    the right, of the exact two-site action X = H Theta_F, formed by streaming the full fusion
    basis of the opposite frame (a unitary change of basis of its legs) through the action."""
    from .bug_step import fusion_probe, stream_left, stream_right
    W1, W2 = cx.mpo[bond], cx.mpo[bond + 1]
    if side == "left":
        Om = fusion_probe(fr.V0, "right")
        if Om is None:
            return 0.0
        Y = stream_left(lenv, W1, W2, renv, T.site_times_core(fr.U0, fr.S0), fr.V0, Om, cx.prec)
        return T.norm(T.perp_left(fr.U0, Y))
    Om = fusion_probe(fr.U0, "left")
    if Om is None:
        return 0.0
    Y = stream_right(lenv, W1, W2, renv, fr.U0, T.core_times_site(fr.S0, fr.V0), Om, cx.prec)
    return T.norm(T.perp_right(fr.V0, Y))


def _leak(tsf, cx: _Ctx, bond: int, lenv, renv, side: str, width: int, fold: bool) -> float:
    if fold:
        return sampled_leak(tsf, cx, bond, lenv, renv, side, max(8, -(-width // 4)))
    return exact_leak(tsf, cx, bond, lenv, renv, side)


def _enrich_bond(tsf, cx: _Ctx, bond: int, lch, rch, side: str, Fsite, site_core, amp: float,
                 ledger):
    """CBE + local Krylov on one leaking frontier bond (``_frontier_expand`` + ``_krylov_enrich``
    with the frontier's amplitude and an absolute weight)."""
    from .bug_step import AdaptiveLedger, _record_krylov, dd_cbe_expand, krylov_enrich
    bl = AdaptiveLedger()
    cap = cx.wcap[bond]
    xp = dd_cbe_expand(tsf, cx.mpo, bond, lch, rch, side, cx.alloc, bl, cap, cx.ex, cx.stream,
                       None, cx.prec, cx.exact)
    prec = cx.prec
    if side == "left":
        seed = T.site_times_core(site_core, T.overlap_right(Fsite, xp.V_ex))
        Hl, Hw, Hr = lch, cx.mpo[bond], ev.push_right(rch, xp.V_ex, cx.mpo[bond + 1], precision=prec)
        start = xp.U_ex
    else:
        seed = T.core_times_site(T.overlap_left(xp.U_ex, Fsite), site_core)
        Hl, Hw, Hr = ev.push_left(lch, xp.U_ex, cx.mpo[bond], precision=prec), cx.mpo[bond + 1], rch
        start = xp.V_ex
    es = cx.eshift
    frame, st = krylov_enrich(seed, lambda v: ev.apply_one_site(Hl, Hw, Hr, v, shift=es,
                                                              precision=prec),
                              start, side, cx.tau, cx.alloc, cap, cx.recurrence,
                              hermitian_tol=cx.htol, amp=amp, state_seed=False)
    _record_krylov(bl, st)
    bl.n_h1_closure += st.n_h1
    ledger.merge(bl)
    return frame


def frontier_left(Fm: SymMPS, fr: KLFrames, cx: _Ctx, amp: float, ledger, enrich: bool = True,
                  fold: Optional[bool] = None):
    """The left half-sweep of one frontier: per bond the old frame carried into the grown basis,
    the folded leak, and where it exceeds the bond's share CBE + Krylov growth and a polar
    re-orthonormalisation (``enrich = False`` only measures, on the frames as they are).
    Returns (frames, root environment, leaks, added)."""
    fold = cx.fold if fold is None else fold
    from .bug_step import canonical
    L, rl, mpo, prec = len(Fm), fr.rl, cx.mpo, cx.prec
    canonical(Fm, 0)
    renv: List[Optional[BlockTensor]] = [None] * (L + 1)
    renv[L] = ev.boundary_right(mpo, Fm.bond_leg(L))
    for i in range(L - 1, 1, -1):
        renv[i] = ev.push_right(renv[i + 1], Fm[i], mpo[i], precision=prec)
    lch = ev.boundary_left(mpo, Fm.bond_leg(0))
    lenv = lch
    carry = Pm = None
    Wn: List[Optional[BlockTensor]] = [None] * (rl + 1)
    leak = [0.0] * (rl + 1)
    added = 0
    for b in range(rl + 1):
        old = fr.left[b] if Pm is None else T.core_times_site(Pm, fr.left[b])
        site_core = Fm[b] if carry is None else T.core_times_site(carry, Fm[b])
        tsf = frame_onto_left(old, site_core, Fm[b + 1])
        rch = renv[b + 2]
        width = old.layout.col.total
        leak[b] = _leak(tsf, cx, b, lch, rch, "left", width, fold)
        if not enrich:
            Wn[b] = old
        else:
            if cx.alloc.scale * leak[b] <= cx.alloc.per_bond:
                Wi = old
            else:
                Wi = _enrich_bond(tsf, cx, b, lch, rch, "left", Fm[b + 1], site_core, amp, ledger)
            Wn[b] = polar_frame(Wi, "left")
            added += Wn[b].layout.col.total - width
            Pm = T.overlap_left(Wn[b], old)
        carry = T.overlap_left(Wn[b], site_core)
        if b == rl:
            lenv = lch
        else:
            lch = ev.push_left(lch, Wn[b], mpo[b], precision=prec)
    return Wn, lenv, leak, added


def frontier_right(Fm: SymMPS, fr: KLFrames, cx: _Ctx, amp: float, ledger, enrich: bool = True,
                   fold: Optional[bool] = None):
    """The mirror of ``frontier_left`` over bonds L-2 .. rl."""
    fold = cx.fold if fold is None else fold
    from .bug_step import canonical
    L, rl, mpo, prec = len(Fm), fr.rl, cx.mpo, cx.prec
    canonical(Fm, L - 1)
    lenvs = [ev.boundary_left(mpo, Fm.bond_leg(0))]
    for i in range(L - 2):
        lenvs.append(ev.push_left(lenvs[-1], Fm[i], mpo[i], precision=prec))
    rch = ev.boundary_right(mpo, Fm.bond_leg(L))
    renv = rch
    carry = Pm = None
    Zn: List[Optional[BlockTensor]] = [None] * (L - 1)
    leak = [0.0] * (L - 1)
    added = 0
    for b in range(L - 2, rl - 1, -1):
        old = fr.right[b] if Pm is None else T.site_times_core(fr.right[b], Pm)
        site_core = Fm[b + 1] if carry is None else T.site_times_core(Fm[b + 1], carry)
        tsf = frame_onto_right(Fm[b], site_core, old)
        lch = lenvs[b]
        width = old.layout.row.total
        leak[b] = _leak(tsf, cx, b, lch, rch, "right", width, fold)
        if not enrich:
            Zn[b] = old
        else:
            if cx.alloc.scale * leak[b] <= cx.alloc.per_bond:
                Zj = old
            else:
                Zj = _enrich_bond(tsf, cx, b, lch, rch, "right", Fm[b], site_core, amp, ledger)
            Zn[b] = polar_frame(Zj, "right")
            added += Zn[b].layout.row.total - width
            Pm = T.overlap_right(old, Zn[b])
        carry = T.overlap_right(site_core, Zn[b])
        if b == rl:
            renv = rch
        else:
            rch = ev.push_right(rch, Zn[b], mpo[b + 1], precision=prec)
    return Zn, renv, leak, added


def frontier_sweep(fr: KLFrames, psi: SymMPS, core: BlockTensor, cx: _Ctx, nrm: float,
                   ledger, enrich: bool = True, fold: Optional[bool] = None) -> SweepResult:
    """One sweep of a frontier on both sides.  With ``enrich`` a sweep that finds nothing to add
    leaves the frames and their generation counter as they were; ``enrich = False`` measures the
    leaks and changes nothing.
    ``fold`` None takes the step's leak form."""
    from .bug_step import AdaptiveLedger
    L, rl = len(psi), fr.rl
    nF = T.norm(core)
    eps = np.finfo(float).eps
    if nF <= FRONTIER_FLOOR * max(nrm, eps):
        return SweepResult(0, 0.0, True, [0.0] * (rl + 1), [0.0] * (L - 1))
    Fm = frontier_mps(psi, fr, core)
    FmR = Fm.copy()
    amp = nF / max(nrm, eps)
    lg, lg_l, lg_r = AdaptiveLedger(), AdaptiveLedger(), AdaptiveLedger()
    if cx.parallel:
        # the two halves concurrently, the right one on the side thread and stream with a probe
        # stream split off the step's (frontier_closure.jl:125,137-140)
        from .bug_step import _run_beside
        cxR = cx.with_stream(cx.stream.split())
        (Wn, lenv, leak_l, add_l), (Zn, renv, leak_r, add_r) = _run_beside(
            lambda: frontier_left(Fm, fr, cx, amp, lg_l, enrich, fold),
            lambda: frontier_right(FmR, fr, cxR, amp, lg_r, enrich, fold))
    else:
        Wn, lenv, leak_l, add_l = frontier_left(Fm, fr, cx, amp, lg_l, enrich, fold)
        Zn, renv, leak_r, add_r = frontier_right(FmR, fr, cx, amp, lg_r, enrich, fold)
    lg.merge(lg_l)
    lg.merge(lg_r)
    lg.n_closure_sweeps = 1 if enrich else 0
    ledger.merge(lg)
    lL = math.sqrt(sum(x * x for x in leak_l))
    lR = math.sqrt(sum(x * x for x in leak_r))
    added = add_l + add_r
    if enrich and added > 0:
        for b in range(rl + 1):
            fr.left[b] = Wn[b]
        for b in range(rl, L - 1):
            fr.right[b] = Zn[b]
        fr.lenv, fr.renv, fr.gen = lenv, renv, fr.gen + 1
    met = lg.tolerance_met if enrich else all(
        cx.alloc.scale * x <= cx.alloc.per_bond for x in leak_l + leak_r)
    return SweepResult(added, cx.alloc.scale * math.sqrt(lL * lL + lR * lR), met, leak_l, leak_r,
                       enrich and lg.workcap_reached)


def single_pass_closure(fr: KLFrames, psi: SymMPS, root_start: BlockTensor, cx: _Ctx, nrm: float,
                        ledger, order: int):
    """One enriching sweep per frontier degree 1 .. order-1, each built in the frames the previous
    one grew.  Returns the root coefficients in the final frames and the root applications spent
    on the frontiers; a frontier whose sweep left a bond over its share marks ``closure`` unmet."""
    napply = [0]
    current = root_start
    op = [root_generator(cx.mpo, fr, cx.prec)]

    def apply_B(C):
        napply[0] += 1
        return op[0](C)

    leaks = [float("nan")] * order
    met = True
    for degree in range(1, order):
        core = projected_frontiers(apply_B, current, cx.tau, degree)[degree - 1]
        sr = frontier_sweep(fr, psi, core, cx, nrm, ledger)
        leaks[degree] = sr.leak
        met &= sr.met
        if sr.added > 0:
            op[0] = root_generator(cx.mpo, fr, cx.prec)
            current = project_root(psi, fr)
    ledger.leak = leaks
    ledger.defect = sum(leaks[j] / (j + 1) for j in range(order) if not math.isnan(leaks[j]))
    if not met:
        ledger.fail("closure")
    return current, napply[0]


def measure_frontiers(fr: KLFrames, psi: SymMPS, root_start: BlockTensor, cx: _Ctx, nrm: float,
                      count: int):
    """Non-enriching sweeps of the frontiers j = 0 .. count-1 of ``root_start`` on the frames
    (left untouched).  Returns (leak, rho, met, met0, ledger, n_root): the time-weighted leaks,
    the scale-free ||(I - P) A F_j|| / ||F_j||, whether every frontier j >= 1 met its shares, the
    j = 0 verdict, the ledger of the sweeps and the root applications."""
    from .bug_step import AdaptiveLedger
    napply = [0]
    gen = root_generator(cx.mpo, fr, cx.prec)

    def apply_B(C):
        napply[0] += 1
        return gen(C)

    cores = [root_start] + (projected_frontiers(apply_B, root_start, cx.tau, count - 1)
                            if count > 1 else [])
    lg = AdaptiveLedger()
    leak, rho = [0.0] * count, [0.0] * count
    met = met0 = True
    for j in range(count):
        r = frontier_sweep(fr, psi, cores[j], cx, nrm, lg, enrich=False)
        nF = T.norm(cores[j])
        raw = math.sqrt(sum(x * x for x in r.leaks_left) + sum(x * x for x in r.leaks_right))
        leak[j] = r.leak
        rho[j] = raw / nF if nF > 0 else 0.0
        if j >= 1:
            met &= r.met
        else:
            met0 = r.met
    lg.leak_initial, lg.rho_initial = list(leak), list(rho)
    lg.n_root_frontier = napply[0]
    return leak, rho, met, met0, lg, napply[0]


def closure_loop(frontier, sweep, order: int, maxpass: int, include_zero: bool = True):
    """The termination logic of the measured closure: pass 1 sweeps degrees 0..p-1 (1..p-1
    without degree 0), later passes 1..p-1; converged when a pass adds nothing and meets every
    share, stopped as capped when a pass misses a share with a sweep at the work ceiling.
    Returns (passes, converged, exhausted, capped)."""
    later = list(range(1, order))
    initial = list(range(order)) if include_zero else later
    if not initial:
        return 0, True, False, False
    ok0 = True
    for p in range(1, maxpass + 1):
        grew, allmet, capped = False, True, False
        for degree in (initial if p == 1 else later):
            r = sweep(frontier(degree), degree, p)
            grew |= r.added > 0
            allmet &= r.met
            capped |= r.capped
            if degree == 0:
                ok0 = r.met
        if not grew or not later:
            return p, allmet and ok0, False, False
        if capped and not allmet:
            return p, False, False, True
    return maxpass, False, True, False


def checked_closure(fr: KLFrames, psi: SymMPS, root_start: BlockTensor, cx: _Ctx, nrm: float,
                    ledger, order: int, include_zero: bool = False):
    """The closure with ``closure_checks``: measure every frontier on the incoming frames; if
    they all close, the frames stay as the half-sweeps built them, otherwise loop enriching
    passes (``closure_loop``) and measure degree 0 on the final frames.  Returns the root
    coefficients in the final frames and the root applications."""
    from .bug_step import AdaptiveLedger
    lg = AdaptiveLedger()
    pleak, _, pmet, pmet0, plg, pn = measure_frontiers(fr, psi, root_start, cx, nrm, order)
    lg.merge(plg)
    lg.leak_initial, lg.rho_initial = list(plg.leak_initial), list(plg.rho_initial)
    napply = [pn]
    current = [root_start]
    final_leak = {j: pleak[j] for j in range(order)}
    gen0 = fr.gen
    if pmet and (pmet0 or not include_zero):
        converged, passes = True, 0
    else:
        op = [root_generator(cx.mpo, fr, cx.prec)]

        def apply_B(C):
            napply[0] += 1
            return op[0](C)

        def frontier(degree):
            if degree == 0:
                return current[0]
            return projected_frontiers(apply_B, current[0], cx.tau, degree)[degree - 1]

        def sweep(core, degree, p):
            r = frontier_sweep(fr, psi, core, cx, nrm, lg)
            final_leak[degree] = r.leak
            if r.added > 0:
                op[0] = root_generator(cx.mpo, fr, cx.prec)
                current[0] = project_root(psi, fr)
            return r

        passes, converged, _, _ = closure_loop(frontier, sweep, order, closure_maxpass(order),
                                               include_zero)
        if fr.gen != gen0:
            final_leak[0] = frontier_sweep(fr, psi, current[0], cx, nrm, lg, enrich=False).leak
    lg.n_passes = passes
    lg.leak = [final_leak[j] for j in range(order)]
    lg.defect = sum(lg.leak[j] / (j + 1) for j in range(order))
    lg.n_root_frontier = napply[0]
    if not converged:
        lg.fail("closure")
    ledger.merge(lg)
    ledger.leak = list(lg.leak)
    ledger.defect = max(ledger.defect, lg.defect)
    ledger.leak_initial = list(ledger.leak_initial) + list(lg.leak_initial)
    ledger.rho_initial = list(ledger.rho_initial) + list(lg.rho_initial)
    ledger.n_passes += lg.n_passes
    return current[0], napply[0]


def complement_action(psi: SymMPS, fr: KLFrames, core: BlockTensor, cx: _Ctx) -> dict:
    """The outside-Galerkin action (I - P) H v of v = (W (x) Z) core from one measuring sweep
    with exact (unfolded) bond actions; for a nearest-neighbour H,
    max(left, right) <= ||(I - P) H v|| <= sqrt(left^2 + right^2) = bound.  Returns left, right,
    bound and the per-bond actions."""
    from .bug_step import AdaptiveLedger
    L, rl = len(psi), fr.rl
    eps = np.finfo(float).eps
    # on a copy of the stream, so the measurement leaves the step's draws untouched
    r = frontier_sweep(fr, psi, core, cx.with_stream(copy.copy(cx.stream)), max(T.norm(core), eps),
                       AdaptiveLedger(), enrich=False, fold=False)
    pl, pr = list(r.leaks_left), list(r.leaks_right[rl:L - 1])
    left = math.sqrt(sum(x * x for x in pl))
    right = math.sqrt(sum(x * x for x in pr))
    return dict(left=left, right=right, bound=math.sqrt(left * left + right * right),
                per_left=pl, per_right=pr)
