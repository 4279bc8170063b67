"""Oracle (test infrastructure only): the order-p closure of the root frontiers, restated in numpy.

Only the tests, the smoke check and the benchmark's CPU baseline use this module, and only as the
checker.  The default order-p path of the reference step (``closure_checks = false``,
``frontier_tol = 0``, folded sketches): after the two half-sweeps, ONE enriching frontier sweep per
degree j = 1 .. p-1 grows the frames by the directions the projected Krylov frontier
F_j = (tau^j / j!) B^j S0 needs outside them.  ``closure_checks = true`` first measures every
frontier on the incoming frames and loops enriching passes until one adds nothing; measuring
(non-enriching) sweeps also give the outside-Galerkin action of a root vector.

The reference runs some measurements concurrently when Julia has four or more threads
(``_frontier_parallel``); this restatement takes the reference's path with fewer threads, where
every measurement runs in order on the step's own stream.

Reference files (madhavmen/BondUpdateGalerkin):
  - src/BUG/Core/RankAdaptation/krylov_closure_loop.jl:18-194    _krylov_closure! (checks = true)
  - src/BUG/Core/RankAdaptation/krylov_closure_loop.jl:203-281   _single_pass_closure!
  - src/BUG/Core/Sweeps/frontier_closure.jl:238-287              estimate_global_complement_action
  - src/BUG/Core/Sweeps/frontier_closure.jl:289-358              _measure_frontiers
  - src/BUG/Core/Sweeps/frontier_closure.jl:370-410              _closure_loop!
  - src/BUG/Core/RankAdaptation/krylov_closure_frames.jl         KLFrames, _projected_frontiers,
        _frontier_mps (:331-356), _project_root (:284-290), _root_generator (:292-299),
        _FRONTIER_FLOOR (:9)
  - src/BUG/Core/Sweeps/frontier_closure.jl:42-236               _frontier_sweep!
  - src/BUG/Core/Sweeps/frontier_sweeps.jl:20-33, 46-199, 201-354
        _polar_frame, _frontier_left, _frontier_right
  - src/BUG/Core/RankAdaptation/krylov_frontier_bonds.jl:13-51, 80-106
        _frame_onto_left/_frame_onto_right, _bond_action (fold = true)
  - src/BUG/Randomised/cbe.jl:62-132, 355-391                    _fold_sketch_closures, sampled_leak
  - src/BUG/Randomised/sketch.jl:14-54, 141-180                  sector_graded_sketch, _sample_residual

Sites are 0-based; bond b joins sites b and b+1; ``rl`` is the root bond.  Gauge moves are unique
QRs (the reference's cutoff-0 SVDs span the same spaces); the probes come from the counter-based
stream shared with the device path.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import blocks as ob
from . import linalg as la

FRONTIER_FLOOR = 1e-14


@dataclass
class KLFrames:
    """Frames of one step (krylov_closure_frames.jl:33-45): ``left[0..rl]``, ``right[rl..L-2]``
    (``right[b]`` is site b+1), ``lenv`` the environment on link rl through ``left[0..rl-1]``,
    ``renv`` the one on link rl+2 through ``right[rl+1..L-2]``; ``gen`` counts enrichments."""
    left: List
    right: List
    rl: int
    lenv: ob.Fam
    renv: ob.Fam
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


def root_generator(mpo, fr: KLFrames):
    """B = P H P - eshift at the root (``_root_generator``): two environment pushes, then
    zero-site applies."""
    Lr = ob.push_left(fr.lenv, fr.left[fr.rl], mpo[fr.rl])
    Rr = ob.push_right(fr.renv, fr.right[fr.rl], mpo[fr.rl + 1])
    es = fr.eshift
    return lambda C: ob.apply_zero_site(Lr, Rr, C, shift=es)


def project_root(psi, fr: KLFrames) -> ob.Fam:
    """S = (W (x) Z)^dag psi (``_project_root``, ``_overlap_left``/``_overlap_right``)."""
    L = len(psi)
    AL = None
    for i in range(fr.rl + 1):
        c = psi.sites[i] if AL is None else la.core_times_site(AL, psi.sites[i])
        AL = la.overlap_left(fr.left[i], c)
    AR = None
    for b in range(L - 2, fr.rl - 1, -1):
        c = psi.sites[b + 1] if AR is None else la.site_times_core(psi.sites[b + 1], AR)
        AR = la.overlap_right(c, fr.right[b])
    return la.product(AL.row, AR.col, [0], AL, AR, [(0, 0, 0, la.OP_N, la.OP_N, 1.0)])


def projected_frontiers(apply_B, root_start, tau: complex, count: int):
    """F_j = (tau / j) B F_{j-1}, j = 1..count (``_projected_frontiers``)."""
    out, F = [], root_start
    for j in range(1, count + 1):
        F = apply_B(F).scale(tau / j)
        out.append(F)
    return out


def frontier_mps(psi, fr: KLFrames, core: ob.Fam):
    """(W (x) Z) core as an MPS with its centre on site rl (``_frontier_mps``)."""
    from .bug import MPS
    L = len(psi)
    sites = [fr.left[i] for i in range(fr.rl)]
    sites.append(la.site_times_core(fr.left[fr.rl], core))
    sites += [fr.right[b] for b in range(fr.rl, L - 1)]
    return MPS(sites, center=fr.rl)


def polar_frame(frame: ob.Fam, side: str) -> ob.Fam:
    """U V^dag of frame = U S V^dag (``_polar_frame``): same span and width, orthonormal."""
    if side == "left":
        r = la.svd_left(frame, cutoff=0.0)
        return la.site_times_core(r.U, r.Vh)
    r = la.svd_right(frame, cutoff=0.0)
    return la.core_times_site(r.U, r.Vh)


def frame_onto_left(left_iso: ob.Fam, carry: ob.Fam, right: ob.Fam):
    """``_frame_onto_left``: the left isometry prescribed, right = L Q split off by a QR."""
    from .bug import BondFrame
    Lr, V0 = la.qr_right(right)
    ov = la.overlap_left(left_iso, carry)
    S0 = la.product(ov.row, Lr.col, [0], ov, Lr, [(0, 0, 0, la.OP_N, la.OP_N, 1.0)])
    return BondFrame(left_iso, S0, V0, sum(left_iso.col.values()))


def frame_onto_right(left: ob.Fam, carry: ob.Fam, right_iso: ob.Fam):
    """``_frame_onto_right``: the right isometry prescribed, left = Q R split off by a QR."""
    from .bug import BondFrame
    U0, Rl = la.qr_left(left)
    ov = la.overlap_right(carry, right_iso)
    S0 = la.product(Rl.row, ov.col, [0], Rl, ov, [(0, 0, 0, la.OP_N, la.OP_N, 1.0)])
    return BondFrame(U0, S0, right_iso, sum(right_iso.row.values()))


def sampled_leak(fr, mpo, bond: int, lenv, renv, side: str, count: int, rng) -> float:
    """Folded-sketch leak of one frontier bond (``_bond_action`` with fold = true, ``sampled_leak``):
    ``count`` graded Gaussian columns on the full opposite local space, pushed through the
    two-site action without forming it, projected off the frame and normalised per sector."""
    from .bug import graded_probe, sample_residual, stream_left, stream_right
    W1, W2 = mpo[bond], mpo[bond + 1]
    if side == "left":
        Om = graded_probe(fr.V0, "right", count, rng)
        if Om is None:
            return 0.0
        A = la.site_times_core(fr.U0, fr.S0)
        project = lambda T: la.perp_left(fr.U0, T)
        Y = project(stream_left(lenv, W1, W2, renv, A, fr.V0, Om))
    else:
        Om = graded_probe(fr.U0, "left", count, rng)
        if Om is None:
            return 0.0
        Bv = la.core_times_site(fr.S0, fr.V0)
        project = lambda T: la.perp_right(fr.V0, T)
        Y = project(stream_right(lenv, W1, W2, renv, fr.U0, Bv, Om))
    return sample_residual(None, project, Y, Om, side)


def _merge(into, other):
    """merge_ledger! (basis_budget.jl:343-383) on the fields the oracle ledger keeps."""
    into.max_sample = max(into.max_sample, other.max_sample)
    into.max_tail = max(into.max_tail, other.max_tail)
    into.max_krylov = max(into.max_krylov, other.max_krylov)
    into.max_krylov_c = max(into.max_krylov_c, other.max_krylov_c)
    into.max_krylov_d = max(into.max_krylov_d, other.max_krylov_d)
    into.n_probes += other.n_probes
    into.n_selected += other.n_selected
    into.n_krylov += other.n_krylov
    into.n_h1 += other.n_h1
    into.n_holdout += other.n_holdout
    into.n_targeted += other.n_targeted
    into.workcap_reached |= other.workcap_reached
    into.tolerance_met &= other.tolerance_met
    for u in other.unmet:
        if u not in into.unmet:
            into.unmet.append(u)
    into.bond_log.extend(other.bond_log)


def exact_leak(fr, mpo, bond: int, lenv, renv, side: str) -> float:
    """Unfolded leak (``_bond_action`` with fold = false, ``_left_leak``/``_right_leak``,
    krylov_frontier_bonds.jl:57-106): the outside-space norm of the exact two-site action
    X = H Theta_F, formed by streaming the opposite frame's full fusion basis (unitary)."""
    from .bug import fusion_probe, stream_left, stream_right
    W1, W2 = mpo[bond], mpo[bond + 1]
    if side == "left":
        Om = fusion_probe(fr.V0, "right")
        if Om is None:
            return 0.0
        Y = stream_left(lenv, W1, W2, renv, la.site_times_core(fr.U0, fr.S0), fr.V0, Om)
        return la.perp_left(fr.U0, Y).norm()
    Om = fusion_probe(fr.U0, "left")
    if Om is None:
        return 0.0
    Y = stream_right(lenv, W1, W2, renv, fr.U0, la.core_times_site(fr.S0, fr.V0), Om)
    return la.perp_right(fr.V0, Y).norm()


def _leak(tsf, mpo, bond, lenv, renv, side, width, rng, fold: bool) -> float:
    if fold:
        return sampled_leak(tsf, mpo, bond, lenv, renv, side, max(8, -(-width // 4)), rng)
    return exact_leak(tsf, mpo, bond, lenv, renv, side)


def frontier_left(Fm, mpo, fr: KLFrames, tau, alloc, opts, enrich: bool, wcap, amp: float, rng,
                  recurrence: str, fold: bool):
    """``_frontier_left`` (frontier_sweeps.jl:46-199), enriching, folded leaks."""
    from .bug import (Ledger, _cap_at, canonical, dd_cbe_expand, krylov_enrich, left_boundary,
                      record_krylov, right_boundary)
    L, rl = len(Fm), fr.rl
    canonical(Fm, 0)
    renv = [None] * (L + 1)
    renv[L] = right_boundary(mpo, Fm)
    for i in range(L - 1, 1, -1):
        renv[i] = ob.push_right(renv[i + 1], Fm.sites[i], mpo[i])
    lch = left_boundary(mpo, Fm)
    lenv = lch
    carry, Pm = None, None
    Wn = [None] * (rl + 1)
    lg = Ledger()
    leak = [0.0] * (rl + 1)
    added = 0
    for b in range(rl + 1):
        old = fr.left[b] if Pm is None else la.core_times_site(Pm, fr.left[b])
        site_core = Fm.sites[b] if carry is None else la.core_times_site(carry, Fm.sites[b])
        tsf = frame_onto_left(old, site_core, Fm.sites[b + 1])
        rch = renv[b + 2]
        width = sum(old.col.values())
        leak[b] = _leak(tsf, mpo, b, lch, rch, "left", width, rng, fold)
        quiet = alloc.scale * leak[b] <= alloc.per_bond
        if enrich and not quiet:
            bl = Ledger()
            ex = dd_cbe_expand(tsf, mpo, b, lch, rch, "left", alloc, bl, _cap_at(wcap, b), opts,
                               rng)
            Kex = la.site_times_core(site_core, la.overlap_right(Fm.sites[b + 1], ex.V_ex))
            rch1 = ob.push_right(rch, ex.V_ex, mpo[b + 1])
            Hl, Hw, es = lch, mpo[b], fr.eshift
            Wi, st = krylov_enrich(Kex, lambda v: ob.apply_one_site(Hl, Hw, rch1, v, shift=es),
                                   ex.U_ex,
                                   "left", tau, alloc, _cap_at(wcap, b), recurrence, amp=amp,
                                   state_seed=False)
            record_krylov(bl, st, alloc)
            _merge(lg, bl)
        else:
            Wi = old
        Wn[b] = polar_frame(Wi, "left") if enrich else old
        added += sum(Wn[b].col.values()) - width
        if enrich:
            Pm = la.overlap_left(Wn[b], old)
        carry = la.overlap_left(Wn[b], site_core)
        if b == rl:
            lenv = lch
        else:
            # a measuring sweep reads the frames' own environment chain (frontier_sweeps.jl:104)
            lch = ob.push_left(lch, Wn[b], mpo[b])
    return Wn, lenv, leak, added, lg


def frontier_right(Fm, mpo, fr: KLFrames, tau, alloc, opts, enrich: bool, wcap, amp: float, rng,
                   recurrence: str, fold: bool):
    """``_frontier_right`` (frontier_sweeps.jl:201-354), the mirror of ``frontier_left``."""
    from .bug import (Ledger, _cap_at, canonical, dd_cbe_expand, krylov_enrich, left_boundary,
                      record_krylov, right_boundary)
    L, rl = len(Fm), fr.rl
    canonical(Fm, L - 1)
    lenvs = [left_boundary(mpo, Fm)]
    for i in range(L - 2):
        lenvs.append(ob.push_left(lenvs[-1], Fm.sites[i], mpo[i]))
    rch = right_boundary(mpo, Fm)
    renv = rch
    carry, Pm = None, None
    Zn = [None] * (L - 1)
    lg = Ledger()
    leak = [0.0] * (L - 1)
    added = 0
    for b in range(L - 2, rl - 1, -1):
        old = fr.right[b] if Pm is None else la.site_times_core(fr.right[b], Pm)
        site_core = Fm.sites[b + 1] if carry is None else la.site_times_core(Fm.sites[b + 1], carry)
        tsf = frame_onto_right(Fm.sites[b], site_core, old)
        lch = lenvs[b]
        width = sum(old.row.values())
        leak[b] = _leak(tsf, mpo, b, lch, rch, "right", width, rng, fold)
        quiet = alloc.scale * leak[b] <= alloc.per_bond
        if enrich and not quiet:
            bl = Ledger()
            ex = dd_cbe_expand(tsf, mpo, b, lch, rch, "right", alloc, bl, _cap_at(wcap, b), opts,
                               rng)
            Bex = la.core_times_site(la.overlap_left(ex.U_ex, Fm.sites[b]), site_core)
            lch1 = ob.push_left(lch, ex.U_ex, mpo[b])
            Hw, Hr, es = mpo[b + 1], rch, fr.eshift
            Zj, st = krylov_enrich(Bex, lambda v: ob.apply_one_site(lch1, Hw, Hr, v, shift=es),
                                   ex.V_ex,
                                   "right", tau, alloc, _cap_at(wcap, b), recurrence, amp=amp,
                                   state_seed=False)
            record_krylov(bl, st, alloc)
            _merge(lg, bl)
        else:
            Zj = old
        Zn[b] = polar_frame(Zj, "right") if enrich else old
        added += sum(Zn[b].row.values()) - width
        if enrich:
            Pm = la.overlap_right(old, Zn[b])
        carry = la.overlap_right(site_core, Zn[b])
        if b == rl:
            renv = rch
        else:
            rch = ob.push_right(rch, Zn[b], mpo[b + 1])
    return Zn, renv, leak, added, lg


def frontier_sweep(fr: KLFrames, psi, mpo, core: ob.Fam, tau, alloc, opts, wcap, nrm: float, rng,
                   recurrence: str, ledger, enrich: bool = True,
                   fold: Optional[bool] = None) -> SweepResult:
    """``_frontier_sweep!`` (frontier_closure.jl:42-236): both half-sweeps of one frontier, the
    frames replaced (and their generation advanced) only if something was added; ``enrich =
    False`` only measures.  ``fold`` None takes the step's (``resolved_fold``)."""
    from .bug import resolved_fold
    fold = resolved_fold(opts) if fold is None else fold
    L, rl = len(psi), fr.rl
    nF = core.norm()
    eps = np.finfo(float).eps
    if nF <= FRONTIER_FLOOR * max(nrm, eps):
        return SweepResult(0, 0.0, True, [0.0] * (rl + 1), [0.0] * (L - 1))
    Fm = frontier_mps(psi, fr, core)
    FmR = Fm.copy()
    amp = nF / max(nrm, eps)
    # under `parallel` the right half draws from a generator split off the step's
    # (frontier_closure.jl:125); run one after the other here, they share nothing else
    rngR = rng.split() if opts.parallel else rng
    Wn, lenv, leak_l, add_l, lg_l = frontier_left(Fm, mpo, fr, tau, alloc, opts, enrich, wcap,
                                                  amp, rng, recurrence, fold)
    Zn, renv, leak_r, add_r, lg_r = frontier_right(FmR, mpo, fr, tau, alloc, opts, enrich, wcap,
                                                   amp, rngR, recurrence, fold)
    from .bug import Ledger
    lg = Ledger()
    _merge(lg, lg_l)
    _merge(lg, lg_r)
    _merge(ledger, lg)
    lL = math.sqrt(sum(x * x for x in leak_l))
    lR = math.sqrt(sum(x * x for x in leak_r))
    leak = alloc.scale * math.sqrt(lL * lL + lR * lR)
    met = lg.tolerance_met if enrich else all(
        alloc.scale * x <= alloc.per_bond for x in leak_l + leak_r)
    added = add_l + add_r
    if enrich and added > 0:
        for b in range(rl + 1):
            fr.left[b] = Wn[b]
        for b in range(rl, L - 1):
            fr.right[b] = Zn[b]
        fr.lenv, fr.renv, fr.gen = lenv, renv, fr.gen + 1
    return SweepResult(added, leak, met, leak_l, leak_r, enrich and lg.workcap_reached)


def single_pass_closure(fr: KLFrames, psi, mpo, root_start, tau, alloc, opts, wcap, nrm: float,
                        rng, recurrence: str, ledger, order: int):
    """``_single_pass_closure!`` (krylov_closure_loop.jl:203-281): one enriching sweep per
    frontier degree 1..order-1, each built in the frames the previous one grew.  Returns the root
    coefficients in the final frames and the root applications spent on the frontiers."""
    napply = [0]
    current = root_start
    gen_op = [root_generator(mpo, fr)]

    def apply_B(C):
        napply[0] += 1
        return gen_op[0](C)

    leaks = [float("nan")] * order
    met = True
    for degree in range(1, order):
        core = projected_frontiers(apply_B, current, tau, degree)[degree - 1]
        sr = frontier_sweep(fr, psi, mpo, core, tau, alloc, opts, wcap, nrm, rng, recurrence,
                            ledger)
        leaks[degree] = sr.leak
        met &= sr.met
        if sr.added > 0:
            gen_op[0] = root_generator(mpo, fr)
            current = project_root(psi, fr)
    ledger.leak = leaks
    ledger.defect = sum(leaks[j] / (j + 1) for j in range(order) if not math.isnan(leaks[j]))
    if not met:
        ledger.unmet_("closure")
    return current, napply[0]


def measure_frontiers(fr: KLFrames, psi, mpo, root_start, tau, alloc, opts, wcap, nrm: float, rng,
                      recurrence: str, count: int):
    """``_measure_frontiers`` (frontier_closure.jl:289-358): non-enriching sweeps of the frontiers
    j = 0 .. count-1 of ``root_start`` on the frames (left untouched).  Returns (leak, rho, met,
    met0, ledger, n_root): the time-weighted leaks, the scale-free ||(I - P) A F_j|| / ||F_j||,
    whether every frontier j >= 1 met its shares, the j = 0 verdict."""
    from .bug import Ledger
    napply = [0]
    gen = root_generator(mpo, fr)

    def apply_B(C):
        napply[0] += 1
        return gen(C)

    cores = [root_start] + (projected_frontiers(apply_B, root_start, tau, count - 1)
                            if count > 1 else [])
    lg = Ledger()
    leak, rho = [0.0] * count, [0.0] * count
    met, met0 = True, True
    for j in range(count):
        r = frontier_sweep(fr, psi, mpo, cores[j], tau, alloc, opts, wcap, nrm, rng, recurrence,
                           lg, enrich=False)
        nF = cores[j].norm()
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
    """``_closure_loop!`` (frontier_closure.jl:370-410): pass 1 sweeps degrees 0..p-1 (1..p-1
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


def checked_closure(fr: KLFrames, psi, mpo, root_start, tau, alloc, opts, wcap, nrm: float, rng,
                    recurrence: str, ledger, order: int, include_zero: bool = False):
    """``_krylov_closure!`` with ``checks = true`` (krylov_closure_loop.jl:18-194): measure every
    frontier on the incoming frames first; if they all close, the frames stay as the half-sweeps
    built them.  Otherwise loop enriching passes (``closure_loop``) and measure degree 0 on the
    final frames.  Returns the root coefficients in the final frames and the root applications."""
    from .bug import Ledger
    lg = Ledger()
    pleak, _, pmet, pmet0, plg, pn = measure_frontiers(fr, psi, mpo, root_start, tau, alloc, opts,
                                                       wcap, nrm, rng, recurrence, order)
    _merge(lg, plg)
    lg.leak_initial, lg.rho_initial = plg.leak_initial, plg.rho_initial
    napply = [pn]
    current = [root_start]
    final_leak = {j: pleak[j] for j in range(order)}
    gen0 = fr.gen
    if pmet and (pmet0 or not include_zero):
        converged = True
        passes = 0
    else:
        op = [root_generator(mpo, fr)]

        def apply_B(C):
            napply[0] += 1
            return op[0](C)

        def frontier(degree):
            if degree == 0:
                return current[0]
            return projected_frontiers(apply_B, current[0], tau, degree)[degree - 1]

        def sweep(core, degree, p):
            r = frontier_sweep(fr, psi, mpo, core, tau, alloc, opts, wcap, nrm, rng, recurrence,
                               lg)
            final_leak[degree] = r.leak
            if r.added > 0:
                op[0] = root_generator(mpo, fr)
                current[0] = project_root(psi, fr)
            return r

        passes, converged, _, _ = closure_loop(frontier, sweep, order, closure_maxpass(order),
                                               include_zero)
        if fr.gen != gen0:
            r0 = frontier_sweep(fr, psi, mpo, current[0], tau, alloc, opts, wcap, nrm, rng,
                                recurrence, lg, enrich=False)
            final_leak[0] = r0.leak
    lg.n_passes = passes
    lg.leak = [final_leak[j] for j in range(order)]
    lg.defect = sum(lg.leak[j] / (j + 1) for j in range(order))
    lg.n_root_frontier = napply[0]
    if not converged:
        lg.unmet_("closure")
    _merge(ledger, lg)
    ledger.leak = list(lg.leak)
    ledger.defect = max(ledger.defect, lg.defect)
    ledger.leak_initial = list(ledger.leak_initial) + list(lg.leak_initial)
    ledger.rho_initial = list(ledger.rho_initial) + list(lg.rho_initial)
    ledger.n_passes += lg.n_passes
    return current[0], napply[0]


def complement_action(psi, mpo, fr: KLFrames, core: ob.Fam, alloc, opts, wcap, rng,
                      recurrence: str):
    """``estimate_global_complement_action`` (frontier_closure.jl:238-287): the outside-Galerkin
    action (I - P) H v of v = (W (x) Z) core from one measuring sweep with exact (unfolded) bond
    actions.  For a nearest-neighbour H the per-bond pieces are orthogonal, so
    max(left, right) <= ||(I - P) H v|| <= sqrt(left^2 + right^2) = bound.  Returns a dict with
    left, right, bound and the per-bond actions."""
    from .bug import Ledger
    L, rl = len(psi), fr.rl
    eps = np.finfo(float).eps
    # on a copy of the stream, so the measurement leaves the step's draws untouched
    r = frontier_sweep(fr, psi, mpo, core, 0j, alloc, opts, wcap, max(core.norm(), eps),
                       copy.copy(rng), recurrence, Ledger(), enrich=False, fold=False)
    pl, pr = list(r.leaks_left), list(r.leaks_right[rl:L - 1])
    left = math.sqrt(sum(x * x for x in pl))
    right = math.sqrt(sum(x * x for x in pr))
    return dict(left=left, right=right, bound=math.sqrt(left * left + right * right),
                per_left=pl, per_right=pr)
