"""``cbe_bug_step``: one rank-adaptive CBE-BUG step on a device-resident U(1) MPS.

The drop-in for ``cbe_bug_step!(psi, mpo, tau; kwargs...)`` (src/BUG/TimeEvolution/bug_step.jl:
17-513): same stages, options, tolerance shares and ledger verdict.  Every tensor stays on the
device; the host makes the data-dependent decisions (ranks, Krylov stops, budget checks) from
downloaded scalars and spectra.

Stages (0-based sites; bond b joins sites b and b+1; link j sits left of site j):

* options / budget         bug_step_stages.jl:15-142; basis_budget.jl:75-128,454-494,568
* half-sweeps              src/BUG/Core/Sweeps/half_sweeps.jl:218-410
* bond frames              effective_operators.jl:201-220 (unique-QR gauge)
* DD-QR CBE ``:rank_tail`` src/BUG/Randomised/dd_cbe_expand.jl:19-87,252-301,383-419; dd_cbe.jl:
                           38-89,132-161; cbe.jl:188-236; sketch.jl:141-180; streamed sketch
                           mpo_environments.jl:409-481
* local Krylov (C/D)       krylov_enrichment.jl:135-292; krylov_lanczos.jl; krylov_arnoldi.jl
* projection miss          krylov_closure_frames.jl:242-280
* root solve               effective_operators.jl:91-107; krylov.jl:128-263
* assembly + compression   diagnostics.jl:269-337; truncation.jl:49-158

* order-p closure (order > 1)  krylov_closure_loop.jl:18-281 (``closure.single_pass_closure``,
                           and the measured loop under ``closure_checks``)
"""
from __future__ import annotations

import math
import os
import sys
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from . import environments as ev
from . import krylov as K
from . import tensor_ops as T
from .blocks import BlockTensor, Layout, Leg
from .mpo import MPO, MPOSite
from .mps import SymMPS
from .probes import ProbeStream

CBE_EPS = T.CBE_EPS
TOL_DEFAULTS = dict(basis=1e-5, trunc=1e-7, expv=1e-10)
# BUG_PROGRESS=1 prints one line per 16 bonds of each half-sweep to stderr (long steps).
_PROGRESS = os.environ.get("BUG_PROGRESS", "") == "1"


def _progress(msg: str):
    if _PROGRESS:
        print(f"# {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)
DEFAULT_BOND_CEILING_FACTOR = 1.2


@dataclass
class DDQRExpansion:
    """The DD-QR CBE strategy (``DDQRExpansion``, src/BUG/Randomised/cbe.jl:254-290)."""

    mode: str = "rank_tail"
    fixed_width: int = 8
    tail_factor: float = 0.10
    krylov_mode: str = "auto"
    holdout_width: int = 4

    def __post_init__(self):
        if self.mode not in ("rank_tail", "adaptive"):
            raise ValueError(f"DD mode {self.mode!r} is not implemented (rank_tail, adaptive)")
        if self.mode == "rank_tail" and self.fixed_width < 1:
            raise ValueError("rank_tail requires positive fixed_width")
        if self.holdout_width < 1:
            raise ValueError("holdout_width must be positive")
        if not 0 < self.tail_factor <= 1:
            raise ValueError("tail_factor must lie in (0, 1]")
        if self.krylov_mode not in ("auto", "lanczos", "arnoldi"):
            raise ValueError("krylov_mode must be auto, lanczos or arnoldi")


@dataclass
class AdaptiveLedger:
    """The step's verdict (``AdaptiveLedger``, basis_budget.jl:144-299)."""

    max_sample: float = 0.0
    max_tail: float = 0.0
    max_krylov: float = 0.0
    max_krylov_c: float = 0.0
    max_krylov_d: float = 0.0
    discarded: float = 0.0
    expv_err: float = 0.0
    expv_iters: int = 0
    n_probes: int = 0
    n_selected: int = 0
    n_krylov: int = 0
    n_h1: int = 0
    n_holdout: int = 0
    n_targeted: int = 0
    workcap_reached: bool = False
    tolerance_met: bool = True
    maxdim_reached: bool = False
    proj_err: float = 0.0
    r_old: int = 0
    r_aug: int = 0
    r_final: int = 0
    # order-p closure: time-weighted frontier leaks (index = degree; NaN = not measured), their
    # defect sum_j leak_j / (j + 1), and the one-site applies the frontier sweeps spent
    leak: List[float] = field(default_factory=list)
    defect: float = 0.0
    n_h1_closure: int = 0
    n_closure_sweeps: int = 0
    # measured closure (closure_checks) and closure_probe: the frontiers' leaks on the incoming
    # frames (time-weighted and scale-free), the enriching passes, the root applications spent
    leak_initial: List[float] = field(default_factory=list)
    rho_initial: List[float] = field(default_factory=list)
    n_passes: int = 0
    n_root_frontier: int = 0
    unmet: List[str] = field(default_factory=list)
    # one entry per half-sweep bond visit (0-based bond), written by log_bond; after the
    # compression each entry also holds the bond's r_old, r_final and relative discarded norm
    bond_log: List[dict] = field(default_factory=list)

    def fail(self, what: str):
        self.tolerance_met = False
        if what not in self.unmet:
            self.unmet.append(what)

    def merge(self, other: "AdaptiveLedger"):
        """``merge_ledger!`` (basis_budget.jl:343-383) on the fields kept here."""
        self.max_sample = max(self.max_sample, other.max_sample)
        self.max_tail = max(self.max_tail, other.max_tail)
        self.max_krylov = max(self.max_krylov, other.max_krylov)
        self.max_krylov_c = max(self.max_krylov_c, other.max_krylov_c)
        self.max_krylov_d = max(self.max_krylov_d, other.max_krylov_d)
        self.n_probes += other.n_probes
        self.n_selected += other.n_selected
        self.n_krylov += other.n_krylov
        self.n_h1 += other.n_h1
        self.n_holdout += other.n_holdout
        self.n_targeted += other.n_targeted
        self.n_h1_closure += other.n_h1_closure
        self.n_closure_sweeps += other.n_closure_sweeps
        self.workcap_reached |= other.workcap_reached
        self.tolerance_met &= other.tolerance_met
        for u in other.unmet:
            if u not in self.unmet:
                self.unmet.append(u)
        self.bond_log.extend(other.bond_log)


def log_bond(ledger: AdaptiveLedger, bl: AdaptiveLedger, bond: int, side: str, width: int):
    """Close one half-sweep bond visit (``_log_bond!``, basis_budget.jl:390-418): ``bl`` is the
    fresh ledger the bond's Tests A/B/C/D wrote into; its summary is logged and it is merged
    into ``ledger``, so the step's maxima and sums are the aggregate of the logged bonds."""
    bl.bond_log.append(dict(
        bond=bond, side=side, sample=bl.max_sample, tail=bl.max_tail, krylov_c=bl.max_krylov_c,
        krylov_d=bl.max_krylov_d, n_probes=bl.n_probes, n_holdout=bl.n_holdout,
        n_selected=bl.n_selected, n_targeted=bl.n_targeted, n_krylov=bl.n_krylov, n_h1=bl.n_h1,
        workcap=bl.workcap_reached, met=bl.tolerance_met, unmet=list(bl.unmet), width=width))
    ledger.merge(bl)


_SIDE = {}


def _side_executor():
    """The persistent side thread of parallel half-sweeps (one per process, created on first use:
    its library context and stream then live as long as the thread)."""
    ex = _SIDE.get("pool")
    if ex is None:
        import concurrent.futures

        ex = _SIDE["pool"] = concurrent.futures.ThreadPoolExecutor(
            max_workers=1, thread_name_prefix="bug-half-sweep")
    return ex


def _run_beside(main_fn: Callable, side_fn: Callable):
    """Run ``side_fn`` on the side thread's HIP stream and ``main_fn`` on the caller's, ordered at
    both ends: the side stream first waits for everything queued so far on the caller's stream,
    and the caller's stream waits for the side stream's work before the results are used."""
    from .blocks import device

    main_stream = torch.cuda.current_stream()
    fork = torch.cuda.Event()
    fork.record(main_stream)

    def side():
        local = _SIDE.setdefault(threading.get_ident(), {})
        st = local.get("stream")
        if st is None:
            st = local["stream"] = torch.cuda.Stream(device=device())
        st.wait_event(fork)
        with torch.cuda.stream(st):
            out = side_fn()
            done = torch.cuda.Event()
            done.record(st)
        return out, done

    fut = _side_executor().submit(side)
    try:
        a = main_fn()
    finally:
        b, done = fut.result()
    main_stream.wait_event(done)
    return a, b


def budget_status(ledger: AdaptiveLedger) -> str:
    return "met" if ledger.tolerance_met else "budget_unmet"


@dataclass
class CBEBugSweepInfo:
    """What one step reports (``CBEBugSweepInfo``, diagnostics.jl:209-255)."""

    ledger: AdaptiveLedger
    expanded: List[int]
    root_link: int
    root_rank: int
    discarded: float
    root_applications: int
    seconds: Dict[str, float]


@dataclass
class _Alloc:
    per_bond: float
    sample: float
    tail: float
    krylov: float
    scale: float
    krylov_rank: float = 0.0             # > 0: ranked Krylov selection (basis_budget.jl:37-48)


def basis_allocation(tol_basis, nbonds, tau, nrm, nfront, tol_krylov=None,
                     krylov_rank=0.0) -> _Alloc:
    """The ``:linear`` rule the step uses (basis_budget.jl:75-104); ``krylov_rank`` (tol_krylov's
    units) is split over the bonds the same way; it bounds what a ranked Krylov stage may drop."""
    nb, nf = max(nbonds, 1), max(nfront, 1)
    kr = 1.0 if tol_krylov is None else tol_krylov / tol_basis
    if kr <= 0:
        raise ValueError("tol_krylov must be positive")
    if not krylov_rank >= 0:
        raise ValueError("krylov_rank_tol must be >= 0")
    at = abs(tau)
    scale = 1.0 if at == 0 else at / max(nrm, np.finfo(float).eps)
    per_test = tol_basis / (2 * (nb + 1) * nf)
    return _Alloc(2 * per_test, per_test, per_test, kr * per_test, scale,
                  krylov_rank / tol_basis * per_test)


def bond_work_ceilings(ranks, d, maxdim, factor):
    """Per-bond working-rank caps (bond_work_ceilings, basis_budget.jl:484-494)."""
    def one(r):
        return math.ceil(round(factor * d * r, 9)) if factor > 0 else 2 * d * r
    return [max(min(one(r), maxdim), 1) for r in ranks]


def bond_work_ceilings_strict(ranks, maxdim, factor):
    """min(ceil(f D_b), maxdim) per bond, no local-dimension factor (bond_work_ceilings_strict,
    basis_budget.jl:505-512)."""
    if not factor > 1:
        raise ValueError("strict bond factor must exceed 1")
    return [max(1, min(maxdim, math.ceil(round(factor * r, 9)))) for r in ranks]


def bond_frame_root_ceilings(ranks, maxdim, root_bond, frame_factor=2.0, root_factor=1.2):
    """Two tiers (bond_frame_root_ceilings, basis_budget.jl:526-549): ceil(frame_factor D_b)
    away from the root bond and ceil(root_factor D_root) at it (0-based root_bond), each within
    [1, maxdim], all from the bonds at the start of the step."""
    if not 0 <= root_bond < len(ranks):
        raise IndexError("root bond outside the chain")
    if not frame_factor > 1:
        raise ValueError("frame_factor must exceed 1")
    if not root_factor > 1:
        raise ValueError("root_factor must exceed 1")
    if root_factor > frame_factor:
        raise ValueError("root_factor cannot exceed frame_factor")
    if maxdim < 1:
        raise ValueError("maxdim must be positive")
    out = []
    for b, r in enumerate(ranks):
        f = root_factor if b == root_bond else frame_factor
        out.append(max(1, min(maxdim, math.ceil(round(f * r, 9)))))
    return out


# ---- gauge --------------------------------------------------------------------------------

def canonical(psi: SymMPS, site: int) -> SymMPS:
    """canonical! (canonical_form.jl:43-54) with unique-QR moves."""
    if not 0 <= site < len(psi):
        raise IndexError(f"site {site} outside the chain")
    while psi.center < site:
        i = psi.center
        Q, R = T.qr_left(psi[i])
        psi[i] = Q
        psi[i + 1] = T.core_times_site(R, psi[i + 1])
        psi.center += 1
    while psi.center > site:
        i = psi.center
        Lc, Q = T.qr_right(psi[i])
        psi[i] = Q
        psi[i - 1] = T.site_times_core(psi[i - 1], Lc)
        psi.center -= 1
    return psi


# ---- probes and sketches ------------------------------------------------------------------

def _fused_dims(lay: Layout, side: str) -> Dict[int, int]:
    out: Dict[int, int] = {}
    if side == "left":
        for q, d in zip(lay.row.charges, lay.row.dims):
            for sh in lay.shifts:
                out[q + sh] = out.get(q + sh, 0) + d
    else:
        for q, d in zip(lay.col.charges, lay.col.dims):
            for sh in lay.shifts:
                out[q - sh] = out.get(q - sh, 0) + d
    return dict(sorted(out.items()))


def _widths(lay: Layout, side: str) -> Dict[int, int]:
    return lay.col.as_dict() if side == "left" else lay.row.as_dict()


def _width(t: BlockTensor, side: str) -> int:
    return t.layout.col.total if side == "left" else t.layout.row.total


def _graded_draw(frame: BlockTensor, side: str, dfull: Dict[int, int], cols: Dict[int, int],
                 stream: ProbeStream) -> BlockTensor:
    """Gaussian columns on the fused local space of ``frame``, drawn on the host from the step's
    stream and uploaded (``_randomize_like`` + ``_trim_per_sector``, cbe.jl:156-178): per fused
    charge ascending, per slice ascending, one (link x dfull) block column-major of which the
    first ``cols`` columns are kept."""
    lay = frame.layout
    wleg = Leg.make((q, c) for q, c in cols.items() if c > 0)
    link = lay.row if side == "left" else lay.col
    play = Layout(link, wleg, lay.shifts) if side == "left" else Layout(wleg, link, lay.shifts)
    # the stream positions of every kept number, generated in one vectorised pass
    spans, pos = [], stream.position
    for q, df in dfull.items():
        c = cols.get(q, 0)
        for s, sh in enumerate(lay.shifts):
            lq = q - sh if side == "left" else q + sh
            dl = link.dim_of(lq)
            if dl == 0:
                continue
            if c > 0:
                spans.append((s, q, lq, dl, c, pos))
            pos += dl * df
    stream.skip(pos - stream.position)
    blocks = {}
    if spans:
        idx = np.concatenate([np.arange(p0, p0 + dl * c, dtype=np.uint64)
                              for (_, _, _, dl, c, p0) in spans])
        vals = stream.at(idx)
        o = 0
        for s, q, lq, dl, c, _ in spans:
            blk = vals[o: o + dl * c].reshape(c, dl).T
            o += dl * c
            if side == "left":
                blocks[(s, lq)] = blk
            else:
                blocks[(s, q)] = blk.T.copy()
    return BlockTensor.from_blocks(play, blocks)


def _alloc_columns(dims: Dict[int, int], width: int) -> Dict[int, int]:
    """``_alloc_columns`` (cbe.jl:188-195)."""
    tot = sum(dims.values())
    if tot == 0 or width <= 0:
        return {q: 0 for q in dims}
    return {q: min(d, math.ceil(d / tot * width)) for q, d in dims.items()}


def dd_probe(frame: BlockTensor, side: str, width: int, stream: ProbeStream,
             boost: Optional[Dict[int, int]] = None) -> Optional[BlockTensor]:
    """Sector-graded Gaussian probe in the discarded space of ``frame`` (_dd_probe, dd_cbe.jl:132-161):
    columns per fused charge by its discarded dimension (plus a sector ``boost``, ``_boost!``,
    cbe.jl:206-216, capped by the discarded dimension), the frame projected out afterwards."""
    lay = frame.layout
    dfull = _fused_dims(lay, side)
    dT = _widths(lay, side)
    dC = {q: max(d - dT.get(q, 0), 0) for q, d in dfull.items()}
    if sum(dC.values()) == 0:
        return None
    cols = _alloc_columns(dC, width)
    if boost:
        for q, extra in boost.items():
            if q in dC:
                have = cols.get(q, 0)
                cols[q] = have + min(max(extra, 0), max(dC[q] - have, 0))
    Om = _graded_draw(frame, side, dfull, cols, stream)
    proj = T.perp_left(frame, Om) if side == "left" else T.perp_right(frame, Om)
    return proj if T.norm(proj) > 0 else None


def graded_probe(frame: BlockTensor, side: str, width: int, stream: ProbeStream) -> Optional[BlockTensor]:
    """``sector_graded_sketch`` with comp_ratio = nothing (sketch.jl:45-54): columns shared out by
    the FULL fused dimension of each sector, no projector in the probe."""
    dfull = _fused_dims(frame.layout, side)
    cols = _alloc_columns(dfull, width)
    if not any(c > 0 for c in cols.values()):
        return None
    Om = _graded_draw(frame, side, dfull, cols, stream)
    return Om if T.norm(Om) > 0 else None


def _slots(entries):
    slots = {}
    for e in entries:
        slots.setdefault(e, len(slots))
    return slots


def stream_left(L, W1: MPOSite, W2: MPOSite, R, A, V, Om, prec: int = _lib.PREC_FP64) -> BlockTensor:
    """(H Theta) Om^dag for Theta = A V with Om in V's layout: (link_l, s, g)
    (_stream_left, mpo_environments.jl:409-444), four grouped launches."""
    N, Tt, C, Rr = T.OP_N, T.OP_T, T.OP_C, T.OP_R
    Ls, Rs, As, Vs, Os = L.layout.shifts, R.layout.shifts, A.layout.shifts, V.layout.shifts, Om.layout.shifts
    e2 = list(zip(W2.a, W2.b, W2.sout, W2.sin, W2.coeff))
    e1 = list(zip(W1.a, W1.b, W1.sout, W1.sin, W1.coeff))
    fs = _slots([(c, si) for c, b, so, si, _ in e2])
    fsh = {}
    tf = []
    for c, b, so, si, co in e2:
        k = fs[(c, si)]
        fsh[k] = -Rs[b] - Os[so]
        tf.append((k, b, so, Tt, C, co))
    F = T.product(R.layout.col, Om.layout.row, [fsh[k] for k in range(len(fs))], R, Om, tf,
                  precision=prec)
    cs = sorted({c for c, _ in fs})
    ci = {c: i for i, c in enumerate(cs)}
    bsh, tb = {}, []
    for (c, si), k in fs.items():
        tb.append((ci[c], si, k, N, N, 1.0))
        bsh[ci[c]] = Vs[si] + fsh[k]
    B = T.product(V.layout.row, Om.layout.row, [bsh[i] for i in range(len(cs))], V, F, tb,
                  precision=prec)
    cslots, csh, tc = {}, {}, []
    for a, c, so, si, co in e1:
        if c in ci and (c, si) not in cslots:
            k = cslots[(c, si)] = len(cslots)
            tc.append((k, si, ci[c], N, N, 1.0))
            csh[k] = As[si] + bsh[ci[c]]
    Cm = T.product(A.layout.row, Om.layout.row, [csh[k] for k in range(len(cslots))], A, B, tc,
                   precision=prec)
    ty = [(so, a, cslots[(c, si)], N, N, co) for a, c, so, si, co in e1 if (c, si) in cslots]
    return T.product(L.layout.row, Om.layout.row, As, L, Cm, ty, precision=prec)


def stream_right(L, W1: MPOSite, W2: MPOSite, R, U, Bv, Om, prec: int = _lib.PREC_FP64) -> BlockTensor:
    """Om^dag (H Theta) for Theta = U Bv with Om in U's layout: (g, s, link_r)
    (_stream_right, mpo_environments.jl:446-481), four grouped launches."""
    N, Tt, C, Rr = T.OP_N, T.OP_T, T.OP_C, T.OP_R
    Ls, Rs, Us, Bs, Os = L.layout.shifts, R.layout.shifts, U.layout.shifts, Bv.layout.shifts, Om.layout.shifts
    e1 = list(zip(W1.a, W1.b, W1.sout, W1.sin, W1.coeff))
    e2 = list(zip(W2.a, W2.b, W2.sout, W2.sin, W2.coeff))
    fs = _slots([(c, si) for a, c, so, si, _ in e1])
    fsh, tf = {}, []
    for a, c, so, si, co in e1:
        k = fs[(c, si)]
        fsh[k] = -Ls[a] + Os[so]
        tf.append((k, a, so, Tt, Rr, co))
    F = T.product(L.layout.col, Om.layout.col, [fsh[k] for k in range(len(fs))], L, Om, tf,
                  precision=prec)
    cs = sorted({c for c, _ in fs})
    ci = {c: i for i, c in enumerate(cs)}
    bsh, tb = {}, []
    for (c, si), k in fs.items():
        tb.append((ci[c], si, k, Tt, N, 1.0))
        bsh[ci[c]] = -Us[si] + fsh[k]
    B = T.product(U.layout.col, Om.layout.col, [bsh[i] for i in range(len(cs))], U, F, tb,
                  precision=prec)
    cslots, csh, tc = {}, {}, []
    for c, b, so, si, co in e2:
        if c in ci and (c, si) not in cslots:
            k = cslots[(c, si)] = len(cslots)
            tc.append((k, ci[c], si, Tt, N, 1.0))
            csh[k] = -bsh[ci[c]] + Bs[si]
    Cm = T.product(Om.layout.col, Bv.layout.col, [csh[k] for k in range(len(cslots))], B, Bv, tc,
                   precision=prec)
    ty = [(so, cslots[(c, si)], b, N, Tt, co) for c, b, so, si, co in e2 if (c, si) in cslots]
    return T.product(Om.layout.col, R.layout.row, Bs, Cm, R, ty, precision=prec)


# ---- frames -------------------------------------------------------------------------------

@dataclass
class BondFrame:
    """The canonical snapshot at a bond (``BondFrame``, src/BUG/Core/Sweeps/frames.jl:17-30)."""

    U0: BlockTensor
    S0: BlockTensor
    V0: BlockTensor
    old_rank: int


def frame_from(left: BlockTensor, right: BlockTensor) -> BondFrame:
    """``_frame_from`` (effective_operators.jl:201-220)."""
    U0, Rl = T.qr_left(left)
    Lr, V0 = T.qr_right(right)
    return BondFrame(U0, T.core_times_core(Rl, Lr), V0, U0.layout.col.total)


def _perp(side):
    return T.perp_left if side == "left" else T.perp_right


def reortho(frame: BlockTensor, block: Optional[BlockTensor], side: str) -> Optional[BlockTensor]:
    """``_reortho_left/right`` (truncation.jl:374-394)."""
    if block is None:
        return None
    outside = _perp(side)(frame, block)
    if T.norm(outside) <= CBE_EPS * max(T.norm(block), 1.0):
        return None
    if side == "left":
        r = T.svd_left(outside, cutoff=1e-14)
        return r.U if r.kept else None
    r = T.svd_right(outside, cutoff=1e-14)
    return r.Vh if r.kept else None


def cap_to_local_room(frame: BlockTensor, cand: Optional[BlockTensor], side: str):
    """``_cap_to_local_room`` (truncation.jl:332-351)."""
    if cand is None:
        return None
    dfull = _fused_dims(frame.layout, side)
    dT = _widths(frame.layout, side)
    counts, over = {}, False
    for q, du in _widths(cand.layout, side).items():
        if du == 0:
            continue
        counts[q] = min(du, dfull.get(q, 0) - dT.get(q, 0))
        over |= counts[q] < du
    return cand if not over else T.trim(cand, side, counts)


def trim_total(t: BlockTensor, side: str, n: int) -> Optional[BlockTensor]:
    """``_trim_total`` (cbe.jl:219-236): the first sectors in ascending charge order."""
    if n <= 0:
        return None
    w = _widths(t.layout, side)
    if sum(w.values()) <= n:
        return t
    left, keep = n, {}
    for q in sorted(w):
        keep[q] = min(w[q], left)
        left -= keep[q]
    return T.trim(t, side, keep)


def _oplus(frame, new, side):
    return T.oplus_left(frame, new) if side == "left" else T.oplus_right(frame, new)


def sample_residual(basis, project, Y: BlockTensor, Om: BlockTensor, side: str) -> float:
    """Test B's per-width-sector normalised residual (``_sample_residual``, sketch.jl:141-180)."""
    return sample_residual_full(basis, project, Y, Om, side)[0]


def sample_residual_full(basis, project, Y: BlockTensor, Om: BlockTensor, side: str):
    """``_sample_residual`` -> (estimate, {width sector: squared residual / its columns},
    {width sector: holdout columns})."""
    cols = Om.layout.row.as_dict() if side == "left" else Om.layout.col.as_dict()
    cols = {q: n for q, n in cols.items() if n > 0}
    per: Dict[int, float] = {}
    if not cols:
        return 0.0, per, cols
    res = project(Y)
    if T.norm(res) <= CBE_EPS:
        return 0.0, per, cols
    if basis is not None:
        res = project(_perp(side)(basis, res))
    blks, keys = [], []
    for k, q, qc, d, dc, off in res.layout.blocks():
        blks.append((off, d, d, dc))
        keys.append(qc if side == "left" else q)
    sq = T.block_sqnorms(blks, res.data)
    acc: Dict[int, float] = {}
    for w, v in zip(keys, sq):
        acc[w] = acc.get(w, 0.0) + float(v)
    for w, v in acc.items():
        if cols.get(w, 0) <= 0:
            raise RuntimeError("Test B: residual sector without a holdout column")
        per[w] = v / cols[w]
    return math.sqrt(sum(per.values())), per, cols


REORTH_KEEP = 1.0 / math.sqrt(2.0)      # _REORTH_KEEP, dd_cbe.jl:273
ADAPTIVE_GROWTH = 1.2                  # basis_budget.jl:565


def dd_select(frame: BlockTensor, sketch: BlockTensor, side: str) -> Optional[BlockTensor]:
    """``_dd_select`` (dd_cbe.jl:287-321): projected = P^D sketch, its pivoted thin QR at the
    floor 100 eps max(|sketch|, |projected|), then the thin QR of P^D(first pass) keeping the
    pivots >= 1/sqrt 2 ("twice is enough")."""
    perp = _perp(side)
    projected = perp(frame, sketch)
    nz, ny = T.norm(sketch), T.norm(projected)
    if ny <= CBE_EPS * max(nz, 1.0):
        return None
    first = T.thin_qr(projected, side, CBE_EPS * max(nz, ny))
    if first is None:
        return None
    return T.thin_qr(perp(frame, first), side, REORTH_KEEP)


def target_sectors(boost: Dict[int, int], per: Dict[int, float], Om, side: str, alloc) -> int:
    """``_target_sectors!`` (sketch.jl:188-201): a sector whose scaled residual exceeds its equal
    share of the sample allowance gets as many extra probe columns as the holdout gave it."""
    if not per:
        return 0
    have = {} if Om is None else (Om.layout.row.as_dict() if side == "left"
                                  else Om.layout.col.as_dict())
    share = alloc.sample ** 2 / len(per)
    n = 0
    for q, r2 in per.items():
        if alloc.scale ** 2 * r2 > share:
            boost[q] = boost.get(q, 0) + max(have.get(q, 0), 1)
            n += 1
    return n


def dd_controller(frame: BlockTensor, side: str, draw, sketch, project, width0: int, wcap: int,
                  ncap: int, alloc, ledger: "AdaptiveLedger", holdout_width: int = 4,
                  acceptance_factor: float = 1.0):
    """``_dd_controller`` (dd_cbe_controller.jl:26-212; ``batch_blocks = 1``, storage-order cap):
    grow the DD basis block by block until an independent holdout finds the rest within the
    side's share.  Returns (basis or None, status)."""
    st = {"aug": frame, "dd": None, "count": 0, "discovered": 0}
    boost: Dict[int, int] = {}

    def width_of(t):
        return 0 if t is None else _width(t, side)

    def probe_width(Om):
        return Om.layout.row.total if side == "left" else Om.layout.col.total

    def absorb(sketched):
        new = dd_select(st["aug"], sketched, side)
        if new is None:
            return 0
        room = ncap - width_of(st["dd"])
        if room <= 0:
            return 0
        if width_of(new) > room:
            new = trim_total(new, side, room)
            if new is None:
                return 0
        st["dd"] = new if st["dd"] is None else _oplus(st["dd"], new, side)
        st["aug"] = _oplus(st["aug"], new, side)
        return width_of(new)

    def stream(Om, holdout):
        w = probe_width(Om)
        if holdout:
            ledger.n_holdout += w
        else:
            st["discovered"] += w
            ledger.n_probes += w
        return sketch(Om)

    def fresh(columns):
        Om = draw(columns, None)
        return None if Om is None else (Om, stream(Om, True))

    def newsector():
        if st["dd"] is None:
            return False
        fw = _widths(frame.layout, side)
        return any(fw.get(q, 0) == 0 for q, n in _widths(st["dd"].layout, side).items() if n)

    width = min(max(width0, 1), max(min(wcap, ncap), 1))
    last = draw(width, None)
    if last is None:
        return None, "empty"
    st["count"] += absorb(stream(last, False))
    check = fresh(holdout_width)
    escalated = False
    while True:
        if check is None:
            return st["dd"], "met"
        est, per, _ = sample_residual_full(st["dd"], project, check[1], check[0], side)
        ind = alloc.scale * est
        ok = ind <= acceptance_factor * alloc.sample
        if ok and not escalated and (ind > alloc.sample / 2 or newsector()):
            escalated = True
            wide = fresh(2 * holdout_width)
            if wide is not None:
                check = wide
                continue
        capped = st["discovered"] >= wcap or st["count"] >= ncap
        if ok or capped:
            ledger.max_sample = max(ledger.max_sample, ind)
            if ok:
                return st["dd"], "met"
            ledger.fail("sample")
            ledger.workcap_reached = True
            return st["dd"], "work_limited"
        nt = target_sectors(boost, per, check[0], side, alloc)
        ledger.n_targeted += nt
        next_check = fresh(holdout_width)
        added = absorb(check[1])
        if nt == 0:
            width = min(2 * width, wcap)
        remaining = ncap - width_of(st["dd"])
        if remaining > 0:
            blk = draw(min(width, max(remaining + 2, 1)), boost if boost else None)
            if blk is not None:
                last = blk
                added += absorb(stream(blk, False))
        check = next_check
        st["count"] += added
        if added == 0:
            ledger.max_sample = max(ledger.max_sample, ind)
            ledger.fail("sample")
            return st["dd"], "basis_limited"


@dataclass
class _Expansion:
    U_ex: BlockTensor
    V_ex: BlockTensor
    n_new: int


def _dd_cbe_adaptive(fr: BondFrame, frame, opp, side: str, oside: str, lenv, renv, W1, W2,
                     alloc: _Alloc, ledger: "AdaptiveLedger", cap: int, wbind: bool, full_opp: int,
                     workcap: int, ex: "DDQRExpansion", stream: ProbeStream,
                     prec: int) -> "_Expansion":
    """dd_cbe_expand's ``:adaptive`` branch (dd_cbe_expand.jl:88-118,362-433)."""
    left = side == "left"
    dmax = max(fr.U0.layout.row.total, fr.old_rank, fr.V0.layout.col.total)
    width0 = math.ceil(1.2 * max(math.ceil(ADAPTIVE_GROWTH * dmax) - fr.old_rank, 1))
    wcap = min(full_opp, workcap) if workcap > 0 else full_opp
    if left:
        A = T.site_times_core(fr.U0, fr.S0)
        sketch = lambda Om: stream_left(lenv, W1, W2, renv, A, fr.V0, Om, prec)     # noqa: E731
        project = lambda t: T.perp_left(fr.U0, t)                                  # noqa: E731
    else:
        Bv = T.core_times_site(fr.S0, fr.V0)
        sketch = lambda Om: stream_right(lenv, W1, W2, renv, fr.U0, Bv, Om, prec)  # noqa: E731
        project = lambda t: T.perp_right(fr.V0, t)                                 # noqa: E731
    directions, _status = dd_controller(
        frame, side, lambda w, b: dd_probe(opp, oside, w, stream, b), sketch, project, width0,
        wcap, cap, alloc, ledger, holdout_width=ex.holdout_width)
    return _finish_expansion(fr, frame, side, directions, cap, wbind, ledger)


def fusion_probe(frame: BlockTensor, side: str) -> Optional[BlockTensor]:
    """The fusion basis of ``frame``'s fused local space (``fusion_basis``, dd_cbe.jl:143): per
    fused charge the slices stacked in ascending order, one identity block each, in the probe
    layout of ``_graded_draw`` (a unitary change of basis of the fused legs)."""
    lay = frame.layout
    dfull = _fused_dims(lay, side)
    link = lay.row if side == "left" else lay.col
    wl = sorted((q, d) for q, d in dfull.items() if d > 0)
    if not wl:
        return None
    wleg = Leg.make(wl)
    play = Layout(link, wleg, lay.shifts) if side == "left" else Layout(wleg, link, lay.shifts)
    blocks = {}
    for q, df in wl:
        off = 0
        for s, sh in enumerate(lay.shifts):
            lq = q - sh if side == "left" else q + sh
            dl = link.dim_of(lq)
            if dl == 0:
                continue
            blk = np.zeros((dl, df), dtype=complex)
            blk[:, off: off + dl] = np.eye(dl)
            off += dl
            blocks[(s, lq) if side == "left" else (s, q)] = blk if side == "left" else blk.T.copy()
    return BlockTensor.from_blocks(play, blocks)


def exact_probe(frame: BlockTensor, side: str) -> Optional[BlockTensor]:
    """``_dd_probe`` with exact = true (dd_cbe.jl:141-148,155-160): the fusion basis with the
    frame projected out, i.e. the whole discarded space.  No random numbers."""
    lay = frame.layout
    dfull = _fused_dims(lay, side)
    dT = _widths(lay, side)
    if sum(max(d - dT.get(q, 0), 0) for q, d in dfull.items()) == 0:
        return None
    Om = fusion_probe(frame, side)
    proj = T.perp_left(frame, Om) if side == "left" else T.perp_right(frame, Om)
    return proj if T.norm(proj) > 0 else None


def _dd_cbe_exact(fr: BondFrame, frame, opp, side: str, oside: str, lenv, renv, W1, W2,
                  ledger: "AdaptiveLedger", cap: int, wbind: bool, prec: int) -> "_Expansion":
    """dd_cbe_expand with exact = true (dd_cbe_expand.jl:91-97,346-360): ``_certified_range``
    without an allocation is ``_range_find`` (range_finder.jl:40-48, sketch.jl:119-132), whose one
    pass at the full width sketches the exact probe and keeps ``_dd_select`` of it; no probes or
    holdouts are counted."""
    Om = exact_probe(opp, oside)
    if Om is None:
        return _Expansion(fr.U0, fr.V0, 0)
    if side == "left":
        Ysk = stream_left(lenv, W1, W2, renv, T.site_times_core(fr.U0, fr.S0), fr.V0, Om, prec)
    else:
        Ysk = stream_right(lenv, W1, W2, renv, fr.U0, T.core_times_site(fr.S0, fr.V0), Om, prec)
    return _finish_expansion(fr, frame, side, dd_select(frame, Ysk, side), cap, wbind, ledger)


def _finish_expansion(fr: BondFrame, frame, side: str, directions, cap: int, wbind: bool,
                      ledger: "AdaptiveLedger") -> "_Expansion":
    """The common end of dd_cbe_expand (dd_cbe_expand.jl:383-419)."""
    left = side == "left"
    none = _Expansion(fr.U0, fr.V0, 0)
    if directions is None:
        return none
    count = _width(directions, side)
    if count > cap:
        directions = trim_total(directions, side, cap)
        if directions is None:
            return none
    if wbind and count >= cap:
        ledger.workcap_reached = True
    directions = cap_to_local_room(frame, directions, side)
    if directions is None:
        return none
    nq = _width(directions, side)
    ledger.n_selected += nq
    grown = _oplus(frame, directions, side)
    return _Expansion(grown, fr.V0, nq) if left else _Expansion(fr.U0, grown, nq)


def dd_cbe_expand(fr: BondFrame, mpo: MPO, bond: int, lenv, renv, side: str, alloc: _Alloc,
                  ledger: AdaptiveLedger, workcap: int, ex: DDQRExpansion, stream: ProbeStream,
                  probe_hook=None, prec: int = _lib.PREC_FP64, exact: bool = False) -> _Expansion:
    """One side of one bond, ``:rank_tail`` (dd_cbe_expand.jl:19-87,252-301,383-419), the
    adaptive controller, or the deterministic ``exact`` reference."""
    left = side == "left"
    frame, opp = (fr.U0, fr.V0) if left else (fr.V0, fr.U0)
    oside = "right" if left else "left"
    none = _Expansion(fr.U0, fr.V0, 0)
    full_self = sum(_fused_dims(frame.layout, side).values())
    full_opp = sum(_fused_dims(opp.layout, oside).values())
    cap = full_self - fr.old_rank
    wbind = False
    if workcap > 0:
        wk = max(workcap - fr.old_rank, 0)
        wbind = wk < cap
        cap = min(cap, wk)
    if cap <= 0:
        return none
    W1, W2 = mpo[bond], mpo[bond + 1]
    if exact:
        return _dd_cbe_exact(fr, frame, opp, side, oside, lenv, renv, W1, W2, ledger, cap, wbind,
                             prec)
    if ex.mode == "adaptive":
        if probe_hook is not None:
            raise ValueError("probe_hook observes the rank_tail probes only")
        return _dd_cbe_adaptive(fr, frame, opp, side, oside, lenv, renv, W1, W2, alloc, ledger,
                                cap, wbind, full_opp, workcap, ex, stream, prec)
    width = min(ex.fixed_width, full_opp)
    Om = probe_hook(opp, oside, width) if probe_hook else dd_probe(opp, oside, width, stream)
    if Om is None:
        return none
    ledger.n_probes += _width(Om, oside)
    if left:
        Ysk = stream_left(lenv, W1, W2, renv, T.site_times_core(fr.U0, fr.S0), fr.V0, Om, prec)
        project = lambda t: T.perp_left(fr.U0, t)
    else:
        Ysk = stream_right(lenv, W1, W2, renv, fr.U0, T.core_times_site(fr.S0, fr.V0), Om, prec)
        project = lambda t: T.perp_right(fr.V0, t)
    Y = project(Ysk)
    ny = T.norm(Y)
    if not math.isfinite(ny):
        raise FloatingPointError(f"dd_cbe_expand: non-finite sketch at bond {bond} ({side})")
    if ny <= CBE_EPS:
        return none
    rmax = min(cap, _width(Om, oside))
    cache = {}
    fac = T.factor_left(Y) if left else T.factor_right(Y)

    def candidate(k):
        if k <= 0:
            return None
        if k not in cache:
            r = fac.select(CBE_EPS, k)
            cache[k] = (reortho(frame, r.U if left else r.Vh, side) if r.kept else None)
        return cache[k]

    limit = ex.tail_factor * alloc.sample

    def metric(k):
        return alloc.scale * sample_residual(candidate(k), project, Y, Om, side)

    if metric(0) <= limit:
        return none
    hi = rmax
    if metric(hi) > limit:
        directions, count = candidate(hi), hi
    else:
        lo = 0
        while hi - lo > 1:
            k = (hi + lo) // 2
            if metric(k) <= limit:
                hi = k
            else:
                lo = k
        directions = candidate(hi)
        count = 0 if directions is None else _width(directions, side)
    if directions is None:
        return none
    if count > cap:
        directions = trim_total(directions, side, cap)
        if directions is None:
            return none
    if wbind and count >= cap:
        ledger.workcap_reached = True
    directions = cap_to_local_room(frame, directions, side)
    if directions is None:
        return none
    nq = _width(directions, side)
    ledger.n_selected += nq
    grown = _oplus(frame, directions, side)
    return _Expansion(grown, fr.V0, nq) if left else _Expansion(fr.U0, grown, nq)


# ---- local Krylov enrichment ----------------------------------------------------------------

def extend_frame(frame, residual, direction, side, weight, allowance, room):
    """``_extend_frame`` (krylov_enrichment.jl:135-171) -> (frame, added, fits)."""
    rn = T.norm(residual)
    cut = CBE_EPS * max(T.norm(direction), rn) / rn
    fac = T.factor_left(residual) if side == "left" else T.factor_right(residual)
    res = fac.select(cut)
    nk = len(res.kept)
    keep, tail = K.tail_rank(res.kept, weight, allowance, min(nk, room), base=T.sq_weight(res.trunc))
    fits = tail <= allowance
    if keep == 0:
        return frame, 0, fits
    if keep < nk:
        res = fac.select(cut, keep)
    new = reortho(frame, res.U if side == "left" else res.Vh, side)
    new = cap_to_local_room(frame, new, side)
    if new is None:
        return frame, 0, fits
    return _oplus(frame, new, side), _width(new, side), fits


UNION_KEEP = 0.5


def union_fresh(frame, cand, side, tol=0.0, room=1 << 60):
    """``_union_fresh`` (half_sweeps.jl:117-176): orthonormal directions of ``cand`` outside
    ``frame`` that keep at least half of themselves under a second projection (twice is enough),
    or None.  ``tol = 0`` admits everything above the numerical floor; ``tol > 0`` only
    directions carrying about ``tol ||cand||``; at most ``room``, the strongest first."""
    if room <= 0:
        return None
    perp = _perp(side)
    factor = T.factor_left if side == "left" else T.factor_right
    pick = (lambda r: r.U) if side == "left" else (lambda r: r.Vh)          # noqa: E731
    residual = perp(frame, cand)
    rn, cn = T.norm(residual), T.norm(cand)
    if rn <= CBE_EPS * max(cn, 1.0) or rn <= tol * cn:
        return None
    first_cut = max(CBE_EPS * max(cn, rn) / rn, tol * cn / rn)
    basis = pick(factor(T.scale(1.0 / rn, residual)).select(first_cut))
    second = perp(frame, basis)
    fac = factor(second)
    smax = max((e[0] for e in fac.select(1e-13).kept), default=0.0)
    if smax < UNION_KEEP:
        return None
    return pick(fac.select(UNION_KEEP / smax, None if room >= 1 << 60 else room))


def union_frame(frame, cand, side, tol=0.0, room=1 << 60):
    """``_union_left``/``_union_right`` (half_sweeps.jl:184-207): the frame verbatim, then the
    fresh directions of ``cand`` appended."""
    fresh = union_fresh(frame, cand, side, tol, room)
    return frame if fresh is None else _oplus(frame, fresh, side)


def append_ranked(frame, candidates, side, allowance, workcap):
    """``_append_ranked`` (krylov_enrichment.jl:185-196) -> (frame, added, fits): one SVD of the
    direct sum of the weighted residuals, all outside the same frame, keeps the fewest top
    directions whose discarded tail fits ``allowance``."""
    if not candidates:
        return frame, 0, True
    room = workcap - _width(frame, side) if workcap > 0 else 1 << 60
    block = candidates[0]
    for c in candidates[1:]:
        block = _oplus(block, c, side)
    if room <= 0:
        return frame, 0, T.norm(block) <= allowance
    return extend_frame(frame, block, block, side, 1.0, allowance, room)


@dataclass
class _KrylovStats:
    eta_c: float = 0.0
    eta_d: float = 0.0
    added: int = 0
    n_h1: int = 0
    capped: bool = False
    met: bool = True


def krylov_enrich(seed, applyH, start_frame, side, tau, alloc: _Alloc, workcap: int,
                  recurrence: str = "lanczos", maxdepth: int = K.MAXDEPTH,
                  hermitian_tol: float = 1e-9, amp: float = 1.0, state_seed: bool = True):
    """Tests C/D: grow the frame by the local Krylov directions the step needs
    (``_krylov_enrich`` + ``_lanczos_chain``/``_arnoldi_chain``, krylov_enrichment.jl:213-292).
    ``amp`` scales every weight (a frontier seed's norm relative to the state)."""
    perp = _perp(side)
    allow = alloc.krylov
    ranked = alloc.krylov_rank > 0
    candidates = []
    st = _KrylovStats()
    frame = [start_frame]

    def score(direction, weight, first):
        residual = perp(frame[0], direction)
        eta = weight * T.norm(residual)
        if first:
            st.eta_c = eta
        else:
            st.eta_d = max(st.eta_d, eta)
        if not eta > allow:
            return True
        if ranked:                       # measured against start_frame, selected at the end
            candidates.append(T.scale(weight, residual))
            return True
        room = workcap - _width(frame[0], side) if workcap > 0 else 1 << 60
        if room <= 0:
            return False
        frame[0], nadd, fits = extend_frame(frame[0], residual, direction, side, weight, allow, room)
        st.added += nadd
        return fits

    sn = T.norm(seed)
    if sn == 0:
        return frame[0], st
    teff = tau if tau != 0 else 1.0 + 0j
    vs = [T.scale(1.0 / sn, seed)]
    H = np.zeros((maxdepth + 1, maxdepth + 1), dtype=complex)
    closed = capped = False
    prev_beta = 0.0
    for k in range(1, maxdepth + 1):
        img = applyH(vs[k - 1] if recurrence == "lanczos" else vs[-1])
        st.n_h1 += 1
        if recurrence == "lanczos":
            if k > 1:
                T.axpy_(-prev_beta, vs[k - 2], img)
            a = T.inner(vs[k - 1], img)
            if not abs(a.imag) <= hermitian_tol * max(abs(a), 1.0):
                raise ValueError("Lanczos needs a Hermitian one-site generator")
            H[k - 1, k - 1] = a.real
            T.axpy_(-a.real, vs[k - 1], img)
            if k >= 20 and k % 2 == 0:
                cs = T.inner_many(vs, img)
                T.axpy_many_([-c for c in cs], vs, img)
        else:
            for _p in range(2):
                cs = T.inner_many(vs, img)
                for row, c in enumerate(cs):
                    H[row, k - 1] += c
                T.axpy_many_([-c for c in cs], vs, img)
        beta = T.norm(img)
        if beta <= K.BREAKDOWN:
            closed = True
            break
        H[k, k - 1] = beta
        if recurrence == "lanczos":
            H[k - 1, k] = beta
        vs.append(T.scale(1.0 / beta, img))
        w = amp * K.direction_weight(H, k, teff, state_seed)
        if not score(vs[-1], w, k == 1):
            capped = True
            break
        if w <= allow:
            closed = True
            break
        prev_beta = beta
    if not closed:
        capped = True
    if ranked:
        frame[0], st.added, fits = append_ranked(start_frame, candidates, side, alloc.krylov_rank,
                                                 workcap)
        capped = capped or not fits
    st.capped = capped
    st.met = closed and not capped
    return frame[0], st


def _record_krylov(ledger: AdaptiveLedger, st: _KrylovStats):
    ledger.max_krylov_c = max(ledger.max_krylov_c, st.eta_c)
    ledger.max_krylov_d = max(ledger.max_krylov_d, st.eta_d)
    ledger.max_krylov = max(ledger.max_krylov, st.eta_c, st.eta_d)
    ledger.n_krylov += st.added
    ledger.n_h1 += st.n_h1
    if st.capped:
        ledger.workcap_reached = True
    if not st.met:
        ledger.fail("krylov")


# ---- the step -----------------------------------------------------------------------------

def cbe_bug_step(psi: SymMPS, mpo: MPO, tau: complex, *, tol_basis: Optional[float] = None,
                 tol_trunc: Optional[float] = None, tol_expv: Optional[float] = None,
                 tol_root: Optional[float] = None, tol_krylov: Optional[float] = None,
                 krylov_rank_tol: float = 0.0, exact: bool = False,
                 fold_omega: Optional[bool] = None, keep_states: Sequence[SymMPS] = (),
                 keep_tols: Sequence[float] = (), frames_only: bool = False,
                 frames_report: Optional[dict] = None,
                 order: int = 1, maxdim: Optional[int] = None, hermitian: bool = True,
                 work_ceiling_rule: str = "bond", bond_ceiling_factor: Optional[float] = None,
                 root_ceiling_factor: float = 1.2,
                 root: Optional[int] = None, reorth: bool = True, truncate: bool = True,
                 expansion: Optional[DDQRExpansion] = None, rng=None,
                 probe_hook: Optional[Callable] = None, precision: str = "fp64",
                 parallel: bool = False, eshift: float = 0.0, closure_checks: bool = False,
                 closure_probe: int = 0, root_report: Optional[dict] = None,
                 outside_residual: bool = False, root_method: str = "exp",
                 root_shift: bool = True, root_maxiter: int = K.ROOT_MAXITER) -> CBEBugSweepInfo:
    """One order-p step in place (``order > 1`` closes the root frontiers j = 1..p-1 by one
    enriching frontier sweep each, ``closure.py``); ``root`` is the 0-based root bond (the
    reference's ``root`` minus 1),
    ``rng`` the probe seed (default 0x5EED, fresh per call as the reference's default RNG), or a
    ``ProbeStream`` the caller keeps drawing from across steps.

    ``eshift`` makes the step integrate H - eshift (bug_step.jl:88-91): every local Krylov
    generator and the root generator are shifted, so the basis weights measure the change of
    the normalised state; 0 leaves the step unchanged.  ``closure_checks`` (order > 1) measures
    every frontier on the incoming frames first and loops enriching passes until one adds nothing
    (krylov_closure_loop.jl:18-194); ``closure_probe = m`` (order 1) only measures frontiers
    0..m-1 into the ledger's ``leak_initial``/``rho_initial``, on a copy of the probe stream.
    ``root_report``, a dict, receives the root's account (bug_step_stages.jl:242-398): the
    energies before and after the root solve and, with ``outside_residual``, the
    outside-Galerkin defect |tau| ||(I - P) H psi_1|| / ||psi_1|| of the evolved vector, the
    frames re-enriched by its own frontier sweep and the root re-solved while it exceeds
    tol_basis.  ``root_method`` (bug_step.jl:76-87) is the root integrator: "exp" (the Krylov
    exponential) or the stability function of one implicit Runge-Kutta step, "trapezoid",
    "radau2" or "radau3", applied in one stage on a basis of at most ``root_maxiter`` vectors;
    ``root_shift = False`` applies R(tau A_G) unshifted, its decay kept (a linear PDE flow).

    ``krylov_rank_tol > 0`` ranks the local Krylov directions instead of appending them in chain
    order (krylov_enrichment.jl:174-196,209-211): every direction above the allowance is collected
    against the bond's starting frame, and one SVD of their weighted direct sum keeps the fewest
    top ones whose discarded tail fits the apportioned ``krylov_rank_tol``.

    ``exact = True`` is the reference's deterministic debug path (bug_step.jl:117-119): each
    bond's DD expansion sketches the whole discarded space of the opposite frame instead of
    Gaussian probes, so the step draws no random numbers and serial and ``parallel`` half-sweeps
    agree.  ``fold_omega`` (default: ``order > 1 and not exact``, bug_step_stages.jl:67) picks the
    closure's frontier leak: folded (a Gaussian sketch of the two-site action, never formed) or
    unfolded (the exact two-site action's outside-space norm).

    ``keep_states`` (with ``keep_tols``, default 0 each) are states the trial space must contain
    (half_sweeps.jl:117-207,280-296,384-395): after each bond's Krylov stage their carried site
    tensors are unioned into the frame, exactly at tol 0 (the MPDO identity, a mass state), at
    tol > 0 only directions carrying about tol of them, within the bond's work ceiling under the
    per-bond rules.

    ``frames_only = True`` (with ``truncate = False``) skips the root solve: the state comes back
    as itself in the step's frames, for a caller that solves its own root; ``frames_report``, a
    dict, receives the frames (``left``, ``right``, ``root_link``).

    ``work_ceiling_rule`` caps every bond's working rank from the bonds before the step
    (bug_step_stages.jl:116-129): ``"bond"`` ceil(f d D_b) (f defaults to 1.2), ``"maxdim"`` one
    cap d max(maxdim, D_max), ``"bond_strict"`` ceil(f D_b) (f > 1 required), ``"frame_root"``
    ceil(f D_b) away from the root bond (f defaults to 2) and ceil(root_ceiling_factor D_root) at
    it; each per-bond cap is clipped to [1, maxdim].

    ``precision = "mixed"`` runs the dense block contractions (environment pushes, one- and
    zero-site applies, streamed sketches) on FP32 MFMA with the complex products split into real
    ones and FP32 accumulation; Krylov orthogonalisation, the small exponentials, QR and SVD stay
    FP64.  The Hermitian check of the local Lanczos chain is then held to FP32 rounding (1e-5).

    ``parallel = True`` runs the two half-sweeps concurrently (bug_step.jl:130-134,222-268): the
    right one on a side host thread with its own library context and HIP stream, drawing from a
    probe stream split off the step's (so its probes differ from the serial step's, as the
    reference's do)."""
    L = len(psi)
    # -- option checks, before anything touches the state (bug_step_stages.jl:41-111) --
    if len(mpo) != L:
        raise ValueError(f"MPO has {len(mpo)} sites, state has {L}")
    if L < 2:
        raise ValueError("cbe_bug_step needs at least two sites")
    rl = (max(1, min(L - 1, L // 2)) - 1) if root is None else root
    if not 0 <= rl <= L - 2:
        raise ValueError(f"root must be a bond in 0:{L - 2}, got {rl}")
    if tol_root is not None and tol_expv is not None:
        raise ValueError("tol_root and tol_expv name the same tolerance; pass one")
    tol_expv = tol_root if tol_root is not None else tol_expv
    tb = TOL_DEFAULTS["basis"] if tol_basis is None else tol_basis
    tt = TOL_DEFAULTS["trunc"] if tol_trunc is None else tol_trunc
    te = TOL_DEFAULTS["expv"] if tol_expv is None else tol_expv
    if not tb > 0 or not tt >= 0 or not te > 0:
        raise ValueError("tolerances must be positive (tol_trunc non-negative)")
    if order < 1:
        raise ValueError(f"order must be >= 1, got {order}")
    if not krylov_rank_tol >= 0:
        raise ValueError(f"krylov_rank_tol must be >= 0, got {krylov_rank_tol}")
    fold = (order > 1 and not exact) if fold_omega is None else bool(fold_omega)
    keep_states = list(keep_states)
    if keep_tols and len(keep_tols) != len(keep_states):
        raise ValueError(f"keep_tols has {len(keep_tols)} entries for {len(keep_states)} "
                         "keep_states")
    keep_tols = [float(t) for t in keep_tols] if keep_tols else [0.0] * len(keep_states)
    if any(not t >= 0 for t in keep_tols):
        raise ValueError("keep_tols must be >= 0")
    if any(len(k) != L for k in keep_states):
        raise ValueError("every kept state needs the chain's length")
    cap_keep = work_ceiling_rule != "maxdim"
    if frames_only and (truncate or root_report is not None):
        raise ValueError("frames_only needs root_report = None and truncate = False")
    K.check_root_method(root_method)
    if root_maxiter < 1:
        raise ValueError("root_maxiter must be positive")
    if closure_probe < 0:
        raise ValueError(f"closure_probe must be a non-negative count, got {closure_probe}")
    eshift = float(eshift)

    def keep_room(b, width, tol):
        return wcap[b] - width if cap_keep and tol > 0 else 1 << 60
    rules = ("maxdim", "bond", "bond_strict", "frame_root")
    if work_ceiling_rule not in rules:
        raise ValueError(f"work_ceiling_rule must be one of {rules}, got {work_ceiling_rule!r}")
    factor = bond_ceiling_factor
    if factor is None:
        factor = DEFAULT_BOND_CEILING_FACTOR if work_ceiling_rule == "bond" else 0.0
    if not (factor == 0 or (factor > 1 and work_ceiling_rule != "maxdim")):
        raise ValueError("bond_ceiling_factor must be 0, or > 1 with a per-bond work_ceiling_rule")
    if not root_ceiling_factor > 1:
        raise ValueError("root_ceiling_factor must exceed one")
    if work_ceiling_rule == "frame_root" and root_ceiling_factor > (2.0 if factor == 0 else factor):
        raise ValueError("root_ceiling_factor must not exceed the frame factor")
    if probe_hook is not None and parallel:
        raise ValueError("probe_hook observes one probe sequence; it is not taken with parallel")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    cp = _lib.PREC_FP32 if precision == "mixed" else _lib.PREC_FP64
    htol = 1e-5 if precision == "mixed" else 1e-9
    ex = expansion or DDQRExpansion()
    maxdim = 128 if maxdim is None else maxdim
    tau = complex(tau)
    d = psi.d
    t0 = time.perf_counter()
    stream = rng if isinstance(rng, ProbeStream) else ProbeStream(0x5EED if rng is None else rng)
    bd0 = psi.bond_dims()
    if work_ceiling_rule == "bond":
        wcap = bond_work_ceilings(bd0, d, maxdim, factor)
    elif work_ceiling_rule == "bond_strict":
        wcap = bond_work_ceilings_strict(bd0, maxdim, factor)
    elif work_ceiling_rule == "frame_root":
        wcap = bond_frame_root_ceilings(bd0, maxdim, rl, 2.0 if factor == 0 else factor,
                                        root_ceiling_factor)
    else:
        wcap = [max(d * maxdim, d * max(bd0 + [0]), 1)] * (L - 1)
    nrm = T.norm(psi[psi.center]) if 0 <= psi.center < L else 1.0
    if not math.isfinite(nrm):
        raise FloatingPointError("cbe_bug_step: the state is not finite")
    alloc = basis_allocation(tb, L - 1, tau, nrm, order, tol_krylov, krylov_rank_tol)
    ledger = AdaptiveLedger(r_old=max(bd0 + [0]))
    recurrence = ex.krylov_mode if ex.krylov_mode != "auto" else ("lanczos" if hermitian else "arnoldi")
    expanded = [0] * (L - 1)
    left_frames: List[Optional[BlockTensor]] = [None] * (rl + 1)
    right_frames: List[Optional[BlockTensor]] = [None] * (L - 1)
    secs = {}

    # ---- the two half-sweeps (half_sweeps.jl:218-410) ----
    # The left sweep works on a copy, so `psi` stays left of the root for the assembly; the right
    # sweep works on `psi` itself.  Under `parallel` they run concurrently, the right one on a side
    # thread with its own library context and HIP stream, its own probe stream split off the
    # step's, and each with its own ledger, merged after the join (bug_step.jl:222-268).
    psiL = psi.copy()

    def run_left(stream, ledger):
        ts = time.perf_counter()
        canonical(psiL, 0)
        renvs: List[Optional[BlockTensor]] = [None] * (L + 1)
        renvs[L] = ev.boundary_right(mpo, psiL.bond_leg(L))
        for i in range(L - 1, 1, -1):
            renvs[i] = ev.push_right(renvs[i + 1], psiL[i], mpo[i], precision=cp)
        lch = ev.boundary_left(mpo, psiL.bond_leg(0))
        lroot = lch
        carry = None
        kcar = [None] * len(keep_states)
        for b in range(rl + 1):
            if b % 16 == 0:
                _progress(f"left half-sweep bond {b}/{rl}")
            carried = psiL[b] if carry is None else T.core_times_site(carry, psiL[b])
            fr = frame_from(carried, psiL[b + 1])
            bl = AdaptiveLedger()
            xp = dd_cbe_expand(fr, mpo, b, lch, renvs[b + 2], "left", alloc, bl, wcap[b], ex,
                               stream, probe_hook, cp, exact)
            Kex = T.site_times_core(carried, T.overlap_right(psiL[b + 1], xp.V_ex))
            rch1 = ev.push_right(renvs[b + 2], xp.V_ex, mpo[b + 1], precision=cp)
            renvs[b + 2] = None
            Hl, Hw, Hr = lch, mpo[b], rch1
            Wk, kst = krylov_enrich(Kex, lambda v: ev.apply_one_site(Hl, Hw, Hr, v, shift=eshift,
                                                                     precision=cp),
                                    xp.U_ex, "left", tau, alloc, wcap[b], recurrence,
                                    hermitian_tol=htol)
            _record_krylov(bl, kst)
            # every kept state unioned before any of their carries is taken
            kxs = [ks[b] if kc is None else T.core_times_site(kc, ks[b])
                   for ks, kc in zip(keep_states, kcar)]
            for kx, tol in zip(kxs, keep_tols):
                Wk = union_frame(Wk, kx, "left", tol, keep_room(b, Wk.layout.col.total, tol))
            kcar = [T.overlap_left(Wk, kx) for kx in kxs]
            left_frames[b] = Wk
            expanded[b] = Wk.layout.col.total
            log_bond(ledger, bl, b, "left", expanded[b])
            carry = T.overlap_left(Wk, carried)
            if b == rl:
                lroot = lch
            else:
                lch = ev.push_left(lch, Wk, mpo[b], precision=cp)
        secs["left_sweep"] = time.perf_counter() - ts
        return lroot

    def run_right(stream, ledger):
        ts = time.perf_counter()
        psiR = psi
        canonical(psiR, L - 1)
        lenvs = [ev.boundary_left(mpo, psiR.bond_leg(0))]
        for i in range(L - 2):
            lenvs.append(ev.push_left(lenvs[-1], psiR[i], mpo[i], precision=cp))
        rch = ev.boundary_right(mpo, psiR.bond_leg(L))
        rroot = rch
        carry = None
        kcar = [None] * len(keep_states)
        for b in range(L - 2, rl - 1, -1):
            if b % 16 == 0:
                _progress(f"right half-sweep bond {b}/{rl}")
            carried = psiR[b + 1] if carry is None else T.site_times_core(psiR[b + 1], carry)
            fr = frame_from(psiR[b], carried)
            bl = AdaptiveLedger()
            xp = dd_cbe_expand(fr, mpo, b, lenvs[b], rch, "right", alloc, bl, wcap[b], ex,
                               stream, probe_hook, cp, exact)
            Bex = T.core_times_site(T.overlap_left(xp.U_ex, psiR[b]), carried)
            lch1 = ev.push_left(lenvs[b], xp.U_ex, mpo[b], precision=cp)
            lenvs[b] = None
            Hl, Hw, Hr = lch1, mpo[b + 1], rch
            Zk, kst = krylov_enrich(Bex, lambda v: ev.apply_one_site(Hl, Hw, Hr, v, shift=eshift,
                                                                     precision=cp),
                                    xp.V_ex, "right", tau, alloc, wcap[b], recurrence,
                                    hermitian_tol=htol)
            _record_krylov(bl, kst)
            bxs = [ks[b + 1] if kc is None else T.site_times_core(ks[b + 1], kc)
                   for ks, kc in zip(keep_states, kcar)]
            for bx, tol in zip(bxs, keep_tols):
                Zk = union_frame(Zk, bx, "right", tol, keep_room(b, Zk.layout.row.total, tol))
            kcar = [T.overlap_right(bx, Zk) for bx in bxs]
            log_bond(ledger, bl, b, "right", Zk.layout.row.total)
            right_frames[b] = Zk
            if b != rl:
                expanded[b] = Zk.layout.row.total
            carry = T.overlap_right(carried, Zk)
            if b == rl:
                rroot = rch
            else:
                rch = ev.push_right(rch, Zk, mpo[b + 1], precision=cp)
        secs["right_sweep"] = time.perf_counter() - ts
        return rroot, carry

    if parallel:
        streamR = stream.split()
        ledgerL, ledgerR = AdaptiveLedger(), AdaptiveLedger()
        keep = list(psi.tensors)        # read by the side stream until the join
        lroot, (rroot, AR) = _run_beside(lambda: run_left(stream, ledgerL),
                                         lambda: run_right(streamR, ledgerR))
        del keep
        ledger.merge(ledgerL)
        ledger.merge(ledgerR)
    else:
        lroot = run_left(stream, ledger)
        rroot, AR = run_right(stream, ledger)

    _progress("root start and projection miss")
    # ---- root start and projection miss ----
    AL = None
    for i in range(rl + 1):
        c = psi[i] if AL is None else T.core_times_site(AL, psi[i])
        AL = T.overlap_left(left_frames[i], c)
    root_start = T.core_times_core(AL, AR)

    # ---- order p > 1: close the root frontiers j = 1..p-1 (bug_step_stages.jl:172-240) ----
    from .closure import KLFrames, _Ctx
    cx = _Ctx(mpo, tau, alloc, wcap, ex, stream, recurrence, cp, htol, fold, exact, parallel,
              eshift)
    if order > 1:
        from .closure import checked_closure, single_pass_closure
        ts = time.perf_counter()
        _progress("frontier closure")
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=eshift)
        close = checked_closure if closure_checks else single_pass_closure
        root_start, _ = close(fr, psi, root_start, cx, nrm, ledger, order)
        lroot, rroot = fr.lenv, fr.renv
        _refresh_expanded(expanded, left_frames, right_frames, rl)
        secs["closure"] = time.perf_counter() - ts
    elif closure_probe > 0:
        # measure only, on a copy of the probe stream: the step itself is unchanged
        import copy
        from .closure import measure_frontiers
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=eshift)
        *_, plg, _ = measure_frontiers(fr, psi, root_start, cx.with_stream(copy.copy(stream)),
                                       nrm, closure_probe)
        ledger.merge(plg)
        ledger.leak_initial = list(ledger.leak_initial) + list(plg.leak_initial)
        ledger.rho_initial = list(ledger.rho_initial) + list(plg.rho_initial)
        ledger.n_root_frontier += plg.n_root_frontier
    npsi = T.norm(psi[psi.center])
    ledger.r_aug = max(max(expanded + [0]), right_frames[rl].layout.row.total)
    ledger.proj_err = projection_miss(psi, left_frames, right_frames, rl) / npsi if npsi > 0 else 0.0
    if ledger.proj_err > tb:
        ledger.fail("projection")

    # ---- root solve ----
    ts = time.perf_counter()
    Lr = ev.push_left(lroot, left_frames[rl], mpo[rl], precision=cp)
    Rr = ev.push_right(rroot, right_frames[rl], mpo[rl + 1], precision=cp)
    napply = [0]
    _progress("root solve")

    def substep(n):
        if n % 50 == 0:
            _progress(f"root solve substep {n}")

    def solve_with(Lr, Rr, start):
        def apply_root(C):
            napply[0] += 1
            return ev.apply_zero_site(Lr, Rr, C, shift=eshift, precision=cp)
        rk = dict(root_method=root_method, rational_shift=root_shift)
        if hermitian:
            return K.lanczos_expv(apply_root, tau, start, tol=1e-15, maxiter=root_maxiter,
                                  reorth=reorth, conv_tol=te, substeps=0, on_substep=substep,
                                  **rk)
        return K.arnoldi_expv(apply_root, tau, start, tol=1e-15, maxiter=root_maxiter,
                              conv_tol=te, substeps=0, on_substep=substep, **rk)

    if frames_only:                      # the caller solves its own root (bug_step.jl:381-382)
        result = root_start
    else:
        result, rep = solve_with(Lr, Rr, root_start)
    if root_report is not None:
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=eshift)
        root_start, result, rep, report = report_root(
            psi, fr, Lr, Rr, root_start, result, rep, solve_with, cx, nrm, ledger, expanded,
            outside_residual, tb, order)
        root_report.clear()
        root_report.update(report)
    if not frames_only:
        ledger.expv_err, ledger.expv_iters = rep.err, rep.iterations
        if rep.err > te * (1 + 1e-8):
            ledger.fail("expv")
    if frames_report is not None:        # bug_step.jl:455-467
        frames_report.update(left=list(left_frames), right=list(right_frames), root_link=rl,
                             lenv=lroot, renv=rroot, eshift=float(eshift))
    secs["root"] = time.perf_counter() - ts

    _progress("assembly and compression")
    # ---- assembly and compression ----
    ts = time.perf_counter()
    res = T.svd_core(result, cutoff=0.0)
    expanded[rl] = left_frames[rl].layout.col.total
    discarded = T.trunc_weight(res)
    for i in range(rl):
        psi[i] = left_frames[i]
    psi[rl] = T.site_times_core(left_frames[rl], res.U)
    psi[rl + 1] = T.core_times_site(T.diag_scale(res.Vh, res.S, "rows"), right_frames[rl])
    for b in range(rl + 1, L - 1):
        psi[b + 1] = right_frames[b]
    psi.center = rl + 1
    root_rank = psi[rl].layout.col.total
    bwts = [0.0] * (L - 1)
    if truncate:
        wtr, cumw, capped = truncate_recursive(psi, rl + 1, maxdim, tt, bwts)
        discarded = max(discarded, wtr)
        ledger.discarded, ledger.maxdim_reached = cumw, capped
        if cumw > tt * (1 + 1e-8):
            ledger.fail("trunc")
    bd1 = psi.bond_dims()
    ledger.r_final = max(bd1 + [0])
    # close the bond log with what exists only after the compression (diagnostics.jl:328-335)
    for e in ledger.bond_log:
        e.update(r_old=bd0[e["bond"]], r_final=bd1[e["bond"]], discarded=bwts[e["bond"]])
    secs["truncation"] = time.perf_counter() - ts
    secs["step"] = time.perf_counter() - t0
    return CBEBugSweepInfo(ledger, expanded, rl, root_rank, discarded, napply[0], secs)


def _refresh_expanded(expanded, left_frames, right_frames, rl):
    """``_refresh_expanded!``: the frame widths off the root bond."""
    for b in range(rl):
        expanded[b] = left_frames[b].layout.col.total
    for b in range(rl + 1, len(expanded)):
        expanded[b] = right_frames[b].layout.row.total


def report_root(psi: SymMPS, fr, Lr, Rr, root_start, result, rep, solve_with, cx, nrm: float,
                ledger: AdaptiveLedger, expanded, outside_residual: bool, tol_basis: float,
                order: int):
    """``_report_root!`` (bug_step_stages.jl:242-398): the energies before and after the root
    solve (Rayleigh quotients of the shifted root generator plus eshift: exact energies, since the
    state lies in the frames' span) and, with ``outside_residual``, the outside-Galerkin defect
    |tau| ||(I - P) H psi_1|| / ||psi_1|| of the evolved root vector.  While the defect exceeds
    tol_basis the frames are re-enriched by the evolved vector's own frontier sweep and the root
    is re-solved (at most 2p+2 passes; a pass that adds nothing or binds the work ceiling ends the
    loop).  Returns the possibly re-solved root start, result and solve report, and the report."""
    from .closure import closure_maxpass, complement_action, frontier_sweep, project_root
    es, tau, prec = fr.eshift, cx.tau, cx.prec
    eps = np.finfo(float).eps

    def rayleigh(S):
        n2 = T.norm(S) ** 2
        if not n2 > 0:
            return float("nan")
        return float(T.inner(S, ev.apply_zero_site(Lr, Rr, S, shift=es, precision=prec)).real) \
            / n2 + es

    e_pre = rayleigh(root_start)
    outside = complement_action(psi, fr, result, cx) if outside_residual else None

    def resid(act, S):
        return float("nan") if act is None else abs(tau) * act["bound"] / max(T.norm(S), eps)

    initial = resid(outside, result)
    passes = added = 0
    if outside_residual:
        while not (resid(outside, result) <= tol_basis) and passes < closure_maxpass(order):
            passes += 1
            sr = frontier_sweep(fr, psi, result, cx, nrm, ledger)
            if sr.added == 0:
                break
            added += sr.added
            root_start = project_root(psi, fr)
            Lr = ev.push_left(fr.lenv, fr.left[fr.rl], cx.mpo[fr.rl], precision=prec)
            Rr = ev.push_right(fr.renv, fr.right[fr.rl], cx.mpo[fr.rl + 1], precision=prec)
            result, rep = solve_with(Lr, Rr, root_start)
            outside = complement_action(psi, fr, result, cx)
            if sr.capped:
                break
        if added > 0:
            _refresh_expanded(expanded, fr.left, fr.right, fr.rl)
            ledger.r_aug = max(max(expanded + [0]), fr.right[fr.rl].layout.row.total)
    e_root = rayleigh(result)
    rn = max(T.norm(result), eps)
    nan = float("nan")
    report = dict(
        eshift=es, e_pre=e_pre, e_root=e_root, resid_initial=initial, outside_passes=passes,
        outside_added=added,
        outside=nan if outside is None else outside["bound"] / rn,
        outside_left=nan if outside is None else outside["left"] / rn,
        outside_right=nan if outside is None else outside["right"] / rn,
        per_left=[] if outside is None else [x / rn for x in outside["per_left"]],
        per_right=[] if outside is None else [x / rn for x in outside["per_right"]],
        resid=nan if outside is None else abs(tau) * outside["bound"] / rn)
    return root_start, result, rep, report


def projection_miss(psi: SymMPS, left_frames, right_frames, rl: int) -> float:
    """``_projection_miss`` (krylov_closure_frames.jl:242-280), square-root factors by QR."""
    L = len(psi)
    rf: List[Optional[BlockTensor]] = [None] * (L + 1)
    for i in range(L - 1, 0, -1):
        carried = psi[L - 1] if i == L - 1 else T.site_times_core(psi[i], rf[i + 1])
        rf[i], _ = T.qr_right(carried)
    miss2 = 0.0
    AL = None
    for i in range(rl + 1):
        carried = psi[i] if AL is None else T.core_times_site(AL, psi[i])
        miss2 += T.norm(T.site_times_core(T.perp_left(left_frames[i], carried), rf[i + 1])) ** 2
        AL = T.overlap_left(left_frames[i], carried)
    lf: List[Optional[BlockTensor]] = [None] * (L - 1)
    _, lf[rl] = T.qr_core(AL)
    for b in range(rl + 1, L - 1):
        _, lf[b] = T.qr_left(T.core_times_site(lf[b - 1], psi[b]))
    AR = None
    for b in range(L - 2, rl - 1, -1):
        carried = psi[b + 1] if AR is None else T.site_times_core(psi[b + 1], AR)
        miss2 += T.norm(T.core_times_site(lf[b], T.perp_right(right_frames[b], carried))) ** 2
        AR = T.overlap_right(carried, right_frames[b])
    return math.sqrt(miss2)


def truncate_recursive(psi: SymMPS, root: int, maxdim: int, weight_tol: float,
                       bond_weights: Optional[List[float]] = None):
    """Budgeted root-to-leaves compression (``truncate_recursive!``, truncation.jl:59-158);
    ``root`` is the 0-based root site.  Returns (worst, cumulative, capped); ``bond_weights``
    (L - 1 entries) receives each bond's discarded norm relative to the state's."""
    if bond_weights is not None and len(bond_weights) != len(psi) - 1:
        raise ValueError("bond_weights needs one entry per bond")
    L = len(psi)
    sdiv = math.sqrt(root) + math.sqrt(L - 1 - root)
    st = {"nrm2": -1.0, "worst": 0.0, "capped": False}
    cum = [0.0, 0.0]

    def cut(t, left_split):
        fac = T.factor_left(t) if left_split else T.factor_right(t)
        res = fac.select(0.0, maxdim)
        tot = T.sq_weight(res.kept) + T.sq_weight(res.trunc)
        if st["nrm2"] < 0:
            st["nrm2"] = tot
        share = weight_tol * math.sqrt(max(st["nrm2"], 0.0)) / sdiv if sdiv > 0 else 0.0
        nk = len(res.kept)
        keep, tail = K.tail_rank(res.kept, 1.0, share, nk, base=T.sq_weight(res.trunc))
        if tail > share and res.trunc:
            st["capped"] = True
        keep = max(keep, 1)
        if keep < nk:
            res = fac.select(0.0, keep)
        return res

    def account(res, side, bond):
        st["worst"] = max(st["worst"], T.trunc_weight(res))
        cum[side] += T.sq_weight(res.trunc)
        if bond_weights is not None:
            bond_weights[bond] = T.sq_weight(res.trunc)

    for site in range(root, 0, -1):
        res = cut(psi[site], False)
        account(res, 0, site - 1)
        psi[site] = res.Vh
        psi[site - 1] = T.site_times_core(psi[site - 1], T.diag_scale(res.U, res.S, "cols"))
        psi.center = site - 1
    canonical(psi, root)
    for site in range(root, L - 1):
        res = cut(psi[site], True)
        account(res, 1, site)
        psi[site] = res.U
        psi[site + 1] = T.core_times_site(T.diag_scale(res.Vh, res.S, "rows"), psi[site + 1])
        psi.center = site + 1
    n2 = st["nrm2"]
    cumulative = (math.sqrt(cum[0]) + math.sqrt(cum[1])) / math.sqrt(n2) if n2 > 0 else 0.0
    if bond_weights is not None:
        for i, w in enumerate(bond_weights):
            bond_weights[i] = math.sqrt(w / n2) if n2 > 0 else 0.0
    return st["worst"], cumulative, st["capped"]
