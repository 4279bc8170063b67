"""ORACLE (test infrastructure): the CBE-BUG step ``cbe_bug_step!`` restated on ``Fam`` tensors.

Follows, stage by stage (0-based sites; bond b joins sites b and b+1; link j sits left of site j):

* options and budget       bug_step_stages.jl:15-142, basis_budget.jl:75-128,454-494,568
* the two half-sweeps       src/BUG/Core/Sweeps/half_sweeps.jl:218-410
* bond frames               effective_operators.jl:201-220 (gauge by unique QR, see linalg.py)
* DD-QR CBE ``:rank_tail``  src/BUG/Randomised/dd_cbe_expand.jl:19-87,252-301,383-419,
                            dd_cbe.jl:38-89,132-161, cbe.jl:188-236, sketch.jl:141-180,
                            the streamed sketch mpo_environments.jl:409-481
* local Krylov, Tests C/D   krylov_enrichment.jl:58-292, krylov_recurrence.jl, krylov_lanczos.jl,
                            krylov_arnoldi.jl
* carries, root start       half_sweeps.jl:300-308,400-408, krylov_closure_frames.jl:207-233
* projection miss           krylov_closure_frames.jl:242-280
* root generator and expv   effective_operators.jl:91-107, krylov.jl:128-263, root_methods.jl
* assembly + compression    diagnostics.jl:269-337, truncation.jl:49-158
* work ceilings             basis_budget.jl:454-549 (:bond, :maxdim, :bond_strict, :frame_root)
* ranked Krylov selection   krylov_enrichment.jl:174-211 (krylov_rank_tol)
* exact DD reference        dd_cbe.jl:141-160, dd_cbe_expand.jl:91-97,346-360, sketch.jl:119-132
* kept states               half_sweeps.jl:117-207,280-296,384-395 (keep_states, keep_tols)
* per-bond log              basis_budget.jl:390-418, diagnostics.jl:328-335
* frames only               bug_step.jl:114-116,381-382

The probe stream is this package's counter-based generator, not Julia's MersenneTwister, and the
probe can be injected; the order-p frontier closure (``order > 1``) is not restated here.

Pinned by tests/test_oracle_bug.py against the exact free-fermion magnetisation of the XX chain
(test/common/free_fermion.jl:25-47), against 2^L exact diagonalisation of the XXZ chain, and
against the reference's frozen runs (test/fixtures/b155550/xx_u1_bug_order1.json) at the
tolerances of test_fixtures_b155550.jl:61-64.  The option paths are pinned there too: the
ceilings on hand-worked values, the ranked selection on a dense SVD of the candidates, the exact
probe on the complement projector, the kept states on the reference's own width checks
(test_step_keep_states.jl:27-57), the bond log on its aggregation into the step's ledger.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import blocks as ob
from . import krylov as ok
from . import linalg as la
from .blocks import Fam
from .linalg import CBE_EPS

TOL_DEFAULTS = dict(basis=1e-5, trunc=1e-7, expv=1e-10)
ADAPTIVE_GROWTH = 1.2
DEFAULT_BOND_CEILING_FACTOR = 1.2


# ---- the state --------------------------------------------------------------------------

class MPS:
    def __init__(self, sites: List[Fam], center: int = 0):
        self.sites = list(sites)
        self.center = center

    def __len__(self):
        return len(self.sites)

    def copy(self):
        return MPS(list(self.sites), self.center)

    def bond_dims(self):
        return [sum(self.sites[b].col.values()) for b in range(len(self.sites) - 1)]


def canonical(psi: MPS, site: int):
    """canonical! (canonical_form.jl:43-54) with QR moves."""
    while psi.center < site:
        i = psi.center
        Q, R = la.qr_left(psi.sites[i])
        psi.sites[i] = Q
        psi.sites[i + 1] = la.core_times_site(R, psi.sites[i + 1])
        psi.center += 1
    while psi.center > site:
        i = psi.center
        Lc, Q = la.qr_right(psi.sites[i])
        psi.sites[i] = Q
        psi.sites[i - 1] = la.site_times_core(psi.sites[i - 1], Lc)
        psi.center -= 1
    return psi


def norm_center(psi: MPS) -> float:
    return psi.sites[psi.center].norm()


def sz_profile(psi: MPS, charges=(1, -1)) -> np.ndarray:
    """<S^z_j>/<psi|psi> for spin-1/2 sites (local_observables.jl:11-31), any gauge.

    Local state 0 is up (S^z = +1/2), state 1 down."""
    L = len(psi)
    sz = [0.5, -0.5]
    q0 = list(psi.sites[0].row)[0]
    left = [Fam({q0: 1}, {q0: 1}, [0], {(0, q0): np.ones((1, 1), dtype=complex)})]
    for i in range(L - 1):
        left.append(_transfer(left[i], psi.sites[i], None))
    qL = list(psi.sites[-1].col)[0]
    right = [None] * (L + 1)
    right[L] = Fam({qL: 1}, {qL: 1}, [0], {(0, qL): np.ones((1, 1), dtype=complex)})
    for i in range(L - 1, 0, -1):
        right[i] = _transfer_right(right[i + 1], psi.sites[i])
    out = np.zeros(L)
    nrm = None
    for i in range(L):
        E = _transfer(left[i], psi.sites[i], sz)
        N = _transfer(left[i], psi.sites[i], None)
        val = sum(np.trace(E.blocks[k] @ right[i + 1].blocks[k]) for k in E.blocks
                  if k in right[i + 1].blocks)
        n = sum(np.trace(N.blocks[k] @ right[i + 1].blocks[k]) for k in N.blocks
                if k in right[i + 1].blocks)
        out[i] = val.real
        nrm = n.real if nrm is None else nrm
    return out / nrm


def _transfer(E: Fam, A: Fam, weights) -> Fam:
    """sum_s w_s A_s^dag E A_s (bra = ket = A), E on the left link of A."""
    out = Fam.zeros(A.col, A.col, [0])
    for (s, ql), blk in A.blocks.items():
        e = E.blocks.get((0, ql))
        if e is None:
            continue
        w = 1.0 if weights is None else weights[s]
        if w == 0:
            continue
        qr = ql + A.shifts[s]
        out.blocks[(0, qr)] += w * (blk.conj().T @ e @ blk)
    return out


def _transfer_right(E: Fam, A: Fam) -> Fam:
    """sum_s A_s E A_s^dag, E on the right link of A; pairs with the left one by a trace."""
    out = Fam.zeros(A.row, A.row, [0])
    for (s, ql), blk in A.blocks.items():
        qr = ql + A.shifts[s]
        e = E.blocks.get((0, qr))
        if e is None:
            continue
        out.blocks[(0, ql)] += blk @ e @ blk.conj().T
    return out


# ---- options ------------------------------------------------------------------------------

@dataclass
class StepOptions:
    tol_basis: float = TOL_DEFAULTS["basis"]
    tol_trunc: float = TOL_DEFAULTS["trunc"]
    tol_expv: float = TOL_DEFAULTS["expv"]
    tol_krylov: Optional[float] = None
    krylov_rank_tol: float = 0.0        # > 0: ranked Krylov selection (bug_step.jl:58)
    exact: bool = False                 # deterministic DD reference (bug_step.jl:117-119)
    fold_omega: Optional[bool] = None   # None: order > 1 and not exact (bug_step_stages.jl:67)
    keep_states: list = field(default_factory=list)   # MPS the trial space must contain
    keep_tols: list = field(default_factory=list)     # per kept state, default 0 (exact)
    frames_only: bool = False           # skip the root solve (bug_step.jl:114-116,381-382)
    order: int = 1
    maxdim: int = 128
    hermitian: bool = True
    work_ceiling_rule: str = "bond"
    bond_ceiling_factor: Optional[float] = None
    root_ceiling_factor: float = 1.2
    root: Optional[int] = None          # 0-based bond of the root (reference root - 1)
    reorth: bool = True
    truncate: bool = True
    mode: str = "rank_tail"             # "rank_tail" (default) or "adaptive"
    fixed_width: int = 8
    holdout_width: int = 4
    tail_factor: float = 0.10
    krylov_mode: str = "auto"
    seed: int = 0x5EED
    parallel: bool = False
    eshift: float = 0.0                 # the step integrates H - eshift (bug_step.jl:88-91)
    closure_checks: bool = False        # measured closure loop (krylov_closure_loop.jl:18-194)
    closure_probe: int = 0              # order 1: measure frontiers 0..m-1 (bug_step.jl:33-38)
    root_report: bool = False           # the root's account (bug_step_stages.jl:242-398)
    outside_residual: bool = False      # with root_report: measure and re-enrich (I - P) H psi_1
    root_method: str = "exp"            # exp | trapezoid | radau2 | radau3 (bug_step.jl:76-81)
    root_shift: bool = True             # rational roots: shift to the extreme Ritz value (:82-86)
    root_maxiter: int = 30              # root Krylov basis cap (_ROOT_MAXITER; bug_step.jl:87)


@dataclass
class Alloc:
    per_bond: float
    sample: float
    tail: float
    krylov: float
    scale: float
    krylov_rank: float = 0.0            # > 0: ranked Krylov selection (basis_budget.jl:37-48)


@dataclass
class Ledger:
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
    leak: List[float] = field(default_factory=list)
    defect: float = 0.0
    leak_initial: List[float] = field(default_factory=list)
    rho_initial: List[float] = field(default_factory=list)
    n_passes: int = 0
    n_root_frontier: int = 0
    unmet: List[str] = field(default_factory=list)
    bond_log: List[dict] = field(default_factory=list)   # half-sweep bond visits, 0-based bonds

    def unmet_(self, what):
        self.tolerance_met = False
        if what not in self.unmet:
            self.unmet.append(what)

    @property
    def status(self):
        return "met" if self.tolerance_met else "budget_unmet"


def basis_allocation(tol_basis, nbonds, tau, nrm, nfront, tol_krylov=None,
                     krylov_rank=0.0) -> Alloc:
    """:linear rule of basis_allocation (basis_budget.jl:75-104); ``krylov_rank`` (tol_krylov's
    units) is split over the bonds the same way; it bounds what a ranked Krylov stage may drop."""
    nb, nf = max(nbonds, 1), max(nfront, 1)
    kr = 1.0 if tol_krylov is None else tol_krylov / tol_basis
    if not krylov_rank >= 0:
        raise ValueError("krylov_rank_tol must be >= 0")
    at = abs(tau)
    scale = 1.0 if at == 0 else at / max(nrm, np.finfo(float).eps)
    per_test = tol_basis / (2 * (nb + 1) * nf)
    return Alloc(2 * per_test, per_test, per_test, kr * per_test, scale,
                 krylov_rank / tol_basis * per_test)


def bond_work_ceilings(ranks, d, maxdim, factor):
    def ceil_of(r):
        return math.ceil(round(factor * d * r, 9)) if factor > 0 else 2 * d * r
    return [max(min(ceil_of(r), maxdim), 1) for r in ranks]


def bond_work_ceilings_strict(ranks, maxdim, factor):
    """bond_work_ceilings_strict (basis_budget.jl:505-512): min(ceil(f D_b), maxdim) per bond,
    no local-dimension factor."""
    if not factor > 1:
        raise ValueError("strict bond factor must exceed 1")
    return [max(1, min(maxdim, math.ceil(round(factor * r, 9)))) for r in ranks]


def bond_frame_root_ceilings(ranks, maxdim, root_bond, frame_factor=2.0, root_factor=1.2):
    """bond_frame_root_ceilings (basis_budget.jl:526-549): ceil(frame_factor D_b) away from the
    root bond, ceil(root_factor D_root) at it, each within [1, maxdim]; root_bond is 0-based."""
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
    return [max(1, min(maxdim, math.ceil(round((root_factor if b == root_bond else frame_factor) * r,
                                                9))))
            for b, r in enumerate(ranks)]


def work_ceilings(rule, bd0, d, maxdim, factor, root_bond, root_factor):
    """The step's work ceiling from the bonds before the step (bug_step_stages.jl:63-85,116-129):
    one number for every bond (rule "maxdim") or one per bond ("bond", "bond_strict",
    "frame_root")."""
    if rule not in ("maxdim", "bond", "bond_strict", "frame_root"):
        raise ValueError(f"work_ceiling_rule must be maxdim, bond, bond_strict or frame_root, "
                         f"got {rule!r}")
    if factor is None:
        factor = DEFAULT_BOND_CEILING_FACTOR if rule == "bond" else 0.0
    if not (factor == 0 or (factor > 1 and rule in ("bond", "bond_strict", "frame_root"))):
        raise ValueError("bond_ceiling_factor must be 0, or > 1 with a per-bond rule")
    if not root_factor > 1:
        raise ValueError("root_ceiling_factor must exceed one")
    if rule == "frame_root" and root_factor > (2.0 if factor == 0 else factor):
        raise ValueError("root_ceiling_factor must not exceed the frame factor")
    if rule == "bond":
        return bond_work_ceilings(bd0, d, maxdim, factor)
    if rule == "bond_strict":
        return bond_work_ceilings_strict(bd0, maxdim, factor)
    if rule == "frame_root":
        return bond_frame_root_ceilings(bd0, maxdim, root_bond, 2.0 if factor == 0 else factor,
                                        root_factor)
    return max(d * maxdim, d * max(bd0 + [0]), 1)


# ---- environments -------------------------------------------------------------------------

def left_boundary(mpo, psi: MPS) -> Fam:
    q0 = list(psi.sites[0].row)[0]
    E = ob.env_fam({q0: 1}, {q0: 1}, mpo[0].cl)
    E.blocks[(0, q0)] = np.ones((1, 1), dtype=complex)
    return E


def right_boundary(mpo, psi: MPS) -> Fam:
    qL = list(psi.sites[-1].col)[0]
    E = ob.env_fam({qL: 1}, {qL: 1}, mpo[-1].cr)
    E.blocks[(0, qL)] = np.ones((1, 1), dtype=complex)
    return E


# ---- probes and sketches ------------------------------------------------------------------

def fused_dims(frame: Fam, side: str) -> Dict[int, int]:
    """Dimensions of the fused local space of a frame per charge (its (link, site) legs)."""
    out: Dict[int, int] = {}
    if side == "left":        # (l, s): charge of l + q(s)
        for ql, d in frame.row.items():
            for sh in frame.shifts:
                out[ql + sh] = out.get(ql + sh, 0) + d
    else:                     # (s, r): charge of r - q(s)
        for qr, d in frame.col.items():
            for sh in frame.shifts:
                out[qr - sh] = out.get(qr - sh, 0) + d
    return dict(sorted(out.items()))


def frame_widths(frame: Fam, side: str) -> Dict[int, int]:
    return dict(frame.col) if side == "left" else dict(frame.row)


def alloc_columns(dims: Dict[int, int], budget: int) -> Dict[int, int]:
    """_alloc_columns, cbe.jl:188-195."""
    tot = sum(dims.values())
    if tot == 0 or budget <= 0:
        return {q: 0 for q in dims}
    return {q: min(d, math.ceil(d / tot * budget)) for q, d in dims.items()}


def _graded_draw(frame: Fam, side: str, dfull: Dict[int, int], cols: Dict[int, int],
                 rng: la.ProbeRNG) -> Fam:
    """Gaussian columns on the fused local space of ``frame`` (``_randomize_like`` then
    ``_trim_per_sector``, cbe.jl:156-178, truncation.jl:283-288).  Stream order: per fused charge
    ascending, per slice ascending, a (link dim x dfull) block drawn column-major of which the
    first ``cols[q]`` columns are kept; the layout is the frame's, the width leg in place of its
    bond leg."""
    wl = {q: c for q, c in cols.items() if c > 0}
    if side == "left":
        Om = Fam(frame.row, wl, frame.shifts)
    else:
        Om = Fam(wl, frame.col, frame.shifts)
    for q, df in dfull.items():
        c = cols.get(q, 0)
        for s, sh in enumerate(frame.shifts):
            lq = q - sh if side == "left" else q + sh
            dl = (frame.row if side == "left" else frame.col).get(lq, 0)
            if dl == 0:
                continue
            n_keep = dl * c
            vals = rng.complex_normal(n_keep)
            rng.skip(dl * df - n_keep)
            if c == 0:
                continue
            blk = vals.reshape(c, dl).T          # column-major (dl x c)
            if side == "left":
                Om.blocks[(s, lq)] = blk
            else:
                Om.blocks[(s, q)] = blk.T.copy()  # (c x dl)
    return Om


def dd_probe(frame: Fam, side: str, width: int, rng: la.ProbeRNG):
    """_dd_probe (dd_cbe.jl:132-161): sector-graded Gaussian columns in the discarded space.

    ``side`` is the side of ``frame`` (the opposite of the expanding frame); columns are shared
    out by the complement dimension of each sector and the frame is projected out afterwards.
    """
    dfull = fused_dims(frame, side)
    dT = frame_widths(frame, side)
    dC = {q: max(d - dT.get(q, 0), 0) for q, d in dfull.items()}
    if sum(dC.values()) == 0:
        return None
    Om = _graded_draw(frame, side, dfull, alloc_columns(dC, width), rng)
    proj = la.perp_left(frame, Om) if side == "left" else la.perp_right(frame, Om)
    return proj if proj.norm() > 0 else None


def graded_probe(frame: Fam, side: str, width: int, rng: la.ProbeRNG):
    """sector_graded_sketch with comp_ratio = nothing (sketch.jl:45-54): ``width`` columns shared
    out by the FULL fused dimension of each sector, no projector in the probe."""
    dfull = fused_dims(frame, side)
    cols = alloc_columns(dfull, width)
    if not any(c > 0 for c in cols.values()):
        return None
    Om = _graded_draw(frame, side, dfull, cols, rng)
    return Om if Om.norm() > 0 else None


def _width_of(T: Fam, side: str) -> int:
    return sum(T.col.values()) if side == "left" else sum(T.row.values())


def stream_left(L: Fam, W1, W2, R: Fam, A: Fam, V: Fam, Om: Fam) -> Fam:
    """Y = (H Theta) Om^dag for Theta = A V, Om in V's layout (g, s, r) -> (l, s, g)
    (_stream_left, mpo_environments.jl:409-444)."""
    d = len(A.shifts)
    # F[(c, s2)](r, g) = sum_{b, s2'} W2[c,b](s2', s2) R_b^T conj(Om_{s2'})^T
    slots, terms = {}, []
    for c, b, so, si, coef in W2.entries:
        k = slots.setdefault((c, si), len(slots))
        terms.append((k, b, so, la.OP_T, la.OP_C, coef))
    sh = {}
    for c, b, so, si, coef in W2.entries:
        k = slots[(c, si)]
        sh[k] = -R.shifts[b] - Om.shifts[so]
    F = la.product(R.col, Om.row, [sh[k] for k in range(len(slots))], R, Om, terms)
    # B_c(m, g) = sum_s2 V_{s2}(m, r) F[(c, s2)](r, g)
    cs = sorted({c for (c, _) in slots})
    cidx = {c: i for i, c in enumerate(cs)}
    tB, shB = [], {}
    for (c, si), k in slots.items():
        tB.append((cidx[c], si, k, la.OP_N, la.OP_N, 1.0))
        shB[cidx[c]] = V.shifts[si] + sh[k]
    B = la.product(V.row, Om.row, [shB[i] for i in range(len(cs))], V, F, tB)
    # C[(c, s1)](l, g) = A_{s1}(l, m) B_c(m, g)
    cslots, tC, shC = {}, [], {}
    for a, c, so, si, coef in W1.entries:
        if c not in cidx:
            continue
        k = cslots.setdefault((c, si), len(cslots))
        if k == len(tC):
            tC.append((k, si, cidx[c], la.OP_N, la.OP_N, 1.0))
            shC[k] = A.shifts[si] + shB[cidx[c]]
    C = la.product(A.row, Om.row, [shC[k] for k in range(len(cslots))], A, B, tC)
    # Y_{s1'}(l', g) = sum W1[a,c](s1', s1) L_a(l', l) C[(c, s1)](l, g)
    tY = []
    for a, c, so, si, coef in W1.entries:
        if (c, si) in cslots:
            tY.append((so, a, cslots[(c, si)], la.OP_N, la.OP_N, coef))
    return la.product(L.row, Om.row, A.shifts, L, C, tY)


def stream_right(L: Fam, W1, W2, R: Fam, U: Fam, Bv: Fam, Om: Fam) -> Fam:
    """Y = Om^dag (H Theta) for Theta = U Bv, Om in U's layout (l, s, g) -> (g, s, r)
    (_stream_right, mpo_environments.jl:446-481)."""
    # G/F[(c, s1)](l, g) = sum_{a, s1'} W1[a,c](s1', s1) L_a^T conj(Om_{s1'})
    slots, terms, sh = {}, [], {}
    for a, c, so, si, coef in W1.entries:
        k = slots.setdefault((c, si), len(slots))
        terms.append((k, a, so, la.OP_T, la.OP_R, coef))
        sh[k] = -L.shifts[a] + Om.shifts[so]
    F = la.product(L.col, Om.col, [sh[k] for k in range(len(slots))], L, Om, terms)
    # B_c(m, g) = sum_s1 U_{s1}^T(m, l) F[(c, s1)](l, g)
    cs = sorted({c for (c, _) in slots})
    cidx = {c: i for i, c in enumerate(cs)}
    tB, shB = [], {}
    for (c, si), k in slots.items():
        tB.append((cidx[c], si, k, la.OP_T, la.OP_N, 1.0))
        shB[cidx[c]] = -U.shifts[si] + sh[k]
    B = la.product(U.col, Om.col, [shB[i] for i in range(len(cs))], U, F, tB)
    # C[(c, s2)](g, r) = B_c^T(g, m) Bv_{s2}(m, r)
    cslots, tC, shC = {}, [], {}
    for c, b, so, si, coef in W2.entries:
        if c not in cidx:
            continue
        k = cslots.setdefault((c, si), len(cslots))
        if k == len(tC):
            tC.append((k, cidx[c], si, la.OP_T, la.OP_N, 1.0))
            shC[k] = -shB[cidx[c]] + Bv.shifts[si]
    C = la.product(Om.col, Bv.col, [shC[k] for k in range(len(cslots))], B, Bv, tC)
    # Y_{s2'}(g, r') = sum W2[c,b](s2', s2) C[(c, s2)](g, r) R_b^T(r, r')
    tY = []
    for c, b, so, si, coef in W2.entries:
        if (c, si) in cslots:
            tY.append((so, cslots[(c, si)], b, la.OP_N, la.OP_T, coef))
    return la.product(Om.col, R.row, Bv.shifts, C, R, tY)


# ---- frames -------------------------------------------------------------------------------

@dataclass
class BondFrame:
    U0: Fam
    S0: Fam
    V0: Fam
    old_rank: int


def frame_from(left: Fam, right: Fam) -> BondFrame:
    """_frame_from (effective_operators.jl:201-220)."""
    U0, Rl = la.qr_left(left)
    Lr, V0 = la.qr_right(right)
    S0 = la.product(Rl.row, Lr.col, [0], Rl, Lr, [(0, 0, 0, la.OP_N, la.OP_N, 1.0)])
    return BondFrame(U0, S0, V0, sum(U0.col.values()))


def reortho(frame: Fam, block: Optional[Fam], side: str) -> Optional[Fam]:
    """_reortho_left/_reortho_right (truncation.jl:374-394)."""
    if block is None:
        return None
    perp = la.perp_left if side == "left" else la.perp_right
    outside = perp(frame, block)
    if outside.norm() <= CBE_EPS * max(block.norm(), 1.0):
        return None
    if side == "left":
        r = la.svd_left(outside, cutoff=1e-14)
        return r.U if r.U.blocks else None
    r = la.svd_right(outside, cutoff=1e-14)
    return r.Vh if r.Vh.blocks else None


def cap_to_local_room(frame: Fam, cand: Optional[Fam], side: str) -> Optional[Fam]:
    """_cap_to_local_room (truncation.jl:332-351)."""
    if cand is None:
        return None
    dfull = fused_dims(frame, side)
    dT = frame_widths(frame, side)
    dU = frame_widths(cand, side)
    counts, over = {}, False
    for q, du in dU.items():
        if du == 0:
            continue
        room = dfull.get(q, 0) - dT.get(q, 0)
        counts[q] = min(du, room)
        over |= counts[q] < du
    if not over:
        return cand
    return trim_per_sector(cand, side, counts)


def trim_per_sector(T: Fam, side: str, counts: Dict[int, int]) -> Optional[Fam]:
    if not any(c > 0 for c in counts.values()):
        return None
    if side == "left":
        col = {q: c for q, c in counts.items() if c > 0}
        out = Fam(T.row, col, T.shifts)
        for (s, ql), b in T.blocks.items():
            qc = ql + T.shifts[s]
            if col.get(qc, 0) > 0:
                out.blocks[(s, ql)] = b[:, :col[qc]]
        return out
    row = {q: c for q, c in counts.items() if c > 0}
    out = Fam(row, T.col, T.shifts)
    for (s, ql), b in T.blocks.items():
        if row.get(ql, 0) > 0:
            out.blocks[(s, ql)] = b[:row[ql], :]
    return out


def trim_total(T: Fam, side: str, n: int) -> Optional[Fam]:
    """_trim_total (cbe.jl:219-236): the first sectors in (ascending charge) order."""
    if n <= 0:
        return None
    widths = frame_widths(T, side)
    if sum(widths.values()) <= n:
        return T
    left = n
    keep = {}
    for q in sorted(widths):
        keep[q] = min(widths[q], left)
        left -= keep[q]
    return trim_per_sector(T, side, keep)


def oplus(frame: Fam, new: Fam, side: str) -> Fam:
    return la.oplus_left(frame, new) if side == "left" else la.oplus_right(frame, new)


def sample_residual(basis, project, Y: Fam, Om: Fam, side: str) -> float:
    """_sample_residual (sketch.jl:141-180) with the sketch already in hand."""
    cols = dict(Om.row) if side == "left" else dict(Om.col)
    cols = {q: n for q, n in cols.items() if n > 0}
    if not cols:
        return 0.0
    residual = project(Y)
    if residual.norm() <= CBE_EPS:
        return 0.0
    if basis is not None:
        off = la.perp_left(basis, residual) if side == "left" else la.perp_right(basis, residual)
        residual = project(off)
    per = {}
    for (s, q), b in residual.blocks.items():
        w = q + residual.shifts[s] if side == "left" else q
        per[w] = per.get(w, 0.0) + float(np.vdot(b, b).real)
    tot = 0.0
    for w, v in per.items():
        n = cols.get(w, 0)
        if n <= 0:
            raise RuntimeError("Test B: residual sector without holdout column")
        tot += v / n
    return math.sqrt(tot)


# ---- one side of a bond expansion (dd_cbe_expand, :rank_tail) ------------------------------

@dataclass
class Expansion:
    U_ex: Fam
    V_ex: Fam
    n_new: int


def dd_cbe_expand(fr: BondFrame, mpo, bond: int, lenv: Fam, renv: Fam, side: str, alloc: Alloc,
                  ledger: Ledger, workcap: int, opts: StepOptions, rng, probe_hook=None) -> Expansion:
    left = side == "left"
    frame, opp = (fr.U0, fr.V0) if left else (fr.V0, fr.U0)
    oside = "right" if left else "left"
    old_rank = fr.old_rank
    none = Expansion(fr.U0, fr.V0, 0)
    full_self = sum(fused_dims(frame, side).values())
    full_opp = sum(fused_dims(opp, oside).values())
    cap = full_self - old_rank
    wbind = False
    if workcap > 0:
        wk = max(workcap - old_rank, 0)
        wbind = wk < cap
        cap = min(cap, wk)
    if cap <= 0:
        return none
    W1, W2 = mpo[bond], mpo[bond + 1]
    if opts.exact:
        return _dd_cbe_exact(fr, frame, opp, side, oside, lenv, renv, W1, W2, ledger, cap, wbind)
    if opts.mode == "adaptive":
        return _dd_cbe_adaptive(fr, frame, opp, side, oside, lenv, renv, W1, W2, alloc, ledger,
                                cap, wbind, full_opp, workcap, opts, rng)
    if opts.mode != "rank_tail":
        raise ValueError(f"unknown DD mode {opts.mode!r}")
    width = min(opts.fixed_width, full_opp)
    Om = probe_hook(opp, oside, width) if probe_hook is not None else dd_probe(opp, oside, width, rng)
    if Om is None:
        return none
    ledger.n_probes += sum(Om.row.values()) if left else sum(Om.col.values())
    if left:
        A = la.site_times_core(fr.U0, fr.S0)
        Ysk = stream_left(lenv, W1, W2, renv, A, fr.V0, Om)
        project = lambda T: la.perp_left(fr.U0, T)
    else:
        Bv = la.core_times_site(fr.S0, fr.V0)
        Ysk = stream_right(lenv, W1, W2, renv, fr.U0, Bv, Om)
        project = lambda T: la.perp_right(fr.V0, T)
    Y = project(Ysk)
    ny = Y.norm()
    if ny <= CBE_EPS:
        return none
    width_om = sum(Om.row.values()) if left else sum(Om.col.values())
    rmax = min(cap, width_om)
    cache = {}

    def candidate(k):
        if k <= 0:
            return None
        if k in cache:
            return cache[k]
        if left:
            r = la.svd_left(Y, cutoff=CBE_EPS, nkeep=k)
            B = None if not r.kept else reortho(frame, r.U, side)
        else:
            r = la.svd_right(Y, cutoff=CBE_EPS, nkeep=k)
            B = None if not r.kept else reortho(frame, r.Vh, side)
        cache[k] = B
        return B

    limit = opts.tail_factor * alloc.sample

    def metric(k):
        return alloc.scale * sample_residual(candidate(k), project, Y, Om, side)

    if metric(0) <= limit:
        return none
    hi = rmax
    if metric(hi) > limit:
        directions = candidate(hi)
        count = hi
    else:
        lo = 0
        while hi - lo > 1:
            k = (hi + lo) // 2
            if metric(k) <= limit:
                hi = k
            else:
                lo = k
        directions = candidate(hi)
        count = 0 if directions is None else _width_of(directions, side)
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
    nq = _width_of(directions, side)
    ledger.n_selected += nq
    grown = oplus(frame, directions, side)
    return Expansion(grown, fr.V0, nq) if left else Expansion(fr.U0, grown, nq)


def _dd_cbe_adaptive(fr: BondFrame, frame: Fam, opp: Fam, side: str, oside: str, lenv, renv, W1,
                     W2, alloc: Alloc, ledger: Ledger, cap: int, wbind: bool, full_opp: int,
                     workcap: int, opts: StepOptions, rng) -> Expansion:
    """dd_cbe_expand's ``:adaptive`` branch (dd_cbe_expand.jl:88-118,362-433): the
    holdout-certified controller (oracle/dd_adaptive.py) picks the new directions."""
    from . import dd_adaptive as da

    left = side == "left"
    width0 = da.initial_width(fr.U0, fr.V0, fr.old_rank)
    wcap = full_opp
    if workcap > 0:
        wcap = min(wcap, workcap)
    if left:
        A = la.site_times_core(fr.U0, fr.S0)
        sketch = lambda Om: stream_left(lenv, W1, W2, renv, A, fr.V0, Om)       # noqa: E731
        project = lambda T: la.perp_left(fr.U0, T)                              # noqa: E731
    else:
        Bv = la.core_times_site(fr.S0, fr.V0)
        sketch = lambda Om: stream_right(lenv, W1, W2, renv, fr.U0, Bv, Om)     # noqa: E731
        project = lambda T: la.perp_right(fr.V0, T)                             # noqa: E731

    def draw(width, boost):
        return da.dd_probe(opp, oside, width, rng, boost)

    directions, _status = da.dd_controller(frame, side, draw, sketch, project, width0, wcap, cap,
                                           alloc, ledger, holdout_width=opts.holdout_width)
    return _finish_expansion(fr, frame, side, directions, cap, wbind, ledger)


def fusion_probe(frame: Fam, side: str) -> Optional[Fam]:
    """fusion_basis (dd_cbe.jl:143) of ``frame``'s fused local space: per fused charge the
    slices stacked in ascending order, one identity block each, in the probe layout of
    ``_graded_draw`` (a unitary change of basis of the fused legs)."""
    dfull = fused_dims(frame, side)
    link = frame.row if side == "left" else frame.col
    wl = {q: d for q, d in sorted(dfull.items()) if d > 0}
    if not wl:
        return None
    Om = Fam(frame.row, wl, frame.shifts) if side == "left" else Fam(wl, frame.col, frame.shifts)
    for q, df in wl.items():
        off = 0
        for s, sh in enumerate(frame.shifts):
            lq = q - sh if side == "left" else q + sh
            dl = link.get(lq, 0)
            if dl == 0:
                continue
            blk = np.zeros((dl, df), dtype=complex)
            blk[:, off: off + dl] = np.eye(dl)
            off += dl
            if side == "left":
                Om.blocks[(s, lq)] = blk
            else:
                Om.blocks[(s, q)] = blk.T.copy()
    return Om


def exact_probe(frame: Fam, side: str) -> Optional[Fam]:
    """_dd_probe with exact = true (dd_cbe.jl:141-148,155-160): the fusion basis with the frame
    projected out, i.e. the whole discarded space.  No random numbers."""
    dfull = fused_dims(frame, side)
    dT = frame_widths(frame, side)
    if sum(max(d - dT.get(q, 0), 0) for q, d in dfull.items()) == 0:
        return None
    Om = fusion_probe(frame, side)
    proj = la.perp_left(frame, Om) if side == "left" else la.perp_right(frame, Om)
    return proj if proj.norm() > 0 else None


def resolved_fold(opts) -> bool:
    """fold_omega's default, order > 1 and not exact (bug_step_stages.jl:67)."""
    return (opts.order > 1 and not opts.exact) if opts.fold_omega is None else bool(opts.fold_omega)


def _dd_cbe_exact(fr: BondFrame, frame: Fam, opp: Fam, side: str, oside: str, lenv: Fam,
                  renv: Fam, W1, W2, ledger: Ledger, cap: int, wbind: bool) -> Expansion:
    """dd_cbe_expand with exact = true (dd_cbe_expand.jl:91-97,346-360): _certified_range without
    an allocation is _range_find (range_finder.jl:40-48, sketch.jl:119-132), whose one pass at
    the full width sketches the exact probe and keeps _dd_select of it; no probes or holdouts
    are counted."""
    from . import dd_adaptive as da

    Om = exact_probe(opp, oside)
    if Om is None:
        return Expansion(fr.U0, fr.V0, 0)
    if side == "left":
        Ysk = stream_left(lenv, W1, W2, renv, la.site_times_core(fr.U0, fr.S0), fr.V0, Om)
    else:
        Ysk = stream_right(lenv, W1, W2, renv, fr.U0, la.core_times_site(fr.S0, fr.V0), Om)
    directions = da.dd_select(frame, Ysk, side)
    return _finish_expansion(fr, frame, side, directions, cap, wbind, ledger)


def _finish_expansion(fr: BondFrame, frame: Fam, side: str, directions: Optional[Fam], cap: int,
                      wbind: bool, ledger: Ledger) -> Expansion:
    """The common end of dd_cbe_expand (dd_cbe_expand.jl:383-419)."""
    left = side == "left"
    none = Expansion(fr.U0, fr.V0, 0)
    if directions is None:
        return none
    count = _width_of(directions, side)
    if count > cap:
        directions = trim_total(directions, side, cap)
        if directions is None:
            return none
    if wbind and count >= cap:
        ledger.workcap_reached = True
    directions = cap_to_local_room(frame, directions, side)
    if directions is None:
        return none
    nq = _width_of(directions, side)
    ledger.n_selected += nq
    grown = oplus(frame, directions, side)
    return Expansion(grown, fr.V0, nq) if left else Expansion(fr.U0, grown, nq)


# ---- local Krylov enrichment (Tests C/D) -------------------------------------------------

def extend_frame(frame: Fam, residual: Fam, direction: Fam, side: str, weight: float,
                 allowance: float, room: int):
    """_extend_frame (krylov_enrichment.jl:135-171) -> (frame, n_added, fits)."""
    rn = residual.norm()
    cut = CBE_EPS * max(direction.norm(), rn) / rn
    svd = la.svd_left if side == "left" else la.svd_right
    res = svd(residual, cutoff=cut)
    nk = len(res.kept)
    keep, tail = ok.tail_rank(res.kept, weight, allowance, min(nk, room), base=la.sq_weight(res.trunc))
    fits = tail <= allowance
    if keep == 0:
        return frame, 0, fits
    if keep < nk:
        res = svd(residual, cutoff=cut, nkeep=keep)
    newb = reortho(frame, res.U if side == "left" else res.Vh, side)
    newb = cap_to_local_room(frame, newb, side)
    if newb is None:
        return frame, 0, fits
    return oplus(frame, newb, side), _width_of(newb, side), fits


UNION_KEEP = 0.5


def union_fresh(frame: Fam, cand: Fam, side: str, tol: float = 0.0, room: int = 1 << 60):
    """_union_fresh (half_sweeps.jl:117-176): orthonormal directions of ``cand`` outside
    ``frame`` that keep at least UNION_KEEP of themselves under a second projection, or None."""
    if room <= 0:
        return None
    perp = la.perp_left if side == "left" else la.perp_right
    svd = la.svd_left if side == "left" else la.svd_right
    pick = (lambda r: r.U) if side == "left" else (lambda r: r.Vh)          # noqa: E731
    residual = perp(frame, cand)
    rn, cn = residual.norm(), cand.norm()
    if rn <= CBE_EPS * max(cn, 1.0) or rn <= tol * cn:
        return None
    first_cut = max(CBE_EPS * max(cn, rn) / rn, tol * cn / rn)
    basis = pick(svd(residual.scale(1.0 / rn), cutoff=first_cut))
    second = perp(frame, basis)
    smax = max((e[0] for e in svd(second, cutoff=1e-13).kept), default=0.0)
    if smax < UNION_KEEP:
        return None
    return pick(svd(second, cutoff=UNION_KEEP / smax, nkeep=None if room >= 1 << 60 else room))


def union_frame(frame: Fam, cand: Fam, side: str, tol: float = 0.0, room: int = 1 << 60) -> Fam:
    """_union_left/_union_right (half_sweeps.jl:184-207)."""
    fresh = union_fresh(frame, cand, side, tol, room)
    return frame if fresh is None else oplus(frame, fresh, side)


def append_ranked(frame: Fam, candidates: List[Fam], side: str, allowance: float, workcap: int):
    """_append_ranked (krylov_enrichment.jl:185-196) -> (frame, n_added, fits): one SVD of the
    direct sum of the weighted residuals keeps the fewest top directions whose tail fits."""
    if not candidates:
        return frame, 0, True
    room = workcap - _width_of(frame, side) if workcap > 0 else 1 << 60
    block = candidates[0]
    for c in candidates[1:]:
        block = oplus(block, c, side)
    if room <= 0:
        return frame, 0, block.norm() <= allowance
    return extend_frame(frame, block, block, side, 1.0, allowance, room)


@dataclass
class KrylovStats:
    eta_c: float = 0.0
    eta_d: float = 0.0
    added: int = 0
    n_h1: int = 0
    capped: bool = False
    met: bool = True


def krylov_enrich(seed: Fam, applyH, start_frame: Fam, side: str, tau: complex, alloc: Alloc,
                  workcap: int, recurrence: str = "lanczos", maxdepth: int = ok.KRY_MAXDEPTH,
                  amp: float = 1.0, state_seed: bool = True):
    """_krylov_enrich + _krylov_basis + the chains (krylov_enrichment.jl:213-292,
    krylov_recurrence.jl:27-67, krylov_lanczos.jl:18-71, krylov_arnoldi.jl:13-58).  ``amp``
    scales every direction's weight (a frontier seed's norm relative to the state) and
    ``state_seed = False`` weighs a non-unitary step absolutely (_residual_integral)."""
    perp = la.perp_left if side == "left" else la.perp_right
    allow = alloc.krylov
    ranked = alloc.krylov_rank > 0
    candidates: List[Fam] = []
    st = KrylovStats()
    state = {"frame": start_frame}

    def score(direction, weight, first):
        residual = perp(state["frame"], direction)
        eta = weight * residual.norm()
        if first:
            st.eta_c = eta
        else:
            st.eta_d = max(st.eta_d, eta)
        if not eta > allow:
            return True
        if ranked:                       # measured against start_frame, selected at the end
            candidates.append(residual.scale(weight))
            return True
        room = workcap - _width_of(state["frame"], side) if workcap > 0 else 1 << 60
        if room <= 0:
            return False
        state["frame"], nadd, fits = extend_frame(state["frame"], residual, direction, side,
                                                  weight, allow, room)
        st.added += nadd
        return fits

    snorm = seed.norm()
    if snorm == 0:
        return state["frame"], st
    teff = tau if tau != 0 else 1.0 + 0j
    v0 = seed.scale(1.0 / snorm)
    closed, capped = False, False
    Hm = np.zeros((maxdepth + 1, maxdepth + 1), dtype=complex)
    vs = [v0]
    if recurrence == "lanczos":
        prev_beta = 0.0
        for k in range(1, maxdepth + 1):
            img = applyH(vs[k - 1])
            st.n_h1 += 1
            if k > 1:
                img = img.axpy(-prev_beta, vs[k - 2])
            alpha = ob.inner(vs[k - 1], img)
            if not abs(alpha.imag) <= 1e-9 * max(abs(alpha), 1.0):
                raise ValueError("Lanczos needs a Hermitian one-site generator")
            Hm[k - 1, k - 1] = alpha.real
            img = img.axpy(-alpha.real, vs[k - 1])
            if k >= 20 and k % 2 == 0:
                coeffs = [ob.inner(v, img) for v in vs]
                for c, v in zip(coeffs, vs):
                    img = img.axpy(-c, v)
            beta = img.norm()
            if beta <= ok.KRY_BREAKDOWN:
                closed = True
                break
            Hm[k, k - 1] = beta
            Hm[k - 1, k] = beta
            nv = img.scale(1.0 / beta)
            vs.append(nv)
            w = amp * ok.kry_weight(Hm, k, teff, state_seed)
            if not score(nv, w, k == 1):
                capped = True
                break
            if w <= allow:
                closed = True
                break
            prev_beta = beta
    else:
        for k in range(1, maxdepth + 1):
            img = applyH(vs[-1])
            st.n_h1 += 1
            for _p in range(2):
                coeffs = [ob.inner(v, img) for v in vs]
                for row, c in enumerate(coeffs):
                    Hm[row, k - 1] += c
                for c, v in zip(coeffs, vs):
                    img = img.axpy(-c, v)
            rn = img.norm()
            if rn <= ok.KRY_BREAKDOWN:
                closed = True
                break
            Hm[k, k - 1] = rn
            nv = img.scale(1.0 / rn)
            vs.append(nv)
            w = amp * ok.kry_weight(Hm, k, teff, state_seed)
            if not score(nv, w, k == 1):
                capped = True
                break
            if w <= allow:
                closed = True
                break
    if not closed:
        capped = True
    if ranked:
        state["frame"], st.added, fits = append_ranked(start_frame, candidates, side,
                                                       alloc.krylov_rank, workcap)
        capped = capped or not fits
    st.capped = capped
    st.met = closed and not capped
    return state["frame"], st


def log_bond(ledger: Ledger, bl: Ledger, bond: int, side: str, width: int):
    """_log_bond! (basis_budget.jl:390-418): log the bond's own indicators and counts from its
    fresh ledger ``bl``, then merge ``bl`` into ``ledger``."""
    from .closure import _merge
    bl.bond_log.append(dict(
        bond=bond, side=side, sample=bl.max_sample, tail=bl.max_tail, krylov_c=bl.max_krylov_c,
        krylov_d=bl.max_krylov_d, n_probes=bl.n_probes, n_holdout=bl.n_holdout,
        n_selected=bl.n_selected, n_targeted=bl.n_targeted, n_krylov=bl.n_krylov, n_h1=bl.n_h1,
        workcap=bl.workcap_reached, met=bl.tolerance_met, unmet=list(bl.unmet), width=width))
    _merge(ledger, bl)


def record_krylov(ledger: Ledger, st: KrylovStats, alloc: Alloc):
    ledger.max_krylov_c = max(ledger.max_krylov_c, st.eta_c)
    ledger.max_krylov_d = max(ledger.max_krylov_d, st.eta_d)
    ledger.max_krylov = max(ledger.max_krylov, st.eta_c, st.eta_d)
    ledger.n_krylov += st.added
    ledger.n_h1 += st.n_h1
    if st.capped:
        ledger.workcap_reached = True
    if not st.met:
        ledger.unmet_("krylov")


# ---- the step -----------------------------------------------------------------------------

@dataclass
class StepInfo:
    ledger: Ledger
    expanded: List[int]
    root_rank: int
    discarded: float
    n_root: int
    root_report: Optional[dict] = None


def _cap_at(wcap, b):
    return wcap if isinstance(wcap, int) else wcap[b]


def cbe_bug_step(psi: MPS, mpo, tau: complex, opts: Optional[StepOptions] = None,
                 probe_hook=None, rng: Optional[la.ProbeRNG] = None,
                 frames_report: Optional[dict] = None) -> StepInfo:
    """One order-p CBE-BUG step in place (bug_step.jl:17-510); order > 1 closes the root
    frontiers by one enriching sweep each (oracle/closure.py), or by the measured loop under
    ``closure_checks``.  ``rng`` is a probe stream the caller keeps across steps (the reference's
    ``rng`` keyword); None draws from a fresh stream seeded by ``opts.seed``."""
    opts = opts or StepOptions()
    L = len(psi)
    if len(mpo) != L:
        raise ValueError("MPO and state lengths differ")
    if L < 2:
        raise ValueError("cbe_bug_step needs at least two sites")
    if opts.order < 1:
        raise ValueError("order must be >= 1")
    rl = (max(1, min(L - 1, L // 2)) - 1) if opts.root is None else opts.root
    keeps = list(opts.keep_states)
    if opts.keep_tols and len(opts.keep_tols) != len(keeps):
        raise ValueError("keep_tols and keep_states differ in length")
    ktols = [float(t) for t in opts.keep_tols] if opts.keep_tols else [0.0] * len(keeps)
    if any(not t >= 0 for t in ktols):
        raise ValueError("keep_tols must be >= 0")
    cap_keep = opts.work_ceiling_rule != "maxdim"
    if opts.frames_only and (opts.truncate or opts.root_report):
        raise ValueError("frames_only needs root_report off and truncate = False")
    if opts.closure_probe < 0:
        raise ValueError("closure_probe must be a non-negative count")
    ok.check_root_method(opts.root_method)
    if opts.root_maxiter < 1:
        raise ValueError("root_maxiter must be positive")
    d = len(psi.sites[0].shifts)
    tau = complex(tau)
    es = float(opts.eshift)
    rng = la.ProbeRNG(opts.seed) if rng is None else rng
    bd0 = psi.bond_dims()
    wcap = work_ceilings(opts.work_ceiling_rule, bd0, d, opts.maxdim, opts.bond_ceiling_factor,
                         rl, opts.root_ceiling_factor)
    nrm = psi.sites[psi.center].norm() if 0 <= psi.center < L else 1.0
    alloc = basis_allocation(opts.tol_basis, L - 1, tau, nrm, opts.order, opts.tol_krylov,
                             opts.krylov_rank_tol)
    ledger = Ledger(r_old=max(bd0 + [0]))
    recurrence = "lanczos" if (opts.krylov_mode == "auto" and opts.hermitian) else (
        "arnoldi" if opts.krylov_mode == "auto" else opts.krylov_mode)
    expanded = [0] * (L - 1)
    left_frames: List[Optional[Fam]] = [None] * (rl + 1)
    right_frames: List[Optional[Fam]] = [None] * (L - 1)

    # ---- per-half-sweep generators and ledgers under `parallel` (bug_step.jl:234-268): the
    # right sweep draws from a generator seeded by the step's next word, each sweep logs into its
    # own ledger, merged left then right after both; run here one after the other ----
    if opts.parallel:
        rngL, rngR = rng, rng.split()
        ledL, ledR = Ledger(), Ledger()
    else:
        rngL = rngR = rng
        ledL = ledR = ledger

    # ---- left half-sweep (half_sweeps.jl:218-310) ----
    psiL = psi.copy()
    canonical(psiL, 0)
    renvs = [None] * (L + 1)
    renvs[L] = right_boundary(mpo, psiL)
    for i in range(L - 1, 1, -1):
        renvs[i] = ob.push_right(renvs[i + 1], psiL.sites[i], mpo[i])
    lch = left_boundary(mpo, psiL)
    lroot = lch
    carry = None
    kcar = [None] * len(keeps)
    for b in range(rl + 1):
        carried = psiL.sites[b] if carry is None else la.core_times_site(carry, psiL.sites[b])
        fr = frame_from(carried, psiL.sites[b + 1])
        bl = Ledger()
        ex = dd_cbe_expand(fr, mpo, b, lch, renvs[b + 2], "left", alloc, bl,
                           _cap_at(wcap, b), opts, rngL, probe_hook)
        right_projected = la.overlap_right(psiL.sites[b + 1], ex.V_ex)
        Kex = la.site_times_core(carried, right_projected)
        rch1 = ob.push_right(renvs[b + 2], ex.V_ex, mpo[b + 1])
        Hl, Hw, Hr = lch, mpo[b], rch1
        Wk, kst = krylov_enrich(Kex, lambda v: ob.apply_one_site(Hl, Hw, Hr, v, shift=es), ex.U_ex,
                                "left", tau, alloc, _cap_at(wcap, b), recurrence)
        record_krylov(bl, kst, alloc)
        # half_sweeps.jl:280-296: every kept state unioned before any carry is taken
        kxs = [ks.sites[b] if kc is None else la.core_times_site(kc, ks.sites[b])
               for ks, kc in zip(keeps, kcar)]
        for kx, tol in zip(kxs, ktols):
            room = _cap_at(wcap, b) - sum(Wk.col.values()) if cap_keep and tol > 0 else 1 << 60
            Wk = union_frame(Wk, kx, "left", tol, room)
        kcar = [la.overlap_left(Wk, kx) for kx in kxs]
        left_frames[b] = Wk
        expanded[b] = sum(Wk.col.values())
        log_bond(ledL, bl, b, "left", expanded[b])
        carry = la.overlap_left(Wk, carried)
        if b == rl:
            lroot = lch
        else:
            lch = ob.push_left(lch, Wk, mpo[b])

    # ---- right half-sweep (half_sweeps.jl:323-410) ----
    psiR = psi
    canonical(psiR, L - 1)
    lenvs = [left_boundary(mpo, psiR)]
    for i in range(L - 2):
        lenvs.append(ob.push_left(lenvs[-1], psiR.sites[i], mpo[i]))
    rch = right_boundary(mpo, psiR)
    rroot = rch
    carry = None
    kcar = [None] * len(keeps)
    for b in range(L - 2, rl - 1, -1):
        carried = psiR.sites[b + 1] if carry is None else la.site_times_core(psiR.sites[b + 1], carry)
        fr = frame_from(psiR.sites[b], carried)
        bl = Ledger()
        ex = dd_cbe_expand(fr, mpo, b, lenvs[b], rch, "right", alloc, bl,
                           _cap_at(wcap, b), opts, rngR, probe_hook)
        left_projected = la.overlap_left(ex.U_ex, psiR.sites[b])
        Bex = la.core_times_site(left_projected, carried)
        lch1 = ob.push_left(lenvs[b], ex.U_ex, mpo[b])
        Hl, Hw, Hr = lch1, mpo[b + 1], rch
        Zk, kst = krylov_enrich(Bex, lambda v: ob.apply_one_site(Hl, Hw, Hr, v, shift=es), ex.V_ex,
                                "right", tau, alloc, _cap_at(wcap, b), recurrence)
        record_krylov(bl, kst, alloc)
        bxs = [ks.sites[b + 1] if kc is None else la.site_times_core(ks.sites[b + 1], kc)
               for ks, kc in zip(keeps, kcar)]                       # half_sweeps.jl:384-395
        for bx, tol in zip(bxs, ktols):
            room = _cap_at(wcap, b) - sum(Zk.row.values()) if cap_keep and tol > 0 else 1 << 60
            Zk = union_frame(Zk, bx, "right", tol, room)
        kcar = [la.overlap_right(bx, Zk) for bx in bxs]
        log_bond(ledR, bl, b, "right", sum(Zk.row.values()))
        right_frames[b] = Zk
        if b != rl:
            expanded[b] = sum(Zk.row.values())
        carry = la.overlap_right(carried, Zk)
        if b == rl:
            rroot = rch
        else:
            rch = ob.push_right(rch, Zk, mpo[b + 1])
    AR = carry
    if opts.parallel:
        from .closure import _merge
        _merge(ledger, ledL)
        _merge(ledger, ledR)

    # ---- root start, projection miss (bug_step.jl:278-327) ----
    AL = None
    for i in range(rl + 1):
        c = psi.sites[i] if AL is None else la.core_times_site(AL, psi.sites[i])
        AL = la.overlap_left(left_frames[i], c)
    root_start = la.product(AL.row, AR.col, [0], AL, AR, [(0, 0, 0, la.OP_N, la.OP_N, 1.0)])

    # ---- order p > 1: close the root frontiers j = 1..p-1 (bug_step_stages.jl:172-240) ----
    from .closure import KLFrames
    if opts.order > 1:
        from .closure import checked_closure, single_pass_closure
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=es)
        close = checked_closure if opts.closure_checks else single_pass_closure
        root_start, _ = close(fr, psi, mpo, root_start, tau, alloc, opts, wcap, nrm, rng,
                              recurrence, ledger, opts.order)
        lroot, rroot = fr.lenv, fr.renv
        _refresh_expanded(expanded, left_frames, right_frames, rl)
    elif opts.closure_probe > 0:
        # measure only, on a copy of the stream: the step itself is unchanged
        import copy
        from .closure import _merge, measure_frontiers
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=es)
        *_, plg, _ = measure_frontiers(fr, psi, mpo, root_start, tau, alloc, opts, wcap, nrm,
                                       copy.copy(rng), recurrence, opts.closure_probe)
        _merge(ledger, plg)
        ledger.leak_initial = list(ledger.leak_initial) + list(plg.leak_initial)
        ledger.rho_initial = list(ledger.rho_initial) + list(plg.rho_initial)
        ledger.n_root_frontier += plg.n_root_frontier
    npsi = psi.sites[psi.center].norm()
    ledger.r_aug = max(max(expanded + [0]), sum(right_frames[rl].row.values()))
    ledger.proj_err = projection_miss(psi, left_frames, right_frames, rl) / npsi if npsi > 0 else 0.0
    if ledger.proj_err > opts.tol_basis:
        ledger.unmet_("projection")

    # ---- root generator and solve (bug_step.jl:366-382) ----
    Lr = ob.push_left(lroot, left_frames[rl], mpo[rl])
    Rr = ob.push_right(rroot, right_frames[rl], mpo[rl + 1])
    napply = [0]
    vec_ops = dict(inner=ob.inner, norm=lambda x: x.norm(), axpy=lambda a, x, y: y.axpy(a, x),
                   scale=lambda a, x: x.scale(a))

    def solve_with(Lr, Rr, start):
        def apply_root(C):
            napply[0] += 1
            return ob.apply_zero_site(Lr, Rr, C, shift=es)
        rk = dict(root_method=opts.root_method, rational_shift=opts.root_shift)
        if opts.hermitian:
            return ok.lanczos_expv(apply_root, tau, start, tol=1e-15, maxiter=opts.root_maxiter,
                                   reorth=opts.reorth, conv_tol=opts.tol_expv, substeps=0,
                                   **rk, **vec_ops)
        return ok.arnoldi_expv(apply_root, tau, start, tol=1e-15, maxiter=opts.root_maxiter,
                               conv_tol=opts.tol_expv, substeps=0, **rk, **vec_ops)

    if opts.frames_only:
        result = root_start
    else:
        result, rep = solve_with(Lr, Rr, root_start)
    report = None
    if opts.root_report:
        fr = KLFrames(left_frames, right_frames, rl, lroot, rroot, eshift=es)
        root_start, result, rep, report = report_root(
            psi, mpo, fr, Lr, Rr, root_start, result, rep, solve_with, tau, alloc, opts, wcap,
            nrm, rng, recurrence, ledger, expanded)
    if frames_report is not None:       # bug_step.jl:455-467
        frames_report.update(left=list(left_frames), right=list(right_frames), root_link=rl,
                             lenv=lroot, renv=rroot, eshift=es)
    if not opts.frames_only:
        ledger.expv_err = rep.err
        ledger.expv_iters = rep.iterations
        if rep.err > opts.tol_expv * (1 + 1e-8):
            ledger.unmet_("expv")

    # ---- assembly and compression (diagnostics.jl:269-337) ----
    res = la.svd_core(result, cutoff=0.0)
    expanded[rl] = sum(left_frames[rl].col.values())
    discarded = la.trunc_weight(res)
    for i in range(rl):
        psi.sites[i] = left_frames[i]
    psi.sites[rl] = la.site_times_core(left_frames[rl], res.U)
    SV = la.diag_times(res.S, res.Vh, "left")
    psi.sites[rl + 1] = la.core_times_site(SV, right_frames[rl])
    for b in range(rl + 1, L - 1):
        psi.sites[b + 1] = right_frames[b]
    psi.center = rl + 1
    root_rank = sum(psi.sites[rl].col.values())
    bwts = [0.0] * (L - 1)
    if opts.truncate:
        wtr, cumw, capped = truncate_recursive(psi, rl + 1, opts.maxdim, opts.tol_trunc, bwts)
        discarded = max(discarded, wtr)
        ledger.discarded = cumw
        ledger.maxdim_reached = capped
        if cumw > opts.tol_trunc * (1 + 1e-8):
            ledger.unmet_("trunc")
    bd1 = psi.bond_dims()
    ledger.r_final = max(bd1 + [0])
    for e in ledger.bond_log:            # diagnostics.jl:328-335
        e.update(r_old=bd0[e["bond"]], r_final=bd1[e["bond"]], discarded=bwts[e["bond"]])
    return StepInfo(ledger, expanded, root_rank, discarded, napply[0], report)


def _refresh_expanded(expanded, left_frames, right_frames, rl):
    """``_refresh_expanded!``: the frame widths off the root bond."""
    for b in range(rl):
        expanded[b] = sum(left_frames[b].col.values())
    for b in range(rl + 1, len(expanded)):
        expanded[b] = sum(right_frames[b].row.values())


def report_root(psi, mpo, fr, Lr, Rr, root_start, result, rep, solve_with, tau, alloc, opts, wcap,
                nrm, rng, recurrence, ledger, expanded):
    """``_report_root!`` (bug_step_stages.jl:242-398): the energies before and after the root
    solve (Rayleigh quotients in the Galerkin space plus eshift, exact energies of the state
    since it lies in the frames' span) and, with ``outside_residual``, the outside-Galerkin
    defect |tau| ||(I - P) H psi_1|| / ||psi_1|| of the evolved root vector
    (``complement_action``).  While the defect exceeds tol_basis the frames are re-enriched by
    the evolved vector's own frontier sweep and the root re-solved (at most 2p+2 passes, stopping
    when a pass adds nothing or binds the work ceiling).  Returns the possibly re-solved
    root_start, result and solve report, and the report."""
    from .closure import closure_maxpass, complement_action, frontier_sweep, project_root
    es = fr.eshift

    def rayleigh(S):
        n2 = S.norm() ** 2
        if not n2 > 0:
            return float("nan")
        return float(ob.inner(S, ob.apply_zero_site(Lr, Rr, S, shift=es)).real) / n2 + es

    e_pre = rayleigh(root_start)
    outside = (complement_action(psi, mpo, fr, result, alloc, opts, wcap, rng, recurrence)
               if opts.outside_residual else None)

    def resid(act, S):
        return float("nan") if act is None else abs(tau) * act["bound"] / max(
            S.norm(), np.finfo(float).eps)

    initial = resid(outside, result)
    passes = added = 0
    if opts.outside_residual:
        while not (resid(outside, result) <= opts.tol_basis) and passes < closure_maxpass(
                opts.order):
            passes += 1
            sr = frontier_sweep(fr, psi, mpo, result, tau, alloc, opts, wcap, nrm, rng, recurrence,
                                ledger)
            if sr.added == 0:
                break
            added += sr.added
            root_start = project_root(psi, fr)
            Lr = ob.push_left(fr.lenv, fr.left[fr.rl], mpo[fr.rl])
            Rr = ob.push_right(fr.renv, fr.right[fr.rl], mpo[fr.rl + 1])
            result, rep = solve_with(Lr, Rr, root_start)
            outside = complement_action(psi, mpo, fr, result, alloc, opts, wcap, rng, recurrence)
            if sr.capped:
                break
        if added > 0:
            _refresh_expanded(expanded, fr.left, fr.right, fr.rl)
            ledger.r_aug = max(max(expanded + [0]), sum(fr.right[fr.rl].row.values()))
    e_root = rayleigh(result)
    rn = max(result.norm(), np.finfo(float).eps)
    report = dict(
        eshift=es, e_pre=e_pre, e_root=e_root, resid_initial=initial, outside_passes=passes,
        outside_added=added,
        outside=float("nan") if outside is None else outside["bound"] / rn,
        outside_left=float("nan") if outside is None else outside["left"] / rn,
        outside_right=float("nan") if outside is None else outside["right"] / rn,
        per_left=[] if outside is None else [x / rn for x in outside["per_left"]],
        per_right=[] if outside is None else [x / rn for x in outside["per_right"]],
        resid=float("nan") if outside is None else abs(tau) * outside["bound"] / rn)
    return root_start, result, rep, report


def projection_miss(psi: MPS, left_frames, right_frames, rl: int) -> float:
    """_projection_miss (krylov_closure_frames.jl:242-280); square-root factors by QR."""
    L = len(psi)
    rf = [None] * (L + 1)
    for i in range(L - 1, 0, -1):
        carried = psi.sites[L - 1] if i == L - 1 else la.site_times_core(psi.sites[i], rf[i + 1])
        Lc, _ = la.qr_right(carried)
        rf[i] = Lc
    miss2 = 0.0
    AL = None
    for i in range(rl + 1):
        carried = psi.sites[i] if AL is None else la.core_times_site(AL, psi.sites[i])
        outside = la.perp_left(left_frames[i], carried)
        miss2 += la.site_times_core(outside, rf[i + 1]).norm() ** 2
        AL = la.overlap_left(left_frames[i], carried)
    lf = [None] * (L - 1)
    _, R = qr_core(AL)
    lf[rl] = R
    for b in range(rl + 1, L - 1):
        _, R = la.qr_left(la.core_times_site(lf[b - 1], psi.sites[b]))
        lf[b] = R
    AR = None
    for b in range(L - 2, rl - 1, -1):
        carried = psi.sites[b + 1] if AR is None else la.site_times_core(psi.sites[b + 1], AR)
        outside = la.perp_right(right_frames[b], carried)
        miss2 += la.core_times_site(lf[b], outside).norm() ** 2
        AR = la.overlap_right(carried, right_frames[b])
    return math.sqrt(miss2)


def qr_core(C: Fam):
    """C = Q R per sector of a one-slice core (unique QR)."""
    qb, rb, bond = {}, {}, {}
    floor = la.ZERO_REL * C.norm()
    for (k, q), M in C.blocks.items():
        if not np.linalg.norm(M) > floor:
            continue
        Q, R = la._unique_qr(M)
        bond[q] = Q.shape[1]
        qb[(0, q)] = Q
        rb[(0, q)] = R
    return Fam(C.row, bond, [0], qb), Fam(bond, C.col, [0], rb)


def truncate_recursive(psi: MPS, root: int, maxdim: int, weight_tol: float, bond_weights=None):
    """truncate_recursive! with the budgeted rule (truncation.jl:59-158); root = 0-based site;
    ``bond_weights`` (L - 1 entries) receives each bond's relative discarded norm."""
    L = len(psi)
    sdiv = math.sqrt(root) + math.sqrt(L - 1 - root)
    state = {"nrm2": -1.0, "worst": 0.0, "capped": False}
    cum = [0.0, 0.0]

    def cut(T, left_split):
        svd = la.svd_left if left_split else la.svd_right
        res = svd(T, cutoff=0.0, nkeep=maxdim)
        tot = la.sq_weight(res.kept) + la.sq_weight(res.trunc)
        if state["nrm2"] < 0:
            state["nrm2"] = tot
        share = weight_tol * math.sqrt(max(state["nrm2"], 0.0)) / sdiv if sdiv > 0 else 0.0
        nk = len(res.kept)
        keep, tail = ok.tail_rank(res.kept, 1.0, share, nk, base=la.sq_weight(res.trunc))
        if tail > share and res.trunc:
            state["capped"] = True
        keep = max(keep, 1)
        if keep < nk:
            res = svd(T, cutoff=0.0, nkeep=keep)
        return res

    def account(res, side, bond):
        state["worst"] = max(state["worst"], la.trunc_weight(res))
        cum[side] += la.sq_weight(res.trunc)
        if bond_weights is not None:
            bond_weights[bond] = la.sq_weight(res.trunc)

    for site in range(root, 0, -1):
        res = cut(psi.sites[site], False)
        account(res, 0, site - 1)
        psi.sites[site] = res.Vh
        US = la.diag_times(res.S, res.U, "right")
        psi.sites[site - 1] = la.site_times_core(psi.sites[site - 1], US)
        psi.center = site - 1
    canonical(psi, root)
    for site in range(root, L - 1):
        res = cut(psi.sites[site], True)
        account(res, 1, site)
        psi.sites[site] = res.U
        SV = la.diag_times(res.S, res.Vh, "left")
        psi.sites[site + 1] = la.core_times_site(SV, psi.sites[site + 1])
        psi.center = site + 1
    nrm2 = state["nrm2"]
    cumulative = (math.sqrt(cum[0]) + math.sqrt(cum[1])) / math.sqrt(nrm2) if nrm2 > 0 else 0.0
    if bond_weights is not None:
        for i, w in enumerate(bond_weights):
            bond_weights[i] = math.sqrt(w / nrm2) if nrm2 > 0 else 0.0
    return state["worst"], cumulative, state["capped"]
