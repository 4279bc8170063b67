"""MPO environments and the one-/zero-site effective Hamiltonians, on the device.

Reference: ``MPOLink``, ``_mpo_left_step``/``_mpo_right_step``, ``push_left_channels``/
``push_right_channels``, ``left_env_stack``/``right_env_stack``
(src/BUG/Core/Environments/mpo_environments.jl:47-175) and ``one_site_h``/``apply_one_site``,
``zero_site_h``/``apply_zero_site``, ``shift_h`` (effective_operators.jl:15-129).

An environment on link b is the family ``E_w`` over the MPO channels of that link, a matrix
(bra bond x ket bond) per channel with shift ``-charge(w)``.  The chain boundaries are explicit
1x1 environments (the reference uses ``nothing`` there and drops the dim-1 MPO leg instead).
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .blocks import BlockTensor, Layout, Leg, current_stream_handle, device
from .mpo import MPO, MPOSite
from .mps import SymMPS


def env_layout(bra: Leg, ket: Leg, channel_charges) -> Layout:
    return Layout(bra, ket, [-int(c) for c in channel_charges])


def boundary_left(mpo: MPO, leg: Leg) -> BlockTensor:
    """Left boundary environment: channel ``id`` on the dim-1 vacuum link."""
    if leg.total != 1:
        raise ValueError("the left boundary link must be one-dimensional")
    lay = env_layout(leg, leg, mpo[0].cl)
    return BlockTensor.from_blocks(lay, {(0, leg.charges[0]): np.ones((1, 1))})


def boundary_right(mpo: MPO, leg: Leg) -> BlockTensor:
    """Right boundary environment: channel ``done`` on the dim-1 total-charge link."""
    if leg.total != 1:
        raise ValueError("the right boundary link must be one-dimensional")
    lay = env_layout(leg, leg, mpo[len(mpo) - 1].cr)
    return BlockTensor.from_blocks(lay, {(0, leg.charges[0]): np.ones((1, 1))})


def _stream():
    return current_stream_handle()


def ctypes_stream(s) -> int:
    return int(s.cuda_stream)


def push_left(L: BlockTensor, A: BlockTensor, W: MPOSite, bra: Optional[BlockTensor] = None,
              precision: int = _lib.PREC_FP64) -> BlockTensor:
    """Environment on link i+1 from link i through site tensor ``A`` (mpo_environments.jl:65-90)."""
    ctx = _lib.context()
    bra_t = A if bra is None else bra
    out = BlockTensor.empty(env_layout(bra_t.layout.col, A.layout.col, W.cr))
    bd = bra.desc() if bra is not None else None
    ctx.check(ctx._lib.bug_env_push_left(
        ctx.handle, L.desc(), A.desc(), bd, W.descriptor(), out.desc(), precision, _stream()))
    return out


def push_right(R: BlockTensor, B: BlockTensor, W: MPOSite, bra: Optional[BlockTensor] = None,
               precision: int = _lib.PREC_FP64) -> BlockTensor:
    """Environment on link i from link i+1 through site tensor ``B`` (mpo_environments.jl:92-108)."""
    ctx = _lib.context()
    bra_t = B if bra is None else bra
    out = BlockTensor.empty(env_layout(bra_t.layout.row, B.layout.row, W.cl))
    bd = bra.desc() if bra is not None else None
    ctx.check(ctx._lib.bug_env_push_right(
        ctx.handle, R.desc(), B.desc(), bd, W.descriptor(), out.desc(), precision, _stream()))
    return out


def left_env_stack(psi: SymMPS, mpo: MPO, upto: Optional[int] = None,
                   precision: int = _lib.PREC_FP64) -> List[BlockTensor]:
    """Environments on links 0..upto (``left_env_stack``, mpo_environments.jl:141-158)."""
    upto = len(psi) if upto is None else upto
    envs = [boundary_left(mpo, psi.bond_leg(0))]
    for i in range(upto):
        envs.append(push_left(envs[-1], psi[i], mpo[i], precision=precision))
    return envs


def right_env_stack(psi: SymMPS, mpo: MPO, downto: int = 0,
                    precision: int = _lib.PREC_FP64) -> List[Optional[BlockTensor]]:
    """Environments on links downto..L, indexed by link (``right_env_stack``, :161-175)."""
    L = len(psi)
    envs: List[Optional[BlockTensor]] = [None] * (L + 1)
    envs[L] = boundary_right(mpo, psi.bond_leg(L))
    for i in range(L - 1, downto - 1, -1):
        envs[i] = push_right(envs[i + 1], psi[i], mpo[i], precision=precision)
    return envs


def apply_one_site(L: BlockTensor, W: MPOSite, R: BlockTensor, X: BlockTensor,
                   shift: float = 0.0, precision: int = _lib.PREC_FP64) -> BlockTensor:
    """``(H_eff - shift) X`` for a site tensor (effective_operators.jl:39-64,118-129)."""
    ctx = _lib.context()
    Y = BlockTensor.empty(X.layout)
    ctx.check(ctx._lib.bug_apply_one_site(
        ctx.handle, L.desc(), W.descriptor(), R.desc(), X.desc(), Y.desc(), float(shift),
        precision, _stream()))
    return Y


def apply_one_site_batch(items, precision: int = _lib.PREC_FP64) -> List[BlockTensor]:
    """``(H_eff,i - shift_i) X_i`` for independent instances ``items = [(L, W, R, X, shift), ...]``
    with one grouped-GEMM launch per product stage for all of them (their block products packed
    into one task list) instead of two launches per instance; each result equals the
    single-instance ``apply_one_site`` to rounding."""
    k = len(items)
    if k == 0:
        return []
    ctx = _lib.context()
    ys = [BlockTensor.empty(it[3].layout) for it in items]
    Ld = (_lib.BugBsm * k)(*[it[0].desc() for it in items])
    Wd = (_lib.BugMpoSite * k)(*[it[1].descriptor() for it in items])
    Rd = (_lib.BugBsm * k)(*[it[2].desc() for it in items])
    Xd = (_lib.BugBsm * k)(*[it[3].desc() for it in items])
    Yd = (_lib.BugBsm * k)(*[y.desc() for y in ys])
    sh = np.array([float(it[4]) for it in items], dtype=np.float64)
    ctx.check(ctx._lib.bug_apply_one_site_batch(
        ctx.handle, k, Ld, Wd, Rd, Xd, Yd, sh.ctypes.data_as(_lib._f64p), precision, _stream()))
    return ys


class PackedApply:
    """Independent one-site actions (L_i, W_i, R_i fixed, i < K) applied together with one
    grouped-GEMM launch per product stage.  The inputs live in one packed device buffer
    (``x[i]`` are views into it) and the outputs in another (``y[i]``), so the merged launch plan
    is built on the first call and reused on every later one; a Krylov recurrence advancing K
    instances in lockstep keeps its current vectors in ``x``."""

    def __init__(self, ops, layouts, shifts=None, precision: int = _lib.PREC_FP64):
        if len(ops) != len(layouts):
            raise ValueError("one layout per operator")
        self.k = len(ops)
        self.precision = precision
        offs, tot = [], 0
        for lay in layouts:
            offs.append(tot)
            tot += -(-max(lay.nelem, 1) // 16) * 16            # 256-byte aligned instances
        self._xbuf = torch.zeros(max(tot, 1), dtype=torch.complex128, device=device())
        self._ybuf = torch.zeros(max(tot, 1), dtype=torch.complex128, device=device())
        self.x = [BlockTensor(lay, self._xbuf[o: o + lay.nelem]) for lay, o in zip(layouts, offs)]
        self.y = [BlockTensor(lay, self._ybuf[o: o + lay.nelem]) for lay, o in zip(layouts, offs)]
        self._ops = list(ops)
        k = self.k
        self._Ld = (_lib.BugBsm * k)(*[op[0].desc() for op in ops])
        self._Wd = (_lib.BugMpoSite * k)(*[op[1].descriptor() for op in ops])
        self._Rd = (_lib.BugBsm * k)(*[op[2].desc() for op in ops])
        self._Xd = (_lib.BugBsm * k)(*[t.desc() for t in self.x])
        self._Yd = (_lib.BugBsm * k)(*[t.desc() for t in self.y])
        self._sh = np.array([0.0] * k if shifts is None else [float(v) for v in shifts])

    def __call__(self) -> List[BlockTensor]:
        """y[i] = (H_eff,i - shift_i) x[i] for every i (outputs overwritten by the next call)."""
        ctx = _lib.context()
        ctx.check(ctx._lib.bug_apply_one_site_batch(
            ctx.handle, self.k, self._Ld, self._Wd, self._Rd, self._Xd, self._Yd,
            self._sh.ctypes.data_as(_lib._f64p), self.precision, _stream()))
        return self.y

    def apply(self, vs: List[BlockTensor]) -> List[BlockTensor]:
        """Copy vs into the packed inputs and apply."""
        for dst, v in zip(self.x, vs):
            if v.layout != dst.layout:
                raise ValueError("input layout differs from the packed one")
            dst.data.copy_(v.data)
        return self()


def apply_zero_site(L: BlockTensor, R: BlockTensor, C: BlockTensor, shift: float = 0.0,
                    precision: int = _lib.PREC_FP64) -> BlockTensor:
    """``(H0 - shift) C`` for the root core between two pushed environments (:78-107)."""
    ctx = _lib.context()
    Y = BlockTensor.empty(C.layout)
    ctx.check(ctx._lib.bug_apply_zero_site(
        ctx.handle, L.desc(), R.desc(), C.desc(), Y.desc(), float(shift), precision, _stream()))
    return Y
