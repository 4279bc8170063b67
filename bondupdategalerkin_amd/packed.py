"""Krylov exponentials of many independent instances advanced in lockstep over packed vectors.

K independent Hermitian one-site problems (their own environments, MPO tensors and layouts) keep
their Krylov vectors side by side in one device buffer (the segments of a ``PackedApply``), so
one Lanczos iteration of all K costs one packed H_eff apply (one grouped launch per product
stage), one segmented multi-dot with one host read for the coefficients, and segmented
multi-axpys, instead of K of each.  Each instance follows the recurrence of ``lanczos_expv``
(src/BUG/Core/LocalSolvers/krylov.jl:128-263: one Krylov pass, substeps = 1, classical
Gram-Schmidt applied twice against its whole basis) and stops growing its basis at its own
breakdown; the others go on.  ``arnoldi_expv_packed`` does the same for non-Hermitian problems
(the Heston roots, ``arnoldi_expv``, krylov.jl:343-442): upper-Hessenberg Arnoldi with CGS2, the
Hessenberg exponential per instance on the host.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .blocks import BlockTensor, current_stream_handle
from .environments import PackedApply
from .krylov import MAX_NORM, ROOT_MAXITER, Hessenberg, Tridiagonal


class Segments:
    """Segment table (offset, length) of a PackedApply's instances in its packed buffers."""

    def __init__(self, packed: PackedApply):
        base = packed._xbuf.data_ptr()
        rows = []
        for t in packed.x:
            off = (t.data.data_ptr() - base) // 16
            rows += [off, t.layout.nelem]
        self.table = np.ascontiguousarray(rows, dtype=np.int64)
        self.nseg = len(packed.x)
        self.n = packed._xbuf.numel()

    def dots(self, xs: Sequence[torch.Tensor], y: torch.Tensor) -> np.ndarray:
        """(len(xs), nseg) complex: <xs[j]_i | y_i>, 64 vectors per launch."""
        out = np.empty((len(xs), self.nseg), dtype=complex)
        ctx = _lib.context()
        for lo in range(0, len(xs), 64):
            chunk = xs[lo: lo + 64]
            res = np.empty(2 * len(chunk) * self.nseg)
            ptrs = (ctypes.c_void_p * len(chunk))(*[x.data_ptr() for x in chunk])
            ctx.check(ctx._lib.bug_dotc_segments(
                ctx.handle, self.n, self.nseg, self.table.ctypes.data_as(_lib._i64p), len(chunk),
                ptrs, y.data_ptr(), res.ctypes.data_as(_lib._f64p), current_stream_handle()))
            out[lo: lo + len(chunk)] = (res[0::2] + 1j * res[1::2]).reshape(len(chunk), self.nseg)
        return out

    def axpy(self, alphas: np.ndarray, xs: Sequence[torch.Tensor], y: torch.Tensor):
        """y_i += sum_j alphas[j, i] xs[j]_i in place, the terms in j order."""
        ctx = _lib.context()
        for lo in range(0, len(xs), 64):
            chunk = xs[lo: lo + 64]
            a = np.ascontiguousarray(alphas[lo: lo + 64], dtype=complex)
            coef = np.empty(2 * a.size)
            coef[0::2], coef[1::2] = a.real.reshape(-1), a.imag.reshape(-1)
            ptrs = (ctypes.c_void_p * len(chunk))(*[x.data_ptr() for x in chunk])
            ctx.check(ctx._lib.bug_axpy_segments(
                ctx.handle, self.n, self.nseg, self.table.ctypes.data_as(_lib._i64p), len(chunk),
                coef.ctypes.data_as(_lib._f64p), ptrs, y.data_ptr(), current_stream_handle()))


def _pack_helpers(packed: PackedApply, seg: Segments):
    def pack(vs):
        buf = torch.zeros_like(packed._xbuf)
        for t, v in zip(packed.x, vs):
            if v.layout != t.layout:
                raise ValueError("start vector layout differs from the packed one")
            off = (t.data.data_ptr() - packed._xbuf.data_ptr()) // 16
            buf[off: off + t.layout.nelem].copy_(v.data)
        return buf

    def apply(v):
        packed._xbuf.copy_(v)
        packed()
        return packed._ybuf.clone()

    def scaled(v, s):
        out = torch.zeros_like(v)
        seg.axpy(np.asarray(s, dtype=complex)[None, :], [v], out)
        return out

    return pack, apply, scaled


def lanczos_expv_packed(packed: PackedApply, tau: complex, v0s: Sequence[BlockTensor], *,
                        tol: float = 1e-15, maxiter: int = ROOT_MAXITER,
                        reorth: bool = True) -> Tuple[List[BlockTensor], List[int]]:
    """exp(tau H_i) v0_i for every instance of ``packed`` (its shifts included in H_i); returns
    the results (new tensors) and each instance's Krylov dimension."""
    seg = Segments(packed)
    K = seg.nseg
    if len(v0s) != K:
        raise ValueError("one start vector per instance")
    pack, apply, scaled = _pack_helpers(packed, seg)
    v0 = pack(v0s)
    beta0 = np.sqrt(np.maximum(seg.dots([v0], v0)[0].real, 0.0))
    active = beta0 > 0
    alpha = [[] for _ in range(K)]
    beta = [[] for _ in range(K)]
    basis = [scaled(v0, np.where(active, 1.0 / np.where(active, beta0, 1.0), 0.0))]
    r = apply(basis[0])
    a = seg.dots([basis[0]], r)[0].real
    for i in np.flatnonzero(active):
        alpha[i].append(float(a[i]))
    seg.axpy(-np.where(active, a, 0.0)[None, :], [basis[0]], r)
    for _ in range(2, maxiter + 1):
        b = np.sqrt(np.maximum(seg.dots([r], r)[0].real, 0.0))
        active &= b >= tol
        if not active.any():
            break
        for i in np.flatnonzero(active):
            beta[i].append(float(b[i]))
        v = scaled(r, np.where(active, 1.0 / np.where(active, b, 1.0), 0.0))
        basis.append(v)
        r = apply(v)
        a = seg.dots([v], r)[0].real
        for i in np.flatnonzero(active):
            alpha[i].append(float(a[i]))
        seg.axpy(np.stack([-np.where(active, a, 0.0), -np.where(active, b, 0.0)]), [v, basis[-2]], r)
        if reorth:
            for _p in range(2):
                cs = seg.dots(basis, r)
                cs[:, ~active] = 0.0
                seg.axpy(-cs, basis, r)
    coef = np.zeros((len(basis), K), dtype=complex)
    dims = []
    for i in range(K):
        dims.append(len(alpha[i]))
        if beta0[i] == 0 or not alpha[i]:
            continue
        c, ls = Tridiagonal(alpha[i], beta[i]).coeffs(tau)
        sc = np.exp(ls)
        scale = beta0[i]
        if np.isfinite(abs(sc)) and sc != 0 and abs(sc) * beta0[i] * np.linalg.norm(c) <= MAX_NORM:
            scale = beta0[i] * sc
        coef[: len(c), i] = c * scale
    out = torch.zeros_like(v0)
    seg.axpy(coef, basis, out)
    res = []
    for t, v, i in zip(packed.x, v0s, range(K)):
        off = (t.data.data_ptr() - packed._xbuf.data_ptr()) // 16
        y = BlockTensor(t.layout, out[off: off + t.layout.nelem].clone())
        res.append(v.clone() if beta0[i] == 0 else y)
    return res, dims


def arnoldi_expv_packed(packed: PackedApply, tau: complex, v0s: Sequence[BlockTensor], *,
                        tol: float = 1e-13, maxiter: int = ROOT_MAXITER
                        ) -> Tuple[List[BlockTensor], List[int]]:
    """exp(tau A_i) v0_i for every instance of ``packed`` by one Arnoldi pass each, in lockstep:
    per iteration one packed apply, two segmented Gram-Schmidt passes against the whole basis
    (one multi-dot launch and one host read each) and segmented multi-axpys; an instance stops
    growing its basis at its own breakdown (residual below ``tol``).  Returns the results (new
    tensors) and each instance's Krylov dimension, as ``arnoldi_expv`` with substeps = 1 gives
    them one instance at a time."""
    seg = Segments(packed)
    K = seg.nseg
    if len(v0s) != K:
        raise ValueError("one start vector per instance")
    pack, apply, scaled = _pack_helpers(packed, seg)
    v0 = pack(v0s)
    beta0 = np.sqrt(np.maximum(seg.dots([v0], v0)[0].real, 0.0))
    active = beta0 > 0
    H = np.zeros((K, maxiter + 1, maxiter), dtype=complex)
    dims = [0] * K
    basis = [scaled(v0, np.where(active, 1.0 / np.where(active, beta0, 1.0), 0.0))]
    for k in range(1, maxiter + 1):
        r = apply(basis[k - 1])
        for _p in range(2):
            cs = seg.dots(basis, r)
            cs[:, ~active] = 0.0
            H[:, :len(basis), k - 1] += cs.T
            seg.axpy(-cs, basis, r)
        for i in np.flatnonzero(active):
            dims[i] = k
        b = np.sqrt(np.maximum(seg.dots([r], r)[0].real, 0.0))
        active &= b >= tol
        if k == maxiter or not active.any():
            break
        for i in np.flatnonzero(active):
            H[i, k, k - 1] = b[i]
        basis.append(scaled(r, np.where(active, 1.0 / np.where(active, b, 1.0), 0.0)))
    coef = np.zeros((len(basis), K), dtype=complex)
    for i in range(K):
        d = dims[i]
        if beta0[i] == 0 or d == 0:
            continue
        c, ls = Hessenberg(H[i, :d, :d].copy()).coeffs(tau)
        sc = np.exp(ls)
        scale = beta0[i]
        if np.isfinite(abs(sc)) and sc != 0 and abs(sc) * beta0[i] * np.linalg.norm(c) <= MAX_NORM:
            scale = beta0[i] * sc
        coef[:d, i] = c[:d] * scale
    out = torch.zeros_like(v0)
    seg.axpy(coef, basis, out)
    res = []
    for t, v, i in zip(packed.x, v0s, range(K)):
        off = (t.data.data_ptr() - packed._xbuf.data_ptr()) // 16
        y = BlockTensor(t.layout, out[off: off + t.layout.nelem].clone())
        res.append(v.clone() if beta0[i] == 0 else y)
    return res, dims
