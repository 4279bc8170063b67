"""Device tensor algebra on block-sparse families: products, projections, direct sums, QR, SVD.

Replaces the Telum operations the step uses (``contract``, ``oplus``, ``getsub``, ``svd``;
deps/Telum/src/contract.jl, tlarray/oplus.jl, tlarray/getsub.jl, svd.jl) with launches of
``libbug.so``.  Conventions (the same ones the step's checker uses):

* left matricisation of charge ``qr`` stacks the slices (s = 0 first), each contributing its
  block at row charge ``qr - q(s)``; right matricisation of ``ql`` puts the slice blocks side by
  side.
* gauge moves are the unique thin QR (real non-negative R diagonal); an all-zero sector block is
  dropped, as the cutoff-0 SVD's ``sv > 0`` rule drops it.
* SVD selection is Telum's (svd.jl:1813-1921): keep ``sv > cutoff * sv_max`` over all sectors,
  then the largest ``nkeep`` by a stable descending sort in sector order; each singular pair's
  phase makes the first entry of u with ``|u_i| >= 0.5 max|u|`` real positive.

Small decisions (the global singular-value selection) are made on the host from the downloaded
spectra; everything with the size of a tensor stays on the device.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .blocks import BlockTensor, Layout, Leg, bsm_product, current_stream_handle, device

OP_N, OP_T, OP_C, OP_R = _lib.OP_N, _lib.OP_T, _lib.OP_C, _lib.OP_R
CBE_EPS = 100 * np.finfo(float).eps
# A sector block (or singular value) below this fraction of the tensor's norm (or of the largest
# singular value) is roundoff: gauge moves drop such blocks and selections such values, so the
# block structure of a result never depends on how the arithmetic rounded.
ZERO_REL = 1e-14


def _ctx():
    return _lib.context()


def _stream() -> int:
    return current_stream_handle()


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def zeros(layout: Layout) -> BlockTensor:
    return BlockTensor(layout)


def empty_buffer(n: int) -> torch.Tensor:
    return torch.empty(max(int(n), 1), dtype=torch.complex128, device=device())


# ---- products ------------------------------------------------------------------------------

def product(row: Leg, col: Leg, shifts, P: BlockTensor, Q: BlockTensor, terms,
            out: Optional[BlockTensor] = None, precision: int = _lib.PREC_FP64) -> BlockTensor:
    """out[o] (+)= sum alpha op_a(P[p]) op_b(Q[q]); a fresh tensor unless ``out`` is given."""
    lay = Layout(row, col, shifts) if out is None else out.layout
    return bsm_product(lay, P, Q, terms, accumulate_into=out, precision=precision)


def core_times_site(C: BlockTensor, A: BlockTensor) -> BlockTensor:
    d = A.layout.nslice
    return product(C.layout.row, A.layout.col, A.layout.shifts, C, A,
                   [(s, 0, s, OP_N, OP_N, 1.0) for s in range(d)])


def site_times_core(A: BlockTensor, C: BlockTensor) -> BlockTensor:
    d = A.layout.nslice
    return product(A.layout.row, C.layout.col, A.layout.shifts, A, C,
                   [(s, s, 0, OP_N, OP_N, 1.0) for s in range(d)])


def overlap_left(U: BlockTensor, A: BlockTensor) -> BlockTensor:
    """U^dag A over (link, site): (bond_U, bond_A)."""
    d = A.layout.nslice
    return product(U.layout.col, A.layout.col, [0], U, A,
                   [(0, s, s, OP_C, OP_N, 1.0) for s in range(d)])


def overlap_right(A: BlockTensor, V: BlockTensor) -> BlockTensor:
    """A V^dag over (site, link): (bond_A, bond_V)."""
    d = A.layout.nslice
    return product(A.layout.row, V.layout.row, [0], A, V,
                   [(0, s, s, OP_N, OP_C, 1.0) for s in range(d)])


def core_times_core(A: BlockTensor, B: BlockTensor) -> BlockTensor:
    return product(A.layout.row, B.layout.col, [0], A, B, [(0, 0, 0, OP_N, OP_N, 1.0)])


def perp_left(U: BlockTensor, K: BlockTensor) -> BlockTensor:
    """K - U (U^dag K)  (perp_component, src/BUG/Core/TensorNetworks/sectors.jl:215-221)."""
    ov = overlap_left(U, K)
    out = K.clone()
    d = K.layout.nslice
    return bsm_product(K.layout, U, ov, [(s, s, 0, OP_N, OP_N, -1.0) for s in range(d)],
                       accumulate_into=out)


def perp_right(V: BlockTensor, K: BlockTensor) -> BlockTensor:
    """K - (K V^dag) V  (perp_component_right, sectors.jl:230-238)."""
    ov = overlap_right(K, V)
    out = K.clone()
    d = K.layout.nslice
    return bsm_product(K.layout, ov, V, [(s, 0, s, OP_N, OP_N, -1.0) for s in range(d)],
                       accumulate_into=out)


# ---- vector algebra (K10) -------------------------------------------------------------------

def _same(a: BlockTensor, b: BlockTensor):
    if a.layout != b.layout:
        raise ValueError("vector operation on tensors of different block layouts")


def inner(a: BlockTensor, b: BlockTensor) -> complex:
    """<a|b> (tensor_inner, symmetric_mps.jl:83-93)."""
    _same(a, b)
    return inner_many([a], b)[0]


def inner_many(xs: Sequence[BlockTensor], y: BlockTensor) -> List[complex]:
    """[<x_j|y>] for up to 64 vectors in one fused launch."""
    out: List[complex] = []
    ctx = _ctx()
    for c0 in range(0, len(xs), 64):
        chunk = xs[c0:c0 + 64]
        for x in chunk:
            _same(x, y)
        arr = (ctypes.c_void_p * len(chunk))(*[_ptr(x.data) for x in chunk])
        res = np.zeros(2 * len(chunk))
        ctx.check(ctx._lib.bug_dotc(ctx.handle, y.layout.nelem, len(chunk), arr, _ptr(y.data),
                                    res.ctypes.data_as(_lib._f64p), _stream()))
        out += [complex(res[2 * j], res[2 * j + 1]) for j in range(len(chunk))]
    return out


def norm(a: BlockTensor) -> float:
    return math.sqrt(max(inner(a, a).real, 0.0))


def axpy_(alpha: complex, x: BlockTensor, y: BlockTensor) -> BlockTensor:
    """y += alpha x in place."""
    _same(x, y)
    alpha = complex(alpha)
    ctx = _ctx()
    ctx.check(ctx._lib.bug_axpy(ctx.handle, y.layout.nelem, alpha.real, alpha.imag, _ptr(x.data),
                                _ptr(y.data), _stream()))
    return y


def axpy_many_(alphas: Sequence[complex], xs: Sequence[BlockTensor], y: BlockTensor) -> BlockTensor:
    """y += sum_j alphas[j] xs[j] in place, the terms added in order (64 per launch), with the
    per-element arithmetic of that many ``axpy_`` calls."""
    if len(alphas) != len(xs):
        raise ValueError("axpy_many_: as many coefficients as vectors")
    for x in xs:
        _same(x, y)
    ctx = _ctx()
    for lo in range(0, len(xs), 64):
        chunk = xs[lo: lo + 64]
        k = len(chunk)
        coef = np.empty(2 * k)
        for j, a in enumerate(alphas[lo: lo + 64]):
            a = complex(a)
            coef[2 * j], coef[2 * j + 1] = a.real, a.imag
        ptrs = (ctypes.c_void_p * k)(*[_ptr(x.data) for x in chunk])
        ctx.check(ctx._lib.bug_axpy_many(ctx.handle, y.layout.nelem, k,
                                         coef.ctypes.data_as(_lib._f64p), ptrs, _ptr(y.data),
                                         _stream()))
    return y


def axpy(alpha: complex, x: BlockTensor, y: BlockTensor) -> BlockTensor:
    """y + alpha x (new tensor)."""
    return axpy_(alpha, x, y.clone())


def scale_(alpha: complex, x: BlockTensor) -> BlockTensor:
    alpha = complex(alpha)
    ctx = _ctx()
    ctx.check(ctx._lib.bug_scal(ctx.handle, x.layout.nelem, alpha.real, alpha.imag, _ptr(x.data),
                                _stream()))
    return x


def scale(alpha: complex, x: BlockTensor) -> BlockTensor:
    return scale_(alpha, x.clone())


# ---- block movement ----------------------------------------------------------------------

class CopyList:
    """Accumulates block copies for one launch of ``bug_copy_blocks``."""

    def __init__(self):
        self.ops: List[tuple] = []

    def add(self, src_off, src_ld, dst_off, dst_ld, rows, cols, mode=0, scale_mode=0, scale_off=0):
        if rows > 0 and cols > 0:
            self.ops.append((int(src_off), int(dst_off), int(src_ld), int(dst_ld), int(rows),
                             int(cols), int(mode), int(scale_mode), int(scale_off)))

    def run(self, src: torch.Tensor, dst: torch.Tensor, scale: Optional[torch.Tensor] = None,
            accumulate: bool = False):
        if not self.ops:
            return
        arr = (_lib.BugCopy * len(self.ops))()
        for i, o in enumerate(self.ops):
            arr[i] = _lib.BugCopy(o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[8])
        ctx = _ctx()
        sp = None if scale is None else ctypes.cast(scale.data_ptr(), _lib._f64p)
        ctx.check(ctx._lib.bug_copy_blocks(ctx.handle, len(self.ops), arr, _ptr(src), src.numel(),
                                           _ptr(dst), dst.numel(), sp, 1 if accumulate else 0,
                                           _stream()))


def block_sqnorms(blocks: List[Tuple[int, int, int, int]], src: torch.Tensor) -> np.ndarray:
    """Squared Frobenius norms of (off, ld, rows, cols) blocks of ``src`` (host array)."""
    out = np.zeros(len(blocks))
    if not blocks:
        return out
    arr = (_lib.BugCopy * len(blocks))()
    for i, (off, ld, r, c) in enumerate(blocks):
        arr[i] = _lib.BugCopy(int(off), 0, int(max(ld, 1)), 1, int(r), int(c), 0, 0, 0)
    ctx = _ctx()
    ctx.check(ctx._lib.bug_block_sqnorms(ctx.handle, len(blocks), arr, _ptr(src), src.numel(),
                                         out.ctypes.data_as(_lib._f64p), _stream()))
    return out


def _block_table(layout: Layout):
    """{(slice, row charge): (offset, rows, cols, col charge)}."""
    return {(k, q): (off, d, dc, qc) for k, q, qc, d, dc, off in layout.blocks()}


def left_sectors(layout: Layout) -> Dict[int, List[Tuple[int, int, int, int]]]:
    """qr -> [(s, ql, rows, offset)] in stacking order."""
    tab = _block_table(layout)
    out = {}
    for qr in layout.col.charges:
        parts = []
        for s, sh in enumerate(layout.shifts):
            e = tab.get((s, qr - sh))
            if e is not None:
                parts.append((s, qr - sh, e[1], e[0]))
        if parts:
            out[qr] = parts
    return out


def right_sectors(layout: Layout) -> Dict[int, List[Tuple[int, int, int, int]]]:
    """ql -> [(s, qr, cols, offset)] in stacking order."""
    tab = _block_table(layout)
    out = {}
    for ql in layout.row.charges:
        parts = []
        for s, sh in enumerate(layout.shifts):
            e = tab.get((s, ql))
            if e is not None:
                parts.append((s, ql + sh, e[2], e[0]))
        if parts:
            out[ql] = parts
    return out


def oplus_left(F: BlockTensor, N: BlockTensor) -> BlockTensor:
    """[F | N] on the column (bond) leg, per sector."""
    fc, nc = F.layout.col.as_dict(), N.layout.col.as_dict()
    col = Leg.make(list(fc.items()) + list(nc.items()))
    out = zeros(Layout(F.layout.row, col, F.layout.shifts))
    ot = _block_table(out.layout)
    cf, cn = CopyList(), CopyList()
    for (s, q), (off, d, dc, qc) in _block_table(F.layout).items():
        o = ot[(s, q)]
        cf.add(off, d, o[0], o[1], d, dc)
    for (s, q), (off, d, dc, qc) in _block_table(N.layout).items():
        o = ot[(s, q)]
        cn.add(off, d, o[0] + o[1] * fc.get(qc, 0), o[1], d, dc)
    cf.run(F.data, out.data)
    cn.run(N.data, out.data)
    return out


def oplus_right(F: BlockTensor, N: BlockTensor) -> BlockTensor:
    """[F; N] on the row (bond) leg, per sector."""
    fr, nr = F.layout.row.as_dict(), N.layout.row.as_dict()
    row = Leg.make(list(fr.items()) + list(nr.items()))
    out = zeros(Layout(row, F.layout.col, F.layout.shifts))
    ot = _block_table(out.layout)
    cf, cn = CopyList(), CopyList()
    for (s, q), (off, d, dc, qc) in _block_table(F.layout).items():
        o = ot[(s, q)]
        cf.add(off, d, o[0], o[1], d, dc)
    for (s, q), (off, d, dc, qc) in _block_table(N.layout).items():
        o = ot[(s, q)]
        cn.add(off, d, o[0] + fr.get(q, 0), o[1], d, dc)
    cf.run(F.data, out.data)
    cn.run(N.data, out.data)
    return out


def trim(T: BlockTensor, side: str, counts: Dict[int, int]) -> Optional[BlockTensor]:
    """Keep the leading ``counts[q]`` columns (side 'left') or rows ('right') of sector q."""
    if not any(c > 0 for c in counts.values()):
        return None
    if side == "left":
        col = Leg.make((q, c) for q, c in counts.items() if c > 0)
        lay = Layout(T.layout.row, col, T.layout.shifts)
    else:
        row = Leg.make((q, c) for q, c in counts.items() if c > 0)
        lay = Layout(row, T.layout.col, T.layout.shifts)
    out = zeros(lay)
    src = _block_table(T.layout)
    cl = CopyList()
    for (s, q), (off, d, dc, qc) in _block_table(lay).items():
        e = src[(s, q)]
        cl.add(e[0], e[1], off, d, d, dc)
    cl.run(T.data, out.data)
    return out


def diag_scale(T: BlockTensor, S: Dict[int, np.ndarray], side: str) -> BlockTensor:
    """S T (scale rows by S[row charge]) for side 'rows', T S (columns by S[col charge]) for
    side 'cols'."""
    keys = sorted(S)
    offs, flat, pos = {}, [], 0
    for q in keys:
        offs[q] = pos
        flat.append(np.asarray(S[q], dtype=np.float64))
        pos += len(S[q])
    sbuf = torch.from_numpy(np.concatenate(flat) if flat else np.zeros(1)).to(device())
    out = zeros(T.layout)
    cl = CopyList()
    for (s, q), (off, d, dc, qc) in _block_table(T.layout).items():
        if side == "rows":
            cl.add(off, d, off, d, d, dc, 0, 1, offs[q])
        else:
            cl.add(off, d, off, d, d, dc, 0, 2, offs[qc])
    cl.run(T.data, out.data, scale=sbuf)
    return out


# ---- QR -----------------------------------------------------------------------------------

def _nonzero(sqnorms) -> List[int]:
    """Indices of the sector blocks above the roundoff floor (squared norms in)."""
    floor = ZERO_REL * ZERO_REL * float(sum(sqnorms))
    return [i for i, v in enumerate(sqnorms) if v > floor]


def _run_qr(mats: List[Tuple[int, int]], ws: torch.Tensor, offs: List[int]):
    """QR of the listed (m, n) matrices stored at offs in ws -> (Qbuf, Rbuf, qoffs, roffs, ks)."""
    qoffs, roffs, ks = [], [], []
    qp = rp = 0
    for (m, n) in mats:
        k = min(m, n)
        qoffs.append(qp)
        roffs.append(rp)
        ks.append(k)
        qp += m * k
        rp += k * n
    Qb, Rb = empty_buffer(qp), empty_buffer(rp)
    if mats:
        arr = (_lib.BugQrJob * len(mats))()
        for i, (m, n) in enumerate(mats):
            arr[i] = _lib.BugQrJob(offs[i], qoffs[i], roffs[i], m, n, max(m, 1), max(m, 1),
                                   max(ks[i], 1), 0)
        ctx = _ctx()
        ctx.check(ctx._lib.bug_block_qr(ctx.handle, len(mats), arr, _ptr(ws), ws.numel(),
                                        _ptr(Qb), Qb.numel(), _ptr(Rb), Rb.numel(), _stream()))
    return Qb, Rb, qoffs, roffs, ks


def qr_left(A: BlockTensor) -> Tuple[BlockTensor, BlockTensor]:
    """A = Q R over (link, site) | bond: Q left isometry (new bond), R (bond, A.col)."""
    lay = A.layout
    secs = left_sectors(lay)
    cold = lay.col.as_dict()
    gather, offs, mats, pos = CopyList(), [], [], 0
    qlist = list(secs.keys())
    for qr in qlist:
        parts = secs[qr]
        m = sum(p[2] for p in parts)
        n = cold[qr]
        r0 = 0
        for s, ql, rows, off in parts:
            gather.add(off, rows, pos + r0, m, rows, n)
            r0 += rows
        offs.append(pos)
        mats.append((m, n))
        pos += m * n
    ws = empty_buffer(pos)
    gather.run(A.data, ws)
    nz = block_sqnorms([(offs[i], mats[i][0], mats[i][0], mats[i][1]) for i in range(len(mats))], ws)
    keep = _nonzero(nz)
    Qb, Rb, qo, ro, ks = _run_qr([mats[i] for i in keep], ws, [offs[i] for i in keep])
    bond = Leg.make((qlist[i], ks[j]) for j, i in enumerate(keep))
    Q = zeros(Layout(lay.row, bond, lay.shifts))
    R = zeros(Layout(bond, lay.col, [0]))
    qt, rt = _block_table(Q.layout), _block_table(R.layout)
    cq, cr = CopyList(), CopyList()
    for j, i in enumerate(keep):
        qr = qlist[i]
        m, n = mats[i]
        k = ks[j]
        r0 = 0
        for s, ql, rows, off in secs[qr]:
            o = qt[(s, ql)]
            cq.add(qo[j] + r0, m, o[0], o[1], rows, k)
            r0 += rows
        o = rt[(0, qr)]
        cr.add(ro[j], k, o[0], o[1], k, n)
    cq.run(Qb, Q.data)
    cr.run(Rb, R.data)
    return Q, R


def qr_right(A: BlockTensor) -> Tuple[BlockTensor, BlockTensor]:
    """A = L Q over link | (site, link): L (A.row, bond), Q right isometry (bond, s, r)."""
    lay = A.layout
    secs = right_sectors(lay)
    rowd = lay.row.as_dict()
    gather, offs, mats, pos = CopyList(), [], [], 0
    qlist = list(secs.keys())
    for ql in qlist:
        parts = secs[ql]
        ncol = sum(p[2] for p in parts)
        rows = rowd[ql]
        c0 = 0
        for s, qr, cols, off in parts:
            # M^H rows c0..c0+cols = (A block)^H
            gather.add(off, rows, pos + c0, ncol, cols, rows, mode=2)
            c0 += cols
        offs.append(pos)
        mats.append((ncol, rows))
        pos += ncol * rows
    ws = empty_buffer(pos)
    gather.run(A.data, ws)
    nz = block_sqnorms([(offs[i], mats[i][0], mats[i][0], mats[i][1]) for i in range(len(mats))], ws)
    keep = _nonzero(nz)
    Qb, Rb, qo, ro, ks = _run_qr([mats[i] for i in keep], ws, [offs[i] for i in keep])
    bond = Leg.make((qlist[i], ks[j]) for j, i in enumerate(keep))
    Lc = zeros(Layout(lay.row, bond, [0]))
    Q = zeros(Layout(bond, lay.col, lay.shifts))
    lt, qt = _block_table(Lc.layout), _block_table(Q.layout)
    cl, cq = CopyList(), CopyList()
    for j, i in enumerate(keep):
        ql = qlist[i]
        mq, nq = mats[i]            # M^H is (mq = sum cols) x (nq = rows)
        k = ks[j]
        o = lt[(0, ql)]
        cl.add(ro[j], k, o[0], o[1], nq, k, mode=2)          # L = R'^H (rows x k)
        c0 = 0
        for s, qr, cols, off in secs[ql]:
            o = qt[(s, ql)]
            cq.add(qo[j] + c0, mq, o[0], o[1], k, cols, mode=2)   # Q block = Q'[c0:,:]^H
            c0 += cols
    cl.run(Rb, Lc.data)
    cq.run(Qb, Q.data)
    return Lc, Q


def _run_qrp(mats: List[Tuple[int, int]], ws: torch.Tensor, offs: List[int]):
    """Pivoted QR of the listed (m, n) matrices at offs in ws -> (Qbuf, qoffs, ks, |R| diagonals
    on the host)."""
    qoffs, doffs, ks = [], [], []
    qp = dp = 0
    for (m, n) in mats:
        k = min(m, n)
        qoffs.append(qp)
        doffs.append(dp)
        ks.append(k)
        qp += m * k
        dp += k
    Qb = empty_buffer(qp)
    rd = torch.zeros(max(dp, 1), dtype=torch.float64, device=device())
    if mats:
        arr = (_lib.BugQrJob * len(mats))()
        for i, (m, n) in enumerate(mats):
            arr[i] = _lib.BugQrJob(offs[i], qoffs[i], doffs[i], m, n, max(m, 1), max(m, 1), 1, 0)
        ctx = _ctx()
        ctx.check(ctx._lib.bug_block_qrp(ctx.handle, len(mats), arr, _ptr(ws), ws.numel(),
                                         _ptr(Qb), Qb.numel(), ctypes.cast(rd.data_ptr(), _lib._f64p),
                                         rd.numel(), _stream()))
    host = rd.cpu().numpy()
    return Qb, qoffs, ks, [host[doffs[i]: doffs[i] + ks[i]] for i in range(len(mats))]


def thin_qr(A: BlockTensor, side: str, thr: float) -> Optional[BlockTensor]:
    """Orthonormal basis of the span of the new-leg columns of ``A`` (``_thin_qr``,
    src/BUG/Randomised/dd_cbe.jl:184-259): per sector of that leg, the column-pivoted QR of the
    stacked blocks (left: the left matricisation; right: the conjugate transpose of the right
    one), the first columns kept as many as pivots exceed ``thr``.  None when nothing is kept."""
    lay = A.layout
    gather, offs, mats, pos = CopyList(), [], [], 0
    if side == "left":
        secs = left_sectors(lay)
        cold = lay.col.as_dict()
        qlist = list(secs.keys())
        for qc in qlist:
            parts = secs[qc]
            m, n = sum(p[2] for p in parts), cold[qc]
            r0 = 0
            for s, ql, rows, off in parts:
                gather.add(off, rows, pos + r0, m, rows, n)
                r0 += rows
            offs.append(pos)
            mats.append((m, n))
            pos += m * n
    else:
        secs = right_sectors(lay)
        rowd = lay.row.as_dict()
        qlist = list(secs.keys())
        for ql in qlist:
            parts = secs[ql]
            m, n = sum(p[2] for p in parts), rowd[ql]
            c0 = 0
            for s, qr, cols, off in parts:
                gather.add(off, n, pos + c0, m, cols, n, mode=2)
                c0 += cols
            offs.append(pos)
            mats.append((m, n))
            pos += m * n
    if not mats:
        return None
    ws = empty_buffer(pos)
    gather.run(A.data, ws)
    Qb, qo, ks, rdiag = _run_qrp(mats, ws, offs)
    kept = [int(np.sum(r > thr)) for r in rdiag]
    if not any(kept):
        return None
    bond = Leg.make((qlist[i], kept[i]) for i in range(len(mats)) if kept[i] > 0)
    if side == "left":
        out = zeros(Layout(lay.row, bond, lay.shifts))
        tab = _block_table(out.layout)
        cq = CopyList()
        for i, qc in enumerate(qlist):
            if kept[i] == 0:
                continue
            m = mats[i][0]
            r0 = 0
            for s, ql, rows, off in secs[qc]:
                o = tab[(s, ql)]
                cq.add(qo[i] + r0, m, o[0], o[1], rows, kept[i])
                r0 += rows
        cq.run(Qb, out.data)
        return out
    out = zeros(Layout(bond, lay.col, lay.shifts))
    tab = _block_table(out.layout)
    cq = CopyList()
    for i, ql in enumerate(qlist):
        if kept[i] == 0:
            continue
        m = mats[i][0]
        c0 = 0
        for s, qr, cols, off in secs[ql]:
            o = tab[(s, ql)]
            cq.add(qo[i] + c0, m, o[0], o[1], kept[i], cols, mode=2)
            c0 += cols
    cq.run(Qb, out.data)
    return out


def qr_core(C: BlockTensor) -> Tuple[BlockTensor, BlockTensor]:
    """C = Q R per sector of a one-slice core."""
    lay = C.layout
    tab = _block_table(lay)
    keys = list(tab.keys())
    nz = block_sqnorms([(tab[k][0], tab[k][1], tab[k][1], tab[k][2]) for k in keys], C.data)
    keep = _nonzero(nz)
    mats = [(tab[keys[i]][1], tab[keys[i]][2]) for i in keep]
    ws = C.data.clone()
    Qb, Rb, qo, ro, ks = _run_qr(mats, ws, [tab[keys[i]][0] for i in keep])
    bond = Leg.make((keys[i][1], ks[j]) for j, i in enumerate(keep))
    Q = zeros(Layout(lay.row, bond, [0]))
    R = zeros(Layout(bond, lay.col, [0]))
    qt, rt = _block_table(Q.layout), _block_table(R.layout)
    cq, cr = CopyList(), CopyList()
    for j, i in enumerate(keep):
        q = keys[i][1]
        m, n = mats[j]
        k = ks[j]
        o = qt[(0, q)]
        cq.add(qo[j], m, o[0], o[1], m, k)
        o = rt[(0, q)]
        cr.add(ro[j], k, o[0], o[1], k, n)
    cq.run(Qb, Q.data)
    cr.run(Rb, R.data)
    return Q, R


# ---- SVD ----------------------------------------------------------------------------------

class SVDResult:
    __slots__ = ("U", "S", "Vh", "kept", "trunc")

    def __init__(self, U, S, Vh, kept, trunc):
        self.U, self.S, self.Vh, self.kept, self.trunc = U, S, Vh, kept, trunc


def select(spectra: Dict[int, np.ndarray], cutoff: float, nkeep: Optional[int]):
    """Telum's global selection (svd.jl:1813-1921) -> (keep per sector, kept, trunc)."""
    entries = []
    smax = 0.0
    for q, sv in spectra.items():
        for j, s in enumerate(sv):
            smax = max(smax, float(s))
            entries.append((float(s), q, j))
    full = sorted(entries, key=lambda e: -e[0])
    thresh = max(cutoff, ZERO_REL) * smax
    sel = [e for e in full if e[0] > thresh]
    if nkeep is not None:
        sel = sel[:max(nkeep, 0)]
    chosen = {(q, j) for _, q, j in sel}
    keep = {q: 0 for q in spectra}
    for _, q, j in sel:
        keep[q] += 1
    kept = [e for e in full if (e[1], e[2]) in chosen]
    trunc = [e for e in full if (e[1], e[2]) not in chosen]
    return keep, kept, trunc


class _SvdBatch:
    """Per-sector SVDs of matrices M_q (m x n) assembled by the caller's gather list."""

    def __init__(self, mats: List[Tuple[int, int]]):
        self.mats = mats
        self.tall = [m >= n for (m, n) in mats]
        self.goffs, pos = [], 0
        for (m, n) in mats:
            self.goffs.append(pos)
            pos += m * n
        self.ws = empty_buffer(pos)

    def gather_offset(self, i: int):
        return self.goffs[i]

    def run(self):
        jobs, uo, so, vo = [], [], [], []
        up = sp = vp = 0
        for i, (m, n) in enumerate(self.mats):
            mj, nj = (m, n) if self.tall[i] else (n, m)
            uo.append(up)
            so.append(sp)
            vo.append(vp)
            jobs.append((self.goffs[i], up, sp, vp, mj, nj, max(mj, 1), max(mj, 1), max(nj, 1),
                         0 if self.tall[i] else 1))
            up += mj * nj
            sp += nj
            vp += nj * nj
        self.U, self.Vh = empty_buffer(up), empty_buffer(vp)
        self.S = torch.empty(max(sp, 1), dtype=torch.float64, device=device())
        self.uo, self.so, self.vo = uo, so, vo
        if jobs:
            arr = (_lib.BugSvdJob * len(jobs))()
            for i, j in enumerate(jobs):
                arr[i] = _lib.BugSvdJob(*j)
            ctx = _ctx()
            ctx.check(ctx._lib.bug_block_svd(
                ctx.handle, len(jobs), arr, _ptr(self.ws), self.ws.numel(), _ptr(self.U),
                self.U.numel(), ctypes.cast(self.S.data_ptr(), _lib._f64p), self.S.numel(),
                _ptr(self.Vh), self.Vh.numel(), 0, _stream()))
        host = self.S.cpu().numpy()
        if not np.isfinite(host).all():
            raise FloatingPointError("SVD of a block with non-finite entries: the state or the "
                                     "operator is not finite")
        self.spectra = []
        for i, (m, n) in enumerate(self.mats):
            k = min(m, n)
            self.spectra.append(host[so[i]: so[i] + k].copy())

    def add_u_rows(self, cl: CopyList, i: int, r0: int, rows: int, k: int, dst_off: int, dst_ld: int):
        """Copy rows r0..r0+rows of the first k columns of U_i."""
        m, n = self.mats[i]
        if self.tall[i]:
            cl.add(self.uo[i] + r0, m, dst_off, dst_ld, rows, k)
        else:                        # U = Vh'^H (Vh' is m x m)
            cl.add(self.vo[i] + r0 * m, m, dst_off, dst_ld, rows, k, mode=2)

    def add_vh_cols(self, cl: CopyList, i: int, c0: int, cols: int, k: int, dst_off: int, dst_ld: int):
        """Copy columns c0..c0+cols of the first k rows of Vh_i."""
        m, n = self.mats[i]
        if self.tall[i]:             # Vh is n x n
            cl.add(self.vo[i] + c0 * n, n, dst_off, dst_ld, k, cols)
        else:                        # Vh = U'^H, U' is n x m
            cl.add(self.uo[i] + c0, n, dst_off, dst_ld, k, cols, mode=2)


def _gather_plain(cl: CopyList, batch: _SvdBatch, i: int, off: int, ld: int, r0: int, c0: int,
                  rows: int, cols: int):
    """Place block (rows x cols, ld) at (r0, c0) of matrix i (stored as M or M^H)."""
    m, n = batch.mats[i]
    g = batch.gather_offset(i)
    if batch.tall[i]:
        cl.add(off, ld, g + r0 + c0 * m, m, rows, cols)
    else:
        cl.add(off, ld, g + c0 + r0 * n, n, cols, rows, mode=2)


class BlockSVD:
    """Per-sector SVD factors of one matricisation, kept on the device.

    Any number of selections (``select``) can be read from one factorization: a truncated SVD is
    a prefix of the full one, so the reference's repeated ``svd(...; Nkeep = k)`` calls on the
    same tensor (rank bisections, tail-rank re-cuts) cost one factorization here."""

    def __init__(self, kind: str, layout: Layout, qlist, secs, mats, batch: _SvdBatch):
        self.kind, self.layout, self.qlist, self.secs = kind, layout, qlist, secs
        self.mats, self.batch = mats, batch
        self.spectra = {q: batch.spectra[i] for i, q in enumerate(qlist)}

    def select(self, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
        keep, kept, trunc = select(self.spectra, cutoff, nkeep)
        lay, batch = self.layout, self.batch
        bond = Leg.make((q, keep[q]) for q in self.qlist if keep[q] > 0)
        if self.kind == "left":
            U = zeros(Layout(lay.row, bond, lay.shifts))
            Vh = zeros(Layout(bond, lay.col, [0]))
        elif self.kind == "right":
            U = zeros(Layout(lay.row, bond, [0]))
            Vh = zeros(Layout(bond, lay.col, lay.shifts))
        else:
            U = zeros(Layout(lay.row, bond, [0]))
            Vh = zeros(Layout(bond, lay.col, [0]))
        ut, vt = _block_table(U.layout), _block_table(Vh.layout)
        cu_u, cu_v, cv_u, cv_v = CopyList(), CopyList(), CopyList(), CopyList()
        S = {}
        for i, q in enumerate(self.qlist):
            k = keep[q]
            if k == 0:
                continue
            S[q] = batch.spectra[i][:k]
            tall = batch.tall[i]
            m, n = self.mats[i]
            if self.kind == "left":
                r0 = 0
                for s, ql, rows, off in self.secs[q]:
                    o = ut[(s, ql)]
                    batch.add_u_rows(cu_u if tall else cu_v, i, r0, rows, k, o[0], o[1])
                    r0 += rows
                o = vt[(0, q)]
                batch.add_vh_cols(cv_v if tall else cv_u, i, 0, n, k, o[0], o[1])
            elif self.kind == "right":
                o = ut[(0, q)]
                batch.add_u_rows(cu_u if tall else cu_v, i, 0, m, k, o[0], o[1])
                c0 = 0
                for s, qr, cols, off in self.secs[q]:
                    o = vt[(s, q)]
                    batch.add_vh_cols(cv_v if tall else cv_u, i, c0, cols, k, o[0], o[1])
                    c0 += cols
            else:
                o = ut[(0, q)]
                batch.add_u_rows(cu_u if tall else cu_v, i, 0, m, k, o[0], o[1])
                o = vt[(0, q)]
                batch.add_vh_cols(cv_v if tall else cv_u, i, 0, n, k, o[0], o[1])
        cu_u.run(batch.U, U.data)
        cu_v.run(batch.Vh, U.data)
        cv_v.run(batch.Vh, Vh.data)
        cv_u.run(batch.U, Vh.data)
        return SVDResult(U, S, Vh, kept, trunc)


def factor_left(A: BlockTensor) -> BlockSVD:
    """Per-sector SVDs of the left matricisation (link, site) | bond."""
    lay = A.layout
    secs = left_sectors(lay)
    cold = lay.col.as_dict()
    qlist = list(secs.keys())
    mats = [(sum(p[2] for p in secs[q]), cold[q]) for q in qlist]
    batch = _SvdBatch(mats)
    g = CopyList()
    for i, q in enumerate(qlist):
        r0 = 0
        for s, ql, rows, off in secs[q]:
            _gather_plain(g, batch, i, off, rows, r0, 0, rows, mats[i][1])
            r0 += rows
    g.run(A.data, batch.ws)
    batch.run()
    return BlockSVD("left", lay, qlist, secs, mats, batch)


def factor_right(A: BlockTensor) -> BlockSVD:
    """Per-sector SVDs of the right matricisation bond | (site, link)."""
    lay = A.layout
    secs = right_sectors(lay)
    rowd = lay.row.as_dict()
    qlist = list(secs.keys())
    mats = [(rowd[q], sum(p[2] for p in secs[q])) for q in qlist]
    batch = _SvdBatch(mats)
    g = CopyList()
    for i, q in enumerate(qlist):
        c0 = 0
        for s, qr, cols, off in secs[q]:
            _gather_plain(g, batch, i, off, mats[i][0], 0, c0, mats[i][0], cols)
            c0 += cols
    g.run(A.data, batch.ws)
    batch.run()
    return BlockSVD("right", lay, qlist, secs, mats, batch)


def factor_core(C: BlockTensor) -> BlockSVD:
    """Per-sector SVDs of a one-slice core."""
    lay = C.layout
    tab = _block_table(lay)
    keys = list(tab)
    mats = [(tab[k][1], tab[k][2]) for k in keys]
    batch = _SvdBatch(mats)
    g = CopyList()
    for i, k in enumerate(keys):
        off, d, dc, qc = tab[k]
        _gather_plain(g, batch, i, off, d, 0, 0, d, dc)
    g.run(C.data, batch.ws)
    batch.run()
    return BlockSVD("core", lay, [k[1] for k in keys], None, mats, batch)


def svd_left(A: BlockTensor, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """A = U S Vh with U over (link, site): U (A.row, bond), S {q: sv}, Vh core (bond, A.col)."""
    return factor_left(A).select(cutoff, nkeep)


def svd_right(A: BlockTensor, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """A = U S Vh with Vh over (site, link): U core (A.row, bond), Vh (bond, s, r)."""
    return factor_right(A).select(cutoff, nkeep)


def svd_core(C: BlockTensor, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """C = U S Vh per sector of a one-slice core."""
    return factor_core(C).select(cutoff, nkeep)


def sq_weight(lst) -> float:
    return float(sum(e[0] ** 2 for e in lst))


def trunc_weight(res: SVDResult) -> float:
    k, c = sq_weight(res.kept), sq_weight(res.trunc)
    t = k + c
    return 0.0 if t == 0 else math.sqrt(c / t)
