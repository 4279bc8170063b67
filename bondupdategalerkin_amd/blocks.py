"""Device-resident block-sparse U(1) tensors.

The reference holds every tensor as a Telum ``TLArray`` (per-sector dense payloads,
deps/Telum/src/TLArray.jl:517-600) and contracts them sector pair by sector pair.  Here a tensor
is a *family of block-diagonal matrices with charge shifts*, stored in ONE device buffer:

* ``Leg``      -- sectors ``(charge, dim)`` sorted by charge (U(1) charge = 2 S^z; a dense leg is
                  one charge-0 sector).
* ``Layout``   -- row leg, column leg and one charge shift per slice.  Slice ``k`` has a block at
                  row sector ``i`` iff the column leg holds ``charge_i + shift_k``; blocks are
                  column-major and packed slice by slice, sector by sector.
* ``BlockTensor`` -- a layout plus a flat complex128 device buffer.

An MPS site tensor ``A(l, s, r)`` is the family ``A_s`` (shift = charge of local state ``s``), an
MPO environment ``E(bra, w, ket)`` is ``E_w`` (shift = minus the charge of channel ``w``), a
zero-site core is one slice with shift 0.  The fused left matrix ``A[(l, s), r]`` of the
reference (``symmetric_mps.jl:44-51``) is the slices stacked in ``s`` order.

Torch is used only to own device memory and streams; arithmetic goes through ``libbug.so``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterable, Optional, Sequence, Tuple

import ctypes
import numpy as np
import torch

from . import _lib

if hasattr(torch._C, "_cuda_getCurrentRawStream"):
    def current_stream_handle() -> int:
        """Raw handle of torch's current stream on the current device, the value of
        ``torch.cuda.current_stream().cuda_stream`` without building a Stream object (the
        entry-point wrappers ask for it on every call)."""
        return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())
else:                                                    # pragma: no cover
    def current_stream_handle() -> int:
        return int(torch.cuda.current_stream().cuda_stream)


@dataclass(frozen=True)
class Leg:
    charges: Tuple[int, ...]
    dims: Tuple[int, ...]

    def __post_init__(self):
        if len(self.charges) != len(self.dims):
            raise ValueError("charges and dims differ in length")
        if list(self.charges) != sorted(set(self.charges)):
            raise ValueError("leg charges must be strictly increasing")
        if any(d < 0 for d in self.dims):
            raise ValueError("negative sector dimension")

    @staticmethod
    def make(pairs: Iterable[Tuple[int, int]]) -> "Leg":
        """Leg from (charge, dim) pairs; zero-dim sectors dropped, charges merged by sum."""
        acc: Dict[int, int] = {}
        for q, d in pairs:
            if d > 0:
                acc[int(q)] = acc.get(int(q), 0) + int(d)
        qs = sorted(acc)
        return Leg(tuple(qs), tuple(acc[q] for q in qs))

    @staticmethod
    def dense(n: int) -> "Leg":
        return Leg((0,), (int(n),))

    @property
    def nsec(self) -> int:
        return len(self.charges)

    @property
    def total(self) -> int:
        return int(sum(self.dims))

    def index(self, q: int) -> int:
        try:
            return self.charges.index(q)
        except ValueError:
            return -1

    def dim_of(self, q: int) -> int:
        i = self.index(q)
        return 0 if i < 0 else self.dims[i]

    def as_dict(self) -> Dict[int, int]:
        return dict(zip(self.charges, self.dims))

    def starts(self) -> Dict[int, int]:
        """Dense row/column start of each sector (sectors in charge order)."""
        out, s = {}, 0
        for q, d in zip(self.charges, self.dims):
            out[q] = s
            s += d
        return out


class Layout:
    """Block structure of a family of shifted block-diagonal matrices (see module docstring)."""

    __slots__ = ("row", "col", "shifts", "offsets", "nelem", "_key", "_c")

    def __init__(self, row: Leg, col: Leg, shifts: Sequence[int]):
        self.row = row
        self.col = col
        self.shifts = tuple(int(s) for s in shifts)
        off = np.full((len(self.shifts), row.nsec), -1, dtype=np.int64)
        pos = 0
        cold = col.as_dict()
        for k, sh in enumerate(self.shifts):
            for i, (q, d) in enumerate(zip(row.charges, row.dims)):
                dc = cold.get(q + sh, 0)
                if d == 0 or dc == 0:
                    continue
                off[k, i] = pos
                pos += d * dc
        self.offsets = off
        self.nelem = int(pos)
        self._key = (row, col, self.shifts)
        self._c = None

    @property
    def nslice(self) -> int:
        return len(self.shifts)

    def key(self):
        return self._key

    def __eq__(self, other):
        return isinstance(other, Layout) and self._key == other._key

    def __hash__(self):
        return hash(self._key)

    def blocks(self):
        """Yield (slice, row charge, col charge, rows, cols, offset) for every stored block."""
        cold = self.col.as_dict()
        for k, sh in enumerate(self.shifts):
            for i, (q, d) in enumerate(zip(self.row.charges, self.row.dims)):
                off = int(self.offsets[k, i])
                if off < 0:
                    continue
                yield k, q, q + sh, d, cold[q + sh], off

    def c_arrays(self):
        """Host tables for the C descriptor, built once per layout."""
        if self._c is None:
            self._c = (
                _lib.i32(self.row.charges),
                _lib.i32(self.row.dims),
                _lib.i32(self.col.charges),
                _lib.i32(self.col.dims),
                _lib.i32(self.shifts),
                np.ascontiguousarray(self.offsets.reshape(-1), dtype=np.int64),
            )
        return self._c

    def descriptor(self, data_ptr: int, nelem: Optional[int] = None) -> _lib.BugBsm:
        rq, rd, cq, cd, sh, off = self.c_arrays()
        d = _lib.BugBsm()
        d.row = _lib.BugLeg(self.row.nsec, _lib.ptr32(rq), _lib.ptr32(rd))
        d.col = _lib.BugLeg(self.col.nsec, _lib.ptr32(cq), _lib.ptr32(cd))
        d.nslice = self.nslice
        d.shift = _lib.ptr32(sh)
        d.offset = _lib.ptr64(off)
        d.data = ctypes.c_void_p(data_ptr)
        d.nelem = self.nelem if nelem is None else nelem
        return d


def device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


class BlockTensor:
    """A layout plus its flat complex128 device buffer."""

    __slots__ = ("layout", "data")

    def __init__(self, layout: Layout, data: Optional[torch.Tensor] = None):
        self.layout = layout
        if data is None:
            data = torch.zeros(max(layout.nelem, 1), dtype=torch.complex128, device=device())
        if data.dtype != torch.complex128 or data.dim() != 1 or data.numel() < layout.nelem:
            raise ValueError("BlockTensor data must be a flat complex128 buffer of nelem elements")
        if not data.is_cuda:
            raise ValueError("BlockTensor data must live on the GPU")
        self.data = data

    @staticmethod
    def empty(layout: Layout) -> "BlockTensor":
        return BlockTensor(
            layout, torch.empty(max(layout.nelem, 1), dtype=torch.complex128, device=device())
        )

    def desc(self) -> _lib.BugBsm:
        return self.layout.descriptor(self.data.data_ptr())

    # ---- host transfer (tests and state download) ----
    def to_blocks(self) -> Dict[Tuple[int, int], np.ndarray]:
        """{(slice, row charge): block} on the host."""
        host = self.data.detach().cpu().numpy()
        out = {}
        for k, q, qc, d, dc, off in self.layout.blocks():
            out[(k, q)] = host[off: off + d * dc].reshape(dc, d).T.copy()
        return out

    @staticmethod
    def from_blocks(layout: Layout, blocks: Dict[Tuple[int, int], np.ndarray]) -> "BlockTensor":
        host = np.zeros(max(layout.nelem, 1), dtype=np.complex128)
        for k, q, qc, d, dc, off in layout.blocks():
            b = blocks.get((k, q))
            if b is None:
                continue
            b = np.asarray(b, dtype=np.complex128)
            if b.shape != (d, dc):
                raise ValueError(f"block {(k, q)} has shape {b.shape}, layout wants {(d, dc)}")
            host[off: off + d * dc] = b.T.reshape(-1)
        return BlockTensor(layout, torch.from_numpy(host).to(device()))

    def to_dense(self) -> np.ndarray:
        """(nslice, rows, cols) dense array, sectors in charge order."""
        rs, cs = self.layout.row.starts(), self.layout.col.starts()
        out = np.zeros((self.layout.nslice, self.layout.row.total, self.layout.col.total),
                       dtype=np.complex128)
        for (k, q), b in self.to_blocks().items():
            qc = q + self.layout.shifts[k]
            out[k, rs[q]: rs[q] + b.shape[0], cs[qc]: cs[qc] + b.shape[1]] = b
        return out

    def clone(self) -> "BlockTensor":
        return BlockTensor(self.layout, self.data.clone())


def bsm_product(out_layout: Layout, P: BlockTensor, Q: BlockTensor, terms, accumulate_into=None,
                precision: int = _lib.PREC_FP64) -> BlockTensor:
    """Grouped block product ``out[o] (+)= sum alpha op_a(P[p]) op_b(Q[q])`` in one launch.

    ``terms`` is a list of (o, p, q, op_a, op_b, alpha).  With ``accumulate_into`` the result is
    added onto that tensor (which must have ``out_layout``) and returned.
    """
    ctx = _lib.context()
    if accumulate_into is not None:
        if accumulate_into.layout != out_layout:
            raise ValueError("accumulate_into has a different layout")
        out = accumulate_into
    else:
        out = BlockTensor.empty(out_layout)
    arr = (_lib.BugTerm * max(len(terms), 1))()
    for n, (o, p, q, oa, ob, alpha) in enumerate(terms):
        alpha = complex(alpha)
        arr[n] = _lib.BugTerm(o, p, q, oa, ob, alpha.real, alpha.imag)
    ctx.check(ctx._lib.bug_bsm_product(
        ctx.handle, out.desc(), P.desc(), Q.desc(), len(terms), arr,
        1 if accumulate_into is not None else 0, precision, current_stream_handle()))
    return out
