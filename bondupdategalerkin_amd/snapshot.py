"""State snapshots for serving and resume: a device-resident ``SymMPS`` written to and read from one
``.npz`` file.

The file holds what a U(1) site tensor is made of (the reference's ``TLArray`` state I/O,
deps/Telum/src/io_hdf5.jl:724-797, keeps the same content in HDF5 groups): per site its row and
column legs (sector charges and dimensions), the local states' charges, and the sector blocks in
the library's packing (slice by slice, row sectors in charge order, column-major); plus the
chain's local charges, the canonical centre, a schema version and caller metadata (JSON).  Arrays
only, read back with ``allow_pickle=False``; a load checks every table against the data length and
uploads the blocks in one copy per site.
"""
from __future__ import annotations

import json
from typing import Optional, Tuple

import numpy as np
import torch

from .blocks import BlockTensor, Layout, Leg, device
from .mps import SymMPS

SCHEMA = "bug-mps-1"


def save_state(path: str, psi: SymMPS, meta: Optional[dict] = None) -> str:
    """Write ``psi`` (and ``meta``, JSON-serialisable) to ``path``; returns the path."""
    arrays = {
        "schema": np.frombuffer(SCHEMA.encode(), dtype=np.uint8),
        "charges": np.asarray(psi.charges, dtype=np.int32),
        "center": np.asarray([psi.center], dtype=np.int64),
        "length": np.asarray([len(psi)], dtype=np.int64),
        "meta": np.frombuffer(json.dumps(meta or {}).encode(), dtype=np.uint8),
    }
    for i, t in enumerate(psi.tensors):
        lay = t.layout
        arrays[f"row_q{i}"] = np.asarray(lay.row.charges, dtype=np.int32)
        arrays[f"row_d{i}"] = np.asarray(lay.row.dims, dtype=np.int32)
        arrays[f"col_q{i}"] = np.asarray(lay.col.charges, dtype=np.int32)
        arrays[f"col_d{i}"] = np.asarray(lay.col.dims, dtype=np.int32)
        arrays[f"shift{i}"] = np.asarray(lay.shifts, dtype=np.int32)
        arrays[f"data{i}"] = t.data[:lay.nelem].detach().cpu().numpy()
    with open(path, "wb") as f:
        np.savez(f, **arrays)
    return path


def load_state(path: str) -> Tuple[SymMPS, dict]:
    """Read a snapshot written by ``save_state``: (the state on the current device, meta)."""
    with np.load(path, allow_pickle=False) as z:
        schema = bytes(z["schema"]).decode()
        if schema != SCHEMA:
            raise ValueError(f"unsupported snapshot schema {schema!r}")
        L = int(z["length"][0])
        charges = tuple(int(c) for c in z["charges"])
        center = int(z["center"][0])
        meta = json.loads(bytes(z["meta"]).decode() or "{}")
        tensors = []
        for i in range(L):
            row = Leg(tuple(int(x) for x in z[f"row_q{i}"]), tuple(int(x) for x in z[f"row_d{i}"]))
            col = Leg(tuple(int(x) for x in z[f"col_q{i}"]), tuple(int(x) for x in z[f"col_d{i}"]))
            shifts = tuple(int(x) for x in z[f"shift{i}"])
            if sorted(set(shifts)) != sorted(set(charges)):
                raise ValueError(f"site {i}: local charges {shifts} differ from the chain's")
            lay = Layout(row, col, shifts)
            data = np.ascontiguousarray(z[f"data{i}"], dtype=np.complex128)
            if data.shape != (lay.nelem,):
                raise ValueError(f"site {i}: {data.size} elements stored, the legs need {lay.nelem}")
            buf = torch.zeros(max(lay.nelem, 1), dtype=torch.complex128, device=device())
            if lay.nelem:
                buf[:lay.nelem].copy_(torch.from_numpy(data))
            tensors.append(BlockTensor(lay, buf))
    if not 0 <= center < L:
        raise ValueError(f"centre {center} outside 0:{L - 1}")
    for i in range(L - 1):
        if tensors[i].layout.col != tensors[i + 1].layout.row:
            raise ValueError(f"bond {i}: the site legs do not match")
    return SymMPS(tensors, charges, center), meta
