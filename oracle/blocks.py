"""ORACLE (test infrastructure): block-sparse U(1) families on the host, their contractions.

A tensor is ``Fam``: a row leg and a column leg ({charge: dim}), one charge shift per slice and a
dict {(slice, row charge): dense block}.  The environment pushes and effective-Hamiltonian
applies follow the reference's contraction chains:

* left push   ``_mpo_left_step``   src/BUG/Core/Environments/mpo_environments.jl:65-90
* right push  ``_mpo_right_step``  mpo_environments.jl:92-108
* one-site    ``apply_one_site``   src/BUG/Core/Environments/effective_operators.jl:39-64
* zero-site   ``apply_zero_site``  effective_operators.jl:103-107 (built by ``zero_site_h`` :91-94)
* shift       ``shift_h``          effective_operators.jl:118-129

Pinned by ``tests/test_oracle_contractions.py`` against plain dense ``numpy.einsum`` of the same
contractions on the densified tensors, and by <psi|H|psi> against the 2^L dense Hamiltonian built
from Pauli matrices (``oracle/dense.py``), so a wrong leg, conjugation or charge shift fails there.
"""
from __future__ import annotations

from typing import Dict, Iterable, Tuple

import numpy as np


class Fam:
    __slots__ = ("row", "col", "shifts", "blocks")

    def __init__(self, row: Dict[int, int], col: Dict[int, int], shifts, blocks=None):
        self.row = dict(sorted(row.items()))
        self.col = dict(sorted(col.items()))
        self.shifts = [int(s) for s in shifts]
        self.blocks: Dict[Tuple[int, int], np.ndarray] = {} if blocks is None else dict(blocks)

    @staticmethod
    def zeros(row, col, shifts) -> "Fam":
        f = Fam(row, col, shifts)
        for k, s in enumerate(f.shifts):
            for q, d in f.row.items():
                dc = f.col.get(q + s, 0)
                if d > 0 and dc > 0:
                    f.blocks[(k, q)] = np.zeros((d, dc), dtype=np.complex128)
        return f

    def get(self, k: int, q: int):
        return self.blocks.get((k, q))

    def add(self, k: int, q: int, m: np.ndarray):
        cur = self.blocks.get((k, q))
        if cur is None:
            self.blocks[(k, q)] = np.array(m, dtype=np.complex128)
        else:
            cur += m

    def copy(self) -> "Fam":
        return Fam(self.row, self.col, self.shifts, {k: v.copy() for k, v in self.blocks.items()})

    def scale(self, a) -> "Fam":
        return Fam(self.row, self.col, self.shifts, {k: a * v for k, v in self.blocks.items()})

    def axpy(self, a, other: "Fam") -> "Fam":
        out = self.copy()
        for key, v in other.blocks.items():
            out.add(key[0], key[1], a * v)
        return out

    def norm(self) -> float:
        return float(np.sqrt(sum(np.vdot(v, v).real for v in self.blocks.values())))

    def dense(self) -> np.ndarray:
        rs, cs, r, c = {}, {}, 0, 0
        for q, d in self.row.items():
            rs[q] = r
            r += d
        for q, d in self.col.items():
            cs[q] = c
            c += d
        out = np.zeros((len(self.shifts), r, c), dtype=np.complex128)
        for (k, q), b in self.blocks.items():
            qc = q + self.shifts[k]
            out[k, rs[q]: rs[q] + b.shape[0], cs[qc]: cs[qc] + b.shape[1]] = b
        return out


def inner(a: Fam, b: Fam) -> complex:
    """<a|b> over matching blocks (``tensor_inner``, symmetric_mps.jl:83-93)."""
    s = 0j
    for key, v in a.blocks.items():
        w = b.blocks.get(key)
        if w is not None:
            s += np.vdot(v, w)
    return complex(s)


class MpoSite:
    """Entries W[a, b] += coeff |sout><sin| with channel charges cl/cr."""

    def __init__(self, cl, cr, entries: Iterable[Tuple[int, int, int, int, complex]]):
        self.cl = [int(c) for c in cl]
        self.cr = [int(c) for c in cr]
        self.entries = [(int(a), int(b), int(so), int(si), complex(c)) for a, b, so, si, c in entries]

    def dense(self, d: int) -> np.ndarray:
        w = np.zeros((len(self.cl), len(self.cr), d, d), dtype=np.complex128)
        for a, b, so, si, c in self.entries:
            w[a, b, so, si] += c
        return w


def env_fam(bra: Dict[int, int], ket: Dict[int, int], channel_charges) -> Fam:
    return Fam.zeros(bra, ket, [-c for c in channel_charges])


def push_left(L: Fam, A: Fam, W: MpoSite, bra: Fam = None) -> Fam:
    """Lout_b = sum W[a,b](so,si) bra_so^H L_a A_si  (mpo_environments.jl:65-90)."""
    bra = A if bra is None else bra
    out = env_fam(bra.col, A.col, W.cr)
    for a, b, so, si, c in W.entries:
        for (k, qk), ablk in A.blocks.items():
            if k != si:
                continue
            qb = qk + W.cl[a]              # L_a: bra charge = ket charge + c_a
            lblk = L.get(a, qb)
            bblk = bra.get(so, qb)
            if lblk is None or bblk is None:
                continue
            out.add(b, qb + bra.shifts[so], c * (bblk.conj().T @ lblk @ ablk))
    return out


def push_right(R: Fam, B: Fam, W: MpoSite, bra: Fam = None) -> Fam:
    """Rout_a = sum W[a,b](so,si) conj(bra_so) R_b B_si^T  (mpo_environments.jl:92-108)."""
    bra = B if bra is None else bra
    out = env_fam(bra.row, B.row, W.cl)
    for a, b, so, si, c in W.entries:
        for (k, ql), bblk in B.blocks.items():
            if k != si:
                continue
            qk = ql + B.shifts[si]          # ket charge on the right link
            qb = qk + W.cr[b]               # R_b: bra charge = ket charge + c_b
            rblk = R.get(b, qb)
            if rblk is None:
                continue
            ql_bra = qb - bra.shifts[so]
            brablk = bra.get(so, ql_bra)
            if brablk is None:
                continue
            out.add(a, ql_bra, c * (brablk.conj() @ rblk @ bblk.T))
    return out


def apply_one_site(L: Fam, W: MpoSite, R: Fam, X: Fam, shift: float = 0.0) -> Fam:
    """Y_so = sum W[a,b](so,si) L_a X_si R_b^T - shift X  (effective_operators.jl:39-64,118-129)."""
    Y = Fam.zeros(X.row, X.col, X.shifts)
    for a, b, so, si, c in W.entries:
        for (k, ql), xblk in X.blocks.items():
            if k != si:
                continue
            qr = ql + X.shifts[si]
            qlb = ql + W.cl[a]
            qrb = qr + W.cr[b]
            lblk = L.get(a, qlb)
            rblk = R.get(b, qrb)
            if lblk is None or rblk is None:
                continue
            if (so, qlb) not in Y.blocks:
                continue
            Y.add(so, qlb, c * (lblk @ xblk @ rblk.T))
    if shift != 0.0:
        Y = Y.axpy(-shift, X)
    return Y


def apply_zero_site(L: Fam, R: Fam, C: Fam, shift: float = 0.0) -> Fam:
    """Y = sum_w L_w C R_w^T - shift C  (effective_operators.jl:103-107)."""
    Y = Fam.zeros(C.row, C.col, [0])
    for (k, ql), cblk in C.blocks.items():
        qr = ql + C.shifts[0]
        for w in range(len(L.shifts)):
            qlb = ql - L.shifts[w]
            qrb = qr - R.shifts[w]
            lblk = L.get(w, qlb)
            rblk = R.get(w, qrb)
            if lblk is None or rblk is None or (0, qlb) not in Y.blocks:
                continue
            Y.add(0, qlb, lblk @ cblk @ rblk.T)
    if shift != 0.0:
        Y = Y.axpy(-shift, C)
    return Y
