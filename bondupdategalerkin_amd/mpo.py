"""The spin-1/2 local space, the XXZ chain and its MPO as sparse operator-valued matrices.

Reference: ``xxz_chain``/``xxz_mpo`` (src/QMB/Models/ClosedSystem/xxz.jl:15-28,61-62) and the
term-list automaton ``mpo_from_terms`` (src/BUG/Core/TensorNetworks/mpo.jl:208-256): automaton
states ``id``, one state per term, ``done``; site 1 keeps only the ``id`` row, site L only the
``done`` column; each term opens with its left operator and closes with its (weighted) right one.

The reference stores each site as a rank-4 ``TLArray`` (w_l, s_ket, s_bra, w_r).  Here a site is
the list of its nonzero entries ``W[a, b] += coeff |sout><sin|`` plus the channel charges (the
cumulative charge change), which is exactly what the grouped block contractions consume: each
entry becomes one weighted block-product term.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

# local states of a spin-1/2 site: 0 = up (charge 2 S^z = +1), 1 = down (charge -1)
UP, DOWN = 0, 1
SPIN_CHARGES_U1 = (1, -1)
SPIN_CHARGES_NONE = (0, 0)

# sparse local operators: list of (sout, sin, value)
OP_I = ((UP, UP, 1.0), (DOWN, DOWN, 1.0))
OP_SP = ((UP, DOWN, 1.0),)           # S^+ = |up><down|
OP_SM = ((DOWN, UP, 1.0),)           # S^- = |down><up|
OP_SZ = ((UP, UP, 0.5), (DOWN, DOWN, -0.5))


def op_charge(op, charges) -> int:
    qs = {charges[so] - charges[si] for so, si, _ in op}
    if len(qs) != 1:
        raise ValueError("operator does not carry a definite charge")
    return qs.pop()


@dataclass
class MPOSite:
    """One MPO site: channel charges and the nonzero entries W[a, b] += coeff |sout><sin|."""

    cl: np.ndarray
    cr: np.ndarray
    a: np.ndarray
    b: np.ndarray
    sout: np.ndarray
    sin: np.ndarray
    coeff: np.ndarray  # complex128
    _c: Optional[tuple] = field(default=None, repr=False, compare=False)

    @property
    def wl(self) -> int:
        return len(self.cl)

    @property
    def wr(self) -> int:
        return len(self.cr)

    @property
    def nent(self) -> int:
        return len(self.a)

    def dense(self, d: int) -> np.ndarray:
        """W as a dense (wl, wr, d_out, d_in) array."""
        w = np.zeros((self.wl, self.wr, d, d), dtype=np.complex128)
        for a, b, so, si, c in zip(self.a, self.b, self.sout, self.sin, self.coeff):
            w[a, b, so, si] += c
        return w

    def descriptor(self) -> _lib.BugMpoSite:
        if self._c is None:
            arrs = tuple(_lib.i32(x) for x in (self.cl, self.cr, self.a, self.b, self.sout, self.sin))
            co = np.ascontiguousarray(
                np.stack([self.coeff.real, self.coeff.imag], axis=1).reshape(-1), dtype=np.float64)
            d = _lib.BugMpoSite()
            d.wl, d.wr = self.wl, self.wr
            d.cl, d.cr = _lib.ptr32(arrs[0]), _lib.ptr32(arrs[1])
            d.nent = self.nent
            d.a, d.b, d.sout, d.sin = (_lib.ptr32(x) for x in arrs[2:])
            d.coeff = _lib.ptrf(co)
            self._c = (d, arrs, co)
        return self._c[0]


@dataclass
class MPO:
    """An MPO on L sites of one local space (``charges[s]`` = U(1) charge of local state s)."""

    sites: List[MPOSite]
    charges: Tuple[int, ...]

    def __len__(self):
        return len(self.sites)

    def __getitem__(self, i: int) -> MPOSite:
        return self.sites[i]

    @property
    def d(self) -> int:
        return len(self.charges)

    def channel_dims(self) -> List[int]:
        return [self.sites[0].wl] + [s.wr for s in self.sites]


@dataclass
class Term:
    """One nearest-neighbour term coeff * left_i right_{i+1} (reference ``XXZTerm``, mpo.jl:160-170)."""

    left: Sequence[Tuple[int, int, float]]
    right: Sequence[Tuple[int, int, float]]
    coeff: float


def xxz_terms(J: float = 1.0, delta: float = 1.0) -> List[Term]:
    """J (Sx Sx + Sy Sy + delta Sz Sz) as S+S-/2 + S-S+/2 + delta Sz Sz (xxz.jl:15-28)."""
    terms = [Term(OP_SP, OP_SM, 0.5 * J), Term(OP_SM, OP_SP, 0.5 * J)]
    if delta != 0.0:
        terms.append(Term(OP_SZ, OP_SZ, J * delta))
    return terms


def mpo_from_terms(L: int, terms: Sequence[Term], charges: Sequence[int]) -> MPO:
    """The automaton MPO of a uniform nearest-neighbour term list (mpo.jl:208-256)."""
    if L < 2:
        raise ValueError(f"an MPO chain needs at least two sites, got {L}")
    T = len(terms)
    ID, DONE = 0, T + 1
    ch = [0] + [op_charge(t.left, charges) for t in terms] + [0]
    for t in terms:
        if op_charge(t.left, charges) + op_charge(t.right, charges) != 0:
            raise ValueError("a term must be charge-neutral")
    sites = []
    for i in range(L):
        ents = []
        if i < L - 1:
            ents += [(ID, ID, so, si, v) for so, si, v in OP_I]
            for t, term in enumerate(terms):
                ents += [(ID, 1 + t, so, si, v) for so, si, v in term.left]
        if i > 0:
            ents += [(DONE, DONE, so, si, v) for so, si, v in OP_I]
            for t, term in enumerate(terms):
                ents += [(1 + t, DONE, so, si, term.coeff * v) for so, si, v in term.right]
        rows = [ID] if i == 0 else list(range(T + 2))
        cols = [DONE] if i == L - 1 else list(range(T + 2))
        rmap = {c: k for k, c in enumerate(rows)}
        cmap = {c: k for k, c in enumerate(cols)}
        ents = [e for e in ents if e[0] in rmap and e[1] in cmap]
        sites.append(MPOSite(
            cl=np.array([ch[c] for c in rows], dtype=np.int32),
            cr=np.array([ch[c] for c in cols], dtype=np.int32),
            a=np.array([rmap[e[0]] for e in ents], dtype=np.int32),
            b=np.array([cmap[e[1]] for e in ents], dtype=np.int32),
            sout=np.array([e[2] for e in ents], dtype=np.int32),
            sin=np.array([e[3] for e in ents], dtype=np.int32),
            coeff=np.array([e[4] for e in ents], dtype=np.complex128),
        ))
    return MPO(sites, tuple(int(c) for c in charges))


def xxz_mpo(L: int, J: float = 1.0, delta: float = 1.0, symmetry: str = "U1") -> MPO:
    """``xxz_mpo(L; J, delta)`` (xxz.jl:61-62): w = 4 for XX, 5 for XXZ."""
    charges = {"U1": SPIN_CHARGES_U1, "none": SPIN_CHARGES_NONE}[symmetry]
    return mpo_from_terms(L, xxz_terms(J, delta), charges)
