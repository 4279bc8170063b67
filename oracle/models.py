"""ORACLE (test infrastructure): XXZ chain MPO, product and random states, dense references.

* ``xxz_mpo``       -- the term-list automaton of ``mpo_from_terms``
                       (src/BUG/Core/TensorNetworks/mpo.jl:208-256) for ``xxz_chain``
                       (src/QMB/Models/ClosedSystem/xxz.jl:15-28): states id / term t / done,
                       site 1 keeps the id row, site L the done column.
* ``product_state`` -- ``product_state``/``domain_wall_state``/``neel_state``
                       (src/QMB/Tooling/initial_states.jl:15-80,137-146).
* ``dense_hamiltonian`` -- H = J sum (Sx Sx + Sy Sy + delta Sz Sz) on 2^L states from Pauli
                       matrices, site 1 most significant and local index 0 = up, the convention of
                       test/common/dense_reference.jl:12-53.

Pinned: the MPO contracted to a 2^L matrix equals ``dense_hamiltonian`` (built without the
automaton), and the dense H spectrum of the XX chain equals the free-fermion single-particle
sums (test/common/free_fermion.jl:20-27).
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from .blocks import Fam, MpoSite

UP, DN = 0, 1
Q_U1 = (1, -1)
Q_NONE = (0, 0)

_SP = np.array([[0, 1], [0, 0]], dtype=np.complex128)   # |up><down|, index 0 = up
_SM = _SP.T.copy()
_SZ = np.diag([0.5, -0.5]).astype(np.complex128)
_ID = np.eye(2, dtype=np.complex128)


def _entries(m: np.ndarray, a: int, b: int, scale: complex = 1.0):
    return [(a, b, so, si, scale * m[so, si]) for so in range(2) for si in range(2)
            if m[so, si] != 0]


def xxz_mpo(L: int, J: float = 1.0, delta: float = 1.0, symmetry: str = "U1") -> List[MpoSite]:
    q = Q_U1 if symmetry == "U1" else Q_NONE
    terms = [(_SP, _SM, J / 2), (_SM, _SP, J / 2)]
    if delta != 0.0:
        terms.append((_SZ, _SZ, J * delta))
    T = len(terms)

    def charge(m):
        so, si = np.argwhere(m != 0)[0]
        return q[so] - q[si]

    ch = [0] + [charge(t[0]) for t in terms] + [0]
    sites = []
    for i in range(L):
        rows = [0] if i == 0 else list(range(T + 2))
        cols = [T + 1] if i == L - 1 else list(range(T + 2))
        full = []
        if i < L - 1:
            full += _entries(_ID, 0, 0)
            for t, (left, _, _) in enumerate(terms):
                full += _entries(left, 0, 1 + t)
        if i > 0:
            full += _entries(_ID, T + 1, T + 1)
            for t, (_, right, c) in enumerate(terms):
                full += _entries(right, 1 + t, T + 1, c)
        ents = [(rows.index(a), cols.index(b), so, si, c) for a, b, so, si, c in full
                if a in rows and b in cols]
        sites.append(MpoSite([ch[r] for r in rows], [ch[c] for c in cols], ents))
    return sites


def mpo_dense(sites: List[MpoSite], d: int = 2) -> np.ndarray:
    """Contract the MPO to its 2^L x 2^L matrix (site 1 most significant)."""
    acc = sites[0].dense(d)[0]                         # (wr, d, d)
    acc = acc.reshape(acc.shape[0], d, d)
    for s in sites[1:]:
        w = s.dense(d)                                 # (wl, wr, d, d)
        n = acc.shape[1]
        acc = np.einsum("aij,abkl->bikjl", acc, w).reshape(w.shape[1], n * d, n * d)
    return acc[0]


def dense_hamiltonian(L: int, J: float = 1.0, delta: float = 1.0) -> np.ndarray:
    sx = np.array([[0, 1], [1, 0]], dtype=np.complex128) / 2
    sy = np.array([[0, -1j], [1j, 0]], dtype=np.complex128) / 2
    sz = _SZ

    def op(o, i):
        m = np.eye(1, dtype=np.complex128)
        for j in range(L):
            m = np.kron(m, o if j == i else _ID)
        return m

    H = np.zeros((2 ** L, 2 ** L), dtype=np.complex128)
    for i in range(L - 1):
        H += J * (op(sx, i) @ op(sx, i + 1) + op(sy, i) @ op(sy, i + 1)
                  + delta * op(sz, i) @ op(sz, i + 1))
    return H


def product_state(spins: Sequence[str], symmetry: str = "U1") -> List[Fam]:
    q = Q_U1 if symmetry == "U1" else Q_NONE
    out, cur = [], 0
    for s in spins:
        st = UP if s == "up" else DN
        f = Fam({cur: 1}, {cur + q[st]: 1}, q)
        f.blocks[(st, cur)] = np.ones((1, 1), dtype=np.complex128)
        out.append(f)
        cur += q[st]
    return out


def domain_wall_state(L: int, symmetry: str = "U1") -> List[Fam]:
    return product_state(["up" if i < L // 2 else "down" for i in range(L)], symmetry)


def neel_state(L: int, symmetry: str = "U1") -> List[Fam]:
    return product_state(["up" if i % 2 == 0 else "down" for i in range(L)], symmetry)


def random_state(legs: List[dict], rng: np.random.Generator, symmetry: str = "U1") -> List[Fam]:
    """Complex-Gaussian MPS on the given link legs (legs[b] = {charge: dim}, b = 0..L)."""
    q = Q_U1 if symmetry == "U1" else Q_NONE
    out = []
    for i in range(len(legs) - 1):
        f = Fam.zeros(legs[i], legs[i + 1], q)
        for key, b in f.blocks.items():
            f.blocks[key] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
        out.append(f)
    return out


def mps_dense(sites: List[Fam]) -> np.ndarray:
    """The 2^L amplitude vector of an MPS (boundary links one-dimensional)."""
    v = sites[0].dense()                               # (d, 1, D)
    v = v[:, 0, :]                                     # (d, D)
    for s in sites[1:]:
        a = s.dense()                                  # (d, Dl, Dr)
        v = np.einsum("xl,slr->xsr", v, a).reshape(-1, a.shape[2])
    return v[:, 0]
