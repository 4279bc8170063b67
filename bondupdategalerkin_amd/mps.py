"""Matrix product states on the device: container, product states, synthetic states.

Reference container: ``SymMPS`` (src/BUG/Core/TensorNetworks/symmetric_mps.jl:44-54), tensors
``(l, s, r)`` with a centre; product states ``product_state``/``domain_wall_state``/``neel_state``
(src/QMB/Tooling/initial_states.jl:15-80,137-146).  Sites are 0-based here (site ``i`` of the
reference is ``i-1``); ``center`` follows the same shift.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

from .blocks import BlockTensor, Layout, Leg
from .mpo import DOWN, SPIN_CHARGES_NONE, SPIN_CHARGES_U1, UP


class SymMPS:
    """L site tensors ``A_s`` (row = left bond, col = right bond, shift = local charge)."""

    def __init__(self, tensors: List[BlockTensor], charges: Sequence[int], center: int = 0):
        self.tensors = list(tensors)
        self.charges = tuple(int(c) for c in charges)
        self.center = int(center)
        for i, t in enumerate(self.tensors):
            if t.layout.shifts != self.charges:
                raise ValueError(f"site {i}: slice shifts {t.layout.shifts} != local charges")
            if i > 0 and t.layout.row != self.tensors[i - 1].layout.col:
                raise ValueError(f"bond {i}: left leg of site {i} != right leg of site {i - 1}")

    def __len__(self):
        return len(self.tensors)

    def __getitem__(self, i: int) -> BlockTensor:
        return self.tensors[i]

    def __setitem__(self, i: int, t: BlockTensor):
        self.tensors[i] = t

    @property
    def d(self) -> int:
        return len(self.charges)

    def bond_leg(self, b: int) -> Leg:
        """Leg of link b (b = 0 .. L): link b sits left of site b."""
        return self.tensors[b].layout.row if b < len(self) else self.tensors[-1].layout.col

    def bond_dims(self) -> List[int]:
        """Dimensions of the L-1 internal bonds (reference ``bond_dims``, symmetric_mps.jl:64)."""
        return [self.tensors[i].layout.col.total for i in range(len(self) - 1)]

    def copy(self) -> "SymMPS":
        """Shares the site tensors, as ``copy(psi)`` does (operations replace, never mutate)."""
        return SymMPS(list(self.tensors), self.charges, self.center)


def site_layout(left: Leg, right: Leg, charges: Sequence[int]) -> Layout:
    return Layout(left, right, charges)


def product_state(spins: Sequence[str], symmetry: str = "U1") -> SymMPS:
    """Product state of ``:up``/``:down`` spins, bond dimension 1 (initial_states.jl:15-80)."""
    charges = SPIN_CHARGES_U1 if symmetry == "U1" else SPIN_CHARGES_NONE
    tensors = []
    q = 0
    for i, s in enumerate(spins):
        if s not in ("up", "down"):
            raise ValueError(f"spin {i} must be 'up' or 'down', got {s!r}")
        st = UP if s == "up" else DOWN
        left = Leg((q,), (1,))
        q2 = q + charges[st]
        right = Leg((q2,), (1,))
        lay = Layout(left, right, charges)
        blocks = {(st, q): np.ones((1, 1), dtype=np.complex128)}
        tensors.append(BlockTensor.from_blocks(lay, blocks))
        q = q2
    return SymMPS(tensors, charges, center=0)


def domain_wall_state(L: int, symmetry: str = "U1") -> SymMPS:
    """up...up down...down with the wall in the middle (initial_states.jl:137-140)."""
    return product_state(["up" if i < L // 2 else "down" for i in range(L)], symmetry)


def neel_state(L: int, symmetry: str = "U1") -> SymMPS:
    """up down up down ... (initial_states.jl:142-146)."""
    return product_state(["up" if i % 2 == 0 else "down" for i in range(L)], symmetry)


def synthetic_bond_legs(L: int, dmax: int, sigma: float = 4.0) -> List[Leg]:
    """Bond legs of the synthetic U(1) state used for steady-state timing.

    Link b (b = 0..L) holds the charges reachable after b spins that can still return to total
    charge 0, dimension min(2^min(b, L-b), dmax) shared over the charges with weights
    exp(-q^2 / (2 sigma^2)), each sector at least 1 and at most its number of basis states.
    """
    from math import comb

    legs = []
    for b in range(L + 1):
        n = min(b, L - b)
        total = min(2 ** n if n < 62 else dmax, dmax)
        qs = [q for q in range(-b, b + 1, 2) if abs(q) <= L - b]
        # basis-state counts on the smaller side bound each sector
        def count(q):
            left = comb(b, (b + q) // 2) if (b + q) % 2 == 0 else 0
            right = comb(L - b, (L - b - q) // 2) if (L - b - q) % 2 == 0 else 0
            return min(left, right)
        w = np.array([np.exp(-q * q / (2 * sigma * sigma)) for q in qs])
        w = w / w.sum()
        dims = [max(1, min(count(q), int(round(total * wi)))) for q, wi in zip(qs, w)]
        # fix the total on the largest sectors
        order = np.argsort(-w)
        k = 0
        while sum(dims) != total and k < 100000:
            j = order[k % len(order)]
            if sum(dims) > total and dims[j] > 1:
                dims[j] -= 1
            elif sum(dims) < total and dims[j] < count(qs[j]):
                dims[j] += 1
            k += 1
        legs.append(Leg.make(zip(qs, dims)))
    return legs


def column_weights(leg: Leg, tail: float) -> Dict[int, np.ndarray]:
    """Per-sector column weights that fall geometrically from 1 to ``tail`` across the leg.

    Sectors are interleaved in proportion to their size: column j of a sector of dimension d_q on
    a leg of total dimension D sits at rank (j + 1/2) D / d_q of the whole leg and gets the weight
    tail ** (rank / D).  Scaling the right-bond columns of a random site tensor by these weights
    gives the state a Schmidt spectrum that decays about as fast (the columns with small weight
    carry little of the state), instead of the flat spectrum of a plain Gaussian MPS.  The weights
    are scaled to a mean square of one (a uniform factor per site, which leaves the normalised
    state unchanged and keeps products along the chain clear of underflow)."""
    out = {}
    for q, d in zip(leg.charges, leg.dims):
        rank = (np.arange(d) + 0.5) * (leg.total / d)
        out[q] = tail ** (rank / leg.total)
    ms = sum(float(np.sum(w * w)) for w in out.values()) / max(leg.total, 1)
    return {q: w / np.sqrt(ms) for q, w in out.items()}


def random_mps(L: int, dmax: int, seed: int = 0, sigma: float = 4.0,
               symmetry: str = "U1", tail: Optional[float] = None) -> SymMPS:
    """Seeded complex-Gaussian U(1) MPS on the synthetic bond legs (not canonicalised).

    ``tail`` (e.g. 1e-8): the columns of every right bond are scaled by ``column_weights``, so the
    Schmidt values fall by about that factor across each bond; None keeps the flat spectrum."""
    rng = np.random.default_rng(seed)
    charges = SPIN_CHARGES_U1 if symmetry == "U1" else SPIN_CHARGES_NONE
    if symmetry == "U1":
        legs = synthetic_bond_legs(L, dmax, sigma)
    else:
        legs = [Leg.dense(min(2 ** min(b, L - b), dmax)) for b in range(L + 1)]
    if tail is not None and not 0 < tail <= 1:
        raise ValueError("tail must lie in (0, 1]")
    tensors = []
    for i in range(L):
        lay = Layout(legs[i], legs[i + 1], charges)
        cw = column_weights(legs[i + 1], tail) if tail is not None and i < L - 1 else None
        blocks = {}
        for k, q, qc, d, dc, off in lay.blocks():
            blk = (rng.standard_normal((d, dc)) + 1j * rng.standard_normal((d, dc))) \
                / np.sqrt(2.0 * max(1, legs[i].total))
            if cw is not None:
                blk = blk * cw[qc][None, :]
            blocks[(k, q)] = blk
        tensors.append(BlockTensor.from_blocks(lay, blocks))
    return SymMPS(tensors, charges, center=0)


def magnetisation(psi: SymMPS, sz=(0.5, -0.5)):
    """<S^z_j> / <psi|psi> for every site (``magnetisation``/``sz_expectation``,
    src/QMB/Observables/local_observables.jl:11-31), any gauge.

    The centre is walked across a copy of the state with unique-QR moves (the copy shares the
    tensors, so ``psi`` itself is untouched); at each site the expectation is read from the
    per-slice weights of the centre tensor."""
    import numpy as np

    from . import tensor_ops as T
    from .bug_step import canonical

    work = psi.copy()
    canonical(work, 0)
    out = np.zeros(len(work))
    for j in range(len(work)):
        if j > 0:
            canonical(work, j)
        A = work[j]
        blks, slices = [], []
        for k, q, qc, d, dc, off in A.layout.blocks():
            blks.append((off, d, d, dc))
            slices.append(k)
        w = T.block_sqnorms(blks, A.data)
        per = np.zeros(A.layout.nslice)
        for k, v in zip(slices, w):
            per[k] += v
        out[j] = float(np.dot(per, sz[:len(per)]) / per.sum()) if per.sum() > 0 else 0.0
    return out
