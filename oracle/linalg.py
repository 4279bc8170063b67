"""ORACLE (test infrastructure): block products, factorizations and the probe RNG on ``Fam``.

Conventions shared with the device library (both sides implement them independently):

* A site-like tensor ``A(l, s, r)`` is a ``Fam`` with slices ``s`` (shift = charge of s); a
  left frame (l, s, bond) and a right frame (bond, s, r) are the same kind of object.
* Left matricisation of charge sector ``qr``: rows stacked slice by slice (s = 0 first), each
  slice contributing its block at row charge ``qr - q(s)``; right matricisation of ``ql``: the
  slice blocks side by side (s = 0 first).  These fix the fused-basis order of the probes.
* Gauge moves (``move_right!``/``move_left!``/``canonical!``, canonical_form.jl:10-54, and
  ``_frame_from``, effective_operators.jl:201-220) use the unique thin QR with a real positive
  R diagonal instead of Telum's cutoff-0 SVD: both give an isometry spanning the same range and the
  same represented tensor; an all-zero sector block is dropped, as the SVD's ``sv > 0`` rule
  drops it.
* Block SVD follows Telum's selection (deps/Telum/src/svd.jl:1813-1921): per-sector SVDs, keep
  ``sv > cutoff * sv_max`` over all sectors, then the largest ``nkeep`` (stable descending sort
  in sector order); ``kept``/``trunc`` lists are (sv, sector, index) descending.  Each singular
  pair's phase is fixed so that the first entry of ``u`` with ``|u_i| >= 0.5 max |u|`` is real
  positive (LAPACK leaves it arbitrary).
* The probe RNG is counter based (splitmix64 of seed and position, Box-Muller), so a block can be
  drawn at full width and trimmed without materialising the skipped columns.

Pinned by tests/test_oracle_linalg.py: products against dense einsum, QR/SVD reconstruction,
isometry and uniqueness, the selection rule on hand-built spectra, the RNG's moments.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .blocks import Fam

OP_N, OP_T, OP_C, OP_R = 0, 1, 2, 3
CBE_EPS = 100 * np.finfo(float).eps          # _CBE_EPS, truncation.jl:173
# A sector block (or singular value) below this fraction of the tensor's norm (or of the largest
# singular value) is roundoff: gauge moves drop such blocks and selections such values, so that the
# block structure of a result never depends on how the arithmetic rounded (a sector that cancels
# to exactly zero in one evaluation order and to 1e-17 in another).
ZERO_REL = 1e-14


def _op(m, o):
    if o == OP_N:
        return m
    if o == OP_T:
        return m.T
    if o == OP_C:
        return m.conj().T
    return m.conj()


def product(row: Dict[int, int], col: Dict[int, int], shifts: Sequence[int], P: Fam, Q: Fam,
            terms, out: Optional[Fam] = None) -> Fam:
    """out[o] += alpha op_a(P[p]) op_b(Q[q]) over every block of ``out`` (a fresh zero Fam if None)."""
    if out is None:
        out = Fam.zeros(row, col, shifts)
    by_o: Dict[int, list] = {}
    for t in terms:
        by_o.setdefault(t[0], []).append(t)
    for (o, qi), blk in out.blocks.items():
        qj = qi + out.shifts[o]
        for _, p, q, oa, ob, alpha in by_o.get(o, ()):
            if oa in (OP_N, OP_R):
                pb = P.blocks.get((p, qi))
                qk = qi + P.shifts[p]
            else:
                qk = qi - P.shifts[p]
                pb = P.blocks.get((p, qk))
            if pb is None:
                continue
            if ob in (OP_N, OP_R):
                if qk + Q.shifts[q] != qj:
                    continue
                qb = Q.blocks.get((q, qk))
            else:
                if qj + Q.shifts[q] != qk:
                    continue
                qb = Q.blocks.get((q, qj))
            if qb is None:
                continue
            blk += alpha * (_op(pb, oa) @ _op(qb, ob))
    return out


# ---- elementary tensor algebra ------------------------------------------------------------

def core_times_site(C: Fam, A: Fam) -> Fam:
    """(a, b) x (b, s, c) -> (a, s, c)."""
    return product(C.row, A.col, A.shifts, C, A,
                   [(s, 0, s, OP_N, OP_N, 1.0) for s in range(len(A.shifts))])


def site_times_core(A: Fam, C: Fam) -> Fam:
    """(a, s, b) x (b, c) -> (a, s, c)."""
    return product(A.row, C.col, A.shifts, A, C,
                   [(s, s, 0, OP_N, OP_N, 1.0) for s in range(len(A.shifts))])


def overlap_left(U: Fam, A: Fam) -> Fam:
    """U^dag A over (l, s): (bond_U, bond_A)."""
    return product(U.col, A.col, [0], U, A, [(0, s, s, OP_C, OP_N, 1.0) for s in range(len(A.shifts))])


def overlap_right(A: Fam, V: Fam) -> Fam:
    """A V^dag over (s, r): (bond_A, bond_V)."""
    return product(A.row, V.row, [0], A, V, [(0, s, s, OP_N, OP_C, 1.0) for s in range(len(A.shifts))])


def perp_left(U: Fam, K: Fam) -> Fam:
    """K - U (U^dag K)  (perp_component, sectors.jl:215-221)."""
    ov = overlap_left(U, K)
    return K.axpy(-1.0, core_times_site_left(U, ov, K))


def core_times_site_left(U: Fam, ov: Fam, like: Fam) -> Fam:
    """U (l, s, b) x ov (b, c) in the layout of ``like``."""
    out = Fam.zeros(like.row, like.col, like.shifts)
    return product(like.row, like.col, like.shifts, U, ov,
                   [(s, s, 0, OP_N, OP_N, 1.0) for s in range(len(U.shifts))], out=out)


def perp_right(V: Fam, K: Fam) -> Fam:
    """K - (K V^dag) V  (perp_component_right, sectors.jl:230-238)."""
    ov = overlap_right(K, V)
    proj = product(K.row, K.col, K.shifts, ov, V,
                   [(s, 0, s, OP_N, OP_N, 1.0) for s in range(len(V.shifts))],
                   out=Fam.zeros(K.row, K.col, K.shifts))
    return K.axpy(-1.0, proj)


def oplus_left(F: Fam, N: Fam) -> Fam:
    """Direct sum on the column (bond) leg: [F | N] per sector."""
    col = dict(F.col)
    for q, d in N.col.items():
        col[q] = col.get(q, 0) + d
    out = Fam.zeros(F.row, col, F.shifts)
    for (k, q), blk in out.blocks.items():
        qc = q + F.shifts[k]
        a = F.blocks.get((k, q))
        b = N.blocks.get((k, q))
        na = F.col.get(qc, 0)
        if a is not None:
            blk[:, :na] = a
        if b is not None:
            blk[:, na:] = b
    return out


def oplus_right(F: Fam, N: Fam) -> Fam:
    """Direct sum on the row (bond) leg: [F; N] per sector."""
    row = dict(F.row)
    for q, d in N.row.items():
        row[q] = row.get(q, 0) + d
    out = Fam.zeros(row, F.col, F.shifts)
    for (k, q), blk in out.blocks.items():
        a = F.blocks.get((k, q))
        b = N.blocks.get((k, q))
        na = F.row.get(q, 0)
        if a is not None:
            blk[:na, :] = a
        if b is not None:
            blk[na:, :] = b
    return out


# ---- matricisations -------------------------------------------------------------------------

def left_sectors(A: Fam) -> Dict[int, List[Tuple[int, int, int]]]:
    """qr -> [(s, ql, rows)] in stacking order (only slices whose row sector exists)."""
    out: Dict[int, list] = {}
    for qr in A.col:
        parts = []
        for s, sh in enumerate(A.shifts):
            ql = qr - sh
            d = A.row.get(ql, 0)
            if d > 0:
                parts.append((s, ql, d))
        if parts:
            out[qr] = parts
    return out


def right_sectors(A: Fam) -> Dict[int, List[Tuple[int, int, int]]]:
    """ql -> [(s, qr, cols)] in stacking order."""
    out: Dict[int, list] = {}
    for ql in A.row:
        parts = []
        for s, sh in enumerate(A.shifts):
            qr = ql + sh
            d = A.col.get(qr, 0)
            if d > 0:
                parts.append((s, qr, d))
        if parts:
            out[ql] = parts
    return out


def left_matrix(A: Fam, qr: int, parts) -> np.ndarray:
    mats = []
    for s, ql, d in parts:
        b = A.blocks.get((s, ql))
        mats.append(np.zeros((d, A.col[qr]), dtype=complex) if b is None else b)
    return np.vstack(mats)


def right_matrix(A: Fam, ql: int, parts) -> np.ndarray:
    mats = []
    for s, qr, d in parts:
        b = A.blocks.get((s, ql))
        mats.append(np.zeros((A.row[ql], d), dtype=complex) if b is None else b)
    return np.hstack(mats)


def _unique_qr(M: np.ndarray):
    Q, R = np.linalg.qr(M, mode="reduced")
    d = np.diag(R).copy()
    ph = np.where(np.abs(d) > 0, d / np.where(np.abs(d) > 0, np.abs(d), 1), 1.0)
    Q = Q * ph[None, :]
    R = ph.conj()[:, None] * R
    return Q, R


def qr_left(A: Fam) -> Tuple[Fam, Fam]:
    """A = Q R over (l, s) | r with Q a left isometry (new bond leg per sector), R (bond, r)."""
    secs = left_sectors(A)
    qblocks, rblocks, bond = {}, {}, {}
    mats = {qr: left_matrix(A, qr, parts) for qr, parts in secs.items()}
    floor = ZERO_REL * math.sqrt(sum(np.vdot(M, M).real for M in mats.values()))
    for qr, parts in secs.items():
        M = mats[qr]
        if not np.linalg.norm(M) > floor:
            continue
        Q, R = _unique_qr(M)
        k = Q.shape[1]
        bond[qr] = k
        r0 = 0
        for s, ql, d in parts:
            qblocks[(s, ql)] = Q[r0: r0 + d]
            r0 += d
        rblocks[(0, qr)] = R
    Qf = Fam(A.row, bond, A.shifts, qblocks)
    Rf = Fam(bond, A.col, [0], rblocks)
    return Qf, Rf


def qr_right(A: Fam) -> Tuple[Fam, Fam]:
    """A = L Q over l | (s, r) with Q a right isometry: L (l, bond), Q (bond, s, r)."""
    secs = right_sectors(A)
    qblocks, lblocks, bond = {}, {}, {}
    mats = {ql: right_matrix(A, ql, parts) for ql, parts in secs.items()}
    floor = ZERO_REL * math.sqrt(sum(np.vdot(M, M).real for M in mats.values()))
    for ql, parts in secs.items():
        M = mats[ql]
        if not np.linalg.norm(M) > floor:
            continue
        Q, R = _unique_qr(M.conj().T)            # M^H = Q R  ->  M = R^H Q^H
        k = Q.shape[1]
        bond[ql] = k
        Qh = Q.conj().T
        c0 = 0
        for s, qr, d in parts:
            qblocks[(s, ql)] = Qh[:, c0: c0 + d]
            c0 += d
        lblocks[(0, ql)] = R.conj().T
    Lf = Fam(A.row, bond, [0], lblocks)
    Qf = Fam(bond, A.col, A.shifts, qblocks)
    return Lf, Qf


def _fix_phase(U: np.ndarray, Vh: np.ndarray):
    for j in range(U.shape[1]):
        u = U[:, j]
        a = np.abs(u)
        m = a.max() if a.size else 0.0
        if m == 0:
            continue
        i = int(np.argmax(a >= 0.5 * m))
        ph = u[i] / a[i]
        U[:, j] = u * np.conj(ph)
        Vh[j, :] = Vh[j, :] * ph
    return U, Vh


def _select(spectra: Dict[int, np.ndarray], cutoff: float, nkeep: Optional[int]):
    """Telum's global selection; returns (keep count per sector, kept list, trunc list)."""
    entries = []
    smax = 0.0
    for q, sv in spectra.items():
        for j, s in enumerate(sv):
            smax = max(smax, s)
            entries.append((float(s), q, j))
    full = sorted(entries, key=lambda e: -e[0])      # Python sort is stable
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


class SVDResult:
    __slots__ = ("U", "S", "Vh", "kept", "trunc")

    def __init__(self, U, S, Vh, kept, trunc):
        self.U, self.S, self.Vh, self.kept, self.trunc = U, S, Vh, kept, trunc


def _block_svd(M):
    U, s, Vh = np.linalg.svd(M, full_matrices=False)
    U, Vh = _fix_phase(U.copy(), Vh.copy())
    return U, s, Vh


def svd_left(A: Fam, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """A = U S Vh with U over (l, s): U (l, s, bond) left isometry, S {q: sv}, Vh (bond, r)."""
    secs = left_sectors(A)
    facs, spectra = {}, {}
    for qr, parts in secs.items():
        M = left_matrix(A, qr, parts)
        U, s, Vh = _block_svd(M)
        facs[qr] = (U, s, Vh)
        spectra[qr] = s
    keep, kept, trunc = _select(spectra, cutoff, nkeep)
    ub, vb, bond, S = {}, {}, {}, {}
    for qr, (U, s, Vh) in facs.items():
        k = keep[qr]
        if k == 0:
            continue
        bond[qr] = k
        S[qr] = s[:k]
        r0 = 0
        for sl, ql, d in secs[qr]:
            ub[(sl, ql)] = U[r0: r0 + d, :k]
            r0 += d
        vb[(0, qr)] = Vh[:k]
    return SVDResult(Fam(A.row, bond, A.shifts, ub), S, Fam(bond, A.col, [0], vb), kept, trunc)


def svd_right(A: Fam, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """A = U S Vh with Vh over (s, r): U (l, bond), S {q: sv}, Vh (bond, s, r) right isometry."""
    secs = right_sectors(A)
    facs, spectra = {}, {}
    for ql, parts in secs.items():
        M = right_matrix(A, ql, parts)
        U, s, Vh = _block_svd(M)
        facs[ql] = (U, s, Vh)
        spectra[ql] = s
    keep, kept, trunc = _select(spectra, cutoff, nkeep)
    ub, vb, bond, S = {}, {}, {}, {}
    for ql, (U, s, Vh) in facs.items():
        k = keep[ql]
        if k == 0:
            continue
        bond[ql] = k
        S[ql] = s[:k]
        ub[(0, ql)] = U[:, :k]
        c0 = 0
        for sl, qr, d in secs[ql]:
            vb[(sl, ql)] = Vh[:k, c0: c0 + d]
            c0 += d
    return SVDResult(Fam(A.row, bond, [0], ub), S, Fam(bond, A.col, A.shifts, vb), kept, trunc)


def svd_core(C: Fam, cutoff: float = 0.0, nkeep: Optional[int] = None) -> SVDResult:
    """C (a, b) = U S Vh per sector (one slice, shift 0): U (a, bond), Vh (bond, b)."""
    facs, spectra = {}, {}
    for (k, q), M in C.blocks.items():
        U, s, Vh = _block_svd(M)
        facs[q] = (U, s, Vh)
        spectra[q] = s
    keep, kept, trunc = _select(spectra, cutoff, nkeep)
    ub, vb, bond, S = {}, {}, {}, {}
    for q, (U, s, Vh) in facs.items():
        k = keep[q]
        if k == 0:
            continue
        bond[q] = k
        S[q] = s[:k]
        ub[(0, q)] = U[:, :k]
        vb[(0, q)] = Vh[:k]
    return SVDResult(Fam(C.row, bond, [0], ub), S, Fam(bond, C.col, [0], vb), kept, trunc)


def diag_times(S: Dict[int, np.ndarray], F: Fam, side: str = "left") -> Fam:
    """S F (scale rows of a core/right tensor) or F S (scale columns) by singular values."""
    out = Fam(F.row, F.col, F.shifts)
    for (k, q), b in F.blocks.items():
        if side == "left":
            out.blocks[(k, q)] = S[q][:, None] * b
        else:
            out.blocks[(k, q)] = b * S[q + F.shifts[k]][None, :]
    return out


def sq_weight(lst) -> float:
    """_sq_weight, truncation.jl:163."""
    return float(sum(e[0] ** 2 for e in lst))


def trunc_weight(res: SVDResult) -> float:
    """_trunc_weight, truncation.jl:166-170."""
    k, c = sq_weight(res.kept), sq_weight(res.trunc)
    t = k + c
    return 0.0 if t == 0 else math.sqrt(c / t)


# ---- probe RNG ------------------------------------------------------------------------------

_M64 = (1 << 64) - 1


def _splitmix(x: np.ndarray) -> np.ndarray:
    z = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


class ProbeRNG:
    """Counter-based complex Gaussian stream: draw k = position of the next number.

    Value ``i`` of the stream uses two uniforms from splitmix64(seed_key + 2i) and
    splitmix64(seed_key + 2i + 1) (53-bit), Box-Muller, scaled to E|z|^2 = 1 like Julia's
    ``randn(ComplexF64)``.
    """

    def __init__(self, seed: int):
        self.key = np.uint64(int(_splitmix(np.array([seed], dtype=np.uint64))[0]))
        self.pos = 0

    def complex_normal(self, n: int) -> np.ndarray:
        i = np.arange(self.pos, self.pos + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            a = _splitmix(self.key + np.uint64(2) * i)
            b = _splitmix(self.key + np.uint64(2) * i + np.uint64(1))
        u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
        u2 = ((b >> np.uint64(11)).astype(np.float64)) * (1.0 / 9007199254740992.0)
        r = np.sqrt(-np.log(u1))
        self.pos += n
        return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)

    def skip(self, n: int):
        self.pos += n

    def split(self) -> "ProbeRNG":
        """A generator seeded by the next 64-bit word of this one, which advances by one number
        (the parallel right half-sweep's ``MersenneTwister(rand(rng, UInt64))``,
        src/BUG/TimeEvolution/bug_step.jl:238)."""
        with np.errstate(over="ignore"):
            word = _splitmix(np.array([self.key + np.uint64(2) * np.uint64(self.pos)],
                                      dtype=np.uint64))[0]
        self.pos += 1
        return ProbeRNG(int(word))
