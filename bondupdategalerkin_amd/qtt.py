"""Quantics grids, 1D finite-difference chains, the Heston generator as an MPO, and QTT states.

These are the input builders of the Heston pricing workload: one-off host FP64 work per instance,
whose outputs (an ``MPO`` with no symmetry and a canonical ``SymMPS``) feed ``cbe_bug_step``.

Reference surface:

* ``qtt_grid``/``qtt_coords`` (src/QTTS/Tooling/grid.jl:152-177): a uniform 2D grid
  ``x_i = x0 + i dx``, ``dx = (x1 - x0) / 2^n`` (periodic) or ``/(2^n - 1)`` (end points on the
  grid), bits in the GROUPED order with each coordinate least significant bit first: site ``k``
  holds bit ``k`` of ``i_x`` and site ``nx + k`` bit ``k`` of ``i_y`` (grid.jl:6-11).  The other
  orderings of the reference are diagnostics and are refused here.
* ``qtt_shift``/``qtt_derivative`` (src/QTTS/Tooling/derivatives.jl:23-97): the forward shift
  ``(S f)_i = f_{i+1}`` as the binary carry automaton (bond = the carry), central differences
  ``(S - S^T)/2h`` and ``(S - 2I + S^T)/h^2``, and with ``onesided`` boundaries rank-1 patches that
  turn the first and last rows into the second-order one-sided stencils.
* ``qtt_diag_1d``/``qtt_coord_diag`` (derivatives.jl:99-105, operator_assembly.jl:27): the
  coordinate as a rank-2 diagonal operator (built here from the affine automaton, no samples).
* ``heston_op``/``heston_mpo``/``HestonParams``/``feller`` (src/QTTS/Models/Heston/heston.jl:16-69):
  A_H = v/2 d_xx + rho sigma v d_xv + sigma^2 v/2 d_vv + (r - q - v/2) d_x + kappa (theta - v) d_v
  - r, with the variance factor on the rows of the v-derivatives.
* ``tt_from_samples``/``qtt_separable`` (src/QTTS/Tooling/cores.jl:52-82,265-270) and
  ``qtt_mps``/``qtt_to_mpo``/``qtt_cores`` (src/QTTS/Tooling/state_adapter.jl:72-109).

Operator chains are lists of cores ``(w_l, out, in, w_r)``; sums are direct sums followed by one
TT rounding at relative tolerance ``tol`` (the reference compresses at 1e-13).  Because the grid
index of a grouped lsb-first configuration is ``sum_s 2^s b_s`` (site 0 fastest), a full-chain
operator ``X (x) Y`` is the x-sites' chain followed by the y-sites' chain.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

Chain = List[np.ndarray]          # operator cores (w_l, out, in, w_r)
TT = List[np.ndarray]             # state cores (chi_l, 2, chi_r)

OP_TOL = 1e-13
TT_TOL = 1e-14


# ---------------------------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------------------------

@dataclass(frozen=True)
class QTTGrid:
    """``x_i = x0 + i dx`` (i < 2^nx), ``y_j = y0 + j dy`` (j < 2^ny), grouped lsb-first bits."""

    nx: int
    ny: int
    x0: float
    dx: float
    y0: float
    dy: float
    periodic_x: bool
    periodic_y: bool

    @property
    def length(self) -> int:
        return self.nx + self.ny


def qtt_grid(nx: int, ny: int, *, x=(-1.0, 1.0), y=(-1.0, 1.0), periodic_x: bool = True,
             periodic_y: bool = True, order: str = "grouped",
             endian: str = "lsb_first") -> QTTGrid:
    """The grid on ``x[0] <= x < x[1]`` (periodic) or ``x[0] <= x <= x[1]`` (end points on the
    grid), likewise in ``y`` (grid.jl:152-168)."""
    if order != "grouped" or endian != "lsb_first":
        raise ValueError(f"only the grouped lsb-first ordering is built, got {order}/{endian}")
    if nx < 1 or ny < 1:
        raise ValueError("every coordinate of a quantics grid must have n >= 1 bits")
    if not (x[1] > x[0] and y[1] > y[0]):
        raise ValueError(f"empty domain {x} x {y}")
    dx = (x[1] - x[0]) / (2 ** nx if periodic_x else 2 ** nx - 1)
    dy = (y[1] - y[0]) / (2 ** ny if periodic_y else 2 ** ny - 1)
    return QTTGrid(nx, ny, float(x[0]), dx, float(y[0]), dy, periodic_x, periodic_y)


def qtt_coords(g: QTTGrid) -> Tuple[np.ndarray, np.ndarray]:
    return g.x0 + g.dx * np.arange(2 ** g.nx), g.y0 + g.dy * np.arange(2 ** g.ny)


# ---------------------------------------------------------------------------------------------
# TT rounding and operator algebra
# ---------------------------------------------------------------------------------------------

def _rank(S: np.ndarray, delta: float, maxdim: int) -> int:
    """Smallest rank whose discarded tail of squared singular values stays within delta^2
    (``_qtt_rank``, cores.jl:18-26)."""
    r = len(S)
    tail = 0.0
    while r > 1 and tail + S[r - 1] ** 2 <= delta * delta:
        tail += S[r - 1] ** 2
        r -= 1
    return max(min(r, maxdim), 1)


def tt_round(cores: Sequence[np.ndarray], tol: float, maxdim: int = 1 << 30) -> List[np.ndarray]:
    """Quasi-optimal TT rounding: right-to-left orthogonalisation by QR, then a left-to-right
    SVD sweep with the per-bond allowance tol * ||A|| / sqrt(L - 1).  Cores may be 3-dim
    (states) or 4-dim (operators; the two physical legs travel together)."""
    shapes = [c.shape for c in cores]
    C = [c.reshape(c.shape[0], -1, c.shape[-1]) for c in cores]
    L = len(C)
    if L == 1:
        return [c.copy() for c in cores]
    for k in range(L - 1, 0, -1):
        r_l, m, r_r = C[k].shape
        Q, R = np.linalg.qr(C[k].reshape(r_l, m * r_r).T)
        C[k] = Q.T.reshape(-1, m, r_r)
        C[k - 1] = np.einsum("amb,cb->amc", C[k - 1], R)
    nrm = np.linalg.norm(C[0])
    delta = tol * nrm / math.sqrt(L - 1)
    for k in range(L - 1):
        r_l, m, r_r = C[k].shape
        U, S, Vh = np.linalg.svd(C[k].reshape(r_l * m, r_r), full_matrices=False)
        r = _rank(S, delta, maxdim)
        C[k] = U[:, :r].reshape(r_l, m, r)
        C[k + 1] = np.einsum("ab,bmc->amc", S[:r, None] * Vh[:r], C[k + 1])
    out = []
    for c, s in zip(C, shapes):
        out.append(c.reshape((c.shape[0],) + s[1:-1] + (c.shape[-1],)))
    return out


def op_identity(n: int) -> Chain:
    return [np.eye(2).reshape(1, 2, 2, 1) for _ in range(n)]


def op_transpose(A: Chain) -> Chain:
    return [c.transpose(0, 2, 1, 3).copy() for c in A]


def op_shift(n: int, periodic: bool) -> Chain:
    """``(S f)_i = f_{i+1}`` on n bits: the carry automaton, +1 entering at bit 0.  State c is the
    carry into a bit: the input bit is (out + c) mod 2 and the carry leaving is (out + c) div 2.
    Past the top bit the carry either re-enters at bit 0 (periodic) or is dropped, which zeroes
    row 2^n - 1 (open boundary)."""
    if n < 1:
        raise ValueError("a shift needs at least one bit")
    core = np.zeros((2, 2, 2, 2))
    for c in range(2):
        for o in range(2):
            t = o + c
            core[c, o, t % 2, t // 2] = 1.0
    cores = [core.copy() for _ in range(n)]
    cores[0] = cores[0][1:2]
    cores[-1] = cores[-1].sum(axis=3, keepdims=True) if periodic else cores[-1][..., 0:1]
    return cores


def op_ket_bra(n: int, row: int, col: int) -> Chain:
    """The rank-1 operator |row><col| on n bits."""
    cores = []
    for k in range(n):
        c = np.zeros((1, 2, 2, 1))
        c[0, (row >> k) & 1, (col >> k) & 1, 0] = 1.0
        cores.append(c)
    return cores


def op_coordinate(n: int, origin: float, step: float) -> Chain:
    """diag(origin + step i) on n bits: state 1 = 'not yet added', state 0 = 'added'; bit k
    adds step 2^k when set, the origin enters at bit 0.  Exact rank 2."""
    nb = np.diag([0.0, 1.0])
    eye = np.eye(2)
    cores = []
    for k in range(n):
        c = np.zeros((2, 2, 2, 2))
        c[0, :, :, 0] = eye
        c[1, :, :, 1] = eye
        c[1, :, :, 0] = step * (2 ** k) * nb + (origin * eye if k == 0 else 0.0)
        cores.append(c)
    cores[0] = cores[0][1:2]
    cores[-1] = cores[-1][..., 0:1]
    return cores


def op_direct_sum(terms: Sequence[Tuple[float, Chain]]) -> Chain:
    """sum_t c_t A_t as one chain whose bonds are the direct sums of the terms' bonds."""
    terms = [(c, A) for c, A in terms if c != 0]
    if not terms:
        raise ValueError("empty operator sum")
    n = len(terms[0][1])
    if any(len(A) != n for _, A in terms):
        raise ValueError("operator sum over chains of different lengths")
    if n == 1:
        return [sum(c * A[0] for c, A in terms)]
    out = []
    for k in range(n):
        cores = [A[k] * (c if k == 0 else 1.0) for c, A in terms]
        wl = 1 if k == 0 else sum(x.shape[0] for x in cores)
        wr = 1 if k == n - 1 else sum(x.shape[3] for x in cores)
        big = np.zeros((wl, 2, 2, wr), dtype=np.result_type(*cores))
        a = b = 0
        for x in cores:
            ra = slice(0, 1) if k == 0 else slice(a, a + x.shape[0])
            rb = slice(0, 1) if k == n - 1 else slice(b, b + x.shape[3])
            big[ra, :, :, rb] += x
            a += x.shape[0]
            b += x.shape[3]
        out.append(big)
    return out


def op_lincomb(terms: Sequence[Tuple[float, Chain]], tol: float = OP_TOL) -> Chain:
    return tt_round(op_direct_sum(terms), tol)


def op_compose(A: Chain, B: Chain, tol: float = OP_TOL) -> Chain:
    """The product A B (B applied first), rounded."""
    out = []
    for a, b in zip(A, B):
        c = np.einsum("aomb,cmid->acoibd", a, b)
        out.append(c.reshape(a.shape[0] * b.shape[0], 2, 2, a.shape[3] * b.shape[3]))
    return tt_round(out, tol)


def op_dense(A: Chain) -> np.ndarray:
    """The 2^n x 2^n matrix of a chain, index sum_k 2^k bit_k (small chains only)."""
    n = len(A)
    if n > 14:
        raise ValueError("op_dense is a small-chain check")
    acc = A[0][0]                                            # (out, in, w)
    dim = 2
    for k in range(1, n):
        c = A[k]
        acc = np.einsum("oia,apqb->poqib", acc, c).reshape(2 * dim, 2 * dim, c.shape[3])
        dim *= 2
    return acc[:, :, 0]


def qtt_derivative(n: int, h: float, order: int = 1, bc: str = "periodic") -> Chain:
    """Second-order central differences on 2^n points of spacing h (derivatives.jl:53-97):
    ``periodic`` wraps, ``dirichlet0`` is zero outside the box, ``onesided`` patches the first
    and last rows to the second-order one-sided stencils [-3, 4, -1]/2h, [2, -5, 4, -1]/h^2 and
    their mirror images."""
    if bc not in ("periodic", "dirichlet0", "onesided"):
        raise ValueError(f"bc must be periodic, dirichlet0 or onesided, got {bc!r}")
    if order not in (1, 2):
        raise ValueError(f"order must be 1 or 2, got {order}")
    fwd = op_shift(n, bc == "periodic")
    bwd = op_transpose(fwd)
    if order == 1:
        terms = [(1 / (2 * h), fwd), (-1 / (2 * h), bwd)]
    else:
        terms = [(1 / h ** 2, fwd), (-2 / h ** 2, op_identity(n)), (1 / h ** 2, bwd)]
    if bc == "onesided":
        N = 2 ** n
        if N < 4:
            raise ValueError(f"onesided stencils need at least 4 points, got {N}")
        e = N - 1
        # rows 0 and N - 1: replace what the central stencil left there by the one-sided row
        if order == 1:
            patch = [(0, 0, -3.0), (0, 1, 3.0), (0, 2, -1.0),
                     (e, e, 3.0), (e, e - 1, -3.0), (e, e - 2, 1.0)]
            scale = 1 / (2 * h)
        else:
            patch = [(0, 0, 4.0), (0, 1, -6.0), (0, 2, 4.0), (0, 3, -1.0),
                     (e, e, 4.0), (e, e - 1, -6.0), (e, e - 2, 4.0), (e, e - 3, -1.0)]
            scale = 1 / h ** 2
        terms += [(v * scale, op_ket_bra(n, r, c)) for r, c, v in patch]
    return op_lincomb(terms)


# ---------------------------------------------------------------------------------------------
# the Heston generator
# ---------------------------------------------------------------------------------------------

@dataclass(frozen=True)
class HestonParams:
    """Rate, dividend yield, mean reversion, long-run variance, vol of variance, correlation."""

    r: float
    q: float
    kappa: float
    theta: float
    sigma: float
    rho: float


def feller(p: HestonParams) -> bool:
    """2 kappa theta >= sigma^2 (the variance stays positive)."""
    return 2 * p.kappa * p.theta >= p.sigma ** 2


def _embed(g: QTTGrid, X: Optional[Chain], Y: Optional[Chain]) -> Chain:
    return list(X if X is not None else op_identity(g.nx)) + \
        list(Y if Y is not None else op_identity(g.ny))


def heston_op(g: QTTGrid, p: HestonParams, adjoint: bool = False, bc: str = "onesided",
              tol: float = OP_TOL) -> Chain:
    """A_H on the grid (x = log price, y = variance from v = 0, end points on the grid), or with
    ``adjoint`` the forward-density operator A_H^T without the discount (heston.jl:36-60).

    The one rounding of the eight-term sum brings the bond between the x and v sites down to
    rank 3: the terms have only three distinct x factors (d_xx, d_x, I)."""
    if g.periodic_x or g.periodic_y:
        raise ValueError("heston_op: build the grid with periodic_x = periodic_y = False "
                         "(end points on the grid)")
    if abs(g.y0) >= 1e-14:
        raise ValueError("heston_op: the variance grid must start at v = 0")
    d1x, d2x = qtt_derivative(g.nx, g.dx, 1, bc), qtt_derivative(g.nx, g.dx, 2, bc)
    d1v, d2v = qtt_derivative(g.ny, g.dy, 1, bc), qtt_derivative(g.ny, g.dy, 2, bc)
    var = op_coordinate(g.ny, g.y0, g.dy)
    var_d1v = op_compose(var, d1v, tol)
    var_d2v = op_compose(var, d2v, tol)
    terms = [
        (0.5, d2x, var),
        (p.rho * p.sigma, d1x, var_d1v),
        (0.5 * p.sigma ** 2, None, var_d2v),
        (p.r - p.q, d1x, None),
        (-0.5, d1x, var),
        (p.kappa * p.theta, None, d1v),
        (-p.kappa, None, var_d1v),
    ]
    if not adjoint:
        terms.append((-p.r, None, None))
    gen = op_lincomb([(c, _embed(g, X, Y)) for c, X, Y in terms], tol)
    return op_transpose(gen) if adjoint else gen


def qtt_to_mpo(W: Chain):
    """The solver MPO (no symmetry, local dimension 2) of operator cores (w_l, out, in, w_r):
    every nonzero core entry becomes one sparse MPO entry (state_adapter.jl:92-106)."""
    from .mpo import MPO, MPOSite, SPIN_CHARGES_NONE

    if len(W) < 2:
        raise ValueError("qtt_to_mpo needs at least two sites")
    sites = []
    for c in W:
        a, so, si, b = np.nonzero(c)
        coeff = np.asarray(c[a, so, si, b], dtype=np.complex128)
        z = lambda n: np.zeros(n, dtype=np.int64)                         # noqa: E731
        sites.append(MPOSite(z(c.shape[0]), z(c.shape[3]), a.astype(np.int64),
                             b.astype(np.int64), so.astype(np.int64), si.astype(np.int64), coeff))
    return MPO(sites, SPIN_CHARGES_NONE)


def heston_mpo(g: QTTGrid, p: HestonParams, **kw):
    """``heston_op`` as the solver MPO (heston.jl:62-69)."""
    return qtt_to_mpo(heston_op(g, p, **kw))


# ---------------------------------------------------------------------------------------------
# states
# ---------------------------------------------------------------------------------------------

def tt_from_samples(vals, tol: float = TT_TOL, maxdim: int = 1 << 30) -> TT:
    """The 1D quantics TT of ``vals[i] = f(x_i)`` (bit 0 first) by TT-SVD with the per-bond
    allowance tol ||vals|| / sqrt(n - 1) (cores.jl:52-82)."""
    vals = np.asarray(vals)
    n = int(round(math.log2(len(vals))))
    if 2 ** n != len(vals):
        raise ValueError(f"need 2^n samples, got {len(vals)}")
    A = vals.reshape((2,) * n, order="F")                    # axis k = bit k
    if n == 1:
        return [A.reshape(1, 2, 1).copy()]
    delta = tol * np.linalg.norm(vals) / math.sqrt(n - 1)
    cores = []
    M = A.reshape(2, -1, order="F")
    chi = 1
    for k in range(n - 1):
        M = M.reshape(chi * 2, -1, order="F")
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        r = _rank(S, delta, maxdim)
        cores.append(U[:, :r].reshape(chi, 2, r, order="F"))
        M = S[:r, None] * Vh[:r]
        chi = r
    cores.append(M.reshape(chi, 2, 1, order="F"))
    return cores


def qtt_separable(g: QTTGrid, fx: Optional[Callable], fy: Optional[Callable],
                  tol: float = TT_TOL, maxdim: int = 1 << 30) -> TT:
    """``fx(x) fy(y)`` from two 1D TT-SVDs; ``None`` is the constant 1 (cores.jl:265-270)."""
    xs, ys = qtt_coords(g)
    ones = lambda n: [np.ones((1, 2, 1)) for _ in range(n)]                  # noqa: E731
    tx = tt_from_samples(np.vectorize(fx)(xs), tol, maxdim) if fx is not None else ones(g.nx)
    ty = tt_from_samples(np.vectorize(fy)(ys), tol, maxdim) if fy is not None else ones(g.ny)
    return tx + ty


def tt_dense(cores: TT) -> np.ndarray:
    """The grid vector of a TT, index sum_s 2^s b_s (site 0 fastest); small chains only."""
    v = cores[0][0]                                          # (2, chi)
    dim = 2
    for c in cores[1:]:
        v = np.einsum("xa,asb->sxb", v, c).reshape(2 * dim, c.shape[2])
        dim *= 2
    return v[:, 0]


def qtt_values(cores: TT, g: QTTGrid) -> np.ndarray:
    """The field as a matrix ``M[ix, iy]`` (small grids only)."""
    return tt_dense(cores).reshape(2 ** g.ny, 2 ** g.nx).T


def qtt_mps(cores: TT, center: int = 0):
    """The state of TT cores (chi_l, 2, chi_r) as a device ``SymMPS`` with no symmetry, gauged
    to ``center`` by QR moves from the right end (state_adapter.jl:72-86)."""
    from .blocks import BlockTensor, Layout, Leg
    from .bug_step import canonical
    from .mps import SymMPS
    from .mpo import SPIN_CHARGES_NONE

    L = len(cores)
    if not 0 <= center < L:
        raise ValueError(f"center {center} outside 0..{L - 1}")
    tensors = []
    for i, c in enumerate(cores):
        if i > 0 and c.shape[0] != cores[i - 1].shape[2]:
            raise ValueError(f"bond {i}: core dimensions do not chain")
        lay = Layout(Leg.dense(c.shape[0]), Leg.dense(c.shape[2]), SPIN_CHARGES_NONE)
        blocks = {(s, 0): np.asarray(c[:, s, :], dtype=np.complex128) for s in range(2)}
        tensors.append(BlockTensor.from_blocks(lay, blocks))
    psi = SymMPS(tensors, SPIN_CHARGES_NONE, center=L - 1)
    canonical(psi, 0)
    canonical(psi, center)
    return psi


def qtt_cores(psi) -> TT:
    """The TT cores (chi_l, 2, chi_r) of a no-symmetry ``SymMPS``, any gauge (downloaded)."""
    out = []
    for t in psi.tensors:
        lay = t.layout
        blocks = t.to_blocks()
        c = np.zeros((lay.row.total, 2, lay.col.total), dtype=np.complex128)
        for (s, _q), b in blocks.items():
            c[:, s, :] = b
        out.append(c)
    return out
