"""ORACLE (test infrastructure): the Heston generator as a dense matrix, and QTT conversions.

* ``stencil``      -- the dense 1D second-order difference matrix with ``onesided``,
                      ``periodic`` or ``dirichlet0`` rows, the dense form of ``qtt_derivative``
                      (src/QTTS/Tooling/derivatives.jl:53-97) as the reference's own tests state it
                      (test/common/qtt_fixtures.jl:39-66).
* ``heston_dense`` -- A_H term by term as Kronecker products ``kron(A_v, A_x)`` on the grid index
                      ``ix + 2^nx iy`` (test/QTTS/Models/Heston/test_heston.jl:14-25, the operator
                      of src/QTTS/Models/Heston/heston.jl:36-60).
* ``affine_rate``  -- the manufactured solution A_H exp(a x + b v) = lambda(v) exp(a x + b v)
                      (test_heston.jl:62-88).
* conversions between TT cores / operator cores and the oracle's block families, and the dense
  grid vector of an oracle MPS (site 0 the least significant bit).

Pinned: ``heston_dense`` with the frozen backward-pricing run's parameters, grid and payoff,
exponentiated densely over the run's time span, reproduces its 1024 frozen final amplitudes
(test/fixtures/b155550/qtt_heston_backward.json; the run's setup is in
test/QTTS/Tooling/test_nonlinear_step.jl:159-176); every stencil annihilates constants and is exact
on quadratics, and the manufactured affine solution converges at second order.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from .blocks import Fam, MpoSite


def stencil(N: int, h: float, order: int, bc: str) -> np.ndarray:
    """Dense first (order 1) or second (order 2) difference matrix on N points of spacing h."""
    D = np.zeros((N, N))
    central = [(1, 1 / (2 * h)), (-1, -1 / (2 * h))] if order == 1 else \
        [(1, 1 / h ** 2), (0, -2 / h ** 2), (-1, 1 / h ** 2)]
    for row in range(N):
        for off, w in central:
            col = row + off
            if bc == "periodic":
                col %= N
            elif not 0 <= col < N:
                continue
            D[row, col] += w
    if bc == "onesided":
        D[0, :] = 0
        D[N - 1, :] = 0
        if order == 1:
            D[0, :3] = np.array([-3, 4, -1]) / (2 * h)
            D[N - 1, N - 3:] = np.array([1, -4, 3]) / (2 * h)
        else:
            D[0, :4] = np.array([2, -5, 4, -1]) / h ** 2
            D[N - 1, N - 4:] = np.array([-1, 4, -5, 2]) / h ** 2
    return D


def grid(nx: int, ny: int, x=(-1.0, 1.0), y=(0.0, 0.5)):
    """Coordinates and spacings of the end-points-on-the-grid grid."""
    dx = (x[1] - x[0]) / (2 ** nx - 1)
    dy = (y[1] - y[0]) / (2 ** ny - 1)
    return x[0] + dx * np.arange(2 ** nx), y[0] + dy * np.arange(2 ** ny), dx, dy


def heston_dense(nx: int, ny: int, params, x=(-1.0, 1.0), y=(0.0, 0.5)) -> np.ndarray:
    """A_H on the grid index ix + 2^nx iy; params = (r, q, kappa, theta, sigma, rho)."""
    r, q, kappa, theta, sigma, rho = params
    xs, vs, dx, dv = grid(nx, ny, x, y)
    Nx, Ny = len(xs), len(vs)
    D1x, D2x = stencil(Nx, dx, 1, "onesided"), stencil(Nx, dx, 2, "onesided")
    D1v, D2v = stencil(Ny, dv, 1, "onesided"), stencil(Ny, dv, 2, "onesided")
    V = np.diag(vs)
    Ix, Iv = np.eye(Nx), np.eye(Ny)
    A = (0.5 * np.kron(V, D2x) + rho * sigma * np.kron(V @ D1v, D1x)
         + 0.5 * sigma ** 2 * np.kron(V @ D2v, Ix) + (r - q) * np.kron(Iv, D1x)
         - 0.5 * np.kron(V, D1x) + kappa * theta * np.kron(D1v, Ix)
         - kappa * np.kron(V @ D1v, Ix) - r * np.eye(Nx * Ny))
    return A


def affine_rate(v, a: float, b: float, params):
    """lambda(v) with A_H exp(a x + b v) = lambda(v) exp(a x + b v) for the continuous A_H."""
    r, q, kappa, theta, sigma, rho = params
    return ((a * a - a) * v / 2 + rho * sigma * a * b * v + sigma ** 2 * b * b * v / 2
            + (r - q) * a + kappa * (theta - v) * b - r)


def tt_grid_vector(cores: Sequence[np.ndarray]) -> np.ndarray:
    """Grid vector (index sum_s 2^s b_s) of TT cores (chi_l, 2, chi_r)."""
    v = np.asarray(cores[0])[0]
    for c in cores[1:]:
        v = np.einsum("xa,asb->sxb", v, c).reshape(-1, c.shape[2])
    return v[:, 0]


def mps_grid_vector(sites: List[Fam]) -> np.ndarray:
    """Grid vector of an oracle MPS with no symmetry (each site one family of dense blocks)."""
    cores = []
    for f in sites:
        Dl, Dr = sum(f.row.values()), sum(f.col.values())
        c = np.zeros((Dl, 2, Dr), dtype=np.complex128)
        for (s, _q), b in f.blocks.items():
            c[:, s, :] = b
        cores.append(c)
    return tt_grid_vector(cores)


def fams_from_cores(cores: Sequence[np.ndarray]) -> List[Fam]:
    """TT cores as oracle block families (one dense block per local state, charge 0)."""
    out = []
    for c in cores:
        blocks = {(s, 0): np.array(c[:, s, :], dtype=np.complex128) for s in range(2)}
        out.append(Fam({0: c.shape[0]}, {0: c.shape[2]}, [0, 0], blocks))
    return out


def mpo_from_cores(W: Sequence[np.ndarray]) -> List[MpoSite]:
    """Operator cores (w_l, out, in, w_r) as oracle MPO sites, one entry per nonzero."""
    sites = []
    for c in W:
        ents = [(a, b, so, si, c[a, so, si, b]) for a, so, si, b in zip(*np.nonzero(c))]
        sites.append(MpoSite([0] * c.shape[0], [0] * c.shape[3], ents))
    return sites


def mpo_grid_matrix(W: Sequence[np.ndarray]) -> np.ndarray:
    """The dense matrix of operator cores on the grid index (small chains only)."""
    acc = np.asarray(W[0])[0]
    dim = 2
    for c in W[1:]:
        acc = np.einsum("oia,apqb->poqib", acc, c).reshape(2 * dim, 2 * dim, c.shape[3])
        dim *= 2
    return acc[:, :, 0]
