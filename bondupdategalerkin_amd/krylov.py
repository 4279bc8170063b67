"""Host FP64 pieces of the Krylov solvers (the k x k side; k <= 65) and the device Krylov loops.

Small dense work stays on the host by design, as in the reference: the tridiagonal
eigendecomposition and coefficient evaluation of ``lanczos_expv`` (src/BUG/Core/LocalSolvers/
krylov.jl:44-117,128-263), the substep search (root_methods.jl:86-130) and the residual integral
that weighs each local Krylov direction (src/BUG/Core/RankAdaptation/krylov_enrichment.jl:58-120).
Every vector operation (applies, inner products, axpys, recombination) runs on the device through
``tensor_ops``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np
import scipy.linalg as sla

from . import tensor_ops as T

BREAKDOWN = 1e-13            # invariant-subspace threshold of the local chains
MAXDEPTH = 64                # local Krylov depth bound
ROOT_MAXITER = 30            # root basis per substep
SKEW_TOL = 1e-12             # |tau H + (tau H)^dag| below which the propagator counts as unitary
MAX_PANELS = 4096            # Simpson panels of one residual integral
TAIL_NODES = 8               # nodes of the real-time tail guard
SHIFT_TRIGGER = 30.0         # exponent past which the spectral shift switches on
MAX_NORM = 1e150


def _exp(x: float) -> float:
    """exp that saturates to inf instead of raising (a weight past the double range is infinite,
    as the IEEE exponential gives it)."""
    return math.inf if x > 709.0 else math.exp(x)


def residual_integral(index: int, M: np.ndarray, tau: complex, relative: bool) -> float:
    """|tau| int_0^1 |(exp(sM) e1)[index]| ds by composite Simpson in log space.

    The nodes c_j = S^j e1 (S = exp(h (M - sigma))) are produced a block of B = 64 at a time:
    with row `index` of S, S^2, ..., S^B at hand (vector-matrix products) and S^B (by squaring),
    the integrand at the next B nodes is those rows times the block's first node, and the next
    block starts from S^B times it, so a stiff generator's thousands of panels cost a few hundred
    small products, not one per node.  The integrand and the Simpson weights are those of the node-by-node recurrence;
    the nodes agree with it to rounding."""
    n = M.shape[0]
    unitary = np.linalg.norm(M + M.conj().T) <= SKEW_TOL * np.linalg.norm(M)
    sigma = 0.0 if unitary else float(np.max(np.linalg.eigvals(M).real))
    A = M - sigma * np.eye(n)
    panels = 2 * min(max(math.ceil(2 * np.abs(A).sum(axis=0).max()), 8), MAX_PANELS)
    h = 1.0 / panels
    step = sla.expm(h * A)
    B = 64 if panels >= 64 else 1
    # row `index` of S^1 .. S^B (vector-matrix products) and S^B (by squaring)
    rows = np.empty((B, n), dtype=complex)
    rows[0] = step[index]
    for b in range(1, B):
        rows[b] = rows[b - 1] @ step
    big = step
    for _ in range(B.bit_length() - 1):
        big = big @ big
    mag = np.empty(panels + 1)
    c = np.zeros(n, dtype=complex)
    c[0] = 1.0
    mag[0] = abs(c[index])
    j = 0
    while j + B <= panels:
        mag[j + 1: j + 1 + B] = np.abs(rows @ c)
        c = big @ c
        j += B
    while j < panels:                                     # the last partial block, node by node
        c = step @ c
        j += 1
        mag[j] = abs(c[index])
    w = np.full(panels + 1, 2.0)
    w[1::2] = 4.0
    w[0] = w[panels] = 1.0
    nodes = np.arange(panels + 1) * h * sigma
    with np.errstate(divide="ignore"):
        logw = np.where(mag == 0, -np.inf, np.log(w * (h / 3) * mag) + nodes)
    top = logw.max()
    if top == -np.inf:
        return 0.0
    lw = math.log(abs(tau)) + top + math.log(np.exp(logw - top).sum())
    if unitary or not relative:
        return _exp(lw)
    return _exp(lw - sigma - math.log(max(np.linalg.norm(c), np.finfo(float).tiny)))


def direction_weight(H: np.ndarray, k: int, tau: complex, state_seed: bool = True) -> float:
    """Weight of the (k+1)-th Krylov vector in exp(tau H) v1 (``_kry_weight``); a state seed is
    weighed relative to its own growth on a non-unitary step, a frontier seed absolutely."""
    beta = abs(H[k, k - 1])
    if beta == 0:
        return 0.0
    return beta * residual_integral(k - 1, tau * H[:k, :k], tau, state_seed)


def tail_rank(kept, scale: float, allowance: float, kmax: int, base: float = 0.0):
    """Smallest rank <= kmax whose scaled discarded tail fits (Test A, truncation.jl:262-278)."""
    kmax = min(kmax, len(kept))
    w = [e[0] * e[0] for e in kept]
    if kmax <= 0:
        return 0, scale * math.sqrt(base + sum(w))
    tail = base + sum(w[kmax:])
    r = kmax
    while r > 0 and scale * math.sqrt(tail + w[r - 1]) <= allowance:
        tail += w[r - 1]
        r -= 1
    return r, scale * math.sqrt(tail)


# The rational roots (root_methods.jl:5-71,143-157): one implicit Runge-Kutta step's stability
# function R(z) = N(z) / D(z) applied on the Krylov basis in one stage (coefficients constant first).
ROOT_METHODS = ("exp", "trapezoid", "radau2", "radau3")
ROOT_PADE = {
    "trapezoid": ((1.0, 1 / 2), (1.0, -1 / 2)),
    "radau2": ((1.0, 1 / 3), (1.0, -2 / 3, 1 / 6)),
    "radau3": ((1.0, 2 / 5, 1 / 20), (1.0, -3 / 5, 3 / 20, -1 / 60)),
}


def check_root_method(method: str):
    if method not in ROOT_METHODS:
        raise ValueError(f"root_method must be one of {ROOT_METHODS}, got {method!r}")


def root_stability(method: str, z: complex) -> complex:
    """R(z) of ``method`` (exp(z) for "exp")."""
    check_root_method(method)
    z = complex(z)
    if method == "exp":
        return complex(np.exp(z))
    num, den = ROOT_PADE[method]
    return np.polyval(num[::-1], z) / np.polyval(den[::-1], z)


def rational_apply(method: str, M: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """D(M)^{-1} N(M) rhs for a small dense M (polynomials by Horner)."""
    num, den = ROOT_PADE[method]
    eye = np.eye(M.shape[0], dtype=complex)

    def poly(c):
        out = c[-1] * eye
        for k in range(len(c) - 2, -1, -1):
            out = out @ M + c[k] * eye
        return out

    return np.linalg.solve(poly(den), poly(num) @ rhs)


def rational_shift(vals, tau: complex, shift: bool) -> complex:
    """The Ritz value maximising Re(tau lambda); none without ``shift`` or in real time."""
    if not shift or len(vals) == 0 or complex(tau).real == 0:
        return 0j
    v = np.asarray(vals, dtype=complex)
    return complex(v[int(np.argmax(np.real(tau * v)))])


class Tridiagonal:
    """Eigendecomposition of the Lanczos tridiagonal, evaluated for any trial step: the exponential
    (``method`` "exp") or a rational root's R(tau (T - sigma)) e1 (no scale returned)."""

    def __init__(self, alpha: List[float], beta: List[float], method: str = "exp",
                 shift: bool = True):
        self.method, self.shift = method, shift
        a, b = np.asarray(alpha, float), np.asarray(beta, float)
        if not (np.isfinite(a).all() and np.isfinite(b).all()):
            raise ValueError("Lanczos tridiagonal has non-finite entries: the input vector or "
                             "the operator is not finite")
        if len(a) == 1:
            self.vals, self.vecs = a.copy(), np.ones((1, 1))
        else:
            try:
                self.vals, self.vecs = sla.eigh_tridiagonal(a, b)
            except np.linalg.LinAlgError:
                # the MRRR driver can fail on the graded, nearly decoupled tridiagonal a chain
                # leaves after a near-breakdown; the dense symmetric solver does not
                Tm = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
                self.vals, self.vecs = np.linalg.eigh(Tm)

    def coeffs(self, tau: complex):
        if self.method != "exp":
            sigma = rational_shift(self.vals, tau, self.shift)
            w = np.array([root_stability(self.method, tau * (v - sigma)) for v in self.vals]) \
                * self.vecs[0, :]
            return self.vecs.astype(complex) @ w, 0j
        ex = np.real(tau * self.vals)
        i = int(np.argmax(ex))
        sigma = float(self.vals[i]) if ex[i] > SHIFT_TRIGGER else 0.0
        w = np.exp(tau * (self.vals - sigma)) * self.vecs[0, :]
        return self.vecs.astype(complex) @ w, tau * sigma


class Hessenberg:
    """The leading Arnoldi block, evaluated for any trial step (``_hess_coeffs``,
    root_methods.jl:136-141): exp(tau (Hm - sigma I)) e1 with the spectral shift sigma taken from
    the real part of the eigenvalue that grows fastest under tau."""

    def __init__(self, Hm: np.ndarray, method: str = "exp", shift: bool = True):
        self.method, self.shift = method, shift
        if not np.isfinite(Hm).all():
            raise ValueError("Arnoldi Hessenberg matrix has non-finite entries: the input vector "
                             "or the operator is not finite")
        self.Hm = Hm
        self.vals = np.linalg.eigvals(Hm) if Hm.size else np.zeros(0, complex)

    def coeffs(self, tau: complex):
        if self.Hm.size == 0:
            return np.zeros(0, dtype=complex), 0j
        if self.method != "exp":
            sigma = rational_shift(self.vals, tau, self.shift)
            n = self.Hm.shape[0]
            e1 = np.zeros(n, dtype=complex)
            e1[0] = 1.0
            return rational_apply(self.method, tau * (self.Hm - sigma * np.eye(n)), e1), 0j
        ex = np.real(tau * self.vals)
        i = int(np.argmax(ex))
        sigma = float(self.vals[i].real) if ex[i] > SHIFT_TRIGGER else 0.0
        prop = sla.expm(tau * (self.Hm - sigma * np.eye(self.Hm.shape[0])))
        if not np.isfinite(prop[:, 0]).all():
            raise FloatingPointError("the Hessenberg exponential overflowed")
        return prop[:, 0], tau * sigma


def tail_sup(tri, step: complex, betak: float, above: float = math.inf) -> float:
    tail = betak * abs(tri.coeffs(step)[0][-1])
    if step.real != 0:
        return tail
    for j in range(1, TAIL_NODES):
        if tail >= above:
            return tail
        tail = max(tail, betak * abs(tri.coeffs(step * (j / TAIL_NODES))[0][-1]))
    return tail


def pick_tau(tri, tau_rem: complex, betak: float, conv_tol: float) -> complex:
    def ok(t):
        v = tail_sup(tri, t, betak, conv_tol if conv_tol > 0 else math.inf)
        return math.isfinite(v) and (conv_tol <= 0 or v <= conv_tol)

    if ok(tau_rem):
        return tau_rem
    lo, hi = 0.0, 1.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if ok(mid * tau_rem) else (lo, mid)
    return max(lo, 1e-6) * tau_rem


@dataclass
class ExpvReport:
    iterations: int
    substeps: int
    err: float


def lanczos_expv(apply, tau: complex, v0, *, tol: float = 1e-15, maxiter: int = ROOT_MAXITER,
                 reorth: bool = True, conv_tol: float = 0.0, substeps: int = 0,
                 on_substep: Optional[Callable[[int], None]] = None, root_method: str = "exp",
                 rational_shift: bool = True):
    """exp(tau A) v0 for a Hermitian device action ``apply`` (``lanczos_expv``); a rational
    ``root_method`` applies that implicit method's stability function in one stage."""
    check_root_method(root_method)
    if root_method != "exp":
        substeps = 1
    rk = (root_method, rational_shift)
    beta0 = T.norm(v0)
    if beta0 == 0:
        return v0, ExpvReport(0, 0, 0.0)
    cur, tau_rem, logscale = v0, complex(tau), 0j
    total = nsub = 0
    err = 0.0
    while True:
        nsub += 1
        if on_substep is not None:
            on_substep(nsub)
        cn = T.norm(cur)
        if cn == 0:
            break
        basis = [T.scale(1.0 / cn, cur)]
        alpha: List[float] = []
        beta: List[float] = []
        r = apply(basis[0])
        total += 1
        a = T.inner(basis[0], r).real
        alpha.append(a)
        T.axpy_(-a, basis[0], r)
        for _ in range(2, maxiter + 1):
            b = T.norm(r)
            if b < tol:
                break
            if conv_tol > 0 and tail_sup(Tridiagonal(alpha, beta, *rk), tau_rem, b,
                                         conv_tol) < conv_tol:
                break
            beta.append(b)
            basis.append(T.scale(1.0 / b, r))
            r = apply(basis[-1])
            total += 1
            a = T.inner(basis[-1], r).real
            alpha.append(a)
            T.axpy_(-a, basis[-1], r)
            T.axpy_(-b, basis[-2], r)
            if reorth:
                for _p in range(2):
                    cs = T.inner_many(basis, r)
                    T.axpy_many_([-c for c in cs], basis, r)
        tri = Tridiagonal(alpha, beta, *rk)
        betak = T.norm(r)
        if substeps == 1:
            tau_s = tau_rem
        elif substeps > 1:
            tau_s = tau_rem / (substeps - nsub + 1)
        else:
            tau_s = pick_tau(tri, tau_rem, betak, conv_tol)
        coeff, ls = tri.coeffs(tau_s)
        err += tail_sup(tri, tau_s, betak) * cn / beta0
        out = T.scale(coeff[0] * cn, basis[0])
        m = min(len(coeff), len(basis))
        T.axpy_many_([c * cn for c in coeff[1:m]], basis[1:m], out)
        cur = out
        logscale += ls
        tau_rem -= tau_s
        if substeps == 1 or abs(tau_rem) <= 1e-14 * max(abs(tau), 1.0):
            break
        if nsub >= 10000:
            raise RuntimeError("expv: substep search failed to advance")
    sc = np.exp(logscale)
    if np.isfinite(abs(sc)) and sc != 0:
        n = T.norm(cur)
        if not (np.isfinite(n) and n > 0 and abs(sc) * n > MAX_NORM):
            cur = T.scale(sc, cur)
    return cur, ExpvReport(total, nsub, err)


def arnoldi_expv(apply, tau: complex, v0, *, tol: float = 1e-13, maxiter: int = ROOT_MAXITER,
                 conv_tol: float = 0.0, substeps: int = 0,
                 on_substep: Optional[Callable[[int], None]] = None, root_method: str = "exp",
                 rational_shift: bool = True):
    """exp(tau A) v0 for a non-Hermitian device action ``apply`` (``arnoldi_expv``,
    src/BUG/Core/LocalSolvers/krylov.jl:343-442): upper-Hessenberg Arnoldi with classical
    Gram-Schmidt applied twice per vector (each pass one batched inner-product launch against the
    whole basis), the Lanczos path's tail estimate, substep search and scale bookkeeping; a
    rational ``root_method`` takes the whole step in one stage (krylov.jl:365-370)."""
    check_root_method(root_method)
    if root_method != "exp":
        substeps = 1
    rk = (root_method, rational_shift)
    beta0 = T.norm(v0)
    if beta0 == 0:
        return v0, ExpvReport(0, 0, 0.0)
    cur, tau_rem, logscale = v0, complex(tau), 0j
    total = nsub = 0
    err = 0.0
    while True:
        nsub += 1
        if on_substep is not None:
            on_substep(nsub)
        cn = T.norm(cur)
        if cn == 0:
            break
        basis = [T.scale(1.0 / cn, cur)]
        H = np.zeros((maxiter + 1, maxiter), dtype=complex)
        dim = 0
        betak = 0.0
        for k in range(1, maxiter + 1):
            r = apply(basis[k - 1])
            total += 1
            for _p in range(2):
                cs = T.inner_many(basis, r)
                H[:len(basis), k - 1] += cs
                T.axpy_many_([-c for c in cs], basis, r)
            dim = k
            betak = T.norm(r)
            if betak < tol:
                break
            if conv_tol > 0 and tail_sup(Hessenberg(H[:dim, :dim].copy(), *rk), tau_rem, betak,
                                         conv_tol) < conv_tol:
                break
            H[k, k - 1] = betak
            if k == maxiter:
                break
            basis.append(T.scale(1.0 / betak, r))
        hess = Hessenberg(H[:dim, :dim].copy(), *rk)
        if substeps == 1:
            tau_s = tau_rem
        elif substeps > 1:
            tau_s = tau_rem / (substeps - nsub + 1)
        else:
            tau_s = pick_tau(hess, tau_rem, betak, conv_tol)
        coeff, ls = hess.coeffs(tau_s)
        err += tail_sup(hess, tau_s, betak) * cn / beta0
        out = T.scale(coeff[0] * cn, basis[0])
        T.axpy_many_([c * cn for c in coeff[1:dim]], basis[1:dim], out)
        cur = out
        logscale += ls
        tau_rem -= tau_s
        if substeps == 1 or abs(tau_rem) <= 1e-14 * max(abs(tau), 1.0):
            break
        if nsub >= 10000:
            raise RuntimeError("expv: substep search failed to advance")
    sc = np.exp(logscale)
    if np.isfinite(abs(sc)) and sc != 0:
        n = T.norm(cur)
        if not (np.isfinite(n) and n > 0 and abs(sc) * n > MAX_NORM):
            cur = T.scale(sc, cur)
    return cur, ExpvReport(total, nsub, err)
