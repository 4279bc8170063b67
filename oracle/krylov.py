"""ORACLE (test infrastructure): the small dense pieces of the Krylov solvers (host FP64).

* ``residual_integral``/``kry_weight``  -- krylov_enrichment.jl:58-120
* ``tail_rank``                          -- truncation.jl:262-278
* ``tridiag_eigen``, ``coeffs_from_eigen``, ``shift`` -- src/BUG/Core/LocalSolvers/krylov.jl:44-117
* ``tail_sup``, ``pick_tau``             -- src/BUG/Core/LocalSolvers/root_methods.jl:86-130
* ``lanczos_expv``                       -- krylov.jl:128-263 (``:exp`` root, ``reorth``, adaptive
                                            substeps), on any vector type through callbacks
* rational roots ``root_stability``, ``rational_apply``, ``rational_coeffs``, ``hess_rational``
                                         -- root_methods.jl:5-71,143-157 (trapezoid, Radau IIA)

Pinned by tests/test_oracle_krylov.py: expv against ``scipy.linalg.expm`` on random Hermitian
matrices, weights against a dense exponential, the 2x2 cancellation case of
test/BUG/Core/LocalSolvers/test_krylov_weights.jl:42-59, the tail-rank rule on hand spectra.
"""
from __future__ import annotations

import math
from typing import Callable, List

import numpy as np
import scipy.linalg as sla

KRY_BREAKDOWN = 1e-13          # _KRY_BREAKDOWN, krylov_enrichment.jl:13
KRY_MAXDEPTH = 64              # _KRY_MAXDEPTH, :24
ROOT_MAXITER = 30              # _ROOT_MAXITER, :28
KRY_SKEW_TOL = 1e-12           # _KRY_SKEW_TOL, :33
RESIDUAL_MAX_PANELS = 4096     # :36
TAIL_NODES = 8                 # _TAIL_NODES, root_methods.jl:83
EXPV_SHIFT_TRIGGER = 30.0      # krylov.jl:26
EXPV_MAX_NORM = 1e150


def _exp(x: float) -> float:
    """exp that saturates to inf instead of raising (a weight past the double range is infinite,
    as the IEEE exponential gives it)."""
    return math.inf if x > 709.0 else math.exp(x)


def residual_integral(defect_norm: Callable, M: np.ndarray, tau: complex,
                      state_seed: bool = True) -> float:
    nM = np.linalg.norm(M)
    unitary = np.linalg.norm(M + M.conj().T) <= KRY_SKEW_TOL * nM
    sigma = 0.0 if unitary else float(np.max(np.linalg.eigvals(M).real))
    shifted = M - sigma * np.eye(M.shape[0])
    intervals = 2 * min(max(math.ceil(2 * np.abs(shifted).sum(axis=0).max()), 8),
                        RESIDUAL_MAX_PANELS)
    h = 1.0 / intervals
    prop = sla.expm(h * shifted)
    col = np.zeros(M.shape[0], dtype=complex)
    col[0] = 1.0
    logs = np.empty(intervals + 1)
    for node in range(intervals + 1):
        if node > 0:
            col = prop @ col
        w = 1.0 if node in (0, intervals) else (4.0 if node % 2 == 1 else 2.0)
        mag = defect_norm(col)
        logs[node] = -np.inf if mag == 0 else math.log(w * h / 3 * mag) + node * h * sigma
    peak = logs.max()
    if peak == -np.inf:
        return 0.0
    lw = math.log(abs(tau)) + peak + math.log(np.exp(logs - peak).sum())
    if unitary or not state_seed:
        return _exp(lw)
    return _exp(lw - sigma - math.log(max(np.linalg.norm(col), np.finfo(float).tiny)))


def kry_weight(Hm: np.ndarray, k: int, tau: complex, state_seed: bool = True) -> float:
    """Weight of v_{k+1} (k = 1-based step), Hm[k, k-1] (0-based) = beta_k."""
    beta = abs(Hm[k, k - 1])
    if beta == 0:
        return 0.0
    return beta * residual_integral(lambda c: abs(c[k - 1]), tau * Hm[:k, :k], tau, state_seed)


def tail_rank(kept, scale: float, allowance: float, kmax: int, base: float = 0.0):
    """_tail_rank: smallest keep <= kmax whose scaled tail fits; (keep, tail)."""
    n = len(kept)
    kmax = min(kmax, n)
    w = [e[0] ** 2 for e in kept]
    if kmax <= 0:
        return 0, scale * math.sqrt(base + sum(w))
    tail = base + sum(w[kmax:])
    keep = kmax
    while keep > 0:
        trial = tail + w[keep - 1]
        if not scale * math.sqrt(trial) <= allowance:
            break
        tail = trial
        keep -= 1
    return keep, scale * math.sqrt(tail)


def tridiag_eigen(alpha: List[float], beta: List[float]):
    a = np.asarray(alpha, dtype=float)
    b = np.asarray(beta, dtype=float)
    if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
        raise ValueError("Lanczos tridiagonal has non-finite entries")
    T = np.diag(a)
    if len(b):
        T += np.diag(b, 1) + np.diag(b, -1)
    return np.linalg.eigh(T)


def shift(vals, tau: complex) -> float:
    if len(vals) == 0:
        return 0.0
    ex = np.real(tau * np.asarray(vals, dtype=complex))
    i = int(np.argmax(ex))
    return float(np.real(vals[i])) if ex[i] > EXPV_SHIFT_TRIGGER else 0.0


def coeffs_from_eigen(vals, vecs, tau: complex):
    sigma = shift(vals, tau)
    w = np.exp(tau * (vals - sigma)) * vecs[0, :]
    return vecs.astype(complex) @ w, tau * sigma


def _tail(coeff, betak):
    return 0.0 if len(coeff) == 0 else betak * abs(coeff[-1])


def tail_sup(evalcoeff, step: complex, betak: float, above: float = math.inf) -> float:
    tail = _tail(evalcoeff(step)[0], betak)
    if step.real != 0:
        return tail
    for node in range(1, TAIL_NODES):
        if tail >= above:
            return tail
        tail = max(tail, _tail(evalcoeff(step * (node / TAIL_NODES))[0], betak))
    return tail


def pick_tau(evalcoeff, tau_rem: complex, betak: float, conv_tol: float) -> complex:
    def accepted(t):
        tl = tail_sup(evalcoeff, t, betak, above=conv_tol if conv_tol > 0 else math.inf)
        return math.isfinite(tl) and (conv_tol <= 0 or tl <= conv_tol)

    if accepted(tau_rem):
        return tau_rem
    lo, hi = 0.0, 1.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if accepted(mid * tau_rem):
            lo = mid
        else:
            hi = mid
    return max(lo, 1e-6) * tau_rem


# ---- the rational roots (root_methods.jl:5-71,143-157) ----------------------------------------------
# One implicit Runge-Kutta step's stability function R(z) = N(z) / D(z), applied on the Krylov
# basis in one stage; coefficients constant term first.
ROOT_METHODS = ("exp", "trapezoid", "radau2", "radau3")
ROOT_PADE = {
    "trapezoid": ([1.0, 1 / 2], [1.0, -1 / 2]),
    "radau2": ([1.0, 1 / 3], [1.0, -2 / 3, 1 / 6]),
    "radau3": ([1.0, 2 / 5, 1 / 20], [1.0, -3 / 5, 3 / 20, -1 / 60]),
}


def check_root_method(method: str):
    if method not in ROOT_METHODS:
        raise ValueError(f"root_method must be one of {ROOT_METHODS}, got {method!r}")


def root_stability(method: str, z: complex) -> complex:
    """R(z) of ``method`` (exp(z) for "exp")."""
    check_root_method(method)
    if method == "exp":
        return complex(np.exp(complex(z)))
    num, den = ROOT_PADE[method]
    z = complex(z)
    return sum(c * z ** k for k, c in enumerate(num)) / sum(c * z ** k for k, c in enumerate(den))


def rational_apply(method: str, M: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """D(M)^{-1} N(M) rhs for a small dense M (the polynomials by Horner)."""
    num, den = ROOT_PADE[method]
    n = M.shape[0]

    def poly(c):
        out = c[-1] * np.eye(n, dtype=complex)
        for k in range(len(c) - 2, -1, -1):
            out = out @ M + c[k] * np.eye(n)
        return out

    return np.linalg.solve(poly(den), poly(num) @ rhs)


def rational_shift(vals, tau: complex, shift: bool = True) -> complex:
    """The Ritz value maximising Re(tau lambda); none without ``shift``, for no values, or when tau
    is purely imaginary (real time)."""
    if not shift or len(vals) == 0 or complex(tau).real == 0:
        return 0j
    v = np.asarray(vals, dtype=complex)
    return complex(v[int(np.argmax(np.real(tau * v)))])


def rational_coeffs(vals, vecs, tau: complex, method: str, shift: bool = True):
    """R(tau (T - sigma)) e_1 from the tridiagonal's eigendecomposition; no scale returned."""
    sigma = rational_shift(vals, tau, shift)
    w = np.array([root_stability(method, tau * (complex(v) - sigma)) for v in vals]) \
        * vecs[0, :].astype(complex)
    return vecs.astype(complex) @ w, 0j


def hess_rational(Hm: np.ndarray, tau: complex, method: str, shift: bool = True):
    """R(tau (Hm - sigma)) e_1 on the Arnoldi Hessenberg; no scale returned."""
    if Hm.size == 0:
        return np.zeros(0, dtype=complex), 0j
    sigma = rational_shift(np.linalg.eigvals(Hm), tau, shift)
    n = Hm.shape[0]
    e1 = np.zeros(n, dtype=complex)
    e1[0] = 1.0
    return rational_apply(method, tau * (Hm - sigma * np.eye(n)), e1), 0j


class ExpvReport:
    __slots__ = ("iterations", "substeps", "err")

    def __init__(self, it, ns, err):
        self.iterations, self.substeps, self.err = it, ns, err


def lanczos_expv(apply, tau: complex, vector, *, inner, norm, axpy, scale,
                 tol: float = 1e-13, maxiter: int = 60, reorth: bool = False,
                 conv_tol: float = 0.0, substeps: int = 1, root_method: str = "exp",
                 rational_shift: bool = True):
    """exp(tau A) vector for Hermitian ``apply``; vector algebra through the callbacks.

    ``inner(a, b)`` = <a|b>, ``norm(a)``, ``axpy(alpha, x, y)`` = y + alpha x (new object),
    ``scale(alpha, x)`` = alpha x.  ``root_method`` other than "exp" applies that implicit
    method's stability function in one stage (krylov.jl:153-160).  Returns (result, ExpvReport).
    """
    check_root_method(root_method)
    if root_method != "exp":
        substeps = 1

    def tri_coeffs(vals, vecs, t):
        if root_method == "exp":
            return coeffs_from_eigen(vals, vecs, t)
        return rational_coeffs(vals, vecs, t, root_method, rational_shift)
    beta0 = norm(vector)
    if beta0 == 0:
        return vector, ExpvReport(0, 0, 0.0)
    cur = vector
    tau_rem = complex(tau)
    logscale = 0j
    total = nsub = 0
    err = 0.0
    while True:
        nsub += 1
        cur_norm = norm(cur)
        if cur_norm == 0:
            break
        v = scale(1.0 / cur_norm, cur)
        basis = [v]
        alpha: List[float] = []
        betas: List[float] = []
        res = apply(v)
        total += 1
        a = inner(v, res).real
        alpha.append(a)
        res = axpy(-a, v, res)
        for _ in range(2, maxiter + 1):
            b = norm(res)
            if b < tol:
                break
            if conv_tol > 0:
                vals, vecs = tridiag_eigen(alpha, betas)
                tt = tail_sup(lambda t: tri_coeffs(vals, vecs, t), tau_rem, b, above=conv_tol)
                if tt < conv_tol:
                    break
            betas.append(b)
            v = scale(1.0 / b, res)
            basis.append(v)
            res = apply(v)
            total += 1
            a = inner(v, res).real
            alpha.append(a)
            res = axpy(-b, basis[-2], axpy(-a, v, res))
            if reorth:
                for _p in range(2):
                    coeffs = [inner(p, res) for p in basis]
                    for c, p in zip(coeffs, basis):
                        res = axpy(-c, p, res)
        vals, vecs = tridiag_eigen(alpha, betas)

        def evalcoeff(t, vals=vals, vecs=vecs):
            return tri_coeffs(vals, vecs, complex(t))

        betak = norm(res)
        if substeps == 1:
            tau_s = tau_rem
        elif substeps > 1:
            tau_s = tau_rem / (substeps - nsub + 1)
        else:
            tau_s = pick_tau(evalcoeff, tau_rem, betak, conv_tol)
        coeff, ls = evalcoeff(tau_s)
        err += tail_sup(evalcoeff, tau_s, betak) * cur_norm / beta0
        out = scale(coeff[0] * cur_norm, basis[0])
        for c, p in zip(coeff[1:], basis[1:]):
            out = axpy(c * cur_norm, p, out)
        cur = out
        logscale += ls
        tau_rem -= tau_s
        if substeps == 1 or abs(tau_rem) <= 1e-14 * max(abs(tau), 1.0):
            break
        if nsub >= 10000:
            raise RuntimeError("expv: substep search failed to advance")
    sc = np.exp(logscale)
    if np.isfinite(abs(sc)) and sc != 0:
        cn = norm(cur)
        if not (np.isfinite(cn) and cn > 0 and abs(sc) * cn > EXPV_MAX_NORM):
            cur = scale(sc, cur)
    return cur, ExpvReport(total, nsub, err)


def hess_coeffs(Hm: np.ndarray, tau: complex):
    """_hess_coeffs (root_methods.jl:136-141): exp(tau Hm) e_1 for an upper-Hessenberg Hm, with the
    spectral shift of the Hermitian path; returns (coefficients, log of the factored scale)."""
    if Hm.size == 0:
        return np.zeros(0, dtype=complex), 0j
    sigma = shift(np.linalg.eigvals(Hm), tau)
    prop = sla.expm(tau * (Hm - sigma * np.eye(Hm.shape[0])))
    return prop[:, 0], tau * sigma


def arnoldi_expv(apply, tau: complex, vector, *, inner, norm, axpy, scale,
                 tol: float = 1e-13, maxiter: int = 60, conv_tol: float = 0.0, substeps: int = 1,
                 root_method: str = "exp", rational_shift: bool = True):
    """exp(tau A) vector for a non-Hermitian ``apply`` (arnoldi_expv, krylov.jl:343-442): the
    upper-Hessenberg Arnoldi factorisation with every vector orthogonalised twice (CGS2,
    krylov.jl:321-332), the same tail estimate, substep search and scale bookkeeping as
    ``lanczos_expv``; a rational ``root_method`` takes the whole step in one stage
    (krylov.jl:365-370).  Returns (result, ExpvReport)."""
    check_root_method(root_method)
    if root_method != "exp":
        substeps = 1

    def hcoeffs(Hm, t):
        if root_method == "exp":
            return hess_coeffs(Hm, t)
        return hess_rational(Hm, t, root_method, rational_shift)
    beta0 = norm(vector)
    if beta0 == 0:
        return vector, ExpvReport(0, 0, 0.0)
    cur = vector
    tau_rem = complex(tau)
    logscale = 0j
    total = nsub = 0
    err = 0.0
    while True:
        nsub += 1
        cur_norm = norm(cur)
        if cur_norm == 0:
            break
        basis = [scale(1.0 / cur_norm, cur)]
        H = np.zeros((maxiter + 1, maxiter), dtype=complex)
        dim = 0
        betak = 0.0
        for k in range(1, maxiter + 1):
            res = apply(basis[k - 1])
            total += 1
            for _p in range(2):
                coeffs = [inner(p, res) for p in basis]
                for row, c in enumerate(coeffs):
                    H[row, k - 1] += c
                for c, p in zip(coeffs, basis):
                    res = axpy(-c, p, res)
            dim = k
            rn = norm(res)
            betak = rn
            if rn < tol:
                break
            if conv_tol > 0:
                Hk = H[:dim, :dim].copy()
                tt = tail_sup(lambda t: hcoeffs(Hk, t), tau_rem, rn, above=conv_tol)
                if tt < conv_tol:
                    break
            H[k, k - 1] = rn
            if k == maxiter:
                break
            basis.append(scale(1.0 / rn, res))
        Hm = H[:dim, :dim].copy()

        def evalcoeff(t, Hm=Hm):
            return hcoeffs(Hm, complex(t))

        if substeps == 1:
            tau_s = tau_rem
        elif substeps > 1:
            tau_s = tau_rem / (substeps - nsub + 1)
        else:
            tau_s = pick_tau(evalcoeff, tau_rem, betak, conv_tol)
        coeff, ls = evalcoeff(tau_s)
        err += tail_sup(evalcoeff, tau_s, betak) * cur_norm / beta0
        out = scale(coeff[0] * cur_norm, basis[0])
        for c, p in zip(coeff[1:], basis[1:dim]):
            out = axpy(c * cur_norm, p, out)
        cur = out
        logscale += ls
        tau_rem -= tau_s
        if substeps == 1 or abs(tau_rem) <= 1e-14 * max(abs(tau), 1.0):
            break
        if nsub >= 10000:
            raise RuntimeError("expv: substep search failed to advance")
    sc = np.exp(logscale)
    if np.isfinite(abs(sc)) and sc != 0:
        cn = norm(cur)
        if not (np.isfinite(cn) and cn > 0 and abs(sc) * cn > EXPV_MAX_NORM):
            cur = scale(sc, cur)
    return cur, ExpvReport(total, nsub, err)
