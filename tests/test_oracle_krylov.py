"""Pin the oracle's small dense Krylov pieces (CPU only)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import krylov as ok


def _arnoldi(Hm_gen, seed, depth):
    n = len(seed)
    V = np.zeros((n, depth + 1), dtype=complex)
    H = np.zeros((depth + 1, depth + 1), dtype=complex)
    V[:, 0] = seed / np.linalg.norm(seed)
    for k in range(depth):
        w = Hm_gen @ V[:, k]
        for _ in range(2):
            c = V[:, :k + 1].conj().T @ w
            H[:k + 1, k] += c
            w = w - V[:, :k + 1] @ c
        H[k + 1, k] = np.linalg.norm(w)
        V[:, k + 1] = w / H[k + 1, k]
    return V, H


def test_residual_integral_does_not_cancel_on_2x2():
    # test/BUG/Core/LocalSolvers/test_krylov_weights.jl:42-59: weight 1.0 to 1e-10
    G = np.array([[2 * np.pi, 1], [1, 0]], dtype=complex)
    V, H = _arnoldi(G, np.array([1, 0], dtype=complex), 1)
    w = ok.kry_weight(H, 1, -1j)
    assert abs(w - 1.0) <= 1e-10
    exact = sla.expm(-1j * G) @ np.array([1, 0])
    err = np.linalg.norm(exact - V[:, :1] @ sla.expm(-1j * H[:1, :1])[:, 0])
    assert err > 0.1 and w >= err


def test_weight_bounds_true_error_on_random_hermitian():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(3):
        for tau in (-0.1j, -1j, -3j):
            raw = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
            G = (raw + raw.conj().T) / 2
            seed = rng.standard_normal(30) + 1j * rng.standard_normal(30)
            seed /= np.linalg.norm(seed)
            V, H = _arnoldi(G, seed, 16)
            for step in range(1, 16):
                exact = sla.expm(tau * G) @ seed
                red = sla.expm(tau * H[:step, :step])[:, 0]
                err = np.linalg.norm(exact - V[:, :step] @ red)
                if err < 1e-13:
                    break
                checked += 1
                assert err <= (1 + 1e-3) * ok.kry_weight(H, step, tau) + 1e-14
    assert checked > 30


def test_tail_sup_guards_the_endpoint():
    vals, vecs = ok.tridiag_eigen([0.0, 0.0], [math.pi])
    ev = lambda t: ok.coeffs_from_eigen(vals, vecs, complex(t))
    assert ok._tail(ev(-1j)[0], 1.0) < 1e-12
    assert ok.tail_sup(ev, -1j, 1.0) >= 0.9
    for step in (-0.5, -0.5 - 1j):
        assert ok.tail_sup(ev, step, 1.0) == ok._tail(ev(step)[0], 1.0)


def test_tail_rank_rules():
    # test/BUG/Randomised/test_expansion.jl:17-58
    lst = [(s, 0, 0) for s in (10.0, 9.0, 3.0, 0.1)]
    assert ok.tail_rank(lst, 1.0, 1.0, 4)[0] == 3
    assert ok.tail_rank(lst, 1.0, 4.0, 4)[0] == 2
    assert ok.tail_rank(lst, 1.0, 9.6, 4)[0] == 1
    assert ok.tail_rank(lst, 1.0, 0.05, 4)[0] == 4
    lst = [(s, 0, 0) for s in (1.0, 1e-2, 1e-4, 1e-6)]
    k, tail = ok.tail_rank(lst, 1.0, 2e-4, 4)
    assert k == 2 and tail <= 2e-4
    assert ok.tail_rank(lst, 1.0, 5e-5, 4)[0] == 3
    lst = [(s, 0, 0) for s in (1.0, 0.5, 0.2, 0.05, 0.01, 1e-3)]
    ks = [ok.tail_rank(lst, 1.0, a, 6)[0] for a in (1e-6, 1e-3, 1e-2, 0.1, 0.5, 2.0)]
    assert ks == sorted(ks, reverse=True) and ks[0] == 6 and ks[-1] == 0


@pytest.mark.parametrize("tau", [-0.3j, -2.0j, -0.5])
def test_lanczos_expv_matches_expm(tau):
    rng = np.random.default_rng(5)
    raw = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    A = (raw + raw.conj().T) / 4
    v = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    out, rep = ok.lanczos_expv(lambda x: A @ x, tau, v, inner=np.vdot, norm=np.linalg.norm,
                               axpy=lambda a, x, y: y + a * x, scale=lambda a, x: a * x,
                               tol=1e-15, maxiter=30, reorth=True, conv_tol=1e-10, substeps=0)
    ref = sla.expm(tau * A) @ v
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-9
    assert rep.err <= 1e-10 * (1 + 1e-8) * rep.substeps


def test_tridiag_eigen_refuses_nonfinite():
    with pytest.raises(ValueError):
        ok.tridiag_eigen([0.0, float("nan")], [1.0])
