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


_VEC = dict(inner=np.vdot, norm=np.linalg.norm, axpy=lambda a, x, y: y + a * x,
            scale=lambda a, x: a * x)


def _sector_fixture():
    """A vector over four 'sectors' of unequal size and the sector label of each entry: a
    generator that multiplies each sector by its own number spans exactly four Krylov vectors."""
    rng = np.random.default_rng(7)
    labels = np.repeat(np.array([-2, -1, 1, 2]), [3, 5, 2, 4])
    x = rng.standard_normal(labels.size) + 1j * rng.standard_normal(labels.size)
    return x, labels


@pytest.mark.parametrize("tau", [-0.4 + 0j, -0.7j])
def test_arnoldi_matches_sector_closed_form(tau):
    # test/BUG/Core/LocalSolvers/test_krylov.jl:81-87 (atol 1e-11)
    x, lab = _sector_fixture()
    s = 0.5 * lab + 1.0
    y, rep = ok.arnoldi_expv(lambda v: s * v, tau, x, **_VEC)
    assert np.linalg.norm(y - np.exp(tau * s) * x) <= 1e-11
    assert rep.iterations == 4          # the true Krylov dimension (test_krylov.jl:166-173)


def test_arnoldi_terminates_on_non_hermitian_generator():
    # test/BUG/Core/LocalSolvers/test_krylov.jl:135-164: complex sector scales
    x, lab = _sector_fixture()
    s = (0.5 * lab + 1.0) + 0.3j * lab
    y, rep = ok.arnoldi_expv(lambda v: s * v, -0.35, x, maxiter=60, **_VEC)
    assert np.linalg.norm(y - np.exp(-0.35 * s) * x) <= 1e-12
    assert rep.iterations == 4


def test_arnoldi_matches_lanczos_on_hermitian_generator():
    # test/BUG/Core/LocalSolvers/test_krylov.jl:53-57, on a dense Hermitian generator
    rng = np.random.default_rng(3)
    raw = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
    A = (raw + raw.conj().T) / 4
    v = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    kw = dict(tol=1e-15, maxiter=30, conv_tol=1e-12, substeps=0)
    ya, _ = ok.arnoldi_expv(lambda z: A @ z, -0.2j, v, **kw, **_VEC)
    yl, _ = ok.lanczos_expv(lambda z: A @ z, -0.2j, v, reorth=True, **kw, **_VEC)
    assert np.linalg.norm(ya - yl) <= 1e-11


@pytest.mark.parametrize("tau", [0.3, 0.05 + 0.02j, -1.0j, 4.0])
def test_arnoldi_expv_matches_expm_with_substeps(tau):
    # a non-normal generator; tau = 4 forces the substep search and the spectral shift
    rng = np.random.default_rng(9)
    A = (rng.standard_normal((60, 60)) + 1j * rng.standard_normal((60, 60))) / np.sqrt(60)
    A += 10 * np.eye(60) if tau == 4.0 else 0
    v = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    y, rep = ok.arnoldi_expv(lambda z: A @ z, tau, v, tol=1e-15, maxiter=30, conv_tol=1e-10,
                             substeps=0, **_VEC)
    ref = sla.expm(tau * A) @ v
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
    assert rep.err <= 1e-10 * (1 + 1e-8) * rep.substeps


def test_arnoldi_zero_input_and_shift_uses_real_part():
    z = np.zeros(5, dtype=complex)
    y, rep = ok.arnoldi_expv(lambda v: 2 * v, -0.1j, z, **_VEC)
    assert np.linalg.norm(y) == 0 and rep.iterations == 0
    # eigenvalues 40 +- 3i under tau = 1: the shift is the real part of the top one
    assert ok.shift(np.array([40 + 3j, 40 - 3j, 1.0]), 1.0) == 40.0
    assert ok.shift(np.array([20 + 3j]), 1.0) == 0.0
