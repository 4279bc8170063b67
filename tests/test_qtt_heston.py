"""The Heston inputs (quantics grid, difference chains, generator MPO, QTT states) and the oracle's
non-Hermitian step on them (CPU only).

Mirrors test/QTTS/Models/Heston/test_heston.jl:27-106: the generator against dense stencils, the
forward adjoint's mass conservation, the degenerate v = 0 rows, a manufactured affine solution at
second order, and BUG steps (hermitian = False) against the dense exponential.  The dense
generator itself is pinned on the frozen backward-pricing run.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from bondupdategalerkin_amd import qtt as Q
from oracle import bug as obug
from oracle import heston as oh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
HP = Q.HestonParams(0.03, 0.01, 2.0, 0.09, 0.4, -0.7)        # test/common/qtt_fixtures.jl:78
PT = (HP.r, HP.q, HP.kappa, HP.theta, HP.sigma, HP.rho)


def _grid(nx, ny):
    return Q.qtt_grid(nx, ny, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---- the oracle's dense generator, pinned ------------------------------------------------------

def test_dense_generator_reproduces_frozen_backward_pricing_run():
    fx = json.load(open(os.path.join(GOLDEN, "heston_backward.json")))
    n, p = fx["bits"], fx["params"]
    params = (p["r"], p["q"], p["kappa"], p["theta"], p["sigma"], p["rho"])
    xs, _, _, _ = oh.grid(n, n, fx["x"], fx["v"])
    softplus = lambda z: z if z > 30 else np.log1p(np.exp(z))              # noqa: E731
    u0 = np.kron(np.ones(2 ** n), [0.02 * softplus((np.exp(x) - 1) / 0.02) for x in xs])
    A = oh.heston_dense(n, n, params, fx["x"], fx["v"])
    exact = sla.expm(fx["h"] * fx["steps"] * A) @ u0
    frozen = np.array(fx["final_state"]["re"]) + 1j * np.array(fx["final_state"]["im"])
    # the frozen run's own integration error is ~1e-7; a wrong sign, coefficient, stencil row or
    # grid spacing moves the state by orders of magnitude more
    assert _relerr(frozen, exact) < 1e-6


@pytest.mark.parametrize("order", [1, 2])
def test_stencils_exact_on_quadratics(order):
    N, h = 16, 0.37
    x = 0.2 + h * np.arange(N)
    D = oh.stencil(N, h, order, "onesided")
    for f, df, d2f in [(np.ones_like(x), 0 * x, 0 * x), (x, 1 + 0 * x, 0 * x),
                       (x * x, 2 * x, 2 + 0 * x)]:
        want = df if order == 1 else d2f
        assert np.abs(D @ f - want).max() < 1e-10
    # the periodic central stencil on a Fourier mode: the textbook symbols
    k = 2 * np.pi * 3 / N
    mode = np.exp(1j * k * np.arange(N))
    Dp = oh.stencil(N, h, order, "periodic")
    sym = 1j * np.sin(k) / h if order == 1 else (2 * np.cos(k) - 2) / h ** 2
    assert np.abs(Dp @ mode - sym * mode).max() < 1e-10


def test_dense_generator_is_pure_discount_on_constants():
    A = oh.heston_dense(3, 4, PT)
    assert np.abs(A @ np.ones(A.shape[0]) + HP.r).max() < 1e-9


# ---- the package's chains and MPO against the dense restatement -------------------------------

@pytest.mark.parametrize("n", [2, 3, 5])
@pytest.mark.parametrize("bc", ["onesided", "periodic", "dirichlet0"])
@pytest.mark.parametrize("order", [1, 2])
def test_derivative_chain_matches_dense_stencil(n, bc, order):
    D = Q.op_dense(Q.qtt_derivative(n, 0.13, order, bc))
    R = oh.stencil(2 ** n, 0.13, order, bc)
    assert np.abs(D - R).max() <= 1e-11 * np.abs(R).max()


def test_shift_and_coordinate_chains():
    for periodic in (False, True):
        S = Q.op_dense(Q.op_shift(4, periodic))
        want = np.eye(16, k=1)
        if periodic:
            want[15, 0] = 1
        assert np.array_equal(S, want)
    C = Q.op_dense(Q.op_coordinate(5, -0.25, 0.1))
    assert np.allclose(C, np.diag(-0.25 + 0.1 * np.arange(32)), atol=1e-14)
    assert max(c.shape[3] for c in Q.op_coordinate(5, -0.25, 0.1)) == 2


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 4), (3, 5)])
def test_generator_and_forward_adjoint_match_dense(nx, ny):
    g = _grid(nx, ny)
    W = Q.heston_op(g, HP)
    ref = oh.heston_dense(nx, ny, PT)
    A = Q.op_dense(W)
    assert _relerr(A, ref) < 1e-11
    assert W[nx - 1].shape[3] <= 3                       # the x|v bond: d_xx, d_x, I
    fwd = Q.op_dense(Q.heston_op(g, HP, adjoint=True))
    assert _relerr(fwd, (ref + HP.r * np.eye(len(ref))).T) < 1e-11
    assert np.linalg.norm(np.ones(len(fwd)) @ fwd) < 1e-9 * np.linalg.norm(fwd)
    # at v = 0 (the first 2^nx grid rows) no diffusion or mixed term survives
    flat = Q.HestonParams(HP.r, HP.q, HP.kappa, HP.theta, 0.0, 0.0)
    deg = Q.op_dense(Q.heston_op(g, flat))
    assert _relerr(A[:2 ** nx], deg[:2 ** nx]) < 1e-11


def test_mpo_sites_carry_the_same_operator():
    g = _grid(3, 3)
    W = Q.heston_op(g, HP)
    mpo = Q.qtt_to_mpo(W)
    cores = [s.dense(2).transpose(0, 2, 3, 1) for s in mpo.sites]
    assert _relerr(oh.mpo_grid_matrix(cores), oh.heston_dense(3, 3, PT)) < 1e-11
    assert all(s.nent == np.count_nonzero(c) for s, c in zip(mpo.sites, W))


def test_heston_op_refuses_bad_grids():
    with pytest.raises(ValueError):
        Q.heston_op(Q.qtt_grid(3, 3), HP)
    with pytest.raises(ValueError):
        Q.heston_op(Q.qtt_grid(3, 3, y=(0.1, 0.5), periodic_x=False, periodic_y=False), HP)
    with pytest.raises(ValueError):
        Q.qtt_derivative(1, 0.1, 1, "onesided")
    with pytest.raises(ValueError):
        Q.qtt_grid(3, 3, order="interleaved")


def test_manufactured_affine_solution_at_second_order():
    # test_heston.jl:62-88
    a, b = 0.5, -1.0
    errs = []
    for n in (4, 5):
        g = _grid(n, n)
        xs, vs = Q.qtt_coords(g)
        field = Q.qtt_separable(g, lambda x: np.exp(a * x), lambda v: np.exp(b * v))
        applied = (Q.op_dense(Q.heston_op(g, HP)) @ Q.tt_dense(field)).reshape(2 ** n, 2 ** n).T
        exact = np.array([[oh.affine_rate(v, a, b, PT) * np.exp(a * x + b * v) for v in vs]
                          for x in xs])
        errs.append(np.abs(applied - exact).max() / np.abs(exact).max())
    assert errs[1] < 0.35 * errs[0] and errs[1] < 1e-2


# ---- states ------------------------------------------------------------------------------------

def test_tt_from_samples_and_separable_fields():
    rng = np.random.default_rng(2)
    vals = rng.standard_normal(64)
    assert np.allclose(Q.tt_dense(Q.tt_from_samples(vals)), vals, atol=1e-13)
    lin = Q.tt_from_samples(0.3 + 0.7 * np.arange(64))
    assert max(c.shape[2] for c in lin) == 2                # a linear function has rank 2
    g = _grid(4, 3)
    xs, vs = Q.qtt_coords(g)
    f = Q.qtt_separable(g, np.cos, lambda v: 1 + v * v)
    assert np.allclose(Q.qtt_values(f, g), np.outer(np.cos(xs), 1 + vs * vs), atol=1e-13)
    const_v = Q.qtt_separable(g, np.sin, None)
    assert np.allclose(Q.qtt_values(const_v, g), np.outer(np.sin(xs), np.ones(len(vs))))


# ---- the oracle's step with the Arnoldi root on the Heston generator ---------------------------

def _oracle_heston_run(tol_basis, order=1, n=4, steps=5, tau=0.02):
    g = _grid(n, n)
    init = Q.qtt_separable(g, lambda x: np.exp(0.5 * x), lambda v: np.exp(-2 * v) * (1 + v))
    mpo = oh.mpo_from_cores(Q.heston_op(g, HP))
    psi = obug.MPS(oh.fams_from_cores(init), center=len(init) - 1)
    obug.canonical(psi, 0)
    infos = []
    for _ in range(steps):
        infos.append(obug.cbe_bug_step(psi, mpo, tau, obug.StepOptions(
            tol_basis=tol_basis, tol_trunc=1e-12, maxdim=64, hermitian=False, order=order)))
    exact = sla.expm(steps * tau * oh.heston_dense(n, n, PT)) @ oh.tt_grid_vector(init)
    return psi, infos, exact


def test_oracle_step_matches_dense_exponential():
    # test_heston.jl:90-106
    errors = []
    for tb in (1e-3, 1e-8):
        psi, infos, exact = _oracle_heston_run(tb)
        errors.append(_relerr(oh.mps_grid_vector(psi.sites), exact))
        assert all(i.ledger.expv_iters > 0 and i.ledger.proj_err < 1e-12 for i in infos)
    assert errors[1] <= max(errors[0], 1e-10)
    assert errors[1] < 1e-5
    assert max(psi.bond_dims()) > 1                         # the x|v bond grows


def test_oracle_order2_step_on_heston():
    psi, infos, exact = _oracle_heston_run(1e-8, order=2, steps=3)
    assert _relerr(oh.mps_grid_vector(psi.sites), exact) < 1e-5
    assert all(i.ledger.expv_iters > 0 for i in infos)


def test_arnoldi_root_equals_lanczos_root_on_hermitian_chain():
    from oracle import models as om

    out = []
    for herm in (True, False):
        psi = obug.MPS(om.neel_state(8))
        obug.cbe_bug_step(psi, om.xxz_mpo(8, delta=0.5), -0.05j,
                          obug.StepOptions(hermitian=herm, krylov_mode="arnoldi"))
        out.append(om.mps_dense(psi.sites))
    assert np.linalg.norm(out[0] - out[1]) < 1e-10
