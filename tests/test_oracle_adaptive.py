"""The oracle's holdout-certified adaptive DD controller (CPU): its pivoted QR against LAPACK's
geqp3, the two-pass selection, and whole steps in the mode the reference's frozen runs were made
with (test/BUG/TimeEvolution/test_fixtures_b155550.jl:35-38)."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import bug as obug
from oracle import dd_adaptive as da
from oracle import models as om
from oracle.blocks import Fam
from tests.helpers import free_fermion_sz

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "b155550_xx_u1.json")
FROZEN = dict(mode="adaptive", fixed_width=0, krylov_mode="arnoldi", work_ceiling_rule="maxdim")


@pytest.mark.parametrize("m,n,rank", [(30, 8, 8), (40, 12, 5), (9, 14, 9), (50, 6, 1)])
def test_pivoted_qr_matches_geqp3(m, n, rank):
    rng = np.random.default_rng(m * n)
    A = (rng.standard_normal((m, rank)) + 1j * rng.standard_normal((m, rank))) @ \
        (rng.standard_normal((rank, n)) + 1j * rng.standard_normal((rank, n)))
    A *= np.exp(np.linspace(0, -6, n))[None, :]            # graded column norms
    Q, r, perm = da.pivoted_qr(A)
    Qs, Rs, Ps = sla.qr(A, mode="economic", pivoting=True)
    k = min(m, n)
    assert np.array_equal(perm[:rank], Ps[:rank])           # the pivots that matter agree
    assert np.allclose(r[:rank], np.abs(np.diag(Rs))[:rank], rtol=1e-10)
    assert np.abs(Q.conj().T @ Q - np.eye(k)).max() < 1e-12
    assert np.all(r >= 0)
    # Q R reproduces the pivoted columns (R = Q^H A[:, perm] is upper triangular)
    R = Q.conj().T @ A[:, perm]
    assert np.abs(np.tril(R, -1)).max() < 1e-10 * np.abs(A).max()
    assert np.allclose(np.abs(np.diag(R)), r, atol=1e-10 * np.abs(A).max())


def test_dd_select_spans_the_projected_sketch_outside_the_frame():
    rng = np.random.default_rng(3)
    row, shifts = {-1: 3, 1: 4}, [1, -1]
    frame = Fam(row, {0: 2, 2: 2, -2: 1}, shifts)
    M = {}
    U = Fam.zeros(row, {0: 2, 2: 2, -2: 1}, shifts)
    for key, b in U.blocks.items():
        U.blocks[key] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    from oracle import linalg as la

    frame, _ = la.qr_left(U)
    sk = Fam.zeros(row, {0: 5, 2: 3, -2: 2}, shifts)
    for key, b in sk.blocks.items():
        sk.blocks[key] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    B = da.dd_select(frame, sk, "left")
    assert B is not None
    # orthonormal, orthogonal to the frame, and it captures the projected sketch entirely
    G = la.overlap_left(B, B)
    for (s, q), blk in G.blocks.items():
        assert np.abs(blk - np.eye(blk.shape[0])).max() < 1e-12
    assert la.overlap_left(frame, B).norm() < 1e-12
    proj = la.perp_left(frame, sk)
    rest = la.perp_left(B, proj)
    assert rest.norm() < 1e-12 * proj.norm()


def test_adaptive_steps_replay_frozen_order1_run():
    """xx_u1_bug_order1 in its own mode (:adaptive, :arnoldi, :maxdim): S^z within 1e-5 of the
    frozen trajectory and within twice its free-fermion error + 1e-6, statuses as frozen."""
    fx = json.load(open(GOLDEN))["xx_u1_bug_order1"]
    L, dt = fx["L"], fx["dt"]
    psi = obug.MPS(om.neel_state(L))
    mpo = om.xxz_mpo(L, delta=0.0)
    for n, row in enumerate(fx["steps"][:6], start=1):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt, obug.StepOptions(**FROZEN))
        sz = obug.sz_profile(psi)
        assert np.abs(sz - np.array(row["sz"])).max() <= 1e-5, n
        ff = free_fermion_sz(L, n * dt, range(0, L, 2))
        assert np.abs(sz - ff).max() <= 2 * row["err_exact"] + 1e-6, n
        assert info.ledger.status == row["status"], n
        assert info.ledger.n_holdout > 0


def test_adaptive_steps_replay_frozen_order2_run():
    fx = json.load(open(GOLDEN))["xx_u1_bug_order2"]
    L, dt = fx["L"], fx["dt"]
    psi = obug.MPS(om.neel_state(L))
    mpo = om.xxz_mpo(L, delta=0.0)
    for n, row in enumerate(fx["steps"][:4], start=1):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt, obug.StepOptions(order=2, **FROZEN))
        sz = obug.sz_profile(psi)
        assert np.abs(sz - np.array(row["sz"])).max() <= 1e-5, n
        ff = free_fermion_sz(L, n * dt, range(0, L, 2))
        assert np.abs(sz - ff).max() <= 2 * row["err_exact"] + 1e-6, n
        assert info.ledger.status == row["status"], n
