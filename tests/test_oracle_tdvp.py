"""Pins of the oracle's two-site TDVP (CPU): the fused pair MPO against the dense two-site
operator, the two-site split against its input, the step against the exact evolution (second
order in the step, energy and norm conserved in real time without truncation), the energy-shift
invariance, and the budgeted cut."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import bug as obug
from oracle import linalg as la
from oracle import models as om
from oracle import tdvp
from tests.helpers import random_legs


@pytest.mark.parametrize("delta", [0.0, 0.5, 1.0])
def test_fused_mpo_is_the_dense_pair_operator(delta):
    W = om.xxz_mpo(2, delta=delta)
    W12 = tdvp.fused_mpo(W[0], W[1], 2)
    assert np.allclose(om.mpo_dense([W12], 4), om.mpo_dense(W, 2), atol=1e-14)
    assert np.allclose(om.mpo_dense(W, 2), om.dense_hamiltonian(2, delta=delta), atol=1e-14)


def test_two_site_split_reconstructs_and_is_isometric():
    rng = np.random.default_rng(3)
    sites = om.random_state(random_legs(rng, 6, 5), rng)
    Th = tdvp.two_site(sites[2], sites[3])
    A, S, B, kept, trunc = tdvp.svd_two_site(Th, [1, -1], [1, -1])
    back = tdvp.two_site(A, la.diag_times(S, B, "left"))
    diff = sum(np.linalg.norm(back.blocks[k] - Th.blocks[k]) ** 2 for k in Th.blocks)
    assert np.sqrt(diff) <= 1e-12 * Th.norm()
    ov = la.overlap_left(A, A)
    for (k, qq), b in ov.blocks.items():
        assert np.allclose(b, np.eye(b.shape[0]), atol=1e-12)
    ov = la.overlap_right(B, B)
    for (k, qq), b in ov.blocks.items():
        assert np.allclose(b, np.eye(b.shape[0]), atol=1e-12)
    assert not trunc or max(e[0] for e in trunc) <= 1e-14 * max(e[0] for e in kept)


def _run(L, delta, dt, steps, **kw):
    mpo = om.xxz_mpo(L, delta=delta)
    psi = obug.MPS(om.neel_state(L))
    for _ in range(steps):
        info = tdvp.tdvp2_step(psi, mpo, -1j * dt, **kw)
    return psi, info


def test_tdvp2_follows_the_exact_evolution_to_second_order():
    L, delta, T = 8, 0.5, 0.4
    H = om.dense_hamiltonian(L, delta=delta)
    v0 = om.mps_dense(om.neel_state(L))
    ex = sla.expm(-1j * T * H) @ v0
    errs = []
    for n in (8, 16):
        psi, info = _run(L, delta, T / n, n, maxdim=64)
        v = om.mps_dense(psi.sites)
        errs.append(np.linalg.norm(v - ex))
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert abs(np.vdot(v, H @ v).real - np.vdot(v0, H @ v0).real) < 1e-12
        assert info.discarded < 1e-12 and info.max_bond <= 16
    assert errs[1] < errs[0] / 3.5 and errs[0] < 1e-4


def test_energy_shift_changes_only_the_scale():
    """exp(-tau (H - E)) = exp(tau E) exp(-tau H) for every local update, and the cuts are
    relative: the normalised states agree with and without the shift."""
    L = 8
    mpo = om.xxz_mpo(L, delta=1.0)
    out = []
    for es in (0.0, -3.1):
        psi = obug.MPS(om.neel_state(L))
        for _ in range(3):
            tdvp.tdvp2_step(psi, mpo, -0.1, eshift=es, maxdim=6)
            psi.sites[psi.center] = psi.sites[psi.center].scale(1.0 / psi.sites[psi.center].norm())
        out.append(om.mps_dense(psi.sites))
    assert np.linalg.norm(out[0] - out[1]) < 1e-10


def test_budgeted_cut_keeps_the_discarded_norm_within_its_share():
    L, tol = 10, 1e-3
    psi, info = _run(L, 0.5, 0.1, 6, maxdim=64, weight_tol=tol)
    assert 0 < info.cumulative <= np.sqrt(2) * tol * (1 + 1e-9)
    psi2, info2 = _run(L, 0.5, 0.1, 6, maxdim=64, trunc_thresh=1e-12)
    assert max(psi.bond_dims()) <= max(psi2.bond_dims())
    with pytest.raises(ValueError):
        tdvp.tdvp2_step(psi, om.xxz_mpo(L), -0.1j, weight_tol=-1.0)


def test_backward_growth_is_nonnegative_in_imaginary_time_and_zero_in_real_time():
    L = 8
    mpo = om.xxz_mpo(L, delta=1.0)
    psi = obug.MPS(om.neel_state(L))
    info = tdvp.tdvp2_step(psi, mpo, -0.2, maxdim=16, backward_report=True)
    assert info.backward_growth >= -1e-12
    psi = obug.MPS(om.neel_state(L))
    tdvp.tdvp2_step(psi, mpo, -0.2j, maxdim=16)
    info = tdvp.tdvp2_step(psi, mpo, -0.2j, maxdim=16, backward_report=True)
    assert abs(info.backward_growth) < 1e-10
