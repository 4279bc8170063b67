"""Pin the oracle's whole CBE-BUG step against closed forms, ED and the reference's frozen runs."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import bug as obug
from oracle import models as om

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "b155550_xx_u1.json")


def free_fermion_sz(L, t, occupied, J=1.0):
    """xx_free_fermion_sz (test/common/free_fermion.jl:37-40)."""
    h = np.diag(np.full(L - 1, J / 2), 1) + np.diag(np.full(L - 1, J / 2), -1)
    U = sla.expm(-1j * t * h)
    return np.array([sum(abs(U[j, k]) ** 2 for k in occupied) - 0.5 for j in range(L)])


def test_replays_frozen_order1_run():
    """xx_u1_bug_order1 (L=10 Neel, dt 0.05): S^z to 1e-5 of the frozen run, error vs free
    fermions <= 2 err + 1e-6, every status as frozen (test_fixtures_b155550.jl:40-66; the frozen
    run used the :maxdim ceiling and the Arnoldi recurrence)."""
    fx = json.load(open(GOLDEN))["xx_u1_bug_order1"]
    L, dt = fx["L"], fx["dt"]
    psi = obug.MPS(om.neel_state(L))
    mpo = om.xxz_mpo(L, delta=0.0)
    opts = obug.StepOptions(work_ceiling_rule="maxdim", krylov_mode="arnoldi")
    occ = list(range(0, L, 2))
    for n, row in enumerate(fx["steps"], start=1):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt, opts)
        sz = obug.sz_profile(psi)
        assert np.abs(sz - np.array(row["sz"])).max() <= 1e-5, n
        assert np.abs(sz - free_fermion_sz(L, n * dt, occ)).max() <= 2 * row["err_exact"] + 1e-6, n
        assert info.ledger.status == row["status"], n


def test_default_step_tracks_free_fermions_domain_wall():
    """The default step (rank_tail, Lanczos, per-bond 1.2 d D ceiling) on the L = 20 domain wall."""
    L, dt = 20, 0.05
    psi = obug.MPS(om.domain_wall_state(L))
    mpo = om.xxz_mpo(L, delta=0.0)
    for n in range(1, 6):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt)
        err = np.abs(obug.sz_profile(psi) - free_fermion_sz(L, n * dt, range(L // 2))).max()
        assert err < 1e-4, (n, err)
        assert abs(psi.sites[psi.center].norm() - 1.0) < 1e-8


def test_xxz_step_local_error_is_second_order_against_ed():
    """One order-1 step from the Neel state against exp(-i dt H) on 2^L states: the step is exact
    in the energy and its local error scales as dt^2 (the order-1 BUG integrator)."""
    L, delta = 8, 0.5
    H = om.dense_hamiltonian(L, delta=delta)
    errs = []
    for dt in (0.05, 0.01):
        psi = obug.MPS(om.neel_state(L))
        v0 = om.mps_dense(psi.sites)
        info = obug.cbe_bug_step(psi, om.xxz_mpo(L, delta=delta), -1j * dt,
                                 obug.StepOptions(work_ceiling_rule="maxdim"))
        got = om.mps_dense(psi.sites)
        exact = sla.expm(-1j * dt * H) @ v0
        errs.append(np.linalg.norm(got - exact))
        assert abs(np.vdot(got, H @ got) - np.vdot(v0, H @ v0)) < 1e-10
        assert info.ledger.status == "met"
    assert errs[0] < 5e-3
    assert 20 < errs[0] / errs[1] < 30            # (0.05 / 0.01)^2 = 25


def test_replays_frozen_order2_run():
    """xx_u1_bug_order2 (L=10 Neel, dt 0.05, order 2): S^z to 1e-5 of the frozen run, error vs
    free fermions <= 2 err + 1e-6, every status as frozen (test_fixtures_b155550.jl:40-66)."""
    fx = json.load(open(GOLDEN))["xx_u1_bug_order2"]
    L, dt = fx["L"], fx["dt"]
    psi = obug.MPS(om.neel_state(L))
    mpo = om.xxz_mpo(L, delta=0.0)
    opts = obug.StepOptions(order=2, work_ceiling_rule="maxdim", krylov_mode="arnoldi")
    occ = list(range(0, L, 2))
    for n, row in enumerate(fx["steps"], start=1):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt, opts)
        sz = obug.sz_profile(psi)
        assert np.abs(sz - np.array(row["sz"])).max() <= 1e-5, n
        assert np.abs(sz - free_fermion_sz(L, n * dt, occ)).max() <= 2 * row["err_exact"] + 1e-6, n
        assert info.ledger.status == row["status"], n
        assert len(info.ledger.leak) == 2 and info.ledger.leak[1] >= 0.0


def test_replays_frozen_order2_run_without_symmetry():
    """xx_none_bug_order2 (L=8 Neel, no symmetry, 10 steps): every norm within 1e-8 of frozen,
    the final 256 amplitudes within 1e-5 of the frozen state and the error against the dense
    propagator within 2 err_dense + 1e-6 (test_fixtures_b155550.jl:68-98)."""
    fx = json.load(open(GOLDEN))["xx_none_bug_order2"]
    L, dt = fx["L"], fx["dt"]
    psi = obug.MPS(om.neel_state(L, "none"))
    v0 = om.mps_dense(psi.sites)
    mpo = om.xxz_mpo(L, delta=0.0, symmetry="none")
    opts = obug.StepOptions(order=2, work_ceiling_rule="maxdim", krylov_mode="arnoldi")
    for n, row in enumerate(fx["steps"], start=1):
        info = obug.cbe_bug_step(psi, mpo, -1j * dt, opts)
        assert abs(np.linalg.norm(om.mps_dense(psi.sites)) - row["norm"]) <= 1e-8, n
        assert info.ledger.status == row["status"], n
    v = om.mps_dense(psi.sites)
    frozen = np.array(fx["final_state"]["re"]) + 1j * np.array(fx["final_state"]["im"])
    assert np.linalg.norm(v - frozen) <= 1e-5
    exact = sla.expm(-1j * len(fx["steps"]) * dt * om.dense_hamiltonian(L, delta=0.0)) @ v0
    assert np.linalg.norm(v - exact) / np.linalg.norm(exact) <= 2 * fx["err_dense"] + 1e-6


def test_order2_local_error_is_third_order_against_ed():
    """One order-2 step (one enriching frontier sweep of F_1 = tau B S_0) from the Neel state of the
    XXZ chain against exp(-i dt H): the local error falls 8x per halving of dt (O(dt^3)), where
    the order-1 step falls 4x, and the energy is conserved."""
    L, delta = 8, 0.5
    H = om.dense_hamiltonian(L, delta=delta)
    errs = []
    for dt in (0.1, 0.05, 0.025):
        psi = obug.MPS(om.neel_state(L))
        v0 = om.mps_dense(psi.sites)
        info = obug.cbe_bug_step(psi, om.xxz_mpo(L, delta=delta), -1j * dt,
                                 obug.StepOptions(order=2, work_ceiling_rule="maxdim"))
        got = om.mps_dense(psi.sites)
        errs.append(np.linalg.norm(got - sla.expm(-1j * dt * H) @ v0))
        assert abs(np.vdot(got, H @ got) - np.vdot(v0, H @ v0)) < 1e-10
        assert info.ledger.status == "met"
    assert errs[0] < 2e-4
    assert 7.5 < errs[0] / errs[1] < 8.5 and 7.5 < errs[1] / errs[2] < 8.5


def test_invalid_options_are_refused_before_touching_the_state():
    psi = obug.MPS(om.neel_state(6))
    before = [s.copy() for s in psi.sites]
    with pytest.raises(ValueError):
        obug.cbe_bug_step(psi, om.xxz_mpo(6), -0.1j, obug.StepOptions(work_ceiling_rule="nope"))
    with pytest.raises(ValueError):
        obug.cbe_bug_step(psi, om.xxz_mpo(5), -0.1j)
    for a, b in zip(before, psi.sites):
        assert a.blocks.keys() == b.blocks.keys()


def test_strict_and_frame_root_ceilings_closed_forms():
    """bond_work_ceilings_strict and bond_frame_root_ceilings (basis_budget.jl:497-549) against
    hand-worked values: ceil(f D_b) clipped to [1, maxdim], no local-dimension factor, the root
    bond on its own factor; 1.1 * 10 must give 11, not 12 (the product is rounded to 9 digits
    before the ceiling)."""
    assert obug.bond_work_ceilings_strict([1, 3, 10, 100], 64, 1.5) == [2, 5, 15, 64]
    assert obug.bond_work_ceilings_strict([10, 0], 64, 1.1) == [11, 1]
    assert obug.bond_frame_root_ceilings([4, 8, 8, 4], 100, 1, 2.0, 1.2) == [8, 10, 16, 8]
    assert obug.bond_frame_root_ceilings([4, 8, 8, 4], 12, 2, 2.0, 1.2) == [8, 12, 10, 8]
    with pytest.raises(ValueError):
        obug.bond_work_ceilings_strict([4], 64, 1.0)
    with pytest.raises(ValueError):
        obug.bond_frame_root_ceilings([4, 4], 64, 0, 1.5, 1.6)
    with pytest.raises(IndexError):
        obug.bond_frame_root_ceilings([4, 4], 64, 2)
    # Rule resolution (bug_step_stages.jl:63-85, 116-129): strict needs an explicit factor above
    # one; frame_root takes 2 as its frame factor when none is given, and refuses a root factor
    # above the frame factor.
    with pytest.raises(ValueError):
        obug.work_ceilings("bond_strict", [2, 2], 2, 64, None, 0, 1.2)
    assert obug.work_ceilings("frame_root", [2, 3], 2, 64, None, 1, 1.5) == [4, 5]
    with pytest.raises(ValueError):
        obug.work_ceilings("frame_root", [2, 3], 2, 64, None, 1, 2.5)
    with pytest.raises(ValueError):
        obug.work_ceilings("maxdim", [2, 3], 2, 64, 1.5, 1, 1.2)
    assert obug.work_ceilings("maxdim", [2, 3], 2, 64, None, 1, 1.2) == 128


def test_product_ceiling_helpers_equal_the_oracle():
    from bondupdategalerkin_amd import bug_step as B

    rng = np.random.default_rng(7)
    for _ in range(200):
        ranks = [int(x) for x in rng.integers(0, 300, size=int(rng.integers(1, 12)))]
        maxdim = int(rng.integers(1, 400))
        f = float(rng.choice([1.05, 1.1, 1.2, 1.5, 2.0, 2.5]))
        rf = min(f, float(rng.choice([1.01, 1.1, 1.2])))
        root = int(rng.integers(0, len(ranks)))
        assert B.bond_work_ceilings_strict(ranks, maxdim, f) == \
            obug.bond_work_ceilings_strict(ranks, maxdim, f)
        assert B.bond_frame_root_ceilings(ranks, maxdim, root, f, rf) == \
            obug.bond_frame_root_ceilings(ranks, maxdim, root, f, rf)


def test_ceiling_rules_bound_the_step():
    """A step under :bond_strict or :frame_root keeps every bond within the ceiling computed
    from the bonds before it."""
    for rule, f, rf in [("bond_strict", 1.5, 1.2), ("frame_root", None, 1.2),
                        ("frame_root", 1.8, 1.5)]:
        psi = obug.MPS(om.domain_wall_state(10, "U1"))
        mpo = om.xxz_mpo(10, delta=0.5)
        for _ in range(3):
            cap = obug.work_ceilings(rule, psi.bond_dims(), 2, 128, f, 4, rf)
            obug.cbe_bug_step(psi, mpo, -0.05j, obug.StepOptions(
                work_ceiling_rule=rule, bond_ceiling_factor=f, root_ceiling_factor=rf, order=2))
            assert all(b <= c for b, c in zip(psi.bond_dims(), cap))


def _left_dense(f):
    d = f.dense()
    return d.reshape(d.shape[0] * d.shape[1], d.shape[2])


def test_ranked_krylov_selection_keeps_the_top_directions():
    """_append_ranked (krylov_enrichment.jl:185-196) against a dense SVD: the directions added
    to the frame span the top-k left singular subspace of the weighted candidates' direct sum,
    k the fewest whose discarded tail fits the allowance."""
    from oracle import linalg as la

    rng = np.random.default_rng(3)
    row, shifts = {-1: 2, 1: 3}, [1, -1]

    def rand(col):
        f = obug.Fam.zeros(row, col, shifts)
        for blk in f.blocks.values():
            blk[...] = rng.standard_normal(blk.shape) + 1j * rng.standard_normal(blk.shape)
        return f

    frame = la.svd_left(rand({0: 2, 2: 1})).U
    cands = [la.perp_left(frame, rand({0: 3, 2: 1, -2: 1})).scale(0.3),
             la.perp_left(frame, rand({0: 1, -2: 1})).scale(0.05)]
    M = np.hstack([_left_dense(c) for c in cands])
    U, s, _ = np.linalg.svd(M)
    tails = np.sqrt(np.cumsum((s ** 2)[::-1]))[::-1]          # tails[k] = ||s[k:]||
    for k_want in (1, 2, 4):
        allowance = float(tails[k_want]) * 1.0001
        grown, added, fits = obug.append_ranked(frame, cands, "left", allowance, 0)
        assert fits and added == k_want
        G, F = _left_dense(grown), _left_dense(frame)
        assert np.abs(G.conj().T @ G - np.eye(G.shape[1])).max() < 1e-12
        p_added = G @ G.conj().T - F @ F.conj().T
        p_svd = U[:, :k_want] @ U[:, :k_want].conj().T
        assert np.abs(p_added - p_svd).max() < 1e-10
    # the work ceiling stops the selection: two columns of room, the allowance not met
    grown, added, fits = obug.append_ranked(frame, cands, "left", float(tails[4]) * 1.0001,
                                            frame.col[0] + frame.col.get(2, 0) + 2)
    assert added == 2 and not fits
    assert obug.append_ranked(frame, [], "left", 1e-3, 0) == (frame, 0, True)


def test_ranked_allocation_share_and_validation():
    """krylov_rank is apportioned by the :linear rule like the other shares
    (basis_budget.jl:88-104): (krylov_rank_tol / tol_basis) * tol_basis / (2 (nb + 1) nf)."""
    a = obug.basis_allocation(1e-5, 9, -0.05j, 1.0, 2, None, 4e-6)
    assert a.krylov_rank == pytest.approx(4e-6 / (2 * 10 * 2), rel=1e-14)
    assert obug.basis_allocation(1e-5, 9, -0.05j, 1.0, 2).krylov_rank == 0.0
    with pytest.raises(ValueError):
        obug.basis_allocation(1e-5, 9, -0.05j, 1.0, 2, None, -1e-6)


def test_ranked_step_tracks_free_fermions():
    """XX domain wall L = 10 under ranked Krylov selection: <S^z_j>(t) within 1e-7 of free
    fermions after 6 order-2 steps of dt = 0.05 (8.5e-9 measured, the same as the sequential
    rule)."""
    L, dt = 10, 0.05
    psi = obug.MPS(om.domain_wall_state(L, "U1"))
    mpo = om.xxz_mpo(L, delta=0.0)
    for _ in range(6):
        obug.cbe_bug_step(psi, mpo, -1j * dt, obug.StepOptions(order=2, krylov_rank_tol=1e-5,
                                                               tol_basis=1e-6))
    occupied = [j for j in range(L // 2)]
    ref = free_fermion_sz(L, 6 * dt, occupied)
    assert np.abs(obug.sz_profile(psi) - ref).max() < 1e-7


@pytest.mark.parametrize("start,delta,order", [("neel", 0.5, 1), ("dw", 0.0, 1), ("neel", 1.0, 2)])
def test_bond_log_aggregates_to_the_step_ledger(start, delta, order):
    """The per-bond log (_log_bond!, basis_budget.jl:390-418; diagnostics.jl:328-335): one entry
    per half-sweep bond visit (the root bond from both sides), the step's maxima and sums are
    exactly the aggregate of the entries, r_old/r_final are the bonds before and after, and the
    per-bond discarded norms recombine into the cumulative one, sqrt(sum over the left subtree)
    + sqrt(sum over the right subtree)."""
    L = 10
    st = {"neel": om.neel_state, "dw": om.domain_wall_state}[start](L, "U1")
    psi = obug.MPS(st)
    mpo = om.xxz_mpo(L, delta=delta)
    for _ in range(3):
        bd0 = psi.bond_dims()
        info = obug.cbe_bug_step(psi, mpo, -0.05j, obug.StepOptions(order=order, maxdim=6))
        led, log = info.ledger, info.ledger.bond_log
        rl = L // 2 - 1
        assert [(e["bond"], e["side"]) for e in log] == \
            [(b, "left") for b in range(rl + 1)] + [(b, "right") for b in range(L - 2, rl - 1, -1)]
        for key, agg in (("n_probes", led.n_probes), ("n_selected", led.n_selected),
                         ("n_krylov", led.n_krylov), ("n_holdout", led.n_holdout),
                         ("n_targeted", led.n_targeted)):
            assert sum(e[key] for e in log) <= agg        # order > 1: closure sweeps add
            if order == 1:
                assert sum(e[key] for e in log) == agg, key
        if order == 1:
            assert max(e["sample"] for e in log) == led.max_sample
            assert max(e["krylov_c"] for e in log) == led.max_krylov_c
            bond_unmet = {u for e in log for u in e["unmet"]}
            assert bond_unmet == set(led.unmet) - {"trunc", "projection", "expv"}
            assert all(e["met"] == (not e["unmet"]) for e in log)
        bd1 = psi.bond_dims()
        assert [e["r_old"] for e in log] == [bd0[e["bond"]] for e in log]
        assert [e["r_final"] for e in log] == [bd1[e["bond"]] for e in log]
        w = {e["bond"]: e["discarded"] for e in log}
        left = math.sqrt(sum(w[b] ** 2 for b in range(rl + 1)))
        right = math.sqrt(sum(w[b] ** 2 for b in range(rl + 1, L - 1)))
        assert left + right == pytest.approx(led.discarded, rel=1e-12, abs=1e-15)
        widths = {(e["bond"], e["side"]): e["width"] for e in log}
        assert all(widths[(b, "left")] == info.expanded[b] for b in range(rl + 1))


def test_exact_probe_spans_the_discarded_space():
    """exact = true probes (dd_cbe.jl:141-160): the fusion basis with the frame projected out is
    the orthogonal projector onto the frame's complement, Om Om^dag = I - F F^dag on the fused
    local space (left frames) and Om^dag Om = I - F^dag F (right frames)."""
    from oracle import linalg as la

    rng = np.random.default_rng(5)
    row, shifts = {-1: 2, 1: 3}, [1, -1]
    A = obug.Fam.zeros(row, {0: 2, 2: 1}, shifts)
    for blk in A.blocks.values():
        blk[...] = rng.standard_normal(blk.shape) + 1j * rng.standard_normal(blk.shape)
    F = la.svd_left(A).U
    Om = obug.exact_probe(F, "left")
    O, Fd = _left_dense(Om), _left_dense(F)
    assert np.abs(O @ O.conj().T - (np.eye(O.shape[0]) - Fd @ Fd.conj().T)).max() < 1e-13
    Vr = la.svd_right(obug.Fam.zeros({0: 2, -2: 1}, row, [-1, 1]).axpy(
        1.0, _random_like(obug.Fam.zeros({0: 2, -2: 1}, row, [-1, 1]), rng))).Vh
    Omr = obug.exact_probe(Vr, "right")
    dr, dv = Omr.dense(), Vr.dense()
    Or = np.concatenate(list(dr), axis=1)          # (width, slices * link)
    Vd = np.concatenate(list(dv), axis=1)
    assert np.abs(Or.conj().T @ Or - (np.eye(Or.shape[1]) - Vd.conj().T @ Vd)).max() < 1e-13


def _random_like(f, rng):
    for blk in f.blocks.values():
        blk[...] = rng.standard_normal(blk.shape) + 1j * rng.standard_normal(blk.shape)
    return f


def test_exact_step_is_deterministic_and_tracks_free_fermions():
    """exact = true (bug_step.jl:117-119) draws no random numbers: serial and parallel half-sweeps
    give the identical state (test_parallel_half_sweeps.jl:6-9), any seed gives the same step,
    no probe is counted, and on the XX domain wall it gives the Gaussian-probe step's state (both
    span the discarded space the sketch needs) and its free-fermion error, 1.8e-5 after 6
    order-1 steps of dt = 0.05."""
    L, dt = 10, 0.05
    runs = []
    for par, seed in ((False, 1), (True, 2)):
        psi = obug.MPS(om.domain_wall_state(L, "U1"))
        mpo = om.xxz_mpo(L, delta=0.0)
        for _ in range(6):
            info = obug.cbe_bug_step(psi, mpo, -1j * dt, obug.StepOptions(
                exact=True, parallel=par, seed=seed, tol_basis=1e-6))
            assert info.ledger.n_probes == 0 and info.ledger.n_holdout == 0
        runs.append(om.mps_dense(psi.sites))
    assert np.array_equal(runs[0], runs[1])
    ref = free_fermion_sz(L, 6 * dt, list(range(L // 2)))
    assert np.abs(obug.sz_profile(psi) - ref).max() < 2e-5
    gauss = obug.MPS(om.domain_wall_state(L, "U1"))
    for _ in range(6):
        obug.cbe_bug_step(gauss, om.xxz_mpo(L, delta=0.0), -1j * dt,
                          obug.StepOptions(tol_basis=1e-6))
    assert np.abs(obug.sz_profile(psi) - obug.sz_profile(gauss)).max() < 1e-12


@pytest.mark.parametrize("kw", [dict(exact=True), dict(fold_omega=False)])
def test_unfolded_closure_steps_track_free_fermions(kw):
    """Order 2 with the unfolded frontier leak (fold_omega = false, the default under exact,
    bug_step_stages.jl:67; _bond_action, krylov_frontier_bonds.jl:80-106): the XX domain wall
    within 1e-7 of free fermions after 6 steps of dt = 0.05 (8.5e-9 measured, as folded); under
    exact the serial and parallel runs with different seeds are identical."""
    L, dt = 10, 0.05
    states = []
    for par, seed in ((False, 1), (True, 2)):
        psi = obug.MPS(om.domain_wall_state(L, "U1"))
        mpo = om.xxz_mpo(L, delta=0.0)
        for _ in range(6):
            info = obug.cbe_bug_step(psi, mpo, -1j * dt, obug.StepOptions(
                order=2, tol_basis=1e-6, parallel=par, seed=seed, **kw))
            assert np.isfinite(info.ledger.leak[1])
        ref = free_fermion_sz(L, 6 * dt, list(range(L // 2)))
        assert np.abs(obug.sz_profile(psi) - ref).max() < 1e-7
        states.append(om.mps_dense(psi.sites))
    if kw.get("exact"):
        assert np.array_equal(states[0], states[1])


def test_fusion_probe_is_unitary():
    """fusion_basis (dd_cbe.jl:143) in the probe layouts: a unitary change of basis of the fused
    legs on both sides, so a sketch with it is the two-site action itself up to that rotation
    (which is what makes the unfolded leak exact)."""
    row, shifts = {-1: 2, 1: 3}, [1, -1]
    F = obug.Fam.zeros(row, {0: 1, 2: 1, -2: 1}, shifts)
    O = _left_dense(obug.fusion_probe(F, "left"))
    assert O.shape[0] == O.shape[1] == 10
    assert np.abs(O @ O.conj().T - np.eye(10)).max() == 0.0
    V = obug.Fam.zeros({0: 1, -2: 1}, row, [-1, 1])
    Or = np.concatenate(list(obug.fusion_probe(V, "right").dense()), axis=1)
    assert Or.shape[0] == Or.shape[1]
    assert np.abs(Or.conj().T @ Or - np.eye(Or.shape[1])).max() == 0.0


def _keep_fixture():
    """test_step_keep_states.jl:11-25: Heisenberg L = 8, exact selection, order 2; the start
    state one step of -0.2i from Neel, the kept state three more steps of -0.3i."""
    mpo = om.xxz_mpo(8, delta=1.0)
    kw = dict(order=2, tol_basis=1e-8, tol_trunc=1e-12, maxdim=64, exact=True)
    start = obug.MPS(om.neel_state(8, "U1"))
    obug.cbe_bug_step(start, mpo, -0.2j, obug.StepOptions(**kw))
    kept = start.copy()
    for _ in range(3):
        obug.cbe_bug_step(kept, mpo, -0.3j, obug.StepOptions(**kw))
    return mpo, kw, start, kept


def test_kept_states_widen_the_frames_within_their_rules():
    """keep_states / keep_tols (half_sweeps.jl:117-207,280-296,384-395), the reference's own
    checks (test_step_keep_states.jl:27-57): a kept state at tol 0 widens some frame, admitted at
    a tolerance it widens no frame more than exactly, under :bond with a small maxdim no frame
    exceeds its ceiling; and at tol 0 the kept state lies in the step's trial space (its
    projection miss onto the frames is at rounding level)."""
    mpo, kw, start, kept = _keep_fixture()
    kw = dict(kw, truncate=False)

    def widths(**extra):
        info = obug.cbe_bug_step(start.copy(), mpo, -0.2j, obug.StepOptions(**dict(kw, **extra)))
        return np.array(info.expanded)

    plain = widths()
    exact = widths(keep_states=[kept])
    admitted = widths(keep_states=[kept], keep_tols=[1e-6])
    assert (exact > plain).any()
    assert (admitted <= exact).all()
    bonds = start.bond_dims()
    small = max(max(bonds), 2)
    ceil = obug.bond_work_ceilings(bonds, 2, small, obug.DEFAULT_BOND_CEILING_FACTOR)
    capped = widths(keep_states=[kept], keep_tols=[1e-12], maxdim=small, work_ceiling_rule="bond")
    assert (capped <= np.array(ceil)).all()
    with pytest.raises(ValueError):
        widths(keep_states=[kept], keep_tols=[1e-6, 1e-6])


def test_union_fresh_contains_the_candidate():
    """_union_fresh at tol 0: the grown frame is an isometry whose span holds the candidate to
    rounding, and a candidate already inside the frame adds nothing."""
    from oracle import linalg as la

    rng = np.random.default_rng(11)
    row, shifts = {-1: 2, 1: 3}, [1, -1]
    F = la.svd_left(_random_like(obug.Fam.zeros(row, {0: 2, 2: 1}, shifts), rng)).U
    C = _random_like(obug.Fam.zeros(row, {0: 1, -2: 1}, shifts), rng)
    G = obug.union_frame(F, C, "left")
    Gd, Cd = _left_dense(G), _left_dense(C)
    assert np.abs(Gd.conj().T @ Gd - np.eye(Gd.shape[1])).max() < 1e-13
    assert np.linalg.norm(Cd - Gd @ (Gd.conj().T @ Cd)) < 1e-13 * np.linalg.norm(Cd)
    assert obug.union_frame(G, C, "left") is G


def test_frames_only_returns_the_state_in_the_frames():
    """frames_only (bug_step.jl:114-116,381-382; test_step_keep_states.jl:59-76): no root solve,
    so the state comes back as the input in the frames (overlap ratio 1 to 1e-12), with the
    full step's frame widths; a truncation with it is refused."""
    mpo, kw, start, _ = _keep_fixture()
    kw = dict(kw, truncate=False)
    full = obug.cbe_bug_step(start.copy(), mpo, -0.2j, obug.StepOptions(**kw))
    psi = start.copy()
    info = obug.cbe_bug_step(psi, mpo, -0.2j, obug.StepOptions(**kw, frames_only=True))
    a, b = om.mps_dense(psi.sites), om.mps_dense(start.sites)
    assert 1 - abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b)) < 1e-12
    assert info.expanded == full.expanded
    with pytest.raises(ValueError):
        obug.cbe_bug_step(start.copy(), mpo, -0.2j, obug.StepOptions(**dict(kw, truncate=True),
                                                                     frames_only=True))
