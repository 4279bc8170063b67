"""Pins of the oracle's imaginary-time driver and the step instrumentation it uses (CPU): the
energy and norm against the dense state, the outside-Galerkin action against the dense projector,
the root report's energies, the measured closure against the single pass, the step schedules, and
the driver reaching the exact ground energy."""
import math

import numpy as np
import pytest

from oracle import bug as obug
from oracle import closure as oc
from oracle import ite
from oracle import linalg as la
from oracle import models as om
from tests.helpers import random_legs


def _random_mps(L, dmax, seed):
    rng = np.random.default_rng(seed)
    return obug.MPS(om.random_state(random_legs(rng, L, dmax), rng), center=0)


@pytest.mark.parametrize("L,delta,seed", [(6, 0.7, 1), (8, 1.0, 2), (4, 0.0, 3)])
def test_energy_and_overlap_match_the_dense_state(L, delta, seed):
    psi = _random_mps(L, 4, seed)
    psi.center = L // 2                      # not canonical: the formulas hold in any gauge
    v = om.mps_dense(psi.sites)
    H = om.dense_hamiltonian(L, delta=delta)
    mpo = om.xxz_mpo(L, delta=delta)
    assert ite.overlap(psi) == pytest.approx(np.vdot(v, v).real, rel=1e-12)
    assert ite.energy(psi, mpo) == pytest.approx((np.vdot(v, H @ v) / np.vdot(v, v)).real,
                                                 rel=1e-11, abs=1e-12)


def _frames(L, rl, dmax, seed):
    """Left isometries on sites 0..rl and right isometries on rl+1..L-1 from two random states."""
    a = _random_mps(L, dmax, seed)
    b = _random_mps(L, dmax, seed + 100)
    obug.canonical(a, rl + 1)
    obug.canonical(b, L - 1)
    obug.canonical(b, rl)
    left = [a.sites[i] for i in range(rl + 1)]
    right = [None] * (L - 1)
    for bd in range(rl, L - 1):
        right[bd] = b.sites[bd + 1]
    return oc.KLFrames(left, right, rl, None, None), a


def _unit_cores(lleg, rleg):
    for q, dl in lleg.items():
        if q not in rleg:
            continue
        for i in range(dl):
            for j in range(rleg[q]):
                f = obug.Fam.zeros(lleg, rleg, [0])
                f.blocks[(0, q)][i, j] = 1.0
                yield f


@pytest.mark.parametrize("L,rl,delta,seed", [(6, 2, 0.5, 4), (8, 3, 1.0, 5), (6, 1, 0.0, 6)])
def test_outside_action_bounds_the_dense_complement(L, rl, delta, seed):
    """max(left, right) <= ||(I - P) H v|| <= sqrt(left^2 + right^2) for nearest-neighbour H,
    with P the projector on the frames' span, formed densely from unit cores."""
    fr, psi = _frames(L, rl, 3, seed)
    mpo = om.xxz_mpo(L, delta=delta)
    lleg, rleg = fr.left[rl].col, fr.right[rl].row
    rng = np.random.default_rng(seed)
    core = obug.Fam.zeros(lleg, rleg, [0])
    for k, b in core.blocks.items():
        core.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    opts = obug.StepOptions()
    alloc = obug.basis_allocation(1e-5, L - 1, -0.1j, 1.0, 1)
    act = oc.complement_action(psi, mpo, fr, core, alloc, opts, 1 << 30, la.ProbeRNG(1), "lanczos")
    H = om.dense_hamiltonian(L, delta=delta)
    V = np.stack([om.mps_dense(oc.frontier_mps(psi, fr, u).sites) for u in _unit_cores(lleg, rleg)],
                 axis=1)
    assert np.allclose(V.conj().T @ V, np.eye(V.shape[1]), atol=1e-12)
    v = om.mps_dense(oc.frontier_mps(psi, fr, core).sites)
    Hv = H @ v
    out = np.linalg.norm(Hv - V @ (V.conj().T @ Hv))
    assert max(act["left"], act["right"]) <= out * (1 + 1e-10) + 1e-13
    assert out <= act["bound"] * (1 + 1e-10) + 1e-13
    assert len(act["per_left"]) == rl + 1 and len(act["per_right"]) == L - 1 - rl


@pytest.mark.parametrize("start,L,order", [("neel", 8, 1), ("dw", 8, 2)])
def test_root_report_energies_are_the_states_energies(start, L, order):
    """e_pre is the energy of the state (it lies in the frames' span) whatever the shift; e_root
    is the energy of the evolved root vector, which the lossless split keeps."""
    mpo = om.xxz_mpo(L, delta=0.5)
    make = om.neel_state if start == "neel" else om.domain_wall_state
    e0 = ite.energy(obug.MPS(make(L)), mpo)
    reps = []
    for es in (0.0, 0.37):
        psi = obug.MPS(make(L))
        info = obug.cbe_bug_step(psi, mpo, -0.1, obug.StepOptions(
            order=order, eshift=es, root_report=True, truncate=False))
        rep = info.root_report
        assert rep["e_pre"] == pytest.approx(e0, abs=1e-10)
        assert rep["e_root"] == pytest.approx(ite.energy(psi, mpo), abs=1e-10)
        assert rep["e_root"] < rep["e_pre"]            # imaginary time lowers the energy
        reps.append(rep)
    assert reps[0]["e_root"] == pytest.approx(reps[1]["e_root"], abs=1e-6)


def test_outside_residual_measures_the_evolved_vector():
    """The reported defect is |tau| ||(I - P) H psi_1|| / ||psi_1|| of the root result in the
    final frames: a step with a loose tol_basis takes no re-enrichment pass."""
    L = 8
    mpo = om.xxz_mpo(L, delta=1.0)
    psi = obug.MPS(om.neel_state(L))
    info = obug.cbe_bug_step(psi, mpo, -0.05, obug.StepOptions(
        order=2, tol_basis=1.0, root_report=True, outside_residual=True))
    rep = info.root_report
    assert rep["outside_passes"] == 0 and rep["outside_added"] == 0
    assert rep["resid"] == pytest.approx(0.05 * rep["outside"], rel=1e-12)
    assert rep["resid"] == rep["resid_initial"]
    assert rep["outside"] >= max(rep["outside_left"], rep["outside_right"])
    # a tight tol_basis re-enriches until the defect fits or nothing is added
    psi = obug.MPS(om.neel_state(L))
    info = obug.cbe_bug_step(psi, mpo, -0.05, obug.StepOptions(
        order=2, tol_basis=1e-9, root_report=True, outside_residual=True, maxdim=64))
    rep2 = info.root_report
    assert rep2["outside_passes"] >= 1
    assert rep2["resid"] <= rep2["resid_initial"] + 1e-15


def test_measured_closure_starts_from_the_single_pass_measurement():
    """With exact leaks (fold_omega = false) and frames that cannot grow (the product state's
    per-bond work ceiling binds), the measured loop's first reading of frontier 1 on the incoming
    frames is the single pass's leak of frontier 1, and both report the closure unmet.  With room
    to grow the loop converges, reports one leak per degree, and the step is as accurate as the
    single pass against the exact evolution (its error is the order-2 step's own, O(dt^3))."""
    import scipy.linalg as sla
    L = 8
    mpo = om.xxz_mpo(L, delta=0.5)
    for rule, closes in (("bond", False), ("maxdim", True)):
        base = dict(order=2, fold_omega=False, tol_basis=1e-6, work_ceiling_rule=rule, maxdim=32)
        a = obug.MPS(om.neel_state(L))
        ia = obug.cbe_bug_step(a, mpo, -0.2j, obug.StepOptions(**base))
        b = obug.MPS(om.neel_state(L))
        ib = obug.cbe_bug_step(b, mpo, -0.2j, obug.StepOptions(closure_checks=True, **base))
        assert len(ib.ledger.leak_initial) == 2 and len(ib.ledger.leak) == 2
        if not closes:
            assert ib.ledger.leak_initial[1] == pytest.approx(ia.ledger.leak[1], rel=1e-10)
            assert "closure" in ib.ledger.unmet and "closure" in ia.ledger.unmet
            continue
        assert "closure" not in ib.ledger.unmet and ib.ledger.n_passes >= 1
        assert max(ib.ledger.leak) < 1e-12                  # every frontier closed at the end
        ex = sla.expm(-0.2j * om.dense_hamiltonian(L, delta=0.5)) @ om.mps_dense(om.neel_state(L))
        err = [np.linalg.norm(om.mps_dense(p.sites) - ex) for p in (a, b)]
        assert err[1] <= err[0] * (1 + 1e-6) + 1e-12 and err[1] < 1e-3


def test_closure_probe_measures_without_changing_the_step():
    L = 8
    mpo = om.xxz_mpo(L, delta=0.5)
    a = obug.MPS(om.neel_state(L))
    obug.cbe_bug_step(a, mpo, -0.1j, obug.StepOptions())
    b = obug.MPS(om.neel_state(L))
    ib = obug.cbe_bug_step(b, mpo, -0.1j, obug.StepOptions(closure_probe=2))
    assert np.array_equal(om.mps_dense(a.sites), om.mps_dense(b.sites))
    assert len(ib.ledger.leak_initial) == 2 and len(ib.ledger.rho_initial) == 2
    with pytest.raises(ValueError):
        obug.cbe_bug_step(b, mpo, -0.1j, obug.StepOptions(closure_probe=-1))


def test_schedules():
    s = ite.log_panel_steps(8.0, 3, per_panel=4)
    assert s == [0.5] * 4 + [0.5] * 4 + [1.0] * 4
    assert math.isclose(sum(ite.log_panel_steps(3.0, 5, per_panel=2)), 3.0)
    assert max(ite.log_panel_steps(3.0, 5, per_panel=2)) == pytest.approx(3.0 / 4)
    assert ite.uniform_beta_steps(2.0, 4) == [0.5] * 4
    assert ite.beta_grid([0.5, 0.25]) == [0.0, 0.5, 0.75]
    assert ite.geometric_schedule(0.01, 2.0, 0.05, 5) == [0.01, 0.02, 0.04, 0.05, 0.05]
    for bad in (lambda: ite.log_panel_steps(0, 2), lambda: ite.log_panel_steps(1, 0),
                lambda: ite.uniform_beta_steps(1, 0)):
        with pytest.raises(ValueError):
            bad()


def test_nsplit_rule():
    a = dict(unmet=["outside"], resid=1e-3)
    assert ite.nsplit(a, 1e-5, 2) == math.ceil(1.2 * (100.0) ** (1 / 3))
    assert ite.nsplit(dict(unmet=["trunc"], resid=1e-3), 1e-5, 2) == 2
    assert ite.nsplit(dict(unmet=["outside"], resid=float("nan")), 1e-5, 2) == 2
    assert ite.nsplit(dict(unmet=["outside"], resid=1e30), 1e-5, 1) == ite.ITE_MAXSPLIT


def test_ite_follows_the_exact_imaginary_time_energy():
    """Heisenberg L = 6 from Neel to beta = 6: the normalised evolution's energy follows the exact
    E(beta) = <psi0|H exp(-2 beta H)|psi0> / <psi0|exp(-2 beta H)|psi0>, each step lowering it,
    with the order-2 error falling about fourfold when the step halves.  The first step from the
    product state binds the per-bond work ceiling, so it is flagged, not split."""
    L = 6
    mpo = om.xxz_mpo(L, delta=1.0)
    w, U = np.linalg.eigh(om.dense_hamiltonian(L, delta=1.0))
    c = U.conj().T @ om.mps_dense(om.neel_state(L))

    def exact(beta):
        p = np.abs(c) ** 2 * np.exp(-2 * beta * (w - w[0]))
        return float(p @ w / p.sum())

    errs = []
    for steps in (24, 48):
        psi = obug.MPS(om.neel_state(L))
        log = ite.ite_evolve(psi, mpo, ite.uniform_beta_steps(6.0, steps), maxdim=16)
        es = [r["e_after"] for r in log]
        assert all(b <= a + 1e-9 for a, b in zip([log[0]["e_before"]] + es, es))
        assert log[-1]["beta"] == pytest.approx(6.0)
        assert abs(psi.sites[psi.center].norm() - 1.0) < 1e-12
        assert log[0]["status"] == "failed" and log[0]["nsub"] == 1
        assert all(r["status"] == "met" for r in log[1:])
        errs.append(abs(es[-1] - exact(6.0)))
    assert errs[0] < 3e-4 and errs[1] < errs[0] / 2.5


def test_ite_retry_and_split_logic(monkeypatch):
    """The driver's control flow on scripted attempts: an unclosed closure is redone with the
    measured loop, an outside-defect failure splits by (resid / tol)^(1/(p+1)), a failure with no
    retries left is flagged, and on_fail = 'error' raises."""
    calls = []

    class _Lg:
        def __init__(self, capped):
            self.maxdim_reached, self.workcap_reached = capped, False

    class _Info:
        def __init__(self, capped):
            self.ledger = _Lg(capped)

    def fake(psi, mpo, dbeta, *, closure_checks, order, tol_basis, capped=False, **kw):
        calls.append((round(dbeta, 6), closure_checks))
        if dbeta > 0.3:                        # big steps miss the outside defect
            unmet, resid = ["outside"], 1e3 * tol_basis * dbeta
        elif dbeta > 0.2 and not closure_checks:
            unmet, resid = ["closure"], 0.0
        else:
            unmet, resid = [], 0.0
        return dict(dbeta=dbeta, ok=not unmet, unmet=unmet, e_before=0.0, e_pre_root=0.0,
                    e_root=0.0, e_after=-dbeta, de_root=0.0, de_compression=0.0, outside=resid,
                    resid=resid, root_err=0.0, discarded=0.0, chi=1, info=_Info(capped))

    def attempt(psi, mpo, dbeta, **kw):
        kw.pop("step_kw")
        for k in ("tol_trunc", "tol_root", "maxdim", "eshift", "outside_residual", "parallel",
                  "rng"):
            kw.pop(k)
        return fake(psi, mpo, dbeta, **kw)

    monkeypatch.setattr(ite, "attempt_step", attempt)
    monkeypatch.setattr(ite, "normalise", lambda psi: 1.0)
    psi = obug.MPS(om.neel_state(4))
    mpo = om.xxz_mpo(4)
    # 0.25: closure fails, the measured loop passes
    rec = ite.ite_step(psi, mpo, 0.25, tol_basis=1e-5)
    assert rec["status"] == "retried" and calls == [(0.25, False), (0.25, True)]
    # 0.5: outside defect 5e-3 -> ceil(1.2 * 500^(1/3)) = 10 sub-steps of 0.05
    calls.clear()
    rec = ite.ite_step(psi, mpo, 0.5, tol_basis=1e-5)
    n = math.ceil(1.2 * (500.0) ** (1 / 3))
    assert rec["status"] == "retried" and rec["nsub"] == n
    assert calls == [(0.5, False)] + [(round(0.5 / n, 6), False)] * n
    assert rec["dbeta"] == pytest.approx(0.5)
    # no retries left: flagged, or raised
    rec = ite.ite_step(psi, mpo, 0.5, tol_basis=1e-5, retries=0)
    assert rec["status"] == "failed" and rec["unmet"] == ["outside"]
    with pytest.raises(ArithmeticError):
        ite.ite_step(psi, mpo, 0.5, tol_basis=1e-5, retries=0, on_fail="error")
    with pytest.raises(ValueError):
        ite.ite_step(psi, mpo, -0.1)
    with pytest.raises(ValueError):
        ite.ite_step(psi, mpo, 0.1, root_method="radau2")
