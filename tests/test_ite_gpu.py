"""GPU parity of the imaginary-time driver and the step instrumentation it uses with the oracle
on identical inputs and probe seeds: the shifted step with its root report and outside-Galerkin
defect, the measured closure loop, the closure probe, and whole ``ite_evolve`` runs.

Every step starts both sides from the same state (the oracle's, uploaded): these paths take
many rank and enrichment decisions on small quantities, and a chained run would compare the
device's state after the oracle's rounding-level differences had crossed one of them.  Bars:
states within 1e-10 (relative, dense amplitudes), energies within 1e-10, the defects and leaks
within 1e-6 relative or 1e-9 absolute (FP64 on both sides, the same probes; a leak is H's
action on the state, so its absolute precision is the state bar times ||H|| <= L), identical
verdicts, pass and sub-step counts."""
import math

import numpy as np
import pytest

from oracle import bug as obug
from oracle import ite as oite
from oracle import models as om
from tests.helpers import fam_from_device

pytestmark = pytest.mark.gpu


def _states(start, L):
    from bondupdategalerkin_amd import mps
    make_d = {"neel": mps.neel_state, "dw": mps.domain_wall_state}[start]
    make_o = {"neel": om.neel_state, "dw": om.domain_wall_state}[start]
    return make_d(L), obug.MPS(make_o(L))


def _upload(psi_o, charges):
    from bondupdategalerkin_amd.mps import SymMPS
    from tests.helpers import device_from_fam
    return SymMPS([device_from_fam(f) for f in psi_o.sites], charges, psi_o.center)


def _dense(psi):
    return om.mps_dense([fam_from_device(t) for t in psi.tensors])


def _close(a, b, rel=1e-6, floor=1e-12):
    if isinstance(a, float) and math.isnan(a):
        return isinstance(b, float) and math.isnan(b)
    return abs(a - b) <= rel * max(abs(a), abs(b)) + floor


def _state_err(psi_d, psi_o):
    vd, vo = _dense(psi_d), om.mps_dense(psi_o.sites)
    return np.linalg.norm(vd - vo) / np.linalg.norm(vo)


@pytest.mark.parametrize("start,L,delta,order,tol_basis,eshift", [
    ("neel", 8, 1.0, 1, 1e-5, 0.0),
    ("dw", 10, 0.5, 2, 1e-5, -1.3),
    ("neel", 10, 0.5, 2, 1e-9, 0.7),
    ("dw", 8, 1.0, 3, 1e-7, -0.5),
])
def test_shifted_step_with_root_report_matches_oracle(gpu, start, L, delta, order, tol_basis,
                                                      eshift):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd.mpo import xxz_mpo

    psi_d, psi_o = _states(start, L)
    mpo_d, mpo_o = xxz_mpo(L, delta=delta), om.xxz_mpo(L, delta=delta)
    print(f"case {start} L={L} order={order} tol_basis={tol_basis} eshift={eshift}", flush=True)
    for n in range(2):
        psi_d = _upload(psi_o, psi_d.charges)
        rep = {}
        info_d = B.cbe_bug_step(psi_d, mpo_d, -0.1, order=order, tol_basis=tol_basis,
                                eshift=eshift, root_report=rep, outside_residual=True,
                                work_ceiling_rule="maxdim", maxdim=32)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -0.1, obug.StepOptions(
            order=order, tol_basis=tol_basis, eshift=eshift, root_report=True,
            outside_residual=True, work_ceiling_rule="maxdim", maxdim=32))
        ro = info_o.root_report
        assert psi_d.bond_dims() == psi_o.bond_dims(), n
        assert _state_err(psi_d, psi_o) <= 1e-10, n
        assert info_d.ledger.unmet == info_o.ledger.unmet, n
        assert (rep["outside_passes"], rep["outside_added"]) == (ro["outside_passes"],
                                                                 ro["outside_added"]), n
        for k in ("e_pre", "e_root"):
            assert abs(rep[k] - ro[k]) <= 1e-10 * max(1.0, abs(ro[k])), (n, k)
        for k in ("resid", "resid_initial", "outside", "outside_left", "outside_right"):
            assert _close(rep[k], ro[k]), (n, k, rep[k], ro[k])
        assert len(rep["per_left"]) == len(ro["per_left"])
        for a, b in zip(rep["per_left"] + rep["per_right"], ro["per_left"] + ro["per_right"]):
            assert _close(a, b, floor=1e-14)


@pytest.mark.parametrize("start,L,delta,order,rule,fold", [
    ("neel", 10, 0.5, 2, "maxdim", None),
    ("dw", 10, 0.0, 2, "bond", None),
    ("neel", 8, 1.0, 3, "maxdim", None),
    ("neel", 10, 0.5, 2, "maxdim", False),
])
def test_measured_closure_matches_oracle(gpu, start, L, delta, order, rule, fold):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd.mpo import xxz_mpo

    psi_d, psi_o = _states(start, L)
    mpo_d, mpo_o = xxz_mpo(L, delta=delta), om.xxz_mpo(L, delta=delta)
    print(f"case {start} L={L} order={order} {rule} fold={fold}", flush=True)
    for n in range(3):
        psi_d = _upload(psi_o, psi_d.charges)
        info_d = B.cbe_bug_step(psi_d, mpo_d, -0.1j, order=order, closure_checks=True,
                                work_ceiling_rule=rule, fold_omega=fold, maxdim=32)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -0.1j, obug.StepOptions(
            order=order, closure_checks=True, work_ceiling_rule=rule, fold_omega=fold,
            maxdim=32))
        ld, lo = info_d.ledger, info_o.ledger
        assert psi_d.bond_dims() == psi_o.bond_dims(), n
        assert _state_err(psi_d, psi_o) <= 1e-10, n
        assert ld.unmet == lo.unmet and ld.n_passes == lo.n_passes, n
        assert len(ld.leak_initial) == len(lo.leak_initial) == order
        for a, b in zip(ld.leak_initial + ld.rho_initial + ld.leak,
                        lo.leak_initial + lo.rho_initial + lo.leak):
            assert _close(a, b, floor=1e-9), n


def test_closure_probe_matches_oracle_and_leaves_the_step_unchanged(gpu):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 10
    psi_d, psi_o = _states("neel", L)
    plain, _ = _states("neel", L)
    mpo_d, mpo_o = xxz_mpo(L, delta=0.5), om.xxz_mpo(L, delta=0.5)
    for n in range(2):
        psi_d = _upload(psi_o, psi_d.charges)
        plain = _upload(psi_o, psi_d.charges)
        info_d = B.cbe_bug_step(psi_d, mpo_d, -0.05j, closure_probe=3)
        B.cbe_bug_step(plain, mpo_d, -0.05j)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -0.05j, obug.StepOptions(closure_probe=3))
        assert np.array_equal(_dense(psi_d), _dense(plain)), n
        assert _state_err(psi_d, psi_o) <= 1e-10, n
        for a, b in zip(info_d.ledger.leak_initial + info_d.ledger.rho_initial,
                        info_o.ledger.leak_initial + info_o.ledger.rho_initial):
            assert _close(a, b, floor=1e-9), n


@pytest.mark.parametrize("start,L,delta,schedule,kw", [
    ("neel", 8, 1.0, ("uniform", 3.0, 12), dict(maxdim=16)),
    ("dw", 10, 0.5, ("panel", 2.0, 3), dict(maxdim=24, order=1)),
    ("neel", 10, 1.0, ("uniform", 1.0, 4), dict(maxdim=32, tol_basis=1e-8,
                                                work_ceiling_rule="maxdim")),
])
def test_ite_evolve_matches_oracle(gpu, start, L, delta, schedule, kw):
    from bondupdategalerkin_amd import ite
    from bondupdategalerkin_amd.mpo import xxz_mpo

    kind, beta, count = schedule
    steps = (ite.uniform_beta_steps(beta, count) if kind == "uniform"
             else ite.log_panel_steps(beta, count, per_panel=2))
    assert steps == (oite.uniform_beta_steps(beta, count) if kind == "uniform"
                     else oite.log_panel_steps(beta, count, per_panel=2))
    psi_d, psi_o = _states(start, L)
    charges = psi_d.charges
    mpo_d, mpo_o = xxz_mpo(L, delta=delta), om.xxz_mpo(L, delta=delta)
    print(f"case {start} L={L} {schedule} {kw}", flush=True)
    # one ite_evolve step at a time, each from the oracle's state: the records are ite_evolve's
    log_d, log_o = [], []
    for n, db in enumerate(steps):
        psi_d = _upload(psi_o, charges)
        d = ite.ite_evolve(psi_d, mpo_d, [db], **kw)[0]
        o = oite.ite_evolve(psi_o, mpo_o, [db], **kw)[0]
        assert (d["status"], d["unmet"], d["nsub"], len(d["attempts"])) == (
            o["status"], o["unmet"], o["nsub"], len(o["attempts"])), n
        assert d["dbeta"] == pytest.approx(o["dbeta"], abs=1e-15)
        for k in ("e_before", "e_pre_root", "e_root", "e_after", "de_root"):
            assert abs(d[k] - o[k]) <= 1e-10 * max(1.0, abs(o[k])), (n, k)
        assert _close(d["resid"], o["resid"]), (n, d["resid"], o["resid"])
        assert d["chi"] == o["chi"], n
        assert psi_d.bond_dims() == psi_o.bond_dims(), n
        assert _state_err(psi_d, psi_o) <= 1e-10, n
        log_d.append(d)
        log_o.append(o)
    # each accepted step lowers the energy
    es = [r["e_after"] for r in log_d]
    assert all(b <= a + 1e-9 for a, b in zip([log_d[0]["e_before"]] + es, es))


def _sparse_heisenberg(L):
    import scipy.sparse as sp
    sp_ = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))
    sm = sp_.T.tocsr()
    sz = sp.csr_matrix(np.diag([0.5, -0.5]))
    eye = sp.identity(2, format="csr")

    def op(o, i):
        m = sp.identity(1, format="csr")
        for j in range(L):
            m = sp.kron(m, o if j == i else eye, format="csr")
        return m

    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L - 1):
        H = H + 0.5 * (op(sp_, i) @ op(sm, i + 1) + op(sm, i) @ op(sp_, i + 1)) \
            + op(sz, i) @ op(sz, i + 1)
    return H


def test_ite_l16_heisenberg_follows_the_exact_evolution(gpu):
    """A longer run on the device alone: Heisenberg L = 16 from Neel over a logarithmic panel
    grid (24 steps) to beta = 8 at maxdim 64.  The energy falls at every step and ends within 1e-3
    of the exact normalised exp(-8 H)|Neel> energy (sparse action on the 2^16 space, local state 0
    up, site 0 most significant); the state stays normalised."""
    import scipy.sparse.linalg as spla

    from bondupdategalerkin_amd import ite, mps
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 16
    psi = mps.neel_state(L)
    log = ite.ite_evolve(psi, xxz_mpo(L, delta=1.0), ite.log_panel_steps(8.0, 4, per_panel=6),
                         maxdim=64)
    es = [r["e_after"] for r in log]
    assert all(np.isfinite(es))
    assert all(b <= a + 1e-8 for a, b in zip([log[0]["e_before"]] + es, es))
    H = _sparse_heisenberg(L)
    v = np.zeros(2 ** L)
    v[int("".join("0" if j % 2 == 0 else "1" for j in range(L)), 2)] = 1.0
    w = spla.expm_multiply(-8.0 * H, v)
    w /= np.linalg.norm(w)
    assert abs(es[-1] - float(w @ (H @ w))) < 1e-3
    assert abs(T.norm(psi[psi.center]) - 1.0) < 1e-12
