"""The step-level C entry point ``bug_step`` (the whole step orchestrated in C++ inside libbug.so)
against the oracle on identical inputs and probe seeds (GPU).

The bar is the step's own: relative state difference <= 1e-10 per step (dense amplitudes for
L <= 12, the direct-sum distance beyond), local magnetisation within 1e-9, identical bond
dimensions, verdicts and per-bond logs."""
import numpy as np
import pytest

from oracle import bug as obug
from oracle import models as om
from tests.helpers import fam_from_device, mps_distance

pytestmark = pytest.mark.gpu

_LOG_EXACT = ("bond", "n_probes", "n_selected", "n_krylov", "n_h1", "width", "r_old", "r_final")


def _states(start, L):
    from bondupdategalerkin_amd import mps

    make_d = {"neel": mps.neel_state, "dw": mps.domain_wall_state}[start]
    make_o = {"neel": om.neel_state, "dw": om.domain_wall_state}[start]
    return make_d(L), obug.MPS(make_o(L))


def _dense(psi):
    return om.mps_dense([fam_from_device(t) for t in psi.tensors])


def _same_log(dev, ora):
    assert len(dev) == len(ora)
    for d, o in zip(dev, ora):
        assert {k: d[k] for k in _LOG_EXACT} == {k: o[k] for k in _LOG_EXACT}
        assert d["side"] == (0 if o["side"] == "left" else 1)
        assert bool(d["met"]) == o["met"] and bool(d["workcap"]) == o["workcap"]
        for k in ("sample", "krylov_c", "krylov_d"):
            assert abs(d[k] - o[k]) <= 1e-6 * max(abs(d[k]), abs(o[k])) + 1e-12, (d["bond"], k)


@pytest.mark.parametrize("start,L,delta,rule,steps,order,parallel", [
    ("neel", 10, 0.0, "maxdim", 3, 1, False),
    ("dw", 10, 0.0, "bond", 3, 1, False),
    ("neel", 8, 0.5, "bond", 3, 1, False),
    ("neel", 10, 0.0, "maxdim", 3, 2, False),
    ("dw", 12, 0.0, "bond", 3, 2, False),
    ("neel", 10, 0.5, "maxdim", 4, 2, False),
    ("neel", 8, 1.0, "bond", 2, 3, False),
    ("neel", 10, 0.5, "maxdim", 3, 1, True),
    ("dw", 12, 0.0, "bond", 3, 2, True),
])
def test_c_step_matches_oracle(gpu, start, L, delta, rule, steps, order, parallel):
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd.mpo import xxz_mpo

    psi_d, psi_o = _states(start, L)
    st, mp = C.CState(psi_d), C.CMPO(xxz_mpo(L, delta=delta))
    mpo_o = om.xxz_mpo(L, delta=delta)
    print(f"case {start} L={L} delta={delta} {rule} order={order} parallel={parallel}", flush=True)
    for n in range(steps):
        info_d = C.bug_step(st, mp, -0.05j, work_ceiling_rule=rule, order=order, parallel=parallel)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -0.05j, obug.StepOptions(
            work_ceiling_rule=rule, order=order, parallel=parallel))
        psi = st.to_mps()
        assert psi.bond_dims() == psi_o.bond_dims(), n
        assert info_d["status"] == info_o.ledger.status, n
        assert info_d["unmet_names"] == sorted(info_o.ledger.unmet,
                                               key=lambda u: [b[1] for b in C.UNMET_BITS].index(u))
        _same_log(info_d["bond_log"], info_o.ledger.bond_log)
        vd, vo = _dense(psi), om.mps_dense(psi_o.sites)
        assert np.linalg.norm(vd - vo) / np.linalg.norm(vo) <= 1e-10, n
        assert np.abs(mps.magnetisation(psi) - obug.sz_profile(psi_o)).max() <= 1e-9, n


def test_c_step_xx_l64_drop_in_configuration(gpu):
    """The FP64 drop-in configuration through the C entry point: XX L = 64 domain wall, dt 0.05,
    order 2, tol_basis 1e-3, maxdim 256, bond_ceiling_factor 1.2, 20 steps; every step the same
    bonds and verdict as the oracle and the state within 1e-10 (direct-sum distance)."""
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L, dt = 64, 0.05
    psi_d, psi_o = _states("dw", L)
    st, mp = C.CState(psi_d), C.CMPO(xxz_mpo(L, delta=0.0))
    mpo_o = om.xxz_mpo(L, delta=0.0)
    kw = dict(order=2, tol_basis=1e-3, maxdim=256, bond_ceiling_factor=1.2)
    for n in range(1, 21):
        info_d = C.bug_step(st, mp, -1j * dt, **kw)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -1j * dt, obug.StepOptions(**kw))
        psi = st.to_mps()
        assert psi.bond_dims() == psi_o.bond_dims(), n
        assert info_d["status"] == info_o.ledger.status == "met", n
        assert mps_distance([fam_from_device(t) for t in psi.tensors], psi_o.sites) <= 1e-10, n


def test_c_step_xx_l20_reference_configuration(gpu):
    """The plain reference configuration: XX L = 20 domain wall, 20 steps of tau = -0.05 i,
    order 2, probe seed 1, against the oracle and the free-fermion magnetisation."""
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from tests.helpers import free_fermion_sz

    L, dt = 20, 0.05
    psi_d, psi_o = _states("dw", L)
    st, mp = C.CState(psi_d), C.CMPO(xxz_mpo(L, delta=0.0))
    mpo_o = om.xxz_mpo(L, delta=0.0)
    for n in range(1, 21):
        info_d = C.bug_step(st, mp, -1j * dt, order=2, rng=1)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, -1j * dt, obug.StepOptions(order=2, seed=1))
        psi = st.to_mps()
        assert psi.bond_dims() == psi_o.bond_dims(), n
        assert info_d["status"] == info_o.ledger.status, n
        assert mps_distance([fam_from_device(t) for t in psi.tensors], psi_o.sites) <= 1e-10, n
    ff = free_fermion_sz(L, 20 * dt, range(L // 2))
    assert np.abs(mps.magnetisation(st.to_mps()) - ff).max() < 1e-5


def test_c_step_arnoldi_root_matches_oracle(gpu):
    """hermitian = false: the Arnoldi root solve and Arnoldi local chains, on a Heston QTT
    generator (4 + 4 bits, no symmetry), against the oracle at 1e-10."""
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import qtt as Q
    from oracle import heston as oh
    from tests.test_heston_gpu import _heston_inputs, _tt_distance

    g, init, W = _heston_inputs(4, False)
    psi_d = Q.qtt_mps(init)
    st, mp = C.CState(psi_d), C.CMPO(Q.qtt_to_mpo(W))
    psi_o = obug.MPS(oh.fams_from_cores(init), center=len(init) - 1)
    obug.canonical(psi_o, 0)
    mpo_o = oh.mpo_from_cores(W)
    opts = dict(tol_basis=1e-8, tol_trunc=1e-12, maxdim=64, hermitian=False, order=2)
    for k in range(3):
        info_d = C.bug_step(st, mp, 0.02, **opts)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, 0.02, obug.StepOptions(**opts))
        psi = st.to_mps()
        assert psi.bond_dims() == psi_o.bond_dims(), k
        assert info_d["status"] == info_o.ledger.status, k
        assert _tt_distance([fam_from_device(t) for t in psi.tensors], psi_o.sites) <= 1e-10, k


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-7)])
def test_c_step_equals_python_device_step(gpu, precision, tol):
    """The C++ orchestration and the package's Python orchestration drive the same kernels with the
    same probe stream (parallel half-sweeps, order 2, XXZ chain): equal to the parity bar, 1e-10,
    in FP64 (their small host solvers differ: scipy's expm and tridiagonal eigensolver against
    the library's own, measured 5e-12 apart after two steps).  In the mixed form the FP32
    contractions round inputs that differ in the last bits differently, so the two agree to the
    FP32 level, 1e-7, as the mixed form agrees with the FP64 one."""
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 16
    mpo = xxz_mpo(L, delta=0.5)
    psi_p = mps.neel_state(L)
    st, mp = C.CState(mps.neel_state(L)), C.CMPO(mpo)
    kw = dict(order=2, precision=precision, parallel=True, work_ceiling_rule="maxdim", maxdim=64)
    for n in range(4):
        info_p = B.cbe_bug_step(psi_p, mpo, -0.05j, **kw)
        info_c = C.bug_step(st, mp, -0.05j, **kw)
        psi_c = st.to_mps()
        assert psi_c.bond_dims() == psi_p.bond_dims(), n
        assert info_c["status"] == B.budget_status(info_p.ledger), n
        assert mps_distance([fam_from_device(t) for t in psi_c.tensors],
                            [fam_from_device(t) for t in psi_p.tensors]) <= tol, n


@pytest.mark.parametrize("method,shift,kind", [("radau2", False, "heston"),
                                                ("trapezoid", False, "heston"),
                                                ("radau3", True, "chain")])
def test_c_step_rational_root_equals_python_device_step(gpu, method, shift, kind):
    """The rational roots (root_methods.jl:5-71,143-157) through the C entry point against the
    package's Python step: an unshifted Radau / trapezoid root on a Heston grid (the linear-flow
    step of qtt_bug_step!) and a shifted Radau IIA(3) root on an XXZ chain at real tau (the
    imaginary-time use, where the shift applies), equal to 1e-10 with the same Krylov size."""
    import math

    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd.mpo import xxz_mpo

    if kind == "heston":
        g = Q.qtt_grid(4, 4, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)
        init = Q.qtt_separable(g, lambda x: math.exp(0.5 * x), lambda v: math.exp(-v))
        make = lambda: Q.qtt_mps(init)                                   # noqa: E731
        mpo = Q.heston_mpo(g, Q.HestonParams(0.03, 0.01, 2.0, 0.09, 0.4, -0.7))
        tau, herm = 0.02, False
    else:
        make = lambda: mps.neel_state(12)                                # noqa: E731
        mpo = xxz_mpo(12, delta=0.5)
        tau, herm = -0.05, True
    kw = dict(order=2, maxdim=32, work_ceiling_rule="maxdim", hermitian=herm, root_method=method,
              root_shift=shift, root_maxiter=200 if kind == "heston" else 30, tol_root=1e-11)
    psi_p = make()
    st, mp = C.CState(make()), C.CMPO(mpo)
    for n in range(3):
        print(f"case {method} shift={shift} {kind} step {n}", flush=True)
        info_p = B.cbe_bug_step(psi_p, mpo, tau, **kw)
        info_c = C.bug_step(st, mp, tau, **kw)
        psi_c = st.to_mps()
        assert psi_c.bond_dims() == psi_p.bond_dims(), n
        assert info_c["expv_iters"] == info_p.ledger.expv_iters, n
        assert info_c["status"] == B.budget_status(info_p.ledger), n
        assert mps_distance([fam_from_device(t) for t in psi_c.tensors],
                            [fam_from_device(t) for t in psi_p.tensors]) <= 1e-10, n


def test_c_step_refuses_options_before_touching_the_state(gpu):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 8
    st, mp = C.CState(mps.neel_state(L)), C.CMPO(xxz_mpo(L))
    before = _dense(st.to_mps())
    bad = [dict(order=0), dict(root=L - 1), dict(exact=True), dict(fold_omega=False, order=2),
           dict(krylov_rank_tol=1e-3), dict(expansion=B.DDQRExpansion(mode="adaptive")),
           dict(work_ceiling_rule="maxdim", bond_ceiling_factor=1.5), dict(tol_basis=0.0)]
    for kw in bad:
        with pytest.raises(_lib.BugError) as e:
            C.bug_step(st, mp, -0.05j, **kw)
        assert e.value.code == _lib.BUG_EINVAL, kw
    assert np.array_equal(_dense(st.to_mps()), before)
    with pytest.raises(_lib.BugError):
        C.bug_step(st, C.CMPO(xxz_mpo(L + 2)), -0.05j)


def test_state_handles_device_host_clone_and_magnetisation(gpu):
    """A state handle made from device tensors equals one made through host copies; a clone shares
    the tensors and evolves independently; the handle's S^z equals the package's magnetisation."""
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 12
    psi = mps.random_mps(L, 16, seed=3, tail=1e-4)
    B.canonical(psi, 0)
    T.scale_(1.0 / T.norm(psi[0]), psi[0])
    a, b = C.CState(psi), C.CState(psi, via_host=True)
    assert np.array_equal(_dense(a.to_mps()), _dense(b.to_mps()))
    assert np.abs(a.expect_diagonal() - mps.magnetisation(psi)).max() <= 1e-12
    c = a.clone()
    mp = C.CMPO(xxz_mpo(L, delta=0.5))
    C.bug_step(a, mp, -0.05j, order=2)
    assert np.array_equal(_dense(c.to_mps()), _dense(b.to_mps()))   # the clone did not move
    assert np.abs(a.expect_diagonal() - mps.magnetisation(a.to_mps())).max() <= 1e-12
    assert a.bond_dims() == a.to_mps().bond_dims()
