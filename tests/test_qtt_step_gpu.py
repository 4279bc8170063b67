"""``qtt_step.qtt_bug_step`` (the linear-flow branch of ``qtt_bug_step!``, the rational root of
the order's implicit Runge-Kutta method on the Galerkin operator) on the device, against the
oracle's step from the same input and probe seed, and against the dense R(hA) psi when the frames
span the whole space (GPU)."""
import math

import numpy as np
import pytest

from oracle import bug as obug
from oracle import heston as oh
from oracle import krylov as ok
from oracle import models as om
from oracle import qtt_step as oq
from tests.helpers import device_from_fam, fam_from_device

pytestmark = pytest.mark.gpu

HP = (0.03, 0.01, 2.0, 0.09, 0.4, -0.7)


def _grid(n):
    from bondupdategalerkin_amd import qtt as Q

    return Q.qtt_grid(n, n, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)


def _device_mps(fams, center):
    from bondupdategalerkin_amd.mpo import SPIN_CHARGES_NONE
    from bondupdategalerkin_amd.mps import SymMPS

    return SymMPS([device_from_fam(f) for f in fams], SPIN_CHARGES_NONE, center=center)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_full_frame_step_is_the_dense_rational_step(gpu, order):
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd import qtt_step as S

    n = 3
    W = Q.heston_op(_grid(n), Q.HestonParams(*HP))
    A = oh.mpo_grid_matrix(W)
    L = 2 * n
    legs = [{0: min(2 ** min(b, L - b), 64)} for b in range(L + 1)]
    psi_o = obug.MPS(om.random_state(legs, np.random.default_rng(11), symmetry="none"))
    obug.canonical(psi_o, 0)
    v0 = oh.mps_grid_vector(psi_o.sites)
    psi = _device_mps(psi_o.sites, 0)
    h = 0.05
    print(f"case full frames order={order}", flush=True)
    rep = S.qtt_bug_step(psi, S.LinearFlow(Q.qtt_to_mpo(W)), 0.0, h, order=order,
                         tol_basis=1e-10, tol_trunc=1e-14, maxdim=64, tol_stage=1e-13,
                         tol_scaling="fixed")
    want = ok.rational_apply(S.qtt_root_for_order(order), h * A, v0)
    got = oh.mps_grid_vector([fam_from_device(t) for t in psi.tensors])
    assert rep["converged"] and rep["frame_defect"] < 1e-12 and rep["miss_input"] < 1e-12
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)


@pytest.mark.parametrize("n,order,steps", [(4, 2, 3), (4, 3, 2), (5, 2, 2)])
def test_qtt_step_matches_oracle(gpu, n, order, steps):
    """Heston backward pricing steps from the smooth exp(0.5 x) exp(-v) datum: the device step
    equals the oracle's (state, bond dimensions, verdict, root iterations) step by step."""
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd import qtt_step as S

    g = _grid(n)
    init = Q.qtt_separable(g, lambda x: math.exp(0.5 * x), lambda v: math.exp(-v))
    W = Q.heston_op(g, Q.HestonParams(*HP))
    psi_d = Q.qtt_mps(init)
    psi_o = obug.MPS(oh.fams_from_cores(init), center=len(init) - 1)
    obug.canonical(psi_o, 0)
    flow_d, flow_o = S.LinearFlow(Q.qtt_to_mpo(W)), oq.LinearFlow(oh.mpo_from_cores(W))
    kw = dict(order=order, maxdim=32, tol_basis=1e-6, tol_trunc=1e-9)
    for k in range(steps):
        print(f"case n={n} order={order} step={k}", flush=True)
        rd = S.qtt_bug_step(psi_d, flow_d, k * 0.02, 0.02, **kw)
        ro = oq.qtt_bug_step(psi_o, flow_o, k * 0.02, 0.02, **kw)
        assert rd["root"] == ro["root"] == S.qtt_root_for_order(order)
        assert rd["status"] == ro["status"] and rd["krylov"] == ro["krylov"], (rd, ro)
        assert psi_d.bond_dims() == psi_o.bond_dims(), k
        a = oh.mps_grid_vector([fam_from_device(t) for t in psi_d.tensors])
        b = oh.mps_grid_vector(psi_o.sites)
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b), k
        assert abs(rd["root_error"] - ro["root_error"]) <= 1e-6 * ro["root_error"] + 1e-15
