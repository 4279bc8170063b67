"""GPU parity of the two-site TDVP with the oracle, step by step from the same state: the fused
pair generator on the one-site H_eff kernel, the sector-batched two-site SVD, the budgeted and
relative cuts, real and imaginary time, Lanczos and Arnoldi.

Bars: states within 1e-10 (relative, dense amplitudes), identical bond dimensions, identical
structural counts (largest block, working set), the discarded weights within 1e-10 absolute.
The operator-application counts are not compared: a local Lanczos chain stops when its residual
falls below 1e-15, which on product-like local problems is a decision taken at rounding level
(the evolved vectors agree either way)."""
import numpy as np
import pytest

from oracle import bug as obug
from oracle import models as om
from oracle import tdvp as otdvp
from tests.helpers import device_from_fam, fam_from_device

pytestmark = pytest.mark.gpu


def _upload(psi_o, charges):
    from bondupdategalerkin_amd.mps import SymMPS
    return SymMPS([device_from_fam(f) for f in psi_o.sites], charges, psi_o.center)


def _dense(psi):
    return om.mps_dense([fam_from_device(t) for t in psi.tensors])


@pytest.mark.parametrize("start,L,delta,tau,kw", [
    ("neel", 8, 0.5, -0.1j, dict(maxdim=64)),
    ("dw", 10, 0.0, -0.2j, dict(maxdim=6)),
    ("neel", 10, 1.0, -0.1j, dict(maxdim=32, weight_tol=1e-4)),
    ("neel", 8, 1.0, -0.2, dict(maxdim=16, eshift=-2.5, backward_report=True)),
    ("dw", 8, 0.5, -0.1j, dict(maxdim=16, hermitian=False)),
    ("neel", 12, 0.5, -0.05j, dict(maxdim=24, trunc_thresh=1e-6)),
])
def test_tdvp2_step_matches_oracle(gpu, start, L, delta, tau, kw):
    from bondupdategalerkin_amd import mps, tdvp
    from bondupdategalerkin_amd.mpo import xxz_mpo

    make_d = {"neel": mps.neel_state, "dw": mps.domain_wall_state}[start]
    make_o = {"neel": om.neel_state, "dw": om.domain_wall_state}[start]
    charges = make_d(L).charges
    psi_o = obug.MPS(make_o(L))
    mpo_d, mpo_o = xxz_mpo(L, delta=delta), om.xxz_mpo(L, delta=delta)
    print(f"case {start} L={L} tau={tau} {kw}", flush=True)
    for n in range(3):
        psi_d = _upload(psi_o, charges)
        info_d = tdvp.tdvp2_step(psi_d, mpo_d, tau, **kw)
        info_o = otdvp.tdvp2_step(psi_o, mpo_o, tau, **kw)
        assert psi_d.bond_dims() == psi_o.bond_dims(), n
        vd, vo = _dense(psi_d), om.mps_dense(psi_o.sites)
        assert np.linalg.norm(vd - vo) <= 1e-10 * np.linalg.norm(vo), n
        assert (info_d.max_bond, info_d.theta_elements, info_d.peak_elements) == (
            info_o.max_bond, info_o.theta_elements, info_o.peak_elements), n
        assert info_d.krylov_dims > 0 and info_o.krylov_dims > 0
        assert abs(info_d.discarded - info_o.discarded) <= 1e-10, n
        assert abs(info_d.cumulative - info_o.cumulative) <= 1e-10, n
        if kw.get("backward_report"):
            assert abs(info_d.backward_growth - info_o.backward_growth) <= 1e-9, n
        # imaginary time: renormalise both as a driver would
        if tau.real != 0:
            c = psi_o.center
            psi_o.sites[c] = psi_o.sites[c].scale(1.0 / psi_o.sites[c].norm())


def test_tdvp2_l32_real_time_conserves_energy_and_norm(gpu):
    """A longer device-only run: XXZ L = 32 from Neel, 10 steps of dt = 0.05 at maxdim 64 with a
    1e-10 relative cut.  The energy and norm stay within the discarded weight's reach of their
    start and the bond dimension grows from 1."""
    from bondupdategalerkin_amd import ite, mps, tdvp
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 32
    mpo = xxz_mpo(L, delta=0.5)
    psi = mps.neel_state(L)
    e0 = ite.energy(psi, mpo)
    for _ in range(10):
        info = tdvp.tdvp2_step(psi, mpo, -0.05j, maxdim=64, trunc_thresh=1e-10)
    assert abs(T.norm(psi[psi.center]) - 1.0) < 1e-8
    assert abs(ite.energy(psi, mpo) - e0) < 1e-6 * L
    assert max(psi.bond_dims()) > 4 and info.cumulative < 1e-6
