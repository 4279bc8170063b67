"""GPU parity of the non-Hermitian path: the device Arnoldi root (closed forms) and whole CBE-BUG
steps on the Heston generator against the oracle's steps and the dense exponential.

The bar is that of the Hermitian step: relative state difference <= 1e-10 per step on identical
inputs and probe seeds, identical bond dimensions and budget verdicts; against the dense
exponential the reference's own 1e-5 (test/QTTS/Models/Heston/test_heston.jl:90-106)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import blocks as ob
from oracle import bug as obug
from oracle import heston as oh
from tests.helpers import device_from_fam, fam_from_device

pytestmark = pytest.mark.gpu

HP = (0.03, 0.01, 2.0, 0.09, 0.4, -0.7)


def _params():
    from bondupdategalerkin_amd import qtt as Q

    return Q.HestonParams(*HP)


def _grid(n):
    from bondupdategalerkin_amd import qtt as Q

    return Q.qtt_grid(n, n, x=(-1.0, 1.0), y=(0.0, 0.5), periodic_x=False, periodic_y=False)


def _sector_core(seed=4):
    rng = np.random.default_rng(seed)
    row = {-3: 2, -1: 3, 1: 1, 3: 4}
    f = ob.Fam.zeros(row, row, [0])
    for k in f.blocks:
        f.blocks[k] = rng.standard_normal(f.blocks[k].shape) + 1j * rng.standard_normal(
            f.blocks[k].shape)
    return f, row


@pytest.mark.parametrize("complex_scale", [False, True])
@pytest.mark.parametrize("tau", [-0.4 + 0j, -0.7j])
def test_device_arnoldi_sector_closed_form(gpu, tau, complex_scale):
    """A generator that scales each charge sector by its own number (complex: non-Hermitian):
    exp(tau G) x in closed form, reached at the true Krylov dimension (test_krylov.jl:81-87,
    135-173)."""
    from bondupdategalerkin_amd import krylov as K
    from bondupdategalerkin_amd import tensor_ops as T

    f, row = _sector_core()
    scale = {q: 0.5 * q + 1.0 for q in row}
    c = (1 + 0.3j) if complex_scale else 1.0
    x = device_from_fam(f)
    S = {q: np.full(d, scale[q]) for q, d in row.items()}

    def apply(v):
        return T.scale(c, T.diag_scale(v, S, "rows"))

    y, rep = K.arnoldi_expv(apply, tau, x, maxiter=60, conv_tol=0.0, substeps=1)
    got = fam_from_device(y)
    for (k, q), b in f.blocks.items():
        assert np.abs(got.blocks[(k, q)] - np.exp(tau * c * scale[q]) * b).max() <= 1e-11
    assert rep.iterations == len(row)


def test_device_arnoldi_matches_oracle_on_heston_root(gpu):
    """The device Arnoldi exponential and the oracle's on the same non-normal zero-site action
    (the Heston generator's root core), with the substep search on."""
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd import krylov as K
    from bondupdategalerkin_amd import qtt as Q
    from oracle import krylov as ok

    g = _grid(3)
    W = Q.heston_op(g, _params())
    mpo_d = Q.qtt_to_mpo(W)
    rng = np.random.default_rng(1)
    cores = [rng.standard_normal((a, 2, b)) for a, b in
             zip([1, 2, 4, 6, 4, 2], [2, 4, 6, 4, 2, 1])]
    psi = Q.qtt_mps(cores, center=3)
    from bondupdategalerkin_amd import tensor_ops as T

    # the root sits on the bond between sites 2 and 3
    C, B3 = T.qr_right(psi[3])
    L = ev.boundary_left(mpo_d, psi.bond_leg(0))
    for i in range(3):
        L = ev.push_left(L, psi[i], mpo_d[i])
    R = ev.boundary_right(mpo_d, psi.bond_leg(len(psi)))
    for i in (5, 4):
        R = ev.push_right(R, psi[i], mpo_d[i])
    R = ev.push_right(R, B3, mpo_d[3])
    x = T.scale(1.0 / T.norm(C), C)
    Lf, Rf, xf = fam_from_device(L), fam_from_device(R), fam_from_device(x)
    dev, rep = K.arnoldi_expv(lambda v: ev.apply_zero_site(L, R, v), 0.3, x, maxiter=30,
                              conv_tol=1e-10, substeps=0)
    ref, rep_o = ok.arnoldi_expv(lambda v: ob.apply_zero_site(Lf, Rf, v), 0.3, xf,
                                 inner=ob.inner, norm=lambda v: v.norm(),
                                 axpy=lambda a, u, v: v.axpy(a, u), scale=lambda a, v: v.scale(a),
                                 maxiter=30, tol=1e-15, conv_tol=1e-10, substeps=0)
    assert (rep.iterations, rep.substeps) == (rep_o.iterations, rep_o.substeps)
    got = fam_from_device(dev)
    num = sum(np.linalg.norm(got.blocks[k] - ref.blocks[k]) ** 2 for k in ref.blocks)
    assert np.sqrt(num) <= 1e-11 * ref.norm()


def _heston_inputs(n, payoff):
    from bondupdategalerkin_amd import qtt as Q

    g = _grid(n)
    if payoff:
        sp = lambda z: z if z > 30 else np.log1p(np.exp(z))                 # noqa: E731
        init = Q.qtt_separable(g, lambda x: 0.02 * sp((np.exp(x) - 1) / 0.02), None)
    else:
        init = Q.qtt_separable(g, lambda x: np.exp(0.5 * x), lambda v: np.exp(-2 * v) * (1 + v))
    W = Q.heston_op(g, _params())
    return g, init, W


@pytest.mark.parametrize("n,payoff,order,steps,tb", [(4, False, 1, 5, 1e-8),
                                                      (4, False, 2, 3, 1e-8),
                                                      (4, True, 1, 4, 1e-3),
                                                      (5, False, 3, 2, 1e-8)])
def test_heston_run_matches_oracle(gpu, n, payoff, order, steps, tb):
    """Multi-step device runs against the oracle's from the same input at 1e-10, and against the
    dense exponential at the reference's 1e-5 (test_heston.jl:90-106)."""
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import qtt as Q

    g, init, W = _heston_inputs(n, payoff)
    psi_d = Q.qtt_mps(init)
    mpo_d = Q.qtt_to_mpo(W)
    psi_o = obug.MPS(oh.fams_from_cores(init), center=len(init) - 1)
    obug.canonical(psi_o, 0)
    mpo_o = oh.mpo_from_cores(W)
    opts = dict(tol_basis=tb, tol_trunc=1e-12, maxdim=64, hermitian=False, order=order)
    for k in range(steps):
        info_d = B.cbe_bug_step(psi_d, mpo_d, 0.02, **opts)
        info_o = obug.cbe_bug_step(psi_o, mpo_o, 0.02, obug.StepOptions(**opts))
        assert psi_d.bond_dims() == psi_o.bond_dims(), k
        assert B.budget_status(info_d.ledger) == info_o.ledger.status, k
        fd = [fam_from_device(t) for t in psi_d.tensors]
        assert _tt_distance(fd, psi_o.sites) <= 1e-10, k
    if tb <= 1e-8:
        vo = oh.mps_grid_vector(psi_o.sites)
        exact = sla.expm(steps * 0.02 * oh.heston_dense(n, n, HP)) @ oh.tt_grid_vector(init)
        assert np.linalg.norm(vo - exact) / np.linalg.norm(exact) < 1e-5


@pytest.mark.parametrize("n,payoff,order,steps", [(6, True, 2, 2), (10, False, 1, 2),
                                                  (10, False, 2, 1)])
def test_stiff_heston_steps_match_oracle(gpu, n, payoff, order, steps):
    """Single steps from identical inputs on stiff grids (||tau A|| ~ 1e3 at 10 + 10 bits).
    There the step is conditioned far above roundoff: the oracle itself moves by ~1e-8 when
    its input is perturbed at 1e-15, and trajectories drift apart by more every step.  So each
    step starts the device from the oracle's current state, the floor is measured on a twin
    oracle step from the perturbed state, and the device is held to 4x it (1e-10 at least)."""
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd.mps import SymMPS
    from bondupdategalerkin_amd.mpo import SPIN_CHARGES_NONE

    g, init, W = _heston_inputs(n, payoff)
    mpo_d = Q.qtt_to_mpo(W)
    mpo_o = oh.mpo_from_cores(W)
    psi_o = obug.MPS(oh.fams_from_cores(init), center=len(init) - 1)
    obug.canonical(psi_o, 0)
    rng = np.random.default_rng(5)
    opts = dict(tol_basis=1e-8, tol_trunc=1e-12, maxdim=64, hermitian=False, order=order)
    for k in range(steps):
        psi_d = SymMPS([device_from_fam(f) for f in psi_o.sites], SPIN_CHARGES_NONE,
                       center=psi_o.center)
        twin = obug.MPS([f.copy() for f in psi_o.sites], psi_o.center)
        for f in twin.sites:
            for key, b in f.blocks.items():
                f.blocks[key] = b * (1 + 1e-15 * rng.standard_normal(b.shape))
        B.cbe_bug_step(psi_d, mpo_d, 0.02, **opts)
        obug.cbe_bug_step(twin, mpo_o, 0.02, obug.StepOptions(**opts))
        obug.cbe_bug_step(psi_o, mpo_o, 0.02, obug.StepOptions(**opts))
        floor = _tt_distance(twin.sites, psi_o.sites)
        gap = _tt_distance([fam_from_device(t) for t in psi_d.tensors], psi_o.sites)
        print(f"step {k}: device-oracle {gap:.2e}, oracle conditioning floor {floor:.2e}",
              flush=True)
        assert gap <= max(1e-10, 4 * floor), k


def _tt_distance(a, b):
    """||a - b|| / ||b|| of two no-symmetry MPS without densifying: the difference as one TT
    (direct-sum cores) swept left to right by QR, so no cancellation between overlaps."""
    def cores(fams, sign):
        out = []
        for f in fams:
            c = np.stack([f.blocks.get((s, 0)) for s in range(2)], axis=1)
            out.append(c)
        out[0] = out[0] * sign
        return out

    ca, cb = cores(a, 1.0), cores(b, -1.0)
    n = len(ca)
    diff = []
    for i, (x, y) in enumerate(zip(ca, cb)):
        if i == 0:
            diff.append(np.concatenate([x, y], axis=2))
        elif i == n - 1:
            diff.append(np.concatenate([x, y], axis=0))
        else:
            z = np.zeros((x.shape[0] + y.shape[0], 2, x.shape[2] + y.shape[2]), dtype=complex)
            z[:x.shape[0], :, :x.shape[2]] = x
            z[x.shape[0]:, :, x.shape[2]:] = y
            diff.append(z)

    def tt_norm(cs):
        carry = np.ones((1, 1), dtype=complex)
        for c in cs:
            m = np.einsum("ab,bsc->asc", carry, c)
            carry = np.linalg.qr(m.reshape(-1, m.shape[2]), mode="r")
        return float(np.linalg.norm(carry))

    return tt_norm(diff) / tt_norm(cores(b, 1.0))


def test_qtt_mps_roundtrip_and_gauge(gpu):
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd import tensor_ops as T

    g, init, _ = _heston_inputs(5, True)
    want = Q.tt_dense(init)
    for center in (0, 4, 9):
        psi = Q.qtt_mps(init, center=center)
        assert psi.center == center
        assert np.linalg.norm(Q.tt_dense(Q.qtt_cores(psi)) - want) <= 1e-12 * np.linalg.norm(want)
        assert abs(T.norm(psi[center]) - np.linalg.norm(want)) <= 1e-12 * np.linalg.norm(want)
