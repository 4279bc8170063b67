"""Many independent instances in lockstep (GPU): the segmented vector operations against their
per-instance forms, and the packed Lanczos and Arnoldi exponentials against one lanczos_expv /
arnoldi_expv per instance."""
import numpy as np
import pytest

from oracle import models as om
from tests.helpers import device_from_fam, random_legs

pytestmark = pytest.mark.gpu


def _instances(specs):
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from tests.test_contractions_gpu import _envs

    ops, xs = [], []
    for seed, (L, dmax, site, delta) in enumerate(specs):
        rng = np.random.default_rng(40 + seed)
        legs = random_legs(rng, L, dmax, symmetric=True)
        psi = om.random_state(legs, rng, "U1")
        mpo_o = om.xxz_mpo(L, delta=delta)
        lenvs, renvs = _envs(psi, mpo_o)
        ops.append((device_from_fam(lenvs[site]), xxz_mpo(L, delta=delta)[site],
                    device_from_fam(renvs[site + 1])))
        xs.append(device_from_fam(psi[site]))
    return ops, xs


SPECS = [(6, 50, 2, 0.5), (8, 30, 4, 1.0), (5, 70, 1, 0.0), (6, 20, 3, 0.5)]


def test_segmented_dots_and_axpys(gpu):
    import torch

    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.environments import PackedApply
    from bondupdategalerkin_amd.packed import Segments

    ops, xs = _instances(SPECS)
    packed = PackedApply(ops, [x.layout for x in xs])
    for t, x in zip(packed.x, xs):
        t.data.copy_(x.data)
    seg = Segments(packed)
    a = packed._xbuf.clone()
    b = torch.randn_like(packed._xbuf)
    d = seg.dots([a, b], b)
    for i, t in enumerate(packed.x):
        off = (t.data.data_ptr() - packed._xbuf.data_ptr()) // 16
        sl = slice(off, off + t.layout.nelem)
        for j, v in enumerate([a, b]):
            want = complex(torch.vdot(v[sl], b[sl]).item())
            assert abs(d[j, i] - want) <= 1e-12 * abs(want), (i, j)
    alphas = np.array([[0.5 + 1j] * 4, [-2.0] * 4])
    y = b.clone()
    seg.axpy(alphas, [a, b], y)
    want = b + (0.5 + 1j) * a - 2.0 * b
    mask = torch.zeros_like(y, dtype=torch.bool)
    for t in packed.x:
        off = (t.data.data_ptr() - packed._xbuf.data_ptr()) // 16
        mask[off: off + t.layout.nelem] = True
    assert torch.allclose(y[mask], want[mask], rtol=0, atol=1e-12)
    assert torch.equal(y[~mask], b[~mask])          # padding untouched
    del T


@pytest.mark.parametrize("tau", [-0.3j, -0.05 + 0j])
def test_packed_lanczos_matches_per_instance(gpu, tau):
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd import krylov as K
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.environments import PackedApply
    from bondupdategalerkin_amd.packed import lanczos_expv_packed

    ops, xs = _instances(SPECS)
    packed = PackedApply(ops, [x.layout for x in xs])
    # the random environments are not normalised: scale tau so that |tau| ||H_i x_i|| / ||x_i||
    # stays below 1 for every instance, where a 12-vector Krylov space resolves the exponential
    gain = max(T.norm(ev.apply_one_site(L, W, R, x)) / T.norm(x) for (L, W, R), x in zip(ops, xs))
    tau = tau / gain
    got, dims = lanczos_expv_packed(packed, tau, xs, maxiter=12)
    for i, ((L, W, R), x) in enumerate(zip(ops, xs)):
        want, rep = K.lanczos_expv(lambda v, L=L, W=W, R=R: ev.apply_one_site(L, W, R, v), tau, x,
                                   maxiter=12, substeps=1, conv_tol=0.0)
        assert dims[i] == rep.iterations, i
        diff = T.norm(T.axpy(-1.0, want, got[i]))
        assert diff <= 1e-11 * T.norm(want), (i, diff)
        if tau.real == 0:                             # unitary: the norm is kept
            assert abs(T.norm(got[i]) - T.norm(x)) <= 1e-10 * T.norm(x)


@pytest.mark.parametrize("tau", [-0.3j, -0.2 + 0.1j])
def test_packed_arnoldi_matches_per_instance(gpu, tau):
    """The lockstep Arnoldi (a shifted, hence non-Hermitian-looking, generator: tau complex) equals
    one arnoldi_expv with substeps = 1 per instance to 1e-11, with the same Krylov dimensions."""
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd import krylov as K
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.environments import PackedApply
    from bondupdategalerkin_amd.packed import arnoldi_expv_packed

    ops, xs = _instances(SPECS)
    packed = PackedApply(ops, [x.layout for x in xs])
    gain = max(T.norm(ev.apply_one_site(L, W, R, x)) / T.norm(x) for (L, W, R), x in zip(ops, xs))
    tau = tau / gain
    got, dims = arnoldi_expv_packed(packed, tau, xs, maxiter=12)
    for i, ((L, W, R), x) in enumerate(zip(ops, xs)):
        want, rep = K.arnoldi_expv(lambda v, L=L, W=W, R=R: ev.apply_one_site(L, W, R, v), tau, x,
                                   maxiter=12, substeps=1, conv_tol=0.0)
        assert dims[i] == rep.iterations, i
        diff = T.norm(T.axpy(-1.0, want, got[i]))
        assert diff <= 1e-11 * T.norm(want), (i, diff)
