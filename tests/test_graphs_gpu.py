"""hipGraph capture of the effective-Hamiltonian actions: every replay equals the eager call bit
for bit (same plans and schedules), with eager work and synchronisations between replays; the
capture entry points refuse misuse (GPU)."""
import numpy as np
import pytest

from oracle import blocks as ob
from oracle import models as om
from tests.helpers import device_from_fam, fam_from_device, random_legs, rel_err

pytestmark = pytest.mark.gpu


def _setup(seed=3, L=6, dmax=90):
    from tests.test_contractions_gpu import _envs

    rng = np.random.default_rng(seed)
    legs = random_legs(rng, L, dmax, symmetric=True)
    psi = om.random_state(legs, rng, "U1")
    mpo_o = om.xxz_mpo(L, delta=0.5)
    lenvs, renvs = _envs(psi, mpo_o)
    return rng, psi, mpo_o, lenvs, renvs


def _rand_like(rng, f):
    g = f.copy()
    for k, b in g.blocks.items():
        g.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    return g


@pytest.mark.parametrize("prec", [0, 1])
def test_captured_one_site_apply_replays_equal_eager(gpu, prec):
    import torch

    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.graphs import CapturedApply
    from bondupdategalerkin_amd.mpo import xxz_mpo

    rng, psi, mpo_o, lenvs, renvs = _setup()
    mpo_d = xxz_mpo(len(psi), delta=0.5)
    site = 2
    Ld, Rd = device_from_fam(lenvs[site]), device_from_fam(renvs[site + 1])

    def action(x):
        return ev.apply_one_site(Ld, mpo_d[site], Rd, x, shift=0.25, precision=prec)

    cap = CapturedApply(action, device_from_fam(psi[site]))
    try:
        for r in range(4):
            v = _rand_like(rng, psi[site])
            vd = device_from_fam(v)
            got = cap(vd).data.clone()
            eager = action(vd).data
            # eager work and a synchronisation between replays
            torch.cuda.synchronize()
            assert torch.equal(got, eager), r
            ref = ob.apply_one_site(lenvs[site], mpo_o[site], renvs[site + 1], v, shift=0.25)
            tol = 1e-12 if prec == 0 else 2e-5
            assert rel_err(fam_from_device(cap.replay()), ref) <= tol, r
    finally:
        cap.close()


def test_captured_zero_site_apply_replays_equal_eager(gpu):
    import torch

    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.graphs import CapturedApply

    rng, psi, mpo_o, lenvs, renvs = _setup(seed=8)
    link = 3
    Ld, Rd = device_from_fam(lenvs[link]), device_from_fam(renvs[link])
    C = ob.Fam.zeros(psi[link].row, psi[link].row, [0])
    C = _rand_like(rng, C)

    def action(x):
        return ev.apply_zero_site(Ld, Rd, x, shift=-0.1)

    cap = CapturedApply(action, device_from_fam(C))
    try:
        for r in range(3):
            vd = device_from_fam(_rand_like(rng, C))
            got = cap(vd).data.clone()
            torch.cuda.synchronize()
            assert torch.equal(got, action(vd).data), r
    finally:
        cap.close()


def test_capture_entry_points_refuse_misuse(gpu):
    from bondupdategalerkin_amd import _lib

    ctx = _lib.context()
    lib = ctx._lib
    assert lib.bug_graph_capture_end(ctx.handle) == _lib.BUG_EINVAL
    assert lib.bug_graph_release(ctx.handle, None) == _lib.BUG_EINVAL
    assert lib.bug_graph_capture_begin(ctx.handle, None) == _lib.BUG_EINVAL
