"""The allocation-free one-site apply (bug_apply_one_site_noalloc) through the C ABI: its output
equals bug_apply_one_site's on two calls in a row, the second with a smaller input, in both
precisions and with a shift; an apply that needs more than the static workspace is refused before
anything runs (GPU)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _operands(L, D, seed, prec):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo
    from bondupdategalerkin_amd.mps import random_mps

    psi = random_mps(L, D, seed=seed)
    mid = L // 2
    B.canonical(psi, mid)
    mpo = xxz_mpo(L, delta=0.5)
    le = ev.left_env_stack(psi, mpo, upto=mid, precision=prec)
    re = ev.right_env_stack(psi, mpo, downto=mid + 1, precision=prec)
    return le[mid], mpo[mid], re[mid + 1], psi[mid]


def _apply(name, Lm, W, Rm, X, shift, prec):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import BlockTensor, current_stream_handle

    ctx = _lib.context()
    Y = BlockTensor.empty(X.layout)
    Y.data.fill_(float("nan"))           # every output element must be written
    fn = getattr(ctx._lib, name)
    rc = fn(ctx.handle, Lm.desc(), W.descriptor(), Rm.desc(), X.desc(), Y.desc(), shift, prec,
            current_stream_handle())
    ctx.check(rc)
    torch.cuda.synchronize()
    return Y.data.cpu().numpy()[:X.layout.nelem]


@pytest.mark.parametrize("prec_name", ["fp64", "fp32"])
def test_noalloc_apply_equals_the_allocating_one(gpu, prec_name):
    from bondupdategalerkin_amd import _lib

    prec = _lib.PREC_FP64 if prec_name == "fp64" else _lib.PREC_FP32
    for L, D, seed, shift in ((20, 256, 1, 0.0), (14, 48, 2, 0.3)):     # the second call smaller
        ops = _operands(L, D, seed, prec)
        ref = _apply("bug_apply_one_site", *ops, shift, prec)
        got = _apply("bug_apply_one_site_noalloc", *ops, shift, prec)
        assert np.isfinite(got).all()
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel <= (1e-14 if prec_name == "fp64" else 1e-6), (L, D, rel)
        # the same plans and schedules: expected equal bit for bit
        assert np.array_equal(got, ref), (L, D)


def test_noalloc_apply_refuses_beyond_its_workspace(gpu):
    """Zero operands of a D = 4096 centre site (the stage-1 intermediate alone is larger than the
    64 MiB workspace): BUG_EINVAL, and the output is left untouched."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, current_stream_handle
    from bondupdategalerkin_amd.environments import env_layout
    from bondupdategalerkin_amd.mpo import SPIN_CHARGES_U1, xxz_mpo
    from bondupdategalerkin_amd.mps import synthetic_bond_legs

    L, D = 26, 4096
    mid = L // 2
    legs = synthetic_bond_legs(L, D)
    W = xxz_mpo(L, delta=0.5)[mid]
    Lm = BlockTensor(env_layout(legs[mid], legs[mid], W.cl))
    Rm = BlockTensor(env_layout(legs[mid + 1], legs[mid + 1], W.cr))
    X = BlockTensor(Layout(legs[mid], legs[mid + 1], SPIN_CHARGES_U1))
    Y = BlockTensor(X.layout)
    Y.data.fill_(2.0)
    ctx = _lib.context()
    rc = ctx._lib.bug_apply_one_site_noalloc(ctx.handle, Lm.desc(), W.descriptor(), Rm.desc(),
                                             X.desc(), Y.desc(), 0.0, _lib.PREC_FP64,
                                             current_stream_handle())
    assert rc == _lib.BUG_EINVAL
    assert b"workspace" in ctx._lib.bug_last_error(ctx.handle)
    torch.cuda.synchronize()
    assert bool((Y.data == 2.0).all())
