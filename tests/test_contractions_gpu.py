"""GPU parity of the grouped block contractions against the oracle (FP64 and FP32 paths)."""
import numpy as np
import pytest

from oracle import blocks as ob
from oracle import models as om
from tests.helpers import device_from_fam, fam_from_device, oracle_mpo_site, random_legs, rel_err

pytestmark = pytest.mark.gpu

TOL64 = 1e-12   # FP64 MFMA vs numpy: reassociation only
TOL32 = 2e-5    # FP32 split-complex MFMA with FP32 accumulation


def _rand_fam(rng, row, col, shifts):
    f = ob.Fam.zeros(row, col, shifts)
    for k, b in f.blocks.items():
        f.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    return f


@pytest.mark.parametrize("opa,opb", [(a, b) for a in range(4) for b in range(4)])
def test_bsm_product_all_ops(gpu, opa, opb):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import Layout, Leg, bsm_product

    rng = np.random.default_rng(10 + 4 * opa + opb)
    # out (rows m, cols n), inner k; charges chosen so shifts line up
    m = {-2: 70, 0: 5, 2: 33}
    kk = {-1: 17, 1: 90, 3: 1}
    n = {0: 129, 2: 3, 4: 64}
    sp, sq = 1, 1          # op(P): m -> k (+1), op(Q): k -> n (+1)
    # stored P is op_a^{-1}: plain = (m x k, shift +1); transposed = (k x m, shift -1)
    P = _rand_fam(rng, m, kk, [sp]) if opa in (0, 3) else _rand_fam(rng, kk, m, [-sp])
    Q = _rand_fam(rng, kk, n, [sq]) if opb in (0, 3) else _rand_fam(rng, n, kk, [-sq])
    alpha = 0.7 - 0.3j
    Pd, Qd = device_from_fam(P), device_from_fam(Q)
    outl = Layout(Leg.make(m.items()), Leg.make(n.items()), [sp + sq])
    Y = bsm_product(outl, Pd, Qd, [(0, 0, 0, opa, opb, alpha)])
    got = fam_from_device(Y).dense()[0]

    def op(x, o):
        return [x, x.T, x.conj().T, x.conj()][o]

    ref = alpha * op(P.dense()[0], opa) @ op(Q.dense()[0], opb)
    assert np.abs(got - ref).max() <= TOL64 * np.abs(ref).max()
    # accumulate onto the previous result
    Y2 = bsm_product(outl, Pd, Qd, [(0, 0, 0, opa, opb, 1.0)], accumulate_into=Y)
    ref2 = ref + op(P.dense()[0], opa) @ op(Q.dense()[0], opb)
    got2 = fam_from_device(Y2).dense()[0]
    assert np.abs(got2 - ref2).max() <= TOL64 * np.abs(ref2).max()


def _chain(seed, L, dmax, sym):
    rng = np.random.default_rng(seed)
    legs = random_legs(rng, L, dmax, symmetric=(sym == "U1"))
    return om.random_state(legs, rng, sym), rng


def _envs(psi, mpo):
    L = len(psi)
    lenvs = [ob.env_fam({0: 1}, {0: 1}, mpo[0].cl)]
    lenvs[0].blocks[(0, 0)] = np.ones((1, 1), dtype=complex)
    for i in range(L):
        lenvs.append(ob.push_left(lenvs[-1], psi[i], mpo[i]))
    qtot = list(psi[-1].col)[0]
    renvs = [None] * (L + 1)
    renvs[L] = ob.env_fam({qtot: 1}, {qtot: 1}, mpo[-1].cr)
    renvs[L].blocks[(0, qtot)] = np.ones((1, 1), dtype=complex)
    for i in reversed(range(L)):
        renvs[i] = ob.push_right(renvs[i + 1], psi[i], mpo[i])
    return lenvs, renvs


@pytest.mark.parametrize("sym,dmax,prec", [("U1", 6, 0), ("U1", 150, 0), ("none", 64, 0),
                                           ("U1", 150, 1)])
def test_env_push_matches_oracle(gpu, sym, dmax, prec):
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 8
    psi, rng = _chain(7, L, dmax, sym)
    mpo_o = om.xxz_mpo(L, delta=0.5, symmetry=sym)
    mpo_d = xxz_mpo(L, delta=0.5, symmetry=sym)
    lenvs, renvs = _envs(psi, mpo_o)
    tol = TOL64 if prec == 0 else TOL32
    for i in range(L):
        Ld = device_from_fam(lenvs[i])
        A = device_from_fam(psi[i])
        got = fam_from_device(ev.push_left(Ld, A, mpo_d[i], precision=prec))
        assert rel_err(got, lenvs[i + 1]) <= tol, i
        Rd = device_from_fam(renvs[i + 1])
        got = fam_from_device(ev.push_right(Rd, A, mpo_d[i], precision=prec))
        assert rel_err(got, renvs[i]) <= tol, i


@pytest.mark.parametrize("sym,dmax,prec", [("U1", 9, 0), ("U1", 120, 0), ("none", 70, 0),
                                           ("U1", 120, 1)])
def test_one_and_zero_site_match_oracle(gpu, sym, dmax, prec):
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 8
    psi, rng = _chain(11, L, dmax, sym)
    mpo_o = om.xxz_mpo(L, delta=0.5, symmetry=sym)
    mpo_d = xxz_mpo(L, delta=0.5, symmetry=sym)
    lenvs, renvs = _envs(psi, mpo_o)
    tol = TOL64 if prec == 0 else TOL32
    for site in range(L):
        X = psi[site]
        ref = ob.apply_one_site(lenvs[site], mpo_o[site], renvs[site + 1], X, shift=0.3)
        got = ev.apply_one_site(device_from_fam(lenvs[site]), mpo_d[site],
                                device_from_fam(renvs[site + 1]), device_from_fam(X),
                                shift=0.3, precision=prec)
        assert rel_err(fam_from_device(got), ref) <= tol, site
    for link in range(1, L):
        C = _rand_fam(rng, psi[link].row, psi[link].row, [0])
        ref = ob.apply_zero_site(lenvs[link], renvs[link], C, shift=-0.2)
        got = ev.apply_zero_site(device_from_fam(lenvs[link]), device_from_fam(renvs[link]),
                                 device_from_fam(C), shift=-0.2, precision=prec)
        assert rel_err(fam_from_device(got), ref) <= tol, link


def test_every_tile_config_agrees(gpu):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 6
    psi, rng = _chain(5, L, 80, "U1")
    mpo_o = om.xxz_mpo(L, delta=1.0)
    mpo_d = xxz_mpo(L, delta=1.0)
    lenvs, renvs = _envs(psi, mpo_o)
    site = 3
    ref = ob.apply_one_site(lenvs[site], mpo_o[site], renvs[site + 1], psi[site])
    ctx = _lib.context()
    try:
        for cfg in range(_lib.lib().bug_tile_config_count()):
            ctx.set_tile_config(cfg)
            got = ev.apply_one_site(device_from_fam(lenvs[site]), mpo_d[site],
                                    device_from_fam(renvs[site + 1]), device_from_fam(psi[site]))
            assert rel_err(fam_from_device(got), ref) <= TOL64, cfg
    finally:
        ctx.set_tile_config(-1)


def test_flop_count_matches_block_sum(gpu):
    """The library's algorithmic flop count equals sum 8 m n k over the multiplied blocks."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import Layout, Leg, bsm_product

    rng = np.random.default_rng(3)
    m, kk, n = {0: 40, 2: 7}, {1: 13, 3: 50}, {2: 9, 4: 21}
    P = _rand_fam(rng, m, kk, [1])
    Q = _rand_fam(rng, kk, n, [1])
    outl = Layout(Leg.make(m.items()), Leg.make(n.items()), [2])
    bsm_product(outl, device_from_fam(P), device_from_fam(Q), [(0, 0, 0, 0, 0, 1.0)])
    expect = 8 * (40 * 13 * 9 + 7 * 50 * 21)
    assert _lib.context().last_flops() == expect


def test_invalid_layout_is_refused(gpu):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, Leg, bsm_product
    import torch

    lay = Layout(Leg.make([(0, 4)]), Leg.make([(0, 4)]), [0])
    small = BlockTensor(lay, torch.zeros(20, dtype=torch.complex128, device="cuda"))
    small.layout = lay
    # a buffer shorter than the layout must be refused by the host-side checks
    d = lay.descriptor(small.data.data_ptr(), nelem=8)
    ctx = _lib.context()
    arr = (_lib.BugTerm * 1)(_lib.BugTerm(0, 0, 0, 0, 0, 1.0, 0.0))
    rc = ctx._lib.bug_bsm_product(ctx.handle, d, d, d, 1, arr, 0, 0, None)
    assert rc == _lib.BUG_EINVARIANT


@pytest.mark.parametrize("prec", [0, 1])
def test_every_schedule_agrees(gpu, prec):
    """Every (tile configuration, split-k factor) schedule against the oracle on the one-site
    apply and on an accumulating product; unsplit configurations agree bit for bit, and each
    schedule is deterministic (a repeat is bitwise identical)."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.blocks import Layout, Leg, bsm_product
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 6
    psi, rng = _chain(7, L, 90, "U1")
    mpo_o = om.xxz_mpo(L, delta=1.0)
    mpo_d = xxz_mpo(L, delta=1.0)
    lenvs, renvs = _envs(psi, mpo_o)
    site = 3
    ref = ob.apply_one_site(lenvs[site], mpo_o[site], renvs[site + 1], psi[site])
    Ld, Rd, Xd = (device_from_fam(lenvs[site]), device_from_fam(renvs[site + 1]),
                  device_from_fam(psi[site]))
    m, kk, n = {0: 70, 2: 5}, {1: 200, 3: 33}, {2: 40, 4: 90}
    P, Q = _rand_fam(rng, m, kk, [1]), _rand_fam(rng, kk, n, [1])
    C0 = _rand_fam(rng, m, n, [2])
    outl = Layout(Leg.make(m.items()), Leg.make(n.items()), [2])
    Pd, Qd = device_from_fam(P), device_from_fam(Q)
    ref2 = C0.dense()[0] + (0.5 + 1j) * P.dense()[0] @ Q.dense()[0]
    tol = TOL64 if prec == 0 else TOL32
    ctx = _lib.context()
    unsplit = None
    try:
        for cfg in range(_lib.lib().bug_tile_config_count()):
            for splits in (1, 2, 3, 8):
                ctx.set_gemm_schedule(cfg, splits)
                runs = []
                for _ in range(2):
                    got = ev.apply_one_site(Ld, mpo_d[site], Rd, Xd, precision=prec)
                    acc = bsm_product(outl, Pd, Qd, [(0, 0, 0, 0, 0, 0.5 + 1j)],
                                      accumulate_into=device_from_fam(C0), precision=prec)
                    runs.append((got.data.cpu().numpy(), acc.data.cpu().numpy()))
                assert all(np.array_equal(a, b) for a, b in zip(runs[0], runs[1])), (cfg, splits)
                assert rel_err(fam_from_device(got), ref) <= tol, (cfg, splits)
                dense = fam_from_device(acc).dense()[0]
                assert np.abs(dense - ref2).max() <= tol * np.abs(ref2).max(), (cfg, splits)
                if splits == 1:
                    if unsplit is None:
                        unsplit = runs[0]
                    assert all(np.array_equal(a, b) for a, b in zip(unsplit, runs[0])), cfg
    finally:
        ctx.set_gemm_schedule(-1, 0)


def test_autotuner_table(gpu, tmp_path):
    """Tuning mode times the candidate schedules of each new plan class and records the winner;
    the table survives save / clear / load, and use mode applies it without timing."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 6
    psi, _ = _chain(9, L, 100, "U1")
    mpo_o = om.xxz_mpo(L, delta=0.5)
    mpo_d = xxz_mpo(L, delta=0.5)
    lenvs, renvs = _envs(psi, mpo_o)
    dev = [(device_from_fam(lenvs[s]), device_from_fam(renvs[s + 1]), device_from_fam(psi[s]))
           for s in range(L)]
    refs = [ob.apply_one_site(lenvs[s], mpo_o[s], renvs[s + 1], psi[s]) for s in range(L)]
    path = str(tmp_path / "schedules.txt")
    _lib.autotune_clear()
    try:
        _lib.set_autotune(_lib.AUTOTUNE_TUNE)
        for s in range(L):
            got = ev.apply_one_site(dev[s][0], mpo_d[s], dev[s][1], dev[s][2])
            assert rel_err(fam_from_device(got), refs[s]) <= TOL64, s
        entries, tuned = _lib.autotune_count()
        assert entries >= 1 and tuned == entries
        _lib.autotune_save(path)
        lines = [ln for ln in open(path) if not ln.startswith("#")]
        assert len(lines) == entries
        _lib.autotune_clear()
        assert _lib.autotune_count()[0] == 0
        _lib.autotune_load(path)
        assert _lib.autotune_count() == (entries, tuned)
        _lib.set_autotune(_lib.AUTOTUNE_USE)
        for s in range(L):
            got = ev.apply_one_site(dev[s][0], mpo_d[s], dev[s][1], dev[s][2])
            assert rel_err(fam_from_device(got), refs[s]) <= TOL64, s
        assert _lib.autotune_count() == (entries, tuned)      # use mode never times
        with open(path, "a") as f:
            f.write("zz 1\n")
        with pytest.raises(ValueError):
            _lib.autotune_load(path)
        with pytest.raises(ValueError):
            _lib.set_autotune(3)
    finally:
        _lib.set_autotune(_lib.AUTOTUNE_OFF)
        _lib.autotune_clear()


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_one_site_apply(gpu, prec):
    """Independent instances (different chains, sites, layouts and shifts) packed into one launch
    per stage: each result equals its single-instance apply to rounding and the oracle's."""
    from bondupdategalerkin_amd import environments as ev
    from bondupdategalerkin_amd.mpo import xxz_mpo

    items, refs, singles = [], [], []
    for seed, (L, dmax, site, delta, shift) in enumerate([(6, 60, 2, 0.5, 0.0), (8, 40, 5, 1.0, 0.3),
                                                          (5, 90, 1, 0.0, -0.2), (6, 25, 4, 0.5, 0.1)]):
        psi, _ = _chain(20 + seed, L, dmax, "U1")
        mpo_o = om.xxz_mpo(L, delta=delta)
        mpo_d = xxz_mpo(L, delta=delta)
        lenvs, renvs = _envs(psi, mpo_o)
        Ld, Rd, Xd = (device_from_fam(lenvs[site]), device_from_fam(renvs[site + 1]),
                      device_from_fam(psi[site]))
        items.append((Ld, mpo_d[site], Rd, Xd, shift))
        refs.append(ob.apply_one_site(lenvs[site], mpo_o[site], renvs[site + 1], psi[site],
                                      shift=shift))
        singles.append(ev.apply_one_site(Ld, mpo_d[site], Rd, Xd, shift=shift, precision=prec))
    got = ev.apply_one_site_batch(items, precision=prec)
    tol = TOL64 if prec == 0 else TOL32
    for i, (g, r, one) in enumerate(zip(got, refs, singles)):
        assert rel_err(fam_from_device(g), r) <= tol, i
        d = (g.data - one.data).abs().max().item()
        assert d <= (1e-13 if prec == 0 else 1e-5) * one.data.abs().max().item(), i
    assert ev.apply_one_site_batch([]) == []
    # the packed operator: persistent input/output buffers, the merged plan reused
    from bondupdategalerkin_amd import _lib

    packed = ev.PackedApply([it[:3] for it in items], [it[3].layout for it in items],
                            shifts=[it[4] for it in items], precision=prec)
    ctx = _lib.context()
    for rep in range(3):
        out = packed.apply([it[3] for it in items])
        before = ctx.stats()["misses"]
        for i, (o, one) in enumerate(zip(out, singles)):
            d = (o.data - one.data).abs().max().item()
            assert d <= (1e-13 if prec == 0 else 1e-5) * one.data.abs().max().item(), (rep, i)
    out = packed()
    assert ctx.stats()["misses"] == before          # no plan built on a repeated call
