"""The benchmark configurations at their own sizes on the device.

* XXZ (delta = 0.5) at L = 128 with bond dimensions up to D = 1024: the mixed-precision step
  against the FP64 drop-in from the same state (the synthetic state the benchmark times).
* The serving mix at its stated sizes: a full-size chain request (L = 128, maxdim 1024) and a
  2^11 x 2^11 Heston request give the same result alone and in the mix.
"""
import numpy as np
import pytest

from oracle import models as om
from tests.helpers import fam_from_device, mps_distance

pytestmark = pytest.mark.gpu


def test_l128_d1024_mixed_matches_fp64(gpu):
    """Two order-2 steps (parallel half-sweeps, tau = -0.02 i, maxdim = 1024, work ceiling
    rule maxdim) in each form from the same L = 128 state whose bonds reach D = 1024: the local
    magnetisation of the mixed form within 1e-6 of the FP64 drop-in (the stated error bound of the
    mixed form), the norms within 1e-8 (the FP32 products of the root exponential are unitary to
    their rounding, ~1e-7 relative, times |tau| per step), both runs at the maxdim cap, and the
    same verdicts except that the truncation verdict may differ where both runs discard within
    20 % of tol_trunc (the compression spends the whole truncation budget by design, so the
    verdict sits on its threshold).  Bond dimensions are not compared bond by bond: where the
    spectrum is flat at the cut, rounding at the 1e-7 level moves the cut by many states (93 on
    one bond here) at a discarded weight that differs only in the last digits."""
    import torch

    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L, D = 128, 1024
    tol_trunc = B.TOL_DEFAULTS["trunc"]
    psi = mps.random_mps(L, D, seed=0, tail=1e-9)
    B.canonical(psi, L - 1)
    T.scale_(1.0 / T.norm(psi[L - 1]), psi[L - 1])
    B.canonical(psi, 0)
    assert max(psi.bond_dims()) >= 1000
    mpo = xxz_mpo(L, delta=0.5)
    runs = {}
    for form in ("mixed", "fp64"):
        st = psi.copy()
        infos = []
        for _ in range(2):
            infos.append(B.cbe_bug_step(st, mpo, -0.02j, maxdim=D, work_ceiling_rule="maxdim",
                                        precision=form, order=2, parallel=True))
        torch.cuda.synchronize()
        runs[form] = (st, infos)
    (pm, im), (pf, iff) = runs["mixed"], runs["fp64"]
    dsz = np.abs(mps.magnetisation(pm) - mps.magnetisation(pf)).max()
    bm, bf = pm.bond_dims(), pf.bond_dims()
    nm, nf = T.norm(pm[pm.center]), T.norm(pf[pf.center])
    report = dict(dsz=dsz, dnorm=abs(nm - nf), bond_diff=max(abs(a - b) for a, b in zip(bm, bf)),
                  max_bond=(max(bm), max(bf)),
                  mixed=[(i.ledger.unmet, i.ledger.discarded) for i in im],
                  fp64=[(i.ledger.unmet, i.ledger.discarded) for i in iff])
    assert dsz <= 1e-6, report
    assert abs(nm - nf) <= 1e-8, report
    assert max(bm) == max(bf) == D, report
    for a, b in zip(im, iff):
        ua, ub = set(a.ledger.unmet), set(b.ledger.unmet)
        assert ua - {"trunc"} == ub - {"trunc"}, report
        if ("trunc" in ua) != ("trunc" in ub):
            for d in (a.ledger.discarded, b.ledger.discarded):
                assert 0.8 * tol_trunc <= d <= 1.2 * tol_trunc, report


def _full_size_mix():
    from bondupdategalerkin_amd.serving import Request

    reqs = [
        Request("chain", {"L": 128, "delta": 0.5, "start": "neel", "maxdim": 1024, "steps": 3,
                          "dt": 0.02}),
        Request("heston", {"bits": 11, "params": (0.03, 0.01, 2.0, 0.09, 0.4, -0.7),
                           "maxdim": 128, "steps": 1, "dt": 0.02}),
        Request("chain", {"L": 64, "delta": 1.0, "start": "dw", "maxdim": 256, "steps": 3,
                          "dt": 0.05}, precision="fp64"),
    ]
    for r in reqs:
        r.keep_state = True
    return reqs


def test_full_size_requests_alone_equal_in_the_mix(gpu):
    """A chain request at the mix's largest sizes (L = 128, maxdim 1024), a Heston request on the
    largest grid (2^11 x 2^11, maxdim 128) and a mid-size FP64 chain: each result alone equals the
    same request run concurrently with the others, to 1e-12 relative (the state distance for the
    chains without densifying, the 2^22 amplitudes for the Heston request)."""
    from bondupdategalerkin_amd import qtt as Q
    from bondupdategalerkin_amd.serving import Engine

    def run(slots):
        eng = Engine(slots=slots)
        try:
            if slots == 1:
                out = []
                for r in _full_size_mix():
                    out.append(eng.wait([eng.submit(r)], timeout=600)[0])
                return out
            rids = [eng.submit(r) for r in _full_size_mix()]
            return eng.wait(rids, timeout=600)
        finally:
            eng.close()

    alone, mixed = run(1), run(3)
    assert all(r.error is None for r in alone + mixed), [r.error for r in alone + mixed]
    spans = sorted((r.admitted, r.finished) for r in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for a, m in zip(alone, mixed):
        assert a.statuses == m.statuses, a.rid
        if a.kind == "chain":
            fa = [fam_from_device(t) for t in a.state.tensors]
            fm = [fam_from_device(t) for t in m.state.tensors]
            assert mps_distance(fm, fa) <= 1e-12, a.rid
        else:
            va, vm = Q.tt_dense(Q.qtt_cores(a.state)), Q.tt_dense(Q.qtt_cores(m.state))
            assert np.linalg.norm(va - vm) <= 1e-12 * np.linalg.norm(va), a.rid
        assert np.array_equal(a.observable, m.observable)
