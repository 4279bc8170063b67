"""The continuous-batching request loop: results in the mix equal results alone, admission by
footprint, and the latency bookkeeping (GPU)."""
import numpy as np
import pytest

from oracle import models as om
from tests.helpers import fam_from_device

pytestmark = pytest.mark.gpu


def _mix():
    from bondupdategalerkin_amd.serving import Request

    hp = (0.03, 0.01, 2.0, 0.09, 0.4, -0.7)
    reqs = [
        Request("chain", {"L": 12, "delta": 0.0, "start": "dw", "maxdim": 32, "steps": 3,
                          "dt": 0.05}),
        Request("heston", {"bits": 4, "params": hp, "maxdim": 16, "steps": 2, "dt": 0.02}),
        Request("chain", {"L": 10, "delta": 0.5, "start": "neel", "maxdim": 16, "steps": 3,
                          "dt": 0.05}, precision="fp64"),
        Request("chain", {"L": 14, "delta": 1.0, "start": "neel", "maxdim": 32, "steps": 2,
                          "dt": 0.05}, order=1),
        Request("heston", {"bits": 5, "params": (0.02, 0.0, 1.0, 0.1, 0.3, -0.5), "maxdim": 32,
                           "steps": 2, "dt": 0.02}, precision="fp64"),
    ]
    for r in reqs:
        r.keep_state = True
    return reqs


def _dense(res):
    from bondupdategalerkin_amd import qtt as Q

    psi = res.state
    if res.kind == "chain":
        return om.mps_dense([fam_from_device(t) for t in psi.tensors])
    return Q.tt_dense(Q.qtt_cores(psi))


def test_results_alone_equal_results_in_the_mix(gpu):
    from bondupdategalerkin_amd.serving import Engine, summarize

    alone = []
    eng = Engine(slots=1)
    try:
        for r in _mix():
            rid = eng.submit(r)
            alone.append(eng.wait([rid], timeout=300)[0])
    finally:
        eng.close()
    eng = Engine(slots=3)
    try:
        rids = [eng.submit(r) for r in _mix()]
        mixed = eng.wait(rids, timeout=300)
    finally:
        eng.close()
    assert all(r.error is None for r in alone + mixed), [r.error for r in alone + mixed]
    assert [r.precision for r in mixed] == ["mixed", "fp64", "fp64", "mixed", "fp64"]
    # the mix really overlapped: some request was admitted before another one finished
    spans = sorted((r.admitted, r.finished) for r in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for a, m in zip(alone, mixed):
        assert a.statuses == m.statuses
        va, vm = _dense(a), _dense(m)
        # the stated tolerance: 1e-12 relative (the kernels are deterministic, so expect 0)
        assert np.linalg.norm(va - vm) <= 1e-12 * np.linalg.norm(va), a.rid
        assert np.array_equal(a.observable, m.observable)
    s = summarize(mixed, max(r.finished for r in mixed) - min(r.arrival for r in mixed))
    assert s["requests"] == 5 and s["errors"] == 0 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_admission_holds_requests_beyond_the_budget(gpu):
    from bondupdategalerkin_amd.serving import Engine, Request, footprint_bytes

    reqs = [Request("chain", {"L": 10, "delta": 0.0, "start": "dw", "maxdim": 16, "steps": 2,
                              "dt": 0.05}) for _ in range(3)]
    fp = footprint_bytes(reqs[0])
    eng = Engine(slots=3, hbm_budget=int(1.5 * fp))       # room for one request at a time
    try:
        rids = [eng.submit(r) for r in reqs]
        res = eng.wait(rids, timeout=300)
        with pytest.raises(ValueError):
            eng.submit(Request("chain", {"L": 64, "delta": 0.0, "start": "dw", "maxdim": 1024,
                                         "steps": 1, "dt": 0.05}))
    finally:
        eng.close()
    spans = sorted((r.admitted, r.finished) for r in res)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))
    assert all(r.error is None and r.statuses for r in res)


def test_c_and_python_drivers_agree(gpu):
    """A request stepped through the library's ``bug_step`` (the default driver) ends where the
    package's ``cbe_bug_step`` ends: the same statuses, states within 1e-10 (both orchestrate
    the same kernels; their host-side eigen- and exponential solvers round differently)."""
    from bondupdategalerkin_amd.serving import run_request

    for r in _mix()[:3]:
        outs = []
        for drv in ("c", "python"):
            r.driver = drv
            psi, statuses, obs = run_request(r)
            res = type("R", (), {"state": psi, "kind": r.kind})
            outs.append((statuses, _dense(res), obs))
        (sa, va, oa), (sb, vb, ob) = outs
        assert sa == sb, r.kind
        assert np.linalg.norm(va - vb) <= 1e-10 * np.linalg.norm(vb), r.kind
        assert np.allclose(oa, ob, atol=1e-10), r.kind
