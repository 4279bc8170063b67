"""Host logic of the request loop: the footprint estimate and the seeded arrival mix (CPU)."""
import numpy as np

from bondupdategalerkin_amd.serving import Request, footprint_bytes, poisson_mix


def test_footprint_grows_with_length_and_maxdim():
    base = {"delta": 0.0, "start": "dw", "steps": 1, "dt": 0.05}
    f = [footprint_bytes(Request("chain", dict(base, L=L, maxdim=D)))
         for L, D in ((32, 64), (64, 64), (64, 256), (128, 1024))]
    assert f == sorted(f) and len(set(f)) == 4
    # the bond can never exceed 2^(L/2): a short chain's estimate ignores a large maxdim
    assert footprint_bytes(Request("chain", dict(base, L=8, maxdim=1024))) == \
        footprint_bytes(Request("chain", dict(base, L=8, maxdim=16)))
    h = footprint_bytes(Request("heston", {"bits": 10, "params": (0,) * 6, "maxdim": 128,
                                           "steps": 1, "dt": 0.02}))
    assert h < footprint_bytes(Request("chain", dict(base, L=128, maxdim=1024)))


def test_poisson_mix_is_seeded_and_mixed():
    a, b = poisson_mix(200, rate=2.0, seed=3), poisson_mix(200, rate=2.0, seed=3)
    assert [t for t, _ in a] == [t for t, _ in b]
    assert [r.params for _, r in a] == [r.params for _, r in b]
    t = np.array([x for x, _ in a])
    assert np.all(np.diff(t) > 0)
    assert abs(np.diff(t).mean() - 0.5) < 0.1                      # rate 2 per second
    frac = np.mean([r.kind == "chain" for _, r in a])
    assert 0.5 < frac < 0.7
    full = poisson_mix(50, rate=1.0, seed=1, scale="full")
    assert all(r.params["L"] in (32, 64, 96, 128) for _, r in full if r.kind == "chain")
    assert all(8 <= r.params["bits"] <= 11 for _, r in full if r.kind == "heston")


def test_precision_policy():
    """Chains run in the requested form; Heston requests always run FP64 (CPU only)."""
    import pytest

    from bondupdategalerkin_amd.serving import Request, precision_used

    chain = {"L": 12, "delta": 0.0, "start": "dw", "maxdim": 16, "steps": 1, "dt": 0.05}
    heston = {"bits": 4, "params": (0.03, 0.0, 1.0, 0.1, 0.3, -0.5), "maxdim": 8, "steps": 1,
              "dt": 0.02}
    assert precision_used(Request("chain", chain, precision="mixed")) == "mixed"
    assert precision_used(Request("chain", chain, precision="fp64")) == "fp64"
    assert precision_used(Request("heston", heston, precision="mixed")) == "fp64"
    with pytest.raises(ValueError):
        precision_used(Request("chain", chain, precision="fp16"))
