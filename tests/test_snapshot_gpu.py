"""State snapshots (GPU): a device state written to a file and read back is the same state bit for
bit, an evolution resumed from a snapshot equals the uninterrupted one bit for bit, and damaged
files are refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert len(a) == len(b) and a.center == b.center and a.charges == b.charges
    for x, y in zip(a.tensors, b.tensors):
        assert x.layout == y.layout
        n = x.layout.nelem
        assert torch.equal(x.data[:n], y.data[:n])


def test_round_trip_is_exact(gpu, tmp_path):
    from bondupdategalerkin_amd import mps, snapshot

    psi = mps.random_mps(12, 24, seed=3)
    psi.center = 5
    p = snapshot.save_state(str(tmp_path / "s.npz"), psi, {"t": 0.25, "model": "xxz"})
    back, meta = snapshot.load_state(p)
    _same(psi, back)
    assert meta == {"t": 0.25, "model": "xxz"}


def test_resumed_evolution_equals_the_uninterrupted_one(gpu, tmp_path):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import mps, snapshot
    from bondupdategalerkin_amd.mpo import xxz_mpo

    L = 16
    mpo = xxz_mpo(L, delta=0.5)
    run = mps.neel_state(L)
    for _ in range(4):
        B.cbe_bug_step(run, mpo, -0.05j, order=2, maxdim=32)
    half = mps.neel_state(L)
    for _ in range(2):
        B.cbe_bug_step(half, mpo, -0.05j, order=2, maxdim=32)
    p = snapshot.save_state(str(tmp_path / "half.npz"), half, {"steps": 2})
    resumed, meta = snapshot.load_state(p)
    assert meta["steps"] == 2
    for _ in range(2):
        B.cbe_bug_step(resumed, mpo, -0.05j, order=2, maxdim=32)
    _same(run, resumed)


def test_damaged_snapshots_are_refused(gpu, tmp_path):
    from bondupdategalerkin_amd import mps, snapshot

    psi = mps.random_mps(6, 8, seed=1)
    p = str(tmp_path / "s.npz")
    snapshot.save_state(p, psi)
    with np.load(p, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    bad = dict(arrays)
    bad["data2"] = bad["data2"][:-1]
    q = str(tmp_path / "short.npz")
    np.savez(q, **bad)
    with pytest.raises(ValueError):
        snapshot.load_state(q)
    bad = dict(arrays)
    bad["schema"] = np.frombuffer(b"other", dtype=np.uint8)
    np.savez(q, **bad)
    with pytest.raises(ValueError):
        snapshot.load_state(q)
    bad = dict(arrays)
    bad["col_d3"] = bad["col_d3"] + 1
    np.savez(q, **bad)
    with pytest.raises(ValueError):
        snapshot.load_state(q)
