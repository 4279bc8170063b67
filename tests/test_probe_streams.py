"""The product's probe stream and the oracle's are the same numbers (CPU only)."""
import numpy as np

from bondupdategalerkin_amd.probes import ProbeStream
from oracle.linalg import ProbeRNG


def test_streams_identical_with_skips():
    a, b = ProbeStream(0x5EED), ProbeRNG(0x5EED)
    for n, skip in ((5, 0), (17, 3), (1, 100), (64, 7)):
        assert np.array_equal(a.draw(n), b.complex_normal(n))
        a.skip(skip)
        b.skip(skip)
    # golden vector: first numbers of seed 1 (pins the generator across refactors)
    z = ProbeStream(1).draw(3)
    assert np.allclose(z, ProbeRNG(1).complex_normal(3))
    assert np.all(np.isfinite(z)) and np.all(np.abs(z) < 10)


def test_split_streams_identical_and_independent():
    a, b = ProbeStream(0x5EED), ProbeRNG(0x5EED)
    a.draw(11)
    b.complex_normal(11)
    ra, rb = a.split(), b.split()
    assert a.position == b.pos == 12                   # the split consumes one number
    assert np.array_equal(ra.draw(40), rb.complex_normal(40))
    assert np.array_equal(a.draw(9), b.complex_normal(9))
    # the split stream is not a shifted copy of its parent
    c = ProbeStream(0x5EED)
    parent = c.draw(200)
    child = ProbeStream(0x5EED)
    child.skip(11)
    child = child.split().draw(40)
    assert np.abs(np.subtract.outer(child, parent)).min() > 1e-12


def test_parallel_oracle_step_uses_split_stream():
    """parallel = True draws the right half-sweep's probes from the split stream, so the step
    differs from the serial one only at the basis tolerance (bug_step.jl:130-134)."""
    from oracle import bug as obug
    from oracle import models as om

    out, status = [], []
    for par in (False, True):
        psi = obug.MPS(om.domain_wall_state(12))
        for _ in range(2):
            info = obug.cbe_bug_step(psi, om.xxz_mpo(12, delta=0.0), -0.05j,
                                     obug.StepOptions(parallel=par, work_ceiling_rule="maxdim"))
        out.append(om.mps_dense(psi.sites))
        status.append(info.ledger.status)
    d = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[0])
    assert d < 1e-4 and status[0] == status[1] == "met"


def test_large_draws_are_chunk_independent():
    """A draw large enough to be evaluated in chunks by the thread pool equals the oracle's
    stream bit for bit, and so does every chunk-sized piece drawn alone."""
    from bondupdategalerkin_amd import probes

    n = 5 * probes._CHUNK + 123
    dev = probes.ProbeStream(77)
    ref = ProbeRNG(77)
    got = dev.draw(n)
    want = ref.complex_normal(n)
    assert np.array_equal(got, want)
    idx = np.arange(n, dtype=np.uint64)[::-1].copy()
    assert np.array_equal(dev.at(idx), want[::-1])
    small = probes.ProbeStream(77)
    parts = [small.draw(1000) for _ in range(n // 1000)]
    assert np.array_equal(np.concatenate(parts), want[: 1000 * (n // 1000)])
