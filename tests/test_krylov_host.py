"""The package's host-side Krylov pieces against the oracle's (CPU only): the direction weight
(the Simpson residual integral, computed a block of nodes at a time in the package) agrees with
the oracle's node-by-node recurrence to rounding, on Hermitian, non-normal and stiff cases."""
import numpy as np
import pytest

from bondupdategalerkin_amd import krylov as K
from oracle import krylov as ok


@pytest.mark.parametrize("scale", [0.1, 1.0, 30.0, 300.0, 3000.0])
def test_direction_weight_matches_oracle(scale):
    rng = np.random.default_rng(int(scale * 10))
    for trial in range(12):
        k = int(rng.integers(2, 40))
        H = np.triu(rng.standard_normal((k + 1, k)) + 1j * rng.standard_normal((k + 1, k)), -1)
        if trial % 3 == 0:                       # a Hermitian (Lanczos) block
            T = np.zeros((k + 1, k))
            a, b = rng.standard_normal(k), np.abs(rng.standard_normal(k))
            T[np.arange(k), np.arange(k)] = a
            T[np.arange(1, k + 1), np.arange(k)] = b
            T[np.arange(k - 1), np.arange(1, k)] = b[:-1]
            H = T.astype(complex)
        for tau, seed in ((-0.05j * scale, True), (0.02 * scale, True), (0.02 * scale, False)):
            got = K.direction_weight(H, k, tau, seed)
            want = ok.kry_weight(H, k, tau, seed)
            assert abs(got - want) <= 1e-12 * abs(want) + 1e-300, (trial, k, tau, seed)
