"""Pin the oracle's block factorizations, selection rule and probe RNG (CPU only)."""
import numpy as np
import pytest

from oracle import blocks as ob
from oracle import linalg as la
from tests.helpers import random_legs
from oracle import models as om


def _rand_site(seed=0, dmax=6, L=6, site=3):
    rng = np.random.default_rng(seed)
    legs = random_legs(rng, L, dmax)
    return om.random_state(legs, rng)[site]


def test_product_matches_dense():
    rng = np.random.default_rng(1)
    P = ob.Fam.zeros({0: 3, 2: 4}, {1: 5, 3: 2}, [1])
    Q = ob.Fam.zeros({1: 5, 3: 2}, {2: 6, 4: 1}, [1])
    for f in (P, Q):
        for k, b in f.blocks.items():
            f.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    out = la.product(P.row, Q.col, [2], P, Q, [(0, 0, 0, la.OP_N, la.OP_N, 0.5j)])
    assert np.allclose(out.dense()[0], 0.5j * P.dense()[0] @ Q.dense()[0])


def test_qr_left_right_reconstruct_and_isometric():
    A = _rand_site(2)
    Q, R = la.qr_left(A)
    assert np.allclose(la.core_times_site(R, ob.Fam(A.row, A.col, A.shifts)).dense().shape, A.dense().shape)
    rec = la.site_times_core(Q, R)
    assert np.allclose(rec.dense(), A.dense(), atol=1e-13)
    G = la.overlap_left(Q, Q)
    assert np.allclose(G.dense()[0], np.eye(G.dense().shape[1]), atol=1e-13)
    for (k, q), r in R.blocks.items():
        d = np.diag(r)
        assert np.all(np.abs(d.imag) < 1e-14) and np.all(d.real >= 0)
    Lc, Qr = la.qr_right(A)
    rec = la.core_times_site(Lc, Qr)
    assert np.allclose(rec.dense(), A.dense(), atol=1e-13)
    G = la.overlap_right(Qr, Qr)
    assert np.allclose(G.dense()[0], np.eye(G.dense().shape[1]), atol=1e-13)


def test_svd_reconstructs_and_selects_globally():
    A = _rand_site(3, dmax=5)
    r = la.svd_left(A)
    rec = la.site_times_core(r.U, la.diag_times(r.S, r.Vh, "left"))
    assert np.allclose(rec.dense(), A.dense(), atol=1e-12)
    allsv = sorted([s for v in r.S.values() for s in v], reverse=True)
    k = 3
    rk = la.svd_left(A, nkeep=k)
    kept = sorted([s for v in rk.S.values() for s in v], reverse=True)
    assert np.allclose(kept, allsv[:k])
    assert [e[0] for e in rk.kept] == kept and len(rk.trunc) == len(allsv) - k
    rc = la.svd_left(A, cutoff=0.5)
    assert all(s > 0.5 * allsv[0] for v in rc.S.values() for s in v)
    # phase convention: first entry with |u| >= 0.5 max|u| is real positive
    for (s, q), b in r.U.blocks.items():
        pass
    for qr in r.U.col:
        parts = la.left_sectors(r.U)[qr]
        M = la.left_matrix(r.U, qr, parts)
        for j in range(M.shape[1]):
            a = np.abs(M[:, j])
            i = int(np.argmax(a >= 0.5 * a.max()))
            assert abs(M[i, j].imag) < 1e-14 and M[i, j].real > 0


def test_selection_rule_on_hand_spectra():
    keep, kept, trunc = la._select({0: np.array([3.0, 1.0, 0.0]), 2: np.array([2.0, 1.0])}, 0.0, None)
    assert keep == {0: 2, 2: 2} and len(trunc) == 1          # the exact zero is dropped (sv > 0)
    keep, kept, trunc = la._select({0: np.array([3.0, 1.0]), 2: np.array([2.0, 1.0])}, 0.0, 3)
    # stable: of the two 1.0 values the one of the first-listed sector survives
    assert keep == {0: 2, 2: 1} and [e[0] for e in kept] == [3.0, 2.0, 1.0]


def test_probe_rng_moments_and_skipping():
    r = la.ProbeRNG(7)
    z = r.complex_normal(200000)
    assert abs(np.mean(z)) < 0.01 and abs(np.mean(np.abs(z) ** 2) - 1) < 0.01
    assert abs(np.mean(z.real ** 2) - 0.5) < 0.01
    a = la.ProbeRNG(7)
    a.skip(100)
    b = la.ProbeRNG(7)
    b.complex_normal(100)
    assert np.array_equal(a.complex_normal(5), b.complex_normal(5))
