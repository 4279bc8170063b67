"""Pin the oracle's MPO and contractions against independent dense references (CPU only)."""
import itertools

import numpy as np
import pytest

from oracle import blocks as ob
from oracle import models as om
from tests.helpers import random_legs


@pytest.mark.parametrize("L,delta,sym", [(4, 0.0, "U1"), (5, 0.5, "U1"), (4, 1.0, "none"), (3, 0.0, "none")])
def test_mpo_equals_dense_hamiltonian(L, delta, sym):
    H = om.dense_hamiltonian(L, J=1.0, delta=delta)
    M = om.mpo_dense(om.xxz_mpo(L, J=1.0, delta=delta, symmetry=sym))
    assert np.allclose(M, H, atol=1e-14)


def test_dense_hamiltonian_is_free_fermion_spectrum():
    # XX chain: many-body levels are sums over occupied single-particle energies of the hopping
    # matrix with off-diagonal J/2 (test/common/free_fermion.jl:20-27)
    L = 6
    H = om.dense_hamiltonian(L, J=1.0, delta=0.0)
    h = np.diag(np.full(L - 1, 0.5), 1) + np.diag(np.full(L - 1, 0.5), -1)
    eps = np.linalg.eigvalsh(h)
    levels = sorted(sum(eps[list(c)]) if c else 0.0
                    for n in range(L + 1) for c in itertools.combinations(range(L), n))
    assert np.allclose(np.linalg.eigvalsh(H), levels, atol=1e-12)


def _random_mps(L, dmax, seed, sym="U1"):
    rng = np.random.default_rng(seed)
    legs = random_legs(rng, L, dmax, symmetric=(sym == "U1"))
    return om.random_state(legs, rng, sym), rng


def _dense_env_left(L, A, W, bra):
    return np.einsum("sxZ,axy,abst,tyz->bZz", bra.dense().conj(), L.dense(), W.dense(2), A.dense())


def _dense_env_right(R, B, W, bra):
    return np.einsum("sXZ,bZz,abst,txz->aXx", bra.dense().conj(), R.dense(), W.dense(2), B.dense())


@pytest.mark.parametrize("sym", ["U1", "none"])
def test_env_pushes_match_dense_einsum(sym):
    L = 7
    psi, rng = _random_mps(L, 5, 1, sym)
    mpo = om.xxz_mpo(L, delta=0.7, symmetry=sym)
    lenv = ob.env_fam({0: 1}, {0: 1}, mpo[0].cl)
    lenv.blocks[(0, 0)] = np.ones((1, 1), dtype=complex)
    for i in range(L):
        new = ob.push_left(lenv, psi[i], mpo[i])
        ref = _dense_env_left(lenv, psi[i], mpo[i], psi[i])
        assert np.allclose(new.dense(), ref, atol=1e-12 * max(1.0, np.abs(ref).max()))
        lenv = new
    qtot = list(psi[-1].col)[0]
    renv = ob.env_fam({qtot: 1}, {qtot: 1}, mpo[-1].cr)
    renv.blocks[(0, qtot)] = np.ones((1, 1), dtype=complex)
    for i in reversed(range(L)):
        new = ob.push_right(renv, psi[i], mpo[i])
        ref = _dense_env_right(renv, psi[i], mpo[i], psi[i])
        assert np.allclose(new.dense(), ref, atol=1e-12 * max(1.0, np.abs(ref).max()))
        renv = new


@pytest.mark.parametrize("sym", ["U1", "none"])
def test_energy_from_environments_matches_dense(sym):
    L = 6
    psi, rng = _random_mps(L, 4, 2, sym)
    mpo = om.xxz_mpo(L, delta=0.5, symmetry=sym)
    lenv = ob.env_fam({0: 1}, {0: 1}, mpo[0].cl)
    lenv.blocks[(0, 0)] = np.ones((1, 1), dtype=complex)
    for i in range(L):
        lenv = ob.push_left(lenv, psi[i], mpo[i])
    e_env = sum(b[0, 0] for b in lenv.blocks.values())
    v = om.mps_dense(psi)
    e_dense = np.vdot(v, om.dense_hamiltonian(L, delta=0.5) @ v)
    assert abs(e_env - e_dense) <= 1e-11 * max(1.0, abs(e_dense))


@pytest.mark.parametrize("sym", ["U1", "none"])
def test_one_and_zero_site_match_dense(sym):
    L = 6
    psi, rng = _random_mps(L, 5, 3, sym)
    mpo = om.xxz_mpo(L, delta=0.3, symmetry=sym)
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
    site = 3
    X = psi[site]
    Y = ob.apply_one_site(lenvs[site], mpo[site], renvs[site + 1], X, shift=0.25)
    ref = np.einsum("aXx,txz,bZz,abst->sXZ", lenvs[site].dense(), X.dense(),
                    renvs[site + 1].dense(), mpo[site].dense(2)) - 0.25 * X.dense()
    assert np.allclose(Y.dense(), ref, atol=1e-12 * np.abs(ref).max())
    # zero-site on link 3: C couples left link 3 and right link 3
    C = ob.Fam.zeros(psi[site].row, psi[site].row, [0])
    for key, b in C.blocks.items():
        C.blocks[key] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    Z = ob.apply_zero_site(lenvs[site], renvs[site], C, shift=-0.5)
    refz = np.einsum("wXx,xz,wZz->XZ", lenvs[site].dense(), C.dense()[0], renvs[site].dense()) \
        + 0.5 * C.dense()[0]
    assert np.allclose(Z.dense()[0], refz, atol=1e-12 * np.abs(refz).max())
    # <psi|H|psi> through the zero-site form at any link is the energy
    v = om.mps_dense(psi)
    e = np.vdot(v, om.dense_hamiltonian(L, delta=0.3) @ v)
    # core of psi at link 3 in the product (left part, right part) basis is identity
    Cid = ob.Fam.zeros(psi[site].row, psi[site].row, [0])
    for key, b in Cid.blocks.items():
        Cid.blocks[key] = np.eye(b.shape[0])
    Zid = ob.apply_zero_site(lenvs[site], renvs[site], Cid)
    assert abs(ob.inner(Cid, Zid) - e) <= 1e-10 * abs(e)
