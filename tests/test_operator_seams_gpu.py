"""The step's operator actions through the C ABI (GPU): the streamed two-site sketches
(``bug_two_site_sketch_left/right``) against the oracle, ``bug_expv`` over a caller-supplied
action against the matrix exponential and the package's own Krylov, and ``bug_svd_select``
against the oracle's selection rule.

Bars: sketches within 1e-12 relative of the oracle (FP64; 1e-5 for the FP32 form), expv within
1e-10 of scipy's expm on dense operators and 1e-12 of the package's Lanczos on the one-site
effective Hamiltonian, the selection identical (counts, order, kept/truncated split)."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import bug as obug
from oracle import linalg as ola
from tests.helpers import fam_from_device, oracle_mpo_site, rel_err

pytestmark = pytest.mark.gpu


def _chain(L, dmax, seed, delta):
    from bondupdategalerkin_amd import mps
    from bondupdategalerkin_amd.environments import left_env_stack, right_env_stack
    from bondupdategalerkin_amd.mpo import xxz_mpo

    psi = mps.random_mps(L, dmax, seed=seed)
    mpo = xxz_mpo(L, delta=delta)
    return psi, mpo, left_env_stack(psi, mpo), right_env_stack(psi, mpo)


def _probe(rng, like, side, width):
    """A complex Gaussian family in the layout of ``like`` with at most ``width`` columns per
    sector on its row leg (side left) or column leg (side right)."""
    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, Leg

    lay = like.layout
    if side == "left":
        out = Layout(Leg(lay.row.charges, tuple(min(width, d) for d in lay.row.dims)), lay.col,
                     lay.shifts)
    else:
        out = Layout(lay.row, Leg(lay.col.charges, tuple(min(width, d) for d in lay.col.dims)),
                     lay.shifts)
    blocks = {(k, q): rng.standard_normal((d, dc)) + 1j * rng.standard_normal((d, dc))
              for k, q, qc, d, dc, off in out.blocks()}
    return BlockTensor.from_blocks(out, blocks)


def _repacked(t, gap=5):
    """The same family stored with its blocks in reverse order and ``gap`` elements between
    them: (descriptor, device buffer, offsets) — a packing that is not the library's own."""
    import torch

    lay = t.layout
    offs = np.full(lay.nslice * lay.row.nsec, -1, np.int64)
    blocks = list(lay.blocks())
    pos = gap
    place = {}
    for k, q, qc, d, dc, off in reversed(blocks):
        place[(k, q)] = (pos, off, d * dc)
        offs[k * lay.row.nsec + lay.row.charges.index(q)] = pos
        pos += d * dc + gap
    buf = torch.zeros(pos, dtype=torch.complex128, device=t.data.device)
    for (k, q), (dst, src, n) in place.items():
        buf[dst:dst + n] = t.data[src:src + n]
    d = lay.descriptor(buf.data_ptr(), pos)
    from bondupdategalerkin_amd import _lib

    d.offset = _lib.ptr64(offs)
    return d, buf, offs


@pytest.mark.parametrize("L,dmax,seed,delta,bond,width", [
    (8, 6, 1, 0.5, 3, 4),
    (10, 12, 2, 1.0, 4, 8),
    (6, 4, 3, 0.0, 0, 3),
    (6, 4, 4, 0.0, 4, 3),
])
def test_two_site_sketches_match_oracle(gpu, L, dmax, seed, delta, bond, width):
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C

    psi, mpo, le, re = _chain(L, dmax, seed, delta)
    rng = np.random.default_rng(seed)
    A, V = psi[bond], psi[bond + 1]
    W1, W2 = mpo[bond], mpo[bond + 1]
    Lv, Rv = le[bond], re[bond + 2]
    oW1, oW2 = oracle_mpo_site(W1), oracle_mpo_site(W2)
    print(f"case L={L} dmax={dmax} bond={bond} width={width}", flush=True)
    Om = _probe(rng, V, "left", width)
    ref = obug.stream_left(fam_from_device(Lv), oW1, oW2, fam_from_device(Rv),
                           fam_from_device(A), fam_from_device(V), fam_from_device(Om))
    py = B.stream_left(Lv, W1, W2, Rv, A, V, Om)
    got = C.two_site_sketch("left", Lv, W1, W2, Rv, A, V, Om, py.layout)
    assert rel_err(fam_from_device(got), ref) <= 1e-12
    assert rel_err(fam_from_device(py), ref) <= 1e-12
    mixed = C.two_site_sketch("left", Lv, W1, W2, Rv, A, V, Om, py.layout, precision="mixed")
    assert rel_err(fam_from_device(mixed), ref) <= 1e-5

    Om = _probe(rng, A, "right", width)
    ref = obug.stream_right(fam_from_device(Lv), oW1, oW2, fam_from_device(Rv),
                            fam_from_device(A), fam_from_device(V), fam_from_device(Om))
    py = B.stream_right(Lv, W1, W2, Rv, A, V, Om)
    got = C.two_site_sketch("right", Lv, W1, W2, Rv, A, V, Om, py.layout)
    assert rel_err(fam_from_device(got), ref) <= 1e-12


def test_two_site_sketch_reads_and_writes_foreign_packings(gpu):
    """Inputs and output whose blocks sit at offsets of the caller's choosing (reverse order,
    gaps between blocks): the entry point gathers and scatters instead of viewing."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd.blocks import BlockTensor, current_stream_handle

    psi, mpo, le, re = _chain(8, 8, 5, 0.5)
    rng = np.random.default_rng(5)
    A, V = psi[3], psi[4]
    Om = _probe(rng, V, "left", 5)
    Lv, Rv = le[3], re[5]
    py = B.stream_left(Lv, mpo[3], mpo[4], Rv, A, V, Om)
    keep = [_repacked(t) for t in (Lv, Rv, A, V, Om)]
    out_t = BlockTensor(py.layout)
    od, obuf, ooffs = _repacked(out_t, gap=3)
    ctx = _lib.context()
    ctx.check(ctx._lib.bug_two_site_sketch_left(
        ctx.handle, keep[0][0], mpo[3].descriptor(), mpo[4].descriptor(), keep[1][0],
        keep[2][0], keep[3][0], keep[4][0], od, _lib.PREC_FP64, current_stream_handle()))
    host = obuf.cpu().numpy()
    lay = py.layout
    want = py.to_blocks()
    for k, q, qc, d, dc, off in lay.blocks():
        o = ooffs[k * lay.row.nsec + lay.row.charges.index(q)]
        blk = host[o:o + d * dc].reshape(dc, d).T
        assert np.abs(blk - want[(k, q)]).max() <= 1e-12 * max(1.0, np.abs(want[(k, q)]).max())
    # the gaps stay untouched
    mask = np.ones(host.size, bool)
    for k, q, qc, d, dc, off in lay.blocks():
        o = ooffs[k * lay.row.nsec + lay.row.charges.index(q)]
        mask[o:o + d * dc] = False
    assert np.all(host[mask] == 0)


def test_two_site_sketch_refuses_a_wrong_output_layout(gpu):
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import bug_step as B
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd.blocks import Layout

    psi, mpo, le, re = _chain(6, 4, 6, 0.5)
    rng = np.random.default_rng(6)
    Om = _probe(rng, psi[3], "left", 3)
    py = B.stream_left(le[2], mpo[2], mpo[3], re[4], psi[2], psi[3], Om)
    bad = Layout(py.layout.row, py.layout.col, tuple(py.layout.shifts[:1]))
    with pytest.raises(_lib.BugError):
        C.two_site_sketch("left", le[2], mpo[2], mpo[3], re[4], psi[2], psi[3], Om, bad)


def _dense_apply(M):
    """apply(x, y, stream): y = M x through the library's grouped product on raw addresses."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd.blocks import Layout, Leg

    n = M.layout.row.dims[0]
    vec = Layout(Leg((0,), (n,)), Leg((0,), (1,)), (0,))
    term = (_lib.BugTerm * 1)(_lib.BugTerm(0, 0, 0, 0, 0, 1.0, 0.0))
    ctx = _lib.context()

    def apply(x, y, stream):
        ctx.check(ctx._lib.bug_bsm_product(ctx.handle, vec.descriptor(y, n), M.desc(),
                                           vec.descriptor(x, n), 1, term, 0, _lib.PREC_FP64,
                                           ctypes.c_void_p(stream)))
    return apply


def _dense_operator(n, seed, hermitian, stiff=False):
    import torch

    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, Leg, device

    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    M = (X + X.conj().T) / 2 if hermitian else X
    M = M / np.linalg.norm(M, 2) * 3.0
    if stiff:
        # a stiff, strongly non-normal generator like a discretised drift-diffusion: decay rates
        # over three decades on the diagonal, one-sided transport above it, a dense weak coupling
        M = (np.diag(-np.exp(np.linspace(0.0, 7.0, n))) + np.diag(np.full(n - 1, 60.0), 1)
             + 0.3 * M)
    lay = Layout(Leg((0,), (n,)), Leg((0,), (n,)), (0,))
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return M, BlockTensor.from_blocks(lay, {(0, 0): M}), v, torch.from_numpy(v).to(device())


@pytest.mark.parametrize("n,hermitian,tau,maxiter,conv_tol", [
    (48, True, -0.7j, 30, 0.0),
    (48, True, -0.5, 30, 0.0),
    (64, True, -2.0j, 12, 1e-13),
    (40, False, -0.3 + 0.2j, 30, 0.0),
    (40, False, -1.5j, 10, 1e-12),
    (80, False, 0.05, 30, 1e-10),
])
def test_expv_matches_the_matrix_exponential(gpu, n, hermitian, tau, maxiter, conv_tol):
    """``bug_expv`` against scipy's expm; the last case is stiff (||tau M|| ~ 55 with a
    non-normal drift) and needs the substep search and the Hessenberg eigenvalues of many
    graded Krylov blocks."""
    from bondupdategalerkin_amd import cstep as C

    M, Md, v, vd = _dense_operator(n, 7 + n, hermitian, stiff=n == 80)
    print(f"case n={n} hermitian={hermitian} tau={tau} maxiter={maxiter}", flush=True)
    out, rep = C.expv(_dense_apply(Md), vd, tau, hermitian=hermitian,
                      tol=1e-15 if hermitian else 1e-13, maxiter=maxiter, conv_tol=conv_tol)
    want = sla.expm(tau * M) @ v
    got = out.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
    assert rep["iterations"] >= 1 and rep["substeps"] >= 1
    # the input stays as it was
    assert np.array_equal(vd.cpu().numpy(), v)


@pytest.mark.parametrize("tau,hermitian", [(-0.1j, True), (-0.2, True), (-0.15j, False)])
def test_expv_of_the_one_site_operator_matches_the_package_krylov(gpu, tau, hermitian):
    """The root solve's shape: expv over the library's one-site effective Hamiltonian (the
    callback calls ``bug_apply_one_site`` on the Krylov vectors' addresses)."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import cstep as C
    from bondupdategalerkin_amd import krylov as K
    from bondupdategalerkin_amd.blocks import BlockTensor
    from bondupdategalerkin_amd.environments import apply_one_site

    psi, mpo, le, re = _chain(10, 10, 9, 0.5)
    j = 5
    X = psi[j]
    lay, n = X.layout, X.layout.nelem
    ctx = _lib.context()
    Ld, Wd, Rd = le[j].desc(), mpo[j].descriptor(), re[j + 1].desc()

    def apply(x, y, stream):
        ctx.check(ctx._lib.bug_apply_one_site(ctx.handle, Ld, Wd, Rd, lay.descriptor(x, n),
                                              lay.descriptor(y, n), 0.0, _lib.PREC_FP64,
                                              ctypes.c_void_p(stream)))

    out, rep = C.expv(apply, X.data[:n].clone(), tau, hermitian=hermitian,
                      tol=1e-15 if hermitian else 1e-13, maxiter=30)
    act = lambda t: apply_one_site(le[j], mpo[j], re[j + 1], t)
    if hermitian:
        want, wrep = K.lanczos_expv(act, tau, X, tol=1e-15, maxiter=30)
    else:
        want, wrep = K.arnoldi_expv(act, tau, X, tol=1e-13, maxiter=30)
    w = want.data[:n].cpu().numpy()
    assert np.linalg.norm(out.cpu().numpy() - w) <= 1e-12 * np.linalg.norm(w)
    assert (rep["iterations"], rep["substeps"]) == (wrep.iterations, wrep.substeps)
    # the tail estimates agree; below 1e-16 both are rounding of negligible coefficients
    assert abs(rep["err"] - wrep.err) <= 1e-6 * wrep.err + 1e-16
    # unitary for real time
    if tau.real == 0:
        assert abs(np.linalg.norm(w) / np.linalg.norm(X.data[:n].cpu().numpy()) - 1) < 1e-10
    _ = BlockTensor


def test_expv_reports_callback_failures_and_bad_arguments(gpu):
    import torch

    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import cstep as C

    M, Md, v, vd = _dense_operator(16, 3, True)

    def broken(x, y, stream):
        raise ValueError("operator unavailable")

    with pytest.raises(ValueError, match="operator unavailable"):
        C.expv(broken, vd, -0.1j)
    with pytest.raises(_lib.BugError):
        C.expv(_dense_apply(Md), vd, -0.1j, maxiter=0)
    ctx = _lib.context()
    rc = ctx._lib.bug_expv(ctx.handle, _lib.APPLY_FN(lambda *a: 0), None, 16, 0.0, -0.1,
                           vd.data_ptr(), vd.data_ptr(), 1, 1e-15, 30, 1, 0.0, None, None)
    assert rc == _lib.BUG_EINVAL
    # the library still works afterwards
    out, _ = C.expv(_dense_apply(Md), vd, -0.1j)
    assert torch.isfinite(out).all()


@pytest.mark.parametrize("seed,cutoff,nkeep", [(0, 0.0, None), (1, 1e-3, None), (2, 0.0, 7),
                                               (3, 1e-2, 4), (4, 0.0, 0)])
def test_svd_select_matches_the_oracle_rule(gpu, seed, cutoff, nkeep):
    from bondupdategalerkin_amd import cstep as C

    rng = np.random.default_rng(seed)
    spectra = []
    for _ in range(5):
        s = np.sort(np.abs(rng.standard_normal(int(rng.integers(0, 6)))))[::-1]
        s = np.round(s, 1)                       # ties across sectors exercise the stable order
        spectra.append(s)
    keep, order, nk, w = C.svd_select(spectra, cutoff, nkeep)
    okeep, kept, trunc = ola._select({i: s for i, s in enumerate(spectra)}, cutoff, nkeep)
    assert keep == [okeep[i] for i in range(len(spectra))]
    assert nk == len(kept)
    assert order == [(q, j) for _, q, j in kept] + [(q, j) for _, q, j in trunc]
    assert w[0] == pytest.approx(sum(s * s for s, _, _ in kept), rel=1e-15, abs=0)
    assert w[1] == pytest.approx(sum(s * s for s, _, _ in trunc), rel=1e-15, abs=0)


def test_svd_select_hand_spectra(gpu):
    from bondupdategalerkin_amd import cstep as C

    keep, order, nk, w = C.svd_select([np.array([3.0, 1.0]), np.array([2.0, 1.0])], 0.0, 3)
    assert keep == [2, 1] and nk == 3
    assert order == [(0, 0), (1, 0), (0, 1), (1, 1)]
    assert w == [14.0, 1.0]
    keep, order, nk, w = C.svd_select([np.array([3.0, 1.0, 0.0]), np.array([2.0, 1.0])], 0.0)
    assert keep == [2, 2] and nk == 4
