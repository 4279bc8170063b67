"""GPU parity of the device tensor algebra (QR, SVD, projections, direct sums) with the oracle."""
import numpy as np
import pytest

from oracle import blocks as ob
from oracle import linalg as la
from oracle import models as om
from tests.helpers import device_from_fam, fam_from_device, random_legs, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _site(seed, dmax, L=8, site=4):
    rng = np.random.default_rng(seed)
    legs = random_legs(rng, L, dmax)
    return om.random_state(legs, rng)[site]


@pytest.mark.parametrize("seed,dmax", [(0, 3), (1, 20), (2, 90)])
def test_qr_matches_oracle(gpu, seed, dmax):
    from bondupdategalerkin_amd import tensor_ops as T

    A = _site(seed, dmax)
    Ad = device_from_fam(A)
    Q, R = T.qr_left(Ad)
    Qo, Ro = la.qr_left(A)
    assert rel_err(fam_from_device(Q), Qo) < TOL and rel_err(fam_from_device(R), Ro) < TOL
    Lc, Q2 = T.qr_right(Ad)
    Lo, Q2o = la.qr_right(A)
    assert rel_err(fam_from_device(Lc), Lo) < TOL and rel_err(fam_from_device(Q2), Q2o) < TOL


@pytest.mark.parametrize("seed,dmax,nkeep,cutoff", [(3, 4, None, 0.0), (4, 25, 7, 0.0),
                                                     (5, 60, None, 1e-2), (6, 130, 40, 0.0)])
def test_svd_matches_oracle(gpu, seed, dmax, nkeep, cutoff):
    from bondupdategalerkin_amd import tensor_ops as T

    A = _site(seed, dmax)
    Ad = device_from_fam(A)
    for dev_fn, ora_fn in ((T.svd_left, la.svd_left), (T.svd_right, la.svd_right)):
        r = dev_fn(Ad, cutoff=cutoff, nkeep=nkeep)
        o = ora_fn(A, cutoff=cutoff, nkeep=nkeep)
        assert [e[1] for e in r.kept] == [e[1] for e in o.kept]
        assert np.allclose([e[0] for e in r.kept], [e[0] for e in o.kept], rtol=1e-12, atol=1e-14)
        assert rel_err(fam_from_device(r.U), o.U) < 1e-10
        assert rel_err(fam_from_device(r.Vh), o.Vh) < 1e-10
    C = ob.Fam.zeros(A.row, A.row, [0])
    rng = np.random.default_rng(seed)
    for k, b in C.blocks.items():
        C.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    r = T.svd_core(device_from_fam(C), cutoff=cutoff, nkeep=nkeep)
    o = la.svd_core(C, cutoff=cutoff, nkeep=nkeep)
    assert rel_err(fam_from_device(r.U), o.U) < 1e-10 and rel_err(fam_from_device(r.Vh), o.Vh) < 1e-10


def test_projections_direct_sums_and_vectors(gpu):
    from bondupdategalerkin_amd import tensor_ops as T

    A = _site(7, 30)
    B = _site(8, 30)
    U, _ = la.qr_left(A)
    Ud, Bd = device_from_fam(U), device_from_fam(B)
    # B needs A's legs for the projection: use a random tensor in A's layout
    rng = np.random.default_rng(9)
    K = ob.Fam.zeros(A.row, {q: 3 for q in A.col}, A.shifts)
    for k, b in K.blocks.items():
        K.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    Kd = device_from_fam(K)
    assert rel_err(fam_from_device(T.perp_left(Ud, Kd)), la.perp_left(U, K)) < TOL
    Lc, V = la.qr_right(A)
    Kr = ob.Fam.zeros({q: 2 for q in A.row}, A.col, A.shifts)
    for k, b in Kr.blocks.items():
        Kr.blocks[k] = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
    assert rel_err(fam_from_device(T.perp_right(device_from_fam(V), device_from_fam(Kr))),
                   la.perp_right(V, Kr)) < TOL
    assert rel_err(fam_from_device(T.oplus_left(Ud, Kd)), la.oplus_left(U, K)) < TOL
    assert rel_err(fam_from_device(T.oplus_right(device_from_fam(V), device_from_fam(Kr))),
                   la.oplus_right(V, Kr)) < TOL
    x, y = device_from_fam(A), device_from_fam(A.scale(0.5 - 1j))
    assert abs(T.inner(x, y) - ob.inner(A, A.scale(0.5 - 1j))) < 1e-10 * A.norm() ** 2
    assert abs(T.norm(x) - A.norm()) < 1e-12 * A.norm()
    z = T.axpy(0.25j, x, y)
    assert rel_err(fam_from_device(z), A.scale(0.5 - 1j).axpy(0.25j, A)) < TOL
    many = T.inner_many([x] * 70, y)
    assert len(many) == 70 and all(abs(v - many[0]) < 1e-12 * abs(many[0]) for v in many)


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("m,n", [(2500, 6), (700, 40), (300, 17), (5, 1), (64, 64), (610, 455)])
def test_svd_kernels_on_tall_blocks(gpu, m, n, variant):
    """Every SVD path (plain one-sided Jacobi; QR-preconditioned Jacobi on R^H in one workgroup, as
    multi-workgroup block Jacobi, Gram block Jacobi, or block Jacobi started from FP32 sweeps) must
    match LAPACK's spectrum, reconstruct the block and return orthonormal singular vectors."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import tensor_ops as T

    ctx = _lib.context()
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, variant, 0))

    rng = np.random.default_rng(m + n)
    C = ob.Fam({0: m}, {0: n}, [0])
    M = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    M[:, -1] = M[:, 0] * 1e-9 + M[:, -1] * 1e-12 if n > 1 else M[:, -1]   # a tiny singular value
    C.blocks[(0, 0)] = M
    r = T.svd_core(device_from_fam(C))
    o = la.svd_core(C)
    assert np.allclose(r.S[0], o.S[0], rtol=1e-12, atol=1e-14 * o.S[0][0])
    U = fam_from_device(r.U).dense()[0]
    Vh = fam_from_device(r.Vh).dense()[0]
    assert np.abs(U * r.S[0][None, :] @ Vh - M).max() < 1e-12 * np.abs(M).max()
    assert np.abs(U.conj().T @ U - np.eye(U.shape[1])).max() < 1e-12
    # the phase convention: first entry with |u| >= 0.5 max|u| real positive
    for j in range(U.shape[1]):
        a = np.abs(U[:, j])
        i = int(np.argmax(a >= 0.5 * a.max()))
        assert abs(U[i, j].imag) < 1e-12 and U[i, j].real > 0
    # singular vectors are determined to ~eps sigma_max / sigma_j: compare the well-separated ones
    well = o.S[0] > 1e-6 * o.S[0][0]
    assert np.allclose(U[:, well], o.U.dense()[0][:, well], atol=1e-9)
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, 0, 0))


@pytest.mark.parametrize("variant", [2, 3, 4, 5, 6])
def test_svd_graded_spectrum_batch(gpu, variant):
    """A batch of blocks of different widths with geometrically graded spectra (1 .. 1e-12): the
    singular values are accurate to 1e-13 sigma_max and U S Vh reconstructs each block."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import tensor_ops as T

    ctx = _lib.context()
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, variant, 0))
    rng = np.random.default_rng(7)
    dims = {-2: (90, 70), 0: (530, 470), 2: (260, 250), 4: (33, 30)}
    C = ob.Fam({q: d[0] for q, d in dims.items()}, {q: d[1] for q, d in dims.items()}, [0])
    sig = {}
    for q, (m, n) in dims.items():
        Qa, _ = np.linalg.qr(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
        Qb, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        sig[q] = 10.0 ** (-12.0 * np.arange(n) / (n - 1)) * (1.0 + 0.1 * q)
        C.blocks[(0, q)] = (Qa * sig[q][None, :]) @ Qb.conj().T
    r = T.svd_core(device_from_fam(C))
    Uf, Vf = fam_from_device(r.U), fam_from_device(r.Vh)
    for q, (m, n) in dims.items():
        s = r.S[q]
        assert np.abs(s - sig[q]).max() < 1e-13 * sig[q][0], q
        M = C.blocks[(0, q)]
        U, Vh = Uf.blocks[(0, q)], Vf.blocks[(0, q)]
        assert np.abs(U * s[None, :] @ Vh - M).max() < 1e-12 * sig[q][0], q
        assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-11, q
        assert np.abs(Vh @ Vh.conj().T - np.eye(n)).max() < 1e-11, q
    ctx.check(ctx._lib.bug_set_svd_variant(ctx.handle, 0, 0))


@pytest.mark.parametrize("variant", [1, 2])
def test_qr_paths_on_wide_blocks(gpu, variant):
    """Both QR paths (one workgroup per block; blocked panels with compact-WY GEMM updates, 16- and
    8-column panels) give the unique thin QR (real non-negative R diagonal) of blocks up to
    1000 x 480, tall and wide, and a rank-deficient block still gives Q R = A with Q^H Q = I."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import tensor_ops as T

    ctx = _lib.context()
    ctx.check(ctx._lib.bug_set_qr_variant(ctx.handle, variant))
    rng = np.random.default_rng(11)
    for dims in ({-2: (300, 140), 0: (500, 480), 2: (20, 50), 4: (61, 61)},
                 {0: (1000, 470), 2: (700, 333), -4: (1, 3)}):
        C = ob.Fam({q: d[0] for q, d in dims.items()}, {q: d[1] for q, d in dims.items()}, [0])
        for q, (m, n) in dims.items():
            C.blocks[(0, q)] = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        C.blocks[(0, 0)][:, 7] = 0.0                       # rank-deficient: one zero column
        Q, R = T.qr_core(device_from_fam(C))
        Qf, Rf = fam_from_device(Q), fam_from_device(R)
        for q, (m, n) in dims.items():
            M = C.blocks[(0, q)]
            Qd, Rd = Qf.blocks[(0, q)], Rf.blocks[(0, q)]
            k = min(m, n)
            assert Qd.shape == (m, k) and Rd.shape == (k, n)
            assert np.abs(Qd @ Rd - M).max() < 1e-12 * np.abs(M).max(), q
            assert np.abs(Qd.conj().T @ Qd - np.eye(k)).max() < 1e-12, q
            assert np.all(np.abs(np.tril(Rd, -1)) == 0) and np.all(np.diag(Rd).real >= 0), q
            assert np.abs(np.diag(Rd).imag).max() == 0, q
            if q != 0:                                       # full rank: the unique factors
                Qo, Ro = la._unique_qr(M)
                assert np.abs(Qd - Qo).max() < 1e-11 and np.abs(Rd - Ro).max() < 1e-11 * np.abs(M).max(), q
    ctx.check(ctx._lib.bug_set_qr_variant(ctx.handle, 0))


@pytest.mark.parametrize("m,n", [(40, 8), (120, 60)])
def test_svd_refuses_non_finite_blocks(gpu, m, n):
    """A block holding a NaN gives NaN singular values, which the host refuses; the device sort of
    the spectrum stays a permutation, so nothing is read outside the block (the narrow block runs
    the one-workgroup kernel, the wide one the preconditioned block Jacobi)."""
    from bondupdategalerkin_amd import tensor_ops as T

    rng = np.random.default_rng(n)
    C = ob.Fam({0: m, 2: m}, {0: n, 2: n}, [0])
    for q in (0, 2):
        C.blocks[(0, q)] = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    C.blocks[(0, 2)][3, 1] = np.nan
    with pytest.raises(FloatingPointError):
        T.svd_core(device_from_fam(C))
    C.blocks[(0, 2)][3, 1] = 1.0
    r = T.svd_core(device_from_fam(C))
    o = la.svd_core(C)
    assert np.allclose(r.S[2], o.S[2], rtol=1e-12)


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("seed,dmax,rank", [(0, 6, None), (1, 20, 3), (2, 40, 1)])
def test_pivoted_thin_qr_matches_oracle(gpu, side, seed, dmax, rank):
    """_thin_qr's sector-wise column-pivoted QR on the device against the oracle's: the same
    kept widths and the same orthonormal columns (both with a non-negative R diagonal), on
    full-rank and rank-deficient sketches."""
    from bondupdategalerkin_amd import tensor_ops as T
    from oracle import dd_adaptive as da

    rng = np.random.default_rng(seed)
    row, col, shifts = {-1: dmax, 1: dmax + 2}, {0: 7, 2: 5, -2: 9}, [1, -1]
    if side == "right":
        row, col = {0: 7, 2: 5, -2: 9}, {-1: dmax, 1: dmax + 2}
    A = ob.Fam.zeros(row, col, shifts)
    for key, b in A.blocks.items():
        M = rng.standard_normal(b.shape) + 1j * rng.standard_normal(b.shape)
        if rank is not None:
            k = min(rank, *b.shape)
            M = M[:, :k] @ (rng.standard_normal((k, b.shape[1])) + 0j) if side == "left" else \
                (rng.standard_normal((b.shape[0], k)) + 0j) @ M[:k, :]
        A.blocks[key] = M
    thr = 1e-12 * A.norm()
    o = da.thin_qr(A, side, thr)
    d = T.thin_qr(device_from_fam(A), side, thr)
    assert (o is None) == (d is None)
    if o is None:
        return
    df = fam_from_device(d)
    assert (df.col if side == "left" else df.row) == (o.col if side == "left" else o.row)
    assert rel_err(df, o) < 1e-10


def test_axpy_many_equals_chained_axpys(gpu):
    """The fused y += sum_j a_j x_j (70 vectors: two launches) against one axpy per vector."""
    import torch

    from bondupdategalerkin_amd import tensor_ops as T
    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, Leg

    rng = np.random.default_rng(12)
    lay = Layout(Leg.make([(0, 37), (2, 5)]), Leg.make([(1, 19), (3, 40)]), [1])
    def rand():
        blocks = {}
        for k, q, qc, d, dc, off in lay.blocks():
            blocks[(k, q)] = rng.standard_normal((d, dc)) + 1j * rng.standard_normal((d, dc))
        return BlockTensor.from_blocks(lay, blocks)
    xs = [rand() for _ in range(70)]
    alphas = list(rng.standard_normal(70) + 1j * rng.standard_normal(70))
    y0 = rand()
    fused = T.axpy_many_(alphas, xs, y0.clone())
    chained = y0.clone()
    for a, x in zip(alphas, xs):
        T.axpy_(a, x, chained)
    torch.cuda.synchronize()
    diff = (fused.data - chained.data).abs().max().item()
    assert diff <= 1e-14 * chained.data.abs().max().item()
    with pytest.raises(ValueError):
        T.axpy_many_(alphas[:3], xs[:2], y0)


@pytest.mark.parametrize("cols", [32, 64])
def test_gram_block_jacobi_pair_widths(gpu, cols):
    """The Gram block Jacobi with block pairs of 32 columns (the default, 512-thread rotate
    workgroups) and of 64 (1024 threads, selectable through bug_set_gram_jacobi_params): blocks
    wide enough to take it, with graded spectra (1 .. 1e-12), give singular values accurate to
    1e-13 sigma_max, reconstruct, and return orthonormal factors."""
    from bondupdategalerkin_amd import _lib
    from bondupdategalerkin_amd import tensor_ops as T

    ctx = _lib.context()
    ctx.check(ctx._lib.bug_set_gram_jacobi_params(ctx.handle, cols, 1, 128))
    try:
        rng = np.random.default_rng(3)
        dims = {-2: (300, 190), 0: (610, 455), 2: (530, 470), 4: (40, 33)}
        C = ob.Fam({q: d[0] for q, d in dims.items()}, {q: d[1] for q, d in dims.items()}, [0])
        sig = {}
        for q, (m, n) in dims.items():
            Qa, _ = np.linalg.qr(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
            Qb, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
            sig[q] = 10.0 ** (-12.0 * np.arange(n) / (n - 1))
            C.blocks[(0, q)] = (Qa * sig[q][None, :]) @ Qb.conj().T
        r = T.svd_core(device_from_fam(C))
        Uf, Vf = fam_from_device(r.U), fam_from_device(r.Vh)
        for q, (m, n) in dims.items():
            s, M = r.S[q], C.blocks[(0, q)]
            U, Vh = Uf.blocks[(0, q)], Vf.blocks[(0, q)]
            assert np.abs(s - sig[q]).max() < 1e-13 * sig[q][0], q
            assert np.abs(U * s[None, :] @ Vh - M).max() < 1e-12, q
            assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-11, q
            assert np.abs(Vh @ Vh.conj().T - np.eye(n)).max() < 1e-11, q
    finally:
        ctx.check(ctx._lib.bug_set_gram_jacobi_params(ctx.handle, 32, 1, 128))
