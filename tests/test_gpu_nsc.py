"""GPU parity tests of the null-space Cholesky build (kernel id 5): for bases definite
on the polynomial null space it solves each stencil without pivot search plus one FP64
refinement step; what it cannot serve (Hermite stencils, breakdowns, a correction that
does not settle) it hands to the rook-pivoted kernels. Every case is checked per stencil
against the extended-precision oracle at max(1e-10, kappa_2(A) u), the bound the rook
kernels meet (test_gpu_build.check_against_truth)."""
import contextlib
import pathlib
from math import comb

import numpy as np
import pytest

from test_gpu_build import _c1_points, _op, check_against_truth

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parent.parent
K_REG32, K_WIDE, K_NSC, K_NSC2 = 0, 2, 5, 6


@contextlib.contextmanager
def pinned(RBF, kernel, buckets=None):
    if buckets is None:
        buckets = {K_REG32: (0,), K_WIDE: (1, 2), K_NSC: (0, 1), K_NSC2: (0,)}[kernel]
    for b in buckets:
        RBF._lib.tune_bucket(b, kernel)
    try:
        yield
    finally:
        for b in buckets:
            RBF._lib.tune_bucket(b, -1)


def _basis(RBF, kind, eps, deg):
    if kind < 10:
        return RBF.PHS(kind, poly_deg=deg)
    return (RBF.IMQ if kind == 10 else RBF.Gaussian)(eps, poly_deg=deg)


# (kind, eps * stencil radius, dim, degree, k): every basis with a degree at or above its
# order, 1-D to 3-D, both built shapes (k <= 20 with m <= 6, k <= 32 with m <= 10)
CASES = [(3, 0, 2, 2, 20), (3, 0, 2, 1, 12), (1, 0, 2, 0, 12), (1, 0, 2, 2, 20), (5, 0, 2, 2, 20),
         (5, 0, 2, 3, 30), (7, 0, 2, 3, 30), (3, 0, 3, 1, 16), (3, 0, 3, 2, 30), (3, 0, 1, 2, 8),
         (10, 1.5, 2, 2, 20), (10, 1.5, 3, 1, 16), (11, 1.5, 2, 2, 20), (11, 1.5, 2, -1, 12)]


# (kind, eps h, dim, degree, k, kernel): every case on the 32-lane kernel, and those
# with k <= 20 and at most 6 monomials on the 8-lane kernel too
def _n_mono(dim, deg):
    return comb(dim + deg, dim) if deg >= 0 else 0


NSC_CASES = [c + (K_NSC,) for c in CASES] + [c + (K_NSC2,) for c in CASES if c[4] <= 20 and _n_mono(c[2], c[3]) <= 6]


@pytest.mark.parametrize("kind,epsh,dim,deg,k,kernel", NSC_CASES)
def test_nsc_parity_against_truth(RBF, O, gpu, kind, epsh, dim, deg, k, kernel):
    rng = np.random.default_rng(1000 + kind * 37 + dim * 11 + deg * 3 + k)
    n = 1500
    pts = rng.random((n, dim))
    eps = 1.0
    if kind >= 10:
        vol = {1: 2.0, 2: np.pi, 3: 4.0 * np.pi / 3.0}[dim]
        eps = epsh / (k / (n * vol)) ** (1.0 / dim)
    L = RBF.Laplacian() + 0.5 * RBF.Partial(1, 1)
    with pinned(RBF, kernel):
        op = RBF.RadialBasisOperator(L, pts, basis=_basis(RBF, kind, eps, deg), k=k)
    nb, kern = RBF._lib.last_build_buckets()
    assert kernel in kern, (nb, kern)
    handed = RBF._lib.last_build_stats()[0]
    assert handed <= n // 20, handed  # breakdowns are rare for these pairings
    adj = op.adj_idx.cpu().numpy()
    terms = [(0, O.abi.OP_LAP, 0, 0, 1.0), (0, O.abi.OP_D1, 0, 0, 0.5)]
    check_against_truth(O, op.vals.cpu().numpy(), pts, pts, adj, O.basis_struct(kind, deg, eps), _op(O, 1, terms), k)


@pytest.mark.parametrize("kernel", [K_NSC, K_NSC2])
def test_nsc_regrid_and_partial(RBF, O, gpu, kernel):
    rng = np.random.default_rng(5)
    pts = rng.random((2500, 2))
    ev = rng.random((900, 2))
    with pinned(RBF, kernel):
        reg = RBF.regrid(pts, ev, basis=RBF.PHS(3, poly_deg=2), k=20)
        par = RBF.partial(pts, 2, 2, basis=RBF.PHS(3, poly_deg=2), k=20)
    b = O.basis_struct(3, 2)
    check_against_truth(O, reg.vals.cpu().numpy(), pts, ev, reg.adj_idx.cpu().numpy(), b,
                        _op(O, 1, [(0, O.abi.OP_VALUE, 0, 0, 1.0)]), 20)
    check_against_truth(O, par.vals.cpu().numpy(), pts, pts, par.adj_idx.cpu().numpy(), b,
                        _op(O, 1, [(0, O.abi.OP_D2, 1, 0, 1.0)]), 20)


@pytest.mark.parametrize("kernel", [K_NSC, K_NSC2])
def test_nsc_c1_laplacian(RBF, O, gpu, kernel):
    """The C1 configuration (10,000 points, PHS3, degree 2, k = 20) on the null-space
    kernels: every stencil within the bound, 99.5% within 1e-10 of the reference's own
    uncentred arithmetic, and the applied Laplacian as accurate as the rook kernel's."""
    pts = _c1_points()
    with pinned(RBF, kernel):
        op = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20)
    assert RBF._lib.last_build_buckets()[1][0] == kernel
    with pinned(RBF, K_REG32):
        rook = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20, adjl=op.adj_idx.cpu().numpy())
    adj = op.adj_idx.cpu().numpy()
    lap = _op(O, 1, [(0, O.abi.OP_LAP, 0, 0, 1.0)])
    b = O.basis_struct(3, 2)
    wg = op.vals.cpu().numpy()
    eg, _ = check_against_truth(O, wg, pts, pts, adj, b, lap, 20)
    assert np.median(eg) <= 1e-12
    _, _, wr, _, _ = O.build(pts, pts, adj, b, lap, scaled=False)
    er = np.abs(wg[0] - wr[0]).reshape(-1, 20).max(1) / np.abs(wr[0]).reshape(-1, 20).max(1)
    assert np.quantile(er, 0.995) <= 1e-10 and er.max() <= 1e-7
    f = np.sin(4 * pts[:, 0]) * np.cos(3 * pts[:, 1])
    e_nsc = np.sqrt(np.mean((op(f) + 25 * f) ** 2))
    e_rook = np.sqrt(np.mean((rook(f) + 25 * f) ** 2))
    assert abs(e_nsc - e_rook) <= 1e-9 * e_rook


@pytest.mark.parametrize("kernel", [K_NSC, K_NSC2])
def test_nsc_hands_back_hermite_stencils(RBF, O, gpu, kernel):
    """Hermite stencils go to the rook kernel; the build matches the oracle and the
    manufactured solution u = x^2 + y^2 + 1 (laplacian_end_to_end.jl:45-78)."""
    from test_oracle import _unit_square_441

    pts, isb, nrm, bcs = _unit_square_441()
    adjl = O.knn(pts, pts, 15)
    with pinned(RBF, kernel):
        op = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), adjl=adjl,
                           hermite=dict(is_boundary=isb, bc=[RBF.BoundaryCondition(a, b) for a, b in bcs],
                                        normals=nrm))
        handed = RBF._lib.last_build_stats()[0]
    assert 0 < handed < len(pts)
    u = pts[:, 0] ** 2 + pts[:, 1] ** 2 + 1
    un = (2 * pts[isb] * nrm).sum(1)
    ub = np.where((bcs[:, 0] == 1) & (bcs[:, 1] == 0), u[isb], bcs[:, 0] * u[isb] + bcs[:, 1] * un)
    uu = u.copy()
    uu[isb] = ub
    res = op(uu)
    assert np.abs(res[~isb] - 4.0).max() < 1e-10
    assert np.abs(res[isb] - ub).max() < 1e-10
    ptr, col, ws, st, rc = O.build(pts, pts, adjl, O.basis_struct(3, 2), _op(O, 1, [(0, 4, 0, 0, 1.0)]),
                                   hermite=(isb, bcs[:, 0], bcs[:, 1], nrm), scaled=True)
    W = op.weights()
    assert np.array_equal(W.indptr, ptr) and np.array_equal(W.indices, col)
    assert np.abs(W.data - ws[0]).max() <= 1e-10 * np.abs(ws[0]).max()


@pytest.mark.parametrize("kernel", [K_NSC, K_NSC2])
def test_nsc_repeated_builds_bit_identical(RBF, gpu, kernel):
    pts = np.random.default_rng(99).random((20000, 2))
    first = None
    with pinned(RBF, kernel):
        for rep in range(8):
            op = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20)
            w = op.vals.clone()
            if first is None:
                first = w
            else:
                assert bool((w == first).all()), rep


@pytest.mark.parametrize("kernel", [K_NSC, K_NSC2])
def test_nsc_singular_stencil_falls_back(RBF, gpu, kernel):
    """A duplicated point makes a stencil singular: the null-space kernel hands it back
    and the rook kernel reports it as the reference does (SingularException)."""
    pts = np.random.default_rng(0).random((50, 2))
    pts[1] = pts[0]
    with pinned(RBF, kernel):
        with pytest.raises(RBF.SingularException):
            RBF.laplacian(pts, k=12)


def test_tuning_map_saved_and_reused(RBF, gpu, tmp_path):
    """The autotuner's choices survive a save / reset / load: the reloaded map serves the
    same kernel with no tuning samples (what a second process does with the file)."""
    path = tmp_path / "tune.txt"
    pts = np.random.default_rng(3).random((30000, 2))
    RBF._lib.tune_reset()
    RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20)
    tuned, kern = RBF._lib.last_build_stats()[1], RBF._lib.last_build_buckets()[1][0]
    assert tuned > 0
    RBF._lib.tune_save(path)
    lines = [l for l in path.read_text().splitlines() if not l.startswith("#")]
    assert lines and all(len(l.split()) == 10 for l in lines)
    RBF._lib.tune_reset()
    RBF._lib.tune_load(path)
    RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20)
    assert RBF._lib.last_build_stats()[1] == 0
    assert RBF._lib.last_build_buckets()[1][0] == kern
    RBF._lib.tune_reset()


def test_tuned_choice_never_given_to_ragged_build(RBF, O, gpu, tmp_path):
    """The null-space kernel takes k and m from the build, so it serves only builds whose
    stencils all have one shape. A choice the autotuner made for a uniform build (k = 30,
    bucket 33..64 rows), or a loaded map that names it for a ragged build of the same
    largest shape, is not used for a ragged adjacency list (k from 27 to 30): that
    build runs a rook kernel and matches the oracle."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(41)
    RBF._lib.tune_reset()
    pts = rng.random((30_000, 2))
    RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=30)
    n = 6000
    pv = rng.random((n, 2))
    _, full = cKDTree(pv).query(pv, 30)
    ks = rng.integers(27, 31, size=n)
    adjl = [full[i, :ks[i]] for i in range(n)]
    path = tmp_path / "ragged.txt"
    # dim n_cols precision bucket kmax m basis hermite uniform kernel
    path.write_text(f"2 1 0 1 30 6 3 0 0 {K_NSC}\n")  # (basis 3: PHS3)
    for load in (False, True):
        if load:
            RBF._lib.tune_load(path)
        op = RBF.laplacian(pv, basis=RBF.PHS(3, poly_deg=2), adjl=adjl)
        nb, kern = RBF._lib.last_build_buckets()
        if RBF._lib.last_build_stats()[2] == 2:  # one launch of the packed LDS kernel
            assert kern == [7, 7, 7], kern
        else:
            assert nb[1] == n and kern[1] != K_NSC, (nb, kern)
        W = op.weights()
        sel = rng.choice(n, 200, replace=False)
        aptr = np.concatenate([[0], np.cumsum([len(adjl[i]) for i in sel])])
        aidx = np.concatenate([adjl[i] for i in sel])
        ptr, col, wr, _, _ = O.build(pv, pv[sel], (aptr, aidx), O.basis_struct(3, 2),
                                     _op(O, 1, [(0, O.abi.OP_LAP, 0, 0, 1.0)]), scaled=True)
        for j, i in enumerate(sel):
            got = W.data[W.indptr[i]:W.indptr[i + 1]]
            ref = wr[0][ptr[j]:ptr[j + 1]]
            assert np.array_equal(W.indices[W.indptr[i]:W.indptr[i + 1]], col[ptr[j]:ptr[j + 1]])
            assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max(), i
    RBF._lib.tune_reset()
