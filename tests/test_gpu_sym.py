"""GPU parity of the packed symmetric LDS build (kernel id 7, k_build_sym): the rook
LDL^T of k_build_wide with only the lower triangle of the saddle-point system in LDS
and the pivots moved by symmetric row-and-column swaps. Every test here pins it for
all three size buckets and re-runs a parity test of test_gpu_build through it: the
weights against the extended-precision oracle per stencil at max(1e-10, kappa_2(A) u),
Hermite rows against the reference's fixtures, ragged stencils, the fused 4-column 3-D
build, singular systems, and bit-identical repeated builds."""
import contextlib

import numpy as np
import pytest

import test_gpu_build as B

pytestmark = pytest.mark.gpu
K_SYM = 7


@contextlib.contextmanager
def pinned_sym(RBF):
    for b in range(3):
        RBF._lib.tune_bucket(b, K_SYM)
    try:
        yield
    finally:
        for b in range(3):
            RBF._lib.tune_bucket(b, -1)


def _ran_sym(RBF):
    n, kern = RBF._lib.last_build_buckets()
    used = [int(k) for c, k in zip(n, kern) if c != 0]
    assert used and all(k == K_SYM for k in used), (list(n), list(kern))


@pytest.mark.parametrize("name,kind,eps", B.BASES)
@pytest.mark.parametrize("dim,deg,k", [(2, 2, 20), (3, 1, 16), (2, 0, 12), (1, 2, 8), (3, 3, 40), (2, 3, 30)])
def test_sym_build_parity_bases(RBF, O, gpu, name, kind, eps, dim, deg, k):
    with pinned_sym(RBF):
        B.test_build_parity_bases(RBF, O, gpu, name, kind, eps, dim, deg, k)
        _ran_sym(RBF)


@pytest.mark.parametrize("which", ["d1", "d2", "mixed", "regrid", "grad", "hess3", "sum"])
def test_sym_build_parity_operators(RBF, O, gpu, which):
    with pinned_sym(RBF):
        B.test_build_parity_operators(RBF, O, gpu, which)
        _ran_sym(RBF)


def test_sym_variable_size_stencils(RBF, O, gpu):
    with pinned_sym(RBF):
        B.test_variable_size_stencils(RBF, O, gpu)
        _ran_sym(RBF)


@pytest.mark.parametrize("case", range(4))
def test_sym_hermite_domain12(RBF, O, gpu, case):
    with pinned_sym(RBF):
        B.test_hermite_domain12(RBF, O, gpu, case)
        _ran_sym(RBF)


def test_sym_hermite_manufactured_441(RBF, O, gpu):
    with pinned_sym(RBF):
        B.test_hermite_manufactured_441(RBF, O, gpu)
        _ran_sym(RBF)


def test_sym_dirichlet_eval_row_without_boundary_neighbours(RBF, O, gpu):
    with pinned_sym(RBF):
        B.test_dirichlet_eval_row_without_boundary_neighbours(RBF, O, gpu)


def test_sym_singular_stencil_raises(RBF, gpu):
    with pinned_sym(RBF):
        B.test_singular_stencil_raises(RBF, gpu)


def test_sym_c2_shape_3d_gradient_laplacian(RBF, O, gpu):
    with pinned_sym(RBF):
        B.test_c2_shape_3d_gradient_laplacian(RBF, O, gpu)
        _ran_sym(RBF)


def test_sym_c3_shape_disk_hermite_variable_k(RBF, O, gpu):
    with pinned_sym(RBF):
        B.test_c3_shape_disk_hermite_variable_k(RBF, O, gpu)
        _ran_sym(RBF)


def test_sym_repeated_builds_bit_identical(RBF, gpu):
    with pinned_sym(RBF):
        B.test_repeated_builds_bit_identical(RBF, gpu)
        _ran_sym(RBF)


def test_sym_matches_full_storage_kernel(RBF, gpu):
    """The packed and the full-storage LDS kernels solve the same systems with the same
    pivot rule; their weights agree to rounding (the packed update keeps one of the two
    mirrored entries, so the two differ by a few ulps times the stencil's conditioning)
    on a 3-D PHS5 degree-3 Laplacian (n = 60) and a 2-D build of 70-row systems."""
    rng = np.random.default_rng(5)
    cases = [(rng.random((3000, 3)), RBF.PHS(5, poly_deg=3), 40), (rng.random((3000, 2)), RBF.PHS(3, poly_deg=3), 60)]
    for pts, basis, k in cases:
        ws = []
        for kern in (2, K_SYM):
            for b in (1, 2):
                RBF._lib.tune_bucket(b, kern)
            try:
                op = RBF.laplacian(pts, basis=basis, k=k)
                assert max(RBF._lib.last_build_buckets()[1]) == kern
            finally:
                for b in (1, 2):
                    RBF._lib.tune_bucket(b, -1)
            ws.append(op.vals.cpu().numpy()[0].reshape(-1, k))
        d = np.abs(ws[0] - ws[1]).max(1) / np.abs(ws[0]).max(1)
        assert np.median(d) <= 1e-10 and np.quantile(d, 0.99) <= 1e-6, (k, np.median(d), np.quantile(d, 0.99), d.max())


@pytest.mark.parametrize("dim,k,deg,which", [(2, 60, 3, "grad"), (3, 50, 3, "lap"), (3, 60, 2, "grad")])
def test_sym_two_slots_multi_column(RBF, O, gpu, dim, k, deg, which):
    """Systems of 70 rows (the second lane slot) with one or several right-hand sides, in
    2-D and 3-D: k_build_sym<2,4,2>, <3,1,2> and <3,4,2>, against the extended-precision
    oracle per stencil."""
    rng = np.random.default_rng(60 + dim * k + deg)
    pts = rng.random((1500, dim))
    b = RBF.PHS(5, poly_deg=deg) if dim == 3 else RBF.PHS(3, poly_deg=deg)
    if which == "grad":
        L, nc = RBF.Jacobian(dim), dim
        terms = [(c, O.abi.OP_D1, c, 0, 1.0) for c in range(dim)]
    else:
        L, nc = RBF.Laplacian(), 1
        terms = [(0, O.abi.OP_LAP, 0, 0, 1.0)]
    with pinned_sym(RBF):
        op = RBF.RadialBasisOperator(L, pts, basis=b, k=k)
        _ran_sym(RBF)
        assert RBF._lib.last_build_buckets()[0][2] == len(pts)  # all in the 65..80-row bucket
    adj = op.adj_idx.cpu().numpy()
    B.check_against_truth(O, op.vals.cpu().numpy(), pts, pts, adj, O.basis_struct(b.kind, b.poly_deg),
                          B._op(O, nc, terms), k)
