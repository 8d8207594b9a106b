"""CPU tests that pin the oracle (oracle/rbf_oracle.c) before it is trusted as the
checker: the reference's own golden values and invariants, closed forms, finite
differences and brute force on small inputs."""
import json
import math
import pathlib

import numpy as np
import pytest

GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_goldens.json").read_text())


def _basis(O, name, eps=1.0, deg=-1):
    kinds = {"PHS1": 1, "PHS3": 3, "PHS5": 5, "PHS7": 7, "IMQ": 10, "Gaussian": 11}
    return O.basis_struct(kinds[name], deg, eps)


# ----------------------------------------------------------------- basis goldens
def test_basis_derivative_goldens(O):
    g = GOLD["basis_derivatives"]
    x1, x2 = g["x1"], g["x2"]
    for c in g["cases"]:
        b = _basis(O, c["basis"], c.get("eps", 1.0))
        assert O.lrbf(b, O.abi.OP_D1, x1, x2, 0) == pytest.approx(c["d1"], rel=1e-12), c["src"]
        assert np.allclose(O.grad(b, x1, x2), c["grad"], rtol=1e-12), c["src"]
        assert O.lrbf(b, O.abi.OP_D2, x1, x2, 0) == pytest.approx(c["d2"], rel=1e-12), c["src"]
        r = math.dist(x1, x2)
        if c["basis"].startswith("PHS"):
            assert O.phi(b, x1, x2) == pytest.approx(r ** int(c["basis"][3]), rel=1e-14)


def test_basis_value_goldens(O):
    # phi itself for every basis, at the points of the reference's "Distances" testsets
    # (PHS r^p, IMQ 1/sqrt((eps r)^2 + 1) = 1/sqrt(21), Gaussian exp(-(eps r)^2) = e^-20)
    g = GOLD["basis_values"]
    for c in g["cases"]:
        b = _basis(O, c["basis"], c.get("eps", 1.0))
        assert O.phi(b, g["x1"], g["x2"]) == pytest.approx(c["value"], rel=1e-14), c["src"]
    # the shape parameter enters as (eps r)^2, not eps r^2: at eps = 2 the two differ
    imq = _basis(O, "IMQ", 2.0)
    assert O.phi(imq, [0.0, 0.0], [0.3, 0.4]) == pytest.approx(1.0 / math.sqrt(2.0), rel=1e-15)
    gau = _basis(O, "Gaussian", 2.0)
    assert O.phi(gau, [0.0, 0.0], [0.3, 0.4]) == pytest.approx(math.exp(-1.0), rel=1e-15)


def _fd_grad(f, x, h=1e-6):
    x = np.asarray(x, float)
    g = np.zeros_like(x)
    for d in range(x.size):
        e = np.zeros_like(x)
        e[d] = h
        g[d] = (f(x + e) - f(x - e)) / (2 * h)
    return g


@pytest.mark.parametrize("name", ["PHS1", "PHS3", "PHS5", "PHS7", "IMQ", "Gaussian"])
def test_hessian_matches_finite_differences(O, name):
    # polyharmonic_spline.jl:74-87 checks H against ForwardDiff; central differences here
    b = _basis(O, name, 2.0 if name in ("IMQ", "Gaussian") else 1.0)
    x1, x2 = np.array([1.0, 2.0]), np.array([1.3, 1.4])
    H = O.hess(b, x1, x2)
    Hfd = np.array([_fd_grad(lambda x: O.grad(b, x, x2)[i], x1) for i in range(2)])
    assert np.allclose(H, Hfd, rtol=1e-6, atol=1e-8)
    assert np.allclose(H, H.T, rtol=1e-12)
    # mixed partial = Hessian entry (basis.jl:87-92)
    assert O.lrbf(b, O.abi.OP_MIXED, x1, x2, 0, 1) == pytest.approx(H[0, 1], rel=1e-13)


@pytest.mark.parametrize("name", ["PHS1", "PHS3", "PHS5", "PHS7", "IMQ", "Gaussian"])
@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_is_derivative_of_phi(O, name, dim):
    # links phi (pinned by its closed form above) to the derivative code (pinned by the
    # reference's derivative goldens): a wrong phi or a wrong derivative fails here
    b = _basis(O, name, 1.7 if name in ("IMQ", "Gaussian") else 1.0)
    x1 = np.array([0.3, -0.2, 0.5])[:dim]
    x2 = np.array([-0.1, 0.4, 0.2])[:dim]
    fd = _fd_grad(lambda x: O.phi(b, x, x2), x1)
    assert np.allclose(O.grad(b, x1, x2), fd, rtol=1e-7, atol=1e-9)
    for d in range(dim):
        assert O.lrbf(b, O.abi.OP_D1, x1, x2, d) == pytest.approx(fd[d], rel=1e-7, abs=1e-9)


@pytest.mark.parametrize("name", ["IMQ", "Gaussian"])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_imq_gaussian_laplacian(O, name, dim):
    # inverse_multiquadric.jl:77-83 and gaussian.jl:160-163 sum d^2/dx_d^2 over the
    # dimensions, so (unlike PHS) the Laplacian is the true one in every dimension:
    # it equals a finite-difference Laplacian of phi and the trace of the Hessian,
    # and each d^2 term equals the Hessian diagonal
    for eps in (0.8, 2.5):
        b = _basis(O, name, eps)
        x1 = np.array([0.3, -0.2, 0.5])[:dim]
        x2 = np.array([-0.1, 0.4, 0.2])[:dim]
        h = 1e-4
        lap_fd = 0.0
        for d in range(dim):
            e = np.zeros(dim)
            e[d] = h
            lap_fd += (O.phi(b, x1 + e, x2) - 2 * O.phi(b, x1, x2) + O.phi(b, x1 - e, x2)) / h ** 2
        lap = O.lrbf(b, O.abi.OP_LAP, x1, x2)
        assert lap == pytest.approx(lap_fd, rel=1e-5)
        H = O.hess(b, x1, x2)
        assert lap == pytest.approx(np.trace(H), rel=1e-13)
        for d in range(dim):
            assert O.lrbf(b, O.abi.OP_D2, x1, x2, d) == pytest.approx(H[d, d], rel=1e-13, abs=1e-15)


def test_phs_laplacian_is_hessian_trace_in_3d(O):
    # in 3-D the reference's PHS constants are the true Laplacian: lap = trace(H)
    for name in ("PHS1", "PHS3", "PHS5", "PHS7"):
        b = _basis(O, name)
        x1, x2 = np.array([0.3, -0.2, 0.5]), np.array([-0.1, 0.4, 0.2])
        assert O.lrbf(b, O.abi.OP_LAP, x1, x2) == pytest.approx(np.trace(O.hess(b, x1, x2)), rel=1e-13)


@pytest.mark.parametrize("name", ["PHS1", "PHS3", "PHS5", "PHS7"])
@pytest.mark.parametrize("kind,d1", [(1, 0), (1, 1), (2, 0), (2, 1), (4, 0)])
def test_hermite_normal_forms_are_directional_derivatives(O, name, kind, d1):
    # the reference checks ∂(x1, x2, n) against FD.gradient(x_2 -> ∂(x1, x_2)) . n
    # (test/basis/polyharmonic_spline.jl:48-72)
    b = _basis(O, name)
    x1, x2 = np.array([1.0, 2.0]), np.array([2.0, 4.0])
    n = np.array([1.0, 1.0]) / math.sqrt(2)
    v = O.lrbf_normal(b, kind, x1, x2, n, d1)
    fd = _fd_grad(lambda y: O.lrbf(b, kind, x1, y, d1), x2) @ n
    assert v == pytest.approx(fd, rel=1e-6)


def test_normal_forms_absent_for_imq_gaussian(O):
    # api.jl:382-397: only PHS implements the normal form
    for name in ("IMQ", "Gaussian"):
        assert O.lrbf_normal(_basis(O, name, 2.0), 4, [1.0, 2.0], [2.0, 4.0], [1.0, 0.0]) is None


def test_phs_laplacian_uses_reference_constants(O):
    # polyharmonic_spline.jl:202-205, 288-291, 373-376, 459-462: 2/r, 12r, 30r^3, 56r^5
    # are the 3-D Laplacians of r, r^3, r^5, r^7 and the reference uses them in every
    # dimension. Pin: the 3-D value equals a finite-difference Laplacian, and the 2-D
    # value equals the 3-D value of the embedded points (not the true 2-D Laplacian).
    for name, p in (("PHS1", 1), ("PHS3", 3), ("PHS5", 5), ("PHS7", 7)):
        b = _basis(O, name)
        x3, y3 = np.array([0.3, -0.2, 0.5]), np.array([-0.1, 0.4, 0.2])
        h = 1e-4
        lap_fd = 0.0
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            lap_fd += (O.phi(b, x3 + e, y3) - 2 * O.phi(b, x3, y3) + O.phi(b, x3 - e, y3)) / h ** 2
        assert O.lrbf(b, 4, x3, y3) == pytest.approx(lap_fd, rel=1e-5)
        x2, y2 = np.array([0.3, -0.2]), np.array([-0.1, 0.4])
        assert O.lrbf(b, 4, x2, y2) == pytest.approx(O.lrbf(b, 4, np.r_[x2, 0], np.r_[y2, 0]), rel=1e-14)
        r = math.dist(x2, y2)
        true2d = p * p * r ** (p - 2)
        assert abs(O.lrbf(b, 4, x2, y2) - true2d) > 1e-3 * abs(true2d)


# ----------------------------------------------------------------- monomials
def test_monomial_goldens(O):
    for c in GOLD["monomials"]:
        dim, deg, x = c["dim"], c["deg"], c["x"]
        if dim > 3:
            # the reference's constant basis in 4-D (monomial.jl:206-207)
            continue
        assert np.allclose(O.mon(dim, deg, x), c["val"]), c["src"]
        for d in range(dim):
            assert np.allclose(O.lmon(dim, deg, 1, x, d), c["d1"][d]), c["src"]
            assert np.allclose(O.lmon(dim, deg, 2, x, d), c["d2"][d]), c["src"]
        assert np.allclose(O.lmon(dim, deg, 4, x), c["lap"]), c["src"]


@pytest.mark.parametrize("dim,deg", [(2, 2), (2, 3), (3, 2), (3, 3), (2, 4)])
def test_monomial_derivatives_match_finite_differences(O, dim, deg):
    x = np.linspace(0.3, 0.9, dim)
    for d in range(dim):
        fd = _fd_grad(lambda y: O.mon(dim, deg, y), x) if False else None
        h = 1e-5
        e = np.zeros(dim)
        e[d] = h
        fd1 = (O.mon(dim, deg, x + e) - O.mon(dim, deg, x - e)) / (2 * h)
        assert np.allclose(O.lmon(dim, deg, 1, x, d), fd1, atol=1e-7)
        fd2 = (O.mon(dim, deg, x + e) - 2 * O.mon(dim, deg, x) + O.mon(dim, deg, x - e)) / h ** 2
        assert np.allclose(O.lmon(dim, deg, 2, x, d), fd2, atol=2e-3)
    if dim >= 2:
        h = 1e-4
        e0, e1 = np.eye(dim)[0] * h, np.eye(dim)[1] * h
        fdm = (O.mon(dim, deg, x + e0 + e1) - O.mon(dim, deg, x + e0 - e1) - O.mon(dim, deg, x - e0 + e1)
               + O.mon(dim, deg, x - e0 - e1)) / (4 * h * h)
        assert np.allclose(O.lmon(dim, deg, 3, x, 0, 1), fdm, atol=1e-5)


def test_hermite_poly_entries_1d(O):
    # compute_hermite_poly_entry!: alpha*P + beta*d_n P (hermite_stencil.jl:108-158)
    g = GOLD["hermite_poly_entries_1d"]
    for c in g["cases"]:
        x = g["points"][c["point"]]
        P = O.mon(1, 1, x)
        if c["boundary"] and not (c["alpha"] == 1.0 and c["beta"] == 0.0):
            P = c["alpha"] * P + c["beta"] * O.mon_normal(1, 1, c["normal"], x)
        assert np.allclose(P, c["expected"]), g["src"]


# ----------------------------------------------------------------- solver
def test_sym_solve_matches_dense_solvers(O):
    rng = np.random.default_rng(1)
    for n in (5, 26, 60):
        A = rng.standard_normal((n, n))
        A = A + A.T
        A[: n // 3, : n // 3] = 0.0  # zero block, like the saddle system
        B = rng.standard_normal((n, 3))
        X = O.sym_solve(A, B)
        assert np.allclose(A @ X, B, atol=1e-10 * np.abs(A).max() * np.abs(X).max())
        assert np.allclose(X, np.linalg.solve(A, B), rtol=1e-8, atol=1e-10)
        Xl = O.sym_solve(A, B, long_double=True)
        assert np.allclose(X, Xl, rtol=1e-8, atol=1e-10)


def test_sym_solve_zero_diagonal_needs_pivoting(O):
    # PHS collocation has phi(0) = 0 on the diagonal: 2x2 pivots are required
    A = np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [2.0, 3.0, 0.0]])
    b = np.array([1.0, 2.0, 3.0])
    assert np.allclose(O.sym_solve(A, b), np.linalg.solve(A, b), rtol=1e-13)


def test_sym_solve_singular_raises(O):
    A = np.zeros((3, 3))
    with pytest.raises(np.linalg.LinAlgError):
        O.sym_solve(A, np.ones(3))


# ----------------------------------------------------------------- k-NN
def test_knn_small_goldens(O):
    g = GOLD["knn_small"]
    pts = np.array(g["points"])
    adj = O.knn(pts, pts, g["k"])
    assert all(adj[i, 0] == i for i in range(len(pts)))
    assert adj[0, 1] in g["second_of_0_in"]
    assert adj[g["farthest_excluded"]["query"], 1] != g["farthest_excluded"]["excluded"]
    ev = g["eval"]
    assert O.knn(pts, np.array(ev["points"]), ev["k"])[0, 0] == ev["first"]


def test_knn_brute_force_matches_kdtree(O):
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(3)
    for dim in (2, 3):
        pts = rng.random((3000, dim))
        q = rng.random((500, dim))
        adj = O.knn(pts, q, 20)
        _, ref = cKDTree(pts).query(q, 20)
        assert all(set(a) == set(r) for a, r in zip(adj, ref))


# ----------------------------------------------------------------- builds
def _lap_op(O):
    return O.abi.make_operator(1, [(0, O.abi.OP_LAP, 0, 0, 1.0)])


def _poly_reproduction_residual(O, pts, ev, ptr, col, val, dim, deg, kind, d1=0, d2=0):
    worst = 0.0
    for e in range(len(ev)):
        idx = col[ptr[e]:ptr[e + 1]]
        w = val[ptr[e]:ptr[e + 1]]
        P = np.array([O.mon(dim, deg, pts[i]) for i in idx])
        LP = O.lmon(dim, deg, kind, ev[e], d1, d2)
        worst = max(worst, np.abs(P.T @ w - LP).max() / max(1.0, np.abs(LP).max()))
    return worst


def test_build_polynomial_reproduction(O):
    # the weights reproduce L on the augmenting polynomials exactly (assembly.jl:107-112)
    rng = np.random.default_rng(5)
    pts = rng.random((400, 2))
    adj = O.knn(pts, pts, 20)
    b = O.basis_struct(3, 2)
    for kind, d1 in ((4, 0), (1, 0), (2, 1)):
        op = O.abi.make_operator(1, [(0, kind, d1, 0, 1.0)])
        for scaled in (False, True):
            ptr, col, val, st, rc = O.build(pts, pts, adj, b, op, scaled=scaled)
            assert rc == 0 and (st == 0).all()
            assert _poly_reproduction_residual(O, pts, pts, ptr, col, val[0], 2, 2, kind, d1) < 1e-9


def test_scaled_frame_matches_reference_arithmetic(O):
    # the centred, radius-scaled system has the same weights in exact arithmetic
    rng = np.random.default_rng(7)
    pts = rng.random((1500, 2))
    adj = O.knn(pts, pts, 20)
    b = O.basis_struct(3, 2)
    _, _, v0, _, _ = O.build(pts, pts, adj, b, _lap_op(O), scaled=False)
    _, _, v1, _, _ = O.build(pts, pts, adj, b, _lap_op(O), scaled=True)
    w0, w1 = v0[0].reshape(-1, 20), v1[0].reshape(-1, 20)
    rel = np.abs(w0 - w1).max(1) / np.abs(w0).max(1)
    assert np.median(rel) < 1e-11
    assert rel.max() < 1e-8


def test_phs_weight_homogeneity(O):
    # w(x_c + h xi) = h^-order w(xi) for PHS: the reference's unscaled arithmetic
    rng = np.random.default_rng(9)
    xi = rng.random((25, 2)) - 0.5
    adj = np.arange(25, dtype=np.int32)[None, :]
    b = O.basis_struct(3, 2)
    for kind, order in ((4, 2), (1, 1), (0, 0)):
        op = O.abi.make_operator(1, [(0, kind, 0, 0, 1.0)])
        _, _, v1, _, _ = O.build(xi, xi[:1], adj, b, op)
        h, c = 0.01, np.array([3.0, -1.0])
        _, _, v2, _, _ = O.build(c + h * xi, (c + h * xi)[:1], adj, b, op)
        ref = v1[0] * h ** -order
        # the uncentred arithmetic loses digits to the offset (cond grows with |c|/h)
        assert np.abs(v2[0] - ref).max() <= 1e-7 * np.abs(ref).max()
        _, _, v3, _, _ = O.build(c + h * xi, (c + h * xi)[:1], adj, b, op, scaled=True)
        assert np.abs(v3[0] - ref).max() <= 1e-10 * np.abs(ref).max()


def test_regrid_reproduces_linear_field(O):
    g = GOLD["regrid_linear"]
    H = GOLD["hermite_domain_12"]
    pts = np.array(H["points"])
    tg = np.array(g["targets"])
    adj = O.knn(pts, tg, 12)
    op = O.abi.make_operator(1, [(0, O.abi.OP_VALUE, 0, 0, 1.0)])
    ptr, col, val, st, rc = O.build(pts, tg, adj, O.basis_struct(3, 2), op)
    a, bx, by = g["field"]
    f = a + bx * pts[:, 0] + by * pts[:, 1]
    out = O.spmv(ptr, col, val[0], f)
    assert np.allclose(out, a + bx * tg[:, 0] + by * tg[:, 1], atol=g["atol"]), g["src"]


@pytest.mark.parametrize("case", range(4))
def test_hermite_exact_polynomial_reproduction(O, case):
    H = GOLD["hermite_domain_12"]
    c = H["cases"][case]
    pts = np.array(H["points"])
    isb = np.array(H["is_boundary"])
    al = np.array([bc[0] for bc in c["bcs"]], float)
    be = np.array([bc[1] for bc in c["bcs"]], float)
    adj = O.knn(pts, pts, 12)
    ptr, col, val, st, rc = O.build(pts, pts, adj, O.basis_struct(3, 2), _lap_op(O),
                                    hermite=(isb, al, be, np.array(H["normals"])))
    assert rc == 0
    u = pts[:, 0] ** 2 + pts[:, 1] ** 2 if c["field"] == "x2+y2" else pts[:, 1] ** 2
    res = O.spmv(ptr, col, val[0], u)
    assert np.abs(res[~isb] - c["interior_laplacian"]).max() < c["tol"], H["src"]
    if c["bcs"][0] == [1, 0]:
        assert res[0] == pytest.approx(u[0])  # Dirichlet identity row


def _unit_square_441():
    # test/solve/integration/hermite_test_utils.jl:7-97 (spacing 0.05, no jitter)
    xs = np.round(np.arange(0, 1.0 + 1e-12, 0.05), 12)
    pts = np.array([[x, y] for x in xs for y in xs])
    isb = (pts[:, 0] == 0) | (pts[:, 0] == 1) | (pts[:, 1] == 0) | (pts[:, 1] == 1)
    normals, bcs = [], []
    for p in pts[isb]:
        n = np.zeros(2)
        n[0] = -1 if p[0] == 0 else (1 if p[0] == 1 else 0)
        n[1] = -1 if p[1] == 0 else (1 if p[1] == 1 else 0)
        normals.append(n / np.linalg.norm(n))
        if p[1] in (0.0, 1.0):
            bcs.append((1.0, 0.0))
        elif p[0] == 0.0:
            bcs.append((0.5, 0.5))
        else:
            bcs.append((0.0, 1.0))
    return pts, isb, np.array(normals), np.array(bcs)


def test_laplacian_end_to_end_manufactured(O):
    # test/solve/integration/laplacian_end_to_end.jl:45-78: u = x^2 + y^2 + 1, Lu = 4
    pts, isb, nrm, bcs = _unit_square_441()
    assert len(pts) == 441 and isb.sum() == 80
    adj = O.knn(pts, pts, 15)
    ptr, col, val, st, rc = O.build(pts, pts, adj, O.basis_struct(3, 2), _lap_op(O),
                                    hermite=(isb, bcs[:, 0], bcs[:, 1], nrm))
    assert rc == 0
    u = pts[:, 0] ** 2 + pts[:, 1] ** 2 + 1
    g = 2 * pts[isb]
    un = (g * nrm).sum(1)
    ub = np.where((bcs[:, 0] == 1) & (bcs[:, 1] == 0), u[isb], bcs[:, 0] * u[isb] + bcs[:, 1] * un)
    uu = u.copy()
    uu[isb] = ub
    res = O.spmv(ptr, col, val[0], uu)
    assert np.abs(res[~isb] - 4.0).max() < 1e-10
    assert np.abs(res[isb] - ub).max() < 1e-10  # boundary rows apply the BC operator


def test_assembled_system_goldens(O):
    # _build_collocation_matrix! on three points, PHS3 with degree 1 (test/solve.jl:7-31):
    # A[i,j] = r_ij^3 in the RBF block and rows [1, x, y] in the polynomial block, with
    # the m x m zero block; the system condition_numbers() measures is this matrix
    import ctypes as C

    g = GOLD["collocation_small"]
    pts = np.array(g["points"])
    ev = np.array([g["eval_point"]])
    op = O.abi.make_operator(1, [(0, O.abi.OP_D1, 0, 0, 1.0)])
    a, keep = O.build(pts, ev, np.arange(3, dtype=np.int32)[None, :], O.basis_struct(3, 1), op, _want_args=True)
    A = np.zeros(36)
    n = O.lib().orc_assemble(C.byref(a), 0, 0, A.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 6
    A = A.reshape(6, 6)
    for i in range(3):
        for j in range(3):
            assert A[i, j] == pytest.approx(math.dist(pts[i], pts[j]) ** 3, rel=1e-14, abs=0), g["src"]
        assert np.allclose(A[i, 3:], [1.0, pts[i, 0], pts[i, 1]]), g["src"]
        assert np.allclose(A[3:, i], [1.0, pts[i, 0], pts[i, 1]]), g["src"]
    assert not A[3:, 3:].any()
    kap = O.condition_numbers(pts, ev, np.arange(3, dtype=np.int32)[None, :], O.basis_struct(3, 1), op,
                              scaled=False)
    s = np.linalg.svd(A, compute_uv=False)
    assert kap[0] == pytest.approx(s[0] / s[-1], rel=1e-10)


def _dirichlet_adjl_without_boundary(O, pts, isb):
    """adjl where each Dirichlet eval node's stencil holds only interior points, the
    other rows are the 10 nearest neighbours (which include boundary points)."""
    adj = O.knn(pts, pts, 10)
    inner = np.nonzero(~isb)[0][:10].astype(np.int32)
    for i in np.nonzero(isb)[0]:
        adj[i] = inner
    return adj


def test_dirichlet_eval_row_without_boundary_neighbours(O):
    # count_nonzeros gives a Dirichlet eval node one entry whatever its stencil holds;
    # the build writes its identity row there and leaves every other row intact
    H = GOLD["hermite_domain_12"]
    pts = np.array(H["points"])
    isb = np.array(H["is_boundary"])
    adj = _dirichlet_adjl_without_boundary(O, pts, isb)
    herm = (isb, np.array([1.0, 1.0]), np.array([0.0, 0.0]), np.array(H["normals"]))
    ptr, col, val, st, rc = O.build(pts, pts, adj, O.basis_struct(3, 2), _lap_op(O), hermite=herm)
    assert rc == 0
    for i in np.nonzero(isb)[0]:
        assert ptr[i + 1] - ptr[i] == 1 and col[ptr[i]] == i and val[0, ptr[i]] == 1.0 and st[i] == 1
    u = pts[:, 1] ** 2
    res = O.spmv(ptr, col, val[0], u)
    assert np.abs(res[~isb] - 2.0).max() < 1e-10
    assert np.allclose(res[isb], u[isb])


def test_count_nonzeros_dirichlet_rows(O):
    # count_nonzeros (execution.jl:39-69): Dirichlet eval rows hold one entry
    import ctypes as C

    isb = np.array([1, 0, 1, 0], dtype=np.uint8)
    bidx = np.array([0, -1, 1, -1], dtype=np.int32)
    al, be = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    ptr = np.zeros(5, dtype=np.int64)
    nnz = O.lib().orc_count_nonzeros(4, None, 3, isb.ctypes.data, bidx.ctypes.data, al.ctypes.data,
                                     be.ctypes.data, ptr.ctypes.data_as(C.POINTER(C.c_int64)))
    assert nnz == 1 + 3 + 3 + 3
    assert list(ptr) == [0, 1, 4, 7, 10]


def test_abi_struct_sizes_match_header(O):
    import ctypes as C

    assert C.sizeof(O.abi.BuildArgs) == O.lib().orc_sizeof_build_args()
    assert C.sizeof(O.abi.SpmvArgs) == O.lib().orc_sizeof_spmv_args()
    assert C.sizeof(O.abi.Operator) == O.lib().orc_sizeof_operator()
