"""GPU tests of the operator API built on the device: custom closures, the `@operator`
notation and Float32 data. They follow the reference's test/operators/custom.jl,
operator_algebra.jl:132-146 and test/float32.jl: same point sets (Halton), bases,
fields and error thresholds; exact equalities where two routes build the same terms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def halton(n, dim=2, start=1):
    """Points start..start+n-1 of the Halton sequence in bases 2, 3, 5."""
    out = np.zeros((n, dim))
    for d, base in enumerate((2, 3, 5)[:dim]):
        for row, i in enumerate(range(start, start + n)):
            f, r, k = 1.0, 0.0, i
            while k > 0:
                f /= base
                r += f * (k % base)
                k //= base
            out[row, d] = r
    return out


def mean_percent_error(test, correct):
    return float(np.mean(np.abs((test - correct) / correct)) * 100)


def f_test(x):
    return 1 + np.sin(4 * x[:, 0]) + np.cos(3 * x[:, 0]) + np.sin(2 * x[:, 1])


def d2f_dxx(x):
    return -16 * np.sin(4 * x[:, 0]) - 9 * np.cos(3 * x[:, 0])


def d2f_dyy(x):
    return -4 * np.sin(2 * x[:, 1])


def df_dx(x):
    return 4 * np.cos(4 * x[:, 0]) - 3 * np.sin(3 * x[:, 0])


def df_dy(x):
    return 2 * np.cos(2 * x[:, 1])


@pytest.fixture(scope="module")
def X10k():
    return halton(10_000)


def _dense(W):
    return W.toarray()


def test_custom_value_closure_equals_identity_weights(RBF, gpu):
    x = halton(1000)
    op = RBF.custom(x, lambda b: (lambda p, q: b(p, q)), rank=0)
    assert isinstance(op, RBF.RadialBasisOperator)
    ref = RBF.Identity()(x)
    assert np.array_equal(op.vals.cpu().numpy(), ref.vals.cpu().numpy())
    # keyword and positional basis, eval points, inferred rank (custom.jl testsets)
    op2 = RBF.custom(x, lambda b: (lambda p, q: b(p, q)), RBF.IMQ(1, poly_deg=2), rank=0)
    assert op2.basis == RBF.IMQ(1, poly_deg=2)
    x2 = halton(100, start=1001)
    op3 = RBF.custom(x, lambda b: (lambda p, q: b(p, q)), eval_points=x2, basis=RBF.PHS(5, poly_deg=3))
    assert op3.n_eval == 100 and op3.rank == 0
    # regrid of a cubic with PHS5 + cubic augmentation is exact
    g = x[:, 0] ** 3 - 2 * x[:, 0] * x[:, 1] ** 2 + 0.5
    assert np.allclose(op3(g), x2[:, 0] ** 3 - 2 * x2[:, 0] * x2[:, 1] ** 2 + 0.5, atol=1e-8)


def test_custom_rank1_gradient_like(RBF, gpu):
    # custom.jl "Custom rank-1 operator (gradient-like)"
    F = RBF.functors
    x = halton(1000)
    op = RBF.custom(x, lambda b: tuple(F.d(b, dim) for dim in (1, 2)), rank=1)
    W = op.weights()
    assert isinstance(W, tuple) and len(W) == 2
    jac = RBF.jacobian(x)
    assert np.array_equal(op.vals.cpu().numpy(), jac.vals.cpu().numpy())
    u = x[:, 0] ** 2 + 2 * x[:, 0] * x[:, 1]
    res = op(u)
    assert res.shape == (1000, 2)
    assert mean_percent_error(res[:, 0], 2 * x[:, 0] + 2 * x[:, 1]) < 10
    assert mean_percent_error(res[:, 1], 2 * x[:, 0]) < 10
    uv = np.stack([u, x[:, 1] ** 2], axis=1)
    rv = op(uv)
    assert rv.shape == (1000, 2, 2)
    assert mean_percent_error(rv[:, 0, 0], 2 * x[:, 0] + 2 * x[:, 1]) < 10
    assert np.abs(rv[:, 1, 0]).max() < 0.1
    assert mean_percent_error(rv[:, 1, 1], 2 * x[:, 1]) < 10


def test_custom_closure_combination_matches_algebra(RBF, gpu):
    """A closure written with the functors builds the same weights as the operator
    algebra it lowers to (same terms, same kernel): bit-identical."""
    F = RBF.functors
    x = halton(2000)
    b = RBF.PHS(5, poly_deg=3)
    op = RBF.custom(x, lambda bb: (lambda p, q: 2.0 * F.d2(bb, 1)(p, q) + 0.5 * F.d2(bb, 2)(p, q) - bb(p, q)),
                    basis=b)
    ref = RBF.RadialBasisOperator(2.0 * RBF.Partial(2, 1) + 0.5 * RBF.Partial(2, 2) - RBF.Identity(), x, basis=b)
    assert np.array_equal(op.vals.cpu().numpy(), ref.vals.cpu().numpy())


def test_operator_div_kappa_grad(RBF, gpu, X10k):
    x, b = X10k, RBF.PHS(5, poly_deg=3)
    op = RBF.custom(x, RBF.operator("∇ ⋅ (κ * ∇)", κ=[3.0, 0.5]), rank=0, basis=b)
    exact = 3.0 * d2f_dxx(x) + 0.5 * d2f_dyy(x)
    assert mean_percent_error(op(f_test(x)), exact) < 5


def test_operator_advection(RBF, gpu, X10k):
    x, b = X10k, RBF.PHS(5, poly_deg=3)
    op = RBF.custom(x, RBF.operator("c ⋅ ∇", c=[1.0, 0.5]), basis=b)
    exact = df_dx(x) + 0.5 * df_dy(x)
    assert mean_percent_error(op(f_test(x)), exact) < 5


def test_operator_composition(RBF, gpu, X10k):
    x, b = X10k, RBF.PHS(5, poly_deg=3)
    op = RBF.custom(x, RBF.operator("∇ ⋅ (κ * ∇) + k² * f", {"κ": 2.0, "k²": 1.0}), rank=0, basis=b)
    exact = 2.0 * (d2f_dxx(x) + d2f_dyy(x)) + f_test(x)
    assert mean_percent_error(op(f_test(x)), exact) < 5


def test_operator_subtraction_and_negation(RBF, gpu, X10k):
    x, b = X10k, RBF.PHS(5, poly_deg=3)
    y = 2 * x[:, 0] + 3 * x[:, 1]
    op = RBF.custom(x, RBF.operator("∂(1) - ∂(2)"), basis=b)
    assert mean_percent_error(op(y), np.full(len(x), -1.0)) < 5
    op_neg = RBF.custom(x, RBF.operator("-∂(1)"), basis=b)
    assert mean_percent_error(op_neg(y), np.full(len(x), -2.0)) < 5


def test_operator_functor_api(RBF, gpu, X10k):
    # custom.jl "AbstractOperator functor API"
    x, b = X10k, RBF.PHS(5, poly_deg=3)
    y = f_test(x)
    op = RBF.Custom(lambda bb: (lambda p, q: bb(p, q)))(x)
    assert isinstance(op, RBF.RadialBasisOperator)
    lap = RBF.operator("∇²")(x, basis=b)
    assert mean_percent_error(lap(y), d2f_dxx(x) + d2f_dyy(x)) < 5
    dx = RBF.operator("∂(1)")(x, basis=b)
    assert mean_percent_error(dx(y), df_dx(x)) < 5
    sc = RBF.operator("3 * ∇²")(x, basis=b)
    assert mean_percent_error(sc(y), 3 * (d2f_dxx(x) + d2f_dyy(x))) < 5
    s = RBF.operator("∂(1) + ∂(2)")(x, basis=b)
    assert mean_percent_error(s(y), df_dx(x) + df_dy(x)) < 5


def test_operator_tree_weights_match_hand_sum(RBF, gpu):
    # operator_algebra.jl:132-146
    pts = halton(400)
    k2 = 1.7
    op = RBF.operator("∇² + k2 * f", k2=k2)(pts)
    hand = _dense(RBF.laplacian(pts).weights()) + k2 * _dense(RBF.Identity()(pts).weights())
    assert np.allclose(_dense(op.weights()), hand, rtol=1e-10, atol=1e-10 * np.abs(hand).max())


def test_custom_hermite_constructs(RBF, gpu):
    # custom.jl "custom() Hermite Boundary Conditions": 1-D, Dirichlet at both ends
    dom = np.arange(0.0, 1.0 + 1e-12, 0.1)[:, None]
    n = len(dom)
    isb = np.zeros(n, dtype=bool)
    isb[0] = isb[-1] = True
    herm = {"is_boundary": isb, "bc": [RBF.Dirichlet(), RBF.Dirichlet()], "normals": np.array([[-1.0], [1.0]])}
    op = RBF.custom(dom, lambda b: (lambda p, q: b(p, q)), basis=RBF.PHS(3, poly_deg=2), hermite=herm, rank=0)
    assert isinstance(op, RBF.RadialBasisOperator)
    W = _dense(op.weights())
    assert W[0, 0] == 1.0 and W[-1, -1] == 1.0   # Dirichlet rows are identity rows


def test_float32_end_to_end_laplacian(RBF, gpu):
    # test/float32.jl:11-32
    import torch

    x64 = halton(200)
    x32 = x64.astype(np.float32)
    x64 = x32.astype(np.float64)
    b = RBF.PHS(3, poly_deg=2)
    op32 = RBF.laplacian(x32, basis=b)
    op64 = RBF.laplacian(x64, basis=b)
    op32.update_weights()
    op64.update_weights()
    W32, W64 = op32.weights(), op64.weights()
    assert W32.dtype == np.float32
    D32, D64 = _dense(W32).astype(np.float64), _dense(W64)
    assert np.linalg.norm(D32 - D64) <= np.sqrt(np.finfo(np.float32).eps) * max(np.linalg.norm(D32),
                                                                                  np.linalg.norm(D64))
    y32 = (np.sin(4 * x32[:, 0]) + np.cos(3 * x32[:, 1])).astype(np.float32)
    out = op32(y32)
    assert isinstance(out, np.ndarray) and out.dtype == np.float32
    # the applied weights are the returned FP32 weights
    assert np.allclose(out, (D32 @ y32.astype(np.float64)).astype(np.float32), rtol=1e-6, atol=1e-6)
    t_out = op32(torch.from_numpy(y32).to(gpu))
    assert t_out.dtype == torch.float32
    # algebra on a Float32 operator keeps FP32 weights
    s = op32 + op32
    assert s.weights().dtype == np.float32
    assert op64(y32.astype(np.float64)).dtype == np.float64
