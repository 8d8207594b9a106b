"""CPU tests of the host-side operator API: custom closures (custom.jl), the `@operator`
notation (operator_macro.jl) and the Float32 element-type detection (test/float32.jl).
They mirror the structural checks of the reference's test/operators/custom.jl and
operator_algebra.jl:132-146; the host closed forms of the basis functors are pinned
against the oracle, which is itself pinned in test_oracle.py."""
import numpy as np
import pytest

BASES = [("PHS1", 1.0), ("PHS3", 1.0), ("PHS5", 1.0), ("PHS7", 1.0), ("IMQ", 1.7), ("Gaussian", 0.9)]
KIND = {"PHS1": 1, "PHS3": 3, "PHS5": 5, "PHS7": 7, "IMQ": 10, "Gaussian": 11}


def _basis(RBF, name, eps):
    if name.startswith("PHS"):
        return RBF.PHS(int(name[3]), poly_deg=2)
    return (RBF.IMQ if name == "IMQ" else RBF.Gaussian)(eps, poly_deg=2)


# ----------------------------------------------------------------- custom closures
def test_custom_struct_and_printing(RBF):
    c = RBF.Custom(lambda b: b, rank=0)
    assert c(RBF.PHS(3, poly_deg=2)) == RBF.PHS(3, poly_deg=2)   # custom.jl "Custom struct"
    assert RBF.print_op(c) == "Custom Operator"


def test_infer_rank(RBF):
    F = RBF.functors
    assert F.infer_rank(lambda b: (lambda x, xi: b(x, xi))) == 0
    assert F.infer_rank(lambda b: tuple(F.d(b, i) for i in (1, 2))) == 1
    assert RBF.Custom(lambda b: (lambda x, xi: b(x, xi))).rank == 0
    assert RBF.Custom(lambda b: [F.d(b, 1), F.d(b, 2)]).rank == 1


def test_closure_lowering_matches_operator_algebra(RBF):
    F, b = RBF.functors, RBF.PHS(3, poly_deg=2)

    def same(closure, op, dim, rank=None):
        n1, t1 = RBF.Custom(closure, rank).lower(dim, b)
        n2, t2 = op.lower(dim, b)
        assert n1 == n2 and sorted(t1) == sorted(t2), (t1, t2)

    same(lambda bb: (lambda x, xi: bb(x, xi)), RBF.Identity(), 2)
    same(lambda bb: bb, RBF.Identity(), 3)
    same(lambda bb: F.d2(bb, 1), RBF.Partial(2, 1), 2)
    same(lambda bb: F.d(bb, 2), RBF.Partial(1, 2), 3)
    same(lambda bb: F.dmixed(bb, 2, 1), RBF.MixedPartial(1, 2), 2)
    same(lambda bb: F.lap(bb), RBF.Laplacian(), 2)
    same(lambda bb: (lambda x, xi: 2.0 * F.lap(bb)(x, xi) - bb(x, xi)), 2.0 * RBF.Laplacian() - RBF.Identity(), 2)
    same(lambda bb: (lambda x, xi: sum(np.float64(c) * F.d(bb, i + 1)(x, xi) for i, c in enumerate((1.0, 0.5)))),
         RBF.Partial(1, 1) + 0.5 * RBF.Partial(1, 2), 2)
    same(lambda bb: (lambda x, xi: F.d2(bb, 1)(x, xi) / 4 + (-F.d2(bb, 2)(x, xi))),
         0.25 * RBF.Partial(2, 1) - RBF.Partial(2, 2), 2)
    same(lambda bb: tuple(F.d(bb, i) for i in (1, 2)), RBF.Jacobian(2), 2)
    same(lambda bb: (F.d(bb, 1), lambda x, xi: F.d(bb, 2)(x, xi)), RBF.Jacobian(2), 2)


@pytest.mark.parametrize("closure, what", [
    (lambda b: (lambda x, xi: x[0] * b(x, xi)), "using the point x"),
    (lambda b: (lambda x, xi: b(x, xi) ** 2), "nonlinear function"),
    (lambda b: (lambda x, xi: np.sin(b(x, xi))), "failed on a basis term"),
    (lambda b: (lambda x, xi: b(x, xi) + 1.0), "adding a constant"),
    (lambda b: (lambda x, xi: b(x, xi) * b(x, xi)), "product of two basis terms"),
    (lambda b: (lambda x, xi: 1.0 / b(x, xi)), "nonlinear function"),
    (lambda b: 3.0, "must return a callable"),
])
def test_closure_lowering_refuses_nonlinear(RBF, closure, what):
    with pytest.raises(RBF.ArgumentError, match=what):
        RBF.Custom(closure, rank=0).lower(2, RBF.PHS(3, poly_deg=2))


def test_closure_rank_and_basis_checks(RBF):
    F, b = RBF.functors, RBF.PHS(3, poly_deg=2)
    with pytest.raises(RBF.ArgumentError, match="rank 0"):
        RBF.Custom(lambda bb: (F.d(bb, 1),), rank=0).lower(2, b)
    with pytest.raises(RBF.ArgumentError, match="rank 1"):
        RBF.Custom(lambda bb: F.d(bb, 1), rank=1).lower(2, b)
    with pytest.raises(RBF.ArgumentError, match="one build assembles one basis"):
        RBF.Custom(lambda bb: F.d(RBF.PHS(5, poly_deg=2), 1)).lower(2, b)
    with pytest.raises(RBF.ArgumentError, match="out of range"):
        RBF.Custom(lambda bb: F.d(bb, 3)).lower(2, b)
    with pytest.raises(RBF.ArgumentError, match="1-based"):
        F.d(b, 0)
    with pytest.raises(RBF.ArgumentError, match="vector-valued"):
        RBF.Custom(lambda bb: (lambda x, xi: F.grad(bb)(x, xi))).lower(2, b)
    with pytest.raises(RBF.ArgumentError):
        RBF.Custom(42)


@pytest.mark.parametrize("name, eps", BASES)
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_host_functors_match_oracle(RBF, O, name, eps, dim):
    """basis(x, xi), d, d2, dmixed, lap, grad and hess evaluated on the host equal the
    oracle's closed forms (pinned against the reference's goldens and finite differences)."""
    F, a = RBF.functors, O.abi
    b = _basis(RBF, name, eps)
    ob = O.basis_struct(KIND[name], 2, eps)
    rng = np.random.default_rng(dim * 7 + KIND[name])
    for _ in range(5):
        x, xi = rng.uniform(-1, 1, dim), rng.uniform(-1, 1, dim)
        assert b(x, xi) == pytest.approx(O.phi(ob, x, xi), rel=1e-13)
        assert F.lap(b)(x, xi) == pytest.approx(O.lrbf(ob, a.OP_LAP, x, xi), rel=1e-12, abs=1e-14)
        assert np.allclose(F.grad(b)(x, xi), O.grad(ob, x, xi), rtol=1e-12, atol=1e-14)
        if dim >= 2:
            assert np.allclose(F.hess(b)(x, xi), O.hess(ob, x, xi), rtol=1e-12, atol=1e-14)
        for i in range(dim):
            assert F.d(b, i + 1)(x, xi) == pytest.approx(O.lrbf(ob, a.OP_D1, x, xi, i), rel=1e-12, abs=1e-14)
            assert F.d2(b, i + 1)(x, xi) == pytest.approx(O.lrbf(ob, a.OP_D2, x, xi, i), rel=1e-12, abs=1e-14)
            for j in range(i + 1, dim):
                assert F.dmixed(b, i + 1, j + 1)(x, xi) == pytest.approx(
                    O.lrbf(ob, a.OP_MIXED, x, xi, i, j), rel=1e-12, abs=1e-14)


@pytest.mark.parametrize("name, eps", BASES)
def test_traced_closure_equals_numeric_closure(RBF, O, name, eps):
    """The term table a closure is lowered to, evaluated term by term with the oracle,
    equals the closure evaluated numerically: the tracing keeps the closure's meaning."""
    F, a = RBF.functors, O.abi
    b = _basis(RBF, name, eps)
    ob = O.basis_struct(KIND[name], 2, eps)

    def closure(bb):
        return lambda x, xi: (0.7 * F.lap(bb)(x, xi) - 2.0 * F.d(bb, 2)(x, xi) + F.dmixed(bb, 1, 3)(x, xi) / 3
                              + 1.5 * bb(x, xi) - F.d2(bb, 3)(x, xi))

    n, terms = RBF.Custom(closure).lower(3, b)
    assert n == 1 and len(terms) == 5
    rng = np.random.default_rng(5)
    for _ in range(5):
        x, xi = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        want = closure(b)(x, xi)
        got = sum(w * O.lrbf(ob, kind, x, xi, d1, d2) for _, kind, d1, d2, w in terms)
        assert got == pytest.approx(want, rel=1e-12)


def test_custom_constructor_keeps_operator_trees(RBF):
    # custom(data, op::AbstractOperator) builds op itself (custom.jl:62-65); a closure is wrapped
    lap = RBF.Laplacian()
    assert RBF.Custom(lap).lower(2) == lap.lower(2)
    assert RBF.Custom(lap, rank=0).rank == 0


# ----------------------------------------------------------------- @operator notation
def test_operator_symbol_variants(RBF):
    op = RBF.operator
    assert isinstance(op("∇²"), RBF.Laplacian)
    assert isinstance(op("Δ"), RBF.Laplacian)
    assert isinstance(op("∇ ⋅ ∇"), RBF.Laplacian)
    assert isinstance(op("I"), RBF.Identity)
    assert isinstance(op("f"), RBF.Identity)
    p = op("∂(1)")
    assert isinstance(p, RBF.Partial) and (p.order, p.dim) == (1, 1)
    p = op("∂²(2)")
    assert isinstance(p, RBF.Partial) and (p.order, p.dim) == (2, 2)


def test_operator_edge_cases(RBF):
    op = RBF.operator
    assert isinstance(op("3 * ∇²"), RBF.ScaledOperator)
    s = op("k^2 * f", k=2.0)
    assert isinstance(s, RBF.ScaledOperator) and s.α == pytest.approx(4.0)
    s = op("∇ ⋅ (a * b * ∇)", a=2.0, b=3.0)
    assert isinstance(s, RBF.ScaledOperator) and s.α == pytest.approx(6.0)
    s = op("∇² / 2")
    assert isinstance(s, RBF.ScaledOperator) and s.α == pytest.approx(0.5)
    assert isinstance(op("∂(1, 2)"), RBF.MixedPartial)
    with pytest.raises(RBF.ArgumentError, match="∂ expects 1 or 2 arguments, got 3"):
        op("∂(1, 2, 3)")
    with pytest.raises(RBF.ArgumentError, match="expects exactly 1 argument, got 2"):
        op("∂²(1, 2)")
    for bad in ["∇", "∇² + 1", "∇² * ∇²", "f / ∇²", "2 + 3", "∇²(", "k * ∇²"]:
        with pytest.raises(RBF.ArgumentError):
            op(bad)
    with pytest.raises(RBF.ArgumentError):
        op("∇² + __import__('os')")   # names are looked up, never executed


def test_operator_trees_are_structure_preserving(RBF):
    # operator_algebra.jl:132-140
    helm = RBF.operator("∇² + k2 * f", k2=1.7)
    assert isinstance(helm, RBF.SumOperator) and len(helm.ops) == 2
    assert isinstance(helm.ops[1], RBF.ScaledOperator)
    assert isinstance(RBF.operator("2.0 * ∂²(1) + 0.5 * ∂²(2)"), RBF.SumOperator)
    assert isinstance(RBF.operator("∇² + k² * f", {"k²": 1.0}), RBF.SumOperator)


def test_operator_expansions_lower_to_terms(RBF):
    a, op = RBF._abi, RBF.operator

    def terms(o, dim=2):
        return sorted((k, d1, d2, float(w)) for _, k, d1, d2, w in o.lower(dim)[1])

    assert terms(op("∇ ⋅ (κ * ∇)", κ=[3.0, 0.5])) == [(a.OP_D2, 0, 0, 3.0), (a.OP_D2, 1, 0, 0.5)]
    assert terms(op("∇ ⋅ (κ * ∇)", κ=2.0)) == [(a.OP_LAP, 0, 0, 2.0)]
    assert terms(op("c ⋅ ∇", c=[1.0, 0.5])) == [(a.OP_D1, 0, 0, 1.0), (a.OP_D1, 1, 0, 0.5)]
    assert terms(op("ν * ∇² - c ⋅ ∇", ν=0.1, c=[1.0, 0.5])) == sorted(
        [(a.OP_LAP, 0, 0, 0.1), (a.OP_D1, 0, 0, -1.0), (a.OP_D1, 1, 0, -0.5)])
    assert terms(op("∇ ⋅ (κ * ∇) + k² * f", {"κ": 2.0, "k²": 1.0})) == sorted(
        [(a.OP_LAP, 0, 0, 2.0), (a.OP_VALUE, 0, 0, 1.0)])
    assert terms(op("∂(1) - ∂(2)")) == [(a.OP_D1, 0, 0, 1.0), (a.OP_D1, 1, 0, -1.0)]
    assert terms(op("-∂(1)")) == [(a.OP_D1, 0, 0, -1.0)]
    assert terms(op("2∇² + κ[2] * f", κ=[1.0, 5.0])) == sorted([(a.OP_LAP, 0, 0, 2.0), (a.OP_VALUE, 0, 0, 5.0)])
    assert terms(op("sqrt(4) * ∇² + [1, 2] ⋅ ∇")) == sorted(
        [(a.OP_LAP, 0, 0, 2.0), (a.OP_D1, 0, 0, 1.0), (a.OP_D1, 1, 0, 2.0)])
    assert terms(op("1.5e-1 * ∇²")) == [(a.OP_LAP, 0, 0, 0.15)]


# ----------------------------------------------------------------- Float32 element type
def test_float32_element_type_detection(RBF):
    import torch

    ft = RBF.operators._float_type
    x32 = np.random.default_rng(0).random((10, 2)).astype(np.float32)
    assert ft(x32) == "float32"
    assert ft(x32.astype(np.float64)) == "float64"
    assert ft(torch.from_numpy(x32)) == "float32"
    assert ft([p for p in x32]) == "float32"
    assert ft([[0.1, 0.2]]) == "float64"
