"""CPU pins of the surrogate-network oracle (oracle/mlp_oracle.py) and of the
invariance the canonical stencil frame relies on.

- backward pass of the float64 network against central finite differences of its loss;
- tanh-GELU derivative against finite differences;
- BF16 rounding against exactly representable values and ties-to-even cases;
- Adam against a hand-computed first step;
- canonical offsets: invariant under rotating, reflecting and scaling the stencil;
- the property the canonical frame needs: the plain Laplacian weights of a stencil
  (oracle build, reference arithmetic) do not change under rotation or reflection.
"""
import numpy as np
import pytest

import mlp_oracle as MO


def small_params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    p = np.zeros(MO.N_PARAMS)
    for l in range(MO.LAYERS):
        n, k = MO.n_of(l), MO.k_of(l)
        p[MO.w_off(l):MO.w_off(l) + n * k] = rng.uniform(-scale, scale, n * k)
        p[MO.b_off(l):MO.b_off(l) + n] = rng.uniform(-0.1, 0.1, n)
    return p


def test_gelu_grad_matches_differences():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (MO.gelu(z + h) - MO.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - MO.gelu_grad(z))) < 1e-8
    # tanh form against its limits: gelu(z) -> z for large z, -> 0 for very negative z
    assert abs(MO.gelu(10.0) - 10.0) < 1e-12 and abs(MO.gelu(-10.0)) < 1e-12


def test_bf16_rounding():
    exact = np.array([1.0, -2.0, 0.5, 1.0 + 2 ** -7, 3.0 * 2 ** -20])
    assert np.array_equal(MO.bf16(exact), exact)
    # halfway between 1 and 1 + 2^-7 rounds to even (1.0); just above rounds up
    assert MO.bf16(1.0 + 2 ** -8) == 1.0
    assert MO.bf16(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    # halfway between 1+2^-7 (odd mantissa) and 1+2^-6 rounds up to even
    assert MO.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = small_params(1)
    M, n_out = 6, 26
    X = np.zeros((M, MO.IN_PAD))
    X[:, :80] = rng.uniform(-1, 1, (M, 80))
    T = rng.standard_normal((M, MO.OUT_PAD))
    rw = rng.uniform(0.5, 2.0, M)
    loss, g = MO.loss_and_grad(p, X, T, n_out, bf16_round=False, row_weight=rw)
    # check a spread of coordinates in every layer (weights and biases)
    idx = []
    for l in range(MO.LAYERS):
        n, k = MO.n_of(l), MO.k_of(l)
        fan_in = 80 if l == 0 else k
        rows = n_out if l == MO.LAYERS - 1 else n
        for _ in range(6):
            idx.append(MO.w_off(l) + rng.integers(rows) * k + rng.integers(fan_in))
        idx.append(MO.b_off(l) + rng.integers(rows))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (MO.loss_and_grad(pp, X, T, n_out, False, rw)[0] - MO.loss_and_grad(pm, X, T, n_out, False, rw)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded input columns and padded output rows carry no gradient
    W0 = g[MO.w_off(0):MO.w_off(0) + 256 * 128].reshape(256, 128)
    assert np.all(W0[:, 80:] == 0)
    W3 = g[MO.w_off(3):MO.w_off(3) + 32 * 256].reshape(32, 256)
    assert np.all(W3[n_out:] == 0)


def test_adam_first_step():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.1, -0.3, 0.0])
    pn, m, v = MO.adam(p, g, np.zeros(3), np.zeros(3), 1e-2, 0.9, 0.999, 1e-8, 1)
    # first step: m^ = g, v^ = g^2 -> update lr * g / (|g| + eps)
    expect = p - 1e-2 * g / (np.abs(g) + 1e-8)
    assert np.allclose(pn, expect, rtol=0, atol=1e-15)


def _stencil(seed, k=20):
    rng = np.random.default_rng(seed)
    pts = rng.random((400, 2))
    e = 123
    d = ((pts - pts[e]) ** 2).sum(1)
    adj = np.argsort(d, kind="stable")[:k]
    return pts, e, adj


def test_canonical_offsets_invariant():
    pts, e, adj = _stencil(5)
    xi, h, r1 = MO.canonical_offsets(pts, pts[[e]], adj[None])
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    for T in (R, np.diag([1.0, -1.0]) @ R, np.diag([-1.0, 1.0])):
        q = 3.0 * (pts - pts[e]) @ T.T + np.array([5.0, -2.0])
        xi2, h2, r12 = MO.canonical_offsets(q, q[[e]], adj[None])
        assert np.allclose(xi2, xi, atol=1e-12)
        assert np.isclose(r12[0], r1[0], rtol=1e-12)
        assert np.isclose(h2[0], 3.0 * h[0], rtol=1e-12)
    # the first nonzero offset lies on +x, and the y offsets sum to >= 0
    assert xi[0, 0, 0] == 0 and xi[0, 0, 1] == 0
    assert abs(xi[0, 1, 1]) < 1e-15 and xi[0, 1, 0] > 0
    assert xi[0, :, 1].sum() >= 0
    assert np.isclose(np.sqrt((xi[0] ** 2).sum(1)).max(), 1.0)


def test_laplacian_weights_rotation_reflection_invariant(O):
    """The canonical frame is sound only because PHS + polynomial Laplacian weights are
    invariant under rotations and reflections of the stencil (radial basis, rotation-
    invariant polynomial space, L P(x_e) invariant). Checked on the oracle build."""
    pts, e, adj = _stencil(9)
    lap = O.abi.make_operator(1, [(0, O.abi.OP_LAP, 0, 0, 1.0)])
    _, _, w0, _, rc = O.build(pts, pts[[e]], adj[None].astype(np.int32), O.basis_struct(3, 2), lap, scaled=True)
    assert rc == 0
    for th, refl in ((0.3, False), (2.1, True), (np.pi / 2, False)):
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        if refl:
            R = R @ np.diag([1.0, -1.0])
        q = (pts - pts[e]) @ R.T + pts[e]
        _, _, w1, _, rc = O.build(q, q[[e]], adj[None].astype(np.int32), O.basis_struct(3, 2), lap, scaled=True)
        assert rc == 0
        assert np.max(np.abs(w1 - w0)) <= 1e-10 * np.max(np.abs(w0))
