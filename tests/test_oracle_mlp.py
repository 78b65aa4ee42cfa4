"""Pins the FP64 oracle of the constitutive-surrogate MLP (no reference
counterpart): central finite differences of the loss for every parameter
block, closed forms of degenerate networks, and the bfloat16 rounding
emulation against exactly representable cases."""
import numpy as np
import pytest

import pyoracle

NP = pyoracle.MLP_NPARAMS
OFF = {"w1": (0, 1024), "b1": (1024, 1088), "w2": (1088, 5184), "b2": (5184, 5248), "w3": (5248, 6272),
       "b3": (6272, 6288)}


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP)
    w1 = rng.uniform(-1, 1, (64, 16)) * np.sqrt(6 / 13) * scale
    w1[:, 13:] = 0.0
    p[0:1024] = w1.ravel()
    p[1024:1088] = rng.uniform(-0.1, 0.1, 64)
    p[1088:5184] = rng.uniform(-1, 1, 4096) * np.sqrt(6 / 64) * scale
    p[5184:5248] = rng.uniform(-0.1, 0.1, 64)
    w3 = rng.uniform(-1, 1, (16, 64)) * np.sqrt(3 / 64) * scale
    w3[7:] = 0.0
    p[5248:6272] = w3.ravel()
    p[6272:6279] = rng.uniform(-0.1, 0.1, 7)
    return p


def rand_scaling(seed):
    rng = np.random.default_rng(seed + 100)
    sc = np.zeros(64)
    sc[0:13] = rng.normal(0, 1e-3, 13)
    sc[16:29] = 1.0 / rng.uniform(1e-3, 2e-3, 13)
    sc[32:39] = rng.normal(0, 10.0, 7)
    sc[48:55] = rng.uniform(5.0, 20.0, 7)
    sc[55:64] = 1.0
    return sc


def rand_rows(seed, n, sc):
    rng = np.random.default_rng(seed + 200)
    x = np.zeros((n, 16), np.float32)
    x[:, :13] = (sc[0:13] + rng.normal(0, 1, (n, 13)) / sc[16:29]).astype(np.float32)
    t = np.zeros((n, 16), np.float32)
    t[:, :7] = (sc[32:39] + rng.normal(0, 1, (n, 7)) * sc[48:55]).astype(np.float32)
    return x, t


def test_gradient_matches_central_differences():
    p, sc = rand_params(1), rand_scaling(1)
    x, t = rand_rows(1, 24, sc)
    loss, g = pyoracle.mlp_grad(p, sc, x, t, bf16=False)
    assert loss > 0
    rng = np.random.default_rng(7)
    eps = 1e-6
    for name, (a, b) in OFF.items():
        idx = rng.choice(np.arange(a, b), size=12, replace=False)
        if name == "w1":  # include live input columns only (padded ones have zero gradient)
            idx = np.array([a + 16 * j + k for j, k in zip(rng.integers(0, 64, 12), rng.integers(0, 13, 12))])
        if name in ("w3", "b3"):
            idx = np.array([i for i in idx if ((i - a) // (64 if name == "w3" else 1)) < 7] or [a])
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += eps
            pm[i] -= eps
            fd = (pyoracle.mlp_grad(pp, sc, x, t, bf16=False)[0] - pyoracle.mlp_grad(pm, sc, x, t, bf16=False)[0]) / (
                2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[i])


def test_padded_outputs_and_inputs_get_zero_gradient():
    p, sc = rand_params(2), rand_scaling(2)
    x, t = rand_rows(2, 16, sc)
    _, g = pyoracle.mlp_grad(p, sc, x, t, bf16=False)
    w1 = g[0:1024].reshape(64, 16)
    w3 = g[5248:6272].reshape(16, 64)
    assert np.all(w1[:, 13:] == 0.0)
    assert np.all(w3[7:] == 0.0) and np.all(g[6279:6288] == 0.0)


def test_zero_weights_output_the_bias():
    sc = rand_scaling(3)
    p = np.zeros(NP)
    p[6272:6279] = np.arange(1, 8) * 0.25
    x, _ = rand_rows(3, 5, sc)
    y = pyoracle.mlp_forward(p, sc, x, bf16=False)
    expect = np.arange(1, 8) * 0.25 * sc[48:55] + sc[32:39]
    assert np.allclose(y[:, :7], expect, rtol=0, atol=1e-12)
    assert np.all(y[:, 7:] == 0.0)


def test_single_relu_path_is_a_product_of_scalars():
    sc = np.zeros(64)
    sc[16:29] = 1.0
    sc[48:64] = 1.0
    p = np.zeros(NP)
    a, b, c = 0.5, -1.5, 2.0
    p[0 * 16 + 0] = a          # w1[0][0]
    p[1088 + 0 * 64 + 0] = b   # w2[0][0]
    p[5248 + 0 * 64 + 0] = c   # w3[0][0]
    x = np.zeros((3, 16), np.float32)
    x[:, 0] = [2.0, -2.0, 0.0]
    y = pyoracle.mlp_forward(p, sc, x, bf16=False)
    # h1 = relu(a x), h2 = relu(b h1) = 0 for b < 0: output 0
    assert np.all(y[:, 0] == 0.0)
    p[1088] = -b
    y = pyoracle.mlp_forward(p, sc, x, bf16=False)
    assert np.allclose(y[:, 0], [c * (-b) * a * 2.0, 0.0, 0.0])


def test_bf16_emulation_exact_on_representable_values():
    # weights, inputs and every partial sum are small dyadic numbers: the
    # bf16 network must agree with the exact one bit for bit
    rng = np.random.default_rng(5)
    p = np.zeros(NP)
    p[0:1024] = rng.integers(-4, 5, 1024) / 8.0
    p[0:1024].reshape(64, 16)[:, 13:] = 0.0
    p[1088:5184] = rng.integers(-2, 3, 4096) / 16.0
    p[5248:6272] = rng.integers(-2, 3, 1024) / 16.0
    p[1024:1088] = rng.integers(-2, 3, 64) / 4.0
    sc = np.zeros(64)
    sc[16:29] = 1.0
    sc[48:64] = 1.0
    x = np.zeros((7, 16), np.float32)
    x[:, :13] = rng.integers(-3, 4, (7, 13)) / 2.0
    ye = pyoracle.mlp_forward(p, sc, x, bf16=False)
    yb = pyoracle.mlp_forward(p, sc, x, bf16=True)
    # hidden sums can exceed 8 significant bits: compare with bf16 resolution
    assert np.allclose(yb, ye, rtol=2 ** -7, atol=2 ** -7)
    # with inputs that are exactly representable and a single layer-1 unit
    # the two agree exactly
    q = np.zeros(NP)
    q[0] = 0.75
    q[1088] = 1.0
    q[5248] = 1.0
    assert np.array_equal(pyoracle.mlp_forward(q, sc, x, bf16=True), pyoracle.mlp_forward(q, sc, x, bf16=False))


def test_bf16_gradient_close_to_exact():
    p, sc = rand_params(4), rand_scaling(4)
    x, t = rand_rows(4, 200, sc)
    le, ge = pyoracle.mlp_grad(p, sc, x, t, bf16=False)
    lb, gb = pyoracle.mlp_grad(p, sc, x, t, bf16=True)
    assert abs(lb - le) <= 2e-2 * le
    for name, (a, b) in OFF.items():
        n = np.linalg.norm(ge[a:b])
        if n > 0:
            assert np.linalg.norm(gb[a:b] - ge[a:b]) <= 5e-2 * n, name


@pytest.mark.parametrize("seed", [11, 12])
def test_loss_is_half_mean_squared_error(seed):
    p, sc = rand_params(seed), rand_scaling(seed)
    x, t = rand_rows(seed, 10, sc)
    y = pyoracle.mlp_forward(p, sc, x, bf16=False)
    yn = (y[:, :7] - sc[32:39]) / sc[48:55]
    tn = (t[:, :7].astype(np.float64) - sc[32:39]) / sc[48:55]
    loss, _ = pyoracle.mlp_grad(p, sc, x, t, bf16=False)
    assert np.isclose(loss, 0.5 * np.sum((yn - tn) ** 2) / len(x), rtol=1e-12)
