"""The warm-start network's oracle (oracle/mlp.py) pinned on the CPU.

* bfloat16 rounding: known values, including ties to even;
* the loss gradient of every parameter kind against central differences of
  the unrounded network, so a dropped term, a wrong sign or a transposed
  operand in the backward formulas fails;
* the loss counts only the constraints a cluster has (its "present" flags).
"""
import numpy as np

import mlp


def test_bf16_rounding_known_values():
    x = np.float32([1.0, 1.0 + 2.0 ** -8, 1.0 + 3.0 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0, 3.0e38])
    want = [1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5, 0.0, 226.0 * 2.0 ** 120]
    got = mlp.bf16(x)
    for g, w in zip(got, want):
        assert g == np.float32(w) or abs(g - w) <= 1e-6 * abs(w)
    # ties go to the even mantissa: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert mlp.bf16(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0
    assert mlp.bf16(np.float32([1.0 + 2.0 ** -7 + 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


def _problem(seed=0, n=12):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 0.3, mlp.NPARAMS)
    X = rng.normal(0.0, 1.0, (n, mlp.F))
    X[:, list(mlp.PRESENT)] = (rng.uniform(size=(n, 4)) > 0.3).astype(float)
    T = rng.normal(0.0, 1.0, (n, 4))
    return theta, X, T


def test_gradients_match_central_differences():
    theta, X, T = _problem()
    loss, g = mlp.loss_and_gradients(X, T, rounding=False, **mlp.unpack(theta))
    G = mlp.pack(g)
    rng = np.random.default_rng(5)
    # a sample of every parameter block
    picks = []
    at = 0
    for name, shape in mlp.LAYOUT:
        size = int(np.prod(shape))
        picks += list(at + rng.choice(size, min(size, 12), replace=False))
        at += size
    eps = 1e-6
    for k in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        lp, _ = mlp.loss_and_gradients(X, T, rounding=False, **mlp.unpack(tp))
        lm, _ = mlp.loss_and_gradients(X, T, rounding=False, **mlp.unpack(tm))
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - G[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, G[k])


def test_absent_constraints_do_not_count():
    theta, X, T = _problem(1)
    X[:, mlp.PRESENT[3]] = 0.0  # no cluster has a fourth constraint
    T2 = T.copy()
    T2[:, 3] += 100.0
    l1, g1 = mlp.loss_and_gradients(X, T, **mlp.unpack(theta))
    l2, g2 = mlp.loss_and_gradients(X, T2, **mlp.unpack(theta))
    assert l1 == l2
    assert np.array_equal(mlp.pack(g1), mlp.pack(g2))
