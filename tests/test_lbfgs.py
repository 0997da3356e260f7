"""The host minimiser's L-BFGS (mdcore.engine.lbfgs) on functions whose
minima are known in closed form, as the reference pins its own search
(test/lbfgs.test.ts upstream): a diagonal quadratic, Rosenbrock's valley from
the classic start, the same under a small per-step displacement cap, and the
stopping rules (already converged; iteration limit)."""
import numpy as np

from mdcore.engine import lbfgs


def _rosenbrock(z):
    x, y = z
    value = (1 - x) ** 2 + 100 * (y - x * x) ** 2
    grad = np.array([-2 * (1 - x) - 400 * x * (y - x * x), 200 * (y - x * x)])
    return value, grad


def test_quadratic_minimum():
    A = np.diag(np.arange(1.0, 11.0))
    b = np.arange(10.0)
    x, _, it = lbfgs(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(10), 0, 1e-10, 1, max_move=1e9)
    assert np.max(np.abs(x - b / np.arange(1.0, 11.0))) < 1e-7
    assert 0 < it < 40


def test_rosenbrock_from_the_classic_start():
    x, value, _ = lbfgs(_rosenbrock, np.array([-1.2, 1.0]), 0, 1e-9, 1, max_move=1e9)
    assert np.allclose(x, [1.0, 1.0], atol=1e-8)
    assert value < 1e-16


def test_displacement_cap_slows_but_does_not_stop_the_search():
    for cap in (0.05, 0.01):
        seen = []

        def f(z):
            seen.append(np.array(z, copy=True))
            return _rosenbrock(z)

        x, _, _ = lbfgs(f, np.array([-1.2, 1.0]), 0, 1e-9, 1, max_move=cap)
        assert np.allclose(x, [1.0, 1.0], atol=1e-8)
        # no evaluated point is further than the cap from the one before it
        # that the search accepted; in particular none is far from the start
        steps = np.abs(np.diff(np.array(seen), axis=0)).max(axis=1)
        assert np.all(steps <= 2 * cap + 1e-12)


def test_stopping_rules():
    x0 = np.array([1.0, 1.0])
    x, _, it = lbfgs(_rosenbrock, x0, 0, 1e-6, 1)
    assert it == 0 and np.array_equal(x, x0)
    _, _, it = lbfgs(_rosenbrock, np.array([-1.2, 1.0]), 5, 1e-12, 1, max_move=1e9)
    assert it == 5
