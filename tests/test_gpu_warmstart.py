"""The learned SHAKE warm start on the GPU.

* The device network's outputs and loss gradients (BF16 matrix-core
  forward and backward kernels) match the float64 oracle with the same
  rounding points, on random parameters and inputs -- which also pins the
  matrix cores' lane layout the kernels assume.
* Recording, training and attaching: the trained network cuts the SHAKE
  passes per cluster-step against plain SHAKE from the same state, while
  every constraint still holds to the tolerance (plain SHAKE finishes the
  job and judges the prediction).
"""
import numpy as np
import pytest

import mlp
from mdcore import native, synthetic
from mdcore.engine import HipEngine, RunOptions, minimise

pytestmark = pytest.mark.gpu


def _random_problem(n=300, seed=4):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0.0, 0.25, mlp.NPARAMS).astype(np.float32)
    X = rng.normal(0.0, 1.0, (n, mlp.F)).astype(np.float32)
    X[:, list(mlp.PRESENT)] = (rng.uniform(size=(n, 4)) > 0.3).astype(np.float32)
    T = rng.normal(0.0, 1.0, (n, 4)).astype(np.float32)
    return theta, X, T


def test_network_matches_oracle():
    ws = native.NativeWarmStart(seed=1, capacity=16)
    theta, X, T = _random_problem()
    ws.set_params(theta)
    assert np.array_equal(ws.params(), theta)
    loss, y, g = ws.evaluate(X, T)
    P = mlp.unpack(theta.astype(np.float64))
    want_y, _ = mlp.forward(X, **P)
    want_loss, want_g = mlp.loss_and_gradients(X, T, **P)
    G = mlp.pack(want_g)
    yerr = np.max(np.abs(y - want_y)) / np.max(np.abs(want_y))
    gerr = np.sqrt(np.sum((g - G) ** 2) / np.sum(G ** 2))
    print("outputs max rel err", yerr, "loss", loss, want_loss, "gradient rel RMS", gerr)
    assert yerr < 1e-2
    assert abs(loss - want_loss) < 1e-2 * abs(want_loss)
    assert gerr < 2e-2
    # every parameter block, not just the total
    at = 0
    for name, shape in mlp.LAYOUT:
        size = int(np.prod(shape))
        a, b = g[at:at + size], G[at:at + size]
        if np.sum(b ** 2) > 0:
            assert np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)) < 3e-2, name
        at += size
    ws.close()


@pytest.fixture(scope="module")
def chain_system():
    b = synthetic.build(600, 2000, 2, "cube", 314)
    e = HipEngine(b.system, b.state, RunOptions(seed=7))
    minimise(e, 100, 10.0)
    e.setVelocitiesToTemperature(300.0, 7)
    e.run(100)
    out = (b, e.positions(), e.velocities(), e.steps)
    e.destroy()
    return out


def _fresh(b, x, v, step, seed=7):
    e = HipEngine(b.system, b.state, RunOptions(seed=seed))
    e.setPositions(x)
    e.setVelocities(v)
    e._native.step = step
    return e


def test_trained_warm_start_saves_shake_passes(chain_system):
    b, x, v, step = chain_system
    ws = native.NativeWarmStart(seed=2, capacity=400_000)
    rec = _fresh(b, x, v, step, seed=11)
    ws.record(rec._native, 300)
    n = ws.samples
    print("recorded samples", n)
    assert n > 10_000
    first = ws.train(1, 4096, 1e-3)
    last = ws.train(1500, 4096, 1e-3)
    print("training loss", first, "->", last)
    assert last < 0.5 * first
    rec.destroy()

    plain = _fresh(b, x, v, step)
    learned = _fresh(b, x, v, step)
    ws.attach(learned._native)
    p0, c0 = native.shake_counts(plain._native)
    q0, d0 = native.shake_counts(learned._native)
    plain.run(200)
    learned.run(200)
    p1, c1 = native.shake_counts(plain._native)
    q1, d1 = native.shake_counts(learned._native)
    plain_rate = (p1 - p0) / (c1 - c0)
    learned_rate = (q1 - q0) / (d1 - d0)
    print(f"SHAKE passes per cluster-step: plain {plain_rate:.3f}, warm-started {learned_rate:.3f}")
    assert learned_rate < 0.8 * plain_rate
    s = b.system
    for e in (plain, learned):
        xx = e.positions().reshape(-1, 3)
        d = np.linalg.norm(xx[s.constraintP1] - xx[s.constraintP2], axis=1)
        assert np.max(np.abs(d - s.constraintDistance) / s.constraintDistance) < 2e-6
        assert np.all(np.isfinite(xx))
    native.NativeWarmStart.detach(learned._native)
    plain.destroy()
    learned.destroy()
    ws.close()


def test_config4_warm_start_on_a_million_samples():
    """Config 4 (250k atoms, 128^3 mesh, about 10^4 H-bond SHAKE clusters):
    10^6 samples recorded from the engine's own SHAKE at 2 fs, the network
    trained, then the same 2 fs steps with and without it.  The warm start
    must cut the SHAKE passes per cluster-step and keep every constraint at
    the tolerance; the integrator's time per step with and without it is
    printed (profiled phase by phase), not gated.  The generated system is
    equilibrated for 600 steps of 0.5 fs before its 2 fs steps: after 300,
    SHAKE diverged near step 340 in one run of four (the runs then differed
    in the last bits of the minimised positions)."""
    b = synthetic.config(4)
    s = b.system
    eq = HipEngine(s, b.state, RunOptions(timeStep=0.0005, friction=10.0, seed=5))
    eq.minimise(200, 10.0)
    eq.setVelocitiesToTemperature(300.0, 5)
    eq.run(600)
    x, v, step = eq.positions(), eq.velocities(), eq.steps
    eq.destroy()

    ws = native.NativeWarmStart(seed=3, capacity=1_100_000)
    rec = _fresh(b, x, v, step, seed=13)
    clusters = None
    while ws.samples < 1_000_000:
        ws.record(rec._native, 20)
        if clusters is None:
            clusters = ws.samples / 20.0
    n = ws.samples
    rec.destroy()
    first = ws.train(1, 4096, 1e-3)
    last = ws.train(1500, 4096, 1e-3)
    print(f"config 4: {s.particleCount} atoms, ~{clusters:.0f} SHAKE clusters, {n} samples, "
          f"loss {first:.3g} -> {last:.3g}")
    assert n >= 1_000_000
    assert last < 0.5 * first

    plain = _fresh(b, x, v, step)
    learned = _fresh(b, x, v, step)
    ws.attach(learned._native)
    rates, integrate = {}, {}
    for name, e in (("plain", plain), ("learned", learned)):
        p0, c0 = native.shake_counts(e._native)
        e.run(100)
        p1, c1 = native.shake_counts(e._native)
        rates[name] = (p1 - p0) / (c1 - c0)
        integrate[name] = e.profileStep(20)["integrate"]
    print(f"config 4 SHAKE passes per cluster-step: plain {rates['plain']:.3f}, learned {rates['learned']:.3f}; "
          f"integrate phase ms: plain {integrate['plain']:.4f}, learned {integrate['learned']:.4f}")
    assert rates["learned"] < 0.8 * rates["plain"]
    for e in (plain, learned):
        xx = e.positions().reshape(-1, 3)
        d = np.linalg.norm(xx[s.constraintP1] - xx[s.constraintP2], axis=1)
        assert np.max(np.abs(d - s.constraintDistance) / s.constraintDistance) < 2e-6
        e.destroy()
