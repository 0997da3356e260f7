"""The pair kernel's erfc approximation against the committed mpmath table (CPU).

The kernel evaluates erfc(z) = exp(-z^2) P(t), t = 1/(1 + z/2), with the
degree-8 coefficients ``MD_ERFC_C0..C8`` defined in ``csrc/nonbonded.hip``.
The coefficients are read from that source file, so this pins the constants
the kernel is built with, not a copy of them.  Bounds:

* the approximation itself (evaluated in f64): within 1.2e-7 relative of
  erfc on [0, 4], the bound upstream sets for its own kernel approximation
  (``test/erf.test.ts:68-89``), at every mpmath table row
  (``test/erf.test.ts:13-25``) and on a dense grid against scipy;
* the f32 evaluation (the kernel's arithmetic, f32 coefficients, f32 exp):
  within 1e-6 relative on [0, 4] -- there the f32 rounding of z^2 inside
  exp(-z^2), amplified by z^2, dominates.
"""
import json
import os
import re

import numpy as np
from scipy.special import erfc

from conftest import GOLDEN, ROOT

SRC = os.path.join(ROOT, "making-it-rain-webgpu_amd", "csrc", "nonbonded.hip")


def _coefficients():
    text = open(SRC).read()
    found = dict((int(k), float(v)) for k, v in re.findall(r"#define MD_ERFC_C(\d)\s+([-+0-9.eE]+)f", text))
    assert sorted(found) == list(range(9)), found
    return np.array([found[k] for k in range(9)], dtype=np.float32)


def _approx_f64(z):
    c = _coefficients().astype(np.float64)
    t = 1.0 / (1.0 + 0.5 * z)
    return np.exp(-z * z) * np.polyval(c[::-1], t)


def _approx_f32(z):
    c = _coefficients()
    z = np.asarray(z, dtype=np.float32)
    t = (np.float32(1) / (np.float32(1) + np.float32(0.5) * z)).astype(np.float32)
    p = np.full_like(t, c[8])
    for k in range(7, -1, -1):
        p = (p * t + c[k]).astype(np.float32)  # the kernel's fmaf chain, rounded per step
    return (np.exp(-(z * z).astype(np.float32)).astype(np.float32) * p).astype(np.float32)


def test_approximation_meets_the_upstream_bound_at_the_mpmath_rows():
    table = json.load(open(os.path.join(GOLDEN, "known_answers.json")))["erf"]
    for z, _erf, erfc_exact in table:
        if z > 4:
            continue
        got = float(_approx_f64(np.array([z]))[0])
        assert abs(got - erfc_exact) <= 1.2e-7 * erfc_exact, (z, got, erfc_exact)


def test_approximation_meets_the_upstream_bound_on_a_dense_grid():
    z = np.linspace(0.0, 4.0, 200001)
    rel = np.abs(_approx_f64(z) - erfc(z)) / erfc(z)
    assert rel.max() <= 1.2e-7, rel.max()


def test_f32_evaluation_within_its_rounding_bound():
    z = np.linspace(0.0, 4.0, 200001)
    zz = z.astype(np.float32).astype(np.float64)
    rel = np.abs(_approx_f32(z).astype(np.float64) - erfc(zz)) / erfc(zz)
    assert rel.max() <= 1e-6, rel.max()
    # at the configurations' alpha * cutoff (units.ts:53-54), the value the
    # pair term takes at the cutoff
    z_c = 2.628260884878466
    assert abs(float(_approx_f32(np.array([z_c]))[0]) - 0.0002016645187154009) <= 1e-6 * 0.0002016645187154009
