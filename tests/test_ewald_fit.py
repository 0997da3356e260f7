"""The pair kernel's Coulomb polynomials (md_probe_ewald_fit, CPU).

The kernel takes the real-space Ewald force and energy per charge product as

    F/r = 1/r^3 + sum_k force[k] v^k,   E = 1/r + sum_k energy[k] v^k,
    v = scale r^2 - 1,

in place of erfc(alpha r)/r^3 + (2 alpha/sqrt(pi)) exp(-alpha^2 r^2)/r^2 and
erfc(alpha r)/r (the screened term of src/gpu/kernels/nonbonded-pairs.wgsl:
203-212).  Checked here against scipy's erfc in double, the polynomials
evaluated as the kernel does (f32 coefficients, v by one fused multiply-add
from the f32 r^2, Horner by fused multiply-adds, each rounded once to f32):

* at the configurations' alpha (Ewald tolerance 5e-4, cutoff 1 nm: alpha rc
  = 2.628, units.ts:53-54) both errors stay within 1e-6 of the bare 1/r^3 and
  1/r at every r up to the cutoff, and the fit is taken (ok);
* the same at tolerance 1e-4 and at another cutoff with the same alpha rc;
* at tolerance 1e-6 (alpha rc = 3.7) the degree-12 fit misses the bound,
  and the probe says so (the kernel then evaluates erfc and exp);
* the probe's own error figures agree with this independent evaluation;
* bad arguments are refused.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

from mdcore import native


def _fma32(a, b, c):
    # a*b of two f32 is exact in f64; one rounding of a*b + c to f64 then
    # f32 differs from a single rounding only in double-rounding ties
    f64 = np.float64
    return (np.asarray(a, f64) * np.asarray(b, f64) + np.asarray(c, f64)).astype(np.float32)


def _horner(c, v):
    p = np.full_like(v, c[-1])
    for k in range(len(c) - 2, -1, -1):
        p = _fma32(p, v, np.float32(c[k]))
    return p.astype(np.float64)


def _errors(fit, alpha, cutoff, n=20001):
    r2 = (np.linspace(0.0, cutoff, n)[1:] ** 2).astype(np.float32)
    v = _fma32(r2, fit["scale"], np.float32(-1.0)).astype(np.float32)
    r = np.sqrt(r2.astype(np.float64))
    ar = alpha * r
    # exact pieces beyond the bare Coulomb terms
    force = (erfc(ar) / r + 2 * alpha / math.sqrt(math.pi) * np.exp(-ar * ar)) / r**2 - 1 / r**3
    energy = erfc(ar) / r - 1 / r
    ef = np.abs(_horner(fit["force"], v) - force) * r**3
    ee = np.abs(_horner(fit["energy"], v) - energy) * r
    return ef.max(), ee.max()


def _alpha(tol, cutoff):
    return math.sqrt(-math.log(2 * tol)) / cutoff


@pytest.mark.parametrize("tol,cutoff", [(5e-4, 1.0), (1e-4, 1.0), (5e-4, 1.2), (5e-4, 0.9)])
def test_fit_within_bound(tol, cutoff):
    alpha = _alpha(tol, cutoff)
    fit = native.probe_ewald_fit(alpha, cutoff)
    ef, ee = _errors(fit, alpha, cutoff)
    print(f"tol {tol} cutoff {cutoff}: force {ef:.2e} energy {ee:.2e} (probe {fit['force_err']:.2e} "
          f"{fit['energy_err']:.2e})")
    assert fit["ok"]
    assert ef <= 1e-6 and ee <= 1e-6
    # the probe's figures come from its own sampling; near the same maximum
    assert abs(ef - fit["force_err"]) <= 0.3 * ef
    assert abs(ee - fit["energy_err"]) <= 0.3 * ee
    assert fit["scale"] == np.float32(2.0 / (cutoff * cutoff * (1 + 1e-4)))


def test_fit_at_the_configurations_alpha_is_near_f32_rounding():
    alpha = 2.628260884878466
    fit = native.probe_ewald_fit(alpha, 1.0)
    ef, ee = _errors(fit, alpha, 1.0)
    assert ef < 4e-7 and ee < 2e-7, (ef, ee)
    # near r = 0 the force piece tends to -(4/3) alpha^3 / sqrt(pi) and the
    # energy piece to -2 alpha / sqrt(pi) (the Taylor limits of erf)
    assert abs(_horner(fit["force"], np.array([-1.0], np.float32))[0] + 4 / 3 * alpha**3 / math.sqrt(math.pi)) \
        < 1e-6 * alpha**3
    assert abs(_horner(fit["energy"], np.array([-1.0], np.float32))[0] + 2 * alpha / math.sqrt(math.pi)) < 1e-6 * alpha


def test_tight_tolerance_falls_back():
    alpha = _alpha(1e-6, 1.0)
    fit = native.probe_ewald_fit(alpha, 1.0)
    ef, _ = _errors(fit, alpha, 1.0)
    assert not fit["ok"]
    assert fit["force_err"] > 1e-6 and ef > 1e-6


def test_bad_arguments_are_refused():
    for alpha, cutoff in ((0.0, 1.0), (3.0, 0.0), (-1.0, 1.0), (float("nan"), 1.0), (3.0, float("inf"))):
        with pytest.raises(native.MdError):
            native.probe_ewald_fit(alpha, cutoff)
