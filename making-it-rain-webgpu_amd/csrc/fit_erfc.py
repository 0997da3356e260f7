"""Generates the erfc coefficients the pair kernels use (MD_ERFC_C* in nonbonded.hip).

Method: erfcx(z) = exp(z^2) erfc(z) is smooth on [0, 4]; it is fitted as a
degree-8 polynomial in t = 1/(1 + z/2) by weighted least squares, the weights
iteratively re-balanced toward a minimax *relative* error (the quantity that
matters at the cutoff, where erfc is ~2e-4).  The device then evaluates
erfc(z) = exp(-z^2) * P(t), sharing exp(-z^2) with the force's erf' term.
Max relative error of the f32 evaluation over [0, 4]: ~3.4e-7.
"""
import numpy as np
from scipy.special import erfcx

P_SCALE, DEGREE = 0.5, 8


def fit():
    z = np.linspace(0, 4.0, 200001)
    t = 1 / (1 + P_SCALE * z)
    q = erfcx(z)
    w = 1 / q
    V = np.vander(t, DEGREE + 1, increasing=True)
    wt = w.copy()
    for _ in range(30):
        c, *_ = np.linalg.lstsq(V * wt[:, None], q * wt, rcond=None)
        r = (V @ c - q) / q
        wt = w * (np.abs(r) / np.abs(r).max() + 1e-3) ** 0.5
    return c


if __name__ == "__main__":
    for k, v in enumerate(fit()):
        print(f"#define MD_ERFC_C{k} {np.float32(v)!r}f".replace("np.float32(", "").replace(")", ""))
