"""Fits the polynomial used by esth_atanf_unit() in esth_math.h.

g(s) = atan(sqrt(s)) / sqrt(s) on s in [0, 1] is interpolated at Chebyshev
nodes (degree DEG), converted to monomials in s and rounded to float32.
Running this prints the coefficient block pasted into esth_math.h together with
the maximum absolute error of t * g(t^2) against numpy's double atan on a dense
grid of t in [0, 1], evaluated the way the header evaluates it (float Horner
with fused multiply-adds).
"""
import numpy as np
from numpy.polynomial import chebyshev as C
from numpy.polynomial import polynomial as P

DEG = 9


def fit(deg=DEG):
    u = np.cos(np.pi * (np.arange(4000) + 0.5) / 4000)
    s = (u + 1) / 2
    g = np.arctan(np.sqrt(s)) / np.sqrt(s)
    cu = C.cheb2poly(C.chebfit(u, g, deg))
    ps = np.zeros(1)
    for k, a in enumerate(cu):
        ps = P.polyadd(ps, a * P.polypow([-1.0, 2.0], k))
    return ps.astype(np.float32)


def horner_f32(coef, s):
    acc = np.full_like(s, coef[-1], dtype=np.float32)
    for c in coef[-2::-1]:
        acc = (acc.astype(np.float64) * s + np.float64(c)).astype(np.float32)  # fma: one rounding
    return acc


if __name__ == "__main__":
    coef = fit()
    t = np.linspace(0, 1, 2_000_001, dtype=np.float32)
    s = (t * t).astype(np.float32)
    val = (horner_f32(coef, s) * t).astype(np.float32)
    err = np.max(np.abs(val.astype(np.float64) - np.arctan(t.astype(np.float64))))
    for i, c in enumerate(coef):
        print(f"#define ESTH_ATAN_C{i} {float(c)!r}f /* {np.float32(c).view(np.uint32):#010x} */")
    print(f"/* max |t*g(t^2) - atan(t)| on [0,1] = {err:.3e} */")
