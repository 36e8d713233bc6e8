"""Fits the polynomial coefficients of sincos_bounded.h.

sin(r) = r + r^3 (S1 + S2 r^2 + S3 r^4) and cos(r) = 1 + r^2 (C1 + C2 r^2 + C3 r^4 + C4 r^6)
on |r| <= pi/4 (+ a margin for the reduction's rounding), minimising the maximum relative error
by Lawson's iteratively reweighted least squares in double precision, then rounding to float.
Run: python fit_sincos.py  (prints the coefficients as C hex-float literals)
"""
import numpy as np

R = np.pi / 4 * (1 + 2 ** -10)


def lawson(basis, target, weight, iters=400):
    w = np.ones_like(target)
    for _ in range(iters):
        A = basis * np.sqrt(w)[:, None]
        c, *_ = np.linalg.lstsq(A, target * np.sqrt(w), rcond=None)
        err = np.abs((basis @ c - target) * weight)
        w = w * err
        w /= w.sum()
    return c, float(np.max(np.abs((basis @ c - target) * weight)))


def main():
    r = np.cos(np.linspace(0, np.pi / 2, 20001)) * R  # Chebyshev-like nodes on (0, R]
    r = r[r > 1e-6]
    t = r * r
    # sin: (sin r - r) / r^3 = S1 + S2 t + S3 t^2; relative error of sin = r^3 * err / sin r
    g = (np.sin(r) - r) / r ** 3
    cs, es = lawson(np.stack([t ** 0, t, t ** 2], 1), g, r ** 3 / np.sin(r))
    # cos: (cos r - 1) / r^2 = C1 + C2 t + C3 t^2 + C4 t^3; relative error of cos = r^2 * err / cos r
    h = (np.cos(r) - 1) / t
    cc, ec = lawson(np.stack([t ** 0, t, t ** 2, t ** 3], 1), h, t / np.cos(r))
    for name, v in zip(["S1", "S2", "S3"], cs):
        print(f"{name} = {float(np.float32(v)).hex()}f  // {np.float32(v)!r}")
    for name, v in zip(["C1", "C2", "C3", "C4"], cc):
        print(f"{name} = {float(np.float32(v)).hex()}f  // {np.float32(v)!r}")
    print(f"max relative error before float rounding: sin {es:.3e}, cos {ec:.3e}")


if __name__ == "__main__":
    main()
