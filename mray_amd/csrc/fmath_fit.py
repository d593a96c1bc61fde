"""Fits the polynomial coefficients of fmath.h (run: python fmath_fit.py).

Each core polynomial is fitted by weighted least squares on Chebyshev nodes
of its reduced interval, in double, against numpy's double-precision
function; the weights make the fit relative.  The printed coefficients are
rounded to float and pasted into fmath.h.  No other library's constants are
used; the range reductions are the textbook ones (Cody and Waite, "Software
Manual for the Elementary Functions", 1980).
"""
import numpy as np


def cheb_nodes(a, b, n):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * n))


def fit(xs, ys, basis, w):
    A = np.stack([f(xs) for f in basis], 1) * w[:, None]
    c, *_ = np.linalg.lstsq(A, ys * w, rcond=None)
    return c


def main():
    np.set_printoptions(precision=17)
    # exp(r) = 1 + r + r^2 * P(r), |r| <= ln2/2
    r = cheb_nodes(-np.log(2) / 2, np.log(2) / 2, 4000)
    y = (np.expm1(r) - r) / (r * r)
    c = fit(r, y, [lambda x, k=k: x ** k for k in range(6)], np.ones_like(r))
    print("exp P:", [float(np.float32(v)) for v in c])
    # atan(u) = u + u^3 * P(u^2), |u| <= tan(pi/8)
    t8 = np.tan(np.pi / 8)
    u = cheb_nodes(1e-4, t8, 4000)
    y = (np.arctan(u) - u) / u ** 3
    c = fit(u, y, [lambda x, k=k: (x * x) ** k for k in range(8)], np.ones_like(u))
    print("atan P:", [float(np.float32(v)) for v in c])


if __name__ == "__main__":
    main()
