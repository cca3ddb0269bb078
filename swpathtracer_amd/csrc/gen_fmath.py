"""Re-derives the constants of spt_fmath.h and prints them as C literals.

Polynomials: least-squares fits of each function's remainder on 6000
Chebyshev nodes, weighted to minimise relative error, rounded to float32.
Splits: pi/2 and ln 2 into parts whose products with small integers are
exact in float; atan(c) for the reduction points as float pairs.
Run: python gen_fmath.py
"""
import numpy as np


def fit(fn, deg, a, b, relw, n=6000):
    k = np.arange(n)
    x = 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * n))
    A = np.stack([x ** i for i in range(deg + 1)], 1) * relw(x)[:, None]
    c, *_ = np.linalg.lstsq(A, fn(x) * relw(x), rcond=None)
    return c.astype(np.float32)


def split(v, bits):
    m, e = np.frexp(v)
    return float(np.ldexp(np.trunc(m * 2 ** bits) / 2 ** bits, e))


def lit(v):
    return repr(float(np.float32(v))) + "f"


def main():
    zs = (np.pi / 4) ** 2
    S = fit(lambda z: (np.sin(np.sqrt(z)) - np.sqrt(z)) / np.sqrt(z) ** 3, 3, 1e-8, zs,
            lambda z: z ** 1.5 / np.sin(np.sqrt(z)))
    Cc = fit(lambda z: (np.cos(np.sqrt(z)) - 1 + z / 2) / z ** 2, 3, 1e-8, zs, lambda z: z ** 2 / np.cos(np.sqrt(z)))
    h = np.log(2.0) / 2
    E = fit(lambda r: (np.expm1(r) - r) / r ** 2, 5, -h, h, lambda r: r ** 2 / np.exp(r))
    sm = (np.sqrt(2) - 1) / (np.sqrt(2) + 1)

    def rfun(z):
        s = np.sqrt(z)
        f = 2 * s / (1 - s)
        hf = 0.5 * f * f
        return ((np.log1p(f) - f + hf) / s - hf) / z

    G = fit(rfun, 3, 1e-10, sm ** 2, lambda z: z)
    A = fit(lambda z: (np.arctan(np.sqrt(z)) - np.sqrt(z)) / np.sqrt(z) ** 3, 5, 1e-10, (7 / 16) ** 2,
            lambda z: z ** 1.5 / np.arctan(np.sqrt(z)))
    p1 = split(np.pi / 2, 16)
    p2 = split(np.pi / 2 - p1, 16)
    p3 = np.pi / 2 - p1 - p2
    l1 = split(np.log(2.0), 16)
    g1 = split(np.log(2.0), 12)
    print("2/pi", lit(2 / np.pi), "pi/2 parts", lit(p1), lit(p2), lit(p3))
    print("sin", [lit(c) for c in S])
    print("cos", [lit(c) for c in Cc])
    print("exp: 1/ln2", lit(1 / np.log(2.0)), "ln2 parts", lit(l1), lit(np.log(2.0) - l1))
    print("exp poly", [lit(c) for c in E])
    print("log: ln2 parts", lit(g1), lit(np.log(2.0) - g1), "sqrt2", lit(np.sqrt(2.0)))
    print("log poly", [lit(c) for c in G])
    for c in (0.5, 1.0, 1.5):
        v = np.arctan(c)
        print(f"atan({c})", lit(v), lit(v - float(np.float32(v))))
    print("pi/2", lit(np.pi / 2), lit(np.pi / 2 - float(np.float32(np.pi / 2))))
    print("atan poly", [lit(c) for c in A])


if __name__ == "__main__":
    main()
