"""Generates kernel_derivs.json: closed-form derivatives G^(k,l) of the softened
kernel G = -(x^2 + y^2 + eps^2)^(-1/2), for k + l <= 2p, by symbolic
differentiation (sympy) evaluated at 40 digits (mpmath). These pin the
oracle's compute_kernel_derivs (galax-core/src/lib.rs:1040-1151) and the HIP
derivative kernels independently of any recurrence.

Run: python tests/golden/gen_kernel_derivs.py
"""
import json
import os

import sympy as sp

POINTS = [  # (dx, dy, eps, p)
    (3.0, 4.0, 0.1, 2),        # the reference's own KAT point (lib.rs:2564)
    (0.3, -0.7, 0.05, 8),
    (-12.5, 3.25, 0.5, 6),
    (0.01, 0.02, 0.1, 8),      # softening-dominated
    (40.0, -25.0, 1e-3, 4),
    (-0.9, -0.4, 0.0, 6),      # unsoftened
]


def main():
    x, y, e = sp.symbols("x y e", real=True)
    order = max(2 * p for *_, p in POINTS)
    d = {(0, 0): -1 / sp.sqrt(x ** 2 + y ** 2 + e ** 2)}
    for n in range(1, order + 1):
        for k in range(n + 1):
            l = n - k
            prev = d[(k - 1, l)] if k > 0 else d[(k, l - 1)]
            d[(k, l)] = sp.factor(sp.together(sp.diff(prev, x if k > 0 else y)))
    out = []
    for dx, dy, eps, p in POINTS:
        vals = {}
        for (k, l), expr in d.items():
            if k + l <= 2 * p:
                v = expr.subs({x: sp.Rational(repr(dx)), y: sp.Rational(repr(dy)), e: sp.Rational(repr(eps))})
                vals[f"{k},{l}"] = float(sp.N(v, 40))
        out.append({"dx": dx, "dy": dy, "eps": eps, "p": p, "G": vals})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_derivs.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
