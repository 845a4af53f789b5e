"""Writes weights.json: exact Lagrange and conservative-prolongation weights,
computed here with fractions.Fraction (the closed forms behind
src/operators.jl:627-682 and src/schedule.jl:1097-1120) and rounded once to
float (float(Fraction) is correctly rounded, as T(::Rational{BigInt}) is).

Run: python tests/golden/make_weights.py   (no reference code is imported)
"""
import json
import os
from fractions import Fraction as F


def lagrange(nodes, x):
    w = []
    for i, ni in enumerate(nodes):
        num, den = F(1), F(1)
        for j, nj in enumerate(nodes):
            if j != i:
                num *= x - nj
                den *= ni - nj
        w.append(num / den)
    return w


def conservative(p):
    r = (p - 1) // 2
    bounds = [F(-r) - F(1, 2) + i for i in range(p + 1)]
    L = lagrange(bounds, F(0))
    tail = [sum(L[t + 1:], F(0)) for t in range(p)]
    low = [2 * (tail[t] - (1 if t < r else 0)) for t in range(p)]
    high = [2 * ((1 if t < r + 1 else 0) - tail[t]) for t in range(p)]
    return low, high


def main():
    out = {"lagrange": [], "conservative": []}
    for p in (2, 3, 4, 5, 6, 8):
        for q in range(0, 4 * (p - 1) + 1):
            xi = F(q, 4)
            w = lagrange([F(j) for j in range(p)], xi)
            out["lagrange"].append({"p": p, "num": xi.numerator, "den": xi.denominator,
                                    "exact": [str(v) for v in w], "float": [float(v) for v in w]})
    for p in (1, 3, 5, 7):
        low, high = conservative(p)
        out["conservative"].append({"p": p, "low": [str(v) for v in low], "high": [str(v) for v in high],
                                    "low_float": [float(v) for v in low],
                                    "high_float": [float(v) for v in high]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
