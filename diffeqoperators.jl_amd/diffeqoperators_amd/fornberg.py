"""Finite-difference weights (Fornberg 1988), host side, FP64.

Follows src/derivative_operators/fornberg.jl:7-62 of DiffEqOperators.jl, including the
sum-to-zero correction of the middle weight (:57-60) that the reference applies to every
stencil of nonzero derivative order.
"""
from __future__ import annotations


def calculate_weights(order: int, x0: float, x) -> list:
    """Weights of the `order`-th derivative at x0 from the values at points x."""
    x = [float(v) for v in x]
    N = len(x)
    if not order < N:
        raise AssertionError("Not enough points for the requested order.")
    M = order
    x0 = float(x0)
    C = [[0.0] * (M + 1) for _ in range(N)]
    C[0][0] = 1.0
    c1 = 1.0
    c4 = x[0] - x0
    for i in range(1, N):
        mn = min(i, M)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - x0
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for s in range(mn, 0, -1):
                    C[i][s] = c1 * (s * C[i - 1][s - 1] - c5 * C[i - 1][s]) / c2
                C[i][0] = -c1 * c5 * C[i - 1][0] / c2
            for s in range(mn, 0, -1):
                C[j][s] = (c4 * C[j][s] - s * C[j][s - 1]) / c3
            C[j][0] = c4 * C[j][0] / c3
        c1 = c2
    w = [C[r][M] for r in range(N)]
    if order != 0:
        tot = 0.0
        for v in w:
            tot += v
        w[N // 2] -= tot
    return w
