"""RODAS4 in numpy (TEST INFRASTRUCTURE ONLY): the checker for the
coefficient table and the stage sequence of the GPU reactor integrator
(astr_amd/csrc/chem.hip).  Nothing in the product imports this module.

RODAS4 is the stiffly accurate, L-stable Rosenbrock method of order 4 with an
embedded order-3 solution of Hairer & Wanner, "Solving Ordinary Differential
Equations II", 2nd ed. (1996), Sect. VI.4, in the transformed form of their
RODAS code: with E = I/(h gamma) - J and an autonomous f,

    E U_i = f(y + sum_{j<i} a_ij U_j) + sum_{j<i} (c_ij / h) U_j,  i = 1..6,

where rows 5 and 6 of a are those of the stage values Y_5 = y + sum a_5j U_j
and Y_6 = Y_5 + U_5; the step ends at y_new = Y_6 + U_6, and U_6 is the
error estimate (y_new minus the embedded order-3 solution Y_6).

The reactor has no reference counterpart (the reference's 0dreactor is
constant-volume and integrated otherwise), so this table is pinned by what
the mathematics fixes (tests/test_chem.py): the observed order of y_new (4)
and of the embedded solution (3) on smooth nonlinear problems with fixed
steps, and L-stability (|R(z)| -> 0 as z -> -infinity).
"""
from __future__ import annotations

import numpy as np

GAMMA = 0.25
# a_ij: stage arguments; c_ij: the stage coupling terms (both strictly lower)
A = np.zeros((6, 6))
C = np.zeros((6, 6))
A[1, 0] = 1.544
A[2, 0], A[2, 1] = 0.9466785280815826, 0.2557011698983284
A[3, 0], A[3, 1], A[3, 2] = 3.314825187068521, 2.896124015972201, 0.9986419139977817
A[4, 0], A[4, 1], A[4, 2], A[4, 3] = 1.221224509226641, 6.019134481288629, 12.53708332932087, -0.6878860361058950
C[1, 0] = -5.6688
C[2, 0], C[2, 1] = -2.430093356833875, -0.2063599157091915
C[3, 0], C[3, 1], C[3, 2] = -0.1073529058151375, -9.594562251023355, -20.47028614809616
C[4, 0], C[4, 1], C[4, 2], C[4, 3] = 7.496443313967647, -10.24680431464352, -33.99990352819905, 11.70890893206160
C[5, 0], C[5, 1], C[5, 2], C[5, 3], C[5, 4] = (8.083246795921522, -7.981132988064893, -31.52159432874371,
                                              16.31930543123136, -6.058818238834054)


def step(f, jac, y, h):
    """One RODAS4 step: (y_new, y_embedded)."""
    n = y.size
    E = np.eye(n) / (h * GAMMA) - jac(y)
    U = np.zeros((6, n))
    for i in range(6):
        if i < 5:
            yi = y + A[i, :i] @ U[:i]
        else:
            yi = y + A[4, :4] @ U[:4] + U[4]   # Y_6 = Y_5 + U_5
        rhs = f(yi) + (C[i, :i] @ U[:i]) / h
        U[i] = np.linalg.solve(E, rhs)
    y6 = y + A[4, :4] @ U[:4] + U[4]
    return y6 + U[5], y6


def integrate_fixed(f, jac, y0, t_end, nsteps):
    """nsteps equal steps; (y_new, y_embedded) of the last step and y_new at t_end."""
    y = np.array(y0, dtype=float)
    h = t_end / nsteps
    emb = y
    for _ in range(nsteps):
        y_prev = y
        y, emb = step(f, jac, y, h)
        del y_prev
    return y, emb


def stability_function(z):
    """R(z) of the method for y' = lambda y, z = h lambda (scalar, complex allowed)."""
    lam = z  # h = 1
    E = 1.0 / GAMMA - lam
    y = 1.0 + 0j
    U = []
    for i in range(6):
        if i < 5:
            yi = y + sum(A[i, j] * U[j] for j in range(i))
        else:
            yi = y + sum(A[4, j] * U[j] for j in range(4)) + U[4]
        rhs = lam * yi + sum(C[i, j] * U[j] for j in range(i))
        U.append(rhs / E)
    return y + sum(A[4, j] * U[j] for j in range(4)) + U[4] + U[5]
