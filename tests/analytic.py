"""The reference's analytical Poisson solutions A-E as numpy-vectorised
functions of (x, y): u, f = -lap u, du/dx, du/dy
(validation_tests/analytical_solutions.hpp:7-96, with its PI = 3.14159265359).

Test infrastructure: the functions a user of the library would pass to
Operator.rhs / Operator.err_norms.  test_oracle_solve.py pins each against
the oracle's orc_u_ex / orc_f / orc_dx_u_ex / orc_dy_u_ex.
"""
import numpy as np

PI = 3.14159265359   # analytical_solutions.hpp:11,34


def _A():
    e = lambda x, y: np.exp(x - y)
    return (lambda x, y: x * y * (y - 1) * (x - 1) * e(x, y),
            lambda x, y: -2 * x * (y - 1) * (y - 2 * x + x * y + 2) * e(x, y),
            lambda x, y: (x * x + x - 1) * (y - 1) * y * e(x, y),
            lambda x, y: -1 * (y * y - 3 * y + 1) * (x - 1) * x * e(x, y))


def _B():
    return (lambda x, y: x * (1 - x) * np.exp(3 * x) * np.sin(2 * PI * y),
            lambda x, y: -np.sin(2 * PI * y) * np.exp(3 * x) * ((4 - 3 * x - 9 * x * x) - 4 * PI * PI * (x - x * x)),
            lambda x, y: (3 * (1 - x) * x - x + (1 - x)) * np.exp(3 * x) * np.sin(2 * PI * y),
            lambda x, y: 2 * PI * (1 - x) * x * np.exp(3 * x) * np.cos(2 * PI * y))


def _C():
    return (lambda x, y: y * (1 - y) * x * x * x,
            lambda x, y: 2 * x * x * x - 6 * x * y * (1 - y),
            lambda x, y: y * (1 - y) * x * x * 3,
            lambda x, y: (1 - 2 * y) * x * x * x)


def _D():
    return (lambda x, y: x * x * y * y,
            lambda x, y: -2 * y * y - 2 * x * x,
            lambda x, y: 2 * x * y * y,
            lambda x, y: 2 * y * x * x)


def _E():
    return (lambda x, y: np.cos(8 * PI * x) + np.cos(8 * PI * y),
            lambda x, y: 64 * PI * PI * np.cos(8 * PI * x) + 64 * PI * PI * np.cos(8 * PI * y),
            lambda x, y: -8 * PI * np.sin(8 * PI * x),
            lambda x, y: -8 * PI * np.sin(8 * PI * y))


# problem kind (the oracle's numbering, 1..5 = A..E) -> (u, f, ux, uy)
SOLUTIONS = {1: _A(), 2: _B(), 3: _C(), 4: _D(), 5: _E()}
