"""Integrator selectors with the reference's names (integrator/__init__.py, _etdrk.py:85-93,
_stable_etdrk/__init__.py:96-105, _rk.py:257-283).

In the reference each member is a factory of an integrator object; here an ETD member selects
the engine scheme whose coefficients are built on the device (tfsm_plan_set_scheme), and a
Runge-Kutta member selects a Butcher tableau that the engine steps with tfsm_apply (L u + N(u))
and tfsm_rk_combine (the stage sums).
"""

from enum import Enum


class ETDRKIntegrator(Enum):
    """Cox-Matthews ETD schemes; coefficients from the reference's complex64 closed forms."""

    ETDRK0 = "ETDRK0"
    ETDRK1 = "ETDRK1"
    ETDRK2 = "ETDRK2"


class SETDRKIntegrator(Enum):
    """Kassam-Trefethen contour-integral ETD schemes (coefficients: float64 contour means on GPU).
    Options accepted by set_integrator: n_integration_points (16), integration_radius (1.0),
    cpu_cached (accepted for compatibility; the precompute runs on the GPU in both cases)."""

    SETDRK1 = "SETDRK1"
    SETDRK2 = "SETDRK2"
    SETDRK3 = "SETDRK3"
    SETDRK4 = "SETDRK4"


class RKIntegrator(Enum):
    """Explicit Runge-Kutta tableaux of the reference (_rk.py:82-255).  Options accepted by
    set_integrator: adaptive (False; adaptive steps are not supported), atol, rtol."""

    Euler = "Euler"
    Midpoint = "Midpoint"
    Heun12 = "Heun12"
    Ralston12 = "Ralston12"
    BogackiShampine23 = "BogackiShampine23"
    RK4 = "RK4"
    RK4_38Rule = "RK4_38Rule"
    Dorpi45 = "Dorpi45"
    Fehlberg45 = "Fehlberg45"
    CashKarp45 = "CashKarp45"


# Butcher tableaux of _rk.py:82-255, kept identical because the engine must reproduce the
# reference's float32 stage sums: each row of CA is [c_i, a_i1, a_i2, ...]; B the update weights.
_TABLEAUX = {
    "Euler": ([[1.0]], [1.0]),
    "Midpoint": ([[1 / 2, 1 / 2]], [0, 1]),
    "Heun12": ([[1, 1]], [1 / 2, 1 / 2]),
    "Ralston12": ([[2 / 3, 2 / 3]], [1 / 4, 3 / 4]),
    "BogackiShampine23": ([[1 / 2, 1 / 2], [3 / 4, 0, 3 / 4], [1, 2 / 9, 1 / 3, 4 / 9]], [2 / 9, 1 / 3, 4 / 9, 0]),
    "RK4": ([[1 / 2, 1 / 2], [1 / 2, 0, 1 / 2], [1, 0, 0, 1]], [1 / 6, 1 / 3, 1 / 3, 1 / 6]),
    "RK4_38Rule": ([[1 / 3, 1 / 3], [2 / 3, -1 / 3, 1], [1, -1, 1, 1]], [1 / 8, 3 / 8, 3 / 8, 1 / 8]),
    "Dorpi45": ([[1 / 5, 1 / 5], [3 / 10, 3 / 40, 9 / 40], [4 / 5, 44 / 45, -56 / 15, 32 / 9],
                 [8 / 9, 19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
                 [1, 9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
                 [1, 35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]],
                [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]),
    "Fehlberg45": ([[1 / 4, 1 / 4], [3 / 8, 3 / 32, 9 / 32], [12 / 13, 1932 / 2197, -7200 / 2197, 7296 / 2197],
                    [1, 439 / 216, -8, 3680 / 513, -845 / 4104], [1 / 2, -8 / 27, 2, -3544 / 2565, 1859 / 4104, -11 / 40]],
                   [16 / 135, 0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55]),
    "CashKarp45": ([[1 / 5, 1 / 5], [3 / 10, 3 / 40, 9 / 40], [3 / 5, 3 / 10, -9 / 10, 6 / 5],
                    [1, -11 / 54, 5 / 2, -70 / 27, 35 / 27],
                    [7 / 8, 1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]],
                   [37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771]),
}


def rk_tableau(integrator: RKIntegrator):
    """(CA, B) of a Runge-Kutta member."""
    return _TABLEAUX[integrator.value]
