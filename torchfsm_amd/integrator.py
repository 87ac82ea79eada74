"""Integrator selectors with the reference's names (integrator/__init__.py, _etdrk.py:85-93,
_stable_etdrk/__init__.py:96-105, _rk.py:257-283).

In the reference each member is a factory of an integrator object; here a member selects the
engine scheme whose coefficients are built on the device (tfsm_plan_set_scheme).  The explicit
Runge-Kutta family is outside the engine's path: selecting it raises ValueError at integration.
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
    """Explicit Runge-Kutta tableaux of the reference; not stepped by the ETD engine."""

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
