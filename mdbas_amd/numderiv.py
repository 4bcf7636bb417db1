"""Central finite-difference stencils of the reference's numerical forces
(numderiv.c), applied to any scalar function of one parameter.

The reference differentiates the potential with respect to one coordinate
at a time and returns the force, -dE/dx.  The same three stencils give
dE/dtheta for a parameter theta (a charge, an LJ eps or Rmin/2), which is
how the adjoint is checked:

  order 2  (numderiv.c:39)   (E(+h) - E(-h)) / 2h
  order 4  (numderiv.c:102)  (8 (E(+h) - E(-h)) - (E(+2h) - E(-2h))) / 12h
  order 6  (numderiv.c:190)  (45 (E(+h) - E(-h)) - 9 (E(+2h) - E(-2h))
                              + (E(+3h) - E(-3h))) / 60h
"""
from __future__ import annotations

from typing import Callable


def derivative(f: Callable[[float], float], h: float, order: int = 4) -> float:
    """dE/dtheta at theta = 0 of f(delta) = E(theta0 + delta)."""
    if order == 2:
        return (f(h) - f(-h)) / (2.0 * h)
    if order == 4:
        return (8.0 * (f(h) - f(-h)) - (f(2.0 * h) - f(-2.0 * h))) / (12.0 * h)
    if order == 6:
        return (45.0 * (f(h) - f(-h)) - 9.0 * (f(2.0 * h) - f(-2.0 * h))
                + (f(3.0 * h) - f(-3.0 * h))) / (60.0 * h)
    raise ValueError("order must be 2, 4 or 6")
