"""Spatial closure weights of the cells: scheme_weights.jl:25-140 (no slowing down), the part of
Radiant's set-up that produces ω for ``flux_1D_BTE`` and ``flux_2D_BTE`` (sn_1D_bte.jl:27-64,
sn_2D_bte.jl:27-116; in 2D the weights of each axis fill the diagonal of ω and DD / DG apply).

For an order K (spatial Legendre moments per cell) the closure relates the outgoing face flux to
the incoming one and the cell moments: f_out = ω[0] f_in + sum_j C_j sx^(j-1) ω[j] psi_j with
C_j = sqrt(2j - 1).  ``DD`` is the high-order diamond difference (exact on polynomials of degree
K), ``DG`` the upwind discontinuous Galerkin closure (exact to degree K - 1), ``DG-`` / ``DG+``
change its last weight.  The adaptive ``AWD`` closure makes the sweep non-linear and is not
provided.
"""
import numpy as np

SCHEMES = ("DD", "DG", "DG-", "DG+")


def weights_1d(scheme, K):
    """ω [K + 1] of the x axis."""
    scheme = scheme.upper()
    if scheme == "AWD":  # the adaptive closure's base weights (scheme_weights.jl:47,52)
        if K > 2:
            raise ValueError("Closure relation not implemented yet.")
        scheme = "DD" if K == 1 else "DG"
    if scheme not in SCHEMES:
        raise ValueError("Closure relation not implemented yet.")
    Mx = K - 1
    w = np.zeros(K + 1)
    n = np.arange(K)
    if scheme == "DD":
        w[0] = (-1.0) ** (Mx + 1)
        w[1:] = 1.0 - (-1.0) ** (Mx + 1 - n)
    else:
        if scheme != "DG" and Mx < 1:
            raise ValueError(f"{scheme} scheme require at least a 2nd-order p expansion.")
        w[1:] = 1.0
        if scheme == "DG-":
            w[K] = Mx / (2.0 * Mx + 1.0)
        elif scheme == "DG+":
            w[K] = 1001.0
    return w


def spatial_moments(prob):
    """Legendre moments per cell of a problem (SN.get_schemes: Nm[5]) over its spatial axes and,
    for the slowing-down types, energy: the product of the orders with full coupling, 1 + the sum
    of (order - 1) without (1D: the x order; 2D Ox*Oy or Ox + Oy - 1; ...)."""
    orders = [max(1, int(getattr(prob, "scheme_order", 1) or 1))]
    if prob.ndims >= 2:
        orders.append(max(1, int(getattr(prob, "scheme_order_y", 1) or 1)))
    if prob.ndims >= 3:
        orders.append(max(1, int(getattr(prob, "scheme_order_z", 1) or 1)))
    if getattr(prob, "csd", False):
        orders.append(max(1, int(getattr(prob, "scheme_order_e", 1) or 1)))
    if getattr(prob, "full_coupling", True):
        return int(np.prod(orders))
    return 1 + sum(o - 1 for o in orders)
