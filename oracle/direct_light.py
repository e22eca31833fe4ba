"""Float64 reference of the surrogate's direct-light term (test infrastructure only).

The light of a point source that reaches the readout plane z = Lz straight, crosses the
crystal -> coupling interface (Fresnel transmission, src/optics.jl:10-19, zero beyond total
internal reflection), survives the bulk over its path (absorption and Rayleigh lengths,
src/transport.jl:127-145) and is detected with the PDE in the SiPM cell under its landing
point (cell rule of src/transport.jl:96-108: i = min(nx, 1 + trunc(x / pitch))).

Method: a brute-force quadrature over the emission directions, independent of the device's
piecewise azimuth integral: midpoints in s (cos theta = uc + (1 - uc) s^2) times midpoints
in phi, each direction's landing point assigned to its cell.  Pinned in the tests by the
closed-form solid angle of a rectangle (T = 1, no attenuation) and by the one-dimensional
integral of the whole cone.
"""
from __future__ import annotations

import math

import numpy as np


def fresnel_T(n1: float, n2: float, u):
    """Unpolarized Fresnel transmission at incidence cosine u (binary64)."""
    u = np.asarray(u, np.float64)
    q = n1 / n2
    s2 = q * q * (1.0 - u * u)
    ok = s2 < 1.0
    ct = np.sqrt(np.where(ok, 1.0 - s2, 1.0))
    rs = (n1 * u - n2 * ct) / (n1 * u + n2 * ct)
    rp = (n1 * ct - n2 * u) / (n1 * ct + n2 * u)
    return np.where(ok, 1.0 - 0.5 * (rs * rs + rp * rp), 0.0)


def critical_cos(n1: float, n2: float) -> float:
    return math.sqrt(1.0 - (n2 / n1) ** 2) if n1 > n2 else 0.0


def direct_light(x0, y0, z, *, L, n1, n2, pde, pitch, nx, ny, inv_lambda=0.0,
                 ns=1500, nphi=3000, transmission=None):
    """D [nx*ny] (cell i + nx j) for a source at (x0, y0, z) in the box L = (Lx, Ly, Lz).
    transmission: optional T(u) replacing the Fresnel transmission (for the pins)."""
    Lx, Ly, Lz = L
    d = max(Lz - z, 1e-9)
    uc = critical_cos(n1, n2) if transmission is None else 0.0
    s = (np.arange(ns) + 0.5) / ns
    u = uc + (1.0 - uc) * s * s
    du = 2.0 * (1.0 - uc) * s / ns                       # du per s-cell
    T = fresnel_T(n1, n2, u) if transmission is None else transmission(u)
    w_u = T * np.exp(-d * inv_lambda / u) * du
    rho = d * np.sqrt(np.maximum(1.0 - u * u, 0.0)) / u
    phi = (np.arange(nphi) + 0.5) * (2.0 * math.pi / nphi)
    x = x0 + np.outer(rho, np.cos(phi))
    y = y0 + np.outer(rho, np.sin(phi))
    inside = (x >= 0) & (x <= Lx) & (y >= 0) & (y <= Ly)
    i = np.minimum(nx - 1, np.floor(x / pitch)).astype(np.int64)
    j = np.minimum(ny - 1, np.floor(y / pitch)).astype(np.int64)
    cell = np.where(inside, i + nx * j, nx * ny)
    w = np.broadcast_to(w_u[:, None], cell.shape) * (2.0 * math.pi / nphi)
    D = np.bincount(cell.ravel(), weights=w.ravel(), minlength=nx * ny + 1)[:nx * ny]
    return pde / (4.0 * math.pi) * D


def rectangle_solid_angle(x1, x2, y1, y2, d):
    """Solid angle of the rectangle [x1, x2] x [y1, y2] (relative to the foot point) seen
    from height d: F(x2, y2) - F(x1, y2) - F(x2, y1) + F(x1, y1),
    F(a, b) = atan(a b / (d sqrt(a^2 + b^2 + d^2)))."""
    def F(a, b):
        return math.atan(a * b / (d * math.sqrt(a * a + b * b + d * d)))
    return F(x2, y2) - F(x1, y2) - F(x2, y1) + F(x1, y1)
