"""Radial grids of the spherical and cylindrical structures, host side.

The reference builds them with Hankel.jl's quasi-discrete Hankel transform
(``Hankel.QDHT(0, 2, ub, N)`` for Uniform1DSphr, ``Hankel.QDHT(0, 1, ub, N)`` for
Uniform1DCyl, base/structure.jl:363-375; Hankel.jl 0.5.10 or later, Project.toml:47,
is not available here).  This module restates the published transform
(Guizar-Sicairos and Gutierrez-Vega, JOSA A 21, 53 (2004)) for what the host needs:
the sample radii r_n = j_n R / j_(N+1) and the quadrature weights of int f dV
(integrals.jl:20-22: 4 pi and 2 pi times integrateR).  j_n are the zeros of J_0
(cylindrical: the 2D radial Fourier transform) or of J_(1/2), n pi (spherical: the
3D one as the order-1/2 transform of sqrt(r) f).  The convolution operators are
built by the library (cdft_system_create).
"""
import numpy as np
from scipy import special

SPHERICAL, CYLINDRICAL = 1, 2


def _zeros(coord, count):
    if coord == SPHERICAL:
        return np.pi * np.arange(1, count + 1, dtype=np.float64)
    return special.jn_zeros(0, count)


def _jp1(coord, x):
    if coord == SPHERICAL:
        return np.sqrt(2.0 / (np.pi * x)) * (np.sin(x) / x - np.cos(x))      # J_(3/2)
    return special.j1(x)


def grid(coord, R, N):
    """Sample radii r_n and angular wavenumbers k_n of the transform over [0, R]."""
    z = _zeros(coord, N + 1)
    S = z[-1]
    return z[:N] * R / S, z[:N] / R


def weights(coord, R, N):
    """w with int f dV = w . f: 2 / (K^2 J_(p+1)(j_n)^2) (the transform's quadrature of
    int_0^R g r dr, g = f r for the sphere, K = j_(N+1) / R) times 4 pi or 2 pi."""
    z = _zeros(coord, N + 1)
    S = z[-1]
    K = S / R
    r = z[:N] * R / S
    wq = 2.0 / (K * K * _jp1(coord, z[:N]) ** 2)
    return 4.0 * np.pi * wq * r if coord == SPHERICAL else 2.0 * np.pi * wq


def domain_volume(coord, lb, ub):
    """_domain_volume (adsorption.jl:23-25): the sphere's or the disk's measure."""
    if coord == SPHERICAL:
        return 4.0 * np.pi / 3.0 * (ub ** 3 - lb ** 3)
    return np.pi * (ub ** 2 - lb ** 2)
