"""Structures (grids + state conditions) and external fields, host side.

Mirrors the reference's constructors:
  Uniform1DCart   src/base/structure.jl:211-213
  TwoPhase1DCart  src/base/structure.jl (TwoPhaseSystem, ``TwoPhase1DCart``)
  Steele          src/structure/external_fields/steele.jl:31-46 (database/Steele/steele_like.csv)
"""
import numpy as np


class Structure:
    """A 1D Cartesian grid with (p, T) conditions and bulk densities."""

    def __init__(self, conditions, rho_bulk, bounds, ngrid, topology="uniform", rho_bulk2=None):
        p, T = conditions
        self.conditions = (float(p), float(T))
        self.rho_bulk = np.asarray(rho_bulk, dtype=np.float64).reshape(-1)
        b = np.asarray(bounds, dtype=np.float64).reshape(-1)
        if b.size != 2:
            raise ValueError("1D bounds must have exactly 2 numbers")   # DimensionMismatch in the reference
        if not b[1] > b[0]:
            raise ValueError("bounds must satisfy ub > lb")
        self.bounds = ((float(b[0]), float(b[1])),)
        n = ngrid[0] if isinstance(ngrid, (tuple, list)) else ngrid
        self.ngrid = (int(n),)
        self.topology = topology
        self.rho_bulk2 = None if rho_bulk2 is None else np.asarray(rho_bulk2, dtype=np.float64).reshape(-1)

    @property
    def ndim(self):
        return 1

    def coords(self):
        """LinRange(lb, ub, N) (get_coords, src/structure/structure.jl:138-151)."""
        lb, ub = self.bounds[0]
        N = self.ngrid[0]
        t = np.arange(N, dtype=np.float64) / (N - 1)
        return (1.0 - t) * lb + t * ub

    def dz(self):
        """(ub - lb)/N (structure_dz, src/structure/structure.jl:187-195)."""
        lb, ub = self.bounds[0]
        return (ub - lb) / self.ngrid[0]


def Uniform1DCart(conditions, rho_bulk, bounds, ngrid):
    return Structure(conditions, rho_bulk, bounds, ngrid, "uniform")


def TwoPhase1DCart(conditions, rho_bulk, rho_bulk2, bounds, ngrid):
    return Structure(conditions, rho_bulk, bounds, ngrid, "two_phase", rho_bulk2)


# database/Steele/steele_like.csv:3 -- eps (K), sigma (A), rho (A^-3), delta (A)
STEELE_DB = {"graphite": (28.0, 3.4, 0.114, 3.35)}


class Steele:
    """Steele 10-4-3 wall(s); a second wall at ``width`` when width != 0."""

    def __init__(self, surface, width=0.0):
        surface = list(surface)
        self.surface = surface
        rows = [STEELE_DB[s] for s in surface]
        self.epsilon = np.array([r[0] for r in rows])
        self.sigma = np.array([r[1] for r in rows]) * 1e-10
        self.rho = np.array([r[2] for r in rows]) * 1e30
        self.delta = np.array([r[3] for r in rows]) * 1e-10
        self.width = float(width)

    def vext(self, structure, model):
        """beta V_ext(z) per component, (N, nc) column-major flattened (steele.jl:49-81)."""
        z = structure.coords()
        T = structure.conditions[1]
        eps_i = np.diag(model.epsilon)
        sig_i = np.diag(model.sigma)
        N, nc = z.size, len(eps_i)
        V = np.zeros((nc, N))
        H = self.width
        for s in range(len(self.surface)):
            esi = np.sqrt(self.epsilon[s] * eps_i)
            ssi = (self.sigma[s] + sig_i) / 2.0
            D = self.delta[s]
            for i in range(nc):
                pre = 2.0 * np.pi * self.rho[s] * esi[i] * D * ssi[i] ** 2
                V[i] += pre * (2.0 / 5.0 * (ssi[i] / z) ** 10 - (ssi[i] / z) ** 4
                               - ssi[i] ** 4 / (3.0 * D * (z + 0.61 * D) ** 3))
                if H != 0:
                    zz = H - z
                    V[i] += pre * (2.0 / 5.0 * (ssi[i] / zz) ** 10 - (ssi[i] / zz) ** 4
                                   - ssi[i] ** 4 / (3.0 * D * (zz + 0.61 * D) ** 3))
        return (V * (1.0 / T)).reshape(-1)
