"""Structures (grids + state conditions) and external fields, host side.

Mirrors the reference's constructors:
  Uniform1DCart      src/base/structure.jl:310
  Uniform2DCart      src/base/structure.jl:315-332
  Uniform3DCart      src/base/structure.jl:335-352
  TwoPhase1DCart     src/base/structure.jl:491 (``TwoPhase1DCart``)
  TwoPhase2DLamCart  src/base/structure.jl:494-513 (lamellar interface along the first axis)
  Uniform1DSphr      src/base/structure.jl:388-410 (radial grid of a spherical transform)
  Uniform1DCyl       src/base/structure.jl:411-433 (radial grid of a cylindrical transform)
  Steele          src/structure/external_fields/steele.jl:31-46 (database/Steele/steele_like.csv)
and adds a field the reference does not have:
  HardSphereMatrix  a periodic matrix of impenetrable spheres on a 3D grid (disks,
                    i.e. infinite cylinders, on a 2D grid)
"""
import numpy as np


class Structure:
    """A 1D, 2D or 3D Cartesian grid with (p, T) conditions and bulk densities."""

    def __init__(self, conditions, rho_bulk, bounds, ngrid, topology="uniform", rho_bulk2=None, ndim=1, coord=0):
        p, T = conditions
        self.conditions = (float(p), float(T))
        self.rho_bulk = np.asarray(rho_bulk, dtype=np.float64).reshape(-1)
        b = np.asarray(bounds, dtype=np.float64).reshape(-1)
        if b.size != 2 * ndim:
            raise ValueError(f"{ndim}D bounds must have exactly {2 * ndim} numbers")   # DimensionMismatch
        b = b.reshape(ndim, 2)
        if not np.all(b[:, 1] > b[:, 0]):
            raise ValueError("bounds must satisfy ub > lb")
        self.bounds = tuple((float(x[0]), float(x[1])) for x in b)
        g = list(ngrid) if isinstance(ngrid, (tuple, list)) else [ngrid]
        if len(g) == 1 and ndim > 1:
            g = g * ndim                                             # one count for every axis
        if len(g) != ndim:
            raise ValueError(f"ngrid must have {ndim} entries")
        self.ngrid = tuple(int(n) for n in g)
        self.topology = topology
        self.rho_bulk2 = None if rho_bulk2 is None else np.asarray(rho_bulk2, dtype=np.float64).reshape(-1)
        self.coord = int(coord)              # 0 Cartesian, 1 spherical, 2 cylindrical
        if self.coord and (ndim != 1 or self.bounds[0][0] < 0):
            raise ValueError("radial structures are 1D with bounds 0 <= lb < ub")

    @property
    def ndim(self):
        return len(self.ngrid)

    @property
    def npoints(self):
        return int(np.prod(self.ngrid))

    def axis_coords(self, a):
        """LinRange(lb, ub, N) of axis a (get_coords, src/structure/structure.jl:138-151);
        the transform's sample radii on a radial grid (structure.jl:203-214)."""
        if self.coord:
            from .radial import grid
            return grid(self.coord, self.bounds[0][1], self.ngrid[0])[0]
        lb, ub = self.bounds[a]
        N = self.ngrid[a]
        t = np.arange(N, dtype=np.float64) / (N - 1)
        return (1.0 - t) * lb + t * ub

    def coords(self):
        """1D: the coordinate vector; 2D and 3D: the list of per-axis coordinate vectors."""
        if self.ndim == 1:
            return self.axis_coords(0)
        return [self.axis_coords(a) for a in range(self.ndim)]

    def dz(self):
        """(ub - lb)/N of the first axis (structure_dz, src/structure/structure.jl:187-195)."""
        return self.dzs()[0]

    def dzs(self):
        return tuple((ub - lb) / n for (lb, ub), n in zip(self.bounds, self.ngrid))

    def volume(self):
        """_domain_volume (src/methods/adsorption.jl:22-25): the product of the bound
        lengths, or the sphere's / disk's measure between lb and ub."""
        if self.coord:
            from .radial import domain_volume
            return float(domain_volume(self.coord, *self.bounds[0]))
        return float(np.prod([ub - lb for lb, ub in self.bounds]))


def Uniform1DCart(conditions, rho_bulk, bounds, ngrid):
    return Structure(conditions, rho_bulk, bounds, ngrid, "uniform")


def Uniform2DCart(conditions, rho_bulk, bounds, ngrid):
    """Uniform2DCart(conditions, rho_bulk, bounds (2x2), ngrid) (base/structure.jl:315-332)."""
    return Structure(conditions, rho_bulk, bounds, ngrid, "uniform", ndim=2)


def TwoPhase2DLamCart(conditions, rho_bulk, rho_bulk2, bounds, ngrid):
    """TwoPhase2DLamCart(conditions, rho_bulk, rho_bulk2, bounds (2x2), ngrid): a lamellar
    two-phase interface along the first axis, uniform along the second
    (base/structure.jl:494-513)."""
    return Structure(conditions, rho_bulk, bounds, ngrid, "two_phase", rho_bulk2, ndim=2)


def Uniform3DCart(conditions, rho_bulk, bounds, ngrid):
    """Uniform3DCart(conditions, rho_bulk, bounds (3x2), ngrid) (base/structure.jl:335-352)."""
    return Structure(conditions, rho_bulk, bounds, ngrid, "uniform", ndim=3)


def Uniform1DSphr(conditions, rho_bulk, bounds, ngrid):
    """Uniform1DSphr(conditions, rho_bulk, [lb, ub] or ub, ngrid): a radial grid over
    [0, ub] (the transform's aperture); lb places walls only (base/structure.jl:388-410)."""
    b = [0.0, float(bounds)] if np.isscalar(bounds) else bounds
    return Structure(conditions, rho_bulk, b, ngrid, "uniform", coord=1)


def Uniform1DCyl(conditions, rho_bulk, bounds, ngrid):
    """Uniform1DCyl(conditions, rho_bulk, [lb, ub] or ub, ngrid) (base/structure.jl:411-433)."""
    b = [0.0, float(bounds)] if np.isscalar(bounds) else bounds
    return Structure(conditions, rho_bulk, b, ngrid, "uniform", coord=2)


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
        """beta V_ext(z) per component, (N, nc) column-major flattened (steele.jl:49-81);
        on a radial grid the single-wall form at the wall distance r - lb (lb > 0: fluid
        outside a sphere or cylinder) or ub - r (a pore), clamped below at half the
        smallest sigma_i (steele.jl:111-142)."""
        if structure.coord:
            return self._vext_radial(structure, model)
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

    def _vext_radial(self, structure, model):
        r = structure.axis_coords(0)
        lb, ub = structure.bounds[0]
        T = structure.conditions[1]
        eps_i = np.diag(model.epsilon)
        sig_i = np.diag(model.sigma)
        zw = np.maximum(r - lb if lb > 0 else ub - r, 0.5 * np.min(sig_i))
        V = np.zeros((len(eps_i), r.size))
        for s in range(len(self.surface)):
            esi = np.sqrt(self.epsilon[s] * eps_i)
            ssi = (self.sigma[s] + sig_i) / 2.0
            D = self.delta[s]
            for i in range(len(eps_i)):
                V[i] += 2.0 * np.pi * self.rho[s] * esi[i] * D * ssi[i] ** 2 * (
                    2.0 / 5.0 * (ssi[i] / zw) ** 10 - (ssi[i] / zw) ** 4 - ssi[i] ** 4 / (3.0 * D * (zw + 0.61 * D) ** 3))
        return (V * (1.0 / T)).reshape(-1)


class HardSphereMatrix:
    """Impenetrable spheres of a porous matrix on a periodic 3D grid (no counterpart
    in the reference): beta V_ext = cap at grid points whose distance (minimum image)
    to a sphere centre is below its radius, 0 elsewhere; the same for every
    component.  Any cap above about 100 gives the same fixed point, because the
    exponent is clamped to [-100, 100] (converge.jl:150)."""

    def __init__(self, centres, radii, cap=500.0, ndim=3):
        self.ndim = int(ndim)
        self.centres = np.asarray(centres, dtype=np.float64).reshape(-1, self.ndim)
        self.radii = np.asarray(radii, dtype=np.float64).reshape(-1)
        self.cap = float(cap)

    @classmethod
    def random(cls, n, rmin, rmax, box, seed=0, cap=500.0, max_tries=200000, ndim=3):
        """n non-overlapping spheres (disks for ndim = 2) with radii ~ U[rmin, rmax] in the
        periodic cube [0, box)^ndim, by random sequential addition from `seed`."""
        rng = np.random.default_rng(seed)
        cs, rs = [], []
        tries = 0
        while len(cs) < n:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("could not place the spheres without overlap")
            r = rng.uniform(rmin, rmax)
            c = rng.uniform(0.0, box, size=ndim)
            ok = True
            for c2, r2 in zip(cs, rs):
                d = c - c2
                d -= box * np.round(d / box)
                if np.dot(d, d) < (r + r2) ** 2:
                    ok = False
                    break
            if ok:
                cs.append(c)
                rs.append(r)
        return cls(np.array(cs), np.array(rs), cap, ndim)

    def mask(self, structure):
        """Boolean (N,) mask of the points inside a sphere, column-major (x fastest)."""
        nd = structure.ndim
        if nd != self.ndim:
            raise ValueError(f"a {self.ndim}D matrix on a {nd}D grid")
        xs = [structure.axis_coords(a) for a in range(nd)]
        lens = [ub - lb for lb, ub in structure.bounds]
        inside = np.zeros(structure.ngrid[::-1], dtype=bool)             # [z,] y, x
        for c, r in zip(self.centres, self.radii):
            d2, idx = [], []
            for a in range(nd):                                          # per-axis window of the sphere
                d = xs[a] - c[a]
                d -= lens[a] * np.round(d / lens[a])
                sel = np.nonzero(d * d < r * r)[0]
                idx.append(sel)
                d2.append(d[sel] ** 2)
            if min(len(i) for i in idx) == 0:
                continue
            if nd == 3:
                blk = d2[2][:, None, None] + d2[1][None, :, None] + d2[0][None, None, :] < r * r
                sub = np.ix_(idx[2], idx[1], idx[0])
            else:
                blk = d2[1][:, None] + d2[0][None, :] < r * r
                sub = np.ix_(idx[1], idx[0])
            inside[sub] |= blk
        return inside.reshape(-1)

    def vext(self, structure, model):
        m = self.mask(structure)
        nc = len(model)
        return np.tile(np.where(m, self.cap, 0.0), nc)
