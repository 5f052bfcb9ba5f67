"""The benchmark workloads, built with the host API.

  methane_wall    methane at 150 K, 1e7 Pa next to a graphite Steele wall pair 5 sigma
                  apart; 201 points over [0.5 sigma, 4.5 sigma] (README example,
                  examples/docs_figures/01_getting_started.jl)
  c1c10_interface methane / n-decane planar vapour-liquid interface at 300 K with liquid
                  composition x = [0.5, 0.5] at its bubble point (surface_tension(model, T, x)
                  semantics, src/methods/surface_tension.jl:13-29); 2048 points over
                  [-20 L, 20 L]; reference cos_prof seed (two_phase.jl:8-35)
  slit_pores      512 methane slit pores between two graphite Steele walls at 150 K,
                  widths W ~ U[2 sigma, 12 sigma], p ~ logU[1e3, 1e7] Pa from seed 0,
                  stable bulk root, N = round(40 W / sigma) points over [0.5 L, W - 0.5 L]
  matrix_3d       methane / ethane / propane (x = 0.6, 0.3, 0.1) at 200 K and 1e7 Pa (a
                  stable liquid: the bubble pressure at 200 K is 3.2e6 Pa, below which the
                  homogeneous liquid is metastable and dries next to the spheres) in a
                  periodic cube of 30 sigma on a 256^3 grid, with a matrix of 64
                  non-overlapping hard spheres, radii ~ U[sigma, 2.5 sigma], from seed 0.
                  Solved with anderson_start=0.5 (MATRIX_3D_SOLVER): the damped Picard
                  warm-up makes slow progress on this system, and Anderson takes over earlier.
  matrix_2d       the same mixture in a periodic square with a matrix of hard disks
                  (infinite cylinders), on a 2D grid
  c1c10_lamellar  the methane / n-decane interface as a TwoPhase2DLamCart: the 1D grid
                  along x, uniform along y
  methane_radial  methane at 150 K, 1e7 Pa inside a spherical (Uniform1DSphr) or
                  cylindrical (Uniform1DCyl) graphite Steele pore of radius 5 sigma
"""
import numpy as np

from .bulk import PCSAFT
from .structure import (HardSphereMatrix, Steele, TwoPhase1DCart, TwoPhase2DLamCart, Uniform1DCart, Uniform1DCyl,
                        Uniform1DSphr, Uniform2DCart, Uniform3DCart)
from .system import DFTOptions, DFTSystem, initialize_profiles


def c1c10_lamellar(options=None, ngrid=(2048, 16), box=20.0, ywidth=2.0, x=(0.5, 0.5)):
    model = PCSAFT(["methane", "n-decane"])
    T = 300.0
    x = np.asarray(x, dtype=np.float64)
    p, vl, vv, y = model.bubble_pressure(T, x)
    L = model.length_scale()
    structure = TwoPhase2DLamCart((p, T), x / vl, y / vv, [[-box * L, box * L], [0.0, ywidth * L]], ngrid)
    system = DFTSystem(model, structure, None, options)
    return system, initialize_profiles(system)


def matrix_2d(options=None, n=256, box=30.0, ndisks=16, seed=0, T=200.0, p=1e7, x=(0.6, 0.3, 0.1),
              components=("methane", "ethane", "propane"), rmin=1.0, rmax=2.5, combining="nocombining"):
    model = PCSAFT(list(components), combining=combining)
    x = np.asarray(x, dtype=np.float64)
    v = model.volume(p, T, x, phase="liquid")
    L = model.length_scale()
    side = box * L
    structure = Uniform2DCart((p, T), x / v, [[0.0, side]] * 2, n)
    field = HardSphereMatrix.random(ndisks, rmin * L, rmax * L, side, seed=seed, ndim=2) if ndisks else None
    system = DFTSystem(model, structure, field, options)
    return system, initialize_profiles(system)


def methane_radial(coord="spherical", options=None, ngrid=151, radius=5.0, wall=True):
    model = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = model.volume(p, T, [1.0], phase="liquid")
    L = model.length_scale()
    mk = Uniform1DSphr if coord == "spherical" else Uniform1DCyl
    structure = mk((p, T), [1 / v], [0.0, radius * L], ngrid)
    system = DFTSystem(model, structure, Steele(["graphite"]) if wall else None, options)
    return system, initialize_profiles(system)


def methane_wall(options=None, ngrid=201):
    model = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = model.volume(p, T, [1.0], phase="liquid")
    L = model.length_scale()
    width = 5 * L
    surface = Steele(["graphite"], width)
    structure = Uniform1DCart((p, T), [1 / v], [0.5 * L, width - 0.5 * L], ngrid)
    system = DFTSystem(model, structure, surface, options)
    return system, initialize_profiles(system)


def c1c10_interface(options=None, ngrid=2048, box=20.0, x=(0.5, 0.5)):
    model = PCSAFT(["methane", "n-decane"])
    T = 300.0
    x = np.asarray(x, dtype=np.float64)
    p, vl, vv, y = model.bubble_pressure(T, x)
    L = model.length_scale()
    structure = TwoPhase1DCart((p, T), x / vl, y / vv, [-box * L, box * L], ngrid)
    system = DFTSystem(model, structure, None, options)
    return system, initialize_profiles(system)


def slit_pore_params(n=512, seed=0):
    rng = np.random.default_rng(seed)
    W = rng.uniform(2.0, 12.0, size=n)            # in units of sigma
    p = 10 ** rng.uniform(3.0, 7.0, size=n)       # Pa
    return W, p


def _slit_pore(model, W, p, T, options):
    sigma = model.length_scale()
    width = W * sigma
    v = model.volume(p, T, [1.0])
    N = int(round(40 * W))
    structure = Uniform1DCart((p, T), [1 / v], [0.5 * sigma, width - 0.5 * sigma], N)
    system = DFTSystem(model, structure, Steele(["graphite"], width), options)
    return system, initialize_profiles(system)


def slit_pores(n=512, seed=0, options=None, T=150.0, indices=None):
    """The n-pore workload, or only the pores listed in indices (same parameters as
    in the full set)."""
    model = PCSAFT(["methane"])
    Ws, ps = slit_pore_params(n, seed)
    idx = range(n) if indices is None else indices
    return [_slit_pore(model, Ws[i], ps[i], T, options) for i in idx]


MATRIX_3D_SOLVER = dict(anderson_start=0.5)


def matrix_3d(options=None, n=256, box=30.0, nspheres=64, seed=0, T=200.0, p=1e7, x=(0.6, 0.3, 0.1),
              components=("methane", "ethane", "propane"), rmin=1.0, rmax=2.5, combining="nocombining"):
    model = PCSAFT(list(components), combining=combining)
    x = np.asarray(x, dtype=np.float64)
    v = model.volume(p, T, x, phase="liquid")
    L = model.length_scale()
    side = box * L
    structure = Uniform3DCart((p, T), x / v, [[0.0, side]] * 3, n)
    field = HardSphereMatrix.random(nspheres, rmin * L, rmax * L, side, seed=seed) if nspheres else None
    system = DFTSystem(model, structure, field, options)
    return system, initialize_profiles(system)
