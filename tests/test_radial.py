"""Radial (spherical / cylindrical) geometries on the host and in the oracle: the
quasi-discrete Hankel transform restated from its published algorithm (Hankel.jl,
a dependency absent here) pinned by closed-form transforms, its quadrature, the
reference's own radial bulk test (test/test_coordinate_systems.jl:60-80), the 3D
Cartesian path on the same spherically symmetric profile, and the radial Steele
wall.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle as po
from cdft_amd import DFTSystem, PCSAFT, Steele, Uniform3DCart
from cdft_amd import radial
from cdft_amd.structure import Uniform1DCyl, Uniform1DSphr
from cdft_amd.system import integrate

DP = C.POINTER(C.c_double)


class _Q(C.Structure):
    _fields_ = [("N", C.c_int), ("sph", C.c_int), ("R", C.c_double), ("S", C.c_double), ("K", C.c_double),
                ("r", DP), ("k", DP), ("jp1", DP), ("C", DP), ("w", DP)]


def _qdht(coord, R, N):
    L = po.lib()
    L.or_qdht_create.restype = C.POINTER(_Q)
    L.or_qdht_create.argtypes = [C.c_int, C.c_double, C.c_int]
    L.or_qdht_forward.argtypes = [C.POINTER(_Q), DP, DP]
    L.or_qdht_backward.argtypes = [C.POINTER(_Q), DP, DP]
    L.or_qdht_free.argtypes = [C.POINTER(_Q)]
    q = L.or_qdht_create(coord, R, N)
    qq = q.contents
    arr = lambda p: np.ctypeslib.as_array(p, (N,)).copy()
    return L, q, arr(qq.r), arr(qq.k), arr(qq.w)


@pytest.mark.parametrize("coord", [radial.SPHERICAL, radial.CYLINDRICAL])
def test_qdht_transforms_gaussians_exactly(coord):
    """F(k) of exp(-r^2/a^2): sqrt(pi)/4 a^3 exp(-k^2 a^2/4) (int f j0(kr) r^2 dr) and
    a^2/2 exp(-k^2 a^2/4) (int f J0(kr) r dr); the inverse restores f; the
    quadrature gives pi^(3/2) a^3 and pi a^2."""
    N, R, a = 200, 10.0, 1.3
    L, q, r, k, w = _qdht(coord, R, N)
    f = np.exp(-r ** 2 / a ** 2)
    F, g = np.zeros(N), np.zeros(N)
    L.or_qdht_forward(q, f.ctypes.data_as(DP), F.ctypes.data_as(DP))
    L.or_qdht_backward(q, F.ctypes.data_as(DP), g.ctypes.data_as(DP))
    if coord == radial.SPHERICAL:
        ref, vol = math.sqrt(math.pi) / 4 * a ** 3 * np.exp(-k ** 2 * a ** 2 / 4), math.pi ** 1.5 * a ** 3
    else:
        ref, vol = a ** 2 / 2 * np.exp(-k ** 2 * a ** 2 / 4), math.pi * a ** 2
    assert np.max(np.abs(F - ref)) < 1e-14 * np.max(ref)
    assert np.max(np.abs(g - f)) < 1e-13
    assert w @ f == pytest.approx(vol, rel=1e-14)
    # the host's grid and weights are the oracle's
    rh, kh = radial.grid(coord, R, N)
    assert np.allclose(rh, r, rtol=1e-14) and np.allclose(kh, k, rtol=1e-14)
    assert np.allclose(radial.weights(coord, R, N), w, rtol=1e-13)
    L.or_qdht_free(q)


def test_spherical_grid_is_the_sine_transform_grid():
    r, k = radial.grid(radial.SPHERICAL, 5.0, 99)
    assert np.allclose(r, 5.0 * np.arange(1, 100) / 100, rtol=1e-15)
    assert np.allclose(k, np.pi * np.arange(1, 100) / 5.0, rtol=1e-15)


def _methane():
    model = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = model.volume(p, T, [1.0], "liquid")
    return model, T, p, 1.0 / v, model.length_scale()


@pytest.mark.parametrize("coord,ub,n", [(radial.SPHERICAL, 10, 151), (radial.CYLINDRICAL, 20, 101)])
def test_bulk_mu_res_at_the_first_radius(coord, ub, n):
    """test_coordinate_systems.jl:60-80: delta F_res / delta rho at the first grid
    point of a uniform profile equals mu_res (the reference's tolerance 1e-5)."""
    model, T, p, rb, L = _methane()
    mk = Uniform1DSphr if coord == radial.SPHERICAL else Uniform1DCyl
    s = DFTSystem(model, mk((p, T), [rb], [0.0, ub * L], n))
    d, _ = po.Oracle(**s.descriptor()).dFdrho_res(np.full(s.N, rb))
    err = abs(d[0] - s.species.chempot_res[0])
    assert err < 1e-7, err


def test_spherical_weighted_densities_match_the_3d_cartesian_path():
    """A spherically symmetric profile: the radial scalar weighted densities equal the
    3D Cartesian ones (the oracle's literal FFT restatement, sampled on its own grid)
    to the radial grid's resolution.  The vector slot is the reference's radial
    construction (weighted_densities.jl:290-328: the kernel in unscaled units, then
    N_A L^3), which comes out 2 pi L^2 times the Cartesian vector weighted density."""
    model, T, p, rb, L = _methane()
    prof = lambda rr: rb * (1 - 0.5 * np.exp(-rr ** 2 / (2 * (1.2 * L) ** 2)))
    ss = DFTSystem(model, Uniform1DSphr((p, T), [rb], [0.0, 12 * L], 601))
    r = ss.structure.axis_coords(0)
    ns = po.Oracle(**ss.descriptor()).weighted_densities(prof(r)).reshape(7, -1)
    n = 32
    sc = DFTSystem(model, Uniform3DCart((p, T), [rb], [[-8 * L, 8 * L]] * 3, n))
    xs = -8 * L + np.arange(n) * 16 * L / n
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    R3 = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).transpose(2, 1, 0).reshape(-1)
    nc3 = po.Oracle(**sc.descriptor()).weighted_densities(prof(R3)).reshape(9, n, n, n)
    for k in (n // 2 + 2, n // 2 + 4, n // 2 + 6):
        rr = abs(xs[k])
        for fs, fc in ((1, 1), (2, 2), (4, 6), (5, 7), (6, 8)):
            a, b = nc3[fc, n // 2, n // 2, k], np.interp(rr, r, ns[fs])
            assert abs(a - b) < 2e-5 * abs(a), (fs, rr / L, a, b)
        ratio = np.interp(rr, r, ns[3]) / nc3[3, n // 2, n // 2, k]
        assert ratio == pytest.approx(2 * math.pi * L ** 2, rel=5e-3)


def test_radial_steele_wall_and_integrals():
    model, T, p, rb, L = _methane()
    pore = Uniform1DSphr((p, T), [rb], [0.0, 5 * L], 101)
    wall = Steele(["graphite"])
    v = wall.vext(pore, model)
    r = pore.axis_coords(0)
    eps_s, sig_s, rho_s, dl = 28.0, 3.4e-10, 0.114e30, 3.35e-10
    esi, ssi = math.sqrt(eps_s * model.epsilon[0, 0]), (sig_s + model.sigma[0, 0]) / 2
    z = np.maximum(5 * L - r, 0.5 * model.sigma[0, 0])
    ref = 2 * math.pi * rho_s * esi * dl * ssi ** 2 * (0.4 * (ssi / z) ** 10 - (ssi / z) ** 4
                                                       - ssi ** 4 / (3 * dl * (z + 0.61 * dl) ** 3)) / T
    assert np.allclose(v, ref, rtol=1e-13)
    outside = Uniform1DCyl((p, T), [rb], [2 * L, 8 * L], 101)
    vo = wall.vext(outside, model)
    ro = outside.axis_coords(0)
    assert vo[np.argmin(np.abs(ro - 2 * L))] > 100 and vo[-1] < 0       # repulsive at the core, attractive beyond
    assert integrate(pore, np.ones(pore.npoints)) == pytest.approx(4 / 3 * math.pi * (5 * L) ** 3, rel=2e-2)
    assert pore.volume() == pytest.approx(4 / 3 * math.pi * (5 * L) ** 3)
    assert outside.volume() == pytest.approx(math.pi * ((8 * L) ** 2 - (2 * L) ** 2))
