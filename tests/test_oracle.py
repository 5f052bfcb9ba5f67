"""Pins the CPU oracle (oracle/oracle.c) against the reference's known answers,
closed forms and invariants.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import pyoracle as po
from cdft_amd.bulk import PCSAFT, RGAS

HERE = os.path.dirname(os.path.abspath(__file__))


def tothk(u):
    g0 = math.cos(0.5 * (u[0] + u[1]))
    return np.array([g0, g0 + 1e-8 * math.sin(u[0] * u[0])])


# --- Anderson: src/utils/anderson.jl:116-160 docstring histories (6 s.f.) -----------
def test_aasol_toth_kelley_beta1():
    a = po.aasol(tothk, np.ones(2), m=2, rtol=1e-10)
    ref = [6.50111e-01, 4.48661e-01, 2.61480e-02, 7.25389e-02, 1.53107e-04, 1.18513e-05, 1.82466e-08, 1.04725e-13]
    assert len(a["history"]) == len(ref)
    np.testing.assert_allclose(a["history"], ref, rtol=6e-6)
    cond = [1.00000e+00, 2.01556e+10, 1.37776e+09, 3.61348e+10, 2.54948e+11, 3.67694e+10]
    alph = [1.00000e+00, 4.61720e+00, 2.15749e+00, 1.18377e+00, 1.00000e+00, 1.00171e+00]
    np.testing.assert_allclose(a["condhist"], cond, rtol=6e-6)
    np.testing.assert_allclose(a["alphanorm"], alph, rtol=6e-6)
    assert a["idid"] and a["errcode"] == 0
    np.testing.assert_allclose(a["solution"], a["functionval"], atol=1e-8)   # test_anderson.jl:19


def test_aasol_toth_kelley_beta_half():
    b = po.aasol(tothk, np.ones(2), m=2, rtol=1e-10, beta=0.5)
    ref = [3.25055e-01, 3.70140e-02, 1.81111e-03, 9.55308e-04, 1.25936e-05, 1.40854e-09, 2.18196e-12]
    np.testing.assert_allclose(b["history"], ref, rtol=6e-6)
    assert b["idid"] and b["errcode"] == 0


def test_aasol_picard_warmup():   # test_anderson.jl:30-36
    a = po.aasol(tothk, np.ones(2), m=2, rtol=1e-10, picard_maxit=5, picard_beta=0.5)
    assert a["idid"] and a["errcode"] == 0 and a["history"][-1] < 1e-10


def test_aasol_errcodes():
    # maxit reached -> 10
    a = po.aasol(tothk, np.ones(2), m=2, rtol=1e-14, atol=1e-14, maxit=2)
    assert a["errcode"] == 10 and not a["idid"]
    # initial iterate already a fixed point -> -1
    x = po.aasol(tothk, np.ones(2), m=2, rtol=1e-10, maxit=50)["solution"]
    b = po.aasol(tothk, x, m=2, rtol=1e-3, atol=1e-3)
    assert b["errcode"] == -1 and b["idid"]
    # divergence guard -> -2: an expanding map
    c = po.aasol(lambda u: 3.0 * u + 1.0, np.ones(3), m=0, maxit=50)
    assert c["errcode"] == -2 and not c["idid"]


# --- Phi and its gradient against 50-digit sympy goldens (tests/golden) ------------
def _golden_cases():
    with open(os.path.join(HERE, "golden", "phi_golden.json")) as fh:
        return json.load(fh)["cases"]


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_phi_gradient_sympy(case):
    nc, nd = case["nc"], case["nd"]
    o = po.Oracle(nc=nc, nd=nd, ngrid=[4] * nd, lb=[0] * nd, ub=[1] * nd, T=case["T"], L=case["L"],
                  m=case["m"], d=case["d"], sigma=case["sigma"], eps=case["eps"], mu_res=[0] * nc,
                  rho_bulk=[1] * nc)
    for p in case["points"]:
        v, g = o.phi(np.array(p["n"]).reshape(-1))
        ref = np.array([float(x) for r in p["dphi"] for x in r])
        assert abs(v - float(p["phi"])) <= 1e-13 * abs(float(p["phi"]))
        np.testing.assert_allclose(g, ref, rtol=1e-12, atol=1e-13 * np.max(np.abs(ref)))


def test_phi_gradient_finite_difference():
    case = _golden_cases()[1]
    nc, nd = case["nc"], case["nd"]
    o = po.Oracle(nc=nc, nd=nd, ngrid=[4], lb=[0], ub=[1], T=case["T"], L=case["L"], m=case["m"],
                  d=case["d"], sigma=case["sigma"], eps=case["eps"], mu_res=[0] * nc, rho_bulk=[1] * nc)
    n0 = np.array(case["points"][0]["n"]).reshape(-1)
    _, g = o.phi(n0)
    for k in range(n0.size):
        h = 1e-5 * max(abs(n0[k]), 1e-3)
        e = np.zeros_like(n0)
        e[k] = h
        fd = (o.phi(n0 + e)[0] - o.phi(n0 - e)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))


# --- bulk consistency (test_coordinate_systems.jl:60-71, test_models.jl:11-27) ------
@pytest.mark.parametrize("comps,T,p,x,nd,N", [
    (["methane"], 150.0, 1e7, [1.0], 1, 101),
    (["methane", "n-decane"], 300.0, 2e7, [0.5, 0.5], 1, 64),
    (["methane", "ethane", "propane"], 200.0, 1e7, [0.6, 0.3, 0.1], 3, 6),
])
def test_bulk_consistency(comps, T, p, x, nd, N):
    m = PCSAFT(comps)
    x = np.array(x)
    v = m.volume(p, T, x, "liquid")
    rb = x / v
    mu = m.VT_chemical_potential_res(v, T, x) / RGAS / T
    L = m.length_scale()
    nc = len(comps)
    o = po.Oracle(nc=nc, nd=nd, ngrid=[N] * nd, lb=[-10 * L] * nd, ub=[10 * L] * nd, T=T, L=L, m=m.segment,
                  d=m.d(T), sigma=m.sigma, eps=m.epsilon, mu_res=mu, rho_bulk=rb)
    df, fv = o.dFdrho_res(np.repeat(rb, o.N))
    df = df.reshape(nc, -1)
    np.testing.assert_allclose(df, np.repeat(mu[:, None], o.N, axis=1), rtol=0, atol=1e-12)
    # Phi of the uniform fluid = a_res * number density (reduced units)
    ares = m.a_res(v, T, x)
    assert abs(fv[0] / L ** 3 - ares * 6.02214076e23 / v) <= 1e-12 * abs(fv[0] / L ** 3)


# --- FFT against numpy.fft (library cross-check), incl. prime and mixed sizes -------
@pytest.mark.parametrize("shape", [(8,), (201,), (97,), (2048,), (12, 10, 7)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    # oracle arrays are column-major: pass the transpose's memory
    af = np.asfortranarray(a)
    re = np.ascontiguousarray(af.real.ravel(order="F"))
    im = np.ascontiguousarray(af.imag.ravel(order="F"))
    ng = (po.C.c_int * 3)(*(list(shape) + [1] * (3 - len(shape))))
    po.lib().or_fft(po._ptr(re), po._ptr(im), len(shape), ng, 0)
    ref = np.fft.fftn(a)
    got = (re + 1j * im).reshape(shape, order="F")
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    po.lib().or_fft(po._ptr(re), po._ptr(im), len(shape), ng, 1)
    back = (re + 1j * im).reshape(shape, order="F")
    assert np.max(np.abs(back - a)) <= 1e-13 * np.max(np.abs(a))


def _c0_oracle(N=201, vext=True):
    from cdft_amd import configs
    s, rho = configs.methane_wall(ngrid=N)
    d = s.descriptor()
    if not vext:
        d["vext"] = None
    return s, rho, po.Oracle(**d)


# --- weights: closed-form Fourier transforms and k = 0 limits -----------------------
def test_omega_closed_forms():
    from scipy.integrate import quad
    s, rho, o = _c0_oracle()
    om = o.omega().reshape(1, 7, -1)[0]
    N = s.N
    lb, ub = s.structure.bounds[0]
    dd = s.species.size[0] / s.L
    fk = np.fft.fftfreq(N) * N * s.L / (ub - lb)          # reduced frequencies
    for k in (0, 1, 7, 50, N // 2):
        f = fk[k]
        # int rho dz (d/2): FT of the indicator of |z| <= w
        w = 0.5 * dd
        ind = quad(lambda z: math.cos(2 * math.pi * f * z), -w, w, epsabs=1e-15, epsrel=1e-13)[0]
        assert abs(om[1][k].real - ind) <= 1e-12 * max(1.0, abs(ind))
        # int rho z^2 dz (d): 3D FT of the sphere indicator, radius d
        R = dd
        sph = quad(lambda r: 4 * math.pi * r * r * (np.sinc(2 * f * r)), 0, R, epsabs=1e-15, epsrel=1e-13)[0]
        assert abs(om[4][k].real - sph) <= 1e-12 * max(1.0, abs(sph))
    # k = 0 limits: 2w and 4 pi R^3/3; vector weight has no real part at k = 0
    assert abs(om[1][0].real - 2 * 0.5 * dd) < 1e-15
    assert abs(om[6][0].real - 4 * math.pi / 3 * (1.3862 * dd) ** 3) < 1e-13
    assert om[3][0].real == 0.0


def test_transposed_is_adjoint():
    """<W_f rho, u> = <rho, W_f^T u> per field, with the reference's sign on the vector slot."""
    s, rho, o = _c0_oracle(vext=False)
    rng = np.random.default_rng(3)
    r = rng.uniform(0.5, 1.5, s.N) * 1e4
    n = o.weighted_densities(r).reshape(7, -1)
    NA = 6.02214076e23 * s.L ** 3
    for f in range(1, 7):
        u = rng.standard_normal(s.N)
        dphi = np.zeros((7, s.N))
        dphi[f] = u
        t = o.transposed(dphi.reshape(-1))          # includes the (zero) rho slot
        lhs = np.dot(n[f] / NA, u)
        rhs = np.dot(r, t)
        assert abs(lhs - rhs) <= 1e-12 * (np.abs(n[f] / NA).sum() * np.abs(u).max())


def test_simpson_pattern():
    """Weights 1,4,2,...,1 times dz/3 (integrals.jl:28-40), also for even N."""
    for N in (5, 6, 201, 2048):
        f = np.zeros(N)
        got = []
        for i in range(N):
            f[:] = 0
            f[i] = 1
            got.append(po.lib().or_simpson(po._ptr(f), 1, (po.C.c_int * 3)(N, 1, 1), po._ptr(np.array([3.0, 1, 1]))))
        ref = [1.0] + [4.0 if (k % 2 == 0) else 2.0 for k in range(2, N)] + [1.0]
        assert got == ref


def test_uniform_bulk_tension_is_zero():
    """gamma of a uniform bulk slab vanishes when F_ideal, F_res, mu and p are consistent."""
    m = PCSAFT(["methane", "n-decane"])
    T, p, x = 300.0, 2e7, np.array([0.5, 0.5])
    v = m.volume(p, T, x, "liquid")
    rb = x / v
    L = m.length_scale()
    mu_res = m.VT_chemical_potential_res(v, T, x) / RGAS / T
    o = po.Oracle(nc=2, nd=1, ngrid=[65], lb=[-5 * L], ub=[5 * L], T=T, L=L, m=m.segment, d=m.d(T),
                  sigma=m.sigma, eps=m.epsilon, mu_res=mu_res, rho_bulk=rb)
    rho = np.repeat(rb, 65)
    mu = m.VT_chemical_potential(v, T, x)
    g = o.surface_tension(rho, mu, p)
    scale = abs(p * 10 * L)
    assert abs(g) <= 1e-9 * scale


def test_uniform_bulk_is_fixed_point():
    """Without a field the bulk profile satisfies G(x) = x: converge! returns errcode -1."""
    m = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = m.volume(p, T, [1.0], "liquid")
    L = m.length_scale()
    mu = m.VT_chemical_potential_res(v, T, [1.0]) / RGAS / T
    o = po.Oracle(nc=1, nd=1, ngrid=[33], lb=[0], ub=[8 * L], T=T, L=L, m=m.segment, d=m.d(T),
                  sigma=m.sigma, eps=m.epsilon, mu_res=mu, rho_bulk=[1 / v])
    rho, res, hist = o.converge(np.full(33, 1 / v), tol=1e-8)
    assert res.errcode == -1 and res.idid == 1
    assert hist[0] < 1e-12


def test_oracle_converge_wall_small():
    """A converged wall profile satisfies the Euler-Lagrange equation to the tolerance."""
    s, rho0, o = _c0_oracle(N=201)
    rho, res, hist = o.converge(rho0.reshape(-1), tol=1e-8)
    assert res.errcode == 0 and res.idid == 1
    lnG = o.get_new_profile(rho)
    x = np.log(rho)
    assert 1e-2 * np.linalg.norm(lnG - x) <= res.tol * 1.5
    # qualitative picture of docs/src/assets/getting_started_profile.png: contact peaks
    # of rho m d^3 N_A near 5.6 at both walls and a mid-pore value near 0.67
    red = rho * s.model.segment[0] * s.species.size[0] ** 3 * 6.02214076e23
    assert 4.5 < red[:40].max() < 7.0 and 4.5 < red[-40:].max() < 7.0
    assert 0.4 < red[100] < 0.9
