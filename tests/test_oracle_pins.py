"""Further pins of the CPU oracle: the reference's stored adsorption answer, the
reference figure of the methane wall, the Steele wall against the exact sum over
Lennard-Jones lattice planes, the seeds against the reference's own seed test, the
NaN semantics of the exponent clamp, and the planned 1D evaluation against the
literal one.  CPU only."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import pyoracle as po
from cdft_amd import configs
from cdft_amd.bulk import PCSAFT
from cdft_amd.structure import STEELE_DB, Steele, TwoPhase1DCart, Uniform1DCart
from cdft_amd.system import DFTSystem, initialize_profiles
from helpers import oracle_of, smooth_profile


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


# --- the reference's stored on-path answer -------------------------------------------
def co2_methane_adsorption_system():
    """adsorption(PCSAFT(["carbon dioxide","methane"]), Steele(["graphite"], 40e-10),
    1e6, 298.15, [0.5, 0.5]) set-up (src/methods/adsorption.jl:48-66)."""
    model = PCSAFT(["carbon dioxide", "methane"])
    surface = Steele(["graphite"], 40e-10)
    p, T = 1e6, 298.15
    n = np.array([0.5, 0.5])
    L = model.length_scale()
    v = model.volume(p, T, n)
    st = Uniform1DCart((p, T), n / v, [0.7 * L, surface.width - 0.7 * L], 201)
    s = DFTSystem(model, st, surface)
    return s, initialize_profiles(s)


@pytest.mark.parametrize("planned", [False, True])
def test_reference_stored_adsorption_answer(planned):
    """test/test_methods.jl:37: ad[1] = 605.2164007273888 (rtol 1e-4) at the default
    tol 1e-4.  k_ij(CO2, CH4) = 0.065 (Gross & Sadowski 2001) reproduces it to ~1e-10."""
    s, rho = co2_methane_adsorption_system()
    assert s.model.kij[0, 1] == 0.065
    o = oracle_of(s)
    x, r, _ = o.converge(rho.reshape(-1), planned=planned)
    assert r.errcode == 0
    ads = o.adsorption(x)
    assert abs(ads[0] / 605.2164007273888 - 1) < 1e-8, ads[0]


# --- the reference figure of the methane wall (docs/src/assets/getting_started_profile.png)
def test_wall_profile_matches_reference_figure():
    """examples/docs_figures/01_getting_started.jl: rho m d^3 N_A (MakiecDFTExt.jl:31-34)
    over z/sigma; read off the figure: contact peaks 5.6 at 0.94 sigma and 4.06 sigma,
    first minima 0.14 at 1.30 sigma, second peaks 1.25 at 1.88 and 3.12 sigma, mid-pore
    0.67, empty below 0.83 sigma."""
    s, rho0 = configs.methane_wall()
    o = oracle_of(s)
    x, r, _ = o.converge(rho0.reshape(-1), planned=True)        # default tol 1e-4, as the figure
    assert r.errcode == 0
    sig = s.model.sigma[0, 0]
    y = x * s.model.segment[0] * s.species.size[0] ** 3 * 6.02214076e23
    z = s.structure.coords() / sig
    half = z < 2.5
    i1 = int(np.argmax(np.where(half, y, 0)))
    assert 0.92 <= z[i1] <= 0.96 and 5.4 <= y[i1] <= 5.85, (z[i1], y[i1])
    i2 = int(np.argmax(np.where(~half, y, 0)))
    assert 4.04 <= z[i2] <= 4.08 and abs(y[i2] - y[i1]) < 1e-6 * y[i1]
    win = (z > 1.1) & (z < 1.6)
    imin = np.nonzero(win)[0][np.argmin(y[win])]
    assert 1.26 <= z[imin] <= 1.34 and 0.11 <= y[imin] <= 0.17, (z[imin], y[imin])
    win = (z > 1.6) & (z < 2.2)
    ip = np.nonzero(win)[0][np.argmax(y[win])]
    assert 1.84 <= z[ip] <= 1.92 and 1.20 <= y[ip] <= 1.30, (z[ip], y[ip])
    mid = np.argmin(np.abs(z - 2.5))
    assert 0.64 <= y[mid] <= 0.70, y[mid]
    assert np.all(y[z < 0.8] < 0.05)


# --- Steele wall: closed form vs the exact sum over lattice planes -------------------
def lj_plane(zr, eps, sig, rho_s, delta):
    """Energy of a site at reduced distance zr = z/sig from one plane of areal density
    rho_s delta, by quadrature of the 12-6 Lennard-Jones potential over the plane."""
    f = lambda x: 4 * ((zr * zr + x * x) ** -6 - (zr * zr + x * x) ** -3) * 2 * math.pi * x
    return rho_s * delta * eps * sig * sig * quad(f, 0, np.inf, epsabs=0, epsrel=1e-12, limit=400)[0]


def test_steele_against_lattice_plane_sum():
    """The 10-4-3 wall (steele.jl:71-73) approximates the sum of 10-4 plane potentials
    over the graphite planes 0, Delta, 2 Delta, ...; with Lorentz-Berthelot mixing on
    sigma (not d) and 1/T it must agree with that sum to ~1.5 % near contact, for the
    host field and the oracle's."""
    s, _ = configs.methane_wall()
    model = s.model
    eps_s, sig_s, rho_s, delta = STEELE_DB["graphite"]
    sig_s, rho_s, delta = sig_s * 1e-10, rho_s * 1e30, delta * 1e-10
    e = math.sqrt(eps_s * model.epsilon[0, 0])
    sg = 0.5 * (sig_s + model.sigma[0, 0])
    H = s.external_field[0].width
    T = s.structure.conditions[1]
    z = s.structure.coords()
    host = s.vext
    orc = oracle_of(s).steele_vext(eps_s, sig_s, delta, rho_s, H)
    assert rel(orc, host) < 1e-13
    d = delta / sg
    for j in range(0, len(z), 10):
        zr, zr2 = z[j] / sg, (H - z[j]) / sg
        if min(zr, zr2) > 2.0 or min(zr, zr2) < 0.9:
            continue
        exact = sum(lj_plane(zr + k * d, e, sg, rho_s, delta) + lj_plane(zr2 + k * d, e, sg, rho_s, delta)
                    for k in range(300)) / T
        assert abs(host[j] - exact) < 0.02 * abs(exact), (z[j] / sg, host[j], exact)


# --- seeds ---------------------------------------------------------------------------
def test_wall_seed_matches_oracle():
    """Uniform seed times exp(-beta V/10), clamped (structure.jl:50-94, external_field.jl:21-26)."""
    for s, rho in [configs.methane_wall()] + configs.slit_pores(n=3, seed=1):
        o = oracle_of(s)
        assert rel(o.init_profiles(), rho.reshape(-1)) < 1e-15
        lo = np.finfo(np.float64).tiny ** (1 / 6)
        assert rho.min() >= lo


def test_two_phase_seed_reference_properties():
    """test/test_two_phase.jl:14-26: methane at its 150 K saturation pressure,
    TwoPhase1DCart over [-10L, 10L] with 51 points: the seed is positive and spans more
    than half the gap between the coexisting densities; and host == oracle."""
    model = PCSAFT(["methane"])
    T = 150.0
    p, vl, vv = model.saturation_pressure(T)
    rl, rv = np.array([1 / vl]), np.array([1 / vv])
    L = model.length_scale()
    s = DFTSystem(model, TwoPhase1DCart((p, T), rl, rv, [-10 * L, 10 * L], 51))
    rho = initialize_profiles(s)
    assert np.all(rho > 0)
    assert rho.max() - rho.min() > 0.5 * abs(rl[0] - rv[0])
    Tc = np.array([model.crit_pure()[0]])
    orc = oracle_of(s).init_two_phase(rv, Tc)
    assert rel(orc, rho.reshape(-1)) < 1e-13


# --- the exponent clamp keeps NaN (Julia's clamp), ending aasol with errcode 10 -------
def test_blow_up_ends_with_nan_residual():
    """Slit pore 40 of the 512-pore workload: the Anderson iterate blows up at the
    211th map evaluation (densities ~1e26, Phi not finite); clamp!(ln_Gx, -100, 100)
    (converge.jl:150) keeps the NaN, so the residual is NaN and AndersonOK returns 10
    (anderson.jl:727-737)."""
    s, rho = configs.slit_pores(n=512, seed=0, indices=[40])[0]
    o = oracle_of(s)
    for planned in (False, True):
        _, r, hist = o.converge(rho.reshape(-1), tol=1e-8, planned=planned)
        assert r.errcode == 10 and r.nevals == 211 and math.isnan(r.resnorm)
        assert math.isnan(hist[-1]) and np.all(np.isfinite(hist[:-1]))


# --- planned 1D evaluation == literal ------------------------------------------------
@pytest.mark.parametrize("N", [80, 97, 201, 256, 383])
def test_planned_matches_literal(N):
    m = PCSAFT(["methane", "n-decane"])
    T, x = 300.0, np.array([0.5, 0.5])
    v = m.volume(1.5e7, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform1DCart((1.5e7, T), x / v, [-8 * L, 8 * L], N))
    o = oracle_of(s)
    for seed in range(2):
        rho = smooth_profile(s, seed)
        a, fa = o.dFdrho_res(rho)
        b, fb = o.dFdrho_res_planned(rho)
        assert rel(b, a) < 1e-12 and rel(fb, fa) < 1e-13


def test_planned_converge_matches_literal():
    s, rho = configs.methane_wall()
    o = oracle_of(s)
    xa, ra, ha = o.converge(rho.reshape(-1), tol=1e-12)
    xb, rb, hb = o.converge(rho.reshape(-1), tol=1e-12, planned=True)
    # near the FP64 floor the two roundings take different Anderson paths: the
    # evaluation counts differ, the fixed point does not
    assert ra.errcode == rb.errcode == 0 and abs(ra.nevals - rb.nevals) < 0.25 * ra.nevals
    assert rel(xb, xa) < 1e-10
    np.testing.assert_allclose(hb[:40], ha[:40], rtol=1e-8)


def test_batch_driver_matches_single_solves():
    pores = configs.slit_pores(n=6, seed=0)
    ors = [oracle_of(s) for s, _ in pores]
    outs, res, secs, wall = po.converge_batch(ors, [r for _, r in pores], tol=1e-8, nthreads=3)
    for (s, r0), o, x, r in zip(pores, ors, outs, res):
        y, ry, _ = o.converge(r0.reshape(-1), tol=1e-8, planned=True)
        assert r.errcode == ry.errcode and r.nevals == ry.nevals
        assert np.array_equal(x, y)
    assert wall > 0 and np.all(secs > 0)
