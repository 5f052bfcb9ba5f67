"""Wertheim association (src/models/DFT/assoc.jl) on the CPU: the oracle's term
against 50-digit goldens, the bulk equation of state, and the reference's two
stored interfacial-tension answers (test/test_methods.jl:3-25)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from cdft_amd.bulk import RGAS, PCSAFT
from cdft_amd.methods import bulk_chemical_potential
from cdft_amd.structure import TwoPhase1DCart, Uniform1DCart
from cdft_amd.system import DFTSystem, initialize_profiles
from helpers import oracle_of

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden_cases():
    with open(os.path.join(HERE, "golden", "assoc_golden.json")) as fh:
        return json.load(fh)["cases"]


def _oracles(case):
    nc, nd = case["nc"], case["nd"]
    common = dict(nc=nc, nd=nd, ngrid=[4] * nd, lb=[0] * nd, ub=[1] * nd, T=case["T"], L=case["L"], m=case["m"],
                  d=case["d"], sigma=case["sigma"], eps=case["eps"], mu_res=[0] * nc, rho_bulk=[1] * nc)
    with_assoc = po.Oracle(**common, assoc_site_comp=case["site_comp"], assoc_site_count=case["site_count"],
                           assoc_pairs=case["pairs"], assoc_eps=case["pair_eps"], assoc_kappa=case["pair_kappa"])
    return with_assoc, po.Oracle(**common)


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_assoc_term_against_50_digit_golden(case):
    """Phi_assoc and its gradient = (oracle with association) - (oracle without)."""
    oa, o0 = _oracles(case)
    for p in case["points"]:
        n = np.array(p["n"])
        va, ga = oa.phi(n)
        v0, g0 = o0.phi(n)
        ref = np.array([float(x) for x in p["dphi"]])
        assert abs((va - v0) - float(p["phi"])) <= 1e-12 * abs(float(p["phi"])) + 1e-13 * abs(va)
        np.testing.assert_allclose(ga - g0, ref, rtol=1e-10, atol=1e-11 * np.max(np.abs(ref)))


def test_closed_form_single_pair_matches_iterative_solver():
    """One pair between two distinct sites takes the closed form (assoc.jl:759-764);
    adding a pair of zero strength sends the same point through the substitution +
    Newton solver (assoc.jl:238-411): same X and the same Phi."""
    case = _golden_cases()[0]
    oa, _ = _oracles(case)
    nc, nd = case["nc"], case["nd"]
    og = po.Oracle(nc=nc, nd=nd, ngrid=[4] * nd, lb=[0] * nd, ub=[1] * nd, T=case["T"], L=case["L"], m=case["m"],
                   d=case["d"], sigma=case["sigma"], eps=case["eps"], mu_res=[0] * nc, rho_bulk=[1] * nc,
                   assoc_site_comp=case["site_comp"], assoc_site_count=case["site_count"],
                   assoc_pairs=case["pairs"] + [[0, 1]], assoc_eps=case["pair_eps"] + [1000.0],
                   assoc_kappa=case["pair_kappa"] + [0.0])
    for p in case["points"]:
        n = np.array(p["n"])
        np.testing.assert_allclose(oa.assoc_X(n), og.assoc_X(n), rtol=1e-13)
        va, ga = oa.phi(n)
        vg, gg = og.phi(n)
        assert abs(va - vg) <= 1e-13 * abs(va)
        np.testing.assert_allclose(ga, gg, rtol=1e-11, atol=1e-12 * np.max(np.abs(ga)))


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_site_fractions_satisfy_mass_action(case):
    """X_s (1 + sum_t Delta_st rho_t n_t X_t) = 1 at the solved point, with the
    mixture quantities evaluated independently here."""
    oa, _ = _oracles(case)
    nc, nd, F = case["nc"], case["nd"], 6 + case["nd"]
    L = case["L"]
    dd = np.array(case["d"]) / L
    m = np.array(case["m"])
    for p in case["points"]:
        n = np.array(p["n"]).reshape(nc, F)
        X = oa.assoc_X(n.reshape(-1))
        n2m = np.sum(np.pi * n[:, 1] * m * dd)
        n3m = np.sum(n[:, 2] * m)
        nv2m = np.array([np.sum(-2 * np.pi * n[:, 3 + j] * m) for j in range(nd)])
        xim = 1 - nv2m @ nv2m / n2m ** 2
        xi = 1 - np.array([np.sum((2 * np.pi * n[i, 3:3 + nd]) ** 2) for i in range(nc)]) / (np.pi * n[:, 1] * dd) ** 2
        w = n[:, 1] / dd * xi
        v = np.array(case["site_count"]) * w[case["site_comp"]]
        C = np.zeros(len(X))
        for (a, b), e, k in zip(case["pairs"], case["pair_eps"], case["pair_kappa"]):
            ic, jc = case["site_comp"][a], case["site_comp"][b]
            sr = case["sigma"][ic * nc + jc] / L
            dij = dd[ic] * dd[jc] / (dd[ic] + dd[jc])
            a1 = 1 / (1 - n3m)
            g = a1 + 0.5 * dij * xim * n2m * a1 ** 2 + dij ** 2 * n2m ** 2 * xim * a1 ** 3 / 18
            delta = g * sr ** 3 * (np.exp(e / case["T"]) - 1) * k
            C[a] += delta * v[b] * X[b]
            C[b] += delta * v[a] * X[a]
        np.testing.assert_allclose(X * (1 + C), 1.0, atol=1e-13)


@pytest.mark.parametrize("comps,x,combining", [
    (["water"], [1.0], "nocombining"),
    (["ethanol"], [1.0], "nocombining"),
    (["ethanol", "hexane"], [0.5, 0.5], "nocombining"),
    (["water", "ethanol"], [0.5, 0.5], "nocombining"),
    (["water", "ethanol"], [0.5, 0.5], "cr1"),
])
def test_bulk_consistency_associating(comps, x, combining):
    """dF/drho_res of the uniform liquid equals the bulk beta mu_res
    (test/test_models.jl:153-188 compares the same pair at rtol 1e-4): the functional's
    association term against the equation of state's, coded independently (the bulk
    derivative is a complex step through the mass-action solve)."""
    p, T = 1e5, 298.15
    m = PCSAFT(comps, combining=combining)
    assert len(m.assoc_pairs) == {"nocombining": len(comps) - (1 if "hexane" in comps else 0), "cr1": 4}[combining]
    x = np.asarray(x)
    v = m.volume(p, T, x)
    L = m.length_scale()
    s = DFTSystem(m, Uniform1DCart((p, T), x / v, [-10 * L, 10 * L], (3,)))
    df, _ = oracle_of(s).dFdrho_res(initialize_profiles(s))
    mu = m.VT_chemical_potential_res(v, T, x) / RGAS / T
    np.testing.assert_allclose(df.reshape(len(comps), -1)[:, 0], mu, rtol=1e-12)


def test_bulk_association_limits():
    """a_assoc -> 0 at vanishing density, and the 2B pure-fluid closed form
    X = 2/(1 + sqrt(1 + 4 rho Delta)) (both sites equal) holds at a liquid density."""
    m = PCSAFT(["water"])
    T = 300.0
    c = m._consts(T, np.array([1.0]))
    assoc = c[-1]
    from cdft_amd.bulk import _a_assoc
    s = c[0]
    tiny = 1e-3
    assert abs(_a_assoc(np.array(tiny), s[2] * tiny, s[3] * tiny, assoc)) < 1e-20
    rho = 3.0e28
    z2, z3 = s[2] * rho, s[3] * rho
    a = _a_assoc(np.array(rho), z2, z3, assoc)
    d = m.d(T)[0]
    dij = d / 2
    g = 1 / (1 - z3) + dij * 3 * z2 / (1 - z3) ** 2 + dij ** 2 * 2 * z2 ** 2 / (1 - z3) ** 3
    delta = g * m.sigma[0, 0] ** 3 * (np.exp(2500.7 / T) - 1) * 0.034868
    X = 2 / (1 + np.sqrt(1 + 4 * rho * delta))
    assert abs(a - 2 * (np.log(X) - X / 2 + 0.5)) < 1e-13 * abs(a)


def test_reference_stored_surface_tension_ethanol_hexane():
    """surface_tension(PCSAFT(["ethanol","hexane"]), 298.15, [0.5,0.5]) ==
    0.019843355568868397 (test/test_methods.jl:3-11, rtol 1e-4 there): the bubble
    point, both critical points, the two-phase seed and converge! at its default
    tolerance, on the oracle."""
    T, x = 298.15, np.array([0.5, 0.5])
    m = PCSAFT(["ethanol", "hexane"])
    p, vl, vv, y = m.bubble_pressure(T, x)
    L = m.length_scale()
    s = DFTSystem(m, TwoPhase1DCart((p, T), x / vl, y / vv, [-10 * L, 10 * L], (201,)))
    o = oracle_of(s)
    rho, res, _ = o.converge(initialize_profiles(s))
    assert res.errcode == 0
    gamma = o.surface_tension(rho, bulk_chemical_potential(s), p) / 2
    assert abs(gamma / 0.019843355568868397 - 1) < 1e-6


def test_reference_stored_interfacial_tension_water_hexane():
    """interfacial_tension(PCSAFT(["water","hexane"]), 1e5, 298.15, x1, x2) ==
    0.030741511241317077 after an LLE flash (test/test_methods.jl:13-25, rtol 1e-3
    there), on the oracle."""
    p, T = 1e5, 298.15
    m = PCSAFT(["water", "hexane"])
    xs, _, _ = m.tp_flash(p, T, [0.5, 0.5], equilibrium="lle")
    assert xs[0, 0] > 0.99 and xs[1, 1] > 0.99
    v1, v2 = m.volume(p, T, xs[0]), m.volume(p, T, xs[1])
    L = m.length_scale()
    s = DFTSystem(m, TwoPhase1DCart((p, T), xs[0] / v1, xs[1] / v2, [-20 * L, 20 * L], (201,)))
    o = oracle_of(s)
    rho, res, _ = o.converge(initialize_profiles(s))
    assert res.errcode == 0
    gamma = o.surface_tension(rho, bulk_chemical_potential(s), p) / 2
    assert abs(gamma / 0.030741511241317077 - 1) < 1e-4
