"""Differentiation on the device through the C ABI: the pullback of Phi (cdft_phi_vjp),
the Hessian-vector product of F_res (cdft_hvp) and parameter sensitivities through
the fixed-point adjoint (cdft_adjoint_solve), checked against the CPU oracle, central
finite differences and finite differences of full converge runs.  Needs a GPU."""
import sys
import time

import numpy as np
import pytest

from cdft_amd import configs, converge_
from cdft_amd import adjoint as adj
from helpers import oracle_of, smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def wall():
    s, rho = configs.methane_wall()
    return s, rho.reshape(-1), oracle_of(s)


def test_phi_vjp_unit_cotangent_is_dfdrho(wall):
    s, rho, _ = wall
    say("phi_vjp with cot = 1 equals dF_res/drho, N=201")
    dfd, _, _ = s.device_system.functional_eval(rho)
    got = s.device_system.phi_vjp(rho, np.ones(s.N))
    e_ = rel(got, dfd)
    assert e_ < 1e-13, e_


def test_phi_vjp_matches_oracle(wall):
    s, rho, o = wall
    say("phi_vjp with Simpson-weight cotangent vs oracle pullback")
    cot = adj.simpson_weights(s.N, s.structure.dz()) / s.structure.dz()
    got = s.device_system.phi_vjp(rho, cot)
    n = o.weighted_densities(rho)
    F, N, nc = o.F, o.N, o.nc
    dphi = np.zeros(N * F * nc)
    for j in range(N):
        nl = np.array([n[(c * F + f) * N + j] for c in range(nc) for f in range(F)])
        _, dl = o.phi(nl)
        for c in range(nc):
            for f in range(F):
                dphi[(c * F + f) * N + j] = dl[c * F + f] * cot[j]
    ref = o.transposed(dphi)
    e_ = rel(got, ref)
    assert e_ < 1e-12, e_


def _fd_dfdrho(fun, rho, v, h):
    """Fourth-order central difference of fun along v (Richardson on h, h/2)."""
    def c(hh):
        return (fun(rho + hh * v) - fun(rho - hh * v)) / (2 * hh)
    return (4 * c(h / 2) - c(h)) / 3


@pytest.mark.parametrize("which", ["wall", "binary"])
def test_hvp_matches_finite_differences(wall, which):
    if which == "wall":
        s, rho, o = wall
    else:
        s, _ = configs.c1c10_interface(ngrid=256, box=8.0)
        rho = smooth_profile(s, seed=3).reshape(-1)
        o = oracle_of(s)
    say(f"hvp vs central differences of dF_res/drho ({which}, N={s.N}, nc={s.nc})")
    rng = np.random.default_rng(7)
    v = smooth_profile(s, seed=11).reshape(-1) * rng.uniform(0.5, 1.5)
    u = smooth_profile(s, seed=12).reshape(-1)
    dev = s.device_system
    hv = dev.hvp(rho, v)
    h = 1e-3 * np.max(rho) / np.max(np.abs(v))
    fd_dev = _fd_dfdrho(lambda r: dev.functional_eval(r)[0], rho, v, h)
    fd_orc = _fd_dfdrho(lambda r: o.dFdrho_res(r)[0], rho, v, h)
    e1, e2 = rel(hv, fd_dev), rel(hv, fd_orc)
    say(f"  rel err vs device FD {e1:.2e}, vs oracle FD {e2:.2e}")
    assert e1 < 1e-7, e1
    assert e2 < 1e-7, e2
    # the Hessian of F_res is symmetric
    a, b = u @ hv, v @ dev.hvp(rho, u)
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b)), (a, b)


def _converged(s, rho0, tol=1e-12, floor=None):
    """Converge from rho0.  A free interface (no external field) drifts along its
    translation mode, where the damped iteration barely contracts; with `floor`
    set, a run that stops at maxit with the residual below it is accepted."""
    r = rho0.copy().reshape(s.nc, s.N)
    converge_(s, r, tol=tol, maxit=20000)
    res = s.last_result
    ok = res["errcode"] == 0 or (floor is not None and res["errcode"] == 10 and res["resnorm"] < floor)
    assert ok, res
    return r.reshape(-1)


def _fd_full(s, rho_star, kind, comp, prm, i, rel_h, tol, floor=None):
    base = adj.model_params(s.model)[s.model.components[i]][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
    h = rel_h * abs(base)
    vals = []
    for sg in (+1, -1):
        sp = adj.perturbed_system(s, i, prm, sg * h)
        r = _converged(sp, rho_star, tol, floor)
        vals.append(adj.objective(sp, r, kind, comp))
    return (vals[0] - vals[1]) / (2 * h)


def test_adjoint_adsorption_sensitivities(wall):
    s, rho0, _ = wall
    say("adsorption sensitivities of the methane wall: adjoint vs FD of full converge runs")
    rho = _converged(s, rho0)
    t = time.time()
    sens, lam, res = adj.sensitivities(s, rho, "adsorption", 0, tol=1e-12)
    t_adj = time.time() - t
    assert res["errcode"] == 0, res
    for key, val in sens.items():
        t = time.time()
        fd = _fd_full(s, rho, "adsorption", 0, key[0], key[1], 1e-5, 1e-12)
        say(f"  d Gamma / d {key}: adjoint {val:.10e}  FD {fd:.10e}  rel {abs(val - fd) / abs(fd):.2e}  "
            f"(adjoint total {t_adj:.2f} s, FD pair {time.time() - t:.2f} s)")
        assert abs(val - fd) <= 2e-5 * abs(fd), (key, val, fd)


def test_adjoint_gamma_sensitivities_binary():
    say("surface tension sensitivities, methane/n-decane interface N=512: adjoint vs FD of full runs")
    s, rho0 = configs.c1c10_interface(ngrid=512, box=10.0)
    rho = _converged(s, rho0.reshape(-1), tol=1e-12)
    sens, lam, res = adj.sensitivities(s, rho, "gamma", tol=1e-10)
    say(f"  adjoint: {res}")
    assert res["errcode"] == 0, res
    for key, val in sens.items():
        fd = _fd_full(s, rho, "gamma", 0, key[0], key[1], 1e-5, 1e-12, floor=1e-9)
        say(f"  d gamma / d {key}: adjoint {val:.10e}  FD {fd:.10e}  rel {abs(val - fd) / abs(fd):.2e}")
        assert abs(val - fd) <= 1e-4 * abs(fd), (key, val, fd)


def test_surface_tension_and_adsorption_abi(wall):
    s, rho, _ = wall
    say("cdft_surface_tension / cdft_adsorption vs the host post-processing")
    from cdft_amd.methods import adsorption_system, surface_tension_system
    st = s.structure
    p, T = st.conditions
    rl = st.rho_bulk
    mu = s.model.VT_chemical_potential(1.0 / rl.sum(), T, rl / rl.sum())
    g_dev = s.device_system.surface_tension(rho, mu, p)
    g_host = surface_tension_system(s, rho)
    assert abs(g_dev - g_host) <= 1e-12 * abs(g_host), (g_dev, g_host)
    a_dev = s.device_system.adsorption(rho)
    a_host = adsorption_system(s, rho)
    assert np.max(np.abs(a_dev - a_host)) <= 1e-14 * np.max(np.abs(a_host))
