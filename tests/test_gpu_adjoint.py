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


def test_adjoint_params_abi_matches_its_restatement():
    """cdft_adjoint_params on the binary interface (two objectives, Gamma_0 and
    Gamma_1, at once; two parameter pairs): its multipliers equal
    cdft_adjoint_solve's for each objective, and each output equals
    lambda_k . (G_plus - G_minus) / (2 h) assembled on the host from the two
    perturbed systems' log maps.  A multiplier given back (g = NULL) gives the same
    outputs without a solve; invalid pairs are refused."""
    from cdft_amd._lib import CDFTError
    say("cdft_adjoint_params: two objectives, two parameter pairs, methane/n-decane N=256")
    s, rho0 = configs.c1c10_interface(ngrid=256, box=10.0)
    rho = _converged(s, rho0.reshape(-1), tol=1e-12, floor=1e-9)
    n = rho.size
    w = adj.integration_weights(s.structure) / s.structure.volume()
    g = np.zeros((2, n))
    for k in range(2):
        g[k, k * s.N:(k + 1) * s.N] = rho[k * s.N:(k + 1) * s.N] * w
    t = adj.translation_mode(s, rho)
    g = g - np.outer(g @ t / (t @ t), t)
    lam = np.zeros((2, n))
    dev = s.device_system
    _, res = dev.adjoint_params(rho, lam, g=g, tol=1e-10)
    assert [r["errcode"] for r in res] == [0, 0], res
    for k in range(2):
        lk, rk = dev.adjoint_solve(rho, g[k], tol=1e-10)
        assert rk["nevals"] == res[k]["nevals"]
        assert np.array_equal(lk, lam[k])
    pairs, hs, want = [], [], []
    for i, prm in ((0, "epsilon"), (1, "sigma")):
        base = adj.model_params(s.model)[s.model.components[i]][{"sigma": 1, "epsilon": 2}[prm]]
        h = 1e-6 * base
        sp, sm = adj.perturbed_system(s, i, prm, h), adj.perturbed_system(s, i, prm, -h)
        pairs.append((sp, sm))
        hs.append(h)
        dG = (adj.log_map(sp, rho) - adj.log_map(sm, rho)) / (2 * h)
        want.append(lam @ dG)
    want = np.array(want).T                       # (objective, parameter)
    got, _ = dev.adjoint_params(rho, lam, pairs=[(a.device_system, b.device_system) for a, b in pairs], h=hs)
    say(f"  implicit parts:\n{got}\n  host restatement:\n{want}")
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.max(np.abs(want)))
    with pytest.raises(CDFTError):
        dev.adjoint_params(rho, lam, pairs=[(pairs[0][0].device_system, dev)], h=[hs[0]])
    with pytest.raises(CDFTError):
        dev.adjoint_params(rho, lam, pairs=[(pairs[0][0].device_system, pairs[0][1].device_system)], h=[0.0])
    other, _ = configs.c1c10_interface(ngrid=128, box=10.0)
    with pytest.raises(CDFTError):
        dev.adjoint_params(rho, lam, pairs=[(pairs[0][0].device_system, other.device_system)], h=[hs[0]])
