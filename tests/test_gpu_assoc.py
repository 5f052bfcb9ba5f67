"""Wertheim association (src/models/DFT/assoc.jl:702-831) on the HIP path: the
functional and its derivatives against the oracle, the bulk invariant, and the
reference's stored interfacial tensions (test/test_methods.jl:3-25) solved on the
device.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import (DFTOptions, DFTSystem, PCSAFT, TwoPhase1DCart, Uniform1DCart, configs, converge_,
                      interfacial_tension, surface_tension)
from cdft_amd.bulk import RGAS
from cdft_amd.methods import bulk_chemical_potential
from cdft_amd.system import initialize_profiles
from helpers import oracle_of, smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


# (components, composition, combining, T, p): one pair in closed form (AS = 1) or the
# general solver (AS = 2: two pairs without cross association, four with cr1)
MIXES = {
    "water": (["water"], [1.0], "nocombining", 350.0, 1e5),
    "ethanol_hexane": (["ethanol", "hexane"], [0.5, 0.5], "nocombining", 298.15, 1e5),
    "water_ethanol": (["water", "ethanol"], [0.5, 0.5], "nocombining", 320.0, 1e5),
    "water_ethanol_cr1": (["water", "ethanol"], [0.5, 0.5], "cr1", 320.0, 1e5),
}


def _system_1d(name, N, box=8.0, options=None):
    comps, x, comb, T, p = MIXES[name]
    m = PCSAFT(comps, combining=comb)
    x = np.asarray(x)
    v = m.volume(p, T, x, "liquid")
    L = m.length_scale()
    return DFTSystem(m, Uniform1DCart((p, T), x / v, [-box * L, box * L], N), None, options)


@pytest.mark.parametrize("name", list(MIXES))
@pytest.mark.parametrize("N", [97, 256])
def test_functional_eval_assoc_1d(name, N):
    say(f"association functional eval 1D: {name}, N={N}")
    s = _system_1d(name, N)
    o = oracle_of(s)
    for seed in range(2):
        rho = smooth_profile(s, seed, amp=0.5)
        got, fres, fval = s.device_system.functional_eval(rho, want_fval=True)
        ref, fv = o.dFdrho_res(rho)
        e_ = rel(got, ref)
        say(f"  seed {seed}: dF/drho rel err {e_:.2e}")
        assert e_ < 1e-12, (seed, e_)
        assert rel(fval, fv) < 1e-12
        assert abs(fres - o.F_res(rho)) <= 1e-11 * abs(fres)


@pytest.mark.parametrize("name", ["ethanol_hexane", "water_ethanol_cr1"])
def test_functional_eval_assoc_3d(name):
    comps, x, comb, T, p = MIXES[name]
    say(f"association functional eval 3D: {name}, 16^3, two hard spheres")
    s, _ = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1, T=T, p=p, x=x, components=comps, combining=comb)
    o = oracle_of(s)
    rho = smooth_profile(s, seed=5, amp=0.5).reshape(-1)
    got, fres, _ = s.device_system.functional_eval(rho)
    ref, _ = o.dFdrho_res(rho)
    e_ = rel(got, ref)
    say(f"  dF/drho rel err {e_:.2e}")
    assert e_ < 1e-11, e_
    assert abs(fres - o.F_res(rho)) <= 1e-11 * abs(fres)


@pytest.mark.parametrize("name", list(MIXES))
def test_bulk_consistency_assoc_on_device(name):
    """Uniform liquid: dF_res/drho equals the equation of state's beta mu_res
    (test/test_models.jl:153-188)."""
    say(f"association bulk consistency on device: {name}")
    s = _system_1d(name, 33)
    comps, x, comb, T, p = MIXES[name]
    rho = initialize_profiles(s)
    got, _, _ = s.device_system.functional_eval(rho)
    rb = s.structure.rho_bulk
    mu = s.model.VT_chemical_potential_res(1 / rb.sum(), T, rb / rb.sum()) / RGAS / T
    np.testing.assert_allclose(got.reshape(len(comps), -1), np.repeat(mu[:, None], 33, axis=1), rtol=1e-11)


def _fd_dfdrho(fun, rho, v, h):
    def c(hh):
        return (fun(rho + hh * v) - fun(rho - hh * v)) / (2 * hh)
    return (4 * c(h / 2) - c(h)) / 3


@pytest.mark.parametrize("name", ["ethanol_hexane", "water_ethanol_cr1"])
def test_hvp_assoc_matches_finite_differences(name):
    say(f"association hvp vs central differences: {name}, N=128")
    s = _system_1d(name, 128)
    rho = smooth_profile(s, seed=3, amp=0.5).reshape(-1)
    v = smooth_profile(s, seed=11).reshape(-1)
    u = smooth_profile(s, seed=12).reshape(-1)
    dev = s.device_system
    hv = dev.hvp(rho, v)
    h = 1e-3 * np.max(rho) / np.max(np.abs(v))
    fd = _fd_dfdrho(lambda r: dev.functional_eval(r)[0], rho, v, h)
    e_ = rel(hv, fd)
    say(f"  rel err vs FD {e_:.2e}")
    assert e_ < 1e-7, e_
    a, b = u @ hv, v @ dev.hvp(rho, u)
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b)), (a, b)


def test_phi_vjp_unit_cotangent_assoc():
    s = _system_1d("water_ethanol_cr1", 128)
    rho = smooth_profile(s, seed=2, amp=0.5)
    dfd, _, _ = s.device_system.functional_eval(rho)
    got = s.device_system.phi_vjp(rho, np.ones(s.N))
    assert rel(got, dfd.reshape(-1)) < 1e-13


def test_reference_stored_surface_tension_on_device():
    """surface_tension(PCSAFT(["ethanol","hexane"]), 298.15, [0.5, 0.5]) ==
    0.019843355568868397 (test/test_methods.jl:3-11), converged by the HIP path."""
    say("stored answer: ethanol/hexane surface tension, converge! on the device")
    gamma = surface_tension(PCSAFT(["ethanol", "hexane"]), 298.15, [0.5, 0.5])
    say(f"  gamma {gamma!r} (stored 0.019843355568868397)")
    assert abs(gamma / 0.019843355568868397 - 1) < 1e-6


def test_reference_stored_interfacial_tension_on_device():
    """interfacial_tension(PCSAFT(["water","hexane"]), 1e5, 298.15, x1, x2) ==
    0.030741511241317077 (test/test_methods.jl:13-25), converged by the HIP path."""
    say("stored answer: water/hexane interfacial tension, converge! on the device")
    m = PCSAFT(["water", "hexane"])
    xs, _, _ = m.tp_flash(1e5, 298.15, [0.5, 0.5], equilibrium="lle")
    gamma = interfacial_tension(m, 1e5, 298.15, xs[0], xs[1])
    say(f"  gamma {gamma!r} (stored 0.030741511241317077)")
    assert abs(gamma / 0.030741511241317077 - 1) < 1e-4


def test_converge_assoc_interface_matches_oracle():
    """The ethanol/hexane interface converged tightly on the device and on the oracle."""
    say("ethanol/hexane interface, N=201, tol 1e-9: device vs oracle")
    T, x = 298.15, np.array([0.5, 0.5])
    m = PCSAFT(["ethanol", "hexane"])
    p, vl, vv, y = m.bubble_pressure(T, x)
    L = m.length_scale()
    s = DFTSystem(m, TwoPhase1DCart((p, T), x / vl, y / vv, [-10 * L, 10 * L], (201,)))
    rho0 = initialize_profiles(s)
    r = rho0.copy()
    converge_(s, r, tol=1e-6)
    o = oracle_of(s)
    ro, res, _ = o.converge(rho0.copy(), tol=1e-6)
    res_d = s.last_result
    say(f"  device errcode {res_d['errcode']} evals {res_d['nevals']} | oracle errcode {res.errcode} evals {res.nevals}")
    assert res_d["errcode"] == 0 and res.errcode == 0
    mu = bulk_chemical_potential(s)
    gd = s.device_system.surface_tension(r, mu, p) / 2
    go = o.surface_tension(ro, mu, p) / 2
    say(f"  gamma device {gd!r} oracle {go!r}")
    assert abs(gd / go - 1) < 1e-8
    assert rel(r.reshape(-1), ro.reshape(-1)) < 1e-5


def test_converge_assoc_3d_mixed_precision():
    """The FP32 map with association (3D mixed precision) reaches the FP64 solution.
    The sphere is small (radius 0.5-0.8 L): at 16^3 over 5 L a sphere of radius
    1-2.5 L is so under-resolved that the seed's surface rings in the spectral
    convolutions (S < 0 inside) and neither precision converges, on the oracle either."""
    say("3D converge with association, mixed vs FP64: water/ethanol cr1, 16^3, one small hard sphere, tol 1e-9")
    kw = dict(n=16, box=5.0, nspheres=1, seed=2, T=320.0, p=1e5, x=(0.5, 0.5), components=("water", "ethanol"),
              combining="cr1", rmin=0.5, rmax=0.8)
    s64, rho0 = configs.matrix_3d(**kw)
    s32, _ = configs.matrix_3d(options=DFTOptions(precision="mixed"), **kw)
    r64, r32 = rho0.copy(), rho0.copy()
    converge_(s64, r64, tol=1e-9)
    converge_(s32, r32, tol=1e-9)
    a, b = s64.last_result, s32.last_result
    say(f"  FP64 errcode {a['errcode']} evals {a['nevals']} | mixed errcode {b['errcode']} evals {b['nevals']} "
        f"(FP32 {b['nevals_fp32']})")
    assert a["errcode"] == 0 and b["errcode"] == 0
    assert b["nevals_fp32"] > 0
    e_ = rel(r32.reshape(-1), r64.reshape(-1))
    say(f"  profile rel diff {e_:.2e}")
    assert e_ < 1e-7, e_


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_associating_systems_match_single_solves(precision):
    """Associating one-component systems (water, 64 and 96 points, seeds perturbed from
    the uniform liquid) in one batched launch run on the single-workgroup bodies.  FP64:
    bit for bit what cdft_converge gives for each alone.  Mixed precision: the FP32
    association code rounds differently in the batched and the single-solve kernel
    (each deterministic), and the FP32 phase's switch point is sensitive to rounding, so
    the two solves agree in termination and, converged at tol 1e-9, to 1e-7 in the
    profile (7e-9 measured) rather than bit for bit."""
    from cdft_amd import _lib
    say(f"batched associating water ({precision}) vs single solves, tol=1e-9")
    opts = DFTOptions(precision=precision)
    items = []
    for N in (64, 96):
        s = _system_1d("water", N, options=opts)
        r = initialize_profiles(s)
        z = np.arange(r.size)
        items.append((s, r * (1.0 + 0.05 * np.sin(2 * np.pi * z / N)).reshape(r.shape)))
    rb = [r.reshape(-1).copy() for _, r in items]
    res = _lib.converge_batch(items[0][0].device_system.ctx, [s.device_system for s, _ in items], rb,
                              tol=1e-9, maxit=5000)
    for (s, r0), b, rr in zip(items, rb, res):
        r1 = r0.copy()
        converge_(s, r1, tol=1e-9, maxit=5000)
        ra = s.last_result
        say(f"  N={s.N}: batch errcode {rr['errcode']} evals {rr['nevals']} (fp32 {rr['nevals_fp32']}) | "
            f"single errcode {ra['errcode']} evals {ra['nevals']} (fp32 {ra['nevals_fp32']})")
        assert ra["errcode"] == 0 and ra["nevals"] > 2 and rr["errcode"] == 0
        if precision == "fp64":
            assert (rr["nevals"], rr["nevals_fp32"]) == (ra["nevals"], ra["nevals_fp32"])
            assert np.array_equal(b, r1.reshape(-1))
        else:
            assert rr["nevals_fp32"] > 0
            assert rel(b, r1.reshape(-1)) < 1e-7
