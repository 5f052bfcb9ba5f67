"""Parity of the plain FP64 HIP path (cdft_functional_eval / cdft_converge) with the
CPU oracle, through the C ABI.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTSystem, PCSAFT, TwoPhase1DCart, Uniform1DCart, configs, converge_
from cdft_amd.bulk import RGAS
from helpers import oracle_of, smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def c0():
    s, rho = configs.methane_wall()
    return s, rho, oracle_of(s)


def test_functional_eval_wall_seed(c0):
    s, rho, o = c0
    say("functional eval: methane wall seed, N=201")
    got, fres, fval = s.device_system.functional_eval(rho, want_fval=True)
    ref, fv = o.dFdrho_res(rho)
    e_ = rel(got, ref)
    assert e_ < 1e-12, e_
    e_ = rel(fval, fv)
    assert e_ < 1e-12, e_
    assert abs(fres - o.F_res(rho)) <= 1e-12 * abs(fres)


@pytest.mark.parametrize("N", [97, 200, 201, 1024, 2048, 3001])
def test_functional_eval_smooth_profiles(N):
    say(f"functional eval: smooth binary profile N={N}")
    m = PCSAFT(["methane", "n-decane"])
    T, x = 300.0, np.array([0.5, 0.5])
    v = m.volume(1.5e7, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform1DCart((1.5e7, T), x / v, [-20 * L, 20 * L], N))
    o = oracle_of(s)
    for seed in range(2):
        rho = smooth_profile(s, seed)
        got, fres, _ = s.device_system.functional_eval(rho)
        ref, _ = o.dFdrho_res(rho)
        e_ = rel(got, ref)
        assert e_ < 1e-12, (seed, e_)
        assert abs(fres - o.F_res(rho)) <= 1e-11 * abs(fres)


def test_functional_eval_interface_seed():
    say("functional eval: methane/decane interface seed, N=2048")
    s, rho = configs.c1c10_interface()
    o = oracle_of(s)
    got, fres, _ = s.device_system.functional_eval(rho)
    ref, _ = o.dFdrho_res(rho)
    e_ = rel(got, ref)
    assert e_ < 1e-12, e_
    assert abs(fres - o.F_res(rho)) <= 1e-11 * abs(fres)


def test_bulk_consistency_on_device():
    """Uniform profile: delta F_res/delta rho equals the EOS mu_res (test_coordinate_systems.jl:60-71)."""
    say("bulk consistency on device")
    m = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = m.volume(p, T, [1.0], "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform1DCart((p, T), [1 / v], [-10 * L, 10 * L], 101))
    got, _, _ = s.device_system.functional_eval(np.full(101, 1 / v))
    mu = m.VT_chemical_potential_res(v, T, [1.0])[0] / RGAS / T
    assert np.max(np.abs(got - mu)) < 1e-12


def test_converge_wall_parity(c0):
    """Same algorithm, same inputs: residual histories agree while rounding has not
    yet separated the paths, and the converged profiles agree to the tolerance the
    tight solve fixes."""
    s, rho0, o = c0
    say("converge: methane wall tol=1e-8 (GPU)")
    rho = rho0.copy()
    converge_(s, rho, tol=1e-8, history=20000)
    r = s.last_result
    say(f"  GPU: errcode {r['errcode']} evals {r['nevals']} res {r['resnorm']:.3e}")
    say("converge: methane wall tol=1e-8 (oracle)")
    rho_o, ro, hist_o = o.converge(rho0.reshape(-1), tol=1e-8)
    assert r["errcode"] == 0 and ro.errcode == 0
    h = r["history"]
    k = min(30, len(h), len(hist_o))
    np.testing.assert_allclose(h[:k], hist_o[:k], rtol=1e-8)
    assert abs(r["nevals"] - ro.nevals) <= max(5, 0.1 * ro.nevals)
    np.testing.assert_allclose(rho.reshape(-1), rho_o, rtol=1e-5)


def test_converge_wall_tight(c0):
    s, rho0, o = c0
    say("converge: methane wall tol=1e-12 (GPU and oracle)")
    rho = rho0.copy()
    converge_(s, rho, tol=1e-12)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-12)
    assert s.last_result["errcode"] == 0 and ro.errcode == 0
    e_ = rel(rho.reshape(-1), rho_o)
    say(f"  GPU evals {s.last_result['nevals']} res {s.last_result['resnorm']:.3e} | oracle evals {ro.nevals} "
        f"res {ro.resnorm:.3e} | profile rel diff {e_:.2e}")
    assert e_ < 1e-10, e_     # the plain FP64 path's parity bar on converged profiles


def test_converge_errcodes(c0):
    s, rho0, o = c0
    say("converge: maxit=2 -> errcode 10")
    rho = rho0.copy()
    converge_(s, rho, tol=1e-8, maxit=2)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-8, maxit=2)
    assert s.last_result["errcode"] == 10 == ro.errcode
    assert s.last_result["nevals"] == ro.nevals
    e_ = rel(rho.reshape(-1), rho_o)
    assert e_ < 1e-10, e_
    say("converge: uniform bulk -> errcode -1")
    m = PCSAFT(["methane"])
    v = m.volume(1e7, 150.0, [1.0], "liquid")
    L = m.length_scale()
    sb = DFTSystem(m, Uniform1DCart((1e7, 150.0), [1 / v], [0, 8 * L], 33))
    rb = np.full((1, 33), 1 / v)
    converge_(sb, rb, tol=1e-8)
    assert sb.last_result["errcode"] == -1


def test_converge_interface_small_parity():
    say("converge: methane/decane interface N=256, tol=1e-11 (GPU and oracle)")
    s, rho0 = configs.c1c10_interface(ngrid=256, box=10.0)
    o = oracle_of(s)
    rho = rho0.copy()
    converge_(s, rho, tol=1e-11)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-11)
    assert s.last_result["errcode"] == 0 and ro.errcode == 0
    e_ = rel(rho.reshape(-1), rho_o)
    say(f"  profile rel diff {e_:.2e}")
    assert e_ < 1e-7, e_   # the interface position drifts at the solver tolerance
