"""Parity of the 3D spectral path (tabulated weights, line FFTs along x, y, z) with the
CPU oracle through the C ABI: functional evaluation, the bulk invariant, the
Hessian-vector product and a small converged solve.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTSystem, PCSAFT, Uniform3DCart, configs, converge_
from cdft_amd.bulk import RGAS
from helpers import oracle_of, smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def ternary16():
    s, rho0 = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1)
    return s, rho0.reshape(-1), oracle_of(s)


def test_functional_eval_3d_smooth(ternary16):
    s, _, o = ternary16
    say("3D functional eval, ternary, 16^3, smooth profile")
    rho = smooth_profile(s, seed=5).reshape(-1)
    got, fres, fval = s.device_system.functional_eval(rho, want_fval=True)
    ref, fv = o.dFdrho_res(rho)
    e_ = rel(got, ref)
    say(f"  dF/drho rel err {e_:.2e}")
    assert e_ < 1e-11, e_
    assert rel(fval, fv) < 1e-11
    assert abs(fres - o.F_res(rho)) <= 1e-11 * abs(fres)


def test_functional_eval_3d_matrix_seed(ternary16):
    """Inside the spheres the seed density is ~1e-18 of the bulk, below the rounding
    level of any FFT of this profile, so the chain term's log(lambda / rho) there is
    rounding noise in both paths (and irrelevant: the exponent is clamped at -100).
    Compared outside the spheres, at the 1e-10 parity bar: next to the sphere
    surfaces the weighted densities still carry that noise at the 1e-11 level."""
    s, rho0, o = ternary16
    say("3D functional eval, ternary, 16^3, hard-sphere matrix seed")
    got, _, _ = s.device_system.functional_eval(rho0)
    ref, _ = o.dFdrho_res(rho0)
    free = s.vext == 0
    e_ = rel(got[free], ref[free])
    say(f"  dF/drho rel err outside the matrix {e_:.2e} (all points {rel(got, ref):.2e})")
    assert e_ < 1e-10, e_


def test_functional_eval_3d_anisotropic():
    say("3D functional eval, methane, 16 x 32 x 64 grid over a non-cubic box")
    m = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    v = m.volume(p, T, [1.0], "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform3DCart((p, T), [1 / v], [[0, 5 * L], [0, 7 * L], [-3 * L, 6 * L]], (16, 32, 64)))
    rho = smooth_profile(s, seed=9).reshape(-1)
    got, fres, _ = s.device_system.functional_eval(rho)
    ref, _ = oracle_of(s).dFdrho_res(rho)
    e_ = rel(got, ref)
    say(f"  dF/drho rel err {e_:.2e}")
    assert e_ < 1e-11, e_


@pytest.mark.parametrize("ngrid", [(16, 32, 128), (16, 16, 256)])
def test_functional_eval_3d_all_pairs_z_pass(ngrid):
    """z lines of 128 or 256 points take the all-pairs z pass on the transposed side
    (one launch accumulates the four field pairs of a component); checked against the
    oracle."""
    say(f"3D functional eval, ternary, {ngrid} grid (all-pairs z pass)")
    m = PCSAFT(["methane", "ethane", "propane"])
    T, p = 200.0, 1e7
    x = np.array([0.6, 0.3, 0.1])
    v = m.volume(p, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform3DCart((p, T), x / v, [[0, 4 * L], [0, 5 * L], [0, 9 * L]], ngrid))
    rho = smooth_profile(s, seed=11).reshape(-1)
    got, fres, _ = s.device_system.functional_eval(rho)
    ref, _ = oracle_of(s).dFdrho_res(rho)
    e_ = rel(got, ref)
    say(f"  dF/drho rel err {e_:.2e}")
    assert e_ < 1e-11, e_
    hv = s.device_system.hvp(rho, smooth_profile(s, seed=12).reshape(-1))
    assert np.all(np.isfinite(hv))


def test_converge_3d_mixed_all_pairs_z_pass():
    say("3D converge, mixed vs FP64 on a 16 x 16 x 128 grid (all-pairs z pass in both maps)")
    from cdft_amd import DFTOptions
    m = PCSAFT(["methane"])
    T, p = 150.0, 1e7
    vb = m.volume(p, T, [1.0], "liquid")
    L = m.length_scale()
    box = [[0, 4 * L], [0, 4 * L], [0, 12 * L]]
    out = {}
    for prec in ("fp64", "mixed"):
        s = DFTSystem(m, Uniform3DCart((p, T), [1 / vb], box, (16, 16, 128)), options=DFTOptions(precision=prec))
        z = (np.arange(s.N) // 256) / 128.0
        rho = (1 / vb) * (1.0 + 0.3 * np.cos(2 * np.pi * z))
        converge_(s, rho, tol=1e-9)
        out[prec] = (rho.reshape(-1), s.last_result)
        say(f"  {prec}: errcode {s.last_result['errcode']} evals {s.last_result['nevals']}")
    assert out["fp64"][1]["errcode"] == 0 and out["mixed"][1]["errcode"] == 0
    e_ = rel(out["mixed"][0], out["fp64"][0])
    say(f"  profile rel diff {e_:.2e}")
    assert e_ < 1e-7, e_


def test_bulk_consistency_3d():
    """Uniform profile: delta F_res/delta rho equals the EOS mu_res for every component."""
    say("3D bulk consistency, ternary, 16^3")
    m = PCSAFT(["methane", "ethane", "propane"])
    T, p = 200.0, 1e7
    x = np.array([0.6, 0.3, 0.1])
    v = m.volume(p, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform3DCart((p, T), x / v, [[0, 6 * L]] * 3, 16))
    rho = np.repeat(x / v, s.N)
    got, _, _ = s.device_system.functional_eval(rho)
    mu = m.VT_chemical_potential_res(v, T, x) / RGAS / T
    err = np.max(np.abs(got.reshape(3, s.N) - mu[:, None]))
    say(f"  max |dF/drho - mu_res| {err:.2e}")
    assert err < 1e-11, err


def test_hvp_3d_matches_finite_differences(ternary16):
    s, _, o = ternary16
    say("3D Hessian-vector product vs central differences, ternary, 16^3")
    rho = smooth_profile(s, seed=5).reshape(-1)
    v = smooth_profile(s, seed=6).reshape(-1)
    hv = s.device_system.hvp(rho, v)
    h = 1e-3 * np.max(rho) / np.max(np.abs(v))

    def c(hh):
        return (o.dFdrho_res(rho + hh * v)[0] - o.dFdrho_res(rho - hh * v)[0]) / (2 * hh)
    fd = (4 * c(h / 2) - c(h)) / 3
    e_ = rel(hv, fd)
    say(f"  rel err vs oracle FD {e_:.2e}")
    assert e_ < 1e-7, e_


def test_converge_3d_matches_oracle():
    say("3D converge: methane, 16^3, one hard sphere, tol 1e-11 (GPU and oracle)")
    s, rho0 = configs.matrix_3d(n=16, box=5.0, nspheres=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
    o = oracle_of(s)
    rho = rho0.copy()
    converge_(s, rho, tol=1e-11)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-11)
    say(f"  GPU {s.last_result['errcode']} evals {s.last_result['nevals']} | oracle {ro.errcode} evals {ro.nevals}")
    assert s.last_result["errcode"] == 0 and ro.errcode == 0
    e_ = rel(rho.reshape(-1), rho_o)
    say(f"  profile rel err {e_:.2e}")
    assert e_ < 1e-8, e_


def test_converge_3d_mixed_precision():
    say("3D converge, mixed precision vs FP64: methane, 16^3, one hard sphere, tol 1e-9")
    from cdft_amd import DFTOptions
    kw = dict(n=16, box=5.0, nspheres=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
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


def test_converge_3d_mixed_falls_back_to_fp64_on_fp32_overflow():
    """A point whose density the FP32 map cannot hold (exp(99.5) mol/m^3: N_A L^3 rho
    overflows FP32 downstream) makes the first FP32 map value non-finite; the solve
    then continues from the last point with a finite FP32 value (here the seed) with
    the FP64 map, and ends where the FP64 solve ends, bit for bit."""
    say("3D mixed precision: FP32 overflow at the seed falls back to the FP64 map")
    from cdft_amd import DFTOptions
    kw = dict(n=16, box=5.0, nspheres=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
    out = {}
    for prec in ("fp64", "mixed"):
        s, rho0 = configs.matrix_3d(options=DFTOptions(precision=prec), **kw)
        r = rho0.copy()
        r[0, 1234] = np.exp(99.5)
        converge_(s, r, tol=1e-9)
        out[prec] = (r, s.last_result)
    (r64, a), (r32, b) = out["fp64"], out["mixed"]
    say(f"  FP64 errcode {a['errcode']} evals {a['nevals']} | mixed errcode {b['errcode']} evals {b['nevals']} "
        f"(FP32 {b['nevals_fp32']})")
    assert a["errcode"] == 0 and b["errcode"] == 0
    assert b["nevals_fp32"] == 1 and b["nevals"] == a["nevals"] + 1
    assert np.array_equal(r32, r64)


def test_adjoint_3d_adsorption_sensitivity():
    say("3D adjoint: d(adsorption)/d(sigma, eps, m), methane around a hard sphere, 16^3, vs FD of full runs")
    from cdft_amd import adjoint as adj
    s, rho0 = configs.matrix_3d(n=16, box=5.0, nspheres=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
    kw = dict(tol=1e-11, maxit=5000)
    rho = rho0.copy()
    converge_(s, rho, **kw)
    assert s.last_result["errcode"] == 0, s.last_result
    rho = rho.reshape(-1)
    sens, lam, res = adj.sensitivities(s, rho, "adsorption", 0, tol=1e-11)
    say(f"  adjoint: errcode {res['errcode']} evals {res['nevals']}")
    assert res["errcode"] == 0, res
    for (prm, i), val in sens.items():
        base = adj.model_params(s.model)[s.model.components[i]][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
        h = 1e-5 * base
        vals = []
        for sg in (+1, -1):
            sp = adj.perturbed_system(s, i, prm, sg * h)
            r = rho0.copy()          # from the seed: aasol's divergence guard is relative to res0
            converge_(sp, r, **kw)
            rr = sp.last_result
            say(f"    FD run {prm} {sg:+d}h: errcode {rr['errcode']} evals {rr['nevals']} resnorm {rr['resnorm']:.2e}")
            # a run that stops at maxit on a residual floor below 1e-9 still fixes Gamma to ~1e-12
            assert rr["errcode"] == 0 or (rr["errcode"] == 10 and rr["resnorm"] < 1e-9), rr
            vals.append(adj.objective(sp, r, "adsorption", 0))
        fd = (vals[0] - vals[1]) / (2 * h)
        say(f"  dGamma/d {prm}: adjoint {val:.10e}  FD {fd:.10e}  rel {abs(val - fd) / abs(fd):.2e}")
        assert abs(val - fd) <= 1e-4 * abs(fd), (i, val, fd)


def test_adjoint_3d_ternary_adsorption_jacobian():
    """The matrix workload's sensitivity d(N_i)/d(eps_j) (all nine entries, three
    adjoint solves) for the ternary methane/ethane/propane at 200 K and 1e7 Pa, on a
    32^3 grid around two small hard spheres, against central differences of full
    converge runs (h = 1e-5 relative, tol 1e-11)."""
    say("3D adjoint: d(N_i)/d(eps_j), ternary around two hard spheres, 32^3, vs FD of full runs")
    from cdft_amd import adjoint as adj
    kw_sys = dict(n=32, box=8.0, nspheres=2, seed=3, rmin=0.6, rmax=1.0)
    s, rho0 = configs.matrix_3d(**kw_sys)
    kw = dict(tol=1e-11, maxit=5000)
    rho = rho0.copy()
    converge_(s, rho, **kw)
    assert s.last_result["errcode"] == 0, s.last_result
    rho = rho.reshape(-1)
    jac, lams, results = adj.jacobian(s, rho, "adsorption", [0, 1, 2], params=("epsilon",), tol=1e-11)
    for c, r in results.items():
        say(f"  adjoint for N_{c}: errcode {r['errcode']} evals {r['nevals']}")
        assert r["errcode"] == 0, (c, r)
    J = np.array([[jac[(c, "epsilon", j)] for j in range(3)] for c in range(3)])
    fd = np.zeros((3, 3))
    for j in range(3):
        h = 1e-5 * adj.model_params(s.model)[s.model.components[j]][2]
        ads = []
        for sg in (+1, -1):
            sp = adj.perturbed_system(s, j, "epsilon", sg * h)
            r = rho0.copy()
            converge_(sp, r, **kw)
            assert sp.last_result["errcode"] == 0, sp.last_result
            ads.append(adj.adsorption_system(sp, r))
        fd[:, j] = (ads[0] - ads[1]) / (2 * h)
    say(f"  adjoint d(N_i)/d(eps_j):\n{J}\n  finite differences:\n{fd}")
    np.testing.assert_allclose(J, fd, rtol=1e-4, atol=1e-4 * np.max(np.abs(fd)))
