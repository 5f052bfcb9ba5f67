"""Mixed precision on the plain 1D path: the FP32 map (compensated FP32 circulant
sums, FP32 Phi gradient with its n3-singular factors in FP64) against the FP64
oracle, mixed solves against FP64 solves, and the differentiation workload's
mixed leg: surface-tension sensitivities of the 2048-point methane / n-decane
interface from a mixed-precision solve, against central differences (h = 1e-6
relative) of FP64 full solves.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, configs, converge_
from cdft_amd import adjoint as adj
from cdft_amd.methods import surface_tension_system
from helpers import oracle_of

pytestmark = pytest.mark.gpu
MIXED = DFTOptions(precision="mixed")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("which", ["wall", "interface"])
def test_fp32_map_matches_oracle_1d(which):
    if which == "wall":
        s, rho0 = configs.methane_wall(options=MIXED)
    else:
        s, rho0 = configs.c1c10_interface(options=MIXED)
    o = oracle_of(s)
    s64 = configs.methane_wall()[0] if which == "wall" else configs.c1c10_interface()[0]
    rho = rho0.copy()
    converge_(s64, rho, tol=1e-8)                 # a converged profile: the map's operating point
    for name, r in (("seed", rho0), ("converged", rho)):
        got = s.device_system.functional_eval_fp32(r.reshape(-1))
        ref, _ = o.dFdrho_res(r.reshape(-1))
        e_ = rel(got, ref)
        say(f"{which} {name}: FP32 map dF/drho rel err vs FP64 oracle {e_:.2e}")
        assert e_ < 2e-6, e_


def test_fp32_map_matches_oracle_3d():
    s, rho0 = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1, options=MIXED)
    o = oracle_of(s)
    from helpers import smooth_profile
    rho = smooth_profile(s, seed=3).reshape(-1)
    got = s.device_system.functional_eval_fp32(rho)
    ref, _ = o.dFdrho_res(rho)
    e_ = rel(got, ref)
    say(f"3D ternary 16^3 smooth profile: FP32 map rel err {e_:.2e}")
    assert e_ < 2e-6, e_


def test_fp32_map_refused_for_fp64_system():
    from cdft_amd._lib import CDFTError
    s, rho0 = configs.methane_wall()
    with pytest.raises(CDFTError):
        s.device_system.functional_eval_fp32(rho0.reshape(-1))


def test_mixed_wall_matches_fp64():
    s64, rho0 = configs.methane_wall()
    s32, _ = configs.methane_wall(options=MIXED)
    a, b = rho0.copy(), rho0.copy()
    converge_(s64, a, tol=1e-10)
    converge_(s32, b, tol=1e-10)
    r64, r32 = s64.last_result, s32.last_result
    e_ = rel(b, a)
    say(f"methane wall tol 1e-10: FP64 evals {r64['nevals']} | mixed evals {r32['nevals']} "
        f"(FP32 {r32['nevals_fp32']}) | profile rel diff {e_:.2e}")
    assert r64["errcode"] == 0 and r32["errcode"] == 0 and r32["nevals_fp32"] > 0
    assert r32["resnorm"] <= r32["tol"]
    assert e_ < 1e-8, e_


def test_mixed_interface_matches_fp64():
    """The interface workload at its own tolerance (1e-8): both solves converge and the
    mixed solve's surface tension is within the stated 1e-6 of the FP64 one."""
    s64, rho0 = configs.c1c10_interface()
    s32, _ = configs.c1c10_interface(options=MIXED)
    a, b = rho0.copy(), rho0.copy()
    converge_(s64, a, tol=1e-8)
    converge_(s32, b, tol=1e-8)
    r64, r32 = s64.last_result, s32.last_result
    g64, g32 = surface_tension_system(s64, a), surface_tension_system(s32, b)
    e_ = abs(g32 - g64) / abs(g64)
    say(f"methane/n-decane interface tol 1e-8: FP64 errcode {r64['errcode']} evals {r64['nevals']} | mixed "
        f"errcode {r32['errcode']} evals {r32['nevals']} (FP32 {r32['nevals_fp32']}) | tension rel diff {e_:.2e}")
    assert r64["errcode"] == 0 and r32["errcode"] == 0 and r32["nevals_fp32"] > 0
    assert e_ < 1e-6, e_


def _converged(s, rho0, tol):
    r = rho0.copy().reshape(s.nc, s.N)
    converge_(s, r, tol=tol, maxit=20000)
    res = s.last_result
    # a free interface's translation mode barely contracts: a run that stops at maxit
    # on a residual floor below 1e-9 is accepted
    assert res["errcode"] == 0 or (res["errcode"] == 10 and res["resnorm"] < 1e-9), res
    return r.reshape(-1), res


def test_c2_gamma_sensitivities_fp64_and_mixed():
    """Differentiation workload at full size (2048 points, 40 sigma box): d gamma /
    d(eps of methane, sigma of n-decane) through the adjoint, from an FP64 and from a
    mixed-precision converged profile, against central differences of FP64 full
    solves with h = 1e-6 relative (the workload's step) and 1e-5.  The full solves
    stop on the interface's residual floor (about 1e-9, its translation mode barely
    contracts), which puts a noise of about 1e-4 relative into the h = 1e-6
    differences and about 1e-5 into the h = 1e-5 ones."""
    s64, rho0 = configs.c1c10_interface()
    s32, _ = configs.c1c10_interface(options=MIXED)
    keys = [("epsilon", 0), ("sigma", 1)]
    bound = {1e-6: 3e-4, 1e-5: 3e-5}
    fd = {}
    rho64, _ = _converged(s64, rho0.reshape(-1), 1e-12)
    for prm, i in keys:
        base = adj.model_params(s64.model)[s64.model.components[i]][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
        for rel_h in bound:
            h = rel_h * abs(base)
            vals = []
            for sg in (+1, -1):
                sp = adj.perturbed_system(s64, i, prm, sg * h)
                r, _ = _converged(sp, rho64, 1e-12)
                vals.append(adj.objective(sp, r, "gamma"))
                del sp
            fd[(prm, i, rel_h)] = (vals[0] - vals[1]) / (2 * h)
    rho32, res32 = _converged(s32, rho0.reshape(-1), 1e-12)
    say(f"mixed forward solve: errcode {res32['errcode']} evals {res32['nevals']} (FP32 {res32['nevals_fp32']})")
    assert res32["nevals_fp32"] > 0
    for label, s, rho in (("FP64", s64, rho64), ("mixed", s32, rho32)):
        sens, _, res = adj.sensitivities(s, rho, "gamma", params=("sigma", "epsilon"), tol=1e-10)
        assert res["errcode"] == 0, res
        for key in keys:
            for rel_h, b in bound.items():
                f = fd[key + (rel_h,)]
                e_ = abs(sens[key] - f) / abs(f)
                say(f"  {label}: d gamma / d {key}: adjoint {sens[key]:.10e}  FD(h={rel_h:g}) {f:.10e}  rel {e_:.2e}")
                assert e_ < b, (label, key, rel_h, e_)
