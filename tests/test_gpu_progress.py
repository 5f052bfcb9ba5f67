"""converge!'s per-evaluation logger (log_interval, save_interval, save_callback;
src/methods/converge.jl:51-60, 256-267) on the device path: the solve pauses after
every k-th map evaluation and hands the host the profile that evaluation used.
Through the C ABI (cdft_solver_opts.progress).  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, configs, converge_, free_energy
from cdft_amd import adjoint as adj

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def residual_norm(system, rho, beta):
    """||beta (G(x) - x)||_2 at x = log rho: the residual the solver records for the
    evaluation at rho (anderson.jl:327-337, GFix! of converge.jl:234-241)."""
    x = np.log(rho.reshape(-1))
    return float(np.linalg.norm(beta * (adj.log_map(system, rho.reshape(-1)) - x)))


def test_save_and_log_callbacks_wall(capsys):
    """Methane wall (201 points, tol 1e-8): save_callback every 5 evaluations and the
    free-energy log every 10.  The solve is bitwise the one without callbacks; each
    saved profile is the one its evaluation used (its recorded residual norm
    reproduces from it); the first evaluation's profile is the seed."""
    say("progress: methane wall, save every 5, log every 10")
    s, rho0 = configs.methane_wall()
    ref = rho0.copy()
    converge_(s, ref, tol=1e-8, history=2000)
    r0 = s.last_result
    saved = []
    got = rho0.copy()
    converge_(s, got, tol=1e-8, history=2000, save_interval=5, log_interval=10,
              save_callback=lambda it, r: saved.append((it, r.copy())))
    r1 = s.last_result
    out = capsys.readouterr().out
    say(f"  {r1['nevals']} evaluations, {len(saved)} saves")
    assert (r1["errcode"], r1["nevals"], r1["iters"]) == (r0["errcode"], r0["nevals"], r0["iters"])
    assert np.array_equal(got, ref)
    assert np.array_equal(r1["history"], r0["history"])
    its = [it for it, _ in saved]
    assert its == list(range(5, r1["nevals"] + 1, 5))
    for it, r in saved:
        assert r.shape == rho0.shape
        want = r1["history"][it - 1]
        have = residual_norm(s, r, 1e-2)
        assert abs(have - want) <= 1e-6 * want + 1e-12, (it, have, want)
    logs = [ln for ln in out.splitlines() if ln.startswith("DFT iter")]
    assert [int(ln.split()[2]) for ln in logs] == list(range(10, r1["nevals"] + 1, 10))
    F10 = float(logs[0].split("F = ")[1])
    F_ref = free_energy(s, dict(saved)[10])
    assert abs(F10 - F_ref) <= 1e-5 * abs(F_ref), (F10, F_ref)


def test_first_evaluation_profile_is_the_seed():
    s, rho0 = configs.methane_wall()
    seen = []
    rho = rho0.copy()
    converge_(s, rho, tol=1e-8, maxit=3, save_interval=1, save_callback=lambda it, r: seen.append((it, r)))
    assert seen[0][0] == 1
    # exp(log rho): a relative error of a few ulps times |log rho| (rho ~ 1e-25 at the wall)
    bound = 4e-16 * (1.0 + np.abs(np.log(rho0))) * rho0
    assert np.all(np.abs(seen[0][1] - rho0) <= bound)
    assert [it for it, _ in seen] == list(range(1, s.last_result["nevals"] + 1))


def test_callback_exception_stops_the_solve():
    """An exception in save_callback stops the solve on the device and propagates;
    rho then holds the profile of the evaluation it saw, and the next solve of the
    same system runs normally."""
    s, rho0 = configs.methane_wall()

    class Stop(Exception):
        pass

    seen = {}

    def cb(it, r):
        seen[it] = r.copy()
        if it == 12:
            raise Stop()

    rho = rho0.copy()
    with pytest.raises(Stop):
        converge_(s, rho, tol=1e-8, save_interval=4, save_callback=cb)
    assert sorted(seen) == [4, 8, 12]
    assert np.array_equal(rho, seen[12])
    ref = rho0.copy()
    converge_(s, ref, tol=1e-8)
    assert s.last_result["errcode"] == 0


def test_progress_every_validated():
    s, rho0 = configs.methane_wall()
    with pytest.raises(ValueError):
        s.device_system.converge(rho0.copy().reshape(-1), progress=lambda it, r: None, progress_every=0)
    with pytest.raises(ValueError):
        converge_(s, rho0.copy(), save_interval=3)


def test_mixed_precision_unchanged_by_pauses():
    """Mixed precision (FP32 map, then FP64 after the switch): pauses every 3
    evaluations, across the switch and the graph chunks, leave the solve bitwise
    unchanged; the saved profiles count every evaluation, FP32 ones included."""
    say("progress: methane wall, mixed precision, save every 3")
    s, rho0 = configs.methane_wall(options=DFTOptions(precision="mixed"))
    ref = rho0.copy()
    converge_(s, ref, tol=1e-8)
    r0 = s.last_result
    its = []
    got = rho0.copy()
    converge_(s, got, tol=1e-8, save_interval=3, save_callback=lambda it, r: its.append(it))
    r1 = s.last_result
    say(f"  {r1['nevals']} evaluations ({r1['nevals_fp32']} FP32)")
    assert (r1["nevals"], r1["nevals_fp32"], r1["errcode"]) == (r0["nevals"], r0["nevals_fp32"], r0["errcode"])
    assert r1["nevals_fp32"] > 0
    assert np.array_equal(got, ref)
    assert its == list(range(3, r1["nevals"] + 1, 3))


CASES_PAUSE = {
    # (system factory, precision)
    "interface_512x2_mixed": (lambda o: configs.c1c10_interface(ngrid=512, box=10.0, options=o), "mixed"),
    "interface_512x2_fp64": (lambda o: configs.c1c10_interface(ngrid=512, box=10.0, options=o), "fp64"),
    "wall_mixed": (lambda o: configs.methane_wall(options=o), "mixed"),
    "wall_fp64": (lambda o: configs.methane_wall(options=o), "fp64"),
}


@pytest.mark.parametrize("every", [1, 2, 5])
@pytest.mark.parametrize("name", list(CASES_PAUSE))
def test_solve_unchanged_by_pauses(name, every):
    """A pause after every (or every 2nd, 5th) evaluation leaves the solve bitwise
    unchanged, on the multi-kernel bodies (methane/n-decane interface, 512 points x 2;
    in mixed precision FP32 bodies until the switch, then FP64 bodies, so the switch
    also falls on a pause) and on the single-workgroup bodies (methane wall)."""
    make, prec = CASES_PAUSE[name]
    s, rho0 = make(DFTOptions(precision="mixed") if prec == "mixed" else None)
    say(f"progress: {name}, save every {every}")
    ref = rho0.copy()
    converge_(s, ref, tol=1e-8)
    r0 = s.last_result
    its = []
    got = rho0.copy()
    converge_(s, got, tol=1e-8, save_interval=every, save_callback=lambda it, r: its.append(it))
    r1 = s.last_result
    say(f"  {r1['nevals']} evaluations ({r1['nevals_fp32']} FP32), errcode {r1['errcode']}")
    assert (r1["nevals"], r1["nevals_fp32"], r1["errcode"]) == (r0["nevals"], r0["nevals_fp32"], r0["errcode"])
    if prec == "mixed":
        assert 0 < r1["nevals_fp32"] < r1["nevals"]
    assert np.array_equal(got, ref)
    assert its == list(range(every, r1["nevals"] + 1, every))


def test_progress_on_3d_grid():
    """The spectral path (3D, 32^3, three components around two hard spheres): saved
    every 7 evaluations, same solve as without."""
    say("progress: 3D ternary matrix 32^3, save every 7")
    s, rho0 = configs.matrix_3d(n=32, box=8.0, nspheres=2, seed=3, rmin=0.6, rmax=1.0)
    ref = rho0.copy()
    converge_(s, ref, tol=1e-8, maxit=5000)
    r0 = s.last_result
    saved = []
    got = rho0.copy()
    converge_(s, got, tol=1e-8, maxit=5000, save_interval=7, save_callback=lambda it, r: saved.append(it))
    r1 = s.last_result
    assert (r1["nevals"], r1["errcode"]) == (r0["nevals"], r0["errcode"])
    assert np.array_equal(got, ref)
    assert saved == list(range(7, r1["nevals"] + 1, 7))

