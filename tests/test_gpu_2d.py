"""2D Cartesian grids on the spectral path (a 2D grid runs as a 3D grid with one z
plane and an identically zero third vector weight): the functional, the bulk
invariant, converged profiles and the DDFT right-hand side against the oracle's
literal 2D restatement (F = 8 slots, 2D FFTs), and the lamellar interface against
the 1D interface it replicates.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, DFTSystem, PCSAFT, Uniform2DCart, configs, converge_
from cdft_amd.bulk import RGAS
from cdft_amd.methods import surface_tension_system
from helpers import oracle_of, smooth_profile

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("ngrid", [(32, 32), (16, 64), (128, 32)])
def test_functional_eval_2d_smooth(ngrid):
    m = PCSAFT(["methane", "n-decane"])
    T, x = 300.0, np.array([0.5, 0.5])
    v = m.volume(1.5e7, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform2DCart((1.5e7, T), x / v, [[0, 6 * L], [-3 * L, 9 * L]], ngrid))
    o = oracle_of(s)
    for seed in range(2):
        rho = smooth_profile(s, seed).reshape(-1)
        got, fres, fval = s.device_system.functional_eval(rho, want_fval=True)
        ref, fv = o.dFdrho_res(rho)
        e_ = rel(got, ref)
        say(f"2D {ngrid} seed {seed}: dF/drho rel err {e_:.2e}, F_res {fres:.12e} vs {o.F_res(rho):.12e}")
        assert e_ < 1e-12, e_
        assert rel(fval, fv) < 1e-12
        assert abs(fres - o.F_res(rho)) <= 1e-12 * abs(fres)


def test_functional_eval_2d_disk_matrix_seed():
    s, rho0 = configs.matrix_2d(n=64, box=10.0, ndisks=3, seed=2)
    o = oracle_of(s)
    rho = rho0.reshape(-1)
    got, _, _ = s.device_system.functional_eval(rho)
    ref, _ = o.dFdrho_res(rho)
    out = np.tile(s.vext.reshape(s.nc, s.N)[0] == 0, s.nc)        # outside the disks
    e_ = np.max(np.abs(got - ref)[out]) / np.max(np.abs(ref[out]))
    say(f"2D ternary disk-matrix seed 64^2: dF/drho rel err outside the disks {e_:.2e}")
    assert e_ < 1e-10, e_


def test_bulk_consistency_2d():
    m = PCSAFT(["methane", "ethane", "propane"])
    T, p = 200.0, 1e7
    x = np.array([0.6, 0.3, 0.1])
    v = m.volume(p, T, x, "liquid")
    L = m.length_scale()
    s = DFTSystem(m, Uniform2DCart((p, T), x / v, [[0, 6 * L], [0, 6 * L]], 32))
    got, _, _ = s.device_system.functional_eval(np.repeat(x / v, s.N))
    mu = m.VT_chemical_potential_res(v, T, x) / RGAS / T
    err = np.max(np.abs(got.reshape(3, s.N) - mu[:, None]))
    say(f"2D bulk consistency: max |dF/drho - mu_res| {err:.2e}")
    assert err < 1e-11, err


def test_converge_2d_matches_oracle():
    s, rho0 = configs.matrix_2d(n=32, box=6.0, ndisks=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
    o = oracle_of(s)
    rho = rho0.copy()
    converge_(s, rho, tol=1e-11)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-11)
    say(f"2D methane around a disk, 32^2: GPU {s.last_result['errcode']} evals {s.last_result['nevals']} | "
        f"oracle {ro.errcode} evals {ro.nevals}")
    assert s.last_result["errcode"] == 0 and ro.errcode == 0
    e_ = rel(rho.reshape(-1), rho_o)
    say(f"  profile rel err {e_:.2e}")
    assert e_ < 1e-8, e_


def test_converge_2d_mixed_precision():
    kw = dict(n=64, box=10.0, ndisks=2, seed=3)
    s64, rho0 = configs.matrix_2d(**kw)
    s32, _ = configs.matrix_2d(options=DFTOptions(precision="mixed"), **kw)
    a, b = rho0.copy(), rho0.copy()
    converge_(s64, a, tol=1e-9, anderson_start=0.5)
    converge_(s32, b, tol=1e-9, anderson_start=0.5)
    r64, r32 = s64.last_result, s32.last_result
    e_ = rel(b.reshape(-1), a.reshape(-1))
    say(f"2D ternary disk matrix 64^2: FP64 errcode {r64['errcode']} evals {r64['nevals']} | mixed errcode "
        f"{r32['errcode']} evals {r32['nevals']} (FP32 {r32['nevals_fp32']}) | profile rel diff {e_:.2e}")
    assert r64["errcode"] == 0 and r32["errcode"] == 0 and r32["nevals_fp32"] > 0
    assert e_ < 1e-6, e_


def test_lamellar_2d_equals_1d_interface():
    """A lamellar 2D interface uniform along y is the 1D interface replicated: the
    converged 2D profile matches the 1D one on every row (up to the small translation
    along x that a free interface's solve leaves, its zero mode), and the 2D tension
    integral is the 1D one times the y integral of 1 (Simpson's weights on the
    periodic y grid: 44/3 dy for 16 points, not the extent 16 dy, as _int has it)."""
    s1, r1 = configs.c1c10_interface(ngrid=256, box=10.0)
    s2, r2 = configs.c1c10_lamellar(ngrid=(256, 16), box=10.0, ywidth=2.0)
    a, b = r1.copy(), r2.copy()
    converge_(s1, a, tol=1e-10, maxit=20000)
    converge_(s2, b, tol=1e-10, maxit=20000)
    say(f"1D: errcode {s1.last_result['errcode']} evals {s1.last_result['nevals']} | 2D: errcode "
        f"{s2.last_result['errcode']} evals {s2.last_result['nevals']}")
    rows = b.reshape(2, 16, 256)
    e_rows = max(rel(rows[:, j, :], a) for j in range(16))
    g1 = surface_tension_system(s1, a)
    g2 = surface_tension_system(s2, b)
    from cdft_amd import simpson
    ly = simpson(np.ones(16), s2.structure.dzs()[1])
    e_g = abs(g2 / ly - g1) / abs(g1)
    say(f"  rows vs 1D profile rel {e_rows:.2e}; tension per y integral vs 1D rel {e_g:.2e}")
    assert e_rows < 1e-6, e_rows
    assert e_g < 1e-9, e_g


def test_ddft_rhs_2d_matches_oracle():
    s, rho0 = configs.matrix_2d(n=32, box=8.0, ndisks=0)
    o = oracle_of(s)
    eta = np.log(smooth_profile(s, 4)).reshape(-1)
    got = s.device_system.ddft_rhs(eta)
    ref = o.ddft_rhs(eta)
    e_ = rel(got, ref)
    say(f"2D DDFT right-hand side rel err {e_:.2e}")
    assert e_ < 1e-10, e_


def test_hvp_2d_matches_finite_differences():
    s, _ = configs.matrix_2d(n=32, box=8.0, ndisks=0)
    o = oracle_of(s)
    rho = smooth_profile(s, seed=5).reshape(-1)
    v = smooth_profile(s, seed=6).reshape(-1)
    hv = s.device_system.hvp(rho, v)
    h = 1e-3 * np.max(rho) / np.max(np.abs(v))

    def c(hh):
        return (o.dFdrho_res(rho + hh * v)[0] - o.dFdrho_res(rho - hh * v)[0]) / (2 * hh)
    fd = (4 * c(h / 2) - c(h)) / 3
    e_ = rel(hv, fd)
    say(f"2D Hessian-vector product vs oracle FD rel err {e_:.2e}")
    assert e_ < 1e-7, e_
