"""Radial (spherical / cylindrical) grids on the device: dense QDHT-convolution
operators built at set-up and applied by k_dense1d, against the oracle's literal
restatement (forward transform, multiply, backward transform per field; the vector
weight as D Q^-1 H Q), the reference's radial bulk test, converged pore profiles
and post-processing.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, DFTSystem, PCSAFT, configs, converge_
from cdft_amd._lib import CDFTError
from cdft_amd.methods import adsorption_system, surface_tension_system
from cdft_amd.structure import Uniform1DCyl, Uniform1DSphr
from helpers import bulk_mu, oracle_of

pytestmark = pytest.mark.gpu
COORDS = ["spherical", "cylindrical"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def binary(coord, n=128):
    m = PCSAFT(["methane", "n-decane"])
    T, x = 300.0, np.array([0.5, 0.5])
    v = m.volume(1.5e7, T, x, "liquid")
    L = m.length_scale()
    mk = Uniform1DSphr if coord == "spherical" else Uniform1DCyl
    return DFTSystem(m, mk((1.5e7, T), x / v, [0.0, 8 * L], n))


@pytest.mark.parametrize("coord", COORDS)
def test_functional_eval_radial_matches_oracle(coord):
    s, rho0 = configs.methane_radial(coord)
    o = oracle_of(s)
    r = s.structure.axis_coords(0)
    for name, rho in (("pore seed", rho0.reshape(-1)),
                      ("smooth", s.species.bulk_density[0] * (1 + 0.3 * np.cos(r / s.L)))):
        got, fres, fval = s.device_system.functional_eval(rho, want_fval=True)
        ref, fv = o.dFdrho_res(rho)
        e_ = rel(got, ref)
        say(f"{coord} methane, {name}: dF/drho rel err {e_:.2e}; F_res {fres:.12e} vs {o.F_res(rho):.12e}")
        assert e_ < 1e-12, e_
        assert rel(fval, fv) < 1e-12
        assert abs(fres - o.F_res(rho)) <= 1e-12 * abs(fres)
    sb = binary(coord)
    ob = oracle_of(sb)
    rb = np.concatenate([sb.species.bulk_density[c] * (1 + 0.2 * np.sin((c + 1) * sb.structure.axis_coords(0) / sb.L))
                         for c in range(2)])
    e_ = rel(sb.device_system.functional_eval(rb)[0], ob.dFdrho_res(rb)[0])
    say(f"{coord} methane/n-decane smooth: dF/drho rel err {e_:.2e}")
    assert e_ < 1e-12, e_


@pytest.mark.parametrize("coord,ub,n", [("spherical", 10, 151), ("cylindrical", 20, 101)])
def test_bulk_mu_res_at_first_radius_on_device(coord, ub, n):
    """test_coordinate_systems.jl:60-80 on the device."""
    s, _ = configs.methane_radial(coord, ngrid=n, radius=ub, wall=False)
    rb = s.species.bulk_density[0]
    got, _, _ = s.device_system.functional_eval(np.full(s.N, rb))
    err = abs(got[0] - s.species.chempot_res[0])
    say(f"{coord}: |dF/drho(r_1) - mu_res| {err:.2e}")
    assert err < 1e-7, err


@pytest.mark.parametrize("coord", COORDS)
def test_converge_radial_pore_matches_oracle(coord):
    s, rho0 = configs.methane_radial(coord)
    o = oracle_of(s)
    rho = rho0.copy()
    converge_(s, rho, tol=1e-11)
    rho_o, ro, _ = o.converge(rho0.reshape(-1), tol=1e-11)
    r = s.last_result
    e_ = rel(rho.reshape(-1), rho_o)
    say(f"{coord} pore: GPU errcode {r['errcode']} evals {r['nevals']} | oracle {ro.errcode} evals {ro.nevals} | "
        f"profile rel err {e_:.2e}")
    assert r["errcode"] == 0 and ro.errcode == 0
    assert e_ < 1e-9, e_
    mu = bulk_mu(s)
    p = s.structure.conditions[0]
    ads = s.device_system.adsorption(rho.reshape(-1))
    assert np.allclose(ads, o.adsorption(rho.reshape(-1)), rtol=1e-12)
    assert np.allclose(ads, adsorption_system(s, rho), rtol=1e-12)
    g = s.device_system.surface_tension(rho.reshape(-1), mu, p)
    assert g == pytest.approx(o.surface_tension(rho.reshape(-1), mu, p), rel=1e-10)
    assert g == pytest.approx(surface_tension_system(s, rho), rel=1e-10)


def test_hvp_radial_matches_finite_differences():
    s, rho0 = configs.methane_radial("spherical", wall=False)
    o = oracle_of(s)
    r = s.structure.axis_coords(0)
    rho = s.species.bulk_density[0] * (1 + 0.3 * np.cos(r / s.L))
    v = s.species.bulk_density[0] * np.sin(2 * r / s.L)
    hv = s.device_system.hvp(rho, v)
    h = 1e-3 * np.max(rho) / np.max(np.abs(v))

    def c(hh):
        return (o.dFdrho_res(rho + hh * v)[0] - o.dFdrho_res(rho - hh * v)[0]) / (2 * hh)
    fd = (4 * c(h / 2) - c(h)) / 3
    e_ = rel(hv, fd)
    say(f"spherical HVP vs oracle FD rel err {e_:.2e}")
    assert e_ < 1e-7, e_


def test_radial_refusals():
    with pytest.raises(CDFTError):
        configs.methane_radial("spherical", options=DFTOptions(precision="mixed"))[0].device_system
    s, rho0 = configs.methane_radial("cylindrical")
    with pytest.raises(CDFTError):
        s.device_system.ddft_rhs(np.log(rho0.reshape(-1)))
    from cdft_amd._lib import converge_batch
    with pytest.raises(CDFTError):
        converge_batch(s.device_system.ctx, [s.device_system], [rho0.reshape(-1).copy()], tol=1e-8)
