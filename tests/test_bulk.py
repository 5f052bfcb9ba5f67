"""Bulk PC-SAFT (host replacement of the Clapeyron calls): thermodynamic consistency
and literature anchors.  CPU only."""
import numpy as np
import pytest

from cdft_amd.bulk import NA, PCSAFT, RGAS


def test_pressure_matches_finite_difference_of_A():
    m = PCSAFT(["methane", "n-decane"])
    T, z = 300.0, np.array([0.4, 0.6])
    V = 2e-4
    h = 1e-7 * V
    A = lambda vv: z.sum() * m.a_res(vv, T, z) * RGAS * T
    p_fd = RGAS * T * z.sum() / V - (A(V + h) - A(V - h)) / (2 * h)
    assert abs(m.pressure(V, T, z) - p_fd) <= 1e-7 * abs(p_fd)


def test_euler_relation():
    """A_res/(nRT) = sum x_i mu_res_i/RT - (Z - 1) at (V, T)."""
    m = PCSAFT(["methane", "ethane", "propane"])
    T, z = 200.0, np.array([0.6, 0.3, 0.1])
    V = 7e-5
    Z = m.pressure(V, T, z) * V / (RGAS * T * z.sum())
    lhs = m.a_res(V, T, z)
    rhs = np.sum(z / z.sum() * m.VT_chemical_potential_res(V, T, z)) / (RGAS * T) - (Z - 1)
    assert abs(lhs - rhs) <= 1e-12 * max(1, abs(lhs))


def test_gibbs_duhem():
    m = PCSAFT(["methane", "n-decane"])
    T, z = 300.0, np.array([0.5, 0.5])
    V, h = 1.5e-4, 1e-8
    dmu = (m.VT_chemical_potential(V + h, T, z) - m.VT_chemical_potential(V - h, T, z)) / (2 * h)
    dp = (m.pressure(V + h, T, z) - m.pressure(V - h, T, z)) / (2 * h)
    assert abs(np.dot(z, dmu) - V * dp) <= 1e-6 * abs(V * dp)


def test_methane_saturation_and_critical_point():
    m = PCSAFT(["methane"])
    p, vl, vv = m.saturation_pressure(150.0)
    assert abs(p - 1.04e6) < 0.03e6                      # NIST: 1.0399 MPa
    z = np.array([1.0])
    assert abs(m.lnphi(vl, 150.0, z)[0] - m.lnphi(vv, 150.0, z)[0]) < 1e-10
    Tc, pc, vc = m.crit_pure()
    assert 189.0 < Tc < 194.0                            # PC-SAFT slightly above 190.56 K


def test_bubble_point_equilibrium():
    m = PCSAFT(["methane", "n-decane"])
    x = np.array([0.5, 0.5])
    p, vl, vv, y = m.bubble_pressure(300.0, x)
    assert abs(y.sum() - 1) < 1e-12
    lhs = m.lnphi(vl, 300.0, x) + np.log(x)
    rhs = m.lnphi(vv, 300.0, y) + np.log(y)
    np.testing.assert_allclose(lhs, rhs, atol=1e-9)
    assert abs(m.pressure(vl, 300.0, x) - p) < 1e-6 * p and abs(m.pressure(vv, 300.0, y) - p) < 1e-6 * p
    assert vl < vv


def test_liquid_root_c0_state():
    m = PCSAFT(["methane"])
    v = m.volume(1e7, 150.0, [1.0], "liquid")
    assert abs(m.pressure(v, 150.0, [1.0]) - 1e7) < 1e-6
    assert 2.2e4 < 1 / v < 2.5e4                         # compressed liquid methane, mol/m^3


def test_hard_sphere_diameter():
    m = PCSAFT(["methane"])
    d = m.d(150.0)[0]
    assert abs(d - 3.7039e-10 * (1 - 0.12 * np.exp(-3 * 150.03 / 150.0))) < 1e-25
