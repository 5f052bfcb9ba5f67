"""Host logic of the 3D structures, the hard-sphere matrix field and the adjoint
assembly that needs no device: integration weights, sphere placement and masks, the
translation mode, perturbed systems.  CPU only."""
import numpy as np
import pytest

from cdft_amd import (DFTSystem, HardSphereMatrix, PCSAFT, Uniform1DCart, Uniform3DCart, configs, integrate,
                      integration_weights, simpson)
from cdft_amd import adjoint as adj


def test_uniform3d_validation():
    with pytest.raises(ValueError):
        Uniform3DCart((1e5, 200.0), [1.0], [[0, 1], [0, 1]], 16)          # 2 axes given
    with pytest.raises(ValueError):
        Uniform3DCart((1e5, 200.0), [1.0], [[0, 1], [0, 1], [1, 0]], 16)  # ub < lb
    st = Uniform3DCart((1e5, 200.0), [1.0], [[0, 1], [0, 2], [0, 3]], (16, 32, 8))
    assert st.ndim == 3 and st.npoints == 16 * 32 * 8 and st.ngrid == (16, 32, 8)
    assert np.allclose(st.dzs(), (1 / 16, 2 / 32, 3 / 8))
    assert st.volume() == pytest.approx(6.0)


def test_integration_weights_match_nested_simpson():
    st = Uniform3DCart((1e5, 200.0), [1.0], [[0, 1.5], [-1, 2], [0, 0.7]], (16, 8, 32))
    rng = np.random.default_rng(0)
    f = rng.random(st.npoints)
    w = integration_weights(st)
    assert w @ f == pytest.approx(integrate(st, f), rel=1e-13)
    # separable f: the 3D rule is the product of the 1D rules (x fastest)
    fx, fy, fz = rng.random(16), rng.random(8), rng.random(32)
    F = (fz[:, None, None] * fy[None, :, None] * fx[None, None, :]).reshape(-1)
    ref = simpson(fx, 1.5 / 16) * simpson(fy, 3 / 8) * simpson(fz, 0.7 / 32)
    assert integrate(st, F) == pytest.approx(ref, rel=1e-13)
    st1 = Uniform1DCart((1e5, 200.0), [1.0], [0, 2.0], 41)
    assert integration_weights(st1) @ np.ones(41) == pytest.approx(simpson(np.ones(41), 2.0 / 41))
    assert np.allclose(adj.simpson_weights(41, 2.0 / 41), integration_weights(st1))


def test_hard_sphere_matrix_placement_and_mask():
    box = 10.0
    m = HardSphereMatrix.random(12, 0.8, 1.6, box, seed=3)
    c, r = m.centres, m.radii
    assert np.all((r >= 0.8) & (r <= 1.6))
    for i in range(len(r)):
        for j in range(i):
            d = c[i] - c[j]
            d -= box * np.round(d / box)
            assert np.sqrt(d @ d) >= r[i] + r[j]          # no overlap under the minimum image
    # a fine grid recovers the excluded volume fraction
    st = Uniform3DCart((1e5, 200.0), [1.0], [[0, box]] * 3, 96)
    mask = m.mask(st)
    frac = (4.0 / 3.0) * np.pi * np.sum(r ** 3) / box ** 3
    assert mask.mean() == pytest.approx(frac, rel=0.05)
    v = m.vext(st, PCSAFT(["methane", "ethane"]))
    assert v.size == 2 * st.npoints
    assert np.array_equal(v[:st.npoints], v[st.npoints:])
    assert set(np.unique(v)) <= {0.0, m.cap}


def test_hard_sphere_mask_wraps_periodically():
    st = Uniform3DCart((1e5, 200.0), [1.0], [[0, 4.0]] * 3, 32)
    m = HardSphereMatrix([[0.1, 0.1, 0.1]], [0.6])
    mask = m.mask(st).reshape(32, 32, 32)                    # [z, y, x]
    assert mask[0, 0, 0] and mask[-1, -1, -1]                 # both corners of the periodic cube


def test_matrix_workload_setup():
    s, rho = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1)
    assert s.N == 16 ** 3 and rho.shape == (3, s.N)
    d = s.descriptor()
    assert d["nd"] == 3 and d["ngrid"] == [16, 16, 16]
    inside = s.vext[:s.N] > 0
    assert inside.any() and not inside.all()
    # seed: bulk times exp(-beta V / 10), clamped (structure.jl:50-94)
    assert np.allclose(rho[:, ~inside], s.species.bulk_density[:, None])
    assert np.all(rho[:, inside] < 1e-15 * s.species.bulk_density[:, None])


def test_translation_mode_is_the_spectral_derivative():
    m = PCSAFT(["methane"])
    st = Uniform1DCart((1e6, 150.0), [1e4], [0, 1.0], 64)
    s = DFTSystem(m, st)
    k = np.arange(64) / 64
    rho = np.exp(0.3 * np.sin(2 * np.pi * 3 * k))
    t = adj.translation_mode(s, rho)      # derivative with respect to the phase 2 pi z / (ub - lb)
    assert np.allclose(t, 0.3 * 3 * np.cos(2 * np.pi * 3 * k), atol=1e-12)


def test_perturbed_system_keeps_grid_and_moves_bulk():
    s, _ = configs.c1c10_interface(ngrid=64, box=6.0)
    sp = adj.perturbed_system(s, 1, "sigma", 0.01)
    assert sp.structure.bounds == s.structure.bounds and sp.structure.ngrid == s.structure.ngrid
    assert sp.model.sigma[1, 1] == pytest.approx(s.model.sigma[1, 1] + 1e-12)
    assert sp.structure.topology == "two_phase"
    assert not np.allclose(sp.structure.rho_bulk, s.structure.rho_bulk)
    s3, _ = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1)
    s3p = adj.perturbed_system(s3, 0, "epsilon", 1.0)
    assert s3p.structure.ndim == 3 and s3p.structure.bounds == s3.structure.bounds
    assert s3p.vext is not None and np.array_equal(s3p.vext > 0, s3.vext > 0)


def test_adsorption_gradient_is_weights_over_volume():
    s, rho = configs.matrix_3d(n=16, box=6.0, nspheres=2, seed=1)
    g = adj.objective_grad_rho(s, rho, "adsorption", 1)
    w = integration_weights(s.structure) / s.structure.volume()
    assert np.allclose(g[s.N:2 * s.N], w) and not g[:s.N].any() and not g[2 * s.N:].any()
    # the objective is linear in rho: its value equals g . rho
    from cdft_amd.methods import adsorption_system
    assert g @ rho.reshape(-1) == pytest.approx(adsorption_system(s, rho)[1], rel=1e-13)


def test_uniform2d_structure_and_weights():
    from cdft_amd import Uniform2DCart
    with pytest.raises(ValueError):
        Uniform2DCart((1e5, 200.0), [1.0], [[0, 1], [0, 1], [0, 1]], 16)      # 3 axes given
    st = Uniform2DCart((1e5, 200.0), [1.0], [[0, 1.5], [-1, 2]], (16, 8))
    assert st.ndim == 2 and st.npoints == 128 and st.volume() == pytest.approx(4.5)
    rng = np.random.default_rng(1)
    fx, fy = rng.random(16), rng.random(8)
    F = (fy[:, None] * fx[None, :]).reshape(-1)
    assert integrate(st, F) == pytest.approx(simpson(fx, 1.5 / 16) * simpson(fy, 3 / 8), rel=1e-13)


def test_hard_disk_matrix_placement_and_mask():
    box = 10.0
    m = HardSphereMatrix.random(6, 0.8, 1.6, box, seed=4, ndim=2)
    c, r = m.centres, m.radii
    assert c.shape == (6, 2)
    for i in range(len(r)):
        for j in range(i):
            d = c[i] - c[j]
            d -= box * np.round(d / box)
            assert np.sqrt(d @ d) >= r[i] + r[j]
    from cdft_amd import Uniform2DCart
    st = Uniform2DCart((1e5, 200.0), [1.0], [[0, box]] * 2, 256)
    frac = np.pi * np.sum(r ** 2) / box ** 2
    assert m.mask(st).mean() == pytest.approx(frac, rel=0.03)
    with pytest.raises(ValueError):
        m.mask(Uniform3DCart((1e5, 200.0), [1.0], [[0, box]] * 3, 16))


def test_lamellar_seed_replicates_the_1d_seed():
    s1, r1 = configs.c1c10_interface(ngrid=128, box=10.0)
    s2, r2 = configs.c1c10_lamellar(ngrid=(128, 16), box=10.0)
    rows = r2.reshape(2, 16, 128)
    for j in range(16):
        assert np.array_equal(rows[:, j, :], r1)
