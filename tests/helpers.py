"""Shared helpers for the tests: build oracle systems from host-API systems."""
import numpy as np

import pyoracle as po


def oracle_of(system):
    d = system.descriptor()
    return po.Oracle(**d)


def smooth_profile(system, seed=0, amp=0.3):
    """A positive, smooth, non-uniform profile around the bulk densities."""
    rng = np.random.default_rng(seed)
    N, nc = system.N, system.nc
    z = np.arange(N) / N
    rho = np.empty((nc, N))
    for c in range(nc):
        a = rng.uniform(-amp, amp, size=4)
        ph = rng.uniform(0, 2 * np.pi, size=4)
        mod = 1 + sum(a[k] * np.sin(2 * np.pi * (k + 1) * z + ph[k]) for k in range(4)) / 2
        rho[c] = system.species.bulk_density[c] * mod
    return rho



def bulk_mu(system):
    """mu_i of the bulk (J/mol) the way surface_tension(system, rho) takes it."""
    from cdft_amd.methods import bulk_chemical_potential
    return bulk_chemical_potential(system)
