"""Post-processing callers of converge!: surface tension and adsorption.

  surface_tension(model, T, x)        src/methods/surface_tension.jl:13-29
  surface_tension(system, rho)        src/methods/surface_tension.jl:31-53
  interfacial_tension(model, p, T, x, xx) / (system, rho)
                                      src/methods/interfacial_tension.jl:17-63
  adsorption(model, surface, p, T, n) src/methods/adsorption.jl:48-66
  adsorption(system, rho)             src/methods/adsorption.jl:22-32
"""
import numpy as np

from .bulk import KB
from .structure import TwoPhase1DCart, Uniform1DCart
from .system import DFTSystem, converge_, free_energy, initialize_profiles, integrate, simpson


def bulk_chemical_potential(system):
    """mu_i (J/mol) of the structure's first bulk phase, as surface_tension(system, rho)
    takes it (src/methods/surface_tension.jl:45)."""
    st = system.structure
    T = st.conditions[1]
    rl = st.rho_bulk
    return np.asarray(system.model.VT_chemical_potential(1.0 / rl.sum(), T, rl / rl.sum()), dtype=np.float64)


def surface_tension_system(system, rho):
    """F k_B T - sum_i mu_i int rho_i + p int 1 (for the periodic box: two interfaces)."""
    rho = np.asarray(rho, dtype=np.float64).reshape(system.nc, system.N)
    st = system.structure
    p, T = st.conditions
    F = free_energy(system, rho)
    mu = bulk_chemical_potential(system)
    chem = sum(mu[i] * integrate(st, rho[i]) for i in range(system.nc))
    return F * KB * T - chem + p * integrate(st, np.ones(system.N))


def surface_tension(model_or_system, T_or_rho, x=None, *, ngrid=201, box=10.0, options=None, **kw):
    """surface_tension(model, T, x=[1.0]) -> gamma (N/m), or surface_tension(system, rho)."""
    if isinstance(model_or_system, DFTSystem):
        return surface_tension_system(model_or_system, T_or_rho)
    model, T = model_or_system, float(T_or_rho)
    x = np.array([1.0]) if x is None else np.asarray(x, dtype=np.float64)
    L = model.length_scale()
    p, vl, vv, y = model.bubble_pressure(T, x)
    rho1 = x / vl
    rho2 = y / vv
    structure = TwoPhase1DCart((p, T), rho1, rho2, [-box * L, box * L], (ngrid,))
    system = DFTSystem(model, structure, None, options)
    rho = initialize_profiles(system)
    converge_(system, rho, **kw)
    return surface_tension_system(system, rho) / 2


def interfacial_tension(model_or_system, p_or_rho, T=None, x=None, xx=None, *, ngrid=201, box=20.0, options=None,
                        **kw):
    """interfacial_tension(model, p, T, x, xx) -> gamma (N/m) of the liquid-liquid
    interface between phases of compositions x and xx at (p, T), or
    interfacial_tension(system, rho), which is surface_tension(system, rho)."""
    if isinstance(model_or_system, DFTSystem):
        return surface_tension_system(model_or_system, p_or_rho)
    model, p = model_or_system, float(p_or_rho)
    x = np.asarray(x, dtype=np.float64)
    xx = np.asarray(xx, dtype=np.float64)
    L = model.length_scale()
    v1 = model.volume(p, T, x)
    v2 = model.volume(p, T, xx)
    structure = TwoPhase1DCart((p, T), x / v1, xx / v2, [-box * L, box * L], (ngrid,))
    system = DFTSystem(model, structure, None, options)
    rho = initialize_profiles(system)
    converge_(system, rho, **kw)
    return surface_tension_system(system, rho) / 2


def adsorption_system(system, rho):
    rho = np.asarray(rho, dtype=np.float64).reshape(system.nc, system.N)
    V = system.structure.volume()
    return np.array([integrate(system.structure, rho[i]) / V for i in range(system.nc)])


def adsorption(model_or_system, surface_or_rho, p=None, T=None, n=(1.0,), *, options=None, **kw):
    """adsorption(model, surface, p, T, n=[1.0]) or adsorption(system, rho)."""
    if isinstance(model_or_system, DFTSystem):
        return adsorption_system(model_or_system, surface_or_rho)
    model, surface = model_or_system, surface_or_rho
    L = model.length_scale()
    width = surface.width
    bounds = [0.7 * L, width - 0.7 * L]
    n = np.asarray(n, dtype=np.float64)
    v = model.volume(p, T, n)
    rho_b = n / v
    structure = Uniform1DCart((p, T), rho_b, bounds, (201,))
    system = DFTSystem(model, structure, surface, options)
    rho = initialize_profiles(system)
    converge_(system, rho, **kw)
    return adsorption_system(system, rho)
