"""Parameter sensitivities of converged profiles through the fixed-point adjoint.

No reference counterpart: the reference's only derivative is dPhi/dn
(src/models/models.jl:355-411).  For an objective J(rho, theta) of a converged
profile rho* = exp(x*), x* = G(x*, theta) the undamped Euler-Lagrange map
(converge.jl:108-153),

    dJ/dtheta = dJ/dtheta|_rho + lambda . dG/dtheta|_x,   (I - dG/dx)^T lambda = rho * dJ/drho,

with lambda and the products lambda . dG/dtheta from cdft_adjoint_params (Hessian-vector
products of F_res on the device, the same damped Picard / Anderson iteration as
converge!).  The two explicit
partials at fixed profile are central differences of one functional evaluation
each, on systems rebuilt with the perturbed parameter over the same grid: two
evaluations per parameter instead of two full solves.

Objectives: ``gamma`` (surface_tension(system, rho) / 2, the planar interface's
tension, surface_tension.jl:28) and ``adsorption`` of one component
(adsorption.jl:22-32).
"""
import numpy as np

from .bulk import KB, NA, PCSAFT
from .methods import adsorption_system, surface_tension_system
from .structure import Structure
from .system import DFTSystem, integration_weights

PARAMS = ("sigma", "epsilon", "segment")


def simpson_weights(n, dz):
    """Weights of _int along one axis (src/utils/integrals.jl:28-40): 1, 4, 2, ..., (4|2), 1 times dz / 3."""
    k = np.arange(1, n + 1)
    c = np.where(k % 2 == 0, 4.0, 2.0)
    c[0] = 1.0
    c[-1] = 1.0
    return c * (dz / 3.0)


def model_params(model):
    """Per-component (m, sigma in Angstrom, eps/k in K, Mw) rows of a PCSAFT model."""
    return {c: (float(model.segment[i]), float(model.sigma[i, i]) * 1e10, float(model.epsilon[i, i]),
                float(model.Mw[i])) for i, c in enumerate(model.components)}


def perturbed_system(system, comp, param, delta):
    """The system rebuilt with parameter `param` of component index `comp` shifted by
    `delta` (sigma in Angstrom, epsilon in K, segment dimensionless), over the same grid
    and bounds, at the same temperature and, for a uniform bulk, the same pressure and
    composition; a two-phase structure takes the new bubble point at the same liquid
    composition (surface_tension.jl:13-29)."""
    model = system.model
    rows = model_params(model)
    name = model.components[comp]
    r = list(rows[name])
    r[{"segment": 0, "sigma": 1, "epsilon": 2}[param]] += delta
    rows[name] = tuple(r)
    new = PCSAFT(model.components, kij=model.kij, params=rows,
                 assoc={c: model.assoc_like.get(c) for c in model.components}, combining=model.combining)
    st = system.structure
    p, T = st.conditions
    x = st.rho_bulk / st.rho_bulk.sum()
    if st.topology == "two_phase":
        p, vl, vv, y = new.bubble_pressure(T, x)
        nst = Structure((p, T), x / vl, list(st.bounds), st.ngrid, topology="two_phase", rho_bulk2=y / vv,
                        ndim=st.ndim)
    else:
        phase = "liquid" if st.rho_bulk.sum() * model.volume(p, T, x, phase="liquid") > 0.5 else "vapour"
        v = new.volume(p, T, x, phase=phase)
        nst = Structure((p, T), x / v, list(st.bounds), st.ngrid, topology=st.topology, ndim=st.ndim)
    return DFTSystem(new, nst, system.external_field or None, system.options)


def objective(system, rho, kind, comp=0):
    if kind == "gamma":
        return 0.5 * surface_tension_system(system, rho)
    if kind == "adsorption":
        return float(adsorption_system(system, rho)[comp])
    raise ValueError(f"unknown objective {kind!r}")


def objective_grad_rho(system, rho, kind, comp=0):
    """dJ/drho at fixed parameters, shape (nc * N,)."""
    st = system.structure
    N, nc = system.N, system.nc
    rho = np.asarray(rho, dtype=np.float64).reshape(nc, N)
    w = integration_weights(st)
    out = np.zeros((nc, N))
    if kind == "adsorption":
        out[comp] = w / st.volume()
        return out.reshape(-1)
    if kind == "gamma":
        p, T = st.conditions
        rl = st.rho_bulk
        mu = system.model.VT_chemical_potential(1.0 / rl.sum(), T, rl / rl.sum())
        dres = NA * system.device_system.phi_vjp(rho.reshape(-1), w).reshape(nc, N)   # d F_res / d rho
        for i in range(nc):
            did = NA * w * (np.log(rho[i]) - 1.5 * np.log(T))                      # d F_ideal / d rho
            out[i] = 0.5 * (KB * T * (did + dres[i]) - mu[i] * w)
        return out.reshape(-1)
    raise ValueError(f"unknown objective {kind!r}")


def log_map(system, rho):
    """G(x) of the undamped map at x = log rho: clamp(mu_res + ln rho_b - dF/drho - V, -100, 100)."""
    dev = system.device_system
    dfd, _, _ = dev.functional_eval(np.asarray(rho, dtype=np.float64).reshape(-1))
    v = system.vext if system.vext is not None else 0.0
    cp = np.repeat(system.species.chempot_res + np.log(system.species.bulk_density), system.N)
    return np.clip(cp - dfd - v, -100.0, 100.0)


def translation_mode(system, rho):
    """d(log rho)/d(phase), phase = 2 pi (z - lb)/(ub - lb): the spectral derivative of
    log rho on the periodic grid (a multiple of d(log rho)/dz), shape (nc * N,).

    Without an external field every operator of the map is circulant, so the
    fixed point comes in a family of translates and I - dG/dx is singular along
    this vector: the adjoint equation is then solvable for the part of g
    orthogonal to it."""
    N, nc = system.N, system.nc
    x = np.log(np.asarray(rho, dtype=np.float64).reshape(nc, N))
    k = np.fft.fftfreq(N) * N
    if N % 2 == 0:
        k[N // 2] = 0.0
    return np.real(np.fft.ifft(1j * k * np.fft.fft(x, axis=1), axis=1)).reshape(-1)


def jacobian(system, rho, kind, comps, params=PARAMS, rel_h=1e-6, tol=1e-10, maxit=10000, beta=1e-2,
             anderson_start=1e-1, m=5, deflate=None):
    """dJ_c/dtheta for the objectives J_c (kind, component c for c in comps) and every
    theta = (param, component): one adjoint solve per objective, and one pair of
    perturbed systems per parameter shared by all objectives.

    deflate (default: when the system has no external field) removes from g its
    component along the translation mode, the solvability condition of the
    adjoint equation for a free interface.  Returns ({(c, param, i): derivative},
    {c: lambda}, {c: adjoint result})."""
    rho = np.asarray(rho, dtype=np.float64).reshape(-1)
    if deflate is None:
        deflate = system.vext is None
    t = translation_mode(system, rho) if deflate else None
    dev = system.device_system
    comps = list(comps)
    g = np.empty((len(comps), rho.size))
    for k, c in enumerate(comps):
        gk = rho * objective_grad_rho(system, rho, kind, c)
        if deflate:
            gk = gk - (gk @ t) / (t @ t) * t
        g[k] = gk
    # the multipliers of every objective (cdft_adjoint_params, nparam = 0)
    lam = np.zeros_like(g)
    _, res = dev.adjoint_params(rho, lam, g=g, maxit=maxit, beta=beta, tol=tol, anderson_start=anderson_start, m=m)
    lams = {c: lam[k] for k, c in enumerate(comps)}
    results = {c: res[k] for k, c in enumerate(comps)}
    out = {}
    rows = model_params(system.model)
    for i, name in enumerate(system.model.components):
        for prm in params:
            base = rows[name][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
            h = rel_h * abs(base)
            sp = perturbed_system(system, i, prm, +h)
            sm = perturbed_system(system, i, prm, -h)
            # lambda_k . dG/dtheta on the device, for all objectives at once
            imp, _ = dev.adjoint_params(rho, lam, pairs=[(sp.device_system, sm.device_system)], h=[h])
            for k, c in enumerate(comps):
                dJ = (objective(sp, rho, kind, c) - objective(sm, rho, kind, c)) / (2 * h)
                out[(c, prm, i)] = dJ + float(imp[k, 0])
            del sp, sm                # release their device systems before the next pair
    return out, lams, results


def sensitivities(system, rho, kind, comp=0, params=PARAMS, rel_h=1e-6, tol=1e-10, maxit=10000, beta=1e-2,
                  anderson_start=1e-1, m=5, deflate=None):
    """dJ/dtheta of one objective for theta = (param, component) pairs through the
    adjoint.  Returns ({(param, i): derivative}, lambda, adjoint result)."""
    out, lams, results = jacobian(system, rho, kind, [comp], params, rel_h, tol, maxit, beta, anderson_start, m,
                                  deflate)
    return {(prm, i): v for (_, prm, i), v in out.items()}, lams[comp], results[comp]
