"""Host-side mirror of cDFT.jl's DFTSystem / initialize_profiles / converge! API
for PC-SAFT on 1D Cartesian grids, running the hot path through the HIP library.

  DFTOptions            src/base/devices.jl:29-40
  DFTSystem             src/base/base.jl:50-89 (get_species, PCSAFT.jl:65-72)
  initialize_profiles   src/structure/structure.jl:50-136, two_phase.jl:1-35,
                        external_field.jl:21-26
  converge_             converge!(system, rho; ...), src/methods/converge.jl:214-254
  dFdrho_res, F_res     src/models/models.jl:86-91, 194-200
  F_ideal, free_energy  src/models/BasicIdeal.jl:10-39, src/utils/base.jl:8-10

Profiles are numpy float64 arrays of shape (nc, N): rho[c, i] is the reference's
rho[i, c] (the same memory layout as the column-major Julia array).
"""
import math

import numpy as np

from . import _lib
from .bulk import KB, NA, RGAS
from .structure import Structure

PSI = 1.3862


class DFTOptions:
    """device: HIP device index; precision: 'fp64' (default) or 'mixed'."""

    def __init__(self, device=0, precision="fp64"):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.device = int(device)
        self.precision = precision

    @property
    def precision_code(self):
        return _lib.CDFT_FP64 if self.precision == "fp64" else _lib.CDFT_MIXED


class Species:
    def __init__(self, size, bulk_density, chempot_res):
        self.nbeads = np.ones(len(size), dtype=np.int64)
        self.size = np.asarray(size, dtype=np.float64)
        self.bulk_density = np.asarray(bulk_density, dtype=np.float64)
        self.chempot_res = np.asarray(chempot_res, dtype=np.float64)


def get_species(model, structure):
    """d(T) and beta mu_res of the bulk (PCSAFT.jl:65-72)."""
    T = structure.conditions[1]
    rb = structure.rho_bulk
    size = model.d(T)
    mu = model.VT_chemical_potential_res(1.0 / rb.sum(), T, rb / rb.sum()) / RGAS / T
    return Species(size, rb, mu)


class DFTSystem:
    def __init__(self, model, structure, external_field=None, options=None):
        if not isinstance(structure, Structure):
            raise TypeError("structure must be a Structure")
        self.model = model
        self.structure = structure
        self.options = options if options is not None else DFTOptions()
        if external_field is None:
            fields = []
        elif isinstance(external_field, (list, tuple)):
            fields = list(external_field)
        else:
            fields = [external_field]
        self.external_field = fields
        self.species = get_species(model, structure)
        self.nc = len(model)
        self.N = structure.npoints
        self.L = model.length_scale()
        self.vext = None
        if fields:
            v = np.zeros(self.N * self.nc)
            for f in fields:
                v += f.vext(structure, model)
            self.vext = v
        self.last_result = None
        self._dev = None

    @property
    def device_system(self):
        if self._dev is None:
            ctx = _lib.default_context(self.options.device)
            p, T = self.structure.conditions
            st = self.structure
            self._dev = _lib.DeviceSystem(
                ctx, nc=self.nc, ndim=st.ndim, ngrid=st.ngrid, lb=[b[0] for b in st.bounds],
                ub=[b[1] for b in st.bounds], T=T, p=p, L=self.L,
                psi=PSI, m=self.model.segment, d=self.species.size, sigma=self.model.sigma,
                eps=self.model.epsilon, rho_bulk=self.species.bulk_density, mu_res=self.species.chempot_res,
                vext=self.vext, precision=self.options.precision_code, coord=st.coord, **self.assoc_descriptor())
        return self._dev

    def descriptor(self):
        """Plain-value description of the system (the inputs every path receives)."""
        p, T = self.structure.conditions
        st = self.structure
        return dict(nc=self.nc, nd=st.ndim, ngrid=list(st.ngrid), lb=[b[0] for b in st.bounds],
                    ub=[b[1] for b in st.bounds], T=T, L=self.L,
                    m=self.model.segment.copy(), d=self.species.size.copy(), sigma=self.model.sigma.copy(),
                    eps=self.model.epsilon.copy(), mu_res=self.species.chempot_res.copy(),
                    rho_bulk=self.species.bulk_density.copy(), vext=None if self.vext is None else self.vext.copy(),
                    coord=st.coord, **self.assoc_descriptor())

    def assoc_descriptor(self):
        """Association sites and bonding pairs (pack_assoc_params, assoc.jl:18-57);
        empty for a non-associating model."""
        mdl = self.model
        pairs = getattr(mdl, "assoc_pairs", [])
        return dict(assoc_site_comp=list(getattr(mdl, "site_comp", [])),
                    assoc_site_count=list(getattr(mdl, "site_count", [])),
                    assoc_pairs=[(a, b) for a, b, _, _ in pairs],
                    assoc_eps=[e for _, _, e, _ in pairs], assoc_kappa=[k for _, _, _, k in pairs])


def _cos_prof(x, start, stop, shift, coef):
    """src/structure/structure.jl:2"""
    c = np.cos((x - shift) * 2 * np.pi)
    return 0.5 * (start - stop) * c * np.sqrt((1 + coef ** 2) / (1 + coef ** 2 * c ** 2)) + 0.5 * (start + stop)


def _clamp_init(rho):
    lo = np.finfo(np.float64).tiny ** (1.0 / 6.0)
    hi = np.finfo(np.float64).max ** (1.0 / 6.0)
    return np.clip(rho, lo, hi)


def initialize_profiles(system, noise=0.0, rng=None):
    """Seed profile (nc, N): uniform bulk, or the TwoPhase1DCart / TwoPhase2DLamCart cos_prof seed, times
    exp(-beta V/10) per external field, then clamped (structure.jl:50-94)."""
    st = system.structure
    N, nc = system.N, system.nc
    rho = np.empty((nc, N))
    if st.topology == "two_phase":
        lb, ub = st.bounds[0]
        H = ub - lb
        T = st.conditions[1]
        x = st.axis_coords(0) / (ub - lb)
        if st.ndim == 2:                    # lamellar: the 1D profile along x, uniform along y (two_phase.jl:37-69)
            x = np.tile(x, st.ngrid[1])
        shift = ub / 4 + 3 * lb / 4
        for i, pure in enumerate(system.model.split_pure()):
            Tc = pure.crit_pure()[0]
            coef = (2.4728 - 2.3625 * T / Tc) * (H / system.L)
            coef = math.sqrt(coef ** 2 - 1) / 4
            rho[i] = _cos_prof(x, st.rho_bulk[i], st.rho_bulk2[i], shift, coef)
    else:
        for i in range(nc):
            rho[i] = st.rho_bulk[i]
    if system.vext is not None:
        for f in system.external_field:
            rho *= np.exp(-f.vext(st, system.model).reshape(nc, N) / 10.0)
    if noise:
        rng = rng or np.random.default_rng()
        tot = rho.sum(axis=0)
        rho *= 1 + noise * (2 * rng.random(rho.shape) - 1)
        rho *= tot / rho.sum(axis=0)
    return _clamp_init(rho)


def two_phase_Tc(system):
    return np.array([p.crit_pure()[0] for p in system.model.split_pure()])


def converge_(system, rho, maxit=10000, beta=1e-2, tol=1e-4, anderson_start=1e-1, anderson_m=5,
              verbose=False, log_interval=0, save_interval=0, save_callback=None, history=0):
    """converge!(system, rho; kwargs...): rho (nc, N) float64 is overwritten with the
    converged profile; returns None like the reference.  The solver outcome
    (idid, errcode, iterations, residual) is kept in system.last_result."""
    if log_interval < 0 or save_interval < 0:
        raise ValueError("log_interval and save_interval must be >= 0")
    if save_interval > 0 and save_callback is None:
        raise ValueError("save_interval > 0 needs a save_callback")
    if not (isinstance(rho, np.ndarray) and rho.dtype == np.float64 and rho.flags.c_contiguous):
        raise TypeError("rho must be a C-contiguous float64 numpy array of shape (nc, N)")
    progress, every = None, 0
    if log_interval > 0 or save_interval > 0:
        # GFix_logger (converge.jl:256-267): after evaluation it, the free energy of the
        # profile that evaluation used every log_interval, save_callback(it, rho) every
        # save_interval; the device solve pauses at multiples of their gcd
        every = math.gcd(log_interval, save_interval)
        shape = rho.shape

        def progress(it, r):
            r = r.reshape(shape)
            if log_interval > 0 and it % log_interval == 0:
                print(f"DFT iter {it:5d} | F = {free_energy(system, r):.6g}")
            if save_interval > 0 and it % save_interval == 0:
                save_callback(it, r)
    res = system.device_system.converge(rho.reshape(-1), maxit=maxit, beta=beta, tol=tol,
                                        anderson_start=anderson_start, m=anderson_m, hist_cap=history,
                                        progress=progress, progress_every=every)
    system.last_result = res
    if verbose:
        msg = "DFT converged" if res["idid"] else "DFT did not converge"
        print(f"{msg} after {res['nevals']} iterations: err = {res['resnorm']:.3g}")
    return None


def dFdrho_res(system, rho):
    """delta F_res / delta rho + beta V_ext (models.jl:194-200); shape (nc, N)."""
    out, _, _ = system.device_system.functional_eval(np.asarray(rho, dtype=np.float64).reshape(-1))
    if system.vext is not None:
        out = out + system.vext
    return out.reshape(system.nc, system.N)


def F_res(system, rho):
    _, fres, _ = system.device_system.functional_eval(np.asarray(rho, dtype=np.float64).reshape(-1))
    return fres


def simpson(f, dz):
    """_int (src/utils/integrals.jl:28-40): weights 1,4,2,...,(4|2),1 times dz/3."""
    f = np.asarray(f, dtype=np.float64)
    n = f.size
    k = np.arange(1, n + 1)
    c = np.where(k % 2 == 0, 4.0, 2.0)
    c[0] = 1.0
    c[-1] = 1.0
    return float(np.sum(c * f) * (dz / 3.0))


def integration_weights(structure):
    """Weights w with _int(f) = w . f over the whole grid (column-major, x fastest)."""
    if structure.coord:
        from .radial import weights
        return weights(structure.coord, structure.bounds[0][1], structure.ngrid[0])
    w = None
    for n, dz in zip(structure.ngrid, structure.dzs()):
        k = np.arange(1, n + 1)
        c = np.where(k % 2 == 0, 4.0, 2.0)
        c[0] = 1.0
        c[-1] = 1.0
        c = c * (dz / 3.0)
        w = c if w is None else np.outer(c, w).reshape(-1)
    return w


def integrate(structure, f):
    """_int over every axis of the structure (src/utils/integrals.jl:28-40), f column-major;
    on a radial grid 4 pi or 2 pi times the transform's quadrature (integrals.jl:20-22)."""
    if structure.coord:
        from .radial import weights
        return float(weights(structure.coord, structure.bounds[0][1], structure.ngrid[0]) @ np.asarray(f).reshape(-1))
    if structure.ndim == 1:
        return simpson(f, structure.dz())
    f = np.asarray(f, dtype=np.float64).reshape(structure.ngrid[::-1])      # [z, y, x]
    out = f
    for a, dz in enumerate(structure.dzs()):                                # x first
        n = structure.ngrid[a]
        k = np.arange(1, n + 1)
        c = np.where(k % 2 == 0, 4.0, 2.0)
        c[0] = 1.0
        c[-1] = 1.0
        out = np.tensordot(out, c * (dz / 3.0), axes=([out.ndim - 1], [0]))
    return float(out)


def F_ideal(system, rho):
    T = system.structure.conditions[1]
    rho = np.asarray(rho, dtype=np.float64).reshape(system.nc, system.N)
    phi = np.sum(NA * rho * (np.log(rho) - 1.5 * math.log(T) - 1.0), axis=0)
    return integrate(system.structure, phi)


def free_energy(system, rho):
    return F_ideal(system, rho) + F_res(system, rho)
