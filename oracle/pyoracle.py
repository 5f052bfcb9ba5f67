"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package never imports
this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAXC = 8
MAXS = 20
MAXP = 20

_dp = C.POINTER(C.c_double)


class OrSystem(C.Structure):
    _fields_ = [
        ("nc", C.c_int), ("nd", C.c_int), ("ngrid", C.c_int * 3),
        ("lb", C.c_double * 3), ("ub", C.c_double * 3),
        ("T", C.c_double), ("L", C.c_double), ("psi", C.c_double),
        ("m", C.c_double * MAXC), ("d", C.c_double * MAXC),
        ("sigma", C.c_double * (MAXC * MAXC)), ("eps", C.c_double * (MAXC * MAXC)),
        ("mu_res", C.c_double * MAXC), ("rho_bulk", C.c_double * MAXC),
        ("vext", _dp),
        ("nsite", C.c_int), ("npair", C.c_int),
        ("site_comp", C.c_int * MAXS), ("site_count", C.c_double * MAXS),
        ("pair_a", C.c_int * MAXP), ("pair_b", C.c_int * MAXP),
        ("pair_eps", C.c_double * MAXP), ("pair_kappa", C.c_double * MAXP),
        ("coord", C.c_int),
    ]


class OrAAOpts(C.Structure):
    _fields_ = [("m", C.c_int), ("maxit", C.c_int), ("picard_maxit", C.c_int),
                ("rtol", C.c_double), ("atol", C.c_double), ("beta", C.c_double),
                ("picard_beta", C.c_double), ("picard_rtol", C.c_double),
                ("picard_atol", C.c_double)]


class OrAAResult(C.Structure):
    _fields_ = [("idid", C.c_int), ("errcode", C.c_int), ("k", C.c_int),
                ("nhist", C.c_int), ("nstats", C.c_int), ("nevals", C.c_int),
                ("resnorm", C.c_double), ("tol", C.c_double)]


GFIX = C.CFUNCTYPE(None, _dp, _dp, C.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = C.CDLL(_LIB)
        P = C.POINTER(OrSystem)
        L.or_fft.argtypes = [_dp, _dp, C.c_int, C.POINTER(C.c_int), C.c_int]
        L.or_omega.argtypes = [P, _dp, _dp]
        L.or_weighted_densities.argtypes = [P, _dp, _dp]
        L.or_phi.argtypes = [P, _dp, _dp]
        L.or_phi.restype = C.c_double
        L.or_assoc_X.argtypes = [P, _dp, _dp]
        L.or_transposed.argtypes = [P, _dp, _dp]
        L.or_dFdrho_res.argtypes = [P, _dp, _dp, _dp]
        L.or_simpson.argtypes = [_dp, C.c_int, C.POINTER(C.c_int), _dp]
        L.or_simpson.restype = C.c_double
        L.or_F_res.argtypes = [P, _dp]
        L.or_F_res.restype = C.c_double
        L.or_F_ideal.argtypes = [P, _dp]
        L.or_F_ideal.restype = C.c_double
        L.or_get_new_profile.argtypes = [P, _dp, _dp]
        L.or_aasol.argtypes = [GFIX, C.c_void_p, C.c_int, _dp, C.POINTER(OrAAOpts), _dp, _dp,
                               _dp, C.c_int, _dp, _dp, C.c_int, C.POINTER(OrAAResult)]
        L.or_converge.argtypes = [P, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(OrAAResult), _dp, C.c_int]
        L.or_steele_vext.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp]
        L.or_init_profiles.argtypes = [P, _dp]
        L.or_init_two_phase.argtypes = [P, _dp, _dp, _dp]
        L.or_surface_tension.argtypes = [P, _dp, _dp, C.c_double]
        L.or_surface_tension.restype = C.c_double
        L.or_adsorption.argtypes = [P, _dp, _dp]
        L.or_plan_create.argtypes = [P]
        L.or_plan_create.restype = C.c_void_p
        L.or_plan_destroy.argtypes = [C.c_void_p]
        L.or_dFdrho_res_planned.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_converge_planned.argtypes = L.or_converge.argtypes
        L.or_converge_batch.argtypes = [P, C.c_int, C.POINTER(_dp), C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_int, C.POINTER(OrAAResult), _dp]
        L.or_wtime.restype = C.c_double
        L.or_ddft_rhs.argtypes = [P, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One oracle system (mirrors the C struct; keeps numpy buffers alive)."""

    def __init__(self, *, nc, nd, ngrid, lb, ub, T, L, m, d, sigma, eps, mu_res, rho_bulk,
                 vext=None, psi=1.3862, assoc_site_comp=(), assoc_site_count=(), assoc_pairs=(),
                 assoc_eps=(), assoc_kappa=(), coord=0):
        s = OrSystem()
        s.nc, s.nd = nc, nd
        s.coord = int(coord)
        for a in range(nd):
            s.ngrid[a] = int(ngrid[a]); s.lb[a] = float(lb[a]); s.ub[a] = float(ub[a])
        for a in range(nd, 3):
            s.ngrid[a] = 1
        s.T, s.L, s.psi = float(T), float(L), float(psi)
        sig = np.asarray(sigma, dtype=np.float64).reshape(nc, nc)
        ep = np.asarray(eps, dtype=np.float64).reshape(nc, nc)
        for i in range(nc):
            s.m[i] = float(m[i]); s.d[i] = float(d[i])
            s.mu_res[i] = float(mu_res[i]); s.rho_bulk[i] = float(rho_bulk[i])
            for j in range(nc):
                s.sigma[i * nc + j] = sig[i, j]
                s.eps[i * nc + j] = ep[i, j]
        s.nsite, s.npair = len(assoc_site_comp), len(assoc_pairs)
        if s.nsite > MAXS or s.npair > MAXP:
            raise ValueError("too many association sites or pairs")
        for t in range(s.nsite):
            s.site_comp[t] = int(assoc_site_comp[t]); s.site_count[t] = float(assoc_site_count[t])
        for q in range(s.npair):
            s.pair_a[q], s.pair_b[q] = int(assoc_pairs[q][0]), int(assoc_pairs[q][1])
            s.pair_eps[q], s.pair_kappa[q] = float(assoc_eps[q]), float(assoc_kappa[q])
        self.N = int(np.prod(ngrid[:nd]))
        self.nc, self.nd, self.F = nc, nd, 6 + nd
        self._vext = None
        if vext is not None:
            self._vext = _f64(vext).reshape(-1)
            s.vext = _ptr(self._vext)
        self.s = s

    def set_vext(self, vext):
        self._vext = _f64(vext).reshape(-1)
        self.s.vext = _ptr(self._vext)

    @property
    def ref(self):
        return C.byref(self.s)

    def omega(self):
        n = self.N * self.F * self.nc
        re = np.zeros(n); im = np.zeros(n)
        lib().or_omega(self.ref, _ptr(re), _ptr(im))
        return re + 1j * im

    def weighted_densities(self, rho):
        rho = _f64(rho).reshape(-1)
        out = np.zeros(self.N * self.F * self.nc)
        lib().or_weighted_densities(self.ref, _ptr(rho), _ptr(out))
        return out

    def phi(self, nloc):
        nloc = _f64(nloc).reshape(-1)
        d = np.zeros(self.F * self.nc)
        v = lib().or_phi(self.ref, _ptr(nloc), _ptr(d))
        return v, d

    def assoc_X(self, nloc):
        X = np.ones(max(self.s.nsite, 1))
        lib().or_assoc_X(self.ref, _ptr(_f64(nloc)), _ptr(X))
        return X[:self.s.nsite]

    def transposed(self, dphi):
        dphi = _f64(dphi).reshape(-1)
        out = np.zeros(self.N * self.nc)
        lib().or_transposed(self.ref, _ptr(dphi), _ptr(out))
        return out

    def dFdrho_res(self, rho):
        rho = _f64(rho).reshape(-1)
        out = np.zeros(self.N * self.nc)
        fv = np.zeros(self.N)
        lib().or_dFdrho_res(self.ref, _ptr(rho), _ptr(out), _ptr(fv))
        return out, fv

    def dFdrho_res_planned(self, rho):
        """The same operator with the kernel spectra computed once (1D only)."""
        rho = _f64(rho).reshape(-1)
        out = np.zeros(self.N * self.nc)
        fv = np.zeros(self.N)
        p = lib().or_plan_create(self.ref)
        if not p:
            raise ValueError("planned evaluation is 1D only")
        try:
            lib().or_dFdrho_res_planned(p, _ptr(rho), _ptr(out), _ptr(fv))
        finally:
            lib().or_plan_destroy(p)
        return out, fv

    def surface_tension_planned(self, rho, mu, p):
        """or_surface_tension with F_res from the planned evaluation (1D only)."""
        rho = _f64(rho).reshape(-1)
        _, fv = self.dFdrho_res_planned(rho)
        ng = (C.c_int * 3)(*self.s.ngrid)
        dz = np.array([(self.s.ub[a] - self.s.lb[a]) / self.s.ngrid[a] for a in range(3)])
        simpson = lambda f: lib().or_simpson(_ptr(_f64(f)), self.nd, ng, _ptr(dz))
        fres = simpson(fv) / self.s.L ** 3
        F = self.F_ideal(rho) + fres
        chem = sum(float(mu[c]) * simpson(rho[c * self.N:(c + 1) * self.N]) for c in range(self.nc))
        KB = 1.380649e-23
        return F * KB * self.s.T - chem + p * simpson(np.ones(self.N))

    def ddft_rhs(self, eta):
        """d(eta)/dt of the DDFT log-density form (ext/SciMLBasecDFTExt.jl:10-61)."""
        eta = _f64(eta).reshape(-1)
        out = np.zeros(self.N * self.nc)
        lib().or_ddft_rhs(self.ref, _ptr(eta), _ptr(out))
        return out

    def F_res(self, rho):
        return lib().or_F_res(self.ref, _ptr(_f64(rho).reshape(-1)))

    def F_ideal(self, rho):
        return lib().or_F_ideal(self.ref, _ptr(_f64(rho).reshape(-1)))

    def get_new_profile(self, rho):
        rho = _f64(rho).reshape(-1)
        out = np.zeros(self.N * self.nc)
        lib().or_get_new_profile(self.ref, _ptr(rho), _ptr(out))
        return out

    def steele_vext(self, eps_s, sigma_s, delta, rho_s, width):
        out = np.zeros(self.N * self.nc)
        lib().or_steele_vext(self.ref, eps_s, sigma_s, delta, rho_s, width, _ptr(out))
        return out

    def init_profiles(self):
        out = np.zeros(self.N * self.nc)
        lib().or_init_profiles(self.ref, _ptr(out))
        return out

    def init_two_phase(self, rho2, Tc):
        out = np.zeros(self.N * self.nc)
        lib().or_init_two_phase(self.ref, _ptr(_f64(rho2)), _ptr(_f64(Tc)), _ptr(out))
        return out

    def converge(self, rho, *, maxit=10000, beta=1e-2, tol=1e-4, anderson_start=1e-1, m=5,
                 hist_cap=100000, planned=False):
        """or_converge (literal operator) or, with planned=True (1D), or_converge_planned."""
        rho = _f64(rho).reshape(-1).copy()
        res = OrAAResult()
        hist = np.zeros(hist_cap)
        fn = lib().or_converge_planned if planned else lib().or_converge
        rc = fn(self.ref, _ptr(rho), maxit, beta, tol, anderson_start, m, C.byref(res), _ptr(hist), hist_cap)
        if rc == -3:
            raise ValueError("planned solve is 1D only")
        return rho, res, hist[:min(res.nhist, hist_cap)]

    def surface_tension(self, rho, mu, p):
        return lib().or_surface_tension(self.ref, _ptr(_f64(rho).reshape(-1)), _ptr(_f64(mu)), p)

    def adsorption(self, rho):
        out = np.zeros(self.nc)
        lib().or_adsorption(self.ref, _ptr(_f64(rho).reshape(-1)), _ptr(out))
        return out


def converge_batch(oracles, rhos, *, maxit=10000, beta=1e-2, tol=1e-4, anderson_start=1e-1, m=5,
                   nthreads=1):
    """or_converge_batch: independent 1D systems, planned solves on nthreads OpenMP
    threads.  Returns (profiles, results, per-system seconds, wall seconds)."""
    n = len(oracles)
    arr = (OrSystem * n)(*[o.s for o in oracles])
    outs = [_f64(r).reshape(-1).copy() for r in rhos]
    ptrs = (_dp * n)(*[_ptr(r) for r in outs])
    res = (OrAAResult * n)()
    secs = np.zeros(n)
    t0 = lib().or_wtime()
    rc = lib().or_converge_batch(arr, n, ptrs, maxit, beta, tol, anderson_start, m, int(nthreads), res, _ptr(secs))
    wall = lib().or_wtime() - t0
    if rc != 0:
        raise ValueError("or_converge_batch failed (only 1D systems are supported)")
    return outs, list(res), secs, wall


def aasol(gfun, x0, *, m, maxit=20, rtol=1e-10, atol=1e-10, beta=1.0, picard_maxit=0,
          picard_beta=1.0, picard_rtol=1e-2, picard_atol=1e-2, cap=10000):
    """Oracle aasol on a Python fixed-point map gfun(x) -> G (numpy)."""
    x0 = _f64(x0).copy()
    n = x0.size

    def cb(Gp, xp, _ctx):
        x = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        g = np.asarray(gfun(x), dtype=np.float64)
        np.ctypeslib.as_array(Gp, shape=(n,))[:] = g

    cfun = GFIX(cb)
    o = OrAAOpts(m, maxit, picard_maxit, rtol, atol, beta, picard_beta, picard_rtol, picard_atol)
    sol = np.zeros(n); gx = np.zeros(n)
    hist = np.zeros(cap); ch = np.zeros(cap); an = np.zeros(cap)
    res = OrAAResult()
    lib().or_aasol(cfun, None, n, _ptr(x0), C.byref(o), _ptr(sol), _ptr(gx), _ptr(hist), cap,
                   _ptr(ch), _ptr(an), cap, C.byref(res))
    return dict(solution=sol, functionval=gx, history=hist[:res.nhist].copy(),
                condhist=ch[:res.nstats].copy(), alphanorm=an[:res.nstats].copy(),
                idid=bool(res.idid), errcode=res.errcode, k=res.k, nevals=res.nevals)
