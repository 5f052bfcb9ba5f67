"""ctypes binding of libcdft_hip.so (declarations: include/cdft_hip.h).

The product path has no CPU fallback: if the HIP library is missing, or no GPU is
visible, every compute call raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(_PKG, "libcdft_hip.so")

CDFT_OK = 0
CDFT_ERR_STOPPED = -5      # a progress callback stopped the solve
CDFT_FP64 = 0
CDFT_MIXED = 1
CDFT_CARTESIAN, CDFT_SPHERICAL, CDFT_CYLINDRICAL = 0, 1, 2
MAXC = 8
NOALLOC_MAX_SYS = 1024     # CDFT_NOALLOC_MAX_SYS

_dp = C.POINTER(C.c_double)

# every symbol include/cdft_hip.h declares
EXPORTS = [
    "cdft_ctx_create", "cdft_ctx_destroy", "cdft_last_error", "cdft_system_create",
    "cdft_system_destroy", "cdft_system_npoints", "cdft_functional_eval", "cdft_functional_eval_fp32",
    "cdft_converge",
    "cdft_converge_batch", "cdft_converge_batch_device", "cdft_converge_batch_device_noalloc",
    "cdft_last_kernel_ms",
    "cdft_surface_tension", "cdft_adsorption", "cdft_phi_vjp", "cdft_hvp", "cdft_adjoint_solve",
    "cdft_adjoint_params",
    "cdft_ddft_rhs", "cdft_system_tune", "cdft_system_get_tuning", "cdft_system_set_tuning",
]


class SystemDesc(C.Structure):
    _fields_ = [
        ("nc", C.c_int), ("ndim", C.c_int), ("ngrid", C.c_int * 3),
        ("lb", C.c_double * 3), ("ub", C.c_double * 3),
        ("T", C.c_double), ("p", C.c_double), ("L", C.c_double), ("psi", C.c_double),
        ("m", _dp), ("d", _dp), ("sigma", _dp), ("eps", _dp),
        ("rho_bulk", _dp), ("mu_res", _dp), ("vext", _dp), ("precision", C.c_int),
        ("assoc_nsite", C.c_int), ("assoc_npair", C.c_int),
        ("assoc_site_comp", C.POINTER(C.c_int)), ("assoc_site_count", _dp),
        ("assoc_pair_site", C.POINTER(C.c_int)), ("assoc_pair_eps", _dp), ("assoc_pair_kappa", _dp),
        ("coord", C.c_int),
    ]


# int (*)(void *user, int iter, const double *rho, long n)
ProgressFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, _dp, C.c_long)


class SolverOpts(C.Structure):
    _fields_ = [("maxit", C.c_int), ("beta", C.c_double), ("tol", C.c_double),
                ("anderson_start", C.c_double), ("m", C.c_int), ("hist_cap", C.c_int),
                ("hist", _dp), ("progress", ProgressFn), ("progress_user", C.c_void_p),
                ("progress_every", C.c_int)]


class Result(C.Structure):
    _fields_ = [("idid", C.c_int), ("errcode", C.c_int), ("iters", C.c_int),
                ("nevals", C.c_int), ("nhist", C.c_int), ("resnorm", C.c_double),
                ("tol", C.c_double), ("nevals_fp32", C.c_int), ("reserved", C.c_int),
                ("device_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("lines_per_tile", C.c_int * 2), ("fused_z", C.c_int * 2), ("ms", C.c_double * 2),
                ("tuned", C.c_int)]

    def as_dict(self):
        return {"lines_per_tile": list(self.lines_per_tile), "fused_z": list(self.fused_z),
                "ms": list(self.ms), "tuned": self.tuned}

    @classmethod
    def make(cls, lines_per_tile, fused_z, tuned=0):
        t = cls()
        t.lines_per_tile[0], t.lines_per_tile[1] = lines_per_tile
        t.fused_z[0], t.fused_z[1] = fused_z
        t.tuned = tuned
        return t


class CDFTError(RuntimeError):
    pass


_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise CDFTError(f"{LIBPATH} not found: build it with __graft_entry__.build() "
                        "(make -C cdft.jl_amd/csrc); there is no CPU fallback")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    L.cdft_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.cdft_ctx_destroy.argtypes = [vp]
    L.cdft_last_error.argtypes = [vp]
    L.cdft_last_error.restype = C.c_char_p
    L.cdft_system_create.argtypes = [vp, C.POINTER(SystemDesc), C.POINTER(vp)]
    L.cdft_system_destroy.argtypes = [vp]
    L.cdft_system_npoints.argtypes = [vp]
    L.cdft_functional_eval.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.cdft_functional_eval_fp32.argtypes = [vp, _dp, _dp]
    L.cdft_system_tune.argtypes = [vp, C.c_int, C.POINTER(Tuning)]
    L.cdft_system_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.cdft_system_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.cdft_converge.argtypes = [vp, _dp, C.POINTER(SolverOpts), C.POINTER(Result)]
    L.cdft_converge_batch.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(_dp),
                                      C.POINTER(SolverOpts), C.POINTER(Result)]
    L.cdft_converge_batch_device.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(_dp),
                                             C.POINTER(SolverOpts), C.POINTER(Result)]
    L.cdft_converge_batch_device_noalloc.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(_dp),
                                                     C.POINTER(SolverOpts), C.POINTER(Result)]
    L.cdft_last_kernel_ms.argtypes = [vp]
    L.cdft_last_kernel_ms.restype = C.c_double
    L.cdft_surface_tension.argtypes = [vp, _dp, _dp, C.c_double, _dp]
    L.cdft_adsorption.argtypes = [vp, _dp, _dp]
    L.cdft_phi_vjp.argtypes = [vp, _dp, _dp, _dp]
    L.cdft_hvp.argtypes = [vp, _dp, _dp, _dp]
    L.cdft_adjoint_solve.argtypes = [vp, _dp, _dp, _dp, C.POINTER(SolverOpts), C.POINTER(Result)]
    L.cdft_adjoint_params.argtypes = [vp, _dp, C.c_int, _dp, _dp, C.c_int, C.POINTER(vp), C.POINTER(vp), _dp,
                                      C.POINTER(SolverOpts), _dp, C.POINTER(Result)]
    L.cdft_ddft_rhs.argtypes = [vp, _dp, _dp]
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(_dp)


def check(rc, ctx=None):
    if rc != CDFT_OK:
        msg = load().cdft_last_error(ctx)
        raise CDFTError(f"cdft error {rc}: {msg.decode() if msg else ''}")


class Context:
    """One device + HIP stream (cdft_ctx)."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        check(L.cdft_ctx_create(int(device), C.byref(h)), None)
        self.handle = h
        self.device = device

    def __del__(self):
        try:
            if self.handle:
                load().cdft_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def last_kernel_ms(self):
        return load().cdft_last_kernel_ms(self.handle)


_default_ctx = {}


def default_context(device=0):
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class DeviceSystem:
    """A cdft_sys handle plus the host arrays its descriptor pointed to."""

    def __init__(self, ctx, *, nc, ndim, ngrid, lb, ub, T, p, L, psi, m, d, sigma, eps,
                 rho_bulk, mu_res, vext, precision, assoc_site_comp=(), assoc_site_count=(),
                 assoc_pairs=(), assoc_eps=(), assoc_kappa=(), coord=0):
        self.ctx = ctx
        keep = {k: np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                for k, v in dict(m=m, d=d, sigma=sigma, eps=eps, rho_bulk=rho_bulk,
                                 mu_res=mu_res).items()}
        desc = SystemDesc()
        desc.nc, desc.ndim = int(nc), int(ndim)
        for a in range(3):
            desc.ngrid[a] = int(ngrid[a]) if a < ndim else 1
            desc.lb[a] = float(lb[a]) if a < ndim else 0.0
            desc.ub[a] = float(ub[a]) if a < ndim else 0.0
        desc.T, desc.p, desc.L, desc.psi = float(T), float(p), float(L), float(psi)
        for k, v in keep.items():
            setattr(desc, k, ptr(v))
        if vext is not None:
            keep["vext"] = np.ascontiguousarray(vext, dtype=np.float64).reshape(-1)
            desc.vext = ptr(keep["vext"])
        desc.precision = int(precision)
        desc.coord = int(coord)
        if len(assoc_pairs):
            keep["site_comp"] = np.ascontiguousarray(assoc_site_comp, dtype=np.int32)
            keep["site_count"] = np.ascontiguousarray(assoc_site_count, dtype=np.float64)
            keep["pair_site"] = np.ascontiguousarray(assoc_pairs, dtype=np.int32).reshape(-1)
            keep["pair_eps"] = np.ascontiguousarray(assoc_eps, dtype=np.float64)
            keep["pair_kappa"] = np.ascontiguousarray(assoc_kappa, dtype=np.float64)
            ip = C.POINTER(C.c_int)
            desc.assoc_nsite, desc.assoc_npair = len(assoc_site_comp), len(assoc_pairs)
            desc.assoc_site_comp = keep["site_comp"].ctypes.data_as(ip)
            desc.assoc_site_count = ptr(keep["site_count"])
            desc.assoc_pair_site = keep["pair_site"].ctypes.data_as(ip)
            desc.assoc_pair_eps = ptr(keep["pair_eps"])
            desc.assoc_pair_kappa = ptr(keep["pair_kappa"])
        h = C.c_void_p()
        check(load().cdft_system_create(ctx.handle, C.byref(desc), C.byref(h)), ctx.handle)
        self.handle = h
        self.N = int(np.prod([int(ngrid[a]) for a in range(ndim)]))
        self.nc = int(nc)

    def __del__(self):
        try:
            if self.handle:
                load().cdft_system_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def functional_eval(self, rho, want_fval=False):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert rho.size == self.N * self.nc
        out = np.zeros(self.N * self.nc)
        fres = np.zeros(1)
        fval = np.zeros(self.N) if want_fval else None
        check(load().cdft_functional_eval(self.handle, ptr(rho), ptr(out), ptr(fres),
                                          ptr(fval) if want_fval else None), self.ctx.handle)
        return out, float(fres[0]), fval

    def tune(self, force=False):
        """Measure the line-FFT variants (2D/3D) and apply the fastest; returns the choice."""
        t = Tuning()
        check(load().cdft_system_tune(self.handle, int(bool(force)), C.byref(t)), self.ctx.handle)
        return t.as_dict()

    def get_tuning(self):
        t = Tuning()
        check(load().cdft_system_get_tuning(self.handle, C.byref(t)), self.ctx.handle)
        return t.as_dict()

    def set_tuning(self, lines_per_tile, fused_z):
        t = Tuning.make(lines_per_tile, fused_z)
        check(load().cdft_system_set_tuning(self.handle, C.byref(t)), self.ctx.handle)

    def functional_eval_fp32(self, rho):
        """delta F_res / delta rho from the FP32 map of a mixed-precision system."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        assert rho.size == self.N * self.nc
        out = np.zeros(self.N * self.nc)
        check(load().cdft_functional_eval_fp32(self.handle, ptr(rho), ptr(out)), self.ctx.handle)
        return out

    def converge(self, rho_inout, *, maxit=10000, beta=1e-2, tol=1e-4, anderson_start=1e-1,
                 m=5, hist_cap=0, progress=None, progress_every=0):
        """cdft_converge.  progress(iter, rho) (rho: a float64 copy of the profile the
        iter-th map evaluation used) is called after every progress_every-th evaluation;
        an exception it raises stops the solve and propagates."""
        assert rho_inout.dtype == np.float64 and rho_inout.flags.c_contiguous
        assert rho_inout.size == self.N * self.nc
        hist = np.zeros(max(hist_cap, 1))
        o = SolverOpts(int(maxit), float(beta), float(tol), float(anderson_start), int(m),
                       int(hist_cap), ptr(hist) if hist_cap else None)
        raised = []
        if progress is not None:
            if int(progress_every) < 1:
                raise ValueError("progress_every must be >= 1")

            def hook(_user, it, rp, n):
                try:
                    progress(int(it), np.ctypeslib.as_array(rp, shape=(int(n),)).copy())
                    return 0
                except BaseException as e:          # carried over the C frames, re-raised below
                    raised.append(e)
                    return 1
            o.progress = ProgressFn(hook)           # referenced by o for the whole call
            o.progress_every = int(progress_every)
        r = Result()
        rc = load().cdft_converge(self.handle, ptr(rho_inout), C.byref(o), C.byref(r))
        if raised:
            raise raised[0]
        check(rc, self.ctx.handle)
        out = r.as_dict()
        if hist_cap:
            out["history"] = hist[:min(r.nhist, hist_cap)].copy()
        return out

    def _vec(self, a, n=None):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == (self.N * self.nc if n is None else n)
        return a

    def surface_tension(self, rho, mu, p):
        rho = self._vec(rho)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        out = np.zeros(1)
        check(load().cdft_surface_tension(self.handle, ptr(rho), ptr(mu), float(p), ptr(out)), self.ctx.handle)
        return float(out[0])

    def adsorption(self, rho):
        rho = self._vec(rho)
        out = np.zeros(self.nc)
        check(load().cdft_adsorption(self.handle, ptr(rho), ptr(out)), self.ctx.handle)
        return out

    def phi_vjp(self, rho, cot):
        rho, cot = self._vec(rho), self._vec(cot, self.N)
        out = np.zeros(self.N * self.nc)
        check(load().cdft_phi_vjp(self.handle, ptr(rho), ptr(cot), ptr(out)), self.ctx.handle)
        return out

    def hvp(self, rho, v):
        rho, v = self._vec(rho), self._vec(v)
        out = np.zeros(self.N * self.nc)
        check(load().cdft_hvp(self.handle, ptr(rho), ptr(v), ptr(out)), self.ctx.handle)
        return out

    def ddft_rhs(self, eta):
        eta = self._vec(eta)
        out = np.zeros(self.N * self.nc)
        check(load().cdft_ddft_rhs(self.handle, ptr(eta), ptr(out)), self.ctx.handle)
        return out

    def adjoint_solve(self, rho, g, lam0=None, *, maxit=10000, beta=1e-2, tol=1e-10, anderson_start=1e-1, m=5):
        rho, g = self._vec(rho), self._vec(g)
        lam = np.zeros(self.N * self.nc) if lam0 is None else self._vec(lam0).copy()
        o = SolverOpts(int(maxit), float(beta), float(tol), float(anderson_start), int(m), 0, None)
        r = Result()
        check(load().cdft_adjoint_solve(self.handle, ptr(rho), ptr(g), ptr(lam), C.byref(o), C.byref(r)),
              self.ctx.handle)
        return lam, r.as_dict()

    def adjoint_params(self, rho, lam, g=None, pairs=(), h=(), *, maxit=10000, beta=1e-2, tol=1e-10,
                       anderson_start=1e-1, m=5):
        """cdft_adjoint_params for nobj = len(lam) objectives: with g (nobj rows) the
        multipliers are solved first, in place in lam (nobj x n float64); then for each
        (plus, minus) device-system pair and step h[p] the implicit sensitivity parts,
        shape (nobj, nparam).  Returns (out, results or None)."""
        rho = self._vec(rho)
        n = self.N * self.nc
        assert isinstance(lam, np.ndarray) and lam.dtype == np.float64 and lam.flags.c_contiguous
        nobj = lam.size // n
        assert nobj >= 1 and lam.size == nobj * n
        gp = None
        if g is not None:
            g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
            assert g.size == nobj * n
            gp = ptr(g)
        npar = len(pairs)
        hs = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        assert hs.size == npar
        plus = (C.c_void_p * max(npar, 1))(*[a.handle.value for a, _ in pairs])
        minus = (C.c_void_p * max(npar, 1))(*[b.handle.value for _, b in pairs])
        out = np.zeros((nobj, max(npar, 1)))
        res = (Result * nobj)()
        o = SolverOpts(int(maxit), float(beta), float(tol), float(anderson_start), int(m), 0, None)
        check(load().cdft_adjoint_params(self.handle, ptr(rho), nobj, gp, ptr(lam), npar, plus, minus,
                                         ptr(hs) if npar else None, C.byref(o), ptr(out), res), self.ctx.handle)
        return out[:, :npar], ([r.as_dict() for r in res] if g is not None else None)


def converge_batch(ctx, systems, rhos, *, maxit=10000, beta=1e-2, tol=1e-4, anderson_start=1e-1,
                   m=5, device_ptrs=False, noalloc=False, hist_cap=0):
    """Batched persistent solve (cdft_converge_batch[_device]): one-component
    mixed-precision 1D systems of at most 512 points with m <= 5 take the
    register-state kernel; other 1D Cartesian systems of one or two components with
    N nc <= 512 (FP64 or mixed, one precision and one association variant per batch)
    run the single-workgroup bodies; other batches raise CDFTError.

    rhos: list of float64 numpy arrays (in/out), or device pointers (ints) when
    device_ptrs is True.  noalloc (device pointers only): the allocation-free entry
    point cdft_converge_batch_device_noalloc.  hist_cap > 0: each result carries
    "history", the residual norm of each of its first hist_cap map evaluations."""
    if noalloc and not device_ptrs:
        raise ValueError("the allocation-free batched entry point takes device pointers")
    n = len(systems)
    hs = (C.c_void_p * n)(*[s.handle.value for s in systems])
    if device_ptrs:
        ps = (_dp * n)(*[C.cast(C.c_void_p(int(p)), _dp) for p in rhos])
    else:
        for r, s in zip(rhos, systems):
            assert r.dtype == np.float64 and r.flags.c_contiguous and r.size == s.N * s.nc
        ps = (_dp * n)(*[ptr(r) for r in rhos])
    hist = np.zeros(max(int(hist_cap), 1) * max(n, 1))
    o = SolverOpts(int(maxit), float(beta), float(tol), float(anderson_start), int(m), int(hist_cap),
                   ptr(hist) if hist_cap > 0 else None)
    res = (Result * n)()
    if noalloc:
        fn = load().cdft_converge_batch_device_noalloc
    else:
        fn = load().cdft_converge_batch_device if device_ptrs else load().cdft_converge_batch
    check(fn(ctx.handle, n, hs, ps, C.byref(o), res), ctx.handle)
    out = [r.as_dict() for r in res]
    if hist_cap > 0 and not noalloc:
        for i, d in enumerate(out):
            d["history"] = hist[i * hist_cap:i * hist_cap + min(d["nhist"], hist_cap)].copy()
    return out
