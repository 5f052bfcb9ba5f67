"""Dynamic DFT on the GPU: the right-hand side of the log-density form behind
cDFT.jl's ODEProblem(system, rho, tspan) (ext/SciMLBasecDFTExt.jl:10-61).

  ddft_rhs(system, eta)        d(eta)/dt, eta = ln rho, shape (nc, N) like rho
  ODEProblem(system, rho, tspan)
                               the problem the reference hands to SciML: f(t, y)
                               (y = ln rho flattened), y0 and tspan, ready for
                               scipy.integrate.solve_ivp(prob.f, prob.tspan, prob.y0)

The right-hand side runs in libcdft_hip.so (cdft_ddft_rhs): rho = exp(clamp(eta, -50,
30)), mu = dF_res/drho + beta V_ext, and per component
lap(eta) + lap(mu) + |grad eta|^2 + grad mu . grad eta with spectral derivatives over
the periodic box in reduced lengths (z / L), as the reference's convolve! does them.
"""
import numpy as np


def ddft_rhs(system, eta):
    """d(eta)/dt at eta (nc, N)."""
    eta = np.ascontiguousarray(eta, dtype=np.float64).reshape(-1)
    return system.device_system.ddft_rhs(eta).reshape(system.nc, system.N)


class ODEProblem:
    """ODEProblem(system, rho, tspan): f(t, y) = d(ln rho)/dt, y0 = ln rho."""

    def __init__(self, system, rho, tspan):
        self.system = system
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.y0 = np.log(np.ascontiguousarray(rho, dtype=np.float64).reshape(-1))
        self.nfev = 0

    def f(self, t, y):
        self.nfev += 1
        return self.system.device_system.ddft_rhs(np.ascontiguousarray(y, dtype=np.float64))

    def rho(self, y):
        """The density profile (nc, N) of a state y."""
        return np.exp(np.clip(np.asarray(y), -50.0, 30.0)).reshape(self.system.nc, self.system.N)
