"""Bulk PC-SAFT equation of state (Gross & Sadowski, Ind. Eng. Chem. Res. 40 (2001) 1244).

Host-side replacement for the Clapeyron.jl calls the reference makes during set-up
and post-processing (Clapeyron is not vendored in the reference):

* hard-sphere diameter ``d``                  (called at PCSAFT.jl:68)
* ``VT_chemical_potential_res``               (PCSAFT.jl:69)
* ``VT_chemical_potential`` (BasicIdeal part)  (surface_tension.jl:45)
* ``volume`` (liquid/vapour/stable root)       (adsorption.jl:54, README example)
* ``bubble_pressure``                          (surface_tension.jl:16)
* ``crit_pure``                                (two_phase.jl:2)
* constants ``N_A``, ``k_B``, ``R``            (cDFT.jl:21)

Derivatives are taken by complex-step differentiation, which is exact to
rounding for these analytic expressions.
"""
import math

import numpy as np
from scipy.optimize import brentq

NA = 6.02214076e23
KB = 1.380649e-23
RGAS = NA * KB  # 8.31446261815324

# Gross & Sadowski (2001) Table 1: segment number m, sigma (Angstrom), eps/k (K), Mw (g/mol).
DATABASE = {
    "methane": (1.0000, 3.7039, 150.03, 16.043),
    "ethane": (1.6069, 3.5206, 191.42, 30.070),
    "propane": (2.0020, 3.6184, 208.11, 44.097),
    "n-butane": (2.3316, 3.7086, 222.88, 58.123),
    "n-hexane": (3.0576, 3.7983, 236.77, 86.177),
    "n-decane": (4.6627, 3.8384, 243.87, 142.285),
    "decane": (4.6627, 3.8384, 243.87, 142.285),
    "carbon dioxide": (2.0729, 2.7852, 169.21, 44.010),
    "nitrogen": (1.2053, 3.3130, 90.96, 28.014),
    "argon": (0.9285, 3.4784, 122.23, 39.948),
}

# Binary interaction parameters k_ij (Gross & Sadowski 2001), used when the caller
# gives none, like Clapeyron's unlike-parameter table.  With CO2/methane = 0.065 the
# reference's stored adsorption answer (test/test_methods.jl:37) is reproduced.
KIJ_DATABASE = {
    frozenset(("carbon dioxide", "methane")): 0.065,
}

# Dispersion constants a_qk, b_qk (G&S 2001 Table 1); identical to PCSAFT.jl:86-103.
_A = np.array([
    [0.9105631445, -0.3084016918, -0.0906148351],
    [0.6361281449, 0.1860531159, 0.4527842806],
    [2.6861347891, -2.5030047259, 0.5962700728],
    [-26.547362491, 21.419793629, -1.7241829131],
    [97.759208784, -65.25588533, -4.1302112531],
    [-159.59154087, 83.318680481, 13.77663187],
    [91.297774084, -33.74692293, -8.6728470368]])
_B = np.array([
    [0.7240946941, -0.5755498075, 0.0976883116],
    [2.2382791861, 0.6995095521, -0.2557574982],
    [-4.0025849485, 3.892567339, -9.155856153],
    [-21.003576815, -17.215471648, 20.642075974],
    [26.855641363, 192.67226447, -38.804430052],
    [206.55133841, -161.82646165, 93.626774077],
    [-355.60235612, -165.20769346, -29.666905585]])

_H = 1e-30  # complex-step size (relative)


class PCSAFT:
    """PC-SAFT model for non-associating components (one bead per component).

    ``sigma`` and ``epsilon`` are stored as full matrices with the Clapeyron
    combining rules: sigma_ij = (sigma_i + sigma_j)/2 and
    eps_ij = (1 - k_ij) sqrt(eps_i eps_j).
    """

    def __init__(self, components, kij=None, params=None):
        self.components = list(components)
        n = len(self.components)
        rows = []
        for c in self.components:
            if params is not None and c in params:
                rows.append(params[c])
            elif c in DATABASE:
                rows.append(DATABASE[c])
            else:
                raise KeyError(f"no PC-SAFT parameters for {c!r}")
        rows = np.array(rows, dtype=np.float64)
        self.segment = rows[:, 0].copy()
        s = rows[:, 1] * 1e-10
        e = rows[:, 2].copy()
        self.Mw = rows[:, 3].copy()
        if kij is None:
            k = np.zeros((n, n))
            for i in range(n):
                for j in range(n):
                    if i != j:
                        k[i, j] = KIJ_DATABASE.get(frozenset((self.components[i], self.components[j])), 0.0)
        else:
            k = np.asarray(kij, dtype=np.float64)
        self.sigma = 0.5 * (s[:, None] + s[None, :])
        self.epsilon = np.sqrt(e[:, None] * e[None, :]) * (1.0 - k)
        self.kij = k

    def __len__(self):
        return len(self.components)

    def length_scale(self):
        """L = max sigma (PCSAFT.jl:79-81)."""
        return float(np.max(self.sigma))

    def d(self, T):
        """Hard-sphere diameter d_i = sigma_i (1 - 0.12 exp(-3 eps_i / T))."""
        s = np.diag(self.sigma)
        e = np.diag(self.epsilon)
        return s * (1.0 - 0.12 * np.exp(-3.0 * e / T))

    # ---------------------------------------------------------------- a_res
    def _consts(self, T, x):
        """Composition-dependent constants of a_res (complex-safe in x)."""
        m = self.segment
        d = self.d(T)
        s = [math.pi / 6.0 * np.sum(x * m * d ** k) for k in range(4)]
        mbar = np.sum(x * m)
        xm = x * m
        s3 = self.sigma ** 3
        eT = self.epsilon / T
        m2es3 = np.sum(np.outer(xm, xm) * eT * s3)
        m2e2s3 = np.sum(np.outer(xm, xm) * eT * eT * s3)
        m2 = (mbar - 1.0) / mbar
        m3 = m2 * (mbar - 2.0) / mbar
        a = _A[:, 0] + m2 * _A[:, 1] + m3 * _A[:, 2]
        b = _B[:, 0] + m2 * _B[:, 1] + m3 * _B[:, 2]
        return s, mbar, m2es3, m2e2s3, a, b, x, m, d

    @staticmethod
    def _ares_rho(rho, c):
        """a_res per molecule over kT at number density rho (array- and complex-safe)."""
        s, mbar, m2es3, m2e2s3, a, b, x, m, d = c
        z0, z1, z2, z3 = (s[k] * rho for k in range(4))
        ahs = (3.0 * z1 * z2 / (1.0 - z3) + z2 ** 3 / (z3 * (1.0 - z3) ** 2)
               + (z2 ** 3 / z3 ** 2 - z0) * np.log(1.0 - z3)) / z0
        ahc = mbar * ahs
        for i in range(len(x)):
            gii = 1.0 / (1.0 - z3) + (d[i] / 2.0) * 3.0 * z2 / (1.0 - z3) ** 2 \
                + (d[i] / 2.0) ** 2 * 2.0 * z2 ** 2 / (1.0 - z3) ** 3
            ahc = ahc - x[i] * (m[i] - 1.0) * np.log(gii)
        eta = z3
        C1 = 1.0 / (1.0 + mbar * (8.0 * eta - 2.0 * eta ** 2) / (1.0 - eta) ** 4
                    + (1.0 - mbar) * (20.0 * eta - 27.0 * eta ** 2 + 12.0 * eta ** 3 - 2.0 * eta ** 4)
                    / ((1.0 - eta) * (2.0 - eta)) ** 2)
        I1 = 0.0
        I2 = 0.0
        pw = 1.0
        for q in range(7):
            I1 = I1 + a[q] * pw
            I2 = I2 + b[q] * pw
            pw = pw * eta
        adisp = -2.0 * math.pi * rho * I1 * m2es3 - math.pi * rho * mbar * C1 * I2 * m2e2s3
        return ahc + adisp

    def a_res(self, V, T, z):
        """Residual Helmholtz energy per mole over RT; V (m^3), z (mol).  Complex-safe."""
        z = np.asarray(z)
        ntot = z.sum()
        return self._ares_rho(NA * ntot / V, self._consts(T, z / ntot))

    # ------------------------------------------------------- derivatives
    def _p_rho(self, rho, T, c):
        """Pressure (Pa) at number density rho (array ok), composition constants c."""
        h = _H * rho
        da = self._ares_rho(rho + 1j * h, c).imag / h
        return rho * KB * T * (1.0 + rho * da)

    def pressure(self, V, T, z):
        z = np.asarray(z, dtype=np.float64)
        ntot = z.sum()
        return self._p_rho(NA * ntot / V, T, self._consts(T, z / ntot))

    def VT_chemical_potential_res(self, V, T, z):
        """mu_res_i at (V, T, z) in J/mol (Clapeyron naming)."""
        z = np.asarray(z, dtype=np.float64)
        out = np.zeros(len(z))
        for i in range(len(z)):
            zc = z.astype(np.complex128)
            h = _H * max(z.sum(), 1e-300)
            zc[i] += 1j * h
            out[i] = (zc.sum() * self.a_res(V, T, zc)).imag / h
        return RGAS * T * out

    def VT_chemical_potential(self, V, T, z):
        """Full chemical potential (J/mol) with the BasicIdeal reference state whose
        Helmholtz density is R T sum rho_i (ln rho_i - 1.5 ln T - 1), rho in mol/m^3,
        the same constant cDFT.jl's F_ideal uses (BasicIdeal.jl:34-39)."""
        z = np.asarray(z, dtype=np.float64)
        rho = z / V
        return self.VT_chemical_potential_res(V, T, z) + RGAS * T * (np.log(rho) - 1.5 * math.log(T))

    def lnphi(self, V, T, z):
        z = np.asarray(z, dtype=np.float64)
        p = self.pressure(V, T, z)
        Z = p * V / (RGAS * T * z.sum())
        return self.VT_chemical_potential_res(V, T, z) / (RGAS * T) - math.log(Z)

    # --------------------------------------------------------- volume
    def volume_roots(self, p, T, z):
        """All mechanically stable roots V (m^3) of p(V) = p, ascending in V."""
        z = np.asarray(z, dtype=np.float64)
        ntot = z.sum()
        c = self._consts(T, z / ntot)
        seg = c[0][3]  # (pi/6) sum x m d^3: eta = seg * rho
        etas = np.concatenate([np.logspace(-14, -2, 300), np.linspace(0.0101, 0.7405, 1200)])
        f = self._p_rho(etas / seg, T, c) - p
        roots = []
        g = lambda e: float(self._p_rho(np.array(e / seg), T, c)) - p
        for i in range(len(etas) - 1):
            if np.isfinite(f[i]) and np.isfinite(f[i + 1]) and f[i] * f[i + 1] < 0 and f[i + 1] > f[i]:
                e = brentq(g, etas[i], etas[i + 1], xtol=1e-17, rtol=1e-15, maxiter=500)
                roots.append(NA * ntot / (e / seg))
        return sorted(roots)

    def gibbs_res(self, V, p, T, z):
        z = np.asarray(z, dtype=np.float64)
        n = z.sum()
        x = z / n
        return self.a_res(V, T, z) + np.sum(x * np.log(z / V)) + p * V / (RGAS * T * n)

    def volume(self, p, T, z=(1.0,), phase="unknown"):
        """Molar volume root (Clapeyron.volume): phase 'liquid', 'vapour' or 'unknown' (stable)."""
        roots = self.volume_roots(p, T, z)
        if not roots:
            raise RuntimeError(f"no volume root at p={p}, T={T}")
        if phase in ("liquid", "l"):
            return roots[0]
        if phase in ("vapour", "vapor", "gas", "v"):
            return roots[-1]
        g = [self.gibbs_res(V, p, T, z) for V in roots]
        return roots[int(np.argmin(g))]

    # ------------------------------------------------------ phase equilibria
    def bubble_pressure(self, T, x, p0=None):
        """Bubble point at T and liquid mole fractions x -> (p, v_l, v_v, y)."""
        x = np.asarray(x, dtype=np.float64)
        x = x / x.sum()

        def inner(p, y0):
            y = y0.copy()
            for _ in range(500):
                vl = self.volume(p, T, x, "liquid")
                vv = self.volume(p, T, y, "vapour")
                K = np.exp(self.lnphi(vl, T, x) - self.lnphi(vv, T, y))
                S = np.sum(K * x)
                ynew = K * x / S
                if np.max(np.abs(ynew - y)) < 1e-15:
                    y = ynew
                    break
                y = ynew
            return S, y, vl, vv

        # pure-component start: highest volatility component enriched in vapour
        y = np.full(len(x), 1e-6)
        y[np.argmin(self.segment * np.diag(self.epsilon))] = 1.0
        y /= y.sum()
        p = 1e6 if p0 is None else p0
        S, y, vl, vv = inner(p, y)
        lp_prev, g_prev = math.log(p), S - 1.0
        p = p * S
        for _ in range(200):
            S, y, vl, vv = inner(p, y)
            g = S - 1.0
            if abs(g) < 1e-14:
                break
            lp = math.log(p)
            if g != g_prev:
                lp_new = lp - g * (lp - lp_prev) / (g - g_prev)
            else:
                lp_new = lp + math.log(S)
            lp_new = min(max(lp_new, lp - 0.5), lp + 0.5)
            lp_prev, g_prev = lp, g
            p = math.exp(lp_new)
        return p, vl, vv, y

    def saturation_pressure(self, T, p0=None):
        """Pure-component vapour pressure -> (p, v_l, v_v)."""
        assert len(self) == 1
        z = np.array([1.0])

        def g(lp):
            p = math.exp(lp)
            vl = self.volume(p, T, z, "liquid")
            vv = self.volume(p, T, z, "vapour")
            if vl == vv:
                return float("nan")
            return self.lnphi(vl, T, z)[0] - self.lnphi(vv, T, z)[0]

        lps = np.linspace(math.log(1e1), math.log(1e8), 200)
        vals = [g(q) for q in lps]
        for i in range(len(lps) - 1):
            if np.isfinite(vals[i]) and np.isfinite(vals[i + 1]) and vals[i] * vals[i + 1] < 0:
                lp = brentq(g, lps[i], lps[i + 1], xtol=1e-15)
                p = math.exp(lp)
                return p, self.volume(p, T, z, "liquid"), self.volume(p, T, z, "vapour")
        raise RuntimeError("no saturation point found")

    def crit_pure(self):
        """Critical point of a pure component -> (Tc, pc, vc)."""
        assert len(self) == 1
        z = np.array([1.0])

        def p_of(rho, T):
            return self.pressure(1.0 / rho, T, z)

        def derivs(rho, T):
            h = 1e-4 * rho
            f0, fp, fm = p_of(rho, T), p_of(rho + h, T), p_of(rho - h, T)
            fpp, fmm = p_of(rho + 2 * h, T), p_of(rho - 2 * h, T)
            d1 = (-fpp + 8 * fp - 8 * fm + fmm) / (12 * h)
            d2 = (-fpp + 16 * fp - 30 * f0 + 16 * fm - fmm) / (12 * h * h)
            return d1, d2

        # initial guess from corresponding states of PC-SAFT's segment parameters
        e = self.epsilon[0, 0]
        T = e * (1.0 + 0.45 * (self.segment[0] - 1.0) ** 0.6) * 1.3
        rho = 0.3 / (NA * math.pi / 6.0 * self.segment[0] * self.d(T)[0] ** 3) * 0.4
        for _ in range(100):
            d1, d2 = derivs(rho, T)
            hT, hr = 1e-5 * T, 1e-5 * rho
            d1T, d2T = derivs(rho, T + hT)
            d1r, d2r = derivs(rho + hr, T)
            J = np.array([[(d1r - d1) / hr, (d1T - d1) / hT], [(d2r - d2) / hr, (d2T - d2) / hT]])
            try:
                step = np.linalg.solve(J, -np.array([d1, d2]))
            except np.linalg.LinAlgError:
                break
            lim = 0.2
            step[0] = max(min(step[0], lim * rho), -lim * rho)
            step[1] = max(min(step[1], lim * T), -lim * T)
            rho += step[0]
            T += step[1]
            if abs(step[0]) < 1e-12 * rho and abs(step[1]) < 1e-12 * T:
                break
        return T, p_of(rho, T), 1.0 / rho

    def split_pure(self):
        out = []
        for i, c in enumerate(self.components):
            mdl = PCSAFT([c])
            mdl.segment = self.segment[i:i + 1].copy()
            mdl.sigma = self.sigma[i:i + 1, i:i + 1].copy()
            mdl.epsilon = np.array([[math.sqrt(self.epsilon[i, i] ** 2)]])
            out.append(mdl)
        return out
