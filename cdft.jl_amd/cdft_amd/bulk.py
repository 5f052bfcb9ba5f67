"""Bulk PC-SAFT equation of state (Gross & Sadowski, Ind. Eng. Chem. Res. 40 (2001) 1244),
with the Wertheim association term of Gross & Sadowski, Ind. Eng. Chem. Res. 41 (2002) 5510.

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
    "hexane": (3.0576, 3.7983, 236.77, 86.177),
    # associating components (Gross & Sadowski 2002, Table 1; 2B scheme)
    "water": (1.0656, 3.0007, 366.51, 18.015),
    "methanol": (1.5255, 3.2300, 188.90, 32.042),
    "ethanol": (2.3827, 3.1771, 198.24, 46.069),
    "1-propanol": (2.9997, 3.2522, 233.40, 60.096),
}

# Association parameters (Gross & Sadowski 2002): eps_AB/k (K), bonding volume kappa_AB,
# and the site types with their multiplicities.  The 2B scheme has one electron-donor
# site "e" and one proton site "H", which bond to each other only.
ASSOC_DATABASE = {
    "water": (2500.7, 0.034868, (("e", 1), ("H", 1))),
    "methanol": (2899.5, 0.035176, (("e", 1), ("H", 1))),
    "ethanol": (2653.4, 0.032384, (("e", 1), ("H", 1))),
    "1-propanol": (2276.8, 0.015268, (("e", 1), ("H", 1))),
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
    """PC-SAFT model, one bead per component, with optional Wertheim association.

    ``sigma`` and ``epsilon`` are stored as full matrices with the Clapeyron
    combining rules: sigma_ij = (sigma_i + sigma_j)/2 and
    eps_ij = (1 - k_ij) sqrt(eps_i eps_j).

    Association follows Clapeyron's sparse layout, which the reference packs in
    ``pack_assoc_params`` (src/models/DFT/assoc.jl:18-57): a list of bonding pairs
    (i, a, j, b) between site a of component i and site b of component j, each with
    eps_AB (K) and kappa_AB.  Like sites bond as the database lists them (e with H for
    the 2B scheme).  Unlike pairs between two associating components exist only with
    ``combining="cr1"``: eps = (eps_i + eps_j)/2 and
    kappa = sqrt(kappa_i kappa_j) (sqrt(sigma_i sigma_j) / sigma_ij)^3; the default,
    ``"nocombining"``, adds none (Clapeyron's AssocOptions default).
    """

    def __init__(self, components, kij=None, params=None, assoc=None, combining="nocombining"):
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
        if combining not in ("nocombining", "cr1"):
            raise ValueError("combining must be 'nocombining' or 'cr1'")
        self.combining = combining
        self.assoc_like = {}
        for c in self.components:
            a = assoc.get(c) if assoc is not None and c in assoc else ASSOC_DATABASE.get(c)
            if a is not None:
                self.assoc_like[c] = (float(a[0]), float(a[1]), tuple((str(x), int(y)) for x, y in a[2]))
        self._build_assoc()

    def _build_assoc(self):
        """Site list and bonding pairs (the data pack_assoc_params flattens)."""
        self.sites = []                      # per component: ((name, count), ...)
        for c in self.components:
            self.sites.append(self.assoc_like[c][2] if c in self.assoc_like else ())
        self.site_comp, self.site_count, base = [], [], []
        for i, st in enumerate(self.sites):
            base.append(len(self.site_comp))
            for _, cnt in st:
                self.site_comp.append(i)
                self.site_count.append(float(cnt))
        pairs = []                           # (global site a, global site b, eps_AB, kappa_AB)

        def donor_acceptor(i, j):
            ia = [k for k, (nm, _) in enumerate(self.sites[i]) if nm == "e"]
            ih = [k for k, (nm, _) in enumerate(self.sites[i]) if nm == "H"]
            ja = [k for k, (nm, _) in enumerate(self.sites[j]) if nm == "e"]
            jh = [k for k, (nm, _) in enumerate(self.sites[j]) if nm == "H"]
            out = [(base[i] + a, base[j] + b) for a in ia for b in jh]
            if i != j:
                out += [(base[i] + a, base[j] + b) for a in ih for b in ja]
            return out

        names = self.components
        for i in range(len(names)):
            if names[i] not in self.assoc_like:
                continue
            eps_i, kap_i, _ = self.assoc_like[names[i]]
            for a, b in donor_acceptor(i, i):
                pairs.append((a, b, eps_i, kap_i))
        if self.combining == "cr1":
            for i in range(len(names)):
                for j in range(i + 1, len(names)):
                    if names[i] not in self.assoc_like or names[j] not in self.assoc_like:
                        continue
                    eps_i, kap_i, _ = self.assoc_like[names[i]]
                    eps_j, kap_j, _ = self.assoc_like[names[j]]
                    si, sj, sij = self.sigma[i, i], self.sigma[j, j], self.sigma[i, j]
                    eps = 0.5 * (eps_i + eps_j)
                    kap = math.sqrt(kap_i * kap_j) * (math.sqrt(si * sj) / sij) ** 3
                    for a, b in donor_acceptor(i, j):
                        pairs.append((a, b, eps, kap))
        self.assoc_pairs = pairs

    @property
    def has_assoc(self):
        return len(self.assoc_pairs) > 0

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
        assoc = None
        if self.has_assoc:
            sc = np.array(self.site_comp)
            pa = np.array([p[0] for p in self.assoc_pairs])
            pb = np.array([p[1] for p in self.assoc_pairs])
            ic, jc = sc[pa], sc[pb]
            strength = np.array([self.sigma[i, j] ** 3 * (math.exp(p[2] / T) - 1.0) * p[3]
                                 for p, i, j in zip(self.assoc_pairs, ic, jc)])
            dij = d[ic] * d[jc] / (d[ic] + d[jc])
            assoc = (pa, pb, strength, dij, x[sc] * np.array(self.site_count), np.array(self.site_count))
        return s, mbar, m2es3, m2e2s3, a, b, x, m, d, assoc

    @staticmethod
    def _ares_rho(rho, c):
        """a_res per molecule over kT at number density rho (array- and complex-safe)."""
        s, mbar, m2es3, m2e2s3, a, b, x, m, d, assoc = c
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
        if assoc is None:
            return ahc + adisp
        return ahc + adisp + _a_assoc(rho, z2, z3, assoc)

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
        if p0 is None:
            # lower the start until the trial vapour has a vapour-like root (Z > 0.3)
            while p > 1.0 and p * self.volume(p, T, y, "vapour") / (RGAS * T) < 0.3:
                p *= 0.1
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

    def tp_flash(self, p, T, z, equilibrium="lle"):
        """Two-phase flash at (p, T, z) by successive substitution on K = phi1/phi2 with
        the Rachford-Rice split (Clapeyron's tp_flash with RRTPFlash).  For
        ``equilibrium="lle"`` both phases take their liquid root and the start splits
        the feed into phases rich in its two most dissimilar components.
        Returns (x, beta, g): x[k] is phase k's composition, beta the fraction in phase 2."""
        z = np.asarray(z, dtype=np.float64)
        z = z / z.sum()
        nc = len(z)
        phase = "liquid" if equilibrium == "lle" else "unknown"
        x1 = np.full(nc, 1e-3)
        x2 = np.full(nc, 1e-3)
        e = self.segment * np.diag(self.epsilon)
        if equilibrium == "lle" and self.has_assoc:
            # the associating component forms one phase, the most apolar one the other
            a = next(i for i, c in enumerate(self.components) if c in self.assoc_like)
            b = int(np.argmax(np.where(np.arange(nc) == a, -np.inf, e)))
        else:
            a, b = int(np.argmax(e)), int(np.argmin(e))
        x1[a] = 1.0
        x2[b] = 1.0
        x1 /= x1.sum()
        x2 /= x2.sum()
        beta = 0.5
        for _ in range(2000):
            v1 = self.volume(p, T, x1, phase)
            v2 = self.volume(p, T, x2, phase if equilibrium == "lle" else "vapour")
            K = np.exp(self.lnphi(v1, T, x1) - self.lnphi(v2, T, x2))

            def rr(b_):
                return np.sum(z * (K - 1.0) / (1.0 + b_ * (K - 1.0)))
            lo = 1.0 / (1.0 - K.max()) + 1e-15
            hi = 1.0 / (1.0 - K.min()) - 1e-15
            lo, hi = max(lo, 0.0), min(hi, 1.0)
            beta = brentq(rr, lo, hi, xtol=1e-16, rtol=1e-15) if rr(lo) * rr(hi) < 0 else beta
            n1 = z / (1.0 + beta * (K - 1.0))
            n2 = K * n1
            n1 /= n1.sum()
            n2 /= n2.sum()
            done = np.max(np.abs(np.log(n1 / x1))) < 1e-13 and np.max(np.abs(np.log(n2 / x2))) < 1e-13
            x1, x2 = n1, n2
            if done:
                break
        g = (1.0 - beta) * self.gibbs_res(v1, p, T, x1) + beta * self.gibbs_res(v2, p, T, x2)
        return np.vstack([x1, x2]), beta, g

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
        """One pure model per component (Clapeyron.split_pure_model), with its own
        association sites."""
        out = []
        for i, c in enumerate(self.components):
            row = (self.segment[i], self.sigma[i, i] * 1e10, self.epsilon[i, i], self.Mw[i])
            assoc = {c: self.assoc_like[c]} if c in self.assoc_like else {c: None}
            out.append(PCSAFT([c], params={c: row}, assoc=assoc))
        return out


def _a_assoc(rho, z2, z3, assoc):
    """Wertheim association Helmholtz energy per molecule over kT at number density rho
    (array- and complex-safe), sum_i x_i sum_a n_ia (ln X_ia - X_ia/2 + 1/2), with the
    site fractions X from the mass-action law
    X_s = 1 / (1 + sum_t rho x_c(t) n_t Delta_st X_t) and
    Delta_p = g_ij sigma_ij^3 (exp(eps_AB/T) - 1) kappa_AB,
    g_ij = 1/(1-z3) + d_ij 3 z2/(1-z3)^2 + d_ij^2 2 z2^2/(1-z3)^3, d_ij = d_i d_j/(d_i + d_j).
    X is solved by successive substitution then Newton in the caller's number type, so a
    complex-step perturbation carries the exact derivative of X through the solve."""
    pa, pb, strength, dij, xn, cnt = assoc
    rho = np.asarray(rho)
    ns = len(xn)
    i1 = 1.0 / (1.0 - z3)
    shape = np.shape(rho)
    M = np.zeros(shape + (ns, ns), dtype=np.result_type(rho, z2, complex if np.iscomplexobj(rho) else float))
    for p in range(len(pa)):
        g = i1 + dij[p] * 3.0 * z2 * i1 ** 2 + dij[p] ** 2 * 2.0 * z2 ** 2 * i1 ** 3
        delta = g * strength[p]
        M[..., pa[p], pb[p]] += delta
        M[..., pb[p], pa[p]] += delta
    K = M * (np.asarray(rho)[..., None, None] * xn[None, :])       # K[s, t] = Delta_st rho x_c(t) n_t
    X = np.full(shape + (ns,), 0.9, dtype=M.dtype)
    for _ in range(10):
        X = 0.5 * (X + 1.0 / (1.0 + np.einsum("...st,...t->...s", K, X)))
    eye = np.eye(ns)
    for _ in range(60):
        C_ = np.einsum("...st,...t->...s", K, X)
        F = X * (1.0 + C_) - 1.0
        J = eye * (1.0 + C_)[..., None] + X[..., :, None] * K
        dX = np.linalg.solve(J, F[..., None])[..., 0]
        Xn = X - dX
        ok = (Xn.real > 0) & (Xn.real < 1)
        X = np.where(ok, Xn, 1.0 / (1.0 + C_))
        if np.all(ok) and np.all(np.abs(dX.real) <= 1e-15 * np.abs(X.real)) and \
                np.all(np.abs(dX.imag) <= 1e-15 * np.abs(X.imag) + 1e-300):
            break
    xs = xn / cnt                                                   # x of each site's component
    return np.sum(xs * cnt * (np.log(X) - 0.5 * X + 0.5), axis=-1)
