"""Generate tests/golden/phi_golden.json: the pointwise PC-SAFT Helmholtz density
Phi(n) and all F*C partials dPhi/dn at seeded random weighted-density vectors.

Phi is restated symbolically (sympy) from the reference formulas
(src/models/DFT/FMT.jl:17-47, src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:110-211) and
differentiated symbolically; values are evaluated with 50 significant digits.
The partials pin the hand-derived analytic gradient in the oracle and in the HIP
kernel (the reference obtains them with Enzyme, models.jl:355-411).

Run:  python tests/golden/gen_phi_golden.py
"""
import json
import os
import random

import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
CORR1 = [(0.9105631445, -0.3084016918, -0.0906148351), (0.6361281449, 0.1860531159, 0.4527842806),
         (2.6861347891, -2.5030047259, 0.5962700728), (-26.547362491, 21.419793629, -1.7241829131),
         (97.759208784, -65.25588533, -4.1302112531), (-159.59154087, 83.318680481, 13.77663187),
         (91.297774084, -33.74692293, -8.6728470368)]
CORR2 = [(0.7240946941, -0.5755498075, 0.0976883116), (2.2382791861, 0.6995095521, -0.2557574982),
         (-4.0025849485, 3.892567339, -9.155856153), (-21.003576815, -17.215471648, 20.642075974),
         (26.855641363, 192.67226447, -38.804430052), (206.55133841, -161.82646165, 93.626774077),
         (-355.60235612, -165.20769346, -29.666905585)]

# (m, sigma[A], eps[K]) and the temperature of each case.
CASES = [
    dict(name="methane_1d", comps=[(1.0, 3.7039, 150.03)], T=150.0, nd=1),
    dict(name="methane_decane_1d", comps=[(1.0, 3.7039, 150.03), (4.6627, 3.8384, 243.87)], T=300.0, nd=1),
    dict(name="c1c2c3_3d", comps=[(1.0, 3.7039, 150.03), (1.6069, 3.5206, 191.42), (2.0020, 3.6184, 208.11)],
         T=200.0, nd=3),
]


def rat(x):
    return sp.Rational(repr(float(x)))


def I_lite(corr, mb, eta):
    m2 = (mb - 1) / mb
    m3 = m2 * (mb - 2) / mb
    return sum((rat(c[0]) + m2 * rat(c[1]) + m3 * rat(c[2])) * eta ** q for q, c in enumerate(corr))


def phi_expr(nsym, nc, nd, m, dd, sig, eps, T):
    F = 6 + nd
    pi = sp.pi
    n0 = n1 = n2 = n3 = 0
    nv1 = [0, 0, 0]
    nv2 = [0, 0, 0]
    for i in range(nc):
        nim = nsym[i][1] * m[i]
        n0 += nim / dd[i]
        n1 += sp.Rational(1, 2) * nim
        n2 += pi * nim * dd[i]
        for j in range(nd):
            nv1[j] += nsym[i][3 + j] * m[i] / dd[i]
            nv2[j] -= 2 * pi * nsym[i][3 + j] * m[i]
        n3 += nsym[i][2] * m[i]
    nv1v2 = sum(nv1[j] * nv2[j] for j in range(3))
    nv2sq = sum(nv2[j] ** 2 for j in range(3))
    dn = 1 - n3
    bkt = sp.log(dn) / (12 * pi * n3 ** 2) + 1 / (12 * pi * n3 * dn ** 2)
    phi_hs = -n0 * sp.log(dn) + (n1 * n2 - nv1v2) / dn + (n2 ** 3 / 3 - n2 * nv2sq) * bkt
    z3 = sum(m[i] * nsym[i][3 + nd] for i in range(nc)) / 8
    z2 = sum(m[i] * nsym[i][3 + nd] / dd[i] for i in range(nc)) / 8
    inv = 1 / (1 - z3)
    phi_hc = 0
    for i in range(nc):
        lam = nsym[i][4 + nd] / (2 * dd[i])
        y = inv + sp.Rational(3, 2) * dd[i] * z2 * inv ** 2 + sp.Rational(1, 2) * dd[i] ** 2 * z2 ** 2 * inv ** 3
        phi_hc += -nsym[i][0] * (m[i] - 1) * sp.log(y * lam / nsym[i][0])
    psi = rat(1.3862)
    factor = 3 / (4 * psi ** 3 * pi)
    rb = [nsym[i][5 + nd] * factor / dd[i] ** 3 for i in range(nc)]
    S = sum(rb)
    mb = sum(rb[i] * m[i] for i in range(nc)) / S
    eta = sum(m[i] * rb[i] * dd[i] ** 3 for i in range(nc)) * pi / 6
    A1 = sum(rb[i] * rb[j] * m[i] * m[j] * sig[i][j] ** 3 * eps[i][j] / T for i in range(nc) for j in range(nc))
    A2 = sum(rb[i] * rb[j] * m[i] * m[j] * sig[i][j] ** 3 * (eps[i][j] / T) ** 2 for i in range(nc) for j in range(nc))
    C1 = 1 + mb * (8 * eta - 2 * eta ** 2) / (1 - eta) ** 4 + \
        (1 - mb) * (20 * eta - 27 * eta ** 2 + 12 * eta ** 3 - 2 * eta ** 4) / ((1 - eta) * (2 - eta)) ** 2
    I1 = I_lite(CORR1, mb, eta)
    I2 = I_lite(CORR2, mb, eta)
    phi_disp = -2 * pi * I1 * A1 - pi * mb * I2 * A2 / C1
    return phi_hs + phi_hc + phi_disp


def d_of(sig, eps, T):
    import math
    return sig * (1 - 0.12 * math.exp(-3 * eps / T))


def main():
    rng = random.Random(20240611)
    out = {"source": "sympy restatement of FMT.jl:17-47 and PCSAFT.jl:110-211, 50 digits", "cases": []}
    NA = 6.02214076e23
    for case in CASES:
        nc = len(case["comps"])
        nd = case["nd"]
        F = 6 + nd
        T = case["T"]
        L = max(c[1] for c in case["comps"]) * 1e-10
        m = [c[0] for c in case["comps"]]
        sig1 = [c[1] * 1e-10 for c in case["comps"]]
        eps1 = [c[2] for c in case["comps"]]
        d = [d_of(sig1[i], eps1[i], T) for i in range(nc)]
        dd = [d[i] / L for i in range(nc)]
        sig = [[(sig1[i] + sig1[j]) / 2 / L for j in range(nc)] for i in range(nc)]
        eps = [[(eps1[i] * eps1[j]) ** 0.5 for j in range(nc)] for i in range(nc)]
        nsym = [[sp.Symbol(f"n_{i}_{f}") for f in range(F)] for i in range(nc)]
        expr = phi_expr(nsym, nc, nd, [rat(v) for v in m], [rat(v) for v in dd],
                        [[rat(v) for v in r] for r in sig], [[rat(v) for v in r] for r in eps], rat(T))
        grads = [[sp.diff(expr, nsym[i][f]) for f in range(F)] for i in range(nc)]
        pts = []
        for _ in range(6):
            # a dense-liquid-like uniform state, perturbed slot by slot
            pack = rng.uniform(0.15, 0.42)
            xs = [rng.uniform(0.2, 1.0) for _ in range(nc)]
            tot = sum(x * m[i] * dd[i] ** 3 for i, x in enumerate(xs)) * 3.141592653589793 / 6
            rho = [x * pack / tot for x in xs]
            vals = {}
            for i in range(nc):
                r = rho[i]
                base = [r, r * dd[i], r * 4 * 3.141592653589793 / 3 * (dd[i] / 2) ** 3] + [0.0] * nd + \
                       [r * 4 * 3.141592653589793 / 3 * dd[i] ** 3, r * 2 * dd[i],
                        r * 4 * 3.141592653589793 / 3 * (1.3862 * dd[i]) ** 3]
                for f in range(F):
                    if 3 <= f < 3 + nd:
                        v = rng.uniform(-0.15, 0.15) * r * dd[i] ** 2
                    else:
                        v = base[f] * (1 + rng.uniform(-0.25, 0.25))
                    vals[nsym[i][f]] = sp.Float(repr(v), 60)
            phi = sp.N(expr.subs(vals), 50)
            g = [[sp.N(grads[i][f].subs(vals), 50) for f in range(F)] for i in range(nc)]
            pts.append({"n": [[float(vals[nsym[i][f]]) for f in range(F)] for i in range(nc)],
                        "phi": str(phi), "dphi": [[str(g[i][f]) for f in range(F)] for i in range(nc)]})
        # dilute and ringing-like points: packing fraction between -3e-4 and 2e-3 (the
        # small-n3 series branch and just above it), the S slot up to 30 times what the
        # packing fraction implies, as next to a sharp wall in a spectral convolution
        rng2 = random.Random(20261204 + nc * 10 + nd)
        for pack, sfac in ((3e-4, 1.0), (8e-4, 30.0), (-3e-4, 10.0), (2e-3, 5.0)):
            xs = [rng2.uniform(0.2, 1.0) for _ in range(nc)]
            tot = sum(x * m[i] * dd[i] ** 3 for i, x in enumerate(xs)) * 3.141592653589793 / 6
            rho = [x * abs(pack) / tot for x in xs]
            vals = {}
            for i in range(nc):
                r = rho[i]
                base = [r, r * dd[i] * sfac, r * 4 * 3.141592653589793 / 3 * (dd[i] / 2) ** 3 * (1 if pack > 0 else -1)] + \
                       [0.0] * nd + [r * 4 * 3.141592653589793 / 3 * dd[i] ** 3, r * 2 * dd[i],
                                     r * 4 * 3.141592653589793 / 3 * (1.3862 * dd[i]) ** 3]
                for f in range(F):
                    if 3 <= f < 3 + nd:
                        v = rng2.uniform(-0.15, 0.15) * r * dd[i] ** 2 * sfac
                    else:
                        v = base[f] * (1 + rng2.uniform(-0.1, 0.1))
                    vals[nsym[i][f]] = sp.Float(repr(v), 60)
            phi = sp.N(expr.subs(vals), 50)
            g = [[sp.N(grads[i][f].subs(vals), 50) for f in range(F)] for i in range(nc)]
            pts.append({"n": [[float(vals[nsym[i][f]]) for f in range(F)] for i in range(nc)],
                        "phi": str(phi), "dphi": [[str(g[i][f]) for f in range(F)] for i in range(nc)]})
        out["cases"].append({"name": case["name"], "nc": nc, "nd": nd, "T": T, "L": L, "m": m,
                             "d": d, "sigma": [[(sig1[i] + sig1[j]) / 2 for j in range(nc)] for i in range(nc)],
                             "eps": [[(eps1[i] * eps1[j]) ** 0.5 for j in range(nc)] for i in range(nc)],
                             "points": pts})
        print(case["name"], "done")
    with open(os.path.join(HERE, "phi_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
