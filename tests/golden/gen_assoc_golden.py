"""Generate tests/golden/assoc_golden.json: the Wertheim association term Phi_assoc(n)
at one grid point and all F*C partials dPhi_assoc/dn, at seeded weighted-density
vectors, with 50 significant digits.

The term is restated from src/models/DFT/assoc.jl:159-221 (n0, xi, Delta) and
:702-831 (f_assoc).  The site fractions X are solved with mpmath's findroot on the
mass-action law; the partials are mpmath high-precision numerical derivatives of the
whole composition (mp.diff), so they pin the hand-derived gradient of the oracle and
of the HIP kernels without using it.

Run:  python tests/golden/gen_assoc_golden.py
"""
import json
import math
import os
import random

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 50

# (m, sigma[A], eps[K]) per component; sites: (component, count); pairs: (a, b, epsAB[K], kappa).
CASES = [
    dict(name="ethanol_hexane_1d", T=298.15, nd=1,
         comps=[(2.3827, 3.1771, 198.24), (3.0576, 3.7983, 236.77)],
         sites=[(0, 1), (0, 1)], pairs=[(0, 1, 2653.4, 0.032384)], rho=[9000.0, 3500.0]),
    dict(name="water_1d", T=350.0, nd=1,
         comps=[(1.0656, 3.0007, 366.51)],
         sites=[(0, 1), (0, 1)], pairs=[(0, 1, 2500.7, 0.034868)], rho=[50000.0]),
    dict(name="water_ethanol_cr1_3d", T=320.0, nd=3,
         comps=[(1.0656, 3.0007, 366.51), (2.3827, 3.1771, 198.24)],
         sites=[(0, 1), (0, 1), (1, 1), (1, 1)],
         pairs=[(0, 1, 2500.7, 0.034868), (2, 3, 2653.4, 0.032384),
                (0, 3, 0.5 * (2500.7 + 2653.4), None), (1, 2, 0.5 * (2500.7 + 2653.4), None)],
         rho=[20000.0, 8000.0]),
]


def d_of(sig, eps, T):
    return sig * (1.0 - 0.12 * math.exp(-3.0 * eps / T))


def setup(case):
    T = case["T"]
    comps = case["comps"]
    nc = len(comps)
    L = max(c[1] for c in comps) * 1e-10
    m = [c[0] for c in comps]
    d = [d_of(c[1], c[2], T) * 1e-10 for c in comps]
    sig = [[0.5 * (comps[i][1] + comps[j][1]) * 1e-10 for j in range(nc)] for i in range(nc)]
    eps = [[math.sqrt(comps[i][2] * comps[j][2]) for j in range(nc)] for i in range(nc)]
    pairs = []
    for a, b, e, k in case["pairs"]:
        if k is None:   # cr1 combining of the two like kappas
            ka, kb = case["pairs"][0][3], case["pairs"][1][3]
            si, sj = sig[0][0], sig[1][1]
            k = math.sqrt(ka * kb) * (math.sqrt(si * sj) / sig[0][1]) ** 3
        pairs.append((a, b, e, k))
    return dict(nc=nc, nd=case["nd"], T=T, L=L, m=m, d=d, sigma=[x for r in sig for x in r],
                eps=[x for r in eps for x in r], site_comp=[s[0] for s in case["sites"]],
                site_count=[float(s[1]) for s in case["sites"]], pairs=pairs)


def phi_assoc(nv, P):
    """Phi_assoc at the flat local field vector nv[(c*F + f)] (mpmath)."""
    nc, nd, F = P["nc"], P["nd"], 6 + P["nd"]
    L = mp.mpf(P["L"])
    dd = [mp.mpf(x) / L for x in P["d"]]
    m = [mp.mpf(x) for x in P["m"]]
    pi = mp.pi
    n2m = sum(pi * nv[i * F + 1] * m[i] * dd[i] for i in range(nc))
    n3m = sum(nv[i * F + 2] * m[i] for i in range(nc))
    nv2m = [sum(-2 * pi * nv[i * F + 3 + j] * m[i] for i in range(nc)) for j in range(nd)]
    xim = 1 - sum(x * x for x in nv2m) / (n2m * n2m)
    w = []
    for i in range(nc):
        S = nv[i * F + 1]
        n0 = S / dd[i]
        n2 = pi * S * dd[i]
        sq = sum((2 * pi * nv[i * F + 3 + j]) ** 2 for j in range(nd))
        w.append(n0 * (1 - sq / (n2 * n2)))
    a1 = 1 / (1 - n3m)
    ns = len(P["site_comp"])
    v = [mp.mpf(P["site_count"][t]) * w[P["site_comp"][t]] for t in range(ns)]
    K = [[mp.mpf(0)] * ns for _ in range(ns)]
    for a, b, e, k in P["pairs"]:
        ic, jc = P["site_comp"][a], P["site_comp"][b]
        sr = mp.mpf(P["sigma"][ic * nc + jc]) / L
        dij = dd[ic] * dd[jc] / (dd[ic] + dd[jc])
        g = a1 + dij * xim * n2m * a1 ** 2 / 2 + dij ** 2 * n2m ** 2 * xim * a1 ** 3 / 18
        delta = g * sr ** 3 * (mp.e ** (mp.mpf(e) / P["T"]) - 1) * mp.mpf(k)
        K[a][b] += delta * v[b]
        K[b][a] += delta * v[a]
    # float Newton start, then findroot to full precision
    X = [0.5] * ns
    for _ in range(200):
        X = [0.5 * (X[s] + 1.0 / (1.0 + sum(float(K[s][t]) * X[t] for t in range(ns)))) for s in range(ns)]
    eqs = lambda *x: [x[s] * (1 + sum(K[s][t] * x[t] for t in range(ns))) - 1 for s in range(ns)]
    X = mp.findroot(eqs, [mp.mpf(x) for x in X], tol=mp.mpf(10) ** -45)
    X = [X] if ns == 1 else [X[s] for s in range(ns)]
    return sum(v[t] * (mp.log(X[t]) - X[t] / 2 + mp.mpf(1) / 2) for t in range(ns))


def points(case, P, rng, npts=3):
    nc, nd, F = P["nc"], P["nd"], 6 + P["nd"]
    NA = 6.02214076e23 * P["L"] ** 3
    out = []
    for _ in range(npts):
        nv = [0.0] * (nc * F)
        for i in range(nc):
            r = case["rho"][i] * NA * rng.uniform(0.6, 1.2)
            di = P["d"][i] / P["L"]
            nv[i * F + 0] = r
            nv[i * F + 1] = r * di * rng.uniform(0.9, 1.1)
            nv[i * F + 2] = r * math.pi * di ** 3 / 6 * rng.uniform(0.9, 1.1)
            for j in range(nd):
                nv[i * F + 3 + j] = r * di * di / 2 * rng.uniform(-0.15, 0.15)
            nv[i * F + 3 + nd] = r * math.pi * di ** 3 * 4 / 3 * rng.uniform(0.9, 1.1)
            nv[i * F + 4 + nd] = 2 * r * di * rng.uniform(0.9, 1.1)
            nv[i * F + 5 + nd] = r * 4 * math.pi * (1.3862 * di) ** 3 / 3 * rng.uniform(0.9, 1.1)
        out.append(nv)
    return out


def main():
    rng = random.Random(20260412)
    cases = []
    for case in CASES:
        P = setup(case)
        F = 6 + P["nd"]
        pts = []
        for nv in points(case, P, rng):
            nm = [mp.mpf(repr(x)) for x in nv]
            val = phi_assoc(nm, P)
            grad = []
            for k in range(len(nm)):
                def f(h, k=k):
                    z = list(nm)
                    z[k] = z[k] + h
                    return phi_assoc(z, P)
                grad.append(mp.diff(f, 0))
            pts.append(dict(n=nv, phi=mp.nstr(val, 40), dphi=[mp.nstr(g, 40) for g in grad]))
        cases.append(dict(name=case["name"], nc=P["nc"], nd=P["nd"], T=P["T"], L=P["L"], m=P["m"], d=P["d"],
                          sigma=P["sigma"], eps=P["eps"], site_comp=P["site_comp"], site_count=P["site_count"],
                          pairs=[[a, b] for a, b, _, _ in P["pairs"]], pair_eps=[e for _, _, e, _ in P["pairs"]],
                          pair_kappa=[k for _, _, _, k in P["pairs"]], F=F, points=pts))
    with open(os.path.join(HERE, "assoc_golden.json"), "w") as fh:
        json.dump(dict(generator="tests/golden/gen_assoc_golden.py", dps=50, cases=cases), fh, indent=1)


if __name__ == "__main__":
    main()
