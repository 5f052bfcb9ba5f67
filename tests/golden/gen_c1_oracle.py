"""Writes c1_interface_oracle.npz: the oracle's solve of the methane / n-decane planar
interface (2048 points over [-20 L, 20 L], T = 300 K, liquid x = [0.5, 0.5] at its
bubble point, cos_prof seed) at tol 1e-12, maxit 20000, and its surface tension.

Test infrastructure: runs the CPU oracle (oracle/), never the product path.
    python tests/golden/gen_c1_oracle.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "cdft.jl_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import pyoracle as po  # noqa: E402
from cdft_amd import configs  # noqa: E402


def main():
    s, rho0 = configs.c1c10_interface()
    o = po.Oracle(**s.descriptor())
    x, r, hist = o.converge(rho0.reshape(-1), tol=1e-12, maxit=20000, planned=True)
    st = s.structure
    p, T = st.conditions
    rl = st.rho_bulk
    mu = s.model.VT_chemical_potential(1.0 / rl.sum(), T, rl / rl.sum())
    gamma = o.surface_tension(x, mu, p) / 2
    ads = o.adsorption(x)
    np.savez(os.path.join(HERE, "c1_interface_oracle.npz"), rho=x, errcode=r.errcode, nevals=r.nevals,
             resnorm=r.resnorm, tol=r.tol, gamma=gamma, mu=mu, adsorption=ads, seed=rho0.reshape(-1))
    print(f"errcode {r.errcode} nevals {r.nevals} resnorm {r.resnorm:.3e} gamma {gamma:.15g}")


if __name__ == "__main__":
    main()
