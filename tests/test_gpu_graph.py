"""The plain path's device-side Anderson solve replayed from a captured hipGraph
against the same bodies enqueued one launch at a time (CDFT_AA_GRAPH=0), at several
chunk sizes: the same kernels run on the same data, so profiles, residual histories
and counts must agree bit for bit.  Needs a GPU."""
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, configs, converge_
from helpers import oracle_of

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(monkeypatch, s, rho0, graph, chunk, **kw):
    monkeypatch.setenv("CDFT_AA_GRAPH", "1" if graph else "0")
    monkeypatch.setenv("CDFT_AA_CHUNK", str(chunk))
    rho = rho0.copy()
    converge_(s, rho, history=40000, **kw)
    return rho, dict(s.last_result)


def same(a, b):
    (ra, xa), (rb, xb) = a, b
    assert np.array_equal(ra, rb)
    for k in ("errcode", "idid", "iters", "nevals", "nevals_fp32", "nhist", "resnorm"):
        assert xa[k] == xb[k], (k, xa[k], xb[k])
    assert np.array_equal(xa["history"], xb["history"])


@pytest.mark.parametrize("kw", [dict(tol=1e-8), dict(tol=1e-8, maxit=3), dict(tol=1e-8, anderson_m=0, maxit=300),
                                dict(tol=1e-8, anderson_m=1), dict(tol=1e-10, anderson_m=7)])
def test_graph_matches_eager_wall(monkeypatch, kw):
    s, rho0 = configs.methane_wall()
    say(f"methane wall {kw}: eager, graph chunk 16, graph chunk 5")
    ref = run(monkeypatch, s, rho0, False, 16, **kw)
    say(f"  eager: errcode {ref[1]['errcode']} evals {ref[1]['nevals']}")
    same(ref, run(monkeypatch, s, rho0, True, 16, **kw))
    same(ref, run(monkeypatch, s, rho0, True, 5, **kw))
    same(ref, run(monkeypatch, s, rho0, True, 16, **kw))     # the cached graph, replayed again


def test_graph_wall_matches_oracle(monkeypatch):
    """The device-side control flow is the reference's: the same errcode and evaluation
    count as the oracle on short runs, where rounding cannot separate the paths."""
    s, rho0 = configs.methane_wall()
    o = oracle_of(s)
    for maxit in (1, 2, 5, 12):
        rho, r = run(monkeypatch, s, rho0, True, 16, tol=1e-8, maxit=maxit)
        rho_o, ro, hist_o = o.converge(rho0.reshape(-1), tol=1e-8, maxit=maxit)
        say(f"  maxit {maxit}: GPU errcode {r['errcode']} evals {r['nevals']} | oracle {ro.errcode} {ro.nevals}")
        assert r["errcode"] == ro.errcode and r["nevals"] == ro.nevals and r["iters"] == ro.k
        np.testing.assert_allclose(r["history"], hist_o, rtol=1e-8)
        # near the walls rho ~ 1e-44 .. 1e-20 = exp(x) with x ~ -100: the sum orders of the
        # convolutions move x by ~1e-9, i.e. rho by as much relative
        np.testing.assert_allclose(rho.reshape(-1), rho_o, rtol=1e-8)


def test_graph_matches_eager_3d_mixed(monkeypatch):
    kw = dict(n=16, box=5.0, nspheres=1, seed=2, T=150.0, x=(1.0,), components=("methane",))
    s, rho0 = configs.matrix_3d(options=DFTOptions(precision="mixed"), **kw)
    say("3D mixed precision 16^3: eager vs graph")
    ref = run(monkeypatch, s, rho0, False, 16, tol=1e-9)
    say(f"  eager: errcode {ref[1]['errcode']} evals {ref[1]['nevals']} (FP32 {ref[1]['nevals_fp32']})")
    assert ref[1]["errcode"] == 0 and ref[1]["nevals_fp32"] > 0
    same(ref, run(monkeypatch, s, rho0, True, 16, tol=1e-9))
    say("3D mixed precision with an FP32 overflow in the seed: eager vs graph")
    r = rho0.copy()
    r[0, 1234] = np.exp(99.5)
    ref = run(monkeypatch, s, r, False, 16, tol=1e-9)
    assert ref[1]["nevals_fp32"] == 1
    same(ref, run(monkeypatch, s, r, True, 3, tol=1e-9))


def test_graph_matches_eager_adjoint(monkeypatch):
    s, rho0 = configs.methane_wall()
    rho = rho0.copy()
    converge_(s, rho, tol=1e-12)
    g = np.random.default_rng(0).standard_normal(rho.size)
    out = []
    for graph in (False, True, True):
        monkeypatch.setenv("CDFT_AA_GRAPH", "1" if graph else "0")
        lam, res = s.device_system.adjoint_solve(rho.reshape(-1), g, tol=1e-11)
        out.append((lam, res))
    say(f"adjoint: errcode {out[0][1]['errcode']} evals {out[0][1]['nevals']}")
    assert out[0][1]["errcode"] == 0
    for lam, res in out[1:]:
        assert np.array_equal(lam, out[0][0])
        assert res["nevals"] == out[0][1]["nevals"]
