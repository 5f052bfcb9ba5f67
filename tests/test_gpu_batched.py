"""Batched persistent path (cdft_converge_batch): same algorithm as the plain path,
checked against it and against the oracle; the mixed-precision mode against FP64.
Needs a GPU."""
import math
import os
import sys

import numpy as np
import pytest

import pyoracle as po
from cdft_amd import DFTOptions, _lib, configs, converge_
from cdft_amd.methods import adsorption_system, surface_tension_system
from helpers import bulk_mu, oracle_of

pytestmark = pytest.mark.gpu

# A launch of the register-state kernel (k_batched) running the FP64 map from the
# first evaluation faulted the GPU once (memory aperture violation, 6 slit pores + the
# methane wall).  A CPU emulation of the same launch under AddressSanitizer is clean
# and the cause is not yet shown, so that kernel's FP64-from-start case stays off the
# GPU.  FP64 batches now run on other code (k_small_batch, tested below); these tests
# were written for k_batched and stay skipped with it.
FP64_FAULT = pytest.mark.skip(reason="k_batched FP64-from-start launch faulted once; cause not yet shown")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def batch(systems, rhos, **kw):
    ctx = systems[0].device_system.ctx
    return _lib.converge_batch(ctx, [s.device_system for s in systems], [r.reshape(-1) for r in rhos], **kw)


@pytest.fixture(scope="module")
def pores():
    return configs.slit_pores(n=6, seed=0)


@FP64_FAULT
def test_batched_fp64_matches_plain(pores):
    say("batched FP64 vs plain FP64: 6 slit pores + methane wall, tol=1e-8")
    items = list(pores) + [configs.methane_wall()]
    systems = [s for s, _ in items]
    seeds = [r for _, r in items]
    rb = [r.copy() for r in seeds]
    res = batch(systems, rb, tol=1e-8)
    for s, r0, r_b, rr in zip(systems, seeds, rb, res):
        rp = r0.copy()
        converge_(s, rp, tol=1e-8)
        rp_res = s.last_result
        say(f"  N={s.N}: batched errcode {rr['errcode']} evals {rr['nevals']} | plain {rp_res['errcode']} "
            f"evals {rp_res['nevals']} | rel {rel(r_b, rp):.2e}")
        assert rr["errcode"] == rp_res["errcode"]
        assert abs(rr["nevals"] - rp_res["nevals"]) <= max(5, 0.1 * rp_res["nevals"])
        if rr["errcode"] == 0:
            e_ = rel(r_b, rp)
            assert e_ < 1e-5, e_


@FP64_FAULT
def test_batched_fp64_tight_matches_oracle(pores):
    say("batched FP64 tol=1e-12 vs oracle tol=1e-12 (2 pores)")
    items = list(pores)[2:4]
    systems = [s for s, _ in items]
    rb = [r.copy() for _, r in items]
    res = batch(systems, rb, tol=1e-12)
    for (s, r0), r_b, rr in zip(items, rb, res):
        rho_o, ro, _ = oracle_of(s).converge(r0.reshape(-1), tol=1e-12)
        assert rr["errcode"] == 0 and ro.errcode == 0
        e_ = rel(r_b.reshape(-1), rho_o)
        assert e_ < 1e-10, e_


@FP64_FAULT
def test_batched_mixed_precision(pores):
    """FP32 map to its noise floor, then FP64 to tol: profile and adsorption agree with FP64."""
    say("batched mixed vs batched FP64 (6 pores), tol=1e-8")
    systems64 = [s for s, _ in pores]
    seeds = [r for _, r in pores]
    r64 = [r.copy() for r in seeds]
    res64 = batch(systems64, r64, tol=1e-8)
    mixed = configs.slit_pores(n=6, seed=0, options=DFTOptions(precision="mixed"))
    sm = [s for s, _ in mixed]
    rm = [r.copy() for _, r in mixed]
    resm = batch(sm, rm, tol=1e-8)
    for s, a, b, ra, rb_ in zip(systems64, r64, rm, res64, resm):
        say(f"  N={s.N}: fp64 evals {ra['nevals']} mixed evals {rb_['nevals']} errcodes {ra['errcode']} {rb_['errcode']}")
        assert rb_["errcode"] == ra["errcode"]
        if ra["errcode"] == 0:
            assert rb_["resnorm"] <= rb_["tol"]
            e_ = rel(b, a)
            assert e_ < 1e-5, e_
            ads_a = adsorption_system(s, a)
            ads_b = adsorption_system(s, b)
            assert abs(ads_b[0] - ads_a[0]) <= 1e-6 * abs(ads_a[0])


@FP64_FAULT
def test_batched_binary_interface():
    say("batched FP64 binary interface N=256 vs plain")
    s, r0 = configs.c1c10_interface(ngrid=256, box=10.0)
    rb = r0.copy()
    rr = batch([s], [rb], tol=1e-8)[0]
    rp = r0.copy()
    converge_(s, rp, tol=1e-8)
    assert rr["errcode"] == 0 == s.last_result["errcode"]
    e_ = rel(rb, rp)
    assert e_ < 1e-5, e_


@FP64_FAULT
def test_batched_errcode_maxit(pores):
    say("batched maxit=2 -> errcode 10, equals plain")
    s, r0 = pores[0]
    rb = r0.copy()
    rr = batch([s], [rb], tol=1e-8, maxit=2)[0]
    rp = r0.copy()
    converge_(s, rp, tol=1e-8, maxit=2)
    assert rr["errcode"] == 10 == s.last_result["errcode"]
    assert rr["nevals"] == s.last_result["nevals"]
    e_ = rel(rb, rp)
    assert e_ < 1e-12, e_


def test_batched_mixed_matches_plain_fp64(pores):
    """Mixed batched solve vs the plain FP64 path: same termination, profiles and
    adsorption within the mixed-precision tolerance."""
    say("batched mixed (6 pores) vs plain FP64, tol=1e-8")
    mixed = configs.slit_pores(n=6, seed=0, options=DFTOptions(precision="mixed"))
    sm = [s for s, _ in mixed]
    rm = [r.copy() for _, r in mixed]
    resm = batch(sm, rm, tol=1e-8)
    for (s64, r0), s, b, rr in zip(pores, sm, rm, resm):
        rp = r0.copy()
        converge_(s64, rp, tol=1e-8)
        ra = s64.last_result
        say(f"  N={s.N}: mixed errcode {rr['errcode']} evals {rr['nevals']} (fp32 {rr['nevals_fp32']}) | "
            f"plain errcode {ra['errcode']} evals {ra['nevals']} | rel {rel(b, rp):.2e}")
        assert rr["errcode"] == ra["errcode"] == 0
        assert rr["resnorm"] <= rr["tol"]
        e_ = rel(b, rp)
        assert e_ < 1e-5, e_
        ads_a = adsorption_system(s64, rp)
        ads_b = adsorption_system(s, b)
        assert abs(ads_b[0] - ads_a[0]) <= 1e-6 * abs(ads_a[0])
        # excess grand potential per area (surface_tension(system, rho)), the tension of the pore walls
        g_a = surface_tension_system(s64, rp)
        g_b = surface_tension_system(s, b)
        say(f"    adsorption rel {abs(ads_b[0] - ads_a[0]) / abs(ads_a[0]):.1e}  tension rel {abs(g_b - g_a) / abs(g_a):.1e}")
        assert abs(g_b - g_a) <= 1e-6 * abs(g_a)


def test_batched_mixed_maxit(pores):
    """maxit bound: a mixed solve stopped early reports errcode 10."""
    say("batched mixed maxit=2 -> errcode 10")
    mixed = configs.slit_pores(n=2, seed=0, options=DFTOptions(precision="mixed"))
    rm = [r.copy() for _, r in mixed]
    res = batch([s for s, _ in mixed], rm, tol=1e-8, maxit=2)
    assert all(r["errcode"] == 10 for r in res)


# The two-component batched kernel (Anderson state in HBM) faulted the GPU with an
# illegal memory access on its first run (mixed precision, methane/n-decane interface,
# 256 points).  A CPU emulation of the launch under AddressSanitizer is clean so far and
# the cause is not yet shown, so the case stays off the GPU.
@pytest.mark.skip(reason="two-component batched launch faulted once (illegal memory access); cause not yet shown")
def test_batched_mixed_binary_interface():
    """Two components (Anderson state in HBM, k_batched<2, false>), mixed precision,
    against the plain FP64 path: same termination, profile and surface tension."""
    say("batched mixed binary interface N=256 vs plain FP64, tol=1e-8")
    s64, r0 = configs.c1c10_interface(ngrid=256, box=10.0)
    sm, _ = configs.c1c10_interface(ngrid=256, box=10.0, options=DFTOptions(precision="mixed"))
    rb = r0.copy()
    rr = batch([sm], [rb], tol=1e-8)[0]
    rp = r0.copy()
    converge_(s64, rp, tol=1e-8)
    ra = s64.last_result
    say(f"  mixed errcode {rr['errcode']} evals {rr['nevals']} (fp32 {rr['nevals_fp32']}) | plain errcode "
        f"{ra['errcode']} evals {ra['nevals']} | rel {rel(rb, rp):.2e}")
    assert rr["errcode"] == ra["errcode"] == 0
    assert rr["nevals_fp32"] > 0
    e_ = rel(rb, rp)
    assert e_ < 1e-5, e_
    g_a = surface_tension_system(s64, rp)
    g_b = surface_tension_system(sm, rb)
    say(f"  tension rel {abs(g_b - g_a) / abs(g_a):.1e}")
    assert abs(g_b - g_a) <= 1e-6 * abs(g_a)


def test_batched_refusals(pores):
    """What neither batched kernel takes raises before any launch: systems above the
    single-workgroup body's 512 points x components (in FP64 and in mixed precision:
    the register-state kernel's HBM-state and multi-component variants share code with
    a kernel that faulted the GPU once, cause not shown), batches mixing association
    variants, and batches mixing precisions."""
    big64, r64 = configs.c1c10_interface(ngrid=512, box=10.0)          # 1024 points x components
    with pytest.raises(_lib.CDFTError):
        batch([big64], [r64.copy()], tol=1e-8)
    bigm, rm = configs.c1c10_interface(ngrid=512, box=10.0, options=DFTOptions(precision="mixed"))
    with pytest.raises(_lib.CDFTError):
        batch([bigm], [rm.copy()], tol=1e-8)
    from test_gpu_small_body import _interface
    sa, ra = _interface(["ethanol", "hexane"], [0.5, 0.5], 298.15, 64)  # one association pair
    sb, rb = configs.c1c10_interface(ngrid=64, box=6.0)                 # none
    with pytest.raises(_lib.CDFTError):
        batch([sa, sb], [ra.copy(), rb.copy()], tol=1e-8)
    sm, rm2 = configs.c1c10_interface(ngrid=64, box=6.0, options=DFTOptions(precision="mixed"))
    with pytest.raises(_lib.CDFTError):
        batch([sm, sb], [rm2.copy(), rb.copy()], tol=1e-8)


# Pores of the 512-pore workload (seed 0) for the tight oracle comparison: smallest and
# largest grids (269: 80 points, 252 and 26: 478-479), lowest and highest pressures
# (343; 287, 18), near the saturation pressure (13, 112), slow to converge (107, 426),
# and one whose Anderson iterate blows up in FP32 as in FP64 (40: the residual turns
# NaN at the 211th evaluation).  Pore 166 blows up in FP64 but not along the FP32
# map's trajectory, so the mixed path converges it; the bench lists such pores.
TIGHT_PORES = [269, 252, 26, 343, 287, 18, 13, 112, 107, 426]
BLOWUP_PORES = [40]


def test_batched_mixed_matches_oracle_tight():
    """Mixed batched solve at tol 1e-12 (FP32 map to its noise floor, then the FP64
    map) against the FP64 oracle at tol 1e-12: same termination; profile within 1e-10,
    adsorption and surface tension within 1e-9 (cdft_adsorption / cdft_surface_tension
    of the batched profile against or_adsorption / or_surface_tension of the oracle's)."""
    idx = TIGHT_PORES + BLOWUP_PORES
    say(f"batched mixed vs FP64 oracle, tol=1e-12, pores {idx}")
    mixed = configs.slit_pores(n=512, seed=0, indices=idx, options=DFTOptions(precision="mixed"))
    sm = [s for s, _ in mixed]
    rm = [r.copy() for _, r in mixed]
    res = batch(sm, rm, tol=1e-12)
    ors = [oracle_of(s) for s in sm]
    outs, ores, _, _ = po.converge_batch(ors, [r for _, r in mixed], tol=1e-12,
                                         nthreads=min(8, os.cpu_count() or 1))
    for i, s, o, b, rr, x, ro in zip(idx, sm, ors, rm, res, outs, ores):
        say(f"  pore {i} N={s.N}: batched errcode {rr['errcode']} evals {rr['nevals']} (fp32 {rr['nevals_fp32']}) "
            f"res {rr['resnorm']:.2e} | oracle errcode {ro.errcode} evals {ro.nevals} res {ro.resnorm:.2e}")
        assert rr["errcode"] == ro.errcode, (i, rr["errcode"], ro.errcode)
        if i in BLOWUP_PORES:
            assert rr["errcode"] == 10 and math.isnan(rr["resnorm"]) and math.isnan(ro.resnorm)
            continue
        assert rr["errcode"] == 0 and rr["resnorm"] <= rr["tol"]
        e_ = rel(b.reshape(-1), x)
        ga = s.device_system.adsorption(b)[0]
        oa = o.adsorption(x)[0]
        mu = bulk_mu(s)
        p = s.structure.conditions[0]
        gt = s.device_system.surface_tension(b, mu, p)
        ot = o.surface_tension(x, mu, p)
        say(f"    profile {e_:.2e}  adsorption {abs(ga - oa) / abs(oa):.2e}  tension {abs(gt - ot) / abs(ot):.2e}")
        assert e_ < 1e-10, e_
        assert abs(ga - oa) <= 1e-9 * abs(oa)
        assert abs(gt - ot) <= 1e-9 * abs(ot)


# grids of 80-230 points; 88, 207 and 257 are pores on which the FP64 oracle's Anderson
# iterate blows up or stalls (reported, not compared)
SMALL_PORES = [269, 11, 403, 13, 18, 88, 207, 257]
# pores whose termination (converged or not within maxit) depends on rounding: the FP64
# oracle and the device part ways on them in either direction
CHAOTIC_PORES = {88, 207, 257}
MIXED_PORES = [269, 13, 0, 4, 252, 426]               # 80, 93, 335, 405, 478, 349 points


@pytest.mark.parametrize("threads", ["256", "512"])
@pytest.mark.parametrize("which", ["small", "mixed_sizes"])
def test_batched_kernels_match_oracle(threads, which, monkeypatch):
    """Both workgroup sizes of the batched kernel (256 threads, two points per thread;
    512 threads, one point per thread) against the FP64 oracle at tol 1e-12, on a
    batch of small grids and on one mixing small and large grids."""
    monkeypatch.setenv("CDFT_BATCH_THREADS", threads)
    idx = SMALL_PORES if which == "small" else MIXED_PORES
    mixed = configs.slit_pores(n=512, seed=0, indices=idx, options=DFTOptions(precision="mixed"))
    say(f"batched mixed ({threads} threads), {len(mixed)} pores ({which}), tol=1e-12")
    sm = [s for s, _ in mixed]
    rm = [r.copy() for _, r in mixed]
    res = batch(sm, rm, tol=1e-12)
    ors = [oracle_of(s) for s in sm]
    outs, ores, _, _ = po.converge_batch(ors, [r for _, r in mixed], tol=1e-12, nthreads=min(8, os.cpu_count() or 1))
    nok = 0
    for i, s, b, rr, x, ro in zip(idx, sm, rm, res, outs, ores):
        e_ = rel(b.reshape(-1), x) if rr["errcode"] == ro.errcode == 0 else float("nan")
        say(f"  pore {i} N={s.N}: batched {rr['errcode']} evals {rr['nevals']} | oracle {ro.errcode} evals {ro.nevals}"
            f" | profile {e_:.2e}")
        if rr["errcode"] == ro.errcode == 0:  # both converge: the same fixed point
            nok += 1
            assert e_ < 1e-10, (i, e_)
        elif i not in CHAOTIC_PORES:          # elsewhere the same termination
            assert rr["errcode"] == ro.errcode, (i, rr["errcode"], ro.errcode)
    assert nok >= 5


def test_batched_results_independent_of_earlier_launches(monkeypatch):
    """The same batch gives the same bits whatever ran before it: a 512-thread launch
    between two 256-thread ones leaves no trace (never-written Anderson history
    columns are selected out, not multiplied by zero)."""
    mixed = configs.slit_pores(n=512, seed=0, indices=MIXED_PORES, options=DFTOptions(precision="mixed"))
    sm = [s for s, _ in mixed]
    outs = []
    for threads in ("256", "512", "256"):
        monkeypatch.setenv("CDFT_BATCH_THREADS", threads)
        rm = [r.copy() for _, r in mixed]
        res = batch(sm, rm, tol=1e-8)
        outs.append(([r["nevals"] for r in res], np.concatenate([r.reshape(-1) for r in rm])))
    assert outs[0][0] == outs[2][0]
    assert np.array_equal(outs[0][1], outs[2][1], equal_nan=True)


def test_batched_residual_history():
    """opts.hist on the batched path: one residual norm per map evaluation for every
    system of the batch (AASol's history, anderson.jl:327-337), in input order.  The
    recording leaves the solve unchanged (same bits as without); each history ends at
    the reported residual; its first entry (the FP32 map at the seed) agrees with the
    FP64 oracle's first residual to FP32 accuracy, and its Picard phase follows the
    oracle's to the same accuracy."""
    mixed = configs.slit_pores(n=512, seed=0, indices=MIXED_PORES, options=DFTOptions(precision="mixed"))
    say(f"batched residual history, {len(mixed)} pores, tol=1e-8")
    sm = [s for s, _ in mixed]
    r0 = [r.copy() for _, r in mixed]
    res0 = batch(sm, r0, tol=1e-8)
    r1 = [r.copy() for _, r in mixed]
    res1 = batch(sm, r1, tol=1e-8, hist_cap=30000)
    for i, (s, a, b, x, y) in enumerate(zip(sm, r0, r1, res0, res1)):
        assert np.array_equal(a, b) and x["nevals"] == y["nevals"] and x["errcode"] == y["errcode"]
        h = y["history"]
        assert y["nhist"] == y["nevals"] == len(h)
        assert h[-1] == y["resnorm"]
        _, ro, ho = oracle_of(s).converge(mixed[i][1].reshape(-1), tol=1e-8, maxit=20, planned=True)
        k = min(10, len(ho))
        say(f"  pore {MIXED_PORES[i]} N={s.N}: {len(h)} entries, first {h[0]:.6e} (oracle {ho[0]:.6e})")
        np.testing.assert_allclose(h[:k], ho[:k], rtol=2e-5)
    # a small cap keeps the first entries
    r2 = [r.copy() for _, r in mixed]
    res2 = batch(sm, r2, tol=1e-8, hist_cap=5)
    for y, z in zip(res1, res2):
        assert z["nhist"] == z["nevals"] and np.array_equal(z["history"], y["history"][:5])


# ---- FP64 batches: the single-workgroup bodies, one workgroup per system (k_small_batch)

def _single(s, r0, **kw):
    r = r0.copy()
    converge_(s, r, **kw)
    return r, s.last_result


def test_fp64_batch_matches_single_solves(pores):
    """An FP64 batch of slit pores of 80 to 478 points and the methane wall gives, bit
    for bit, what cdft_converge gives for each system alone (the same bodies, chunked
    there by the host, run to the end here)."""
    sel = configs.slit_pores(n=512, seed=0, indices=[269, 252, 343, 13, 107, 18]) + [configs.methane_wall()]
    say(f"FP64 batch of {len(sel)} systems vs single solves, tol=1e-8")
    rb = [r.copy() for _, r in sel]
    res = batch([s for s, _ in sel], rb, tol=1e-8, hist_cap=30000)
    for (s, r0), b, rr in zip(sel, rb, res):
        r1, ra = _single(s, r0, tol=1e-8, history=30000)
        say(f"  N={s.N}: batch errcode {rr['errcode']} evals {rr['nevals']} {rr['device_ms']:.2f} ms | "
            f"single errcode {ra['errcode']} evals {ra['nevals']}")
        assert (rr["errcode"], rr["nevals"], rr["iters"]) == (ra["errcode"], ra["nevals"], ra["iters"])
        assert rr["nevals_fp32"] == 0 and rr["nhist"] == rr["nevals"]
        assert np.array_equal(b, r1)
        assert np.array_equal(rr["history"], ra["history"])


def test_fp64_binary_and_assoc_batches_match_single_solves():
    """Two-component FP64 batches, without association (methane/n-decane interfaces of
    100 and 128 points) and with one association pair (ethanol/hexane, 150 and 201
    points): bit for bit what cdft_converge gives for each alone."""
    from test_gpu_small_body import _interface
    batches = {
        "c1c10": [configs.c1c10_interface(ngrid=128, box=10.0), configs.c1c10_interface(ngrid=100, box=8.0)],
        "ethanol_hexane": [_interface(["ethanol", "hexane"], [0.5, 0.5], 298.15, 201),
                           _interface(["ethanol", "hexane"], [0.4, 0.6], 300.0, 150)],
    }
    for name, sel in batches.items():
        say(f"FP64 two-component batch ({name}) vs single solves, tol=1e-8, maxit=3000")
        rb = [r.copy() for _, r in sel]
        res = batch([s for s, _ in sel], rb, tol=1e-8, maxit=3000)
        for (s, r0), b, rr in zip(sel, rb, res):
            r1, ra = _single(s, r0, tol=1e-8, maxit=3000)
            say(f"  N={s.N} nc={s.nc}: batch errcode {rr['errcode']} evals {rr['nevals']} | "
                f"single errcode {ra['errcode']} evals {ra['nevals']}")
            assert (rr["errcode"], rr["nevals"], rr["iters"]) == (ra["errcode"], ra["nevals"], ra["iters"])
            assert np.array_equal(b, r1)


def test_fp64_batch_matches_oracle_tight():
    """The FP64 batch at tol 1e-12 against the FP64 oracle at tol 1e-12 on the tight
    pores: same termination; profile within 1e-10, adsorption and surface tension
    within 1e-9; the blow-up pore ends with a NaN residual on both."""
    idx = TIGHT_PORES + BLOWUP_PORES
    say(f"FP64 batch vs FP64 oracle, tol=1e-12, pores {idx}")
    sel = configs.slit_pores(n=512, seed=0, indices=idx)
    sm = [s for s, _ in sel]
    rb = [r.copy() for _, r in sel]
    res = batch(sm, rb, tol=1e-12)
    ors = [oracle_of(s) for s in sm]
    outs, ores, _, _ = po.converge_batch(ors, [r for _, r in sel], tol=1e-12, nthreads=min(8, os.cpu_count() or 1))
    for i, s, o, b, rr, x, ro in zip(idx, sm, ors, rb, res, outs, ores):
        say(f"  pore {i} N={s.N}: batch errcode {rr['errcode']} evals {rr['nevals']} res {rr['resnorm']:.2e} | "
            f"oracle errcode {ro.errcode} evals {ro.nevals} res {ro.resnorm:.2e}")
        assert rr["errcode"] == ro.errcode, (i, rr["errcode"], ro.errcode)
        if i in BLOWUP_PORES:
            assert rr["errcode"] == 10 and math.isnan(rr["resnorm"]) and math.isnan(ro.resnorm)
            continue
        e_ = rel(b.reshape(-1), x)
        ga, oa = s.device_system.adsorption(b)[0], o.adsorption(x)[0]
        mu, p = bulk_mu(s), s.structure.conditions[0]
        gt, ot = s.device_system.surface_tension(b, mu, p), o.surface_tension(x, mu, p)
        say(f"    profile {e_:.2e}  adsorption {abs(ga - oa) / abs(oa):.2e}  tension {abs(gt - ot) / abs(ot):.2e}")
        assert e_ < 1e-10, e_
        assert abs(ga - oa) <= 1e-9 * abs(oa)
        assert abs(gt - ot) <= 1e-9 * abs(ot)


def test_mixed_batches_on_single_workgroup_bodies_match_single_solves():
    """Mixed-precision batches the register-state kernel does not take run on the
    single-workgroup bodies (FP32 map to the switch, then FP64): two-component
    interfaces, and one-component pores with an Anderson depth of 6.  Bit for bit what
    cdft_converge gives for each system alone, FP32 evaluation counts included."""
    mixed = DFTOptions(precision="mixed")
    batches = {
        "c1c10": ([configs.c1c10_interface(ngrid=128, box=10.0, options=mixed),
                   configs.c1c10_interface(ngrid=100, box=8.0, options=mixed)], {}),
        "pores_m6": (configs.slit_pores(n=512, seed=0, indices=[269, 343, 18], options=mixed), {"m": 6}),
    }
    for name, (sel, kw) in batches.items():
        say(f"mixed batch ({name}) on single-workgroup bodies vs single solves, tol=1e-8")
        rb = [r.copy() for _, r in sel]
        res = batch([s for s, _ in sel], rb, tol=1e-8, maxit=3000, **kw)
        for (s, r0), b, rr in zip(sel, rb, res):
            r1, ra = _single(s, r0, tol=1e-8, maxit=3000, anderson_m=kw.get("m", 5))
            say(f"  N={s.N} nc={s.nc}: batch errcode {rr['errcode']} evals {rr['nevals']} (fp32 {rr['nevals_fp32']}) | "
                f"single errcode {ra['errcode']} evals {ra['nevals']} (fp32 {ra['nevals_fp32']})")
            assert (rr["errcode"], rr["nevals"], rr["nevals_fp32"]) == (ra["errcode"], ra["nevals"], ra["nevals_fp32"])
            assert rr["nevals_fp32"] > 0
            assert np.array_equal(b, r1)
