"""Benchmark: converged profiles per second (residual < 1e-8) of the batched
mixed-precision path on 512 methane slit pores, with the plain FP64 GPU path on the
same 512 pores one after another and the FP64 CPU oracle on all 512 pores with
OpenMP beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--npores 512] [--no-extras]

A step is one batched solve of all pores from their seed profiles (inputs resident
in HBM).  Prints one JSON line.  Besides the headline workload the line carries:
  plain_fp64   the plain FP64 path on the same pores, one after another;
  batched_fp64 the same FP64 pores in one launch, a workgroup per pore running the
               plain path's single-workgroup bodies (bit for bit the plain solves);
  termination  per-pore errcodes of the batched, plain and oracle solves, the pores
               where they differ, and the batched and plain results' largest
               relative differences from the oracle (profile, adsorption, tension);
  c0, c1       the plain FP64 path on the methane wall (201 points) and on the
               methane / n-decane interface (2048 points)
               (each with its mixed-precision solve beside the FP64 one);
  matrix_3d    the 3D hard-sphere-matrix ternary on a 256^3 grid: mixed-precision and
               FP64 solves, stream time per map evaluation, the algorithmic-byte rate,
               the weighted-density + transposed passes against the HBM roofline, and
               the adjoint's d(N_i)/d(eps_j) with one entry checked by full solves;
  matrix_2d    the same mixture with hard disks on a 512^2 grid, FP64 and mixed;
  radial       methane in spherical and cylindrical Steele pores (151 radial points);
  adjoint      surface-tension sensitivities of the methane / n-decane interface
               (2048 points) from FP64 and mixed solves versus finite differences of
               full solves.
(the extras are skipped with --no-extras).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cdft.jl_amd"))

METRIC = "converged profiles/sec (res<1e-8): batched FP32/FP64 vs plain FP64, same GPU"
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_batched_slit_pores_512_traffic.md")


def bytes_per_point_eval(s, m=5, F=7):
    """Compulsory bytes of one map evaluation per grid point per component: weighted
    densities N*C*(1+F)*s, Phi and gradient N*C*2F*s, transposed N*C*(F+1)*s,
    exponent/update N*C*(s+24), Anderson N*C*8*(2m+6)."""
    return (1 + F) * s + 2 * F * s + (F + 1) * s + (s + 24) + 8 * (2 * m + 6)


def measured_traffic():
    """HBM bytes per k_batched launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None when absent)."""
    try:
        txt = open(TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def host_threads():
    """CPU threads this process may use (the GPU box gives one GPU 16 of them)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def relmax(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def run_batched(args, np, torch, _lib, pores):
    systems = [s.device_system for s, _ in pores]
    ctx = systems[0].ctx
    sizes = [s.N * s.nc for s in systems]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    seeds_h = np.concatenate([r.reshape(-1) for _, r in pores])
    seeds = torch.tensor(seeds_h, dtype=torch.float64, device="cuda")
    work = torch.empty_like(seeds)
    base = work.data_ptr()
    ptrs = [base + 8 * int(o) for o in offs[:-1]]

    def step():
        work.copy_(seeds)
        torch.cuda.synchronize()
        return _lib.converge_batch(ctx, systems, ptrs, tol=args.tol, device_ptrs=True)

    if args.copy_probe_mib:
        # a device-to-device copy of known size, the reference for the profiler's
        # FETCH_SIZE / WRITE_SIZE scale
        a = torch.empty(args.copy_probe_mib * 131072, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kern_ms = 0.0
    results = None
    for _ in range(args.steps):
        results = step()
        kern_ms += ctx.last_kernel_ms()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return systems, offs, work.cpu().numpy(), results, dt, kern_ms


def run_plain(args, pores):
    """The plain FP64 path on every pore, one after another (set-up outside the timing)."""
    from cdft_amd import converge_
    out, res = [], []
    t = 0.0
    for s, r0 in pores:
        s.device_system
        rho = r0.copy()
        t1 = time.perf_counter()
        converge_(s, rho, tol=args.tol)
        t += time.perf_counter() - t1
        out.append(rho.reshape(-1))
        res.append(s.last_result)
    return out, res, t


def run_batched_fp64(args, np, torch, _lib, pores, plain_profiles, plain_results):
    """The FP64 pores in one batched launch (one workgroup per pore running the plain
    path's single-workgroup bodies), inputs resident on the device, one warm-up; and
    how many pores give exactly the plain path's bits and counts."""
    systems = [s.device_system for s, _ in pores]
    ctx = systems[0].ctx
    offs = np.concatenate([[0], np.cumsum([s.N * s.nc for s in systems])])
    seeds = torch.tensor(np.concatenate([r.reshape(-1) for _, r in pores]), dtype=torch.float64, device="cuda")
    work = torch.empty_like(seeds)
    ptrs = [work.data_ptr() + 8 * int(o) for o in offs[:-1]]
    res = None
    for it in range(2):
        work.copy_(seeds)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = _lib.converge_batch(ctx, systems, ptrs, tol=args.tol, device_ptrs=True)
        dt = time.perf_counter() - t0
    prof = work.cpu().numpy()
    same = sum(1 for i in range(len(systems))
               if np.array_equal(prof[offs[i]:offs[i + 1]], plain_profiles[i])
               and res[i]["nevals"] == plain_results[i]["nevals"] and res[i]["errcode"] == plain_results[i]["errcode"])
    conv = sum(1 for r in res if r["errcode"] == 0)
    dev = np.array([r["device_ms"] for r in res])
    return {"value": conv / dt, "unit": "profiles/s", "seconds": dt, "kernel_ms": ctx.last_kernel_ms(),
            "converged": conv, "identical_to_plain_fp64": same, "slowest_pore_ms": float(dev.max()),
            "slowest_pore": int(dev.argmax()), "p50_pore_ms": float(np.percentile(dev, 50)),
            "sample": f"all {len(systems)} pores in one launch (k_small_batch: a workgroup per pore, "
                      "the plain path's single-workgroup bodies), FP64"}


def run_cpu_baseline(args, np, pores, mus):
    """cpu_baseline leg: the FP64 oracle (planned 1D evaluation) on all pores with
    OpenMP over pores, and on one thread over a sample; plus the oracle's adsorption
    and tension of each converged profile for the accuracy report."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    ors = [po.Oracle(**s.descriptor()) for s, _ in pores]
    seeds = [r for _, r in pores]
    nthr = host_threads()
    outs, res, secs, wall = po.converge_batch(ors, seeds, tol=args.tol, nthreads=nthr)
    sidx = np.linspace(0, len(pores) - 1, min(args.cpu_1t_sample, len(pores))).round().astype(int)
    _, res1, _, wall1 = po.converge_batch([ors[i] for i in sidx], [seeds[i] for i in sidx], tol=args.tol,
                                          nthreads=1)
    ads, ten = [], []
    for (s, _), o, x, r, mu in zip(pores, ors, outs, res, mus):
        if r.errcode == 0:
            ads.append(o.adsorption(x)[0])
            ten.append(o.surface_tension_planned(x, mu, s.structure.conditions[0]))
        else:
            ads.append(None)
            ten.append(None)
    return dict(profiles=outs, errcodes=[r.errcode for r in res], nevals=[r.nevals for r in res], wall=wall,
                threads=nthr, secs=secs, wall1=wall1, n1=len(sidx), conv1=sum(r.errcode == 0 for r in res1),
                adsorption=ads, tension=ten)


def cpu_oracle_solve(s, rho0, tol, maxit=10000):
    """cpu_baseline leg of C0/C1: the FP64 C oracle's planned 1D solve of the same
    system from the same seed, on one host thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.Oracle(**s.descriptor())
    t = time.perf_counter()
    x, res, _ = o.converge(rho0, tol=tol, maxit=maxit, planned=not s.structure.coord)
    return o, x, res, time.perf_counter() - t


def run_c0(np, configs, converge_, nsolve=20):
    s, rho0 = configs.methane_wall()
    s.device_system
    rho = rho0.copy()
    converge_(s, rho, tol=1e-8)
    t = time.perf_counter()
    for _ in range(nsolve):
        rho = rho0.copy()
        converge_(s, rho, tol=1e-8)
    dt = (time.perf_counter() - t) / nsolve
    r = s.last_result
    o, x, ro, secs = cpu_oracle_solve(s, rho0, 1e-8)
    from cdft_amd import DFTOptions
    sm, _ = configs.methane_wall(options=DFTOptions(precision="mixed"))
    xm = rho0.copy()
    converge_(sm, xm, tol=1e-8)
    t = time.perf_counter()
    for _ in range(nsolve):
        xm = rho0.copy()
        converge_(sm, xm, tol=1e-8)
    dtm = (time.perf_counter() - t) / nsolve
    rm = sm.last_result
    return {"workload": "methane at 150 K, 1e7 Pa, graphite Steele walls 5 sigma apart, 201 points, tol 1e-8; FP64",
            "errcode": r["errcode"], "nevals": r["nevals"], "solve_ms": dt * 1e3, "profiles_per_s": 1.0 / dt,
            "stream_ms": r["device_ms"],
            "cpu_baseline": {"value": 1.0 / secs, "unit": "profiles/s", "cores": 1, "kind": "port",
                             "sample": "the same solve, FP64 C oracle (planned 1D evaluation), 1 thread",
                             "errcode": ro.errcode, "nevals": ro.nevals},
            "max_rel_err_vs_oracle": {"profile": relmax(rho.reshape(-1), x),
                                      "adsorption": relmax(np.asarray(s.device_system.adsorption(rho)),
                                                           o.adsorption(x))},
            "mixed": {"errcode": rm["errcode"], "nevals": rm["nevals"], "nevals_fp32": rm["nevals_fp32"],
                      "solve_ms": dtm * 1e3, "profile_max_rel_err_vs_oracle": relmax(xm.reshape(-1), x)}}


def run_c1(np, configs, converge_):
    from cdft_amd.methods import bulk_chemical_potential as bulk_mu
    s, rho0 = configs.c1c10_interface()
    s.device_system
    rho = rho0.copy()
    t = time.perf_counter()
    converge_(s, rho, tol=1e-8, maxit=20000)
    dt = time.perf_counter() - t
    r = s.last_result
    p = s.structure.conditions[0]
    gamma = s.device_system.surface_tension(rho, bulk_mu(s), p) / 2
    o, x, ro, secs = cpu_oracle_solve(s, rho0, 1e-8, maxit=20000)
    gamma_o = o.surface_tension(x, bulk_mu(s), p) / 2
    from cdft_amd import DFTOptions
    sm, _ = configs.c1c10_interface(options=DFTOptions(precision="mixed"))
    xm = rho0.copy()
    t = time.perf_counter()
    converge_(sm, xm, tol=1e-8, maxit=20000)
    dtm = time.perf_counter() - t
    rm = sm.last_result
    gamma_m = sm.device_system.surface_tension(xm, bulk_mu(sm), p) / 2
    return {"workload": "methane/n-decane interface, x=[0.5,0.5] liquid at its bubble point, T=300 K, 2048 points "
                        "over [-20 L, 20 L], tol 1e-8; FP64",
            "errcode": r["errcode"], "nevals": r["nevals"], "solve_s": dt, "profiles_per_s": 1.0 / dt,
            "stream_ms": r["device_ms"], "surface_tension_N_per_m": gamma,
            "cpu_baseline": {"value": 1.0 / secs, "unit": "profiles/s", "cores": 1, "kind": "port",
                             "sample": "the same solve, FP64 C oracle (planned 1D evaluation), 1 thread",
                             "errcode": ro.errcode, "nevals": ro.nevals},
            "max_rel_err_vs_oracle": {"surface_tension": abs(gamma / gamma_o - 1),
                                      "profile_unaligned": relmax(rho.reshape(-1), x),
                                      "note": "a free interface's position is a zero mode: at tol 1e-8 the two "
                                              "solves stop with the interfaces shifted by a fraction of a grid "
                                              "step, which dominates the unaligned profile difference"},
            "mixed": {"errcode": rm["errcode"], "nevals": rm["nevals"], "nevals_fp32": rm["nevals_fp32"],
                      "solve_s": dtm, "surface_tension_rel_err_vs_oracle": abs(gamma_m / gamma_o - 1)}}


def run_matrix_3d(np, DFTOptions, configs, converge_):
    from cdft_amd.methods import adsorption_system
    # the FP64 solve first: the reference for the mixed-precision result
    s64, rho64 = configs.matrix_3d(n=256)
    s64.device_system.tune()                     # the autotuner's measurement stays out of the timing
    t = time.perf_counter()
    converge_(s64, rho64, tol=1e-8, **configs.MATRIX_3D_SOLVER)
    wall64 = time.perf_counter() - t
    res64 = s64.last_result
    ads64 = adsorption_system(s64, rho64)
    # the workload's sensitivity d(N_i)/d(eps_j) through the adjoint (three adjoint
    # solves, one per component), and one entry against central differences of full
    # FP64 solves (h = 1e-5 relative)
    from cdft_amd import adjoint as adj
    t = time.perf_counter()
    jac, _, jres = adj.jacobian(s64, rho64.reshape(-1), "adsorption", [0, 1, 2], params=("epsilon",), tol=1e-10)
    t_jac = time.perf_counter() - t
    J = [[jac[(c, "epsilon", j)] for j in range(3)] for c in range(3)]
    h = 1e-5 * adj.model_params(s64.model)[s64.model.components[0]][2]
    fd_ads = []
    t = time.perf_counter()
    for sg in (+1, -1):
        sp = adj.perturbed_system(s64, 0, "epsilon", sg * h)
        rr = rho64.copy()
        converge_(sp, rr, tol=1e-10, **configs.MATRIX_3D_SOLVER)
        fd_ads.append(adsorption_system(sp, rr)[0])
        del sp
    t_fd = time.perf_counter() - t
    fd00 = (fd_ads[0] - fd_ads[1]) / (2 * h)
    adjoint = {"dN_deps_mol_m3_per_K": J, "adjoint_errcodes": [int(jres[c]["errcode"]) for c in range(3)],
               "adjoint_nevals": [int(jres[c]["nevals"]) for c in range(3)], "seconds": t_jac,
               "fd_dN0_deps0": fd00, "rel_diff_vs_fd": abs(J[0][0] - fd00) / abs(fd00),
               "fd_two_full_solves_s": t_fd,
               "note": "FP64 adjoint at the FP64 profile; explicit partials by central differences of one "
                       "functional evaluation each; FD of full solves from the converged profile at tol 1e-10"}
    del s64
    s, rho = configs.matrix_3d(n=256, options=DFTOptions(precision="mixed"))
    s.device_system.tune()
    r = rho.copy()
    t = time.perf_counter()
    converge_(s, r, tol=1e-8, **configs.MATRIX_3D_SOLVER)
    wall = time.perf_counter() - t
    res = s.last_result
    n32, n64 = res["nevals_fp32"], res["nevals"] - res["nevals_fp32"]
    F = 9
    alg = s.N * s.nc * (n32 * bytes_per_point_eval(4, F=F) + n64 * bytes_per_point_eval(8, F=F))
    ms = res["device_ms"]
    ads = adsorption_system(s, r)
    tune = s.device_system.get_tuning()
    # compulsory bytes of one weighted-density and one transposed-convolution pass:
    # B_WD + B_T = 2 N C (1 + F) s
    wd = {}
    for pr, (name, sz) in enumerate((("fp64", 8), ("fp32", 4))):
        b = 2.0 * s.N * s.nc * (1 + F) * sz
        tms = tune["ms"][pr]
        wd[name] = {"ms": tms, "compulsory_bytes": b,
                    "frac_of_hbm_peak": (b / (tms * 1e-3) / 1e9 / HBM_PEAK_GBS) if tms > 0 else None,
                    "lines_per_tile": tune["lines_per_tile"][pr], "fused_z": tune["fused_z"][pr]}
    return {
        "workload": "methane/ethane/propane x=[0.6,0.3,0.1], T=200 K, p=1e7 Pa, 30 sigma periodic cube on "
                    "256^3 points, 64 hard spheres r~U[1,2.5] sigma (seed 0); mixed precision, tol 1e-8, "
                    "anderson_start 0.5",
        "errcode": res["errcode"], "nevals": res["nevals"], "nevals_fp32": n32,
        "solve_s": wall, "stream_ms": ms, "ms_per_eval": ms / max(res["nevals"], 1),
        "algorithmic_GBps": alg / (ms * 1e-3) / 1e9,
        "adsorption_mol_m3": [float(v) for v in ads],
        "fp64": {"errcode": res64["errcode"], "nevals": res64["nevals"], "solve_s": wall64,
                 "adsorption_mol_m3": [float(v) for v in ads64]},
        "adsorption_max_rel_diff_mixed_vs_fp64": float(np.max(np.abs(ads - ads64) / np.abs(ads64))),
        "profile_max_rel_diff_mixed_vs_fp64": relmax(r.reshape(-1), rho64.reshape(-1)),
        "weighted_densities_and_transposed": wd,
        "adjoint_dN_deps": adjoint,
        "wd_note": "one weighted-density + transposed-convolution sequence per map precision, timed by the "
                   "autotuner with HIP events for its chosen variant; compulsory bytes (each field read or written once): "
                   "B_WD + B_T = 2 N C (1+F) s",
    }


def run_matrix_2d(np, DFTOptions, configs, converge_):
    from cdft_amd.methods import adsorption_system
    kw = dict(n=512, box=30.0, ndisks=16, seed=0)
    out = {"workload": "methane/ethane/propane x=[0.6,0.3,0.1], T=200 K, p=1e7 Pa, 30 sigma periodic square on "
                       "512^2 points, 16 hard disks r~U[1,2.5] sigma (seed 0), tol 1e-8, anderson_start 0.5"}
    ads = {}
    for prec in ("fp64", "mixed"):
        s, rho = configs.matrix_2d(options=DFTOptions(precision=prec), **kw)
        s.device_system.tune()
        t = time.perf_counter()
        converge_(s, rho, tol=1e-8, **configs.MATRIX_3D_SOLVER)
        wall = time.perf_counter() - t
        res = s.last_result
        ads[prec] = adsorption_system(s, rho)
        out[prec] = {"errcode": res["errcode"], "nevals": res["nevals"], "nevals_fp32": res["nevals_fp32"],
                     "solve_s": wall, "ms_per_eval": res["device_ms"] / max(res["nevals"], 1),
                     "tuning": s.device_system.get_tuning()}
        del s
    out["adsorption_max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(ads["mixed"] - ads["fp64"]) /
                                                                np.abs(ads["fp64"])))
    return out


def run_radial(np, configs, converge_):
    out = {}
    for coord in ("spherical", "cylindrical"):
        s, rho0 = configs.methane_radial(coord)
        r = rho0.copy()
        converge_(s, r, tol=1e-8)
        t = time.perf_counter()
        for _ in range(5):
            r = rho0.copy()
            converge_(s, r, tol=1e-8)
        dt = (time.perf_counter() - t) / 5
        res = s.last_result
        o, x, ro, secs = cpu_oracle_solve(s, rho0.reshape(-1), 1e-8)
        out[coord] = {"workload": f"methane at 150 K, 1e7 Pa in a {coord} graphite Steele pore of radius 5 sigma, "
                                  "151 radial points, tol 1e-8; FP64",
                      "errcode": res["errcode"], "nevals": res["nevals"], "solve_ms": dt * 1e3,
                      "cpu_baseline": {"value": 1.0 / secs, "unit": "profiles/s", "cores": 1, "kind": "port",
                                       "sample": "the same solve, FP64 C oracle (literal QDHT per field, operators "
                                                 "rebuilt at every evaluation), 1 thread",
                                       "errcode": ro.errcode, "nevals": ro.nevals},
                      "profile_max_rel_err_vs_oracle": relmax(r.reshape(-1), x)}
    return out


def run_adjoint(np, configs, converge_):
    from cdft_amd import DFTOptions
    from cdft_amd import adjoint as adj
    s, rho0 = configs.c1c10_interface(ngrid=2048, box=20.0)
    r = rho0.copy()
    t = time.perf_counter()
    converge_(s, r, tol=1e-12, maxit=20000)
    t_solve = time.perf_counter() - t
    rho = r.reshape(-1)
    t = time.perf_counter()
    sens, lam, res = adj.sensitivities(s, rho, "gamma", tol=1e-10)
    t_adj = time.perf_counter() - t
    # the mixed-precision leg: sensitivities at the profile of a mixed solve
    sm, _ = configs.c1c10_interface(ngrid=2048, box=20.0, options=DFTOptions(precision="mixed"))
    rm = rho0.copy()
    converge_(sm, rm, tol=1e-12, maxit=20000)
    sens_m, _, res_m = adj.sensitivities(sm, rm.reshape(-1), "gamma", tol=1e-10)
    # central differences of full FP64 solves for two of the six parameters, at the
    # configured step (1e-6 relative) and at 1e-5
    fd = {}
    t = time.perf_counter()
    for key in (("epsilon", 0), ("sigma", 1)):
        prm, i = key
        b0 = adj.model_params(s.model)[s.model.components[i]][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
        for rel_h in (1e-6, 1e-5):
            h = rel_h * abs(b0)
            vals = []
            for sg in (+1, -1):
                sp = adj.perturbed_system(s, i, prm, sg * h)
                rr = rho.copy().reshape(s.nc, s.N)
                converge_(sp, rr, tol=1e-12, maxit=20000)
                vals.append(adj.objective(sp, rr, "gamma"))
                del sp
            fd[key + (rel_h,)] = (vals[0] - vals[1]) / (2 * h)
    t_fd = (time.perf_counter() - t) / 4

    def diff(sv, rel_h):
        return max(abs(sv[k[:2]] - v) / abs(v) for k, v in fd.items() if k[2] == rel_h)
    return {
        "workload": "d(surface tension)/d(sigma_i, eps_i, m_i), methane/n-decane interface, 2048 points, "
                    "40 sigma box, T=300 K; FP64 and mixed-precision forward solves, FP64 adjoint",
        "solve_s": t_solve, "adjoint_and_partials_s": t_adj, "adjoint_errcode": res["errcode"],
        "adjoint_nevals": res["nevals"], "fd_full_solves_s_per_parameter": t_fd,
        "dgamma": {f"{k[0]}_{k[1]}": v for k, v in sens.items()},
        "dgamma_mixed": {f"{k[0]}_{k[1]}": v for k, v in sens_m.items()},
        "mixed_forward": {"errcode": res_m["errcode"], "adjoint_nevals": res_m["nevals"]},
        "max_rel_diff_vs_fd": {"fp64_h1e-6": diff(sens, 1e-6), "fp64_h1e-5": diff(sens, 1e-5),
                               "mixed_h1e-6": diff(sens_m, 1e-6), "mixed_h1e-5": diff(sens_m, 1e-5)},
        "fd_note": "the full solves stop on the interface's residual floor (about 1e-9, its translation mode "
                   "barely contracts): about 1e-4 relative noise in differences with h = 1e-6, 1e-5 with h = 1e-5",
    }


def termination_report(np, pores, batched, plain, cpu, mus):
    """Per-pore errcodes of the three paths, where they differ, and the largest
    relative differences of the batched and plain results from the oracle's over the
    pores all of them converge."""
    bcode = [r["errcode"] for r in batched["results"]]
    pcode = [r["errcode"] for r in plain["results"]]
    ocode = cpu["errcodes"]

    def hist(c):
        h = {}
        for v in c:
            h[str(v)] = h.get(str(v), 0) + 1
        return h

    def diff(a, b):
        return [i for i in range(len(a)) if a[i] != b[i]]

    err = {k: {"profile": 0.0, "adsorption": 0.0, "surface_tension": 0.0} for k in ("batched_mixed", "plain_fp64")}
    for i, (s, _) in enumerate(pores):
        if not (bcode[i] == pcode[i] == ocode[i] == 0):
            continue
        xo = cpu["profiles"][i]
        ao, to = cpu["adsorption"][i], cpu["tension"][i]
        p = s.structure.conditions[0]
        for key, x in (("batched_mixed", batched["profiles"][i]), ("plain_fp64", plain["profiles"][i])):
            e = err[key]
            e["profile"] = max(e["profile"], relmax(x, xo))
            a = s.device_system.adsorption(x)[0]
            e["adsorption"] = max(e["adsorption"], abs(a - ao) / abs(ao))
            t = s.device_system.surface_tension(x, mus[i], p)
            e["surface_tension"] = max(e["surface_tension"], abs(t - to) / abs(to))
    return {
        "errcodes": {"batched_mixed": hist(bcode), "plain_fp64": hist(pcode), "oracle_fp64": hist(ocode)},
        "mismatch_batched_vs_oracle": [{"pore": i, "batched": bcode[i], "oracle": ocode[i]} for i in diff(bcode, ocode)],
        "mismatch_plain_vs_oracle": [{"pore": i, "plain": pcode[i], "oracle": ocode[i]} for i in diff(pcode, ocode)],
        "mismatch_batched_vs_plain": [{"pore": i, "batched": bcode[i], "plain": pcode[i]} for i in diff(bcode, pcode)],
        "max_rel_err_vs_oracle": err,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--npores", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-1t-sample", type=int, default=32,
                    help="pores (spread over the batch) of the one-thread CPU figure")
    ap.add_argument("--precision", default="mixed", choices=["mixed"],
                    help="the batched path runs mixed precision (FP64 batches are switched off)")
    ap.add_argument("--no-extras", action="store_true", help="skip the c0, c1, 3D and adjoint measurements")
    ap.add_argument("--no-comparators", action="store_true",
                    help="time the batched path only (no plain, CPU or termination legs; for profiler runs)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this many MiB first (profiler counter scale)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the 512 pores are independent and would be "
              "split across GPUs by pore (no collectives), which is not built here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cdft_amd import DFTOptions, _lib, configs, converge_
    from cdft_amd.methods import bulk_chemical_potential as bulk_mu

    # ---------------- headline workload
    pores = configs.slit_pores(n=args.npores, seed=0, options=DFTOptions(precision=args.precision))
    systems, offs, brho, results, dt, kern_ms = run_batched(args, np, torch, _lib, pores)
    nconv = sum(1 for r in results if r["errcode"] == 0)
    codes = {}
    for r in results:
        codes[str(r["errcode"])] = codes.get(str(r["errcode"]), 0) + 1
    value = nconv * args.steps / dt
    dev_ms = np.array([r["device_ms"] for r in results])
    slow = int(np.argmax(dev_ms))

    # ---------------- roofline of the batched kernel (algorithmic bytes / kernel time)
    alg_bytes = 0.0
    for s, r in zip(systems, results):
        n32 = r["nevals_fp32"]
        n64 = r["nevals"] - n32
        alg_bytes += s.N * s.nc * (n32 * bytes_per_point_eval(4) + n64 * bytes_per_point_eval(8))
    kms = kern_ms / args.steps
    achieved = alg_bytes / (kms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "profiles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.npores} PC-SAFT methane slit pores between Steele graphite walls, T=150 K, "
                        "W~U[2,12] sigma, p~logU[1e3,1e7] Pa (seed 0), N=round(40 W/sigma) points",
            "tol": args.tol, "anderson_m": 5, "beta": 0.01, "precision": args.precision,
            "grid_points_total": int(sum(s.N * s.nc for s in systems)), "errcodes": codes,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_batched", "kernel_ms": kms, "algorithmic_bytes": alg_bytes,
                     "limiter": "not HBM: per-system latency of LDS FFT passes, barriers and dependent arithmetic; "
                                "the launch lasts as long as its slowest system"},
        "batched_device_ms": {"p50": float(np.percentile(dev_ms, 50)), "p90": float(np.percentile(dev_ms, 90)),
                              "p99": float(np.percentile(dev_ms, 99)), "max": float(dev_ms.max()),
                              "slowest_pore": slow, "slowest_pore_N": systems[slow].N,
                              "slowest_pore_nevals": results[slow]["nevals"],
                              "slowest_pore_errcode": results[slow]["errcode"]},
    }
    if not args.no_comparators:
        mus = [bulk_mu(s) for s, _ in pores]
        # ---------------- plain FP64 path on the same pores, one after another
        plain_pores = configs.slit_pores(n=args.npores, seed=0, options=DFTOptions(precision="fp64"))
        pprof, pres, t_plain = run_plain(args, plain_pores)
        nplain = sum(1 for r in pres if r["errcode"] == 0)
        plain_rate = nplain / t_plain
        out["plain_fp64"] = {"value": plain_rate, "unit": "profiles/s",
                             "sample": f"all {args.npores} pores, one after another", "seconds": t_plain,
                             "converged": nplain, "speedup_batched_vs_plain": value / plain_rate}
        # ---------------- the same FP64 pores batched (one workgroup each)
        out["batched_fp64"] = run_batched_fp64(args, np, torch, _lib, plain_pores, pprof, pres)
        # ---------------- CPU oracle, OpenMP over all pores and one thread on a sample
        cpu = run_cpu_baseline(args, np, plain_pores, mus)
        ncpu = sum(1 for c in cpu["errcodes"] if c == 0)
        secs = cpu["secs"]
        out["cpu_baseline"] = {
            "value": ncpu / cpu["wall"], "unit": "profiles/s", "cores": cpu["threads"], "kind": "port",
            "sample": f"all {args.npores} pores, FP64 C oracle (planned 1D evaluation), OpenMP over pores "
                      f"(dynamic schedule), {cpu['threads']} threads, {cpu['wall']:.2f} s",
            "one_thread": {"value": cpu["conv1"] / cpu["wall1"], "unit": "profiles/s",
                           "sample": f"{cpu['n1']} pores spread over the batch, 1 thread, {cpu['wall1']:.2f} s",
                           "all_pores_cpu_seconds": float(np.sum(secs))},
            "gpu_over_cpu": value / (ncpu / cpu["wall"]),
        }
        batched = {"results": results, "profiles": [brho[offs[i]:offs[i + 1]] for i in range(len(systems))]}
        plain = {"results": pres, "profiles": pprof}
        out["termination"] = termination_report(np, plain_pores, batched, plain, cpu, mus)
    if not args.no_extras:
        out["c0"] = run_c0(np, configs, converge_)
        out["c1"] = run_c1(np, configs, converge_)
        out["matrix_3d"] = run_matrix_3d(np, DFTOptions, configs, converge_)
        out["matrix_2d"] = run_matrix_2d(np, DFTOptions, configs, converge_)
        out["radial"] = run_radial(np, configs, converge_)
        out["adjoint"] = run_adjoint(np, configs, converge_)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
