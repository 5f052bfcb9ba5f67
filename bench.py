"""Benchmark: converged profiles per second (residual < 1e-8) of the batched
mixed-precision path on 512 methane slit pores, with the plain FP64 path timed on
the same 512 pores one after another and the CPU oracle on a sample of them.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--npores 512] [--no-extras]

A step is one batched solve of all pores from their seed profiles (inputs resident
in HBM).  Prints one JSON line.  Besides the headline workload the line carries
two further measurements of the same library (skipped with --no-extras):
  matrix_3d   the 3D hard-sphere-matrix ternary on a 256^3 grid: one mixed-precision
              solve, stream time per map evaluation and the algorithmic-byte rate;
  adjoint     surface-tension sensitivities of the methane / n-decane interface
              (2048 points): adjoint solve versus finite differences of full solves.
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


def run_c3(args, np, torch, DFTOptions, _lib, configs):
    pores = configs.slit_pores(n=args.npores, seed=0, options=DFTOptions(precision=args.precision))
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
    return pores, systems, sizes, offs, work, results, dt, kern_ms


def run_matrix_3d(np, DFTOptions, configs, converge_):
    s, rho = configs.matrix_3d(n=256, options=DFTOptions(precision="mixed"))
    dev = s.device_system
    r = rho.copy()
    t = time.perf_counter()
    converge_(s, r, tol=1e-8, **configs.MATRIX_3D_SOLVER)
    wall = time.perf_counter() - t
    res = s.last_result
    n32, n64 = res["nevals_fp32"], res["nevals"] - res["nevals_fp32"]
    F = 9
    alg = s.N * s.nc * (n32 * bytes_per_point_eval(4, F=F) + n64 * bytes_per_point_eval(8, F=F))
    ms = res["device_ms"]
    from cdft_amd.methods import adsorption_system
    return {
        "workload": "methane/ethane/propane x=[0.6,0.3,0.1], T=200 K, p=3e6 Pa, 30 sigma periodic cube on "
                    "256^3 points, 64 hard spheres r~U[1,2.5] sigma (seed 0); mixed precision, tol 1e-8, "
                    "anderson_start 0.5",
        "errcode": res["errcode"], "nevals": res["nevals"], "nevals_fp32": n32,
        "solve_s": wall, "stream_ms": ms, "ms_per_eval": ms / max(res["nevals"], 1),
        "algorithmic_GBps": alg / (ms * 1e-3) / 1e9,
        "adsorption_mol_m3": [float(v) for v in adsorption_system(s, r)],
    }


def run_adjoint(np, configs, converge_):
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
    # finite differences of full solves for two of the six parameters
    fd = {}
    t = time.perf_counter()
    for key in (("epsilon", 0), ("sigma", 1)):
        prm, i = key
        b0 = adj.model_params(s.model)[s.model.components[i]][{"segment": 0, "sigma": 1, "epsilon": 2}[prm]]
        h = 1e-5 * abs(b0)
        vals = []
        for sg in (+1, -1):
            sp = adj.perturbed_system(s, i, prm, sg * h)
            rr = rho.copy().reshape(s.nc, s.N)
            converge_(sp, rr, tol=1e-12, maxit=20000)
            vals.append(adj.objective(sp, rr, "gamma"))
        fd[key] = (vals[0] - vals[1]) / (2 * h)
    t_fd = (time.perf_counter() - t) / len(fd)
    return {
        "workload": "d(surface tension)/d(sigma_i, eps_i, m_i), methane/n-decane interface, 2048 points, "
                    "40 sigma box, T=300 K; FP64",
        "solve_s": t_solve, "adjoint_and_partials_s": t_adj, "adjoint_errcode": res["errcode"],
        "adjoint_nevals": res["nevals"], "fd_full_solves_s_per_parameter": t_fd,
        "dgamma": {f"{k[0]}_{k[1]}": v for k, v in sens.items()},
        "max_rel_diff_vs_fd": max(abs(sens[k] - v) / abs(v) for k, v in fd.items()),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--npores", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="plain FP64 comparator on this many pores spread over the batch (0: all, one after another)")
    ap.add_argument("--cpu-sample", type=int, default=8)
    ap.add_argument("--precision", default="mixed", choices=["mixed"],
                    help="the batched path runs mixed precision (FP64 batches are switched off)")
    ap.add_argument("--no-extras", action="store_true", help="skip the 3D and adjoint measurements")
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

    # ---------------- headline workload
    pores, systems, sizes, offs, work, results, dt, kern_ms = run_c3(args, np, torch, DFTOptions, _lib, configs)
    batched_rho = work.cpu().numpy()
    nconv = sum(1 for r in results if r["errcode"] == 0)
    codes = {}
    for r in results:
        codes[str(r["errcode"])] = codes.get(str(r["errcode"]), 0) + 1
    value = nconv * args.steps / dt

    # ---------------- roofline of the batched kernel (algorithmic bytes / kernel time)
    alg_bytes = 0.0
    for s, r in zip(systems, results):
        n32 = r["nevals_fp32"]
        n64 = r["nevals"] - n32
        alg_bytes += s.N * s.nc * (n32 * bytes_per_point_eval(4) + n64 * bytes_per_point_eval(8))
    kms = kern_ms / args.steps
    achieved = alg_bytes / (kms * 1e-3) / 1e9

    # ---------------- plain FP64 path on the same pores, one after another (all by default)
    if args.plain_sample > 0:
        idx = np.linspace(0, args.npores - 1, args.plain_sample).round().astype(int)
    else:
        idx = np.arange(args.npores)
    plain_pores = configs.slit_pores(n=args.npores, seed=0, options=DFTOptions(precision="fp64"))
    t_plain = 0.0
    nplain = 0
    from cdft_amd.methods import adsorption_system, surface_tension_system
    err = {"profile": 0.0, "adsorption": 0.0, "surface_tension": 0.0}
    for i in idx:
        s, r0 = plain_pores[i]
        s.device_system  # set-up outside the timed region
        rho = r0.copy()
        t1 = time.perf_counter()
        converge_(s, rho, tol=args.tol)
        t_plain += time.perf_counter() - t1
        nplain += s.last_result["errcode"] == 0
        if s.last_result["errcode"] == 0 and results[i]["errcode"] == 0:
            # mixed-precision batched result against this FP64 solve of the same pore
            rb = batched_rho[offs[i]:offs[i + 1]]
            ra = rho.reshape(-1)
            err["profile"] = max(err["profile"], float(np.max(np.abs(rb - ra)) / np.max(np.abs(ra))))
            a, b = adsorption_system(s, ra)[0], adsorption_system(s, rb)[0]
            err["adsorption"] = max(err["adsorption"], abs(b - a) / abs(a))
            g, h = surface_tension_system(s, ra), surface_tension_system(s, rb)
            err["surface_tension"] = max(err["surface_tension"], abs(h - g) / abs(g))
    plain_rate = nplain / t_plain

    # ---------------- CPU oracle (1 thread, FP64) on a sample
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cidx = np.linspace(0, args.npores - 1, args.cpu_sample).round().astype(int)
    t_cpu = 0.0
    ncpu = 0
    for i in cidx:
        s, r0 = plain_pores[i]
        o = po.Oracle(**s.descriptor())
        t1 = time.perf_counter()
        _, ro, _ = o.converge(r0.reshape(-1), tol=args.tol)
        t_cpu += time.perf_counter() - t1
        ncpu += ro.errcode == 0
    cpu_rate = ncpu / t_cpu

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
            "grid_points_total": int(sum(sizes)), "errcodes": codes,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_batched", "kernel_ms": kms, "algorithmic_bytes": alg_bytes},
        "plain_fp64": {"value": plain_rate, "unit": "profiles/s",
                       "sample": (f"all {len(idx)} pores, one after another" if len(idx) == args.npores
                                  else f"{len(idx)} pores spread over the batch"),
                       "seconds": t_plain,
                       "converged": int(nplain), "speedup_batched_vs_plain": value / plain_rate,
                       "batched_mixed_max_rel_err": err},
        "cpu_baseline": {"value": cpu_rate, "unit": "profiles/s", "cores": 1, "kind": "port",
                         "sample": f"{len(cidx)} pores spread over the batch, FP64 C oracle, 1 thread"},
    }
    if not args.no_extras:
        out["matrix_3d"] = run_matrix_3d(np, DFTOptions, configs, converge_)
        out["adjoint"] = run_adjoint(np, configs, converge_)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
