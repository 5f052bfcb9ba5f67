/*
 * cdft_hip.h -- C ABI of the MI355X-native PC-SAFT classical-DFT hot path.
 *
 * The library evaluates the PC-SAFT free-energy functional (weighted densities,
 * pointwise Helmholtz density and its gradient, transposed convolutions) and runs
 * the Euler-Lagrange Picard/Anderson fixed-point loop on an AMD Instinct MI355X
 * (gfx950).  It is the drop-in for this call chain of cDFT.jl:
 *
 *   converge!(system, rho; maxit, beta, tol, anderson_start, anderson_m)
 *                                                    src/methods/converge.jl:214-254
 *     -> aasol(GFix!, ln_X0, method)                 src/utils/anderson.jl:301-443
 *       -> get_new_profile!(system, rho, ...)        src/methods/converge.jl:108-153
 *         -> dFdrho_res!(system, rho, ...)            src/models/models.jl:131-186
 *
 * Conventions
 *  - Every function returns CDFT_OK (0) or a negative error code; the message of
 *    the last error on a context is returned by cdft_last_error().
 *  - Arrays are column-major like the reference's Julia arrays: a profile
 *    rho(ngrid..., nc) is rho[c*N + idx] with idx = i0 + n0*(i1 + n1*i2).
 *  - Host pointers are read during the call only; the library keeps no caller
 *    pointer after it returns.  The *_device variants take device pointers.
 *  - One HIP stream per context.  A context is not re-entrant; separate contexts
 *    may be used from separate threads.
 *  - Physical inputs use the reference's units: lengths in m, temperature in K,
 *    densities in mol/m^3, chemical potentials as beta*mu (dimensionless).
 */
#ifndef CDFT_HIP_H
#define CDFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CDFT_OK 0
#define CDFT_ERR_ARG (-1)
#define CDFT_ERR_HIP (-2)
#define CDFT_ERR_UNSUPPORTED (-3)
#define CDFT_ERR_NOMEM (-4)
#define CDFT_ERR_STOPPED (-5)   /* a progress callback asked the solve to stop */

/* precision modes (DFTOptions(...; precision), src/base/devices.jl:29-40) */
#define CDFT_FP64 0   /* everything in FP64 (the reference's default)                */
#define CDFT_MIXED 1  /* FP32 convolutions and Helmholtz density; FP64 n3-singular   */
                      /* terms, exponent assembly, Anderson and residual norms       */

#define CDFT_MAXC 8

typedef struct cdft_ctx cdft_ctx;
typedef struct cdft_sys cdft_sys;

/* Everything DFTSystem(model, structure[, external_field][, options]) carries that
 * the hot path reads (src/base/base.jl:50-89).  The bulk quantities come from the
 * caller's equation of state, as the reference takes them from Clapeyron
 * (get_species, src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:65-72). */
typedef struct {
    int nc;                 /* components, one bead each (PC-SAFT)               */
    int ndim;               /* 1 or 3 Cartesian dimensions                       */
    int ngrid[3];           /* grid points per axis (unused axes ignored)        */
    double lb[3], ub[3];    /* bounds per axis, m                                */
    double T;               /* temperature, K                                    */
    double p;               /* pressure, Pa (carried for post-processing)       */
    double L;               /* length scale max(sigma), m (PCSAFT.jl:79-81)      */
    double psi;             /* dispersion width factor, 1.3862 (PCSAFT.jl:53)   */
    const double *m;        /* [nc] segment numbers                              */
    const double *d;        /* [nc] hard-sphere diameters at T, m                */
    const double *sigma;    /* [nc*nc] sigma_ij, m                               */
    const double *eps;      /* [nc*nc] eps_ij / k_B, K                           */
    const double *rho_bulk; /* [nc] bulk densities, mol/m^3                      */
    const double *mu_res;   /* [nc] beta mu_res of the bulk                      */
    const double *vext;     /* [N*nc] beta V_ext (external_field.jl:6-19), NULL  */
    int precision;          /* CDFT_FP64 or CDFT_MIXED                           */
    /* Wertheim association, the data pack_assoc_params flattens
     * (src/models/DFT/assoc.jl:18-57).  npair = 0 (or a zeroed tail) means none;
     * the device path takes up to 4 sites and 8 pairs. */
    int assoc_nsite;                  /* site types over all components          */
    int assoc_npair;                  /* bonding pairs                           */
    const int *assoc_site_comp;       /* [nsite] component of each site, 0-based */
    const double *assoc_site_count;   /* [nsite] multiplicity of each site type  */
    const int *assoc_pair_site;       /* [2*npair] the two sites (a, b) per pair */
    const double *assoc_pair_eps;     /* [npair] eps_AB / k_B, K                 */
    const double *assoc_pair_kappa;   /* [npair] bonding volume kappa_AB         */
    /* coordinates (base/structure.jl:363-433): CDFT_CARTESIAN, or CDFT_SPHERICAL /
     * CDFT_CYLINDRICAL with ndim = 1: a radial grid of ngrid[0] points over
     * [0, ub[0]] (the sample radii of a quasi-discrete Hankel transform; lb[0]
     * places walls only), FP64 only, at most CDFT_RADIAL_MAXN points */
    int coord;
} cdft_system_desc;

#define CDFT_CARTESIAN 0
#define CDFT_SPHERICAL 1
#define CDFT_CYLINDRICAL 2
#define CDFT_RADIAL_MAXN 2048

/* Per-evaluation progress hook: converge!'s logger GFix_logger
 * (src/methods/converge.jl:256-267), which the reference calls after every map
 * evaluation with the evaluation count and the profile rho = exp(x) that evaluation
 * used, to log the free energy every log_interval evaluations and to call
 * save_callback(iter, rho) every save_interval.  Here the solve pauses on the device
 * after every `every`-th evaluation and calls progress(user, iter, rho, n) on the
 * host with that profile (n = N nc doubles, layout as rho_inout, valid during the
 * call).  The callback may evaluate the functional of the same system
 * (cdft_functional_eval); it must not start another solve of it.  A nonzero return
 * stops the solve: cdft_converge then returns CDFT_ERR_STOPPED with that profile in
 * rho_inout. */
typedef int (*cdft_progress_fn)(void *user, int iter, const double *rho, long n);

/* converge! keywords (src/methods/converge.jl:3-13, :51-60). */
typedef struct {
    int maxit;             /* 10000                                               */
    double beta;           /* 1e-2 damping (Picard and Anderson)                  */
    double tol;            /* 1e-4; used as rtol and atol                         */
    double anderson_start; /* 1e-1; picard_rtol and picard_atol                   */
    int m;                 /* 5 Anderson depth (0 = damped Picard)                */
    int hist_cap;          /* capacity of hist (0: no history)                    */
    double *hist;          /* residual-norm history out (host), or NULL           */
    cdft_progress_fn progress; /* or NULL (the batched path ignores it)           */
    void *progress_user;
    int progress_every;    /* >= 1 when progress is set                            */
} cdft_solver_opts;

/* aasol's outcome (AndersonOK, src/utils/anderson.jl:727-737). */
typedef struct {
    int idid;              /* 1 converged, 0 not                                  */
    int errcode;           /* 0 ok, -1 initial iterate within tol, -2 diverged,   */
                           /* 10 maxit reached                                    */
    int iters;             /* Anderson iteration counter k                        */
    int nevals;            /* fixed-point map evaluations                          */
    int nhist;             /* residual norms recorded                             */
    double resnorm;        /* final ||beta (G(x) - x)||_2                          */
    double tol;            /* rtol*res0 + atol                                    */
    int nevals_fp32;       /* map evaluations done in FP32 (CDFT_MIXED)           */
    int reserved;
    double device_ms;      /* batched: this system's time in its workgroup; plain:
                              stream time of the solve                          */
} cdft_result;

/* Context: device selection and one HIP stream.  (No counterpart: the
 * reference's device is the KernelAbstractions backend, devices.jl:29-40.) */
int cdft_ctx_create(int device, cdft_ctx **out);
int cdft_ctx_destroy(cdft_ctx *ctx);
const char *cdft_last_error(const cdft_ctx *ctx);

/* System set-up: grid, analytic Fourier weights and their real-space kernels,
 * external potential, device buffers.  Replaces DFTSystem(...) + preallocate
 * (src/base/base.jl:61-89, src/base/devices.jl:51-102,
 * src/fields/weighted_densities.jl:28-80,193-246). */
int cdft_system_create(cdft_ctx *ctx, const cdft_system_desc *desc, cdft_sys **out);
int cdft_system_destroy(cdft_sys *sys);
int cdft_system_npoints(const cdft_sys *sys);

/* One functional evaluation: dfdrho = delta F_res / delta rho (no external
 * field), and optionally F_res (Simpson integral of Phi / L^3) and the pointwise
 * Phi (fval, length N).  Replaces dFdrho_res(system, rho) + F_res
 * (src/models/models.jl:86-91,131-200). */
int cdft_functional_eval(cdft_sys *sys, const double *rho, double *dfdrho, double *f_res, double *fval);

/* The same derivative from the FP32 map of a mixed-precision system (weighted
 * densities, Phi's gradient and transposed convolutions in FP32, compensated sums
 * in 1D, the n3-singular factors and the result in FP64): the map the mixed solve
 * iterates before its switch to FP64.  CDFT_ERR_UNSUPPORTED for an FP64 system.
 * (No reference counterpart: the reference runs one precision throughout,
 * src/base/devices.jl:29-40.) */
int cdft_functional_eval_fp32(cdft_sys *sys, const double *rho, double *dfdrho);

/* Euler-Lagrange fixed point: rho is overwritten with the converged profile.
 * Plain path: one system, per-operator kernels; the whole Anderson iteration,
 * its control flow included, runs on the device from a captured hipGraph and the
 * host reads the solver state once per 16 map evaluations (and at each pause for
 * opts->progress).  Replaces converge!(prob, ::AASol, rho) (src/methods/converge.jl:220-267)
 * and aasol (src/utils/anderson.jl:301-443). */
int cdft_converge(cdft_sys *sys, double *rho_inout, const cdft_solver_opts *opts, cdft_result *result);

/* Batched path: nsys independent systems (any mix of 1D grid sizes), one
 * persistent launch, one workgroup per system with its Anderson history resident
 * on the device; systems retire independently.  Each system follows exactly the
 * same algorithm as cdft_converge.  (No batched counterpart in the reference: it
 * replaces a loop of converge! calls, e.g. over adsorption(...) state points,
 * src/methods/adsorption.jl:48-66.)  opts->hist, when set, holds nsys x hist_cap
 * residual norms (system i's from hist + i hist_cap, one per map evaluation,
 * result[i].nhist of them recorded; the allocation-free entry point ignores it);
 * progress is ignored.  Two kernels serve batches: one-component CDFT_MIXED 1D
 * systems of at most 512 points with anderson m <= 5 take the register-state kernel
 * (FP32 map to its noise floor, then FP64); other 1D Cartesian systems of one or two
 * components with N nc <= 512 (CDFT_FP64 or CDFT_MIXED, one precision and one
 * association variant per batch) take one workgroup each running cdft_converge's
 * single-workgroup bodies to the end (the same bits as cdft_converge of each
 * system, except mixed-precision systems with association: same termination, profiles
 * equal to rounding).  Other batches return CDFT_ERR_UNSUPPORTED (use cdft_converge
 * per system). */
int cdft_converge_batch(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_inout,
                        const cdft_solver_opts *opts, cdft_result *results);
/* Same with rho_inout as device pointers (inputs resident in HBM). */
int cdft_converge_batch_device(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_dev,
                               const cdft_solver_opts *opts, cdft_result *results);

/* Allocation-free variant of cdft_converge_batch_device.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc/hipFree: the system descriptors and results of the launch live in a
 * static device array of the library (one per GPU, about 1.4 MB, held until the
 * process ends), and every system must have been prepared for the batched path when
 * it was created (cdft_system_create does so for one-component 1D systems of at most
 * 512 points).  At most CDFT_NOALLOC_MAX_SYS systems per call (CDFT_ERR_ARG beyond).
 * Same arguments, restrictions and results as cdft_converge_batch_device.  Calls
 * that overlap on one GPU (from different contexts or threads) share the array and
 * race; serialise them. */
#define CDFT_NOALLOC_MAX_SYS 1024
int cdft_converge_batch_device_noalloc(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_dev,
                                       const cdft_solver_opts *opts, cdft_result *results);

/* Stream time in ms of the last solve on the context, measured with HIP events
 * on its stream: the persistent kernel of cdft_converge_batch*, or first to last
 * kernel of cdft_converge / cdft_adjoint_solve (their result's device_ms too). */
double cdft_last_kernel_ms(const cdft_ctx *ctx);

/* Tuning of the spectral path's line-FFT passes (2D and 3D grids), per map
 * precision: lines per workgroup tile and, for z lines of 128 or 256 points, the
 * all-pairs z pass of the transposed convolutions.  Every choice gives the same
 * results bit for bit; only the time differs.  cdft_converge tunes a system of at
 * least 2^16 grid points before its first solve (unless CDFT_AUTOTUNE=0), measuring
 * each candidate on the device; a shape measured once is reused by every system of
 * that shape in the context.  (No reference counterpart: FFTW plans are not
 * tuned there, src/structure/transforms.jl:19-22.) */
typedef struct {
    int lines_per_tile[2];   /* [0] FP64 map, [1] FP32 map: 4, 8 or 16               */
    int fused_z[2];          /* 1: the all-pairs z pass                               */
    double ms[2];            /* one weighted-density + transposed-convolution
                                sequence with the choice, ms (0: not measured)        */
    int tuned;               /* 1 once measured for this shape or set explicitly      */
} cdft_tuning;

/* Measure the candidates (if not yet measured for the shape, or force != 0) and
 * apply the fastest; out (optional) receives the choice.  A 1D system has nothing
 * to tune (returns its defaults). */
int cdft_system_tune(cdft_sys *sys, int force, cdft_tuning *out);
int cdft_system_get_tuning(const cdft_sys *sys, cdft_tuning *out);
/* Apply a choice without measuring (CDFT_ERR_ARG for an invalid one); the solves
 * keep it. */
int cdft_system_set_tuning(cdft_sys *sys, const cdft_tuning *in);

/* Post-processing of a converged profile.
 * surface_tension: out = F k_B T - sum_i mu_i int rho_i + p int 1 with F = F_ideal +
 * F_res (Simpson, dz = (ub - lb) / N); mu[nc] in J/mol, p in Pa.  Replaces
 * surface_tension(system, rho) (src/methods/surface_tension.jl:31-53); the planar
 * interface's tension is half of it (surface_tension.jl:28).
 * adsorption: out[i] = int rho_i / (ub - lb).  Replaces adsorption(system, rho)
 * (src/methods/adsorption.jl:22-32). */
int cdft_surface_tension(cdft_sys *sys, const double *rho, const double *mu, double p, double *out);
int cdft_adsorption(cdft_sys *sys, const double *rho, double *out);

/* Differentiation.  The reference's only derivative is dPhi/dn, taken with
 * Enzyme inside dFdrho_res (src/models/models.jl:355-411); these have no
 * reference counterpart.
 * phi_vjp: pullback of sum_i cot_i Phi(n_i) to rho, divided by N_A L^3:
 *   out = cot * dPhi/dn_rho + sum_f W_f^T (cot * dPhi/dn_f)   (cot[N]);
 * with cot = 1 it equals cdft_functional_eval's dfdrho (the rrule of F_res).
 * hvp: out = d(dfdrho)/drho . v at rho, the Hessian of F_res applied to v.
 * adjoint_solve: lambda with (I - J^T) lambda = g at the fixed point rho, J the
 * Jacobian of the undamped map x -> clamp(mu_res + ln rho_b - dfdrho(e^x) - V)
 * (converge.jl:108-153) at x = log rho.  Solved by the same damped Picard /
 * Anderson iteration as cdft_converge (opts, result as there); lambda_inout holds
 * the initial guess.  dJ/dtheta = dJ/dtheta|_rho + lambda . dlnG/dtheta|_rho for an
 * objective J(rho, theta) with g = rho * dJ/drho. */
int cdft_phi_vjp(cdft_sys *sys, const double *rho, const double *cot, double *out);
int cdft_hvp(cdft_sys *sys, const double *rho, const double *v, double *out);
int cdft_adjoint_solve(cdft_sys *sys, const double *rho, const double *g, double *lambda_inout,
                       const cdft_solver_opts *opts, cdft_result *result);

/* adjoint_params: the device part of parameter sensitivities at a converged rho of
 * sys, for nobj objectives J_k at once.  With g (nobj x n, g_k = rho * dJ_k/drho)
 * it first solves lambda_k as cdft_adjoint_solve does (lambda_inout[k n ..] the
 * initial guess in, lambda_k out; results[k] its outcome); with g == NULL
 * lambda_inout already holds the multipliers.  Then for every parameter p, from two
 * systems built with theta_p + h[p] and theta_p - h[p] over the same grid
 * (sys_plus[p], sys_minus[p], same context; their own bulk state and external
 * field), out[k nparam + p] = lambda_k . (G_plus(x) - G_minus(x)) / (2 h[p]) at
 * x = log rho, G the undamped map above: the implicit part of dJ_k/dtheta_p, to which
 * the caller adds the explicit partial dJ_k/dtheta_p at fixed rho.  nparam = 0 only
 * solves.  (No reference counterpart: the reference differentiates no solve.) */
int cdft_adjoint_params(cdft_sys *sys, const double *rho, int nobj, const double *g, double *lambda_inout,
                        int nparam, cdft_sys *const *sys_plus, cdft_sys *const *sys_minus, const double *h,
                        const cdft_solver_opts *opts, double *out, cdft_result *results);

/* Dynamic DFT: the right-hand side d(eta)/dt of the log-density form, eta = ln rho
 * (nc*N values, same layout as rho): rho = exp(clamp(eta, -50, 30)),
 * mu = dF_res/drho + beta V_ext, and per component
 *   deta = lap(eta) + lap(mu) + sum_d grad_d(eta)^2 + grad_d(mu) . grad_d(eta),
 * the derivatives spectral over the periodic box in reduced lengths (z / L).
 * Replaces ddft_rhs_log! of ODEProblem(system, rho, tspan)
 * (ext/SciMLBasecDFTExt.jl:10-61); an ODE integrator calls it per stage. */
int cdft_ddft_rhs(cdft_sys *sys, const double *eta, double *deta);

#ifdef __cplusplus
}
#endif
#endif
