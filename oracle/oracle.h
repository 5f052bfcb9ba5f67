/*
 * oracle.h -- CPU FP64 restatement of cDFT.jl's PC-SAFT functional and its
 * Euler-Lagrange / Anderson fixed-point loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the cDFT.jl repository root).
 *
 * Array conventions follow the reference's column-major Julia arrays:
 *   grid index   idx = i0 + n0*(i1 + n1*i2)
 *   rho          rho[c*N + idx]                    (ngrid..., C)
 *   fields       n[(c*F + f)*N + idx]              (ngrid..., F, C), F = 6 + nd
 *   local fields nloc[c*F + f]                     (F, C) at one grid point
 * Field slots (0-based): 0 rho, 1 int rho dz (d/2), 2 int rho z^2 dz (d/2),
 * 3..2+nd int rho z dz (d/2) vector, 3+nd int rho z^2 dz (d), 4+nd int rho dz (d),
 * 5+nd int rho z^2 dz (psi d)         (PCSAFT.jl:55-61).
 */
#ifndef CDFT_ORACLE_H
#define CDFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXC 8
#define OR_MAXS 20                   /* association sites (assoc.jl pads to 20)  */
#define OR_MAXP 20                   /* association bonding pairs               */

typedef struct {
    int nc;                          /* components (one bead each)          */
    int nd;                          /* spatial dimensions 1..3             */
    int ngrid[3];
    double lb[3], ub[3];             /* bounds, m                           */
    double T;                        /* K                                   */
    double L;                        /* length scale max(sigma), m          */
    double psi;                      /* dispersion width factor (1.3862)    */
    double m[OR_MAXC];               /* segment number                      */
    double d[OR_MAXC];               /* hard-sphere diameter, m             */
    double sigma[OR_MAXC * OR_MAXC]; /* m, sigma[i*nc+j]                    */
    double eps[OR_MAXC * OR_MAXC];   /* K                                   */
    double mu_res[OR_MAXC];          /* beta mu_res (bulk)                  */
    double rho_bulk[OR_MAXC];        /* mol/m^3                             */
    const double *vext;              /* beta V_ext [c*N+idx] or NULL        */
    /* Wertheim association (assoc.jl:18-57 pack_assoc_params); npair = 0: none */
    int nsite, npair;
    int site_comp[OR_MAXS];          /* component of each site (0-based)    */
    double site_count[OR_MAXS];      /* multiplicity n of each site type    */
    int pair_a[OR_MAXP], pair_b[OR_MAXP]; /* global site indices of a pair  */
    double pair_eps[OR_MAXP];        /* eps_AB / k_B, K                     */
    double pair_kappa[OR_MAXP];      /* bonding volume kappa_AB             */
    /* coordinates (base/structure.jl:363-433): 0 Cartesian; 1 spherical, 2
     * cylindrical (nd = 1, radial grid of a quasi-discrete Hankel transform
     * over [0, ub]; lb places walls only) */
    int coord;
} or_system;

typedef struct {
    int m, maxit, picard_maxit;
    double rtol, atol, beta, picard_beta, picard_rtol, picard_atol;
} or_aa_opts;

typedef struct {
    int idid, errcode, k, nhist, nstats, nevals;
    double resnorm, tol;
} or_aa_result;

typedef void (*or_gfix)(double *G, const double *x, void *ctx);

/* physical constants (CODATA 2018 exact values) */
#define OR_NA 6.02214076e23
#define OR_KB 1.380649e-23
#define OR_R  8.31446261815324

/* Quasi-discrete Hankel transform of order 0 (Hankel.jl QDHT, restated from
 * Guizar-Sicairos & Gutierrez-Vega, JOSA A 21, 53 (2004)): cylindrical = the
 * 2D radial Fourier transform (roots of J0), spherical = the 3D one as the
 * order-1/2 transform of sqrt(r) f (roots n pi, a sine transform). */
typedef struct {
    int N, sph;
    double R, S, K;                  /* aperture, j_{N+1}, S / R            */
    double *r, *k;                   /* sample radii and angular wavenumbers */
    double *jp1;                     /* |J_{p+1}(j_n)|                       */
    double *C;                       /* C[m*N+n] = 2/S J_p(j_m j_n / S) / (|J_{p+1}(j_m)| |J_{p+1}(j_n)|) */
    double *w;                       /* integration weights of int f dV (4 pi r^2 dr or 2 pi r dr) */
} or_qdht;
or_qdht *or_qdht_create(int coord, double R, int N);
void   or_qdht_free(or_qdht *q);
void   or_qdht_forward(const or_qdht *q, const double *f, double *F);   /* F(k_m) ~ Fourier transform */
void   or_qdht_backward(const or_qdht *q, const double *F, double *f);  /* its inverse                  */

int    or_nfields(int nd);
int    or_npoints(const or_system *s);

/* FFTW-convention complex FFT along every axis, in place; inverse is scaled 1/N */
void   or_fft(double *re, double *im, int nd, const int *ngrid, int inverse);

void   or_omega(const or_system *s, double *om_re, double *om_im);
void   or_weighted_densities(const or_system *s, const double *rho, double *n);
double or_phi(const or_system *s, const double *nloc, double *dphi);
/* association site fractions X (nsite values) at one point (assoc.jl:702-831) */
void   or_assoc_X(const or_system *s, const double *nloc, double *X);
void   or_transposed(const or_system *s, const double *dphi, double *out);
void   or_dFdrho_res(const or_system *s, const double *rho, double *dfdrho, double *fval);
double or_simpson(const double *f, int nd, const int *ngrid, const double *dz);
double or_F_res(const or_system *s, const double *rho);
double or_F_ideal(const or_system *s, const double *rho);
void   or_get_new_profile(const or_system *s, const double *rho, double *lnG);

int    or_aasol(or_gfix G, void *ctx, int n, const double *x0, const or_aa_opts *o,
                double *sol, double *gx, double *hist, int hist_cap,
                double *condhist, double *alphanorm, int stats_cap, or_aa_result *res);
int    or_converge(const or_system *s, double *rho, int maxit, double beta, double tol,
                   double anderson_start, int m, or_aa_result *res,
                   double *hist, int hist_cap);

void   or_steele_vext(const or_system *s, double eps_s, double sigma_s, double delta,
                      double rho_s, double width, double *vext);
void   or_init_profiles(const or_system *s, double *rho);
void   or_init_two_phase(const or_system *s, const double *rho2, const double *Tc, double *rho);

double or_surface_tension(const or_system *s, const double *rho, const double *mu, double p);
void   or_adsorption(const or_system *s, const double *rho, double *out);

/* Planned 1D evaluation (kernel spectra computed once per system from the
 * literal Omega and FFT; same operator, different rounding) and an OpenMP
 * driver over independent systems.  A plan's buffers serve one call at a time. */
typedef struct or_plan or_plan;
or_plan *or_plan_create(const or_system *s);            /* NULL unless nd == 1 */
void   or_plan_destroy(or_plan *p);
void   or_dFdrho_res_planned(or_plan *p, const double *rho, double *dfdrho, double *fval);
int    or_converge_planned(const or_system *s, double *rho, int maxit, double beta, double tol,
                           double anderson_start, int m, or_aa_result *res, double *hist, int hist_cap);
int    or_converge_batch(const or_system *systems, int nsys, double *const *rhos, int maxit, double beta,
                         double tol, double anderson_start, int m, int nthreads, or_aa_result *res,
                         double *secs);
double or_wtime(void);

/* DDFT right-hand side d(eta)/dt of the log-density form, eta = ln rho
 * (ext/SciMLBasecDFTExt.jl:10-61). */
void   or_ddft_rhs(const or_system *s, const double *eta, double *deta);

#ifdef __cplusplus
}
#endif
#endif
