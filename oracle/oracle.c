/*
 * oracle.c -- CPU FP64 restatement of cDFT.jl's PC-SAFT functional evaluation
 * and Euler-Lagrange / Anderson fixed-point loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  It is
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product library.
 *
 * The convolutions are done exactly as the reference does them: one forward
 * complex FFT, a multiply by the analytic kernel, one inverse FFT and a real()
 * per (field, component) (src/utils/integrals.jl:88-96).  The FFT is a plain
 * mixed-radix decimation-in-time transform with twiddles taken from an 80-bit
 * table, so any grid size works.
 */
#define _POSIX_C_SOURCE 200809L
#define _DEFAULT_SOURCE              /* j0, j1 (glibc's Bessel functions of the first kind) */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define PI 3.14159265358979323846

int or_nfields(int nd) { return 6 + nd; }

double or_wtime(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int or_npoints(const or_system *s)
{
    int N = 1;
    for (int a = 0; a < s->nd; a++) N *= s->ngrid[a];
    return N;
}

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (FFTW sign convention).                           */
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* Twiddle table W[j] = exp(sign * 2 pi i j / N), j = 0..N-1. */
static void twiddles(int N, int sign, double *wr, double *wi)
{
    for (int j = 0; j < N; j++) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)N;
        wr[j] = (double)cosl(a);
        wi[j] = (double)(sign * sinl(a));
    }
}

/* out[k] = sum_j in[j*stride] W_n^{jk}, n divides N, W_n = W_N^{N/n}. */
static void fft_rec(const double *ir, const double *ii, int stride, double *or_, double *oi, int n,
                    int N, const double *wr, const double *wi, double *sr, double *si)
{
    if (n == 1) {
        or_[0] = ir[0];
        oi[0] = ii[0];
        return;
    }
    int p = smallest_factor(n), q = n / p, tw = N / n;
    for (int r = 0; r < p; r++)
        fft_rec(ir + (size_t)r * stride, ii + (size_t)r * stride, stride * p, or_ + r * q, oi + r * q, q, N,
                wr, wi, sr, si);
    /* combine: X[k + s q] = sum_r W_n^{r(k + s q)} Y_r[k] */
    for (int k = 0; k < q; k++) {
        for (int s = 0; s < p; s++) {
            double accr = 0.0, acci = 0.0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * (long)(k + s * q)) % n;
                int j = (int)(e * tw);
                double yr = or_[r * q + k], yi = oi[r * q + k];
                accr += yr * wr[j] - yi * wi[j];
                acci += yr * wi[j] + yi * wr[j];
            }
            sr[k + s * q] = accr;
            si[k + s * q] = acci;
        }
    }
    memcpy(or_, sr, sizeof(double) * n);
    memcpy(oi, si, sizeof(double) * n);
}

static void fft_line(double *re, double *im, int n, int sign)
{
    double *wr = malloc(sizeof(double) * n), *wi = malloc(sizeof(double) * n);
    double *ir = malloc(sizeof(double) * n), *ii = malloc(sizeof(double) * n);
    double *sr = malloc(sizeof(double) * n), *si = malloc(sizeof(double) * n);
    twiddles(n, sign, wr, wi);
    memcpy(ir, re, sizeof(double) * n);
    memcpy(ii, im, sizeof(double) * n);
    fft_rec(ir, ii, 1, re, im, n, n, wr, wi, sr, si);
    free(wr); free(wi); free(ir); free(ii); free(sr); free(si);
}

void or_fft(double *re, double *im, int nd, const int *ngrid, int inverse)
{
    int N = 1;
    for (int a = 0; a < nd; a++) N *= ngrid[a];
    int sign = inverse ? +1 : -1;
    int stride = 1;
    for (int a = 0; a < nd; a++) {
        int n = ngrid[a];
        double *lr = malloc(sizeof(double) * n), *li = malloc(sizeof(double) * n);
        int outer = N / (n * stride);
        for (int o = 0; o < outer; o++)
            for (int in = 0; in < stride; in++) {
                size_t base = (size_t)o * n * stride + in;
                for (int j = 0; j < n; j++) {
                    lr[j] = re[base + (size_t)j * stride];
                    li[j] = im[base + (size_t)j * stride];
                }
                fft_line(lr, li, n, sign);
                for (int j = 0; j < n; j++) {
                    re[base + (size_t)j * stride] = lr[j];
                    im[base + (size_t)j * stride] = li[j];
                }
            }
        free(lr); free(li);
        stride *= n;
    }
    if (inverse) {
        double s = 1.0 / (double)N;
        for (int j = 0; j < N; j++) { re[j] *= s; im[j] *= s; }
    }
}

/* ------------------------------------------------------------------------ */
/* Analytic Fourier weights.                                                 */
/*   frequencies: src/structure/structure.jl:153-185 (fftfreq(N, N/(ub-lb))) */
/*   reduced units: PCSAFT.jl:30,51-54 (omega * L, width = d/L)              */
/*   kernels: src/fields/weighted_densities.jl:48-69 (scalar), 210-227 (vec) */
/* ------------------------------------------------------------------------ */

/* omega component a of the grid point with multi-index (i0,i1,i2). */
static double omega_axis(const or_system *s, int a, int i)
{
    int n = s->ngrid[a];
    double fs = (double)n / (s->ub[a] - s->lb[a]);
    double mult = fs / (double)n;                /* AbstractFFTs.fftfreq */
    int nnonneg = (n + 1) / 2;
    int k = (i < nnonneg) ? i : i - n;
    return ((double)k * mult) * s->L;            /* _scaled_omega */
}

enum { K_RHO = 0, K_S = 1, K_V = 2, K_VEC = 3 };

/* Field table for PC-SAFT (PCSAFT.jl:55-61): kind and width factor of d. */
static void field_desc(int nd, int f, int *kind, double *wfac, int *vdir, double psi)
{
    *vdir = -1;
    if (f == 0) { *kind = K_RHO; *wfac = 0.0; return; }
    if (f == 1) { *kind = K_S; *wfac = 0.5; return; }
    if (f == 2) { *kind = K_V; *wfac = 0.5; return; }
    if (f >= 3 && f < 3 + nd) { *kind = K_VEC; *wfac = 0.5; *vdir = f - 3; return; }
    if (f == 3 + nd) { *kind = K_V; *wfac = 1.0; return; }
    if (f == 4 + nd) { *kind = K_S; *wfac = 1.0; return; }
    *kind = K_V; *wfac = psi;                    /* f == 5 + nd */
}

void or_omega(const or_system *s, double *om_re, double *om_im)
{
    int nd = s->nd, F = 6 + nd, N = or_npoints(s);
    int n0 = s->ngrid[0], n1 = nd > 1 ? s->ngrid[1] : 1;
    const double twopi = 2.0 * PI, fourpi = 4.0 * PI, tp3 = twopi * twopi * twopi;
    for (int c = 0; c < s->nc; c++) {
        double dred = s->d[c] / s->L;
        for (int f = 0; f < F; f++) {
            int kind, vdir;
            double wfac;
            field_desc(nd, f, &kind, &wfac, &vdir, s->psi);
            double width = wfac * dred;
            if (f == 5 + nd) width = dred * s->psi;         /* d .* psi */
            double R = twopi * width;
            double *re = om_re + (size_t)(c * F + f) * N, *im = om_im + (size_t)(c * F + f) * N;
            for (int idx = 0; idx < N; idx++) {
                int i0 = idx % n0, i1 = (idx / n0) % n1, i2 = idx / (n0 * n1);
                int ii[3] = {i0, i1, i2};
                double w[3] = {0, 0, 0}, wb2 = 0.0;
                for (int a = 0; a < nd; a++) {
                    w[a] = omega_axis(s, a, ii[a]);
                    wb2 += w[a] * w[a];
                }
                double wb = sqrt(wb2);
                double wbR = wb * R;
                re[idx] = 0.0;
                im[idx] = 0.0;
                switch (kind) {
                case K_RHO:
                    re[idx] = 1.0;
                    break;
                case K_S:
                    re[idx] = (wb == 0.0 ? 2.0 * R : 2.0 * sin(wbR) / wb) / twopi;
                    break;
                case K_V:
                    re[idx] = (wb == 0.0 ? fourpi * (R * R * R) / 3.0
                                         : fourpi / (wb * wb * wb) * (sin(wbR) - R * wb * cos(wbR))) / tp3;
                    break;
                case K_VEC:
                    im[idx] = (wb == 0.0 ? fourpi * R
                                         : fourpi * w[vdir] / (wb * wb * wb) * (sin(wbR) - R * wb * cos(wbR))) / tp3;
                    break;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Weighted densities: fields.jl:26-41, weighted_densities.jl:118-138,345-366 */
/* ------------------------------------------------------------------------ */

static void convolve(const or_system *s, const double *src, const double *kre, const double *kim,
                     double *dst, double *br, double *bi)
{
    int N = or_npoints(s);
    for (int j = 0; j < N; j++) { br[j] = src[j]; bi[j] = 0.0; }
    or_fft(br, bi, s->nd, s->ngrid, 0);
    for (int j = 0; j < N; j++) {
        double r = br[j] * kre[j] - bi[j] * kim[j];
        double i = br[j] * kim[j] + bi[j] * kre[j];
        br[j] = r;
        bi[j] = i;
    }
    or_fft(br, bi, s->nd, s->ngrid, 1);
    for (int j = 0; j < N; j++) dst[j] = br[j];
}

static void radial_weighted(const or_system *s, const double *rho, double *n);
static void radial_transposed(const or_system *s, const double *df, double *out);

void or_weighted_densities(const or_system *s, const double *rho, double *n)
{
    if (s->coord) {
        radial_weighted(s, rho, n);
        return;
    }
    int N = or_npoints(s), F = 6 + s->nd;
    double NA = OR_NA * (s->L * s->L * s->L);
    double *ore = malloc(sizeof(double) * (size_t)N * F * s->nc);
    double *oim = malloc(sizeof(double) * (size_t)N * F * s->nc);
    double *br = malloc(sizeof(double) * N), *bi = malloc(sizeof(double) * N);
    or_omega(s, ore, oim);
    for (int c = 0; c < s->nc; c++) {
        const double *rc = rho + (size_t)c * N;
        for (int f = 0; f < F; f++) {
            double *dst = n + (size_t)(c * F + f) * N;
            if (f == 0) {
                for (int j = 0; j < N; j++) dst[j] = rc[j] * NA;
                continue;
            }
            convolve(s, rc, ore + (size_t)(c * F + f) * N, oim + (size_t)(c * F + f) * N, dst, br, bi);
            for (int j = 0; j < N; j++) dst[j] *= NA;
        }
    }
    free(ore); free(oim); free(br); free(bi);
}

/* Transposed convolution: weighted_densities.jl:140-163 (scalar: the rho slot
 * assigns, others add), :368-383 (vector: subtract). */
void or_transposed(const or_system *s, const double *dphi, double *out)
{
    if (s->coord) {
        radial_transposed(s, dphi, out);
        return;
    }
    int N = or_npoints(s), F = 6 + s->nd;
    double *ore = malloc(sizeof(double) * (size_t)N * F * s->nc);
    double *oim = malloc(sizeof(double) * (size_t)N * F * s->nc);
    double *br = malloc(sizeof(double) * N), *bi = malloc(sizeof(double) * N);
    double *tmp = malloc(sizeof(double) * N);
    or_omega(s, ore, oim);
    for (int c = 0; c < s->nc; c++) {
        double *oc = out + (size_t)c * N;
        for (int j = 0; j < N; j++) oc[j] = dphi[(size_t)(c * F + 0) * N + j];
        for (int f = 1; f < F; f++) {
            int kind, vdir;
            double wfac;
            field_desc(s->nd, f, &kind, &wfac, &vdir, s->psi);
            convolve(s, dphi + (size_t)(c * F + f) * N, ore + (size_t)(c * F + f) * N,
                     oim + (size_t)(c * F + f) * N, tmp, br, bi);
            if (kind == K_VEC)
                for (int j = 0; j < N; j++) oc[j] -= tmp[j];
            else
                for (int j = 0; j < N; j++) oc[j] += tmp[j];
        }
    }
    free(ore); free(oim); free(br); free(bi); free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Pointwise PC-SAFT Helmholtz density and its gradient.                     */
/*   f_hs   FMT.jl:17-47                                                     */
/*   f_hc   PCSAFT.jl:123-144                                                */
/*   f_disp PCSAFT.jl:150-197, I_lite :199-211, constants :86-103            */
/*   f_res  PCSAFT.jl:110-117 (association = 0 for these models)             */
/* The gradient is the hand-derived analytic one (the reference obtains it   */
/* with Enzyme, models.jl:355-411).                                          */
/* ------------------------------------------------------------------------ */

static const double CORR1[7][3] = {
    {0.9105631445, -0.3084016918, -0.0906148351},
    {0.6361281449, 0.1860531159, 0.4527842806},
    {2.6861347891, -2.5030047259, 0.5962700728},
    {-26.547362491, 21.419793629, -1.7241829131},
    {97.759208784, -65.25588533, -4.1302112531},
    {-159.59154087, 83.318680481, 13.77663187},
    {91.297774084, -33.74692293, -8.6728470368}};
static const double CORR2[7][3] = {
    {0.7240946941, -0.5755498075, 0.0976883116},
    {2.2382791861, 0.6995095521, -0.2557574982},
    {-4.0025849485, 3.892567339, -9.155856153},
    {-21.003576815, -17.215471648, 20.642075974},
    {26.855641363, 192.67226447, -38.804430052},
    {206.55133841, -161.82646165, 93.626774077},
    {-355.60235612, -165.20769346, -29.666905585}};

/* I(mbar, eta) and its partials (I_lite, PCSAFT.jl:199-211). */
static void I_lite(const double corr[7][3], double mb, double eta, double *I, double *Im, double *Ie)
{
    double m2 = (mb - 1.0) / mb;
    double m3 = m2 * (mb - 2.0) / mb;
    double dm2 = 1.0 / (mb * mb);
    double dm3 = (3.0 * mb - 4.0) / (mb * mb * mb);
    double v = 0.0, vm = 0.0, ve = 0.0, ep = 1.0, epm1 = 0.0;
    for (int q = 0; q < 7; q++) {
        double a = corr[q][0] + m2 * corr[q][1] + m3 * corr[q][2];
        double am = dm2 * corr[q][1] + dm3 * corr[q][2];
        v += a * ep;
        vm += am * ep;
        ve += (double)q * a * epm1;
        epm1 = ep;
        ep *= eta;
    }
    *I = v;
    *Im = vm;
    *Ie = ve;
}

/* ---- Wertheim association (f_assoc, assoc.jl:702-831) ----
 * Per component i: n0_i = S_i/d_i, xi_i = 1 - |nv2_i|^2/n2_i^2 with n2_i = pi S_i d_i,
 * nv2_i = -2 pi v_i (assoc.jl:180-221); w_i = n0_i xi_i is the associating density.
 * Per pair p between sites a and b: Delta_p = g_hs sigma_ab^3 (exp(eps/T) - 1) kappa
 * with g_hs = 1/(1-n3) + d_ab xi n2/(2(1-n3)^2) + d_ab^2 n2^2 xi/(18(1-n3)^3) of the
 * mixture n2, n3, xi (assoc.jl:159-168), d_ab = d_i d_j/(d_i + d_j).
 * Site fractions: X_s = 1/(1 + C_s) with
 * C_s = sum_{p: a_p = s} v_b Delta_p X_b + sum_{p: b_p = s} v_a Delta_p X_a, v_t = n_t w_c(t). */
typedef struct {
    int ns, np;
    double v[OR_MAXS], delta[OR_MAXP];
} assoc_pt;

static double assoc_C(const or_system *s, const assoc_pt *A, const double *X, int site)
{
    double c = 0;
    for (int p = 0; p < A->np; p++) {
        if (s->pair_a[p] == site) c += A->v[s->pair_b[p]] * A->delta[p] * X[s->pair_b[p]];
        if (s->pair_b[p] == site) c += A->v[s->pair_a[p]] * A->delta[p] * X[s->pair_a[p]];
    }
    return c;
}

/* X0 = 0.9, 5 NS blended substitutions, then 5 NS safeguarded Newton steps with
 * Gaussian elimination without pivoting (assoc.jl:238-411). */
static void assoc_solve_general(const or_system *s, const assoc_pt *A, double *X)
{
    int ns = A->ns;
    double C[OR_MAXS], Xn[OR_MAXS], J[OR_MAXS][OR_MAXS + 1];
    for (int t = 0; t < ns; t++) X[t] = 0.9;
    for (int it = 0; it < 5 * ns; it++) {
        for (int t = 0; t < ns; t++) Xn[t] = 0.5 * (X[t] + 1.0 / (1.0 + assoc_C(s, A, X, t)));
        memcpy(X, Xn, sizeof(double) * ns);
    }
    for (int it = 0; it < 5 * ns; it++) {
        for (int t = 0; t < ns; t++) C[t] = assoc_C(s, A, X, t);
        for (int a = 0; a < ns; a++) {
            for (int b = 0; b < ns; b++) J[a][b] = (a == b) ? 1.0 + C[a] : 0.0;
            J[a][ns] = X[a] * (1.0 + C[a]) - 1.0;
        }
        for (int p = 0; p < A->np; p++) {
            int a = s->pair_a[p], b = s->pair_b[p];
            J[a][b] += X[a] * A->v[b] * A->delta[p];
            J[b][a] += X[b] * A->v[a] * A->delta[p];
        }
        for (int k = 0; k < ns; k++)
            for (int i = k + 1; i < ns; i++) {
                double f = J[i][k] / J[k][k];
                for (int j = k + 1; j <= ns; j++) J[i][j] -= f * J[k][j];
            }
        double dX[OR_MAXS];
        for (int k = ns - 1; k >= 0; k--) {
            double r = J[k][ns];
            for (int j = k + 1; j < ns; j++) r -= J[k][j] * dX[j];
            dX[k] = r / J[k][k];
        }
        for (int t = 0; t < ns; t++) {
            double cand = X[t] - dX[t];
            Xn[t] = (cand > 0.0 && cand < 1.0) ? cand : 1.0 / (1.0 + C[t]);
        }
        memcpy(X, Xn, sizeof(double) * ns);
    }
}

/* X for the point's associating densities; one pair between two distinct sites has
 * the closed form of assoc.jl:759-764. */
static void assoc_solve(const or_system *s, const assoc_pt *A, double *X)
{
    if (A->np == 1 && s->pair_a[0] != s->pair_b[0]) {
        int a = s->pair_a[0], b = s->pair_b[0];
        for (int t = 0; t < A->ns; t++) X[t] = 1.0;
        double kia = A->v[a] * A->delta[0], kjb = A->v[b] * A->delta[0];
        double bb = 1.0 - kia + kjb;
        X[a] = 2.0 / (bb + sqrt(bb * bb + 4.0 * kia));
        X[b] = 1.0 / (1.0 + kia * X[a]);
        return;
    }
    assoc_solve_general(s, A, X);
}

/* Phi_assoc = sum_s v_s (ln|X_s| - X_s/2 + 1/2) (assoc.jl:811-831) and, when dphi is
 * given, its gradient added into dphi.  At the solution of the mass-action law,
 * dPhi/dw_i = sum_{s in i} n_s ln X_s and dPhi/dDelta_p = -v_a X_a v_b X_b (the
 * stationarity of Michelsen and Hendriks' Q function; the reference obtains the same
 * derivative with an implicit-function-theorem rule, assoc.jl:833-916). */
static double or_assoc(const or_system *s, const double *nl, double n2m, double n3m, const double *nv2m,
                       double *dphi, double *Xout)
{
    int nc = s->nc, nd = s->nd, F = 6 + nd, ns = s->nsite;
    const double twopi = 2.0 * PI;
    double w[OR_MAXC], dwS[OR_MAXC], dwv[OR_MAXC][3];
    for (int i = 0; i < nc; i++) {
        double di = s->d[i] / s->L, S = nl[i * F + 1];
        double n0 = S / di, n2 = PI * S * di, sq = 0, nv[3] = {0, 0, 0};
        for (int j = 0; j < nd; j++) { nv[j] = -twopi * nl[i * F + 3 + j]; sq += nv[j] * nv[j]; }
        double xi = 1.0 - sq / (n2 * n2);
        w[i] = n0 * xi;
        dwS[i] = xi / di + n0 * (2.0 * sq / (n2 * n2 * n2)) * PI * di;
        for (int j = 0; j < 3; j++) dwv[i][j] = n0 * (-2.0 * nv[j] / (n2 * n2)) * (-twopi);
    }
    double nv2sq = 0;
    for (int j = 0; j < nd; j++) nv2sq += nv2m[j] * nv2m[j];
    double xim = 1.0 - nv2sq / (n2m * n2m);
    double a1 = 1.0 / (1.0 - n3m);
    assoc_pt A;
    A.ns = ns;
    A.np = s->npair;
    double cpair[OR_MAXP], dij[OR_MAXP];
    for (int t = 0; t < ns; t++) A.v[t] = s->site_count[t] * w[s->site_comp[t]];
    for (int p = 0; p < A.np; p++) {
        int ic = s->site_comp[s->pair_a[p]], jc = s->site_comp[s->pair_b[p]];
        double sr = s->sigma[ic * nc + jc] / s->L, di = s->d[ic] / s->L, dj = s->d[jc] / s->L;
        cpair[p] = sr * sr * sr * (exp(s->pair_eps[p] / s->T) - 1.0) * s->pair_kappa[p];
        dij[p] = di * dj / (di + dj);
        double g = a1 + 0.5 * dij[p] * xim * n2m * a1 * a1 + dij[p] * dij[p] * n2m * n2m * xim * a1 * a1 * a1 / 18.0;
        A.delta[p] = g * cpair[p];
    }
    double X[OR_MAXS];
    assoc_solve(s, &A, X);
    if (Xout) memcpy(Xout, X, sizeof(double) * ns);
    double phi = 0;
    for (int t = 0; t < ns; t++) phi += A.v[t] * (log(fabs(X[t])) - 0.5 * X[t] + 0.5);
    if (!dphi) return phi;

    double dw[OR_MAXC] = {0}, dn3 = 0, dn2 = 0, dxi = 0;
    for (int t = 0; t < ns; t++) dw[s->site_comp[t]] += s->site_count[t] * log(fabs(X[t]));
    for (int p = 0; p < A.np; p++) {
        int a = s->pair_a[p], b = s->pair_b[p];
        double dg = -A.v[a] * X[a] * A.v[b] * X[b] * cpair[p];
        double dd = dij[p];
        dn3 += dg * (a1 * a1 + dd * xim * n2m * a1 * a1 * a1 + dd * dd * n2m * n2m * xim * a1 * a1 * a1 * a1 / 6.0);
        dn2 += dg * (0.5 * dd * xim * a1 * a1 + dd * dd * n2m * xim * a1 * a1 * a1 / 9.0);
        dxi += dg * (0.5 * dd * n2m * a1 * a1 + dd * dd * n2m * n2m * a1 * a1 * a1 / 18.0);
    }
    dn2 += dxi * 2.0 * nv2sq / (n2m * n2m * n2m);
    for (int i = 0; i < nc; i++) {
        double mi = s->m[i], di = s->d[i] / s->L;
        dphi[i * F + 1] += dn2 * PI * mi * di + dw[i] * dwS[i];
        dphi[i * F + 2] += dn3 * mi;
        for (int j = 0; j < nd; j++)
            dphi[i * F + 3 + j] += dxi * (-2.0 * nv2m[j] / (n2m * n2m)) * (-twopi * mi) + dw[i] * dwv[i][j];
    }
    return phi;
}

double or_phi(const or_system *s, const double *nl, double *dphi)
{
    int nc = s->nc, nd = s->nd, F = 6 + nd;
    double dd[OR_MAXC], m[OR_MAXC];
    for (int i = 0; i < nc; i++) { dd[i] = s->d[i] / s->L; m[i] = s->m[i]; }
    if (dphi) memset(dphi, 0, sizeof(double) * F * nc);

    /* ---- hard sphere (White Bear) ---- */
    double n0 = 0, n1 = 0, n2 = 0, n3 = 0, nv1[3] = {0, 0, 0}, nv2[3] = {0, 0, 0};
    const double twopi = 2.0 * PI;
    for (int i = 0; i < nc; i++) {
        double mi = m[i], di = dd[i];
        double nim = nl[i * F + 1] * mi;
        n0 += nim / di;
        n1 += 0.5 * nim;
        n2 += PI * nim * di;
        for (int j = 0; j < nd; j++) {
            double nvi = nl[i * F + 3 + j];
            nv1[j] += nvi * mi / di;
            nv2[j] -= twopi * nvi * mi;
        }
        n3 += nl[i * F + 2] * mi;
    }
    double nv1v2 = nv1[0] * nv2[0] + nv1[1] * nv2[1] + nv1[2] * nv2[2];
    double nv2sq = nv2[0] * nv2[0] + nv2[1] * nv2[1] + nv2[2] * nv2[2];
    const double twelvepi = 12.0 * PI;
    double dn = 1.0 - n3;
    double log1dn = log(fabs(dn));
    double inv1dn = 1.0 / dn;
    /* bkt = ln(1-n3)/(12 pi n3^2) + 1/(12 pi n3 (1-n3)^2) (FMT.jl:40-41) cancels two
     * O(1/n3) terms, so its rounding error grows as eps/n3^2 (its derivative's as
     * eps/n3^3).  Below |n3| = 1e-3 the same function is evaluated from its Taylor
     * series, sum_k (k+1)(k+3)/(k+2) n3^k / (12 pi), to FP64 accuracy. */
    double bkt, dB;
    if (fabs(n3) < 1e-3) {
        const double c[7] = {1.5, 8.0 / 3.0, 15.0 / 4.0, 24.0 / 5.0, 35.0 / 6.0, 48.0 / 7.0, 63.0 / 8.0};
        bkt = 0.0;
        dB = 0.0;
        for (int k = 6; k >= 0; k--) bkt = bkt * n3 + c[k];
        for (int k = 6; k >= 1; k--) dB = dB * n3 + k * c[k];
        bkt /= twelvepi;
        dB /= twelvepi;
    } else {
        bkt = log1dn / (twelvepi * n3 * n3) + 1.0 / (twelvepi * n3 * dn * dn);
        dB = (-1.0 / (n3 * n3 * dn) - 2.0 * log1dn / (n3 * n3 * n3) + (3.0 * n3 - 1.0) / (n3 * n3 * dn * dn * dn)) /
             twelvepi;
    }
    double phi_hs = -n0 * log1dn + (n1 * n2 - nv1v2) * inv1dn + (n2 * n2 * n2 / 3.0 - n2 * nv2sq) * bkt;

    if (dphi) {
        double d_n0 = -log1dn;
        double d_n1 = n2 * inv1dn;
        double d_n2 = n1 * inv1dn + (n2 * n2 - nv2sq) * bkt;
        double d_n3 = n0 * inv1dn + (n1 * n2 - nv1v2) * inv1dn * inv1dn + (n2 * n2 * n2 / 3.0 - n2 * nv2sq) * dB;
        double d_nv1[3], d_nv2[3];
        for (int j = 0; j < 3; j++) {
            d_nv1[j] = -nv2[j] * inv1dn;
            d_nv2[j] = -nv1[j] * inv1dn - 2.0 * n2 * nv2[j] * bkt;
        }
        for (int i = 0; i < nc; i++) {
            double mi = m[i], di = dd[i];
            dphi[i * F + 1] += mi * (d_n0 / di + 0.5 * d_n1 + PI * di * d_n2);
            dphi[i * F + 2] += mi * d_n3;
            for (int j = 0; j < nd; j++) dphi[i * F + 3 + j] += mi * d_nv1[j] / di - twopi * mi * d_nv2[j];
        }
    }

    double phi_assoc = s->npair > 0 ? or_assoc(s, nl, n2, n3, nv2, dphi, NULL) : 0.0;

    /* ---- hard chain (Sauer-Gross) ---- */
    int iz = 3 + nd, il = 4 + nd, idp = 5 + nd;
    double z3 = 0, z2 = 0;
    for (int i = 0; i < nc; i++) {
        double rh = nl[i * F + iz];
        z3 += m[i] * rh;
        z2 += m[i] * rh / dd[i];
    }
    z3 /= 8.0;
    z2 /= 8.0;
    double inv = 1.0 / (1.0 - z3);
    double phi_hc = 0.0, G3 = 0.0, G2 = 0.0;
    for (int i = 0; i < nc; i++) {
        double rhoi = nl[i * F + 0];
        double lam = nl[i * F + il] / (2.0 * dd[i]);
        double y = inv + 1.5 * dd[i] * z2 * inv * inv + 0.5 * dd[i] * dd[i] * z2 * z2 * inv * inv * inv;
        double lg = log(fabs(y * lam / rhoi));
        phi_hc += -rhoi * (m[i] - 1.0) * lg;
        if (dphi) {
            dphi[i * F + 0] += -(m[i] - 1.0) * (lg - 1.0);
            dphi[i * F + il] += -rhoi * (m[i] - 1.0) / nl[i * F + il];
            double gy = -rhoi * (m[i] - 1.0) / y;
            double dy3 = inv * inv + 3.0 * dd[i] * z2 * inv * inv * inv + 1.5 * dd[i] * dd[i] * z2 * z2 * inv * inv * inv * inv;
            double dy2 = 1.5 * dd[i] * inv * inv + dd[i] * dd[i] * z2 * inv * inv * inv;
            G3 += gy * dy3;
            G2 += gy * dy2;
        }
    }
    if (dphi)
        for (int k = 0; k < nc; k++) dphi[k * F + iz] += m[k] / 8.0 * G3 + m[k] / (8.0 * dd[k]) * G2;

    /* ---- dispersion (psi-averaged density; the reference fixes psi = 1.3862,
     * PCSAFT.jl:53,160, and the descriptor carries it) ---- */
    double psi = s->psi;
    double factor = 3.0 / (4.0 * psi * psi * psi * PI);
    double rb[OR_MAXC], S = 0, mt = 0, es = 0;
    for (int i = 0; i < nc; i++) {
        rb[i] = nl[i * F + idp] * factor / (dd[i] * dd[i] * dd[i]);
        S += rb[i];
        mt += rb[i] * m[i];
        es += m[i] * rb[i] * dd[i] * dd[i] * dd[i];
    }
    double mb = mt / S;
    double eta = es * PI / 6.0;
    double A1 = 0, A2 = 0, dA1[OR_MAXC], dA2[OR_MAXC];
    for (int i = 0; i < nc; i++) { dA1[i] = 0; dA2[i] = 0; }
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            double sij = s->sigma[i * nc + j] / s->L;
            double cm = m[i] * m[j] * sij * sij * sij;
            double eT = s->eps[i * nc + j] / s->T;
            double cij = rb[i] * rb[j] * cm;
            A1 += cij * eT;
            A2 += cij * eT * eT;
            dA1[i] += 2.0 * rb[j] * cm * eT;
            dA2[i] += 2.0 * rb[j] * cm * eT * eT;
        }
    double e2 = eta * eta, om = 1.0 - eta, om4 = om * om * om * om;
    double i1e = 1.0 / (1.0 - eta), i2e = 1.0 / (2.0 - eta);
    double Pn = 8.0 * eta - 2.0 * e2;
    double Qn = 20.0 * eta - 27.0 * e2 + 12.0 * (eta * e2) - 2.0 * (e2 * e2);
    double C1 = 1.0 + mb * Pn / om4 + (1.0 - mb) * Qn * i1e * i1e * i2e * i2e;
    double I1, I1m, I1e, I2, I2m, I2e;
    I_lite(CORR1, mb, eta, &I1, &I1m, &I1e);
    I_lite(CORR2, mb, eta, &I2, &I2m, &I2e);
    double phi_disp = -2.0 * (PI * I1 * A1) - PI * mb * I2 * A2 / C1;

    if (dphi) {
        double P = Pn / om4, Q = Qn * i1e * i1e * i2e * i2e;
        double Pe = (8.0 - 4.0 * eta) / om4 + 4.0 * Pn / (om4 * om);
        double Qe = (20.0 - 54.0 * eta + 36.0 * e2 - 8.0 * eta * e2) * i1e * i1e * i2e * i2e +
                    Qn * (2.0 * i1e * i1e * i1e * i2e * i2e + 2.0 * i1e * i1e * i2e * i2e * i2e);
        double C1m = P - Q, C1e = mb * Pe + (1.0 - mb) * Qe;
        for (int k = 0; k < nc; k++) {
            double dm = (m[k] - mb) / S;
            double de = PI / 6.0 * m[k] * dd[k] * dd[k] * dd[k];
            double g = -2.0 * PI * ((I1m * dm + I1e * de) * A1 + I1 * dA1[k]);
            g -= PI * (dm * I2 * A2 / C1 + mb * (I2m * dm + I2e * de) * A2 / C1 + mb * I2 * dA2[k] / C1 -
                       mb * I2 * A2 * (C1m * dm + C1e * de) / (C1 * C1));
            dphi[k * F + idp] += g * factor / (dd[k] * dd[k] * dd[k]);
        }
    }
    return phi_hs + phi_hc + phi_disp + phi_assoc;
}

void or_assoc_X(const or_system *s, const double *nl, double *X)
{
    int nc = s->nc, nd = s->nd, F = 6 + nd;
    double n2 = 0, n3 = 0, nv2[3] = {0, 0, 0};
    for (int i = 0; i < nc; i++) {
        double di = s->d[i] / s->L;
        n2 += PI * nl[i * F + 1] * s->m[i] * di;
        n3 += nl[i * F + 2] * s->m[i];
        for (int j = 0; j < nd; j++) nv2[j] -= 2.0 * PI * nl[i * F + 3 + j] * s->m[i];
    }
    for (int t = 0; t < s->nsite; t++) X[t] = 1.0;
    if (s->npair > 0) or_assoc(s, nl, n2, n3, nv2, NULL, X);
}

/* ------------------------------------------------------------------------ */
/* Radial (spherical / cylindrical) geometries: base/structure.jl:363-433,     */
/* weighted_densities.jl:92-116 (scalar), :249-343 (vector), integrals.jl:     */
/* 127-137 (convolve! with a QDHT), transforms.jl:35-63 (derivative matrix),   */
/* integrals.jl:20-22 (int = 4 pi / 2 pi integrateR).  The transform itself is  */
/* Hankel.jl's QDHT (a dependency absent here), restated from its published    */
/* algorithm (Guizar-Sicairos & Gutierrez-Vega 2004).                          */
/* ------------------------------------------------------------------------ */

static double bessel_jp(int sph, double x)       /* J_0, or J_{1/2} for the spherical transform */
{
    return sph ? sqrt(2.0 / (PI * x)) * sin(x) : j0(x);
}

static double bessel_jp1(int sph, double x)      /* J_1, or J_{3/2} */
{
    return sph ? sqrt(2.0 / (PI * x)) * (sin(x) / x - cos(x)) : j1(x);
}

static double bessel_zero(int sph, int s)        /* s-th positive zero of J_0 or J_{1/2} */
{
    if (sph) return s * PI;
    double b = (s - 0.25) * PI;                  /* McMahon's expansion, then Newton */
    double x = b + 1.0 / (8.0 * b) - 124.0 / (3.0 * pow(8.0 * b, 3));
    for (int it = 0; it < 50; it++) {
        double dx = j0(x) / j1(x);               /* J0' = -J1 */
        x += dx;
        if (fabs(dx) < 1e-15 * x) break;
    }
    return x;
}

or_qdht *or_qdht_create(int coord, double R, int N)
{
    or_qdht *q = calloc(1, sizeof(or_qdht));
    q->N = N;
    q->sph = (coord == 1);
    q->R = R;
    double *zr = malloc(sizeof(double) * (N + 1));
    for (int n = 0; n < N + 1; n++) zr[n] = bessel_zero(q->sph, n + 1);
    q->S = zr[N];
    q->K = q->S / R;
    q->r = malloc(sizeof(double) * N);
    q->k = malloc(sizeof(double) * N);
    q->jp1 = malloc(sizeof(double) * N);
    q->w = malloc(sizeof(double) * N);
    q->C = malloc(sizeof(double) * (size_t)N * N);
    for (int n = 0; n < N; n++) {
        q->r[n] = zr[n] * R / q->S;
        q->k[n] = zr[n] / R;
        q->jp1[n] = fabs(bessel_jp1(q->sph, zr[n]));
    }
    for (int m = 0; m < N; m++)
        for (int n = 0; n < N; n++)
            q->C[(size_t)m * N + n] = 2.0 / q->S * bessel_jp(q->sph, zr[m] * zr[n] / q->S) / (q->jp1[m] * q->jp1[n]);
    /* the transform's quadrature int_0^R g(r) r dr = sum g(r_n) 2 / (K^2 J_{p+1}(j_n)^2),
     * with g = f (cylindrical) or g = f r (spherical), times the angular factor */
    for (int n = 0; n < N; n++) {
        double wq = 2.0 / (q->K * q->K * q->jp1[n] * q->jp1[n]);
        q->w[n] = q->sph ? 4.0 * PI * wq * q->r[n] : 2.0 * PI * wq;
    }
    free(zr);
    return q;
}

void or_qdht_free(or_qdht *q)
{
    if (!q) return;
    free(q->r); free(q->k); free(q->jp1); free(q->w); free(q->C); free(q);
}

/* forward: F(k_m) = int f j0(k r) r^2 dr (spherical) or int f J0(k r) r dr (cylindrical),
 * from the samples: g = f (times sqrt(r) for the order-1/2 form), scaled by R / |J_{p+1}|,
 * multiplied by C, scaled back by |J_{p+1}| / K (times sqrt(pi / 2k)). */
void or_qdht_forward(const or_qdht *q, const double *f, double *F)
{
    int N = q->N;
    double *a = malloc(sizeof(double) * N);
    for (int n = 0; n < N; n++) a[n] = f[n] * (q->sph ? sqrt(q->r[n]) : 1.0) * q->R / q->jp1[n];
    for (int m = 0; m < N; m++) {
        double acc = 0.0;
        for (int n = 0; n < N; n++) acc += q->C[(size_t)m * N + n] * a[n];
        F[m] = acc * q->jp1[m] / q->K * (q->sph ? sqrt(PI / (2.0 * q->k[m])) : 1.0);
    }
    free(a);
}

void or_qdht_backward(const or_qdht *q, const double *F, double *f)
{
    int N = q->N;
    double *a = malloc(sizeof(double) * N);
    for (int m = 0; m < N; m++) a[m] = F[m] / (q->sph ? sqrt(PI / (2.0 * q->k[m])) : 1.0) * q->K / q->jp1[m];
    for (int n = 0; n < N; n++) {
        double acc = 0.0;
        for (int m = 0; m < N; m++) acc += q->C[(size_t)n * N + m] * a[m];
        f[n] = acc * q->jp1[n] / q->R / (q->sph ? sqrt(q->r[n]) : 1.0);
    }
    free(a);
}

/* convolve!(result, profile, kernel, P::QDHT, iP::QDHT, buf), integrals.jl:127-137 */
static void radial_convolve(const or_qdht *q, const double *src, const double *kern, double *dst)
{
    int N = q->N;
    double *t = malloc(sizeof(double) * N);
    or_qdht_forward(q, src, t);
    for (int m = 0; m < N; m++) t[m] *= kern[m];
    or_qdht_backward(q, t, dst);
    free(t);
}

/* radial_derivative_matrix(r), transforms.jl:43-63 (row-major D[i*N+j]) */
static void radial_derivative_matrix(const double *r, int N, double *D)
{
    memset(D, 0, sizeof(double) * (size_t)N * N);
    for (int i = 1; i < N - 1; i++) {
        double h1 = r[i] - r[i - 1], h2 = r[i + 1] - r[i];
        D[(size_t)i * N + i - 1] = -h2 / (h1 * (h1 + h2));
        D[(size_t)i * N + i] = (h2 - h1) / (h1 * h2);
        D[(size_t)i * N + i + 1] = h1 / (h2 * (h1 + h2));
    }
    double h1 = r[1] - r[0], h2 = r[2] - r[1];
    D[0] = -(2 * h1 + h2) / (h1 * (h1 + h2));
    D[1] = (h1 + h2) / (h1 * h2);
    D[2] = -h1 / (h2 * (h1 + h2));
    h1 = r[N - 2] - r[N - 3];
    h2 = r[N - 1] - r[N - 2];
    D[(size_t)(N - 1) * N + N - 3] = h2 / (h1 * (h1 + h2));
    D[(size_t)(N - 1) * N + N - 2] = -(h1 + h2) / (h1 * h2);
    D[(size_t)(N - 1) * N + N - 1] = (2 * h2 + h1) / (h2 * (h1 + h2));
}

/* The radial kernels of component c: scalar slots sampled at omega = k / 2pi in
 * reduced units (weighted_densities.jl:92-116, no k = 0 sample), and the vector
 * slot's operator T = D Q^-1 H Q materialised column by column in unscaled units
 * (weighted_densities.jl:290-316); kern[f*N + m], T[i*N + j]. */
static void radial_kernels(const or_system *s, const or_qdht *q, int c, double *kern, double *T)
{
    int N = q->N, F = 7;
    const double twopi = 2.0 * PI, fourpi = 4.0 * PI, tp3 = twopi * twopi * twopi;
    double dred = s->d[c] / s->L;
    for (int f = 1; f < F; f++) {
        int kind, vdir;
        double wfac;
        field_desc(1, f, &kind, &wfac, &vdir, s->psi);
        double width = (f == 6) ? dred * s->psi : wfac * dred;
        double R = twopi * width;
        for (int m = 0; m < N; m++) {
            double wb = q->k[m] / twopi * s->L, wbR = wb * R;
            double v = 0.0;
            if (kind == K_S) v = 2.0 * sin(wbR) / wb / twopi;
            else if (kind == K_V || kind == K_VEC) v = fourpi / (wb * wb * wb) * (sin(wbR) - R * wb * cos(wbR)) / tp3;
            kern[(size_t)f * N + m] = v;
        }
    }
    double Rk = twopi * 0.5 * dred * s->L;
    double *H = malloc(sizeof(double) * N), *e = calloc(N, sizeof(double)), *col = malloc(sizeof(double) * N);
    double *M = malloc(sizeof(double) * (size_t)N * N), *D = malloc(sizeof(double) * (size_t)N * N);
    for (int m = 0; m < N; m++) {
        double wb = q->k[m] / twopi;
        H[m] = fourpi / (wb * wb * wb) * (sin(wb * Rk) - Rk * wb * cos(wb * Rk)) / tp3;
    }
    for (int k = 0; k < N; k++) {
        e[k] = 1.0;
        radial_convolve(q, e, H, col);
        e[k] = 0.0;
        for (int i = 0; i < N; i++) M[(size_t)i * N + k] = col[i];
    }
    radial_derivative_matrix(q->r, N, D);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            double acc = 0.0;
            for (int l = 0; l < N; l++) acc += D[(size_t)i * N + l] * M[(size_t)l * N + j];
            T[(size_t)i * N + j] = acc;
        }
    free(H); free(e); free(col); free(M); free(D);
}

/* The radial operators of every component: kernels kern[(c F + f) N + m] and the
 * vector slot's T[c N N + i N + j]. */
static or_qdht *radial_ops(const or_system *s, double **kern, double **T)
{
    int N = s->ngrid[0], F = 7, nc = s->nc;
    or_qdht *q = or_qdht_create(s->coord, s->ub[0], N);
    *kern = malloc(sizeof(double) * (size_t)F * N * nc);
    *T = malloc(sizeof(double) * (size_t)N * N * nc);
    for (int c = 0; c < nc; c++) radial_kernels(s, q, c, *kern + (size_t)c * F * N, *T + (size_t)c * N * N);
    return q;
}

/* evaluate_field! for QDHT plans (weighted_densities.jl:118-138, :318-328) */
static void radial_weighted(const or_system *s, const double *rho, double *n)
{
    int N = s->ngrid[0], F = 7, nc = s->nc;
    double NA = OR_NA * (s->L * s->L * s->L);
    double *kern, *T;
    or_qdht *q = radial_ops(s, &kern, &T);
    for (int c = 0; c < nc; c++) {
        const double *rc = rho + (size_t)c * N;
        const double *Tc = T + (size_t)c * N * N;
        for (int f = 0; f < F; f++) {
            double *dst = n + (size_t)(c * F + f) * N;
            if (f == 0) {
                for (int j = 0; j < N; j++) dst[j] = rc[j] * NA;
            } else if (f == 3) {
                for (int i = 0; i < N; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < N; j++) acc += Tc[(size_t)i * N + j] * rc[j];
                    dst[i] = acc * NA;
                }
            } else {
                radial_convolve(q, rc, kern + (size_t)(c * F + f) * N, dst);
                for (int j = 0; j < N; j++) dst[j] *= NA;
            }
        }
    }
    or_qdht_free(q);
    free(kern); free(T);
}

/* integrate_field! for QDHT plans (weighted_densities.jl:140-163, :330-343): the rho
 * slot assigns, the scalar slots add the same convolution, the vector slot adds
 * transpose(T) dPhi/dn_v. */
static void radial_transposed(const or_system *s, const double *df, double *out)
{
    int N = s->ngrid[0], F = 7, nc = s->nc;
    double *kern, *T;
    or_qdht *q = radial_ops(s, &kern, &T);
    double *tmp = malloc(sizeof(double) * N);
    for (int c = 0; c < nc; c++) {
        double *oc = out + (size_t)c * N;
        const double *Tc = T + (size_t)c * N * N;
        for (int j = 0; j < N; j++) oc[j] = df[(size_t)(c * F) * N + j];
        for (int f = 1; f < F; f++) {
            const double *u = df + (size_t)(c * F + f) * N;
            if (f == 3) {
                for (int j = 0; j < N; j++) {
                    double acc = 0.0;
                    for (int i = 0; i < N; i++) acc += Tc[(size_t)i * N + j] * u[i];
                    oc[j] += acc;
                }
            } else {
                radial_convolve(q, u, kern + (size_t)(c * F + f) * N, tmp);
                for (int j = 0; j < N; j++) oc[j] += tmp[j];
            }
        }
    }
    or_qdht_free(q);
    free(kern); free(T); free(tmp);
}

/* int f dV on the structure: Simpson (Cartesian) or 4 pi / 2 pi integrateR (radial) */
static double or_integrate(const or_system *s, const double *f)
{
    if (s->coord) {
        or_qdht *q = or_qdht_create(s->coord, s->ub[0], s->ngrid[0]);
        double acc = 0.0;
        for (int n = 0; n < q->N; n++) acc += q->w[n] * f[n];
        or_qdht_free(q);
        return acc;
    }
    double dz[3];
    for (int a = 0; a < s->nd; a++) dz[a] = (s->ub[a] - s->lb[a]) / s->ngrid[a];   /* structure.jl:187-195 */
    return or_simpson(f, s->nd, s->ngrid, dz);
}

/* delta F_res / delta rho (models.jl:131-186): fields, pointwise gradient, transposed. */
void or_dFdrho_res(const or_system *s, const double *rho, double *dfdrho, double *fval)
{
    int N = or_npoints(s), F = 6 + s->nd, nc = s->nc;
    double *n = malloc(sizeof(double) * (size_t)N * F * nc);
    double *df = malloc(sizeof(double) * (size_t)N * F * nc);
    double nl[OR_MAXC * 9], dl[OR_MAXC * 9];
    or_weighted_densities(s, rho, n);
    for (int j = 0; j < N; j++) {
        for (int c = 0; c < nc; c++)
            for (int f = 0; f < F; f++) nl[c * F + f] = n[(size_t)(c * F + f) * N + j];
        double v = or_phi(s, nl, dl);
        if (fval) fval[j] = v;
        for (int c = 0; c < nc; c++)
            for (int f = 0; f < F; f++) df[(size_t)(c * F + f) * N + j] = dl[c * F + f];
    }
    or_transposed(s, df, dfdrho);
    free(n);
    free(df);
}

/* Simpson rule exactly as _int (integrals.jl:28-40): per-axis weights 1,4,2,...,
 * with both ends 1, times prod(dz/3). */
double or_simpson(const double *f, int nd, const int *ngrid, const double *dz)
{
    int N = 1;
    for (int a = 0; a < nd; a++) N *= ngrid[a];
    double sum = 0.0;
    for (int idx = 0; idx < N; idx++) {
        int rem = idx;
        double coef = 1.0;
        for (int a = 0; a < nd; a++) {
            int k = rem % ngrid[a] + 1;   /* 1-based */
            rem /= ngrid[a];
            int end = (k == 1 || k == ngrid[a]);
            coef *= end ? 1.0 : ((k % 2 == 0) ? 4.0 : 2.0);
        }
        sum += coef * f[idx];
    }
    double sc = 1.0;
    for (int a = 0; a < nd; a++) sc *= dz[a] / 3.0;
    return sum * sc;
}


/* F_res = int f_val / L^3 (models.jl:86-91, :73) */
double or_F_res(const or_system *s, const double *rho)
{
    int N = or_npoints(s);
    double *fv = malloc(sizeof(double) * N);
    double *tmp = malloc(sizeof(double) * (size_t)N * s->nc);
    or_dFdrho_res(s, rho, tmp, fv);
    double r = or_integrate(s, fv) / (s->L * s->L * s->L);
    free(fv);
    free(tmp);
    return r;
}

/* F_ideal (BasicIdeal.jl:10-39) */
double or_F_ideal(const or_system *s, const double *rho)
{
    int N = or_npoints(s);
    double *phi = malloc(sizeof(double) * N);
    double lnT = log(s->T);
    for (int j = 0; j < N; j++) {
        double acc = 0.0;
        for (int c = 0; c < s->nc; c++) {
            double n = rho[(size_t)c * N + j];
            acc += OR_NA * n * (log(n) - 1.5 * lnT - 1.0);
        }
        phi[j] = acc;
    }
    double r = or_integrate(s, phi);
    free(phi);
    return r;
}

/* get_new_profile! (converge.jl:108-153): lnG = (mu_res + ln rho_b) - dF/drho - beta V,
 * then clamp to [-100, 100]. */
void or_get_new_profile(const or_system *s, const double *rho, double *lnG)
{
    int N = or_npoints(s);
    or_dFdrho_res(s, rho, lnG, NULL);
    if (s->vext)
        for (size_t j = 0; j < (size_t)N * s->nc; j++) lnG[j] += s->vext[j];   /* steele.jl:79 */
    for (int c = 0; c < s->nc; c++) {
        double cp = s->mu_res[c] + log(s->rho_bulk[c]);
        for (int j = 0; j < N; j++) {
            double v = cp - lnG[(size_t)c * N + j];
            if (v < -100.0) v = -100.0;
            if (v > 100.0) v = 100.0;
            lnG[(size_t)c * N + j] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Anderson acceleration with Picard warm-up: anderson.jl:301-443 and the    */
/* helpers at :450-583, :665-737, :769-776.                                  */
/* ------------------------------------------------------------------------ */

static double nrm2(const double *x, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

static double dot(const double *x, const double *y, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* 2-norm condition number by one-sided Jacobi SVD (cond(RA), anderson.jl:412). */
static double cond2(const double *R, int ld, int k)
{
    double A[8 * 8];
    for (int j = 0; j < k; j++)
        for (int i = 0; i < k; i++) A[j * 8 + i] = R[j * ld + i];  /* column j */
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < k - 1; p++)
            for (int q = p + 1; q < k; q++) {
                double a = 0, b = 0, g = 0;
                for (int i = 0; i < k; i++) {
                    a += A[p * 8 + i] * A[p * 8 + i];
                    b += A[q * 8 + i] * A[q * 8 + i];
                    g += A[p * 8 + i] * A[q * 8 + i];
                }
                if (g == 0.0) continue;
                double rel = fabs(g) / sqrt(a * b);
                if (rel > off) off = rel;
                if (rel < 1e-17) continue;
                double zeta = (b - a) / (2.0 * g);
                double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < k; i++) {
                    double x = A[p * 8 + i], y = A[q * 8 + i];
                    A[p * 8 + i] = c * x - sn * y;
                    A[q * 8 + i] = sn * x + c * y;
                }
            }
        if (off < 1e-17) break;
    }
    double smax = 0.0, smin = INFINITY;
    for (int j = 0; j < k; j++) {
        double sv = 0;
        for (int i = 0; i < k; i++) sv += A[j * 8 + i] * A[j * 8 + i];
        sv = sqrt(sv);
        if (sv > smax) smax = sv;
        if (sv < smin) smin = sv;
    }
    if (smax == 0.0) return INFINITY;
    return smax / smin;
}

/* Householder QR (LAPACK dgeqrf/dlarfg conventions) of the m x (m-1) block
 * R[:, 2:m]; returns Rd ((m-1)x(m-1), upper) and the thin Qx (m x (m-1)).
 * R is column-major with leading dimension m. (downdate_aaqr!, anderson.jl:552-583) */
static void householder_qr(const double *R, int m, double *Rd, double *Qx)
{
    int nr = m, ncol = m - 1;
    double A[8 * 8], tau[8], v[8];
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) A[j * 8 + i] = R[(j + 1) * m + i];
    for (int j = 0; j < ncol; j++) {
        double alpha = A[j * 8 + j];
        double xn = 0.0;
        for (int i = j + 1; i < nr; i++) xn += A[j * 8 + i] * A[j * 8 + i];
        xn = sqrt(xn);
        if (xn == 0.0) {
            tau[j] = 0.0;
            continue;
        }
        double beta = -copysign(hypot(alpha, xn), alpha);
        tau[j] = (beta - alpha) / beta;
        double sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i < nr; i++) A[j * 8 + i] *= sc;
        A[j * 8 + j] = beta;
        /* apply H = I - tau v v^T to the trailing columns */
        for (int jj = j + 1; jj < ncol; jj++) {
            double w = A[jj * 8 + j];
            for (int i = j + 1; i < nr; i++) w += A[j * 8 + i] * A[jj * 8 + i];
            w *= tau[j];
            A[jj * 8 + j] -= w;
            for (int i = j + 1; i < nr; i++) A[jj * 8 + i] -= w * A[j * 8 + i];
        }
    }
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < ncol; i++) Rd[j * ncol + i] = (i <= j) ? A[j * 8 + i] : 0.0;
    /* Q = H_0 ... H_{k-1} [I; 0] (thin), applied in reverse order (dorg2r) */
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) Qx[j * nr + i] = (i == j) ? 1.0 : 0.0;
    for (int h = ncol - 1; h >= 0; h--) {
        if (tau[h] == 0.0) continue;
        v[h] = 1.0;
        for (int i = h + 1; i < nr; i++) v[i] = A[h * 8 + i];
        for (int j = 0; j < ncol; j++) {
            double w = 0.0;
            for (int i = h; i < nr; i++) w += v[i] * Qx[j * nr + i];
            w *= tau[h];
            for (int i = h; i < nr; i++) Qx[j * nr + i] -= w * v[i];
        }
    }
}

int or_aasol(or_gfix G, void *ctx, int n, const double *x0, const or_aa_opts *o, double *sol_out,
             double *gx_out, double *hist, int hist_cap, double *condhist, double *alphanorm,
             int stats_cap, or_aa_result *res_out)
{
    int m = o->m;
    double beta = o->beta;
    if (!(fabs(beta) > 0 && fabs(beta) <= 1)) return -1;
    size_t N = (size_t)n;
    double *sol = malloc(sizeof(double) * N), *gx = malloc(sizeof(double) * N);
    double *df = malloc(sizeof(double) * N), *dg = malloc(sizeof(double) * N);
    double *res = malloc(sizeof(double) * N);
    int mm = m > 0 ? m : 1;
    double *QP = calloc(N * mm, sizeof(double)), *DG = calloc(N * mm, sizeof(double));
    double *Qd = calloc(N * mm, sizeof(double));
    double RP[64] = {0}, Rd[64], Qx[64], theta[8], tres[8], alpha[9] = {0};
    int nh = 0, ns = 0, nevals = 0;
    memcpy(sol, x0, sizeof(double) * N);
    memset(gx, 0, sizeof(double) * N);
    memset(df, 0, sizeof(double) * N);
    memset(dg, 0, sizeof(double) * N);
    memset(res, 0, sizeof(double) * N);

#define EVAL(out, x) (G((out), (x), ctx), nevals++)
#define BETAFIX(b) do { if ((b) != 1.0) for (size_t i_ = 0; i_ < N; i_++) gx[i_] = (1.0 - (b)) * sol[i_] + (b) * gx[i_]; } while (0)
#define HIST(v) do { if (nh < hist_cap) hist[nh] = (v); nh++; } while (0)

    int k = 0;
    EVAL(gx, sol);
    BETAFIX(beta);
    for (size_t i = 0; i < N; i++) res[i] = gx[i] - sol[i];
    double resnorm = nrm2(res, n);
    double up_bd = 1.e4 * resnorm;
    double tol = o->rtol * resnorm + o->atol;
    double picard_tol = o->picard_rtol * resnorm + o->picard_atol;
    HIST(resnorm);
    int toosoon = (resnorm <= tol);

    if (!toosoon && o->picard_maxit > 0) {
        memcpy(sol, gx, sizeof(double) * N);
        for (int it = 0; it < o->picard_maxit - 1; it++) {
            EVAL(gx, sol);
            BETAFIX(o->picard_beta);
            for (size_t i = 0; i < N; i++) res[i] = gx[i] - sol[i];
            resnorm = nrm2(res, n);
            HIST(resnorm);
            if (resnorm <= picard_tol || resnorm >= up_bd) break;
            memcpy(sol, gx, sizeof(double) * N);
        }
        toosoon = (resnorm <= tol);
    }

    /* aa_point!: anderson.jl:470-481 */
#define AA_POINT() do {                                                        \
        for (size_t i_ = 0; i_ < N; i_++) dg[i_] = -gx[i_];                    \
        EVAL(gx, sol);                                                         \
        BETAFIX(beta);                                                         \
        for (size_t i_ = 0; i_ < N; i_++) dg[i_] += gx[i_];                    \
        for (size_t i_ = 0; i_ < N; i_++) df[i_] = -res[i_];                   \
        for (size_t i_ = 0; i_ < N; i_++) res[i_] = gx[i_] - sol[i_];          \
        for (size_t i_ = 0; i_ < N; i_++) df[i_] += res[i_];                   \
        resnorm = nrm2(res, n);                                                \
    } while (0)

    if (!toosoon) {
        memcpy(sol, gx, sizeof(double) * N);
        k += 1;
        AA_POINT();
        HIST(resnorm);
    }

    while (k < o->maxit + o->picard_maxit && resnorm > tol && !toosoon && resnorm < up_bd) {
        double condit, alphanrm;
        if (m == 0) {
            alphanrm = 1;
            condit = 1;
            memcpy(sol, gx, sizeof(double) * N);
        } else {
            /* BuildDG!(DG, m, k+1, dg) */
            int kk = k + 1;
            if (m == 1) {
                memcpy(DG, dg, sizeof(double) * N);
            } else if (kk > m + 1) {
                for (int ic = 0; ic < m - 1; ic++) memcpy(DG + ic * N, DG + (ic + 1) * N, sizeof(double) * N);
                memcpy(DG + (m - 1) * N, dg, sizeof(double) * N);
            } else {
                memcpy(DG + (kk - 2) * N, dg, sizeof(double) * N);
            }
            /* aa_qr_update!(QP, RP, df, m, k-1, Qd) */
            int kq = k - 1;
            if (kq == 0) {
                RP[0] = nrm2(df, n);
                double nv = nrm2(df, n);
                for (size_t i = 0; i < N; i++) QP[i] = df[i] / nv;
            } else {
                if (kq > m - 1) {
                    householder_qr(RP, m, Rd, Qx);
                    for (int j = 0; j < m - 1; j++)
                        for (int i = 0; i < m - 1; i++) RP[j * m + i] = Rd[j * (m - 1) + i];
                    for (int i = 0; i < m; i++) RP[(m - 1) * m + i] = 0.0;
                    /* Qd = Q * Qx; Q[:,1:m-1] = Qd; Q[:,m] = 0 */
                    for (int j = 0; j < m - 1; j++)
                        for (size_t i = 0; i < N; i++) {
                            double acc = 0.0;
                            for (int l = 0; l < m; l++) acc += QP[l * N + i] * Qx[j * m + l];
                            Qd[j * N + i] = acc;
                        }
                    memcpy(QP, Qd, sizeof(double) * N * (m - 1));
                    memset(QP + (m - 1) * N, 0, sizeof(double) * N);
                }
                int kqq = kq < m - 1 ? kq : m - 1;
                /* update_aaqr! with Orthogonalize! (CGS2) */
                double h1[8], h2[8];
                for (int j = 0; j < kqq; j++) h1[j] = dot(QP + j * N, df, n);
                for (size_t i = 0; i < N; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < kqq; j++) acc += QP[j * N + i] * h1[j];
                    df[i] -= acc;
                }
                for (int j = 0; j < kqq; j++) h2[j] = dot(QP + j * N, df, n);
                for (size_t i = 0; i < N; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < kqq; j++) acc += QP[j * N + i] * h2[j];
                    df[i] -= acc;
                }
                for (int j = 0; j < kqq; j++) RP[kqq * m + j] = h1[j] + h2[j];
                double hn = nrm2(df, n);
                RP[kqq * m + kqq] = hn;
                for (size_t i = 0; i < N; i++) df[i] /= hn;
                memcpy(QP + kqq * N, df, sizeof(double) * N);
            }
            int mk = m < k ? m : k;
            for (int j = 0; j < mk; j++) tres[j] = dot(QP + j * N, res, n);
            /* theta = RA \ tres (upper triangular back substitution) */
            for (int j = mk - 1; j >= 0; j--) {
                double acc = tres[j];
                for (int l = j + 1; l < mk; l++) acc -= RP[l * m + j] * theta[l];
                theta[j] = acc / RP[j * m + j];
            }
            condit = cond2(RP, m, mk);
            alpha[0] = theta[0];
            for (int ia = 1; ia < mk; ia++) alpha[ia] = theta[ia] - theta[ia - 1];
            alpha[mk] = 1.0 - theta[mk - 1];
            alphanrm = 0.0;
            for (int ia = 0; ia <= m; ia++) alphanrm += fabs(alpha[ia]);
            memcpy(sol, gx, sizeof(double) * N);
            for (size_t i = 0; i < N; i++) {
                double acc = 0.0;
                for (int j = 0; j < mk; j++) acc += DG[j * N + i] * theta[j];
                sol[i] -= acc;
            }
        }
        if (ns < stats_cap) {
            if (condhist) condhist[ns] = condit;
            if (alphanorm) alphanorm[ns] = alphanrm;
        }
        ns++;
        k += 1;
        AA_POINT();
        HIST(resnorm);
    }

    /* AndersonOK (anderson.jl:727-737) */
    int idid, errcode;
    if (toosoon) { idid = 1; errcode = -1; }
    else if (resnorm >= up_bd) { idid = 0; errcode = -2; }
    else if (resnorm <= tol) { idid = 1; errcode = 0; }
    else { idid = 0; errcode = 10; }

    memcpy(sol_out, sol, sizeof(double) * N);
    if (gx_out) memcpy(gx_out, gx, sizeof(double) * N);
    res_out->idid = idid;
    res_out->errcode = errcode;
    res_out->k = k;
    res_out->nhist = nh;
    res_out->nstats = ns;
    res_out->nevals = nevals;
    res_out->resnorm = resnorm;
    res_out->tol = tol;
    free(sol); free(gx); free(df); free(dg); free(res); free(QP); free(DG); free(Qd);
#undef EVAL
#undef BETAFIX
#undef HIST
#undef AA_POINT
    return 0;
}

/* converge!(::DFTProblem, ::AASol, rho): converge.jl:214-254 with the keyword
 * defaults of :3-13 (rtol = atol = tol, picard_maxit = maxit, picard_beta = beta,
 * picard_rtol = picard_atol = anderson_start). */
typedef struct {
    const or_system *s;
    double *rho;
} conv_ctx;

static void gfix(double *G, const double *x, void *vctx)
{
    conv_ctx *c = (conv_ctx *)vctx;
    size_t n = (size_t)or_npoints(c->s) * c->s->nc;
    for (size_t i = 0; i < n; i++) c->rho[i] = exp(x[i]);
    or_get_new_profile(c->s, c->rho, G);
}

int or_converge(const or_system *s, double *rho, int maxit, double beta, double tol, double anderson_start,
                int m, or_aa_result *res, double *hist, int hist_cap)
{
    size_t n = (size_t)or_npoints(s) * s->nc;
    double *x0 = malloc(sizeof(double) * n), *sol = malloc(sizeof(double) * n);
    double *work = malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++) x0[i] = log(rho[i]);
    or_aa_opts o = {m, maxit, maxit, tol, tol, beta, beta, anderson_start, anderson_start};
    conv_ctx c = {s, work};
    int rc = or_aasol(gfix, &c, (int)n, x0, &o, sol, NULL, hist, hist_cap, NULL, NULL, 0, res);
    for (size_t i = 0; i < n; i++) rho[i] = exp(sol[i]);
    free(x0); free(sol); free(work);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* External field and seeds.                                                 */
/* ------------------------------------------------------------------------ */

/* Planar Steele 10-4-3 wall(s): steele.jl:49-81 with LinRange coordinates
 * (structure.jl:138-151, base/structure.jl:178-181).  Pure-component
 * sigma_i, eps_i are the diagonal of the model's matrices. */
void or_steele_vext(const or_system *s, double eps_s, double sigma_s, double delta, double rho_s,
                    double width, double *vext)
{
    int N = s->ngrid[0];
    double lb = s->lb[0], ub = s->ub[0];
    for (int c = 0; c < s->nc; c++) {
        double ei = s->eps[c * s->nc + c], si = s->sigma[c * s->nc + c];
        double esi = sqrt(eps_s * ei), ssi = (sigma_s + si) / 2.0;
        double pre = 2.0 * PI * rho_s * esi * delta * ssi * ssi;
        for (int j = 0; j < N; j++) {
            double t = (double)j / (double)(N - 1);
            double z = (1.0 - t) * lb + t * ub;
            double v = pre * (2.0 / 5.0 * pow(ssi / z, 10) - pow(ssi / z, 4) -
                              pow(ssi, 4) / (3.0 * delta * pow(z + 0.61 * delta, 3)));
            if (width != 0.0) {
                double zz = width - z;
                v += pre * (2.0 / 5.0 * pow(ssi / zz, 10) - pow(ssi / zz, 4) -
                            pow(ssi, 4) / (3.0 * delta * pow(zz + 0.61 * delta, 3)));
            }
            vext[(size_t)c * N + j] = v * (1.0 / s->T);
        }
    }
}

static void clamp_init(double *rho, size_t n)
{
    double lo = pow(DBL_MIN, 1.0 / 6.0), hi = pow(DBL_MAX, 1.0 / 6.0);   /* structure.jl:88-91 */
    for (size_t i = 0; i < n; i++) {
        if (rho[i] < lo) rho[i] = lo;
        if (rho[i] > hi) rho[i] = hi;
    }
}

/* Uniform seed (structure.jl:125-136) times exp(-beta V/10) (external_field.jl:21-26). */
void or_init_profiles(const or_system *s, double *rho)
{
    int N = or_npoints(s);
    for (int c = 0; c < s->nc; c++)
        for (int j = 0; j < N; j++) {
            double r = s->rho_bulk[c];
            if (s->vext) r *= exp(-s->vext[(size_t)c * N + j] / 10.0);
            rho[(size_t)c * N + j] = r;
        }
    clamp_init(rho, (size_t)N * s->nc);
}

/* TwoPhase1DCart seed (two_phase.jl:1-35, cos_prof structure.jl:2). */
void or_init_two_phase(const or_system *s, const double *rho2, const double *Tc, double *rho)
{
    int N = s->ngrid[0];
    double lb = s->lb[0], ub = s->ub[0], H = ub - lb;
    double shift = ub / 4.0 + 3.0 * lb / 4.0;
    for (int c = 0; c < s->nc; c++) {
        double Tr = s->T / Tc[c];
        double coef = (2.4728 - 2.3625 * Tr) * (H / s->L);
        coef = sqrt(coef * coef - 1.0) / 4.0;
        double a = s->rho_bulk[c], b = rho2[c];
        for (int j = 0; j < N; j++) {
            double t = (double)j / (double)(N - 1);
            double X = (1.0 - t) * lb + t * ub;
            double x = X / (ub - lb);
            double cs = cos((x - shift) * 2.0 * PI);
            double v = 0.5 * (a - b) * cs * sqrt((1.0 + coef * coef) / (1.0 + coef * coef * cs * cs)) + 0.5 * (a + b);
            rho[(size_t)c * N + j] = v;
        }
    }
    clamp_init(rho, (size_t)N * s->nc);
}

/* ------------------------------------------------------------------------ */
/* Post-processing: surface_tension.jl:31-53, adsorption.jl:22-32.           */
/* ------------------------------------------------------------------------ */

double or_surface_tension(const or_system *s, const double *rho, const double *mu, double p)
{
    int N = or_npoints(s);
    double F = or_F_ideal(s, rho) + or_F_res(s, rho);
    double chem = 0.0;
    for (int c = 0; c < s->nc; c++) chem += mu[c] * or_integrate(s, rho + (size_t)c * N);
    double *ones = malloc(sizeof(double) * N);
    for (int j = 0; j < N; j++) ones[j] = 1.0;
    double vol = or_integrate(s, ones);
    free(ones);
    return F * OR_KB * s->T - chem + p * vol;
}

void or_adsorption(const or_system *s, const double *rho, double *out)
{
    int N = or_npoints(s);
    double V = 1.0;                                      /* _domain_volume, adsorption.jl:23-25 */
    const double lo = s->lb[0], hi = s->ub[0];
    if (s->coord == 1) V = 4.0 * PI / 3.0 * (hi * hi * hi - lo * lo * lo);
    else if (s->coord == 2) V = PI * (hi * hi - lo * lo);
    else
        for (int a = 0; a < s->nd; a++) V *= (s->ub[a] - s->lb[a]);
    for (int c = 0; c < s->nc; c++) out[c] = or_integrate(s, rho + (size_t)c * N) / V;
}

/* ------------------------------------------------------------------------ */
/* Planned 1D evaluation.  The same operator as or_weighted_densities /      */
/* or_transposed, with its pieces computed once per system instead of once  */
/* per call:                                                                */
/*   h_f = Re IFFT_N(Omega_f)    the real circulant kernel that             */
/*                               Re IFFT(FFT(u) Omega_f) applies to a real u */
/*                               (integrals.jl:88-96), from the literal      */
/*                               Omega and FFT above;                         */
/*   H_f = FFT_M(h_f unrolled on a power-of-two ring M >= 2N - 1, or M = N  */
/*         when N is a power of two): lags -(N-1)..N-1 do not alias, so the  */
/*         length-M circular convolution of the zero-padded u equals the    */
/*         N-periodic one on 0..N-1.                                        */
/* Two real fields share one complex transform; the transposed operator of  */
/* h_f has the spectrum conj(H_f) (the vector slot's h is odd, so this is    */
/* the reference's subtraction, weighted_densities.jl:381).  Used by the    */
/* tests for speed (checked there against the literal functions) and by the */
/* OpenMP batch driver that times the CPU.                                  */
/* ------------------------------------------------------------------------ */

struct or_plan {
    const or_system *s;
    int N, M, lg, nc, F;
    double *twr, *twi;                 /* W_M^t = exp(-2 pi i t / M), t < M/2 */
    int *rev;                          /* bit reversal of log2 M bits         */
    double *Hr, *Hi;                   /* [(c*F + f)*M + k], f >= 1           */
    double *ar, *ai, *pr, *pi_, *sr, *si;
    double *n, *df;                    /* fields and gradients, N*F*nc        */
};

static void fft_pow2(const or_plan *p, double *re, double *im, int inverse)
{
    const int M = p->M;
    for (int i = 0; i < M; i++) {
        int j = p->rev[i];
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len >> 1, step = M / len;
        for (int b = 0; b < M; b += len)
            for (int j = 0; j < half; j++) {
                const double wr = p->twr[j * step], wi = inverse ? -p->twi[j * step] : p->twi[j * step];
                const int a = b + j, c = a + half;
                const double xr = re[c] * wr - im[c] * wi, xi = re[c] * wi + im[c] * wr;
                re[c] = re[a] - xr;
                im[c] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

void or_plan_destroy(or_plan *p)
{
    if (!p) return;
    free(p->twr); free(p->twi); free(p->rev); free(p->Hr); free(p->Hi);
    free(p->ar); free(p->ai); free(p->pr); free(p->pi_); free(p->sr); free(p->si);
    free(p->n); free(p->df);
    free(p);
}

or_plan *or_plan_create(const or_system *s)
{
    if (s->nd != 1 || s->coord) return NULL;
    or_plan *p = calloc(1, sizeof(or_plan));
    const int N = s->ngrid[0], F = 7, nc = s->nc;
    int M = 1, lg = 0;
    if ((N & (N - 1)) == 0) {
        M = N;
    } else {
        while (M < 2 * N - 1) M <<= 1;
    }
    while ((1 << lg) < M) lg++;
    p->s = s; p->N = N; p->M = M; p->lg = lg; p->nc = nc; p->F = F;
    p->twr = malloc(sizeof(double) * M); p->twi = malloc(sizeof(double) * M);
    for (int t = 0; t < M; t++) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)M;
        p->twr[t] = (double)cosl(a);
        p->twi[t] = (double)(-sinl(a));
    }
    p->rev = malloc(sizeof(int) * M);
    for (int i = 0; i < M; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
        p->rev[i] = r;
    }
    size_t nk = (size_t)nc * F * M;
    p->Hr = calloc(nk, sizeof(double)); p->Hi = calloc(nk, sizeof(double));
    double *ore = malloc(sizeof(double) * (size_t)N * F * nc), *oim = malloc(sizeof(double) * (size_t)N * F * nc);
    double *hr = malloc(sizeof(double) * N), *hi = malloc(sizeof(double) * N);
    or_omega(s, ore, oim);
    for (int c = 0; c < nc; c++)
        for (int f = 1; f < F; f++) {
            const size_t o = (size_t)(c * F + f) * N;
            memcpy(hr, ore + o, sizeof(double) * N);
            memcpy(hi, oim + o, sizeof(double) * N);
            or_fft(hr, hi, 1, s->ngrid, 1);                    /* h = Re IFFT_N(Omega) */
            double *Hr = p->Hr + (size_t)(c * F + f) * M, *Hi = p->Hi + (size_t)(c * F + f) * M;
            for (int k = 0; k < M; k++) Hi[k] = 0.0;
            if (M == N) {
                for (int t = 0; t < N; t++) Hr[t] = hr[t];
            } else {
                for (int k = 0; k < M; k++) Hr[k] = 0.0;
                Hr[0] = hr[0];
                for (int t = 1; t < N; t++) {
                    Hr[t] = hr[t];
                    Hr[M - t] = hr[N - t];
                }
            }
            fft_pow2(p, Hr, Hi, 0);
        }
    free(ore); free(oim); free(hr); free(hi);
    p->ar = malloc(sizeof(double) * M); p->ai = malloc(sizeof(double) * M);
    p->pr = malloc(sizeof(double) * M); p->pi_ = malloc(sizeof(double) * M);
    p->sr = malloc(sizeof(double) * M); p->si = malloc(sizeof(double) * M);
    p->n = malloc(sizeof(double) * (size_t)N * F * nc);
    p->df = malloc(sizeof(double) * (size_t)N * F * nc);
    return p;
}

/* dF_res/drho and (optionally) Phi with a plan; same result as or_dFdrho_res up
 * to rounding.  A plan's buffers serve one call at a time. */
void or_dFdrho_res_planned(or_plan *p, const double *rho, double *dfdrho, double *fval)
{
    const or_system *s = p->s;
    const int N = p->N, M = p->M, F = p->F, nc = p->nc;
    const double NA = OR_NA * (s->L * s->L * s->L), invM = 1.0 / (double)M;
    double nl[OR_MAXC * 9], dl[OR_MAXC * 9];
    for (int c = 0; c < nc; c++) {
        const double *rc = rho + (size_t)c * N;
        for (int k = 0; k < M; k++) { p->pr[k] = k < N ? rc[k] : 0.0; p->pi_[k] = 0.0; }
        fft_pow2(p, p->pr, p->pi_, 0);
        for (int j = 0; j < N; j++) p->n[(size_t)(c * F) * N + j] = rc[j] * NA;
        for (int q = 0; q < 3; q++) {
            const int fa = 1 + 2 * q, fb = 2 + 2 * q;
            const double *Har = p->Hr + (size_t)(c * F + fa) * M, *Hai = p->Hi + (size_t)(c * F + fa) * M;
            const double *Hbr = p->Hr + (size_t)(c * F + fb) * M, *Hbi = p->Hi + (size_t)(c * F + fb) * M;
            for (int k = 0; k < M; k++) {
                const double zr = Har[k] - Hbi[k], zi = Hai[k] + Hbr[k];     /* Z = H_a + i H_b */
                p->ar[k] = p->pr[k] * zr - p->pi_[k] * zi;
                p->ai[k] = p->pr[k] * zi + p->pi_[k] * zr;
            }
            fft_pow2(p, p->ar, p->ai, 1);
            double *na = p->n + (size_t)(c * F + fa) * N, *nb = p->n + (size_t)(c * F + fb) * N;
            for (int j = 0; j < N; j++) {
                na[j] = p->ar[j] * invM * NA;
                nb[j] = p->ai[j] * invM * NA;
            }
        }
    }
    for (int j = 0; j < N; j++) {
        for (int c = 0; c < nc; c++)
            for (int f = 0; f < F; f++) nl[c * F + f] = p->n[(size_t)(c * F + f) * N + j];
        double v = or_phi(s, nl, dl);
        if (fval) fval[j] = v;
        for (int c = 0; c < nc; c++)
            for (int f = 0; f < F; f++) p->df[(size_t)(c * F + f) * N + j] = dl[c * F + f];
    }
    for (int c = 0; c < nc; c++) {
        for (int k = 0; k < M; k++) { p->sr[k] = 0.0; p->si[k] = 0.0; }
        for (int q = 0; q < 3; q++) {
            const int fa = 1 + 2 * q, fb = 2 + 2 * q;
            const double *ga = p->df + (size_t)(c * F + fa) * N, *gb = p->df + (size_t)(c * F + fb) * N;
            for (int k = 0; k < M; k++) {
                p->ar[k] = k < N ? ga[k] : 0.0;
                p->ai[k] = k < N ? gb[k] : 0.0;
            }
            fft_pow2(p, p->ar, p->ai, 0);
            const double *Har = p->Hr + (size_t)(c * F + fa) * M, *Hai = p->Hi + (size_t)(c * F + fa) * M;
            const double *Hbr = p->Hr + (size_t)(c * F + fb) * M, *Hbi = p->Hi + (size_t)(c * F + fb) * M;
            for (int k = 0; k < M; k++) {
                const double zr = Har[k] - Hbi[k], zi = Hai[k] + Hbr[k];     /* S += A conj(Z) */
                p->sr[k] += p->ar[k] * zr + p->ai[k] * zi;
                p->si[k] += p->ai[k] * zr - p->ar[k] * zi;
            }
        }
        fft_pow2(p, p->sr, p->si, 1);
        double *oc = dfdrho + (size_t)c * N;
        const double *g0 = p->df + (size_t)(c * F) * N;
        for (int j = 0; j < N; j++) oc[j] = g0[j] + p->sr[j] * invM;
    }
}

typedef struct {
    or_plan *p;
    double *rho;
} conv_ctx_planned;

static void gfix_planned(double *G, const double *x, void *vctx)
{
    conv_ctx_planned *c = (conv_ctx_planned *)vctx;
    const or_system *s = c->p->s;
    const int N = c->p->N;
    const size_t n = (size_t)N * s->nc;
    for (size_t i = 0; i < n; i++) c->rho[i] = exp(x[i]);
    or_dFdrho_res_planned(c->p, c->rho, G, NULL);
    if (s->vext)
        for (size_t j = 0; j < n; j++) G[j] += s->vext[j];
    for (int cc = 0; cc < s->nc; cc++) {
        const double cp = s->mu_res[cc] + log(s->rho_bulk[cc]);
        for (int j = 0; j < N; j++) {
            double v = cp - G[(size_t)cc * N + j];
            if (v < -100.0) v = -100.0;
            if (v > 100.0) v = 100.0;
            G[(size_t)cc * N + j] = v;
        }
    }
}

/* or_converge with a plan (1D only; returns -3 for other grids). */
int or_converge_planned(const or_system *s, double *rho, int maxit, double beta, double tol,
                        double anderson_start, int m, or_aa_result *res, double *hist, int hist_cap)
{
    or_plan *p = or_plan_create(s);
    if (!p) return -3;
    size_t n = (size_t)or_npoints(s) * s->nc;
    double *x0 = malloc(sizeof(double) * n), *sol = malloc(sizeof(double) * n);
    double *work = malloc(sizeof(double) * n);
    for (size_t i = 0; i < n; i++) x0[i] = log(rho[i]);
    or_aa_opts o = {m, maxit, maxit, tol, tol, beta, beta, anderson_start, anderson_start};
    conv_ctx_planned c = {p, work};
    int rc = or_aasol(gfix_planned, &c, (int)n, x0, &o, sol, NULL, hist, hist_cap, NULL, NULL, 0, res);
    for (size_t i = 0; i < n; i++) rho[i] = exp(sol[i]);
    free(x0); free(sol); free(work);
    or_plan_destroy(p);
    return rc;
}

/* Independent 1D systems solved by or_converge_planned on nthreads OpenMP
 * threads (dynamic schedule, one system at a time per thread).  rhos[i] is
 * overwritten like or_converge's rho; secs[i] (optional) receives the wall time
 * of system i. */
int or_converge_batch(const or_system *systems, int nsys, double *const *rhos, int maxit, double beta, double tol,
                      double anderson_start, int m, int nthreads, or_aa_result *res, double *secs)
{
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : bad)
    for (int i = 0; i < nsys; i++) {
        double t0 = or_wtime();
        int rc = or_converge_planned(&systems[i], rhos[i], maxit, beta, tol, anderson_start, m, &res[i], NULL, 0);
        if (secs) secs[i] = or_wtime() - t0;
        bad += (rc < 0);
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* DDFT right-hand side in the log-density form (ext/SciMLBasecDFTExt.jl:   */
/* 10-61): rho = exp(clamp(eta, -50, 30)), mu = dF_res/drho + beta V_ext     */
/* (the ideal propagator adds nothing, propagator/ideal.jl:1-3), then per    */
/* component                                                                 */
/*   deta = lap(eta) + lap(mu) + sum_d grad_d(eta)^2 + grad_d(mu) grad_d(eta) */
/* with every derivative a convolve! (integrals.jl:88-96): real(IFFT(FFT(u)  */
/* .* map)), map_lapl = -(2 pi)^2 |k|^2 and map_grad_d = 2 pi i k_d, k the   */
/* reduced frequencies omega L (structure_omega, structure.jl:153-185).      */
/* ------------------------------------------------------------------------ */

static void ddft_convolve(const or_system *s, const double *src, int kind, double *dst, double *br, double *bi)
{
    /* kind 0: Laplacian; 1 + d: gradient along axis d */
    int N = or_npoints(s), nd = s->nd;
    int n0 = s->ngrid[0], n1 = nd > 1 ? s->ngrid[1] : 1;
    for (int j = 0; j < N; j++) { br[j] = src[j]; bi[j] = 0.0; }
    or_fft(br, bi, nd, s->ngrid, 0);
    const double twopi = 2.0 * PI;
    for (int idx = 0; idx < N; idx++) {
        int ii[3] = {idx % n0, (idx / n0) % n1, idx / (n0 * n1)};
        double mr, mi;
        if (kind == 0) {
            double k2 = 0.0;
            for (int a = 0; a < nd; a++) {
                double k = omega_axis(s, a, ii[a]);
                k2 += k * k;
            }
            mr = -twopi * twopi * k2;
            mi = 0.0;
        } else {
            mr = 0.0;
            mi = twopi * omega_axis(s, kind - 1, ii[kind - 1]);
        }
        double r = br[idx] * mr - bi[idx] * mi, im = br[idx] * mi + bi[idx] * mr;
        br[idx] = r;
        bi[idx] = im;
    }
    or_fft(br, bi, nd, s->ngrid, 1);
    for (int j = 0; j < N; j++) dst[j] = br[j];
}

void or_ddft_rhs(const or_system *s, const double *eta, double *deta)
{
    int N = or_npoints(s), nc = s->nc, nd = s->nd;
    size_t n = (size_t)N * nc;
    double *rho = malloc(sizeof(double) * n), *mu = malloc(sizeof(double) * n);
    double *br = malloc(sizeof(double) * N), *bi = malloc(sizeof(double) * N);
    double *tmp = malloc(sizeof(double) * N), *g = malloc(sizeof(double) * N);
    for (size_t i = 0; i < n; i++) {
        double e = eta[i];
        if (e < -50.0) e = -50.0;
        if (e > 30.0) e = 30.0;
        rho[i] = exp(e);
    }
    or_dFdrho_res(s, rho, mu, NULL);
    if (s->vext)
        for (size_t i = 0; i < n; i++) mu[i] += s->vext[i];
    for (int c = 0; c < nc; c++) {
        const double *ec = eta + (size_t)c * N, *mc = mu + (size_t)c * N;
        double *dc = deta + (size_t)c * N;
        ddft_convolve(s, ec, 0, dc, br, bi);
        ddft_convolve(s, mc, 0, tmp, br, bi);
        for (int j = 0; j < N; j++) dc[j] += tmp[j];
        for (int d = 0; d < nd; d++) {
            ddft_convolve(s, ec, 1 + d, tmp, br, bi);
            for (int j = 0; j < N; j++) dc[j] += tmp[j] * tmp[j];
            ddft_convolve(s, mc, 1 + d, g, br, bi);
            for (int j = 0; j < N; j++) dc[j] += g[j] * tmp[j];
        }
    }
    free(rho); free(mu); free(br); free(bi); free(tmp); free(g);
}
