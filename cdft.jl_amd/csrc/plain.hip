// plain.hip -- kernels of the plain (one system at a time) path for 1D grids.
//
// One Euler-Lagrange map evaluation G(x) (GFix!, src/methods/converge.jl:235-242)
// is four launches:
//   k_circ1d  (forward)  n_f = N_A L^3 * h_f (*) rho          weighted densities
//   k_phi               Phi and dPhi/dn at every point       Helmholtz density + gradient
//   k_circ1d  (adjoint)  p_f = +-h_f (*) dPhi/dn_f             transposed convolutions
//   k_update            ln G = clamp(mu + ln rho_b - sum - beta V) and Anderson book-keeping
// where (*) is the circular convolution over the period of the grid.  The kernel
// h_f is the exact real-space form of the reference's spectral operator
// Re IFFT(FFT(.) * Omega_f) (src/utils/integrals.jl:88-96), so the result equals
// the reference's FFT convolution up to rounding.
#include "common.h"
#include "dual.h"
#include "kernels.h"

namespace cdft {

// Real-space circulant kernels from the analytic weights (set-up, FP64).
// h_f[j] = (1/N) sum_k Omega_herm_f(k) exp(+2 pi i k j / N), with the Hermitian
// part Omega_herm(k) = (Omega(k) + conj Omega(-k))/2 that survives real().
// Frequencies: fftfreq(N, N/(ub-lb)) * L (structure.jl:153-185, PCSAFT.jl:30);
// weights: weighted_densities.jl:48-69 (scalar) and :210-227 (vector).
__global__ void k_setup_h1d(double *h, int N, int nc, const double *dd, double psi, double mult, double L)
{
    const int F = 7;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int cf = blockIdx.y;                 // c*F + f
    int c = cf / F, f = cf % F;
    if (j >= N || f == 0) return;
    double width;
    int kind;                            // 1 S, 2 V, 3 vector
    switch (f) {
    case 1: kind = 1; width = 0.5 * dd[c]; break;
    case 2: kind = 2; width = 0.5 * dd[c]; break;
    case 3: kind = 3; width = 0.5 * dd[c]; break;
    case 4: kind = 2; width = dd[c]; break;
    case 5: kind = 1; width = dd[c]; break;
    default: kind = 2; width = dd[c] * psi; break;
    }
    const double twopi = 2.0 * kPi, fourpi = 4.0 * kPi, tp3 = twopi * twopi * twopi;
    double R = twopi * width;
    int nnonneg = (N + 1) / 2;
    double acc = 0.0;
    for (int k = 0; k < N; k++) {
        int ks = (k < nnonneg) ? k : k - N;
        double w = ((double)ks * mult) * L;
        double wb = fabs(w);
        long r = ((long)k * (long)j) % N;
        double sn, cs;
        sincospi(2.0 * (double)r / (double)N, &sn, &cs);
        double wbR = wb * R;
        if (kind == 1) {
            double om = (wb == 0.0 ? 2.0 * R : 2.0 * sin(wbR) / wb) / twopi;
            acc += om * cs;
        } else if (kind == 2) {
            double om = (wb == 0.0 ? fourpi * (R * R * R) / 3.0
                                   : fourpi / (wb * wb * wb) * (sin(wbR) - R * wb * cos(wbR))) / tp3;
            acc += om * cs;
        } else {
            // Omega = i W(k): zero Hermitian part at k = 0 and at the Nyquist index
            if (ks == 0 || (2 * k == N)) continue;
            double W = fourpi * w / (wb * wb * wb) * (sin(wbR) - R * wb * cos(wbR)) / tp3;
            acc -= W * sn;
        }
    }
    h[(size_t)cf * N + j] = acc / (double)N;
}

// Real-space kernels of the DDFT derivative maps on a 1D grid (set-up, FP64):
// h[(c F + 1) N + j] for the Laplacian map -(2 pi)^2 k^2 and h[(c F + f) N + j],
// f = 2, 3, for the gradient map 2 pi i k (F = 4; the same for every component),
// h = Re IDFT_N(map) with k the reduced frequencies omega L; the gradient's Nyquist
// term is imaginary and real() drops it (ext/SciMLBasecDFTExt.jl:16-19,
// integrals.jl:88-96).
__global__ void k_setup_hddft(double *h, int N, int nc, double mult, double L)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int cf = blockIdx.y;                 // c*4 + f
    const int f = cf % 4;
    if (j >= N || f == 0) return;
    const double twopi = 2.0 * kPi;
    const int nnonneg = (N + 1) / 2;
    double acc = 0.0;
    for (int k = 0; k < N; k++) {
        const int ks = (k < nnonneg) ? k : k - N;
        const double w = ((double)ks * mult) * L;
        const long r = ((long)k * (long)j) % N;
        double sn, cs;
        sincospi(2.0 * (double)r / (double)N, &sn, &cs);
        if (f == 1) {
            acc += -twopi * twopi * w * w * cs;
        } else {
            if (2 * k == N) continue;
            acc -= twopi * w * sn;                // Re(i 2 pi w e^{i theta}) = -2 pi w sin(theta)
        }
    }
    h[(size_t)cf * N + j] = acc / (double)N;
}

// rho = exp(clamp(eta, -50, 30)) with Julia's clamp (NaN stays NaN)
// (ext/SciMLBasecDFTExt.jl:27).
__global__ void k_ddft_rho(const double *eta, double *rho, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double e = eta[i];
    rho[i] = exp(e > 30.0 ? 30.0 : (e < -50.0 ? -50.0 : e));
}

// Sources of the DDFT derivatives: mu = dF_res/drho + beta V_ext; per component the
// slots 1 (eta + mu, for the Laplacian), 2 (eta) and 3 (mu) of src[(c 4 + f) N + i].
__global__ void k_ddft_src(const double *eta, const double *dfdrho, const double *vext, int N, int nc, double *src,
                           double *mu)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * nc) return;
    const int c = t / N, i = t - c * N;
    const double m = dfdrho[t] + (vext ? vext[t] : 0.0);
    const double e = eta[t];
    if (mu) mu[t] = m;
    if (src) {
        src[(size_t)(c * 4 + 1) * N + i] = e + m;
        src[(size_t)(c * 4 + 2) * N + i] = e;
        src[(size_t)(c * 4 + 3) * N + i] = m;
    }
}

// deta = lap(eta + mu) + grad(eta)^2 + grad(mu) grad(eta) from the convolved slots
// d[(c 3 + f - 1) N + i] (1D).
__global__ void k_ddft_combine(const double *d, int N, int nc, double *deta)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * nc) return;
    const int c = t / N, i = t - c * N;
    const double lap = d[(size_t)(c * 3) * N + i], ge = d[(size_t)(c * 3 + 1) * N + i], gm = d[(size_t)(c * 3 + 2) * N + i];
    deta[t] = lap + ge * ge + gm * ge;
}

// Circular convolution dst[i] = scale * sum_k h[k] src[(i - k) mod N] for a batch
// of (src, h, dst) triples selected by blockIdx.y = c*(F-1) + (f-1).  A workgroup
// owns CIRC_TW consecutive outputs; each of its four waves sums one quarter of the
// k range (in LDS chunks: h chunk with broadcast reads, the matching window of src
// with consecutive reads per lane), and the quarters are added in a fixed order,
// so the result is reproducible and many workgroups share the N^2 work.
constexpr int CIRC_TW = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_circ1d(const T *__restrict__ src, long s_cstride, long s_fstride,
                                                const T *__restrict__ h, T *__restrict__ dst, long d_cstride,
                                                long d_fstride, int d_fofs, int N, int F, T scale, int vec_sign)
{
    constexpr int TW = CIRC_TW, NG = 4, KC = 256;
    __shared__ T hs[NG][KC];
    __shared__ T ws[NG][TW + KC];
    __shared__ T part[NG][TW];
    const int y = blockIdx.y;
    const int c = y / (F - 1), f = 1 + y % (F - 1);
    const T *s = src + c * s_cstride + f * s_fstride;
    const T *hk = h + (size_t)(c * F + f) * N;
    T *d = dst + c * d_cstride + (f - d_fofs) * d_fstride;
    T sc = scale;
    if (vec_sign && f == 3) sc = -scale;   // 1D vector slot is f = 3
    const int i0 = blockIdx.x * TW;
    const int g = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int Kg = (N + NG - 1) / NG;                   // taps per wave
    const int kb = g * Kg, ke = min(N, kb + Kg);
    T acc = 0;
    for (int kc0 = 0; kc0 < Kg; kc0 += KC) {            // the same trip count in every wave
        const int kc = kb + kc0;
        const int kcn = max(0, min(KC, ke - kc));
        __syncthreads();
        for (int t = l; t < kcn; t += TW) hs[g][t] = hk[kc + t];
        const int base = i0 - kc - kcn + 1;
        const int wn = kcn > 0 ? TW + kcn - 1 : 0;
        for (int t = l; t < wn; t += TW) {
            int gi = (base + t) % N;
            if (gi < 0) gi += N;
            ws[g][t] = s[gi];
        }
        __syncthreads();
        if (kcn > 0) {
            const T *wp = &ws[g][l + kcn - 1];
            int kk = 0;
            T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (; kk + 3 < kcn; kk += 4) {
                a0 += hs[g][kk] * wp[-kk];
                a1 += hs[g][kk + 1] * wp[-kk - 1];
                a2 += hs[g][kk + 2] * wp[-kk - 2];
                a3 += hs[g][kk + 3] * wp[-kk - 3];
            }
            for (; kk < kcn; kk++) a0 += hs[g][kk] * wp[-kk];
            acc += (a0 + a1) + (a2 + a3);
        }
    }
    part[g][l] = acc;
    __syncthreads();
    const int i = i0 + l;
    if (g == 0 && i < N) d[i] = sc * ((part[0][l] + part[1][l]) + (part[2][l] + part[3][l]));
}

// Phi and dPhi/dn at every grid point.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi(const double *__restrict__ rho, const double *__restrict__ n,
                                             double *__restrict__ dphi, double *__restrict__ fval, int N,
                                             double NA, PhiConst P)
{
    constexpr int F = 6 + ND;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double nl[NC][F], g[NC][F];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        nl[c][0] = rho[(size_t)c * N + i] * NA;
#pragma unroll
        for (int f = 1; f < F; f++) nl[c][f] = n[(size_t)(c * F + f) * N + i];
    }
    double v = phi_grad<double, double, NC, ND, true, AS>(P, nl, g);
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int f = 0; f < F; f++) dphi[(size_t)(c * F + f) * N + i] = g[c][f];
    if (fval) fval[i] = v;
}

// FP32 map (mixed precision): Phi's gradient from FP32 weighted densities, with the
// n3-singular hard-sphere factors in FP64 (phi_grad<float, double>); the chain's
// lambda-slot gradient is 0 where FP32 does not resolve lambda.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi32(const double *__restrict__ rho, const float *__restrict__ n,
                                               float *__restrict__ dphi, int N, double NA, PhiConst P)
{
    constexpr int F = 6 + ND;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float nl[NC][F], g[NC][F];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        nl[c][0] = float(rho[(size_t)c * N + i] * NA);
#pragma unroll
        for (int f = 1; f < F; f++) nl[c][f] = n[(size_t)(c * F + f) * N + i];
    }
    phi_grad<float, double, NC, ND, true, AS, true>(P, nl, g);
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int f = 0; f < F; f++) dphi[(size_t)(c * F + f) * N + i] = g[c][f];
}

// Hessian-vector product of Phi at every point (no reference counterpart: the
// second derivative behind the fixed-point adjoint).  Slot 0 moves along N_A v,
// the other slots along their weighted densities ndot = W v (in ndio on entry);
// on exit ndio holds H_Phi(n) (N_A v, ndot) for every slot, the forward-mode
// derivative of phi_grad taken with dual numbers.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi_hvp(const double *__restrict__ rho, const double *__restrict__ v,
                                                 const double *__restrict__ n, double *__restrict__ ndio, int N,
                                                 double NA, PhiConst P)
{
    constexpr int F = 6 + ND;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    dual nl[NC][F], g[NC][F];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        nl[c][0] = dual(rho[(size_t)c * N + i] * NA, v[(size_t)c * N + i] * NA);
#pragma unroll
        for (int f = 1; f < F; f++) nl[c][f] = dual(n[(size_t)(c * F + f) * N + i], ndio[(size_t)(c * F + f) * N + i]);
    }
    phi_grad<dual, dual, NC, ND, true, AS>(P, nl, g);
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int f = 0; f < F; f++) ndio[(size_t)(c * F + f) * N + i] = g[c][f].d;
}

// mask = 1 where get_new_profile!'s clamp of the exponent to [-100, 100] is inactive
// (converge.jl:150): the exponent's derivative vanishes where it is clamped.
__global__ void k_clamp_mask(const double *dfdrho, const double *vext, const double *cp, int N, int nc, double *mask)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * nc) return;
    const int c = t / N;
    const double lg = cp[c] - dfdrho[t] - (vext ? vext[t] : 0.0);
    mask[t] = (lg > -100.0 && lg < 100.0) ? 1.0 : 0.0;
}

// dphi[(c F + f) N + i] *= cot[i]: the pullback of sum_i cot_i Phi_i (rrule seed).
__global__ void k_scale_slots(double *dphi, const double *cot, int N, int nslots)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double w = cot[i];
    for (int q = 0; q < nslots; q++) dphi[(size_t)q * N + i] *= w;
}

// ---------------------------------------------------------------------------
// Reductions: each block writes one partial per quantity; k_finalize sums the
// partials of each quantity in a fixed order (bitwise reproducible).
// ---------------------------------------------------------------------------
template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], double *part, int nblk)
{
    __shared__ double sm[NQ][16];
#pragma unroll
    for (int q = 0; q < NQ; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) sm[q][wid] = v[q];
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        double s = 0.0;
        for (int w = 0; w < nw; w++) s += sm[threadIdx.x][w];
        part[threadIdx.x * nblk + blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_finalize(const double *part, int nblk, double *out)
{
    __shared__ double sm[16];
    int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[q * nblk + b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < nw; w++) t += sm[w];
        out[q] = t;
    }
}

// sol = log(rho_in); rho = exp(sol)   (converge.jl:244 and GFix! :236)
__global__ void k_init_x(const double *rho_in, double *sol, double *rho, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = log(rho_in[i]);
    sol[i] = x;
    rho[i] = exp(x);
}

// sol = src; rho = exp(sol), or rho = mask sol for the linear adjoint map
__global__ void k_set_point(const double *src, double *sol, double *rho, int n, const double *mask)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = src[i];
    sol[i] = x;
    rho[i] = mask ? mask[i] * x : exp(x);
}

__global__ void k_exp(const double *x, double *rho, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rho[i] = exp(x[i]);
}

// Exponent assembly (get_new_profile!, converge.jl:108-153) fused with the
// Anderson book-keeping that follows each map evaluation.
//  mode 0: out = dF_res/drho (no field, no clamp)                functional only
//  mode 1: gx = betafix(lnG); res = gx - sol; partial |res|^2     first/Picard step
//  mode 2: aa_point! (anderson.jl:470-481); partials |res|^2, |df|^2, df.res
// With adj_g set, the map value is the adjoint's g - rho* (dphi + parts) instead of
// the clamped exponent (dphi, part then hold a Hessian-vector product).
__global__ __launch_bounds__(256) void k_update(const double *__restrict__ dphi, const double *__restrict__ part,
                                                const double *__restrict__ vext, int N, int nc, int F,
                                                const double *__restrict__ cp, double *__restrict__ out,
                                                const double *__restrict__ sol, double *__restrict__ gx,
                                                double *__restrict__ res, double *__restrict__ df,
                                                double *__restrict__ dg, double beta, int mode, double *red,
                                                int nblk, const double *__restrict__ adj_g,
                                                const double *__restrict__ adj_rho)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (t < N * nc) {
        int c = t / N, i = t - c * N;
        double s = dphi[(size_t)(c * F) * N + i];
        for (int f = 1; f < F; f++) s += part[(size_t)(c * (F - 1) + f - 1) * N + i];
        if (mode == 0) {
            out[t] = s;
        } else {
            double lg;
            if (adj_g) {                       // adjoint map: g - rho* (H (mask lambda))
                lg = adj_g[t] - adj_rho[t] * s;
            } else {
                if (vext) s += vext[t];
                lg = cp[c] - s;
                lg = clamp_lng(lg);
            }
            double x = sol[t];
            double g = (beta == 1.0) ? lg : (1.0 - beta) * x + beta * lg;
            if (mode == 1) {
                gx[t] = g;
                double r = g - x;
                res[t] = r;
                v[0] = r * r;
            } else {
                double gold = gx[t], rold = res[t];
                dg[t] = -gold + g;
                gx[t] = g;
                double r = g - x;
                res[t] = r;
                double d = -rold + r;
                df[t] = d;
                v[0] = r * r;
                v[1] = d * d;
                v[2] = d * r;
            }
        }
    }
    if (mode != 0) block_partials<3>(v, red, nblk);
}

// ----- Anderson QR update (aa_qr_update!, anderson.jl:498-583) --------------
// pass 1: optional downdate Q' = Q Qx, then h1 = Q'^T df and t = Q'^T res.
__global__ __launch_bounds__(256) void k_aa_proj1(int n, const double *__restrict__ Q, double *__restrict__ Qout,
                                                  AAMat qx, int downdate, int m, int kq,
                                                  const double *__restrict__ df, const double *__restrict__ res,
                                                  double *red, int nblk)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2 * CDFT_MAXM];
    for (int q = 0; q < 2 * CDFT_MAXM; q++) v[q] = 0.0;
    if (i < n) {
        double qv[CDFT_MAXM];
        if (downdate) {
            double qo[CDFT_MAXM];
            for (int l = 0; l < m; l++) qo[l] = Q[(size_t)l * n + i];
            for (int j = 0; j < m - 1; j++) {
                double a = 0.0;
                for (int l = 0; l < m; l++) a += qo[l] * qx.a[j * m + l];
                qv[j] = a;
                Qout[(size_t)j * n + i] = a;
            }
            Qout[(size_t)(m - 1) * n + i] = 0.0;
        } else {
            for (int j = 0; j < kq; j++) qv[j] = Q[(size_t)j * n + i];
        }
        double dv = df[i], rv = res[i];
        for (int j = 0; j < kq; j++) {
            v[j] = qv[j] * dv;
            v[CDFT_MAXM + j] = qv[j] * rv;
        }
    }
    block_partials<2 * CDFT_MAXM>(v, red, nblk);
}

// pass 2/3: df -= Q' h (h from the finalized reductions); then partials of
// Q'^T df (pass 2) or |df|^2 and df.res (pass 3).
__global__ __launch_bounds__(256) void k_aa_proj2(int n, const double *__restrict__ Q, int kq,
                                                  double *__restrict__ df, const double *__restrict__ res,
                                                  const double *__restrict__ hcoef, int last, double *red, int nblk)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[CDFT_MAXM];
    for (int q = 0; q < CDFT_MAXM; q++) v[q] = 0.0;
    if (i < n) {
        double qv[CDFT_MAXM];
        double a = 0.0;
        for (int j = 0; j < kq; j++) {
            qv[j] = Q[(size_t)j * n + i];
            a += qv[j] * hcoef[j];
        }
        double d = df[i] - a;
        df[i] = d;
        if (!last) {
            for (int j = 0; j < kq; j++) v[j] = qv[j] * d;
        } else {
            v[0] = d * d;
            v[1] = d * res[i];
        }
    }
    block_partials<CDFT_MAXM>(v, red, nblk);
}

// New Q column, the Anderson point sol = gx - DG theta, and rho = exp(sol).
__global__ __launch_bounds__(256) void k_aa_newpoint(int n, const double *__restrict__ df, double nrm,
                                                     double *__restrict__ qcol, double *__restrict__ sol,
                                                     const double *__restrict__ gx, const double *__restrict__ DG,
                                                     AAMat th, int mk, double *__restrict__ rho,
                                                     const double *__restrict__ mask)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (qcol) qcol[i] = df[i] / nrm;   // vnew ./= hv[k+1] (anderson.jl:538)
    double a = 0.0;
    for (int j = 0; j < mk; j++) a += DG[(size_t)th.col[j] * n + i] * th.a[j];
    double x = gx[i] - a;
    sol[i] = x;
    rho[i] = mask ? mask[i] * x : exp(x);
}

}  // namespace cdft
