// batched.hip -- persistent batched Euler-Lagrange solver for 1D grids.
//
// One workgroup owns one system for its whole solve: map evaluations, Anderson
// history and the small QR problems never leave the device, and a system that
// terminates simply ends its workgroup (the next queued system takes the CU).
// The algorithm per system is exactly cdft_converge's (converge!/aasol,
// src/methods/converge.jl:214-254, src/utils/anderson.jl:301-443).
//
// Convolutions run in LDS as FFTs: the reference's spectral operator
// Re IFFT_N(FFT_N(u) * Omega_f) equals the circular convolution h_f (*) u with the
// real kernel h_f = IFFT_N(Omega_herm_f).  With M = N (N a power of two) it is
// applied directly as IFFT_M(FFT_M(u) * H_f), H_f = DFT_N(h_f); otherwise through a
// zero-padded power-of-two length M >= 2N - 1 (linear convolution, then the period-N
// fold out[i] = y[i] + y[i + N]), which is exact for every N.
//
// Precision: CDFT_FP64 runs the whole map in FP64.  CDFT_MIXED runs the FFTs, the
// weighted densities and Phi in FP32 (Phi's n3-singular factors in FP64) until the
// damped residual reaches the FP32 noise floor, then restarts the Anderson phase
// with the FP64 map and finishes to the requested tolerance with it; exponent
// assembly, residual norms and the Anderson solve are FP64 throughout.
#pragma once
#include "common.h"
#include "kernels.h"

namespace cdft {

template <typename T>
struct c2 {
    T x, y;
};

// Per-system descriptor of a batched launch (device memory).
struct BSys {
    int N, nc, M, log2M;
    double NA, beta_unused;
    double cp[CDFT_MAXC];
    const double *vext;        // [c*N + i] or null
    const double *H64;         // spectra [(c*(F-1) + f-1)*M + k] as (re, im) pairs
    const float *H32;
    const double *tw64;        // W_M^t, t < M, (re, im) pairs
    const float *tw32;
    double *rho;               // in: seed, out: converged profile  [c*N + i]
    double *sol, *gx, *df, *dg, *res, *Q, *DG;   // Anderson storage, n = N*nc each (Q, DG: CDFT_MAXM cols)
    PhiConst pc;
};

struct BOpts {
    int maxit, m;
    double beta, tol, anderson_start;
    int mixed;                 // 1: FP32 map until the noise floor, then FP64
    double switch_factor;      // FP32 -> FP64 when resnorm <= switch_factor * tol
    int stall_window;          // ... or when it has not halved in this many iterations
};

struct BResult {
    int idid, errcode, iters, nevals, nevals32, pad;
    double resnorm, tol;
    unsigned long long ticks;   // wall_clock64 ticks spent in the workgroup
};

constexpr int BT = 256;        // threads per workgroup

// ---------------------------------------------------------------- helpers
template <typename T>
__device__ __attribute__((always_inline)) inline c2<T> cmul(c2<T> a, c2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}


#define CDFT_INLINE __device__ __attribute__((always_inline)) inline

// In-place Stockham FFT of length M (power of two, 4 <= M <= 1024 * MAXB) in LDS:
// radix-4 passes and a final radix-2 pass when log2 M is odd.  Each pass loads
// all its butterfly inputs into registers, synchronises, then stores, so one
// buffer suffices.  Twiddles W_M^t come from the LDS table tw.  Unnormalised;
// sign -1 forward, +1 inverse.
template <typename T, int MAXB>
CDFT_INLINE void fft_lds(c2<T> *a, const c2<double> *tw, int M, int log2M, bool inverse)
{
    const int tid = threadIdx.x;
    int Ns = 1;
    const int nr4 = log2M >> 1;
    const T sg = inverse ? T(1) : T(-1);
    for (int pass = 0; pass < nr4; pass++) {
        c2<T> v[MAXB][4];
        const int nb = M >> 2;
        const int step = M / (Ns * 4);
#pragma unroll
        for (int b = 0; b < MAXB; b++) {
            int j = tid + b * BT;
            if (j < nb) {
                int k = j & (Ns - 1);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    c2<T> x = a[j + r * nb];
                    if (r) {
                        c2<double> wd = tw[(r * k * step) & (M - 1)];
                        c2<T> w = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
                        x = cmul(x, w);
                    }
                    v[b][r] = x;
                }
                c2<T> s02 = {v[b][0].x + v[b][2].x, v[b][0].y + v[b][2].y};
                c2<T> d02 = {v[b][0].x - v[b][2].x, v[b][0].y - v[b][2].y};
                c2<T> s13 = {v[b][1].x + v[b][3].x, v[b][1].y + v[b][3].y};
                c2<T> d13 = {v[b][1].x - v[b][3].x, v[b][1].y - v[b][3].y};
                c2<T> id13 = {-sg * d13.y, sg * d13.x};   // (sg i) * d13
                v[b][0] = {s02.x + s13.x, s02.y + s13.y};
                v[b][1] = {d02.x + id13.x, d02.y + id13.y};
                v[b][2] = {s02.x - s13.x, s02.y - s13.y};
                v[b][3] = {d02.x - id13.x, d02.y - id13.y};
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < MAXB; b++) {
            int j = tid + b * BT;
            if (j < nb) {
                int k = j & (Ns - 1);
                int d = (j / Ns) * Ns * 4 + k;
#pragma unroll
                for (int r = 0; r < 4; r++) a[d + r * Ns] = v[b][r];
            }
        }
        __syncthreads();
        Ns <<= 2;
    }
    if (log2M & 1) {
        c2<T> v[2 * MAXB][2];
        const int nb = M >> 1;
#pragma unroll
        for (int b = 0; b < 2 * MAXB; b++) {
            int j = tid + b * BT;
            if (j < nb) {
                int k = j & (Ns - 1);
                c2<double> wd = tw[(k * (M / (Ns * 2))) & (M - 1)];
                c2<T> w = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
                c2<T> x0 = a[j], x1 = cmul(a[j + nb], w);
                v[b][0] = {x0.x + x1.x, x0.y + x1.y};
                v[b][1] = {x0.x - x1.x, x0.y - x1.y};
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 2 * MAXB; b++) {
            int j = tid + b * BT;
            if (j < nb) {
                int k = j & (Ns - 1);
                int d = (j / Ns) * Ns * 2 + k;
                a[d] = v[b][0];
                a[d + Ns] = v[b][1];
            }
        }
        __syncthreads();
    }
}

// Sum of NQ values over the workgroup; every thread receives the same totals
// (fixed summation order, so the result is reproducible and uniform).
template <int NQ>
CDFT_INLINE void wg_sum(double (&v)[NQ], double *red)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[q] = x;
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) red[q * 4 + wid] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = (red[q * 4 + 0] + red[q * 4 + 1]) + (red[q * 4 + 2] + red[q * 4 + 3]);
    __syncthreads();
}

// One system's data, read once from its descriptor into registers.
template <int NC>
struct SysRegs {
    int N, M, lg, n;
    double NA;
    double cp[NC];
    const double *__restrict__ vext;
    const double *__restrict__ H64;
    const float *__restrict__ H32;
    double *__restrict__ rho;
    double *__restrict__ sol;
    double *__restrict__ gx;
    double *__restrict__ df;
    double *__restrict__ dg;
    double *__restrict__ res;
    double *__restrict__ Q;
    double *__restrict__ DG;
};

template <typename T>
struct HSel;
template <>
struct HSel<double> {
    static CDFT_INLINE const double *get(const double *h64, const float *) { return h64; }
};
template <>
struct HSel<float> {
    static CDFT_INLINE const float *get(const double *, const float *h32) { return h32; }
};

// One evaluation of the fixed-point map plus the Anderson book-keeping of mode
// (1: first/Picard step, 2: aa_point!).  out = |res|^2, |df|^2, df.res.
// LDS: tw (M, FP64), work A and accumulator P (M complex T each, at FP64
// stride), fields fld (F*N*NC of T at FP64 stride), red (64 doubles).
template <typename T, int NC, int MAXB>
CDFT_INLINE void evalG(const SysRegs<NC> &S, const PhiConst &P, const c2<double> *tw, c2<T> *A, c2<T> *Pc, T *fld,
                       double *red, int mode, double b, double (&out)[3])
{
    constexpr int F = 7;
    const int tid = threadIdx.x;
    const int N = S.N, M = S.M, lg = S.lg;
    const bool fold = (M != N);
    const T *H = HSel<T>::get(S.H64, S.H32);
    const T invM = T(1) / T(M);

    for (int c = 0; c < NC; c++) {
        // rho = exp(sol): slot 0 of the fields and the zero-padded FFT input
        for (int k = tid; k < M; k += BT) {
            if (k < N) {
                double r = exp(S.sol[c * N + k]);
                fld[(c * F) * N + k] = T(S.NA * r);
                Pc[k] = {T(r), T(0)};
            } else {
                Pc[k] = {T(0), T(0)};
            }
        }
        __syncthreads();
        fft_lds<T, MAXB>(Pc, tw, M, lg, false);
        for (int q = 0; q < 3; q++) {
            const int f1 = 1 + 2 * q, f2 = 2 + 2 * q;
            const T *H1 = H + (size_t)(c * (F - 1) + f1 - 1) * M * 2;
            const T *H2 = H + (size_t)(c * (F - 1) + f2 - 1) * M * 2;
            for (int k = tid; k < M; k += BT) {
                c2<T> p = Pc[k];
                c2<T> a1 = cmul(p, c2<T>{H1[2 * k], H1[2 * k + 1]});
                c2<T> a2 = cmul(p, c2<T>{H2[2 * k], H2[2 * k + 1]});
                A[k] = {a1.x - a2.y, a1.y + a2.x};          // P H1 + i P H2
            }
            __syncthreads();
            fft_lds<T, MAXB>(A, tw, M, lg, true);
            const T sc = T(S.NA) * invM;
            for (int i = tid; i < N; i += BT) {
                c2<T> y = A[i];
                if (fold && i + N < M) { y.x += A[i + N].x; y.y += A[i + N].y; }
                fld[(c * F + f1) * N + i] = sc * y.x;
                fld[(c * F + f2) * N + i] = sc * y.y;
            }
            __syncthreads();
        }
    }
    // Phi and dPhi/dn, in place over the field slots
    for (int i = tid; i < N; i += BT) {
        T nl[NC][F], g[NC][F];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int f = 0; f < F; f++) nl[c][f] = fld[(c * F + f) * N + i];
        phi_grad<T, double, NC, 1>(P, nl, g);
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int f = 0; f < F; f++) fld[(c * F + f) * N + i] = g[c][f];
    }
    __syncthreads();
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int c = 0; c < NC; c++) {
        for (int k = tid; k < M; k += BT) Pc[k] = {T(0), T(0)};
        for (int q = 0; q < 3; q++) {
            const int f1 = 1 + 2 * q, f2 = 2 + 2 * q;
            const T s1 = (f1 == 3) ? T(-1) : T(1), s2 = (f2 == 3) ? T(-1) : T(1);   // vector slot subtracts
            const T *H1 = H + (size_t)(c * (F - 1) + f1 - 1) * M * 2;
            const T *H2 = H + (size_t)(c * (F - 1) + f2 - 1) * M * 2;
            for (int k = tid; k < M; k += BT)
                A[k] = (k < N) ? c2<T>{fld[(c * F + f1) * N + k], fld[(c * F + f2) * N + k]} : c2<T>{T(0), T(0)};
            __syncthreads();
            fft_lds<T, MAXB>(A, tw, M, lg, false);
            for (int k = tid; k < M; k += BT) {
                int km = (M - k) & (M - 1);
                c2<T> a = A[k], bb = A[km];
                c2<T> u1 = {T(0.5) * (a.x + bb.x), T(0.5) * (a.y - bb.y)};     // (a + conj b)/2
                c2<T> u2 = {T(0.5) * (a.y + bb.y), T(-0.5) * (a.x - bb.x)};    // (a - conj b)/(2i)
                c2<T> t1 = cmul(u1, c2<T>{H1[2 * k], H1[2 * k + 1]});
                c2<T> t2 = cmul(u2, c2<T>{H2[2 * k], H2[2 * k + 1]});
                c2<T> acc = Pc[k];
                acc.x += s1 * t1.x + s2 * t2.x;
                acc.y += s1 * t1.y + s2 * t2.y;
                Pc[k] = acc;
            }
            __syncthreads();
        }
        fft_lds<T, MAXB>(Pc, tw, M, lg, true);
        for (int i = tid; i < N; i += BT) {
            T y = Pc[i].x;
            if (fold && i + N < M) y += Pc[i + N].x;
            double s = double(fld[(c * F) * N + i]) + double(y * invM);
            const int t = c * N + i;
            if (S.vext) s += S.vext[t];
            double lgv = S.cp[c] - s;
            lgv = fmin(fmax(lgv, -100.0), 100.0);          // clamp!(ln_Gx, -100, 100)
            double x = S.sol[t];
            double gnew = (b == 1.0) ? lgv : (1.0 - b) * x + b * lgv;
            if (mode == 1) {
                S.gx[t] = gnew;
                double r = gnew - x;
                S.res[t] = r;
                v0 += r * r;
            } else {
                double gold = S.gx[t], rold = S.res[t];
                S.dg[t] = -gold + gnew;
                S.gx[t] = gnew;
                double r = gnew - x;
                S.res[t] = r;
                double d = -rold + r;
                S.df[t] = d;
                v0 += r * r;
                v1 += d * d;
                v2 += d * r;
            }
        }
        __syncthreads();
    }
    double v[3] = {v0, v1, v2};
    wg_sum<3>(v, red);
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

// Householder downdate of the Anderson R (same arithmetic as the host version).
CDFT_INLINE void dev_householder_downdate(double *R, int m, double *Qx)
{
    int nr = m, ncol = m - 1;
    double A[CDFT_MAXM * CDFT_MAXM], tau[CDFT_MAXM], v[CDFT_MAXM];
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) A[j * CDFT_MAXM + i] = R[(j + 1) * m + i];
    for (int j = 0; j < ncol; j++) {
        double alpha = A[j * CDFT_MAXM + j], xn = 0.0;
        for (int i = j + 1; i < nr; i++) xn += A[j * CDFT_MAXM + i] * A[j * CDFT_MAXM + i];
        xn = sqrt(xn);
        if (xn == 0.0) { tau[j] = 0.0; continue; }
        double beta = -copysign(hypot(alpha, xn), alpha);
        tau[j] = (beta - alpha) / beta;
        double sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i < nr; i++) A[j * CDFT_MAXM + i] *= sc;
        A[j * CDFT_MAXM + j] = beta;
        for (int jj = j + 1; jj < ncol; jj++) {
            double w = A[jj * CDFT_MAXM + j];
            for (int i = j + 1; i < nr; i++) w += A[j * CDFT_MAXM + i] * A[jj * CDFT_MAXM + i];
            w *= tau[j];
            A[jj * CDFT_MAXM + j] -= w;
            for (int i = j + 1; i < nr; i++) A[jj * CDFT_MAXM + i] -= w * A[j * CDFT_MAXM + i];
        }
    }
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nr; i++) Qx[j * nr + i] = (i == j) ? 1.0 : 0.0;
    for (int h = ncol - 1; h >= 0; h--) {
        if (tau[h] == 0.0) continue;
        v[h] = 1.0;
        for (int i = h + 1; i < nr; i++) v[i] = A[h * CDFT_MAXM + i];
        for (int j = 0; j < ncol; j++) {
            double w = 0.0;
            for (int i = h; i < nr; i++) w += v[i] * Qx[j * nr + i];
            w *= tau[h];
            for (int i = h; i < nr; i++) Qx[j * nr + i] -= w * v[i];
        }
    }
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < ncol; i++) R[j * m + i] = (i <= j) ? A[j * CDFT_MAXM + i] : 0.0;
    for (int i = 0; i < m; i++) R[(m - 1) * m + i] = 0.0;
}

// Shared scalar state of one system's solve (LDS).
struct AState {
    double R[CDFT_MAXM * CDFT_MAXM];
    double Qx[CDFT_MAXM * CDFT_MAXM];
    double theta[CDFT_MAXM];
    double tres[CDFT_MAXM];
    int ring[CDFT_MAXM];
};

// Dynamic LDS bytes of one system (FP64 layout; the FP32 map uses the same slots).
__host__ __device__ inline size_t batched_lds_layout(int N, int M, int nc)
{
    return (size_t)3 * M * 16 + (size_t)7 * N * nc * 8 + 64 * 8;
}

template <int NC, int MAXB>
__global__ __launch_bounds__(BT) void k_batched(const BSys *__restrict__ systems, BOpts o, BResult *results)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ AState st;
    __shared__ PhiConst pcs;
    constexpr int F = 7;
    const int tid = threadIdx.x;
    const unsigned long long t_start = wall_clock64();

    // ---- descriptor -> registers / LDS, once
    const BSys *D = systems + blockIdx.x;
    SysRegs<NC> S;
    S.N = D->N;
    S.M = D->M;
    S.lg = D->log2M;
    S.n = S.N * NC;
    S.NA = D->NA;
#pragma unroll
    for (int c = 0; c < NC; c++) S.cp[c] = D->cp[c];
    S.vext = D->vext;
    S.H64 = D->H64;
    S.H32 = D->H32;
    S.rho = D->rho;
    S.sol = D->sol;
    S.gx = D->gx;
    S.df = D->df;
    S.dg = D->dg;
    S.res = D->res;
    S.Q = D->Q;
    S.DG = D->DG;
    const double *__restrict__ twg = D->tw64;
    {
        const double *src = reinterpret_cast<const double *>(&D->pc);
        double *dst = reinterpret_cast<double *>(&pcs);
        for (int i = tid; i < (int)(sizeof(PhiConst) / sizeof(double)); i += BT) dst[i] = src[i];
    }
    const int N = S.N, M = S.M, n = S.n;
    c2<double> *tw = reinterpret_cast<c2<double> *>(smem);
    c2<double> *A64 = tw + M;
    c2<double> *P64 = A64 + M;
    double *fld64 = reinterpret_cast<double *>(P64 + M);
    double *red = fld64 + (size_t)F * N * NC;
    c2<float> *A32 = reinterpret_cast<c2<float> *>(A64);
    c2<float> *P32 = reinterpret_cast<c2<float> *>(P64);
    float *fld32 = reinterpret_cast<float *>(fld64);
    for (int t = tid; t < M; t += BT) tw[t] = {twg[2 * t], twg[2 * t + 1]};

    const int m = o.m;
    const double beta = o.beta;
    const int maxit = o.maxit, picard_maxit = o.maxit;
    const double rtol = o.tol, atol = o.tol, picard_beta = beta;
    const double picard_rtol = o.anderson_start, picard_atol = o.anderson_start;
    bool fp32 = (o.mixed != 0);
    int nevals = 0, nevals32 = 0;
    double fo[3];

    for (int i = tid; i < n; i += BT) S.sol[i] = log(S.rho[i]);   // x0 = log(rho)
    __syncthreads();

    // aasol (anderson.jl:301-443) as a phase machine with a single map-evaluation
    // site per precision.  ka is aasol's Anderson counter since the last (re)start;
    // k counts every Anderson step for the maxit bound.
    enum { PH_FIRST, PH_PICARD, PH_RESTART, PH_STEP, PH_SWITCH, PH_DONE };
    int phase = PH_FIRST;
    int k = 0, ka = 0, pit = 0, since_best = 0;
    double resnorm = 0.0, up_bd = 0.0, tol = 0.0, picard_tol = 0.0, best = 0.0, dfn2 = 0.0, dfres = 0.0;
    bool toosoon = false;
    while (phase != PH_DONE) {
        // ---------------- what to evaluate next
        int mode = 1;
        double bb = beta;
        if (phase == PH_PICARD) {
            bb = picard_beta;
        } else if (phase == PH_RESTART) {
            if (tid == 0) {
                for (int j = 0; j < CDFT_MAXM * CDFT_MAXM; j++) st.R[j] = 0.0;
                for (int j = 0; j < CDFT_MAXM; j++) st.ring[j] = j;
            }
            for (int i = tid; i < n; i += BT) S.sol[i] = S.gx[i];   // copy!(sol, gx)
            __syncthreads();
            k += 1;
            ka = 1;
            mode = 2;
        } else if (phase == PH_SWITCH) {
            fp32 = false;                                           // FP32 noise floor: FP64 map from here
        } else if (phase == PH_STEP) {
            if (m == 0) {
                for (int i = tid; i < n; i += BT) S.sol[i] = S.gx[i];
                __syncthreads();
            } else {
                const int kk = ka + 1;
                const int kq = ka - 1;
                if (tid == 0 && m > 1 && kk > m + 1) {               // BuildDG! shift, as a ring
                    int first = st.ring[0];
                    for (int j = 0; j < m - 1; j++) st.ring[j] = st.ring[j + 1];
                    st.ring[m - 1] = first;
                }
                __syncthreads();
                const int slot = (m == 1) ? st.ring[0] : (kk > m + 1 ? st.ring[m - 1] : st.ring[kk - 2]);
                for (int i = tid; i < n; i += BT) S.DG[(size_t)slot * n + i] = S.dg[i];
                double nrm;
                int qcol;
                if (kq == 0) {                                      // aa_qr_update! first column
                    nrm = sqrt(dfn2);
                    if (tid == 0) {
                        st.R[0] = nrm;
                        st.tres[0] = dfres / nrm;
                    }
                    qcol = 0;
                } else {
                    const int down = (kq > m - 1);
                    if (down && tid == 0) dev_householder_downdate(st.R, m, st.Qx);
                    __syncthreads();
                    const int kqq = kq < m - 1 ? kq : m - 1;
                    // pass 1: optional downdate of Q (pointwise, in place); h1 = Q^T df, t = Q^T res
                    double v1[2 * CDFT_MAXM];
#pragma unroll
                    for (int j = 0; j < 2 * CDFT_MAXM; j++) v1[j] = 0.0;
                    for (int i = tid; i < n; i += BT) {
                        double qv[CDFT_MAXM];
                        if (down) {
                            double qo[CDFT_MAXM];
#pragma unroll
                            for (int l = 0; l < CDFT_MAXM; l++) qo[l] = (l < m) ? S.Q[(size_t)l * n + i] : 0.0;
#pragma unroll
                            for (int j = 0; j < CDFT_MAXM - 1; j++) {
                                double a = 0.0;
#pragma unroll
                                for (int l = 0; l < CDFT_MAXM; l++)
                                    if (l < m) a += qo[l] * st.Qx[j * m + l];
                                qv[j] = a;
                                if (j < m - 1) S.Q[(size_t)j * n + i] = a;
                            }
                            S.Q[(size_t)(m - 1) * n + i] = 0.0;
                        } else {
#pragma unroll
                            for (int j = 0; j < CDFT_MAXM - 1; j++) qv[j] = (j < kqq) ? S.Q[(size_t)j * n + i] : 0.0;
                        }
                        const double dv = S.df[i], rv = S.res[i];
#pragma unroll
                        for (int j = 0; j < CDFT_MAXM - 1; j++)
                            if (j < kqq) {
                                v1[j] += qv[j] * dv;
                                v1[CDFT_MAXM + j] += qv[j] * rv;
                            }
                    }
                    wg_sum<2 * CDFT_MAXM>(v1, red);
                    // pass 2: df -= Q h1; h2 = Q^T df
                    double v2[CDFT_MAXM];
#pragma unroll
                    for (int j = 0; j < CDFT_MAXM; j++) v2[j] = 0.0;
                    for (int i = tid; i < n; i += BT) {
                        double qv[CDFT_MAXM], a = 0.0;
#pragma unroll
                        for (int j = 0; j < CDFT_MAXM - 1; j++)
                            if (j < kqq) {
                                qv[j] = S.Q[(size_t)j * n + i];
                                a += qv[j] * v1[j];
                            }
                        const double d = S.df[i] - a;
                        S.df[i] = d;
#pragma unroll
                        for (int j = 0; j < CDFT_MAXM - 1; j++)
                            if (j < kqq) v2[j] += qv[j] * d;
                    }
                    wg_sum<CDFT_MAXM>(v2, red);
                    // pass 3: df -= Q h2; |df|^2, df.res
                    double v3[2] = {0.0, 0.0};
                    for (int i = tid; i < n; i += BT) {
                        double a = 0.0;
#pragma unroll
                        for (int j = 0; j < CDFT_MAXM - 1; j++)
                            if (j < kqq) a += S.Q[(size_t)j * n + i] * v2[j];
                        const double d = S.df[i] - a;
                        S.df[i] = d;
                        v3[0] += d * d;
                        v3[1] += d * S.res[i];
                    }
                    wg_sum<2>(v3, red);
                    nrm = sqrt(v3[0]);
                    if (tid == 0) {
                        for (int j = 0; j < kqq; j++) st.R[kqq * m + j] = v1[j] + v2[j];
                        st.R[kqq * m + kqq] = nrm;
                        for (int j = 0; j < kqq; j++) st.tres[j] = v1[CDFT_MAXM + j];
                        st.tres[kqq] = v3[1] / nrm;
                    }
                    qcol = kqq;
                }
                const int mk = m < ka ? m : ka;
                if (tid == 0) {
                    for (int j = mk - 1; j >= 0; j--) {             // theta = RA \ tres
                        double acc = st.tres[j];
                        for (int l = j + 1; l < mk; l++) acc -= st.R[l * m + j] * st.theta[l];
                        st.theta[j] = acc / st.R[j * m + j];
                    }
                }
                __syncthreads();
                double th[CDFT_MAXM];
                int rg[CDFT_MAXM];
#pragma unroll
                for (int j = 0; j < CDFT_MAXM; j++) {
                    th[j] = (j < mk) ? st.theta[j] : 0.0;
                    rg[j] = (j < mk) ? st.ring[j] : 0;
                }
                // new Q column; sol = gx - DG theta
                for (int i = tid; i < n; i += BT) {
                    S.Q[(size_t)qcol * n + i] = S.df[i] / nrm;
                    double a = 0.0;
#pragma unroll
                    for (int j = 0; j < CDFT_MAXM; j++)
                        if (j < mk) a += S.DG[(size_t)rg[j] * n + i] * th[j];
                    S.sol[i] = S.gx[i] - a;
                }
                __syncthreads();
            }
            k += 1;
            ka += 1;
            mode = 2;
        }

        // ---------------- the one map evaluation
        if (fp32) {
            evalG<float, NC, MAXB>(S, pcs, tw, A32, P32, fld32, red, mode, bb, fo);
            nevals32++;
        } else {
            evalG<double, NC, MAXB>(S, pcs, tw, A64, P64, fld64, red, mode, bb, fo);
        }
        nevals++;
        resnorm = sqrt(fo[0]);

        // ---------------- book-keeping after it
        bool check_loop = false;
        if (phase == PH_FIRST) {
            up_bd = 1.e4 * resnorm;
            tol = rtol * resnorm + atol;
            picard_tol = picard_rtol * resnorm + picard_atol;
            toosoon = (resnorm <= tol);
            if (toosoon) {
                phase = PH_DONE;
            } else if (picard_maxit > 0) {
                for (int i = tid; i < n; i += BT) S.sol[i] = S.gx[i];
                __syncthreads();
                pit = 0;
                phase = (picard_maxit - 1 > 0) ? PH_PICARD : PH_RESTART;
            } else {
                phase = PH_RESTART;
            }
        } else if (phase == PH_PICARD) {
            pit++;
            const bool brk = (resnorm <= picard_tol || resnorm >= up_bd);
            if (!brk) {
                for (int i = tid; i < n; i += BT) S.sol[i] = S.gx[i];
                __syncthreads();
            }
            if (brk || pit >= picard_maxit - 1) {
                toosoon = (resnorm <= tol);
                phase = toosoon ? PH_DONE : PH_RESTART;
            }
        } else if (phase == PH_SWITCH) {
            phase = (resnorm <= tol) ? PH_DONE : PH_RESTART;
        } else {                                                    // PH_RESTART or PH_STEP
            dfn2 = fo[1];
            dfres = fo[2];
            if (phase == PH_RESTART) {
                best = resnorm;
                since_best = 0;
            }
            check_loop = true;
        }
        if (check_loop) {
            // while (k < maxit + picard_maxit && resnorm > tol && !toosoon && resnorm < up_bd)
            phase = PH_STEP;
            if (!(k < maxit + picard_maxit && resnorm < up_bd)) {
                phase = PH_DONE;
            } else if (fp32) {
                if (resnorm < 0.5 * best) {
                    best = resnorm;
                    since_best = 0;
                } else {
                    since_best++;
                }
                if (resnorm <= o.switch_factor * tol || since_best >= o.stall_window) phase = PH_SWITCH;
            } else if (!(resnorm > tol)) {
                phase = PH_DONE;
            }
        }
    }

    // a mixed solve that ran out of iterations in FP32 reports non-convergence
    int idid, errcode;
    if (toosoon) { idid = 1; errcode = -1; }
    else if (resnorm >= up_bd) { idid = 0; errcode = -2; }
    else if (resnorm <= tol && !fp32) { idid = 1; errcode = 0; }
    else { idid = 0; errcode = 10; }
    for (int i = tid; i < n; i += BT) S.rho[i] = exp(S.sol[i]);
    if (tid == 0) {
        BResult r;
        r.idid = idid;
        r.errcode = errcode;
        r.iters = k;
        r.nevals = nevals;
        r.nevals32 = nevals32;
        r.pad = 0;
        r.resnorm = resnorm;
        r.tol = tol;
        r.ticks = wall_clock64() - t_start;
        results[blockIdx.x] = r;
    }
}

// ---------------------------------------------------------------- set-up kernels
// H[(c*(F-1)+f-1)*M + k] = sum_{j<N} h_f[j] exp(-2 pi i j k / M)   (FP64, exact twiddles)
__global__ void k_setup_spec(const double *h, int N, int M, int F, double *H64, float *H32)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int cf = blockIdx.y;                     // c*(F-1) + f-1
    if (k >= M) return;
    int c = cf / (F - 1), f = 1 + cf % (F - 1);
    const double *hk = h + (size_t)(c * F + f) * N;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < N; j++) {
        long r = ((long)j * k) % M;
        double sn, cs;
        sincospi(2.0 * (double)r / (double)M, &sn, &cs);
        re += hk[j] * cs;
        im -= hk[j] * sn;
    }
    size_t o = ((size_t)cf * M + k) * 2;
    H64[o] = re;
    H64[o + 1] = im;
    H32[o] = (float)re;
    H32[o + 1] = (float)im;
}

__global__ void k_setup_tw(int M, double *tw64, float *tw32)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    double sn, cs;
    sincospi(2.0 * (double)t / (double)M, &sn, &cs);
    tw64[2 * t] = cs;
    tw64[2 * t + 1] = -sn;
    tw32[2 * t] = (float)cs;
    tw32[2 * t + 1] = (float)-sn;
}

}  // namespace cdft
