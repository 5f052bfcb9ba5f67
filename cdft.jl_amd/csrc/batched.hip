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
};

constexpr int BT = 256;        // threads per workgroup

// ---------------------------------------------------------------- helpers
template <typename T>
__device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T>
__device__ __forceinline__ const T *tw_of(const BSys &S);
template <>
__device__ __forceinline__ const double *tw_of<double>(const BSys &S) { return S.tw64; }
template <>
__device__ __forceinline__ const float *tw_of<float>(const BSys &S) { return S.tw32; }
template <typename T>
__device__ __forceinline__ const T *H_of(const BSys &S);
template <>
__device__ __forceinline__ const double *H_of<double>(const BSys &S) { return S.H64; }
template <>
__device__ __forceinline__ const float *H_of<float>(const BSys &S) { return S.H32; }

// In-place Stockham FFT of length M (power of two, 4 <= M <= 4096) in LDS:
// radix-4 passes and a final radix-2 pass when log2 M is odd.  Each pass loads
// all its butterfly inputs into registers, synchronises, then stores, so one
// buffer suffices.  Unnormalised; sign -1 forward, +1 inverse.
template <typename T>
__device__ void fft_lds(c2<T> *a, int M, int log2M, const T *tw, bool inverse)
{
    constexpr int MAXB = 4;                       // butterflies per thread (M <= 4096)
    const int tid = threadIdx.x;
    int Ns = 1;
    const int nr4 = log2M >> 1;
    const T sg = inverse ? T(1) : T(-1);
    for (int pass = 0; pass < nr4; pass++) {
        c2<T> v[MAXB][4];
        const int nb = M >> 2;
#pragma unroll
        for (int b = 0; b < MAXB; b++) {
            int j = tid + b * BT;
            if (j < nb) {
                int k = j & (Ns - 1);
                int step = M / (Ns * 4);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    c2<T> x = a[j + r * nb];
                    if (r) {
                        int t = (r * k * step) & (M - 1);
                        c2<T> w = {tw[2 * t], inverse ? -tw[2 * t + 1] : tw[2 * t + 1]};
                        x = cmul(x, w);
                    }
                    v[b][r] = x;
                }
                // y_q = sum_r v_r (sg i)^{r q}
                c2<T> s02 = {v[b][0].x + v[b][2].x, v[b][0].y + v[b][2].y};
                c2<T> d02 = {v[b][0].x - v[b][2].x, v[b][0].y - v[b][2].y};
                c2<T> s13 = {v[b][1].x + v[b][3].x, v[b][1].y + v[b][3].y};
                c2<T> d13 = {v[b][1].x - v[b][3].x, v[b][1].y - v[b][3].y};
                // (sg i) * d13 = (-sg d13.y, sg d13.x)
                c2<T> id13 = {-sg * d13.y, sg * d13.x};
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
                int t = (k * (M / (Ns * 2))) & (M - 1);
                c2<T> w = {tw[2 * t], inverse ? -tw[2 * t + 1] : tw[2 * t + 1]};
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
__device__ __forceinline__ void wg_sum(double (&v)[NQ], double *red)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double x = v[q];
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

// LDS layout (bytes): work[M] c2<T> | acc[M] c2<T> | fld[F*N*nc] T | red[64] double | misc
template <typename T>
struct Lds {
    c2<T> *A, *P;
    T *fld;
    double *red;
};

// One evaluation of the fixed-point map plus the Anderson book-keeping of mode
// (1: first/Picard step, 2: aa_point!).  Returns |res|^2 (and |df|^2, df.res).
template <typename T, int NC>
__device__ void evalG(const BSys &S, const Lds<T> &L, int mode, double b, double (&out)[3])
{
    constexpr int F = 7;
    const int tid = threadIdx.x;
    const int N = S.N, M = S.M, lg = S.log2M;
    const bool fold = (M != N);
    const T *tw = tw_of<T>(S);
    const T *H = H_of<T>(S);
    const T NA = T(S.NA);
    const T invM = T(1) / T(M);

    for (int c = 0; c < NC; c++) {
        // rho = exp(sol); slot 0 of the fields and the padded FFT input
        for (int k = tid; k < M; k += BT) {
            if (k < N) {
                double r = exp(S.sol[c * N + k]);
                L.fld[(c * F) * N + k] = T(S.NA * r);
                L.A[k] = {T(r), T(0)};
            } else {
                L.A[k] = {T(0), T(0)};
            }
        }
        __syncthreads();
        fft_lds<T>(L.A, M, lg, tw, false);
        for (int k = tid; k < M; k += BT) L.P[k] = L.A[k];
        __syncthreads();
        for (int q = 0; q < 3; q++) {
            const int f1 = 1 + 2 * q, f2 = 2 + 2 * q;
            const T *H1 = H + (size_t)(c * (F - 1) + f1 - 1) * M * 2;
            const T *H2 = H + (size_t)(c * (F - 1) + f2 - 1) * M * 2;
            for (int k = tid; k < M; k += BT) {
                c2<T> p = L.P[k];
                c2<T> a1 = cmul(p, c2<T>{H1[2 * k], H1[2 * k + 1]});
                c2<T> a2 = cmul(p, c2<T>{H2[2 * k], H2[2 * k + 1]});
                L.A[k] = {a1.x - a2.y, a1.y + a2.x};
            }
            __syncthreads();
            fft_lds<T>(L.A, M, lg, tw, true);
            for (int i = tid; i < N; i += BT) {
                c2<T> y = L.A[i];
                if (fold && i + N < M) { y.x += L.A[i + N].x; y.y += L.A[i + N].y; }
                L.fld[(c * F + f1) * N + i] = NA * (y.x * invM);
                L.fld[(c * F + f2) * N + i] = NA * (y.y * invM);
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
            for (int f = 0; f < F; f++) nl[c][f] = L.fld[(c * F + f) * N + i];
        phi_grad<T, double, NC, 1>(S.pc, nl, g);
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int f = 0; f < F; f++) L.fld[(c * F + f) * N + i] = g[c][f];
    }
    __syncthreads();
    double v[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < NC; c++) {
        for (int k = tid; k < M; k += BT) L.P[k] = {T(0), T(0)};
        for (int q = 0; q < 3; q++) {
            const int f1 = 1 + 2 * q, f2 = 2 + 2 * q;
            const T s1 = (f1 == 3) ? T(-1) : T(1), s2 = (f2 == 3) ? T(-1) : T(1);   // vector slot subtracts
            const T *H1 = H + (size_t)(c * (F - 1) + f1 - 1) * M * 2;
            const T *H2 = H + (size_t)(c * (F - 1) + f2 - 1) * M * 2;
            for (int k = tid; k < M; k += BT)
                L.A[k] = (k < N) ? c2<T>{L.fld[(c * F + f1) * N + k], L.fld[(c * F + f2) * N + k]} : c2<T>{T(0), T(0)};
            __syncthreads();
            fft_lds<T>(L.A, M, lg, tw, false);
            for (int k = tid; k < M; k += BT) {
                int km = (M - k) & (M - 1);
                c2<T> a = L.A[k], bb = L.A[km];
                // U1 = (a + conj b)/2, U2 = (a - conj b)/(2i)
                c2<T> u1 = {T(0.5) * (a.x + bb.x), T(0.5) * (a.y - bb.y)};
                c2<T> dd = {a.x - bb.x, a.y + bb.y};
                c2<T> u2 = {T(0.5) * dd.y, T(-0.5) * dd.x};
                c2<T> t1 = cmul(u1, c2<T>{H1[2 * k], H1[2 * k + 1]});
                c2<T> t2 = cmul(u2, c2<T>{H2[2 * k], H2[2 * k + 1]});
                c2<T> acc = L.P[k];
                acc.x += s1 * t1.x + s2 * t2.x;
                acc.y += s1 * t1.y + s2 * t2.y;
                L.P[k] = acc;
            }
            __syncthreads();
        }
        fft_lds<T>(L.P, M, lg, tw, true);
        for (int i = tid; i < N; i += BT) {
            T y = L.P[i].x;
            if (fold && i + N < M) y += L.P[i + N].x;
            double s = double(L.fld[(c * F) * N + i]) + double(y * invM);
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
                v[0] += r * r;
            } else {
                double gold = S.gx[t], rold = S.res[t];
                S.dg[t] = -gold + gnew;
                S.gx[t] = gnew;
                double r = gnew - x;
                S.res[t] = r;
                double d = -rold + r;
                S.df[t] = d;
                v[0] += r * r;
                v[1] += d * d;
                v[2] += d * r;
            }
        }
        __syncthreads();
    }
    wg_sum<3>(v, L.red);
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

// Householder downdate of the Anderson R (same arithmetic as the host version).
__device__ void dev_householder_downdate(double *R, int m, double *Qx)
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

template <typename T, int NC>
__device__ double evalG_dispatch(const BSys &S, const Lds<T> &L, const Lds<float> &L32, bool fp32, int mode, double b,
                                 double (&o)[3])
{
    if (fp32) evalG<float, NC>(S, L32, mode, b, o);
    else evalG<T, NC>(S, L, mode, b, o);
    return o[0];
}

template <int NC>
__global__ __launch_bounds__(BT) void k_batched(const BSys *__restrict__ systems, BOpts o, BResult *results)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ AState st;
    const BSys &S = systems[blockIdx.x];
    const int tid = threadIdx.x;
    const int N = S.N, M = S.M, n = N * NC;
    constexpr int F = 7;
    // LDS views for both arithmetic types over the same bytes
    Lds<double> L;
    L.A = reinterpret_cast<c2<double> *>(smem);
    L.P = L.A + M;
    L.fld = reinterpret_cast<double *>(L.P + M);
    L.red = L.fld + (size_t)F * N * NC;
    Lds<float> L32;
    L32.A = reinterpret_cast<c2<float> *>(smem);
    L32.P = L32.A + M;
    L32.fld = reinterpret_cast<float *>(L32.P + M);
    L32.red = L.red;                          // reduction scratch stays FP64, after the FP64 layout

    const int m = o.m;
    const double beta = o.beta;
    const int maxit = o.maxit, picard_maxit = o.maxit;
    const double rtol = o.tol, atol = o.tol, picard_beta = beta;
    const double picard_rtol = o.anderson_start, picard_atol = o.anderson_start;
    bool fp32 = (o.mixed != 0);
    int nevals = 0, nevals32 = 0;
    double fo[3];

    // sol = log(rho)
    for (int i = tid; i < n; i += BT) S.sol[i] = log(S.rho[i]);
    __syncthreads();

#define EVAL(mode, b) do { evalG_dispatch<double, NC>(S, L, L32, fp32, (mode), (b), fo); nevals++; if (fp32) nevals32++; } while (0)
#define SET_SOL_FROM_GX() do { for (int i_ = tid; i_ < n; i_ += BT) S.sol[i_] = S.gx[i_]; __syncthreads(); } while (0)

    int k = 0;
    EVAL(1, beta);
    double resnorm = sqrt(fo[0]);
    const double up_bd = 1.e4 * resnorm;
    const double tol = rtol * resnorm + atol;
    const double picard_tol = picard_rtol * resnorm + picard_atol;
    bool toosoon = (resnorm <= tol);

    if (!toosoon && picard_maxit > 0) {
        SET_SOL_FROM_GX();
        for (int it = 0; it < picard_maxit - 1; it++) {
            EVAL(1, picard_beta);
            resnorm = sqrt(fo[0]);
            if (resnorm <= picard_tol || resnorm >= up_bd) break;
            SET_SOL_FROM_GX();
        }
        toosoon = (resnorm <= tol);
    }

    // Anderson phase; a precision switch restarts it (ka counts Anderson steps
    // since the last (re)start, k counts all of them for maxit).
    double dfn2 = 0.0, dfres = 0.0;
    int ka = 0;
    double best = resnorm;
    int since_best = 0;
    auto restart = [&]() {
        if (tid == 0) {
            for (int j = 0; j < CDFT_MAXM * CDFT_MAXM; j++) st.R[j] = 0.0;
            for (int j = 0; j < CDFT_MAXM; j++) st.ring[j] = j;
        }
        __syncthreads();
        SET_SOL_FROM_GX();
        k += 1;
        ka = 1;
        EVAL(2, beta);
        resnorm = sqrt(fo[0]);
        dfn2 = fo[1];
        dfres = fo[2];
        best = resnorm;
        since_best = 0;
    };
    if (!toosoon) restart();

    while (k < maxit + picard_maxit && !toosoon && resnorm < up_bd) {
        if (fp32) {
            if (resnorm < 0.5 * best) { best = resnorm; since_best = 0; }
            else since_best++;
            if (resnorm <= o.switch_factor * tol || since_best >= o.stall_window) {
                // FP32 noise floor reached: continue with the FP64 map
                fp32 = false;
                EVAL(1, beta);
                resnorm = sqrt(fo[0]);
                if (resnorm <= tol) break;
                restart();
                continue;
            }
        } else if (!(resnorm > tol)) {
            break;
        }
        if (m == 0) {
            SET_SOL_FROM_GX();
        } else {
            int kk = ka + 1;
            int kq = ka - 1;
            int slot, qcol;
            int down = 0;
            if (tid == 0) {
                if (m > 1 && kk > m + 1) {
                    int first = st.ring[0];
                    for (int j = 0; j < m - 1; j++) st.ring[j] = st.ring[j + 1];
                    st.ring[m - 1] = first;
                }
            }
            __syncthreads();
            slot = (m == 1) ? st.ring[0] : (kk > m + 1 ? st.ring[m - 1] : st.ring[kk - 2]);
            // BuildDG!: copy dg into its ring slot
            for (int i = tid; i < n; i += BT) S.DG[(size_t)slot * n + i] = S.dg[i];
            double nrm;
            if (kq == 0) {
                nrm = sqrt(dfn2);
                if (tid == 0) {
                    st.R[0] = nrm;
                    st.tres[0] = dfres / nrm;
                }
                qcol = 0;
            } else {
                down = (kq > m - 1);
                if (down && tid == 0) dev_householder_downdate(st.R, m, st.Qx);
                __syncthreads();
                const int kqq = kq < m - 1 ? kq : m - 1;
                // pass 1: optional downdate of Q (pointwise, in place), h1 = Q^T df, t = Q^T res
                double v1[2 * CDFT_MAXM];
                for (int j = 0; j < 2 * CDFT_MAXM; j++) v1[j] = 0.0;
                for (int i = tid; i < n; i += BT) {
                    double qv[CDFT_MAXM];
                    if (down) {
                        double qo[CDFT_MAXM];
                        for (int l = 0; l < m; l++) qo[l] = S.Q[(size_t)l * n + i];
                        for (int j = 0; j < m - 1; j++) {
                            double a = 0.0;
                            for (int l = 0; l < m; l++) a += qo[l] * st.Qx[j * m + l];
                            qv[j] = a;
                            S.Q[(size_t)j * n + i] = a;
                        }
                        S.Q[(size_t)(m - 1) * n + i] = 0.0;
                    } else {
                        for (int j = 0; j < kqq; j++) qv[j] = S.Q[(size_t)j * n + i];
                    }
                    double dv = S.df[i], rv = S.res[i];
                    for (int j = 0; j < kqq; j++) {
                        v1[j] += qv[j] * dv;
                        v1[CDFT_MAXM + j] += qv[j] * rv;
                    }
                }
                wg_sum<2 * CDFT_MAXM>(v1, L.red);
                // pass 2: df -= Q h1; h2 = Q^T df
                double v2[CDFT_MAXM];
                for (int j = 0; j < CDFT_MAXM; j++) v2[j] = 0.0;
                for (int i = tid; i < n; i += BT) {
                    double a = 0.0, qv[CDFT_MAXM];
                    for (int j = 0; j < kqq; j++) {
                        qv[j] = S.Q[(size_t)j * n + i];
                        a += qv[j] * v1[j];
                    }
                    double d = S.df[i] - a;
                    S.df[i] = d;
                    for (int j = 0; j < kqq; j++) v2[j] += qv[j] * d;
                }
                wg_sum<CDFT_MAXM>(v2, L.red);
                // pass 3: df -= Q h2; |df|^2, df.res
                double v3[2] = {0.0, 0.0};
                for (int i = tid; i < n; i += BT) {
                    double a = 0.0;
                    for (int j = 0; j < kqq; j++) a += S.Q[(size_t)j * n + i] * v2[j];
                    double d = S.df[i] - a;
                    S.df[i] = d;
                    v3[0] += d * d;
                    v3[1] += d * S.res[i];
                }
                wg_sum<2>(v3, L.red);
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
                for (int j = mk - 1; j >= 0; j--) {            // theta = RA \ tres
                    double acc = st.tres[j];
                    for (int l = j + 1; l < mk; l++) acc -= st.R[l * m + j] * st.theta[l];
                    st.theta[j] = acc / st.R[j * m + j];
                }
            }
            __syncthreads();
            // new Q column, sol = gx - DG theta
            for (int i = tid; i < n; i += BT) {
                S.Q[(size_t)qcol * n + i] = S.df[i] / nrm;
                double a = 0.0;
                for (int j = 0; j < mk; j++) a += S.DG[(size_t)st.ring[j] * n + i] * st.theta[j];
                S.sol[i] = S.gx[i] - a;
            }
            __syncthreads();
        }
        k += 1;
        ka += 1;
        EVAL(2, beta);
        resnorm = sqrt(fo[0]);
        dfn2 = fo[1];
        dfres = fo[2];
    }
#undef EVAL
#undef SET_SOL_FROM_GX

    // a mixed solve that ran out of iterations in FP32 still reports FP64 residuals
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
