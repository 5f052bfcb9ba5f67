// batched.hip -- persistent batched Euler-Lagrange solver for 1D grids.
//
// One workgroup owns one system for its whole solve: map evaluations, Anderson
// history and the small QR problems never leave the device, and a system that
// terminates simply ends its workgroup (the next queued system takes the CU).
// The algorithm per system is exactly cdft_converge's (converge!/aasol,
// src/methods/converge.jl:214-254, src/utils/anderson.jl:301-443).
//
// Convolutions run in LDS as FFTs.  The reference's spectral operator
// Re IFFT_N(FFT_N(u) * Omega_f) equals the circular convolution h_f (*) u with the
// real kernel h_f = IFFT_N(Omega_herm_f), which is even (scalar weights) or odd (the
// vector weight).  With M = N (N a power of two) it is applied directly; otherwise
// the N-periodic kernel is unrolled onto a power-of-two ring M >= 2N - 1 at lags
// -(N-1) .. N-1 (h_M[t] = h[t], h_M[M - s] = h[N - s]), and the length-M circular
// convolution of the zero-padded profile then equals the N-periodic one on 0..N-1.
// Either way the kernel spectrum G_f on the ring is real (even kernels) or purely
// imaginary (odd), so it is stored as one real half-spectrum per field, k <= M/2,
// and kept in LDS.  Two real fields travel in one complex FFT (u1 + i u2), and the
// field pairs of a component (three; two for a chain-free system) share one set of
// barrier-separated FFT passes: four FFT rounds per component and evaluation.
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
    const double *G;           // real half-spectra [(c*(F-1) + f-1)*(M/2+1) + k], k <= M/2
    const double *tw64;        // W_M^t, t < M, (re, im) pairs
    double *rho;               // in: seed, out: converged profile  [c*N + i]
    double *sol, *gx, *df, *dg, *res, *Q, *DG;   // Anderson storage, n = N*nc each (Q, DG: CDFT_MAXM cols)
    double *hist;              // residual norm of every evaluation (AASol's history), or null
    int hist_cap;
    PhiConst pc;
};

struct BOpts {
    int maxit, m;
    double beta, tol, anderson_start;
    int mixed;                 // 1: FP32 map until the noise floor, then FP64
    double switch_factor;      // FP32 -> FP64 when resnorm <= switch_factor * tol
    int stall_window;          // ... or when it has not halved in this many iterations
    int g_lds;                 // 1: the spectra are copied into LDS at the start
};

struct BResult {
    int idid, errcode, iters, nevals, nevals32, nhist;
    double resnorm, tol;
    unsigned long long ticks;   // wall_clock64 ticks spent in the workgroup
};

constexpr int BT = 256;        // threads per workgroup of the 256-thread kernel (512 is the other)

// ---------------------------------------------------------------- helpers
template <typename T>
__device__ __attribute__((always_inline)) inline c2<T> cmul(c2<T> a, c2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}


#define CDFT_INLINE __device__ __attribute__((always_inline)) inline

// Radix-4 butterfly (unnormalised DFT of length 4 with sign sg) in registers.
template <typename T>
CDFT_INLINE void radix4(c2<T> (&v)[4], T sg)
{
    const c2<T> s02 = {v[0].x + v[2].x, v[0].y + v[2].y};
    const c2<T> d02 = {v[0].x - v[2].x, v[0].y - v[2].y};
    const c2<T> s13 = {v[1].x + v[3].x, v[1].y + v[3].y};
    const c2<T> d13 = {v[1].x - v[3].x, v[1].y - v[3].y};
    const c2<T> id13 = {-sg * d13.y, sg * d13.x};   // (sg i) * d13
    v[0] = {s02.x + s13.x, s02.y + s13.y};
    v[1] = {d02.x + id13.x, d02.y + id13.y};
    v[2] = {s02.x - s13.x, s02.y - s13.y};
    v[3] = {d02.x - id13.x, d02.y - id13.y};
}

// Stockham FFTs of NARR arrays of length M (power of two, 8 <= M <= 1024) that
// advance together through radix-4 passes and a final radix-2 pass when log2 M is
// odd.  Thread j handles butterfly j of every array (at most M/4 resp. M/2 <= 256
// butterflies per pass), so the twiddles W^{rk} of a pass are one table read and
// two products, shared by the arrays.  The first pass reads its inputs through
// load0(idx, vals[NARR]) and the last pass hands its outputs to
// storeL(idx, vals[NARR]); the passes between work in place on the LDS arrays
// W + q M.  Unnormalised; sign -1 forward, +1 inverse.  Ends with a barrier.
// PAD: array q starts at q (M + M/16) and element x sits at x + x/16, one spare
// slot per 16, which spreads the strided stores of the early passes over the banks.
template <typename T, int NARR, bool PAD = false, class Load0, class StoreL>
CDFT_INLINE void fft_fused(c2<T> *W, const c2<double> *tw, int M, int lg, bool inverse, Load0 &&load0,
                           StoreL &&storeL)
{
    const int MA = PAD ? M + (M >> 4) : M;
    auto wi = [&](int q, int x) __attribute__((always_inline)) { return q * MA + (PAD ? x + (x >> 4) : x); };
    int j = threadIdx.x;
    asm volatile("" : "+v"(j));           // opaque: keeps the index arithmetic inside the solver loop
    const T sg = inverse ? T(1) : T(-1);
    const int nr4 = lg >> 1;
    const bool odd = (lg & 1) != 0;
    int Ns = 1, lNs = 0;
    for (int pass = 0; pass < nr4; pass++) {
        const bool first = (pass == 0), last = (!odd && pass == nr4 - 1);
        const int nb = M >> 2;
        const int k = j & (Ns - 1);
        c2<T> v[NARR][4];
        if (j < nb) {
            if (first) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    c2<T> in[NARR];
                    load0(j + r * nb, in);
#pragma unroll
                    for (int q = 0; q < NARR; q++) v[q][r] = in[q];
                }
            } else {
#pragma unroll
                for (int q = 0; q < NARR; q++)
#pragma unroll
                    for (int r = 0; r < 4; r++) v[q][r] = W[wi(q, j + r * nb)];
                const c2<double> wd = tw[k << (lg - 2 - lNs)];
                const c2<T> w1 = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
                const c2<T> w2 = cmul(w1, w1), w3 = cmul(w1, w2);
#pragma unroll
                for (int q = 0; q < NARR; q++) {
                    v[q][1] = cmul(v[q][1], w1);
                    v[q][2] = cmul(v[q][2], w2);
                    v[q][3] = cmul(v[q][3], w3);
                }
            }
#pragma unroll
            for (int q = 0; q < NARR; q++) radix4<T>(v[q], sg);
        }
        if (!first && !last) __syncthreads();     // in place: every read before any write
        if (j < nb) {
            const int d = ((j >> lNs) << (lNs + 2)) + k;
            if (last) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    c2<T> out[NARR];
#pragma unroll
                    for (int q = 0; q < NARR; q++) out[q] = v[q][r];
                    storeL(d + r * Ns, out);
                }
            } else {
#pragma unroll
                for (int q = 0; q < NARR; q++)
#pragma unroll
                    for (int r = 0; r < 4; r++) W[wi(q, d + r * Ns)] = v[q][r];
            }
        }
        __syncthreads();
        Ns <<= 2;
        lNs += 2;
    }
    if (odd) {
        const int nb = M >> 1;
        const int k = j & (Ns - 1);
        c2<T> v[NARR][2];
        if (j < nb) {
            const c2<double> wd = tw[k << (lg - 1 - lNs)];
            const c2<T> w = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
#pragma unroll
            for (int q = 0; q < NARR; q++) {
                const c2<T> x0 = W[wi(q, j)], x1 = cmul(W[wi(q, j + nb)], w);
                v[q][0] = {x0.x + x1.x, x0.y + x1.y};
                v[q][1] = {x0.x - x1.x, x0.y - x1.y};
            }
            const int d = ((j >> lNs) << (lNs + 1)) + k;
#pragma unroll
            for (int r = 0; r < 2; r++) {
                c2<T> out[NARR];
#pragma unroll
                for (int q = 0; q < NARR; q++) out[q] = v[q][r];
                storeL(d + r * Ns, out);
            }
        }
        __syncthreads();
    }
}

// The same Stockham passes with the butterflies of the NARR arrays spread over the
// workgroup: item it = q nb + j is butterfly j of array q, so with 512 threads and
// two arrays of M = 1024 each thread runs one butterfly per pass (fft_fused gives
// thread j butterfly j of every array and leaves threads >= M/4 idle).  The first
// pass reads its inputs through load0(q, idx); the last pass hands its outputs to
// storeL(q, idx, v) or, with LAST_W, writes them into W like the passes between (a
// barrier then separates that pass's reads from its writes).  The arithmetic of each
// butterfly is fft_fused's.  Ends with a barrier.
template <typename T, int NARR, int BTT, bool LAST_W, class Load0, class StoreL>
CDFT_INLINE void fft_split(c2<T> *W, const c2<double> *tw, int M, int lg, bool inverse, Load0 &&load0,
                           StoreL &&storeL)
{
    constexpr int IT4 = (NARR * 256 + BTT - 1) / BTT;   // radix-4 items per thread (M/4 <= 256)
    constexpr int IT2 = (NARR * 512 + BTT - 1) / BTT;   // radix-2 items per thread (M/2 <= 512)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));         // opaque: keeps the index arithmetic inside the solver loop
    const T sg = inverse ? T(1) : T(-1);
    const int nr4 = lg >> 1;
    const bool odd = (lg & 1) != 0;
    int Ns = 1, lNs = 0;
    for (int pass = 0; pass < nr4; pass++) {
        const bool first = (pass == 0), last = (!odd && pass == nr4 - 1);
        const int lnb = lg - 2, nb = M >> 2, nitems = NARR << lnb;
        c2<T> v[IT4][4];
#pragma unroll
        for (int r = 0; r < IT4; r++) {
            const int it = tid + r * BTT;
            if (it < nitems) {
                const int q = it >> lnb, j = it & (nb - 1);
                if (first) {
#pragma unroll
                    for (int u = 0; u < 4; u++) v[r][u] = load0(q, j + u * nb);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; u++) v[r][u] = W[q * M + j + u * nb];
                    const c2<double> wd = tw[(j & (Ns - 1)) << (lg - 2 - lNs)];
                    const c2<T> w1 = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
                    const c2<T> w2 = cmul(w1, w1), w3 = cmul(w1, w2);
                    v[r][1] = cmul(v[r][1], w1);
                    v[r][2] = cmul(v[r][2], w2);
                    v[r][3] = cmul(v[r][3], w3);
                }
                radix4<T>(v[r], sg);
            }
        }
        if (!first && (!last || LAST_W)) __syncthreads();     // in place: every read before any write
#pragma unroll
        for (int r = 0; r < IT4; r++) {
            const int it = tid + r * BTT;
            if (it < nitems) {
                const int q = it >> lnb, j = it & (nb - 1);
                const int d = ((j >> lNs) << (lNs + 2)) + (j & (Ns - 1));
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (last && !LAST_W) storeL(q, d + u * Ns, v[r][u]);
                    else W[q * M + d + u * Ns] = v[r][u];
                }
            }
        }
        __syncthreads();
        Ns <<= 2;
        lNs += 2;
    }
    if (odd) {
        const int lnb = lg - 1, nb = M >> 1, nitems = NARR << lnb;
        c2<T> v[IT2][2];
#pragma unroll
        for (int r = 0; r < IT2; r++) {
            const int it = tid + r * BTT;
            if (it < nitems) {
                const int q = it >> lnb, j = it & (nb - 1);
                const c2<double> wd = tw[(j & (Ns - 1)) << (lg - 1 - lNs)];
                const c2<T> w = {T(wd.x), inverse ? T(-wd.y) : T(wd.y)};
                const c2<T> x0 = W[q * M + j], x1 = cmul(W[q * M + j + nb], w);
                v[r][0] = {x0.x + x1.x, x0.y + x1.y};
                v[r][1] = {x0.x - x1.x, x0.y - x1.y};
            }
        }
        if (LAST_W) __syncthreads();
#pragma unroll
        for (int r = 0; r < IT2; r++) {
            const int it = tid + r * BTT;
            if (it < nitems) {
                const int q = it >> lnb, j = it & (nb - 1);
                const int d = ((j >> lNs) << (lNs + 1)) + (j & (Ns - 1));
                if (LAST_W) {
                    W[q * M + d] = v[r][0];
                    W[q * M + d + Ns] = v[r][1];
                } else {
                    storeL(q, d, v[r][0]);
                    storeL(q, d + Ns, v[r][1]);
                }
            }
        }
        __syncthreads();
    }
}

// Cross-lane moves on FP64 values, as two 32-bit halves: DPP within a row of 16
// lanes (quad_perm xor 1 = 0xB1, xor 2 = 0x4E, row_half_mirror = 0x141,
// row_mirror = 0x140) and reads of one lane (uniform lane index).
template <int CTRL>
CDFT_INLINE double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

CDFT_INLINE double readlane_f64(double x, int lane)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 8 lanes of a half-row; pairwise commutative steps, so every lane
// of the group ends with the same bits.
CDFT_INLINE double sum8(double x)
{
    x += dpp_f64<0xB1>(x);
    x += dpp_f64<0x4E>(x);
    x += dpp_f64<0x141>(x);
    return x;
}

// Sum of NQ values over the workgroup; every thread receives the same totals
// (fixed summation order, so the result is reproducible and uniform): rows of 16
// lanes by DPP, the four rows of a wave by lane reads, the waves through LDS.
// Consecutive calls alternate between two halves of red (2 x 64 doubles), so one
// barrier per call suffices: a half is rewritten two calls later, after the
// intervening call's barrier, when every thread has read it.
template <int NQ, int BTT = BT>
CDFT_INLINE void wg_sum(double (&v)[NQ], double *red2, int &par)
{
    constexpr int NW = BTT / 64;                  // waves of the workgroup (4 or 8)
    static_assert(NQ * NW <= 64, "reduction slots");
    double *red = red2 + (par ? 64 : 0);
    par ^= 1;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double x = sum8(v[q]);
        x += dpp_f64<0x140>(x);
        v[q] = (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) red[q * NW + wid] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double *r = red + q * NW;
        double t;
        if constexpr (NW == 4) t = (r[0] + r[1]) + (r[2] + r[3]);
        else t = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        v[q] = readlane_f64(t, 0);            // the same bits in every lane: keep it scalar
    }
}

// One system's data, read once from its descriptor into registers.
template <int NC>
struct SysRegs {
    int N, M, lg, n;
    double NA;
    double cp[NC];
    const double *__restrict__ vext;
    double *__restrict__ rho;
};

// Per-point Anderson state of one system (anderson.jl:665-703: sol, gx, res, df,
// dg and the Q / Delta-G history columns).  REG: in registers of the thread that
// owns the points i = tid + b BTT (b < NPT, one component, at most AM history
// columns), so an Anderson step touches no HBM; otherwise in HBM ([i] and
// [col n + i]).  With one point per thread (512 threads) the Delta-G columns sit in
// LDS instead ([col kRegPoints + i], each thread touching only its own points), which
// keeps the kernel within the 256 registers two waves per SIMD allow.  Column
// indices that are not compile-time constants after unrolling go through the set/get
// helpers, which select among the registers.
constexpr int kRegNPT = 2;     // points per thread of the register-resident state (256 threads)
constexpr int kRegAM = 5;      // history columns it holds (Anderson m <= 5)
constexpr int kRegPoints = kRegNPT * BT;   // points of one system the register state holds

template <bool REG, int NPTT = kRegNPT>
struct AStore {
    static constexpr int NPT = REG ? NPTT : 1, AM = REG ? kRegAM : 1;
    static constexpr bool DGL = REG && NPTT == 1;          // Delta-G columns in LDS
    double sol_[NPT], gx_[NPT], res_[NPT], df_[NPT], dg_[NPT], Q_[NPT][AM], DG_[NPT][DGL ? 1 : AM];
    double *gsol, *ggx, *gres, *gdf, *gdg, *gQ, *gDG;
    double *dgl;                                            // LDS Delta-G columns (DGL)
    long n;

    CDFT_INLINE double &sol(int b, int i) { if constexpr (REG) return sol_[b]; else return gsol[i]; }
    CDFT_INLINE double &gx(int b, int i) { if constexpr (REG) return gx_[b]; else return ggx[i]; }
    CDFT_INLINE double &res(int b, int i) { if constexpr (REG) return res_[b]; else return gres[i]; }
    CDFT_INLINE double &df(int b, int i) { if constexpr (REG) return df_[b]; else return gdf[i]; }
    CDFT_INLINE double &dg(int b, int i) { if constexpr (REG) return dg_[b]; else return gdg[i]; }
    // column j known at compile time after unrolling (j < AM for REG)
    CDFT_INLINE double &Q(int b, int i, int j) { if constexpr (REG) return Q_[b][j]; else return gQ[j * n + i]; }
    // Q column j when live (registers: read regardless, the caller selects)
    CDFT_INLINE double Qif(int b, int i, int j, bool live)
    {
        if constexpr (REG) return Q_[b][j];
        else return live ? gQ[j * n + i] : 0.0;
    }
    // Q column j = v when c (registers: a select, so no branch splits the caller)
    CDFT_INLINE void setQif(int b, int i, int j, bool c, double v)
    {
        if constexpr (REG) Q_[b][j] = c ? v : Q_[b][j];
        else if (c) gQ[j * n + i] = v;
    }
    // Delta-G column s (compile-time s < AM) at this thread's point b
    CDFT_INLINE double DGcol(int b, int i, int s)
    {
        if constexpr (DGL) return dgl[s * kRegPoints + i];
        else return DG_[b][s];
    }
    CDFT_INLINE double getDG(int b, int i, int col)
    {
        if constexpr (DGL) {
            return dgl[col * kRegPoints + i];
        } else if constexpr (REG) {
            double v = 0.0;
#pragma unroll
            for (int s = 0; s < AM; s++) v = (s == col) ? DG_[b][s] : v;
            return v;
        } else {
            return gDG[col * n + i];
        }
    }
    CDFT_INLINE void setDG(int b, int i, int col, double v)
    {
        if constexpr (DGL) {
            dgl[col * kRegPoints + i] = v;
        } else if constexpr (REG) {
#pragma unroll
            for (int s = 0; s < AM; s++) DG_[b][s] = (s == col) ? v : DG_[b][s];   // no dynamic register index
        } else {
            gDG[col * n + i] = v;
        }
    }
    CDFT_INLINE void setQ(int b, int i, int col, double v)
    {
        if constexpr (REG) {
#pragma unroll
            for (int s = 0; s < AM; s++) Q_[b][s] = (s == col) ? v : Q_[b][s];
        } else {
            gQ[col * n + i] = v;
        }
    }
};

// a + x y when c, else a: a select rather than a branch, so that the loops over
// history columns stay one basic block (same rounding as the guarded a += x y).
CDFT_INLINE double fma_if(bool c, double a, double x, double y) { return c ? a + x * y : a; }

// Calls fn(b, i) for the points i < n of this thread: i = tid + b BTT.
template <bool REG, int BTT = BT, typename Fn>
CDFT_INLINE void for_points(int n, Fn &&fn)
{
    if constexpr (REG) {
#pragma unroll
        for (int b = 0; b < kRegPoints / BTT; b++) {
            const int i = (int)threadIdx.x + b * BTT;
            if (i < n) fn(b, i);
        }
    } else {
        for (int i = threadIdx.x; i < n; i += BTT) fn(0, i);
    }
}

// Field-slot pairs of the convolutions, slot 2q+1 (real part) and 2q+2 (imaginary
// part) of pair q when NP = 3.  A chain-free system (every segment number 1) needs
// neither the zeta (4) nor the lambda (5) slot: Phi does not read them and their
// gradients vanish, so with NP = 2 the pairs are (1, 2) and (3, 6).
template <int NP>
CDFT_INLINE int pair_slot(int q, int h)
{
    return NP == 3 ? 1 + 2 * q + h : (q == 0 ? 1 + h : (h ? 6 : 3));
}

// Spectral multipliers of the NP field pairs of one component at ring index k:
// Z_q = H_a + i H_b for the pair's slots (a, b), with H_f = G_f (even kernels) or
// i G_f (the vector weight, slot 3), from the half-spectra Gc[(f-1)*G1 + k'] with
// k' = min(k, M-k) and G_f(M-k) = +-G_f(k).  The transposed operator uses
// conj(Z_q): the vector weight's transpose is its negative, the scalar weights are
// symmetric.
template <typename T, int NP>
CDFT_INLINE void pair_multipliers(const double *Gc, int G1, int M, int k, c2<T> (&Z)[NP])
{
    const bool up = k > (M >> 1);
    const int kk = up ? M - k : k;
    const double g3 = up ? -Gc[2 * G1 + kk] : Gc[2 * G1 + kk];
    Z[0] = {T(Gc[kk]), T(Gc[G1 + kk])};
    if constexpr (NP == 3) {
        Z[1] = {T(0), T(g3 + Gc[3 * G1 + kk])};
        Z[2] = {T(Gc[4 * G1 + kk]), T(Gc[5 * G1 + kk])};
    } else {
        Z[1] = {T(0), T(g3 + Gc[5 * G1 + kk])};
    }
}

// Z_q alone (the split FFT handles one pair per butterfly).
template <typename T, int NP>
CDFT_INLINE c2<T> pair_multiplier(const double *Gc, int G1, int M, int k, int q)
{
    const bool up = k > (M >> 1);
    const int kk = up ? M - k : k;
    if (q == 0) return {T(Gc[kk]), T(Gc[G1 + kk])};
    if (NP == 3 && q == 2) return {T(Gc[4 * G1 + kk]), T(Gc[5 * G1 + kk])};
    const double g3 = up ? -Gc[2 * G1 + kk] : Gc[2 * G1 + kk];
    return {T(0), T(g3 + Gc[(NP == 3 ? 3 : 5) * G1 + kk])};
}

// Phi and dPhi/dn at the points tid, tid + nthr, tid + 2 nthr, ... (threads
// tid < nthr; nthr a multiple of 64), in place over the field slots.  A wave whose
// first lane has a second point evaluates two points per lane side by side (a lane
// without one repeats its first and discards it): with CHAIN = false the two
// inlined bodies are branch free and interleave.
template <typename T, int NC, bool CHAIN, int BTT = BT>
CDFT_INLINE void phi_points(const PhiConst &P, T *fld, int N, int nthr)
{
    constexpr int F = 7;
    const int tid = threadIdx.x;
    auto load = [&](T (&nl)[NC][F], int i) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int f = 0; f < F; f++) nl[c][f] = fld[(c * F + f) * N + i];
    };
    auto store = [&](const T (&g)[NC][F], int i) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int f = 0; f < F; f++) fld[(c * F + f) * N + i] = g[c][f];
    };
    if constexpr (BTT > BT) {                    // one point per lane: enough waves to overlap
        for (int i = tid; i < N; i += nthr) {
            T nl[NC][F], g[NC][F];
            load(nl, i);
            phi_grad<T, double, NC, 1, CHAIN>(P, nl, g);
            store(g, i);
        }
        return;
    }
    for (int i = tid; i < N; i += 2 * nthr) {
        T nl[NC][F], g[NC][F];
        load(nl, i);
        if (i - (tid & 63) + nthr < N) {             // uniform over the wave
            const int i2 = i + nthr;
            const bool two = i2 < N;
            T nl2[NC][F], g2[NC][F];
            load(nl2, two ? i2 : i);
            phi_grad<T, double, NC, 1, CHAIN>(P, nl, g);
            phi_grad<T, double, NC, 1, CHAIN>(P, nl2, g2);
            store(g, i);
            if (two) store(g2, i2);
        } else {
            phi_grad<T, double, NC, 1, CHAIN>(P, nl, g);
            store(g, i);
        }
    }
}

// One evaluation of the fixed-point map plus the Anderson book-keeping of mode
// (1: first/Picard step, 2: aa_point!).  out = |res|^2, |df|^2, df.res.  The last
// wave runs side() (whole wave) after its Phi points, before the barrier that ends
// that phase; with side_on it gets no Phi points when the other waves can take
// them all in as many rounds.
// LDS: tw (M, FP64), spectra G (6 NC (M/2+1) FP64, or global), three work arrays W
// and the accumulator P (M complex T each, at FP64 stride), fields fld (F*N*NC of T
// at FP64 stride), red (2 x 64 doubles, used alternately).
template <typename T, int NC, bool REG, int BTT, bool CF, class Side>
CDFT_INLINE void evalG(const SysRegs<NC> &S, AStore<REG, kRegPoints / BTT> &A, const PhiConst &P,
                       const c2<double> *tw, const double *G, c2<T> *W, c2<T> *Pc, T *fld, double *red, int &rpar,
                       int mode, double b, double (&out)[3], bool side_on, Side &&side)
{
    using AS = AStore<REG, kRegPoints / BTT>;
    constexpr int F = 7;
    const int tid = threadIdx.x;
    const int N = S.N, M = S.M, lg = S.lg, G1 = (S.M >> 1) + 1;
    const T invM = T(1) / T(M);
    const T sc = T(S.NA) * invM;
    bool chain = false;                 // CF: every segment number is 1 (known at launch)
#pragma unroll
    for (int c = 0; c < NC; c++) chain = chain || (!CF && P.m[c] != 1.0);
    using NP3 = std::integral_constant<int, 3>;
    using NP2 = std::integral_constant<int, 2>;

    for (int c = 0; c < NC; c++) {
        const double *Gc = G + (size_t)c * (F - 1) * G1;
        // rho = exp(sol): slot 0 of the fields and the zero-padded FFT input
#pragma unroll
        for (int j = 0; j < 1024 / BTT; j++) {
            const int k = tid + j * BTT;
            if (k < M) {
                if (k < N) {
                    const int jb = j < AS::NPT ? j : AS::NPT - 1;
                    const double r = exp(A.sol(jb, c * N + k));
                    fld[(c * F) * N + k] = T(S.NA * r);
                    Pc[k] = {T(r), T(0)};
                } else {
                    Pc[k] = {T(0), T(0)};
                }
            }
        }
        __syncthreads();
        // P = FFT(rho)
        fft_fused<T, 1>(
            W, tw, M, lg, false,
            [&](int idx, c2<T>(&v)[1]) __attribute__((always_inline)) { v[0] = Pc[idx]; },
            [&](int idx, const c2<T>(&v)[1]) __attribute__((always_inline)) { Pc[idx] = v[0]; });
        // n_a + i n_b = IFFT(P Z_q) for the slot pairs: the multiply feeds the first
        // pass, the last pass writes the weighted densities
        auto weighted = [&](auto np) __attribute__((always_inline)) {
            constexpr int NP = decltype(np)::value;
            if constexpr (BTT > BT) {
                fft_split<T, NP, BTT, false>(
                    W, tw, M, lg, true,
                    [&](int q, int idx) __attribute__((always_inline)) {
                        return cmul(Pc[idx], pair_multiplier<T, NP>(Gc, G1, M, idx, q));
                    },
                    [&](int q, int idx, const c2<T> &v) __attribute__((always_inline)) {
                        if (idx < N) {
                            fld[(c * F + pair_slot<NP>(q, 0)) * N + idx] = sc * v.x;
                            fld[(c * F + pair_slot<NP>(q, 1)) * N + idx] = sc * v.y;
                        }
                    });
                return;
            }
            fft_fused<T, NP>(
                W, tw, M, lg, true,
                [&](int idx, c2<T>(&v)[NP]) __attribute__((always_inline)) {
                    c2<T> Z[NP];
                    pair_multipliers<T, NP>(Gc, G1, M, idx, Z);
                    const c2<T> p = Pc[idx];
#pragma unroll
                    for (int q = 0; q < NP; q++) v[q] = cmul(p, Z[q]);
                },
                [&](int idx, const c2<T>(&v)[NP]) __attribute__((always_inline)) {
                    if (idx < N) {
#pragma unroll
                        for (int q = 0; q < NP; q++) {
                            fld[(c * F + pair_slot<NP>(q, 0)) * N + idx] = sc * v[q].x;
                            fld[(c * F + pair_slot<NP>(q, 1)) * N + idx] = sc * v[q].y;
                        }
                    }
                });
        };
        if (chain) weighted(NP3{});
        else weighted(NP2{});
    }
    // Phi and dPhi/dn, in place over the field slots
    {
        constexpr int PPL = BTT > BT ? 1 : 2;         // points per lane and round
        const int r_all = (N + PPL * BTT - 1) / (PPL * BTT);
        const int r_part = (N + PPL * (BTT - 64) - 1) / (PPL * (BTT - 64));
        const int nthr = (side_on && r_part <= r_all) ? BTT - 64 : BTT;
        if (tid < nthr) {
            if (chain) phi_points<T, NC, true, BTT>(P, fld, N, nthr);
            else phi_points<T, NC, false, BTT>(P, fld, N, nthr);
        }
    }
    if ((tid >> 6) == BTT / 64 - 1) side();          // the last wave has the fewest Phi points
    __syncthreads();
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int c = 0; c < NC; c++) {
        const double *Gc = G + (size_t)c * (F - 1) * G1;
        // transposed convolutions: Re IFFT(sum_q FFT(u_a + i u_b) conj(Z_q)) over the slot
        // pairs; the first pass reads the gradient pairs, the last one accumulates the
        // spectrum
        auto transposed = [&](auto np) __attribute__((always_inline)) {
            constexpr int NP = decltype(np)::value;
            if constexpr (BTT > BT) {
                // the pairs' spectra stay in W; their sum with conj(Z_q) is one more
                // pointwise pass
                fft_split<T, NP, BTT, true>(
                    W, tw, M, lg, false,
                    [&](int q, int idx) __attribute__((always_inline)) {
                        return (idx < N) ? c2<T>{fld[(c * F + pair_slot<NP>(q, 0)) * N + idx],
                                                 fld[(c * F + pair_slot<NP>(q, 1)) * N + idx]}
                                         : c2<T>{T(0), T(0)};
                    },
                    [&](int, int, const c2<T> &) __attribute__((always_inline)) {});
                for (int idx = tid; idx < M; idx += BTT) {
                    c2<T> Z[NP];
                    pair_multipliers<T, NP>(Gc, G1, M, idx, Z);
                    c2<T> acc = {T(0), T(0)};
#pragma unroll
                    for (int q = 0; q < NP; q++) {
                        const c2<T> v = W[q * M + idx];
                        acc.x += v.x * Z[q].x + v.y * Z[q].y;      // v * conj(Z)
                        acc.y += v.y * Z[q].x - v.x * Z[q].y;
                    }
                    Pc[idx] = acc;
                }
                __syncthreads();
                return;
            }
            fft_fused<T, NP>(
                W, tw, M, lg, false,
                [&](int idx, c2<T>(&v)[NP]) __attribute__((always_inline)) {
#pragma unroll
                    for (int q = 0; q < NP; q++)
                        v[q] = (idx < N) ? c2<T>{fld[(c * F + pair_slot<NP>(q, 0)) * N + idx],
                                                 fld[(c * F + pair_slot<NP>(q, 1)) * N + idx]}
                                         : c2<T>{T(0), T(0)};
                },
                [&](int idx, const c2<T>(&v)[NP]) __attribute__((always_inline)) {
                    c2<T> Z[NP];
                    pair_multipliers<T, NP>(Gc, G1, M, idx, Z);
                    c2<T> acc = {T(0), T(0)};
#pragma unroll
                    for (int q = 0; q < NP; q++) {
                        acc.x += v[q].x * Z[q].x + v[q].y * Z[q].y;      // v * conj(Z)
                        acc.y += v[q].y * Z[q].x - v[q].x * Z[q].y;
                    }
                    Pc[idx] = acc;
                });
        };
        if (chain) transposed(NP3{});
        else transposed(NP2{});
        fft_fused<T, 1>(
            W, tw, M, lg, true,
            [&](int idx, c2<T>(&v)[1]) __attribute__((always_inline)) { v[0] = Pc[idx]; },
            [&](int idx, const c2<T>(&v)[1]) __attribute__((always_inline)) { Pc[idx] = v[0]; });
#pragma unroll
        for (int j = 0; j < 1024 / BTT; j++) {
            const int i = tid + j * BTT;
            if (i >= N) continue;
            const int jb = j < AS::NPT ? j : AS::NPT - 1;
            double s = double(fld[(c * F) * N + i]) + double(Pc[i].x * invM);
            const int t = c * N + i;
            if (S.vext) s += S.vext[t];
            double lgv = S.cp[c] - s;
            lgv = clamp_lng(lgv);                          // clamp!(ln_Gx, -100, 100)
            const double x = A.sol(jb, t);
            const double gnew = (b == 1.0) ? lgv : (1.0 - b) * x + b * lgv;
            if (mode == 1) {
                A.gx(jb, t) = gnew;
                const double r = gnew - x;
                A.res(jb, t) = r;
                v0 += r * r;
            } else {
                const double gold = A.gx(jb, t), rold = A.res(jb, t);
                A.dg(jb, t) = -gold + gnew;
                A.gx(jb, t) = gnew;
                const double r = gnew - x;
                A.res(jb, t) = r;
                const double d = -rold + r;
                A.df(jb, t) = d;
                v0 += r * r;
                v1 += d * d;
                v2 += d * r;
            }
        }
        __syncthreads();
    }
    double v[3] = {v0, v1, v2};
    wg_sum<3, BTT>(v, red, rpar);
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

// Downdate of the Anderson QR when the oldest column leaves (anderson.jl, qr! of
// R[:, 2:m]): Householder QR of the m x (m-1) matrix R[:, 1:m] with LAPACK geqrf
// conventions (beta = -sign(alpha) |x|, tau = (beta - alpha) / beta, reflector
// scaled to a unit leading entry), its explicit Q factor Qx (m x (m-1), column-major
// ld m) and the new R (upper (m-1) x (m-1), last column zero), written over R in
// LDS.  Run by one whole wave, which holds the matrix one element per lane, row
// i = lane & 7 and column j = lane >> 3, and reduces over the 8 lanes of a column
// with DPP steps.
template <int MX = CDFT_MAXM>
CDFT_INLINE void wave_downdate(double *R, int m, double *Qx)
{
    const int lane = threadIdx.x & 63, i = lane & 7, j = lane >> 3;
    const int nr = m, ncol = m - 1;
    double a = (i < nr && j < ncol) ? R[(j + 1) * m + i] : 0.0;
    double tau[MX];
#pragma unroll
    for (int h = 0; h < MX; h++) {
        tau[h] = 0.0;
        if (h >= ncol) continue;
        const double x = __shfl(a, h * 8 + i, 64);                 // A[i][h]
        const double alpha = readlane_f64(a, h * 9);               // A[h][h]
        const double xn = sqrt(sum8((i > h && i < nr) ? x * x : 0.0));
        if (xn == 0.0) continue;                                    // H_h = I
        const double beta = -copysign(hypot(alpha, xn), alpha);
        const double th = (beta - alpha) / beta;
        tau[h] = th;
        const double sc = 1.0 / (alpha - beta);
        const double v = (i == h) ? 1.0 : ((i > h && i < nr) ? x * sc : 0.0);
        const double w = sum8(v * a) * th;                      // (v^T A[:, j]) tau
        if (j > h && j < ncol) a -= w * v;
        if (j == h && i == h) a = beta;
        if (j == h && i > h && i < nr) a = v;                       // reflector below the diagonal
    }
    // Qx = H_0 H_1 ... H_{ncol-1} I[:, 0:ncol]
    double q = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int h = MX - 1; h >= 0; h--) {
        if (h >= ncol || tau[h] == 0.0) continue;
        const double vh = __shfl(a, h * 8 + i, 64);
        const double v = (i == h) ? 1.0 : ((i > h && i < nr) ? vh : 0.0);
        const double w = sum8(v * q) * tau[h];
        q -= w * v;
    }
    // one wave: every lane read R above, before these writes
    if (i < nr && j < ncol) Qx[j * nr + i] = q;
    if (i < ncol && j < ncol) R[j * m + i] = (i <= j) ? a : 0.0;
    if (j == 0 && i < m) R[(m - 1) * m + i] = 0.0;
}

// theta = R[0:mk, 0:mk] \ tr by column-oriented back substitution.  The last column
// (mk - 1) and tr come in registers (uniform over the workgroup), the other columns
// from R in LDS; lane l of every wave ends holding theta_l (l < mk).
template <int MX = CDFT_MAXM>
CDFT_INLINE double dev_triangular_solve(const double *R, const double (&rc)[MX], const double (&tr)[MX],
                                        int m, int mk)
{
    const int lane = threadIdx.x & 63;
    double x = 0.0, last = 0.0;
#pragma unroll
    for (int l = 0; l < MX; l++) {
        x = (lane == l && l < mk) ? tr[l] : x;
        last = (lane == l) ? rc[l] : last;
    }
    double th = 0.0;
#pragma unroll
    for (int j = MX - 1; j >= 0; j--) {
        if (j >= mk) continue;
        const bool lastc = (j == mk - 1);
        const double rjj = lastc ? rc[j] : R[j * m + j];
        const double tj = readlane_f64(x, j) / rjj;
        if (lane == j) th = tj;
        if (lane < j) x -= (lastc ? last : R[j * m + lane]) * tj;
    }
    return th;
}

// Shared scalar state of one system's solve (LDS).
struct AState {
    double R[CDFT_MAXM * CDFT_MAXM];
    double Qx[CDFT_MAXM * CDFT_MAXM];
};

// Dynamic LDS bytes of one system (FP64 layout; the FP32 map uses the same slots):
// twiddles, three work arrays and the accumulator (M complex each), the spectra when
// they live in LDS, the fields, the reduction slots and, for the 512-thread kernel,
// the Delta-G history columns.
__host__ __device__ inline size_t batched_lds_layout(int N, int M, int nc, bool g_lds, bool dg_lds = false)
{
    return (size_t)5 * M * 16 + (g_lds ? (size_t)6 * nc * ((M >> 1) + 1) * 8 : 0) + (size_t)7 * N * nc * 8 + 128 * 8 +
           (dg_lds ? (size_t)kRegAM * kRegPoints * 8 : 0);
}

// HIST: the variant that records every evaluation's residual norm (a separate
// instantiation: the store's mere presence costs the recording-free kernel about 2 %)
template <int NC, bool REG, int BTT = BT, bool CF = false, bool HIST = false>
__global__ __launch_bounds__(BTT) void k_batched(const BSys *__restrict__ systems, BOpts o, BResult *results)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ AState st;
    __shared__ PhiConst pcs;
    constexpr int F = 7;
    constexpr int AM = REG ? kRegAM : CDFT_MAXM;     // history columns addressed
    constexpr int MX = AM;                             // largest Anderson depth of this kernel
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
    S.rho = D->rho;
    AStore<REG, kRegPoints / BTT> A;
    A.gsol = D->sol;
    A.ggx = D->gx;
    A.gdf = D->df;
    A.gdg = D->dg;
    A.gres = D->res;
    A.gQ = D->Q;
    A.gDG = D->DG;
    A.n = (long)S.N * NC;
    const double *__restrict__ twg = D->tw64;
    double *const hist = HIST ? D->hist : nullptr;
    const int hist_cap = HIST && hist ? D->hist_cap : 0;
    {
        const double *src = reinterpret_cast<const double *>(&D->pc);
        double *dst = reinterpret_cast<double *>(&pcs);
        for (int i = tid; i < (int)(sizeof(PhiConst) / sizeof(double)); i += BTT) dst[i] = src[i];
    }
    const int N = S.N, M = S.M, n = S.n;
    const int ng = (F - 1) * NC * ((M >> 1) + 1);
    c2<double> *tw = reinterpret_cast<c2<double> *>(smem);
    c2<double> *W64 = tw + M;
    c2<double> *P64 = W64 + 3 * M;
    double *Gl = reinterpret_cast<double *>(P64 + M);
    double *fld64 = Gl + (o.g_lds ? ng : 0);
    double *red = fld64 + (size_t)F * N * NC;
    A.dgl = red + 128;                                // Delta-G columns when they live in LDS
    c2<float> *W32 = reinterpret_cast<c2<float> *>(W64);
    c2<float> *P32 = reinterpret_cast<c2<float> *>(P64);
    float *fld32 = reinterpret_cast<float *>(fld64);
    for (int t = tid; t < M; t += BTT) tw[t] = {twg[2 * t], twg[2 * t + 1]};
    const double *G = D->G;
    if (o.g_lds) {
        for (int t = tid; t < ng; t += BTT) Gl[t] = G[t];
        G = Gl;
    }

    const int m = o.m;
    const double beta = o.beta;
    const int maxit = o.maxit, picard_maxit = o.maxit;
    const double rtol = o.tol, atol = o.tol, picard_beta = beta;
    const double picard_rtol = o.anderson_start, picard_atol = o.anderson_start;
    bool fp32 = (o.mixed != 0);
    int nevals = 0, nevals32 = 0;
    int rpar = 0;                                          // which half of red the next sum uses
    double fo[3];

    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.sol(b, i) = log(S.rho[i]); });   // x0 = log(rho)
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
            if (tid == 0)
                for (int j = 0; j < CDFT_MAXM * CDFT_MAXM; j++) st.R[j] = 0.0;
            for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.sol(b, i) = A.gx(b, i); });   // copy!(sol, gx)
            __syncthreads();
            k += 1;
            ka = 1;
            mode = 2;
        } else if (phase == PH_SWITCH) {
            fp32 = false;                                           // FP32 noise floor: FP64 map from here
        } else if (phase == PH_STEP) {
            if (m == 0) {
                for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.sol(b, i) = A.gx(b, i); });
                __syncthreads();
            } else {
                // aa_point! (anderson.jl) with the Delta-G history as a ring: the
                // BuildDG! shift rotates the ring once per step past the m-th, so ring
                // position j holds physical column (rot + j) mod m and the new column
                // goes to (ka - 1) mod m.  Q and R follow aa_qr_update!; when the oldest
                // column leaves, R and the rotation Qx were already downdated by the
                // last wave during the map evaluation.  The new R column and Q^T res
                // stay in registers (uniform), so the solve needs no barrier.
                const int kq = ka - 1;
                const int rot = ka > m ? (ka - m) % m : 0;         // ring offset, < m
                const int pos = ka > m ? m - 1 : ka - 1;            // ring position of the new column
                const int slot = rot + pos - (rot + pos >= m ? m : 0);
                const int mk = m < ka ? m : ka;
                const int kqq = kq < m - 1 ? kq : m - 1;            // the new Q / R column
                const bool down = (kq > m - 1);
                double nrm;
                double rc[MX], tr[MX];                              // new R column, Q^T res
#pragma unroll
                for (int j = 0; j < MX; j++) {
                    rc[j] = 0.0;
                    tr[j] = 0.0;
                }
                if (kq == 0) {                                      // aa_qr_update! first column
                    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.setDG(b, i, slot, A.dg(b, i)); });
                    nrm = sqrt(dfn2);
                    rc[0] = nrm;
                    tr[0] = dfres / nrm;
                } else {
                    // pass 1: store the Delta-G column; rotate Q (Q <- Q Qx) when the
                    // oldest column left; h1 = Q^T df, t = Q^T res
                    double v1[2 * (AM - 1)];
#pragma unroll
                    for (int j = 0; j < 2 * (AM - 1); j++) v1[j] = 0.0;
                    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) {
                        A.setDG(b, i, slot, A.dg(b, i));
                        double qv[AM];
                        if (down) {
                            double qo[AM];
#pragma unroll
                            for (int l = 0; l < AM; l++) qo[l] = (l < m) ? A.Q(b, i, l) : 0.0;
#pragma unroll
                            for (int j = 0; j < AM - 1; j++) {
                                double a = 0.0;
#pragma unroll
                                for (int l = 0; l < AM; l++) a = fma_if(l < m, a, qo[l], st.Qx[j * m + l]);
                                qv[j] = a;
                                A.setQif(b, i, j, j < m - 1, a);
                            }
                            A.setQ(b, i, m - 1, 0.0);
                        } else {
#pragma unroll
                            for (int j = 0; j < AM - 1; j++) qv[j] = (j < kqq) ? A.Qif(b, i, j, j < kqq) : 0.0;
                        }
                        const double dv = A.df(b, i), rv = A.res(b, i);
#pragma unroll
                        for (int j = 0; j < AM - 1; j++) {
                            v1[j] = fma_if(j < kqq, v1[j], qv[j], dv);
                            v1[(AM - 1) + j] = fma_if(j < kqq, v1[(AM - 1) + j], qv[j], rv);
                        }
                    });
                    wg_sum<2 * (AM - 1), BTT>(v1, red, rpar);
                    // pass 2: df -= Q h1; h2 = Q^T df
                    double v2[AM - 1];
#pragma unroll
                    for (int j = 0; j < AM - 1; j++) v2[j] = 0.0;
                    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) {
                        double qv[AM], a = 0.0;
#pragma unroll
                        for (int j = 0; j < AM - 1; j++) {
                            qv[j] = A.Qif(b, i, j, j < kqq);
                            a = fma_if(j < kqq, a, qv[j], v1[j]);
                        }
                        const double d = A.df(b, i) - a;
                        A.df(b, i) = d;
#pragma unroll
                        for (int j = 0; j < AM - 1; j++) v2[j] = fma_if(j < kqq, v2[j], qv[j], d);
                    });
                    wg_sum<AM - 1, BTT>(v2, red, rpar);
                    // pass 3: df -= Q h2; |df|^2, df.res
                    double v3[2] = {0.0, 0.0};
                    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) {
                        double a = 0.0;
#pragma unroll
                        for (int j = 0; j < AM - 1; j++) a = fma_if(j < kqq, a, A.Qif(b, i, j, j < kqq), v2[j]);
                        const double d = A.df(b, i) - a;
                        A.df(b, i) = d;
                        v3[0] += d * d;
                        v3[1] += d * A.res(b, i);
                    });
                    wg_sum<2, BTT>(v3, red, rpar);
                    nrm = sqrt(v3[0]);
#pragma unroll
                    for (int j = 0; j < AM - 1; j++)
                        if (j < kqq) {
                            rc[j] = v1[j] + v2[j];
                            tr[j] = v1[(AM - 1) + j];
                        }
#pragma unroll
                    for (int j = 0; j < AM; j++)
                        if (j == kqq) {
                            rc[j] = nrm;
                            tr[j] = v3[1] / nrm;
                        }
                }
                // R's new column for the next downdate; nothing reads it in this step
                if (tid < MX && tid <= kqq) {
                    double r = 0.0;
#pragma unroll
                    for (int j = 0; j < MX; j++) r = (j == tid) ? rc[j] : r;
                    st.R[kqq * m + tid] = r;
                }
                const double thl = dev_triangular_solve<MX>(st.R, rc, tr, m, mk);
                // coefficient of each physical Delta-G column: theta of its ring position
                double cf[AM], th[MX];
#pragma unroll
                for (int s2 = 0; s2 < AM; s2++) cf[s2] = 0.0;
#pragma unroll
                for (int j = 0; j < MX; j++) {
                    const double thj = readlane_f64(thl, j);
                    th[j] = thj;
                    const int col = (j < mk) ? rot + j - (rot + j >= m ? m : 0) : -1;
#pragma unroll
                    for (int s2 = 0; s2 < AM; s2++) cf[s2] = (s2 == col) ? thj : cf[s2];
                }
                // new Q column; sol = gx - DG theta
                for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) {
                    const double dv = A.df(b, i);
                    A.setQ(b, i, kqq, dv / nrm);
                    double a = 0.0;
                    if constexpr (REG) {
                        // physical columns 0 .. mk-1 hold the ring's mk live columns; the
                        // others were never written (leftover register or LDS bits, possibly
                        // Inf/NaN), so they are selected out rather than multiplied by 0
#pragma unroll
                        for (int s2 = 0; s2 < AM; s2++) a = fma_if(s2 < mk, a, A.DGcol(b, i, s2), cf[s2]);
                    } else {
#pragma unroll
                        for (int j = 0; j < MX; j++)
                            if (j < mk) a += A.getDG(b, i, rot + j - (rot + j >= m ? m : 0)) * th[j];
                    }
                    A.sol(b, i) = A.gx(b, i) - a;
                });
                __syncthreads();
            }
            k += 1;
            ka += 1;
            mode = 2;
        }

        // ---------------- the one map evaluation
        // the QR downdate the next Anderson step needs (it drops the oldest column
        // when ka > m) depends only on R, so the last wave does it meanwhile
        const bool dd = (phase == PH_STEP || phase == PH_RESTART) && m >= 1 && ka > m;
        auto side = [&]() __attribute__((always_inline)) {
            if (dd) wave_downdate<MX>(st.R, m, st.Qx);
        };
        if (fp32) {
            evalG<float, NC, REG, BTT, CF>(S, A, pcs, tw, G, W32, P32, fld32, red, rpar, mode, bb, fo, dd, side);
            nevals32++;
        } else {
            evalG<double, NC, REG, BTT, CF>(S, A, pcs, tw, G, W64, P64, fld64, red, rpar, mode, bb, fo, dd, side);
        }
        nevals++;
        resnorm = sqrt(fo[0]);
        if constexpr (HIST)
            if (tid == 0 && nevals <= hist_cap) hist[nevals - 1] = resnorm;   // every evaluation is recorded

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
                for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.sol(b, i) = A.gx(b, i); });
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
                for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { A.sol(b, i) = A.gx(b, i); });
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
    for_points<REG, BTT>(n, [&](int b, int i) __attribute__((always_inline)) { S.rho[i] = exp(A.sol(b, i)); });
    if (tid == 0) {
        BResult r;
        r.idid = idid;
        r.errcode = errcode;
        r.iters = k;
        r.nevals = nevals;
        r.nevals32 = nevals32;
        r.nhist = HIST ? nevals : 0;
        r.resnorm = resnorm;
        r.tol = tol;
        r.ticks = wall_clock64() - t_start;
        results[blockIdx.x] = r;
    }
}

// ---------------------------------------------------------------- set-up kernels
// Real half-spectra of the kernels on the FFT ring (k <= M/2), FP64, exact twiddles:
//   M == N:  H(k) = sum_{t<N} h[t] e^{-2 pi i t k / N}
//   M > N:   H(k) = h[0] + sum_{s=1}^{N-1} (h[s] e^{-2 pi i s k / M} + h[N-s] e^{+2 pi i s k / M})
// G = Re H for the even kernels and Im H for the odd one (slot 3); the other part
// vanishes by symmetry and is dropped.
__global__ void k_setup_spec(const double *h, int N, int M, int F, double *G)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int cf = blockIdx.y;                       // c*(F-1) + f-1
    const int G1 = (M >> 1) + 1;
    if (k >= G1) return;
    const int c = cf / (F - 1), f = 1 + cf % (F - 1);
    const double *hk = h + (size_t)(c * F + f) * N;
    double re = 0.0, im = 0.0;
    if (M == N) {
        for (int t = 0; t < N; t++) {
            const long r = ((long)t * k) % M;
            double sn, cs;
            sincospi(2.0 * (double)r / (double)M, &sn, &cs);
            re += hk[t] * cs;
            im -= hk[t] * sn;
        }
    } else {
        re = hk[0];
        for (int t = 1; t < N; t++) {
            const long r = ((long)t * k) % M;
            double sn, cs;
            sincospi(2.0 * (double)r / (double)M, &sn, &cs);
            re += (hk[t] + hk[N - t]) * cs;
            im -= (hk[t] - hk[N - t]) * sn;
        }
    }
    G[(size_t)cf * G1 + k] = (f == 3) ? im : re;
}

__global__ void k_setup_tw(int M, double *tw64)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    double sn, cs;
    sincospi(2.0 * (double)t / (double)M, &sn, &cs);
    tw64[2 * t] = cs;
    tw64[2 * t + 1] = -sn;
}

}  // namespace cdft
