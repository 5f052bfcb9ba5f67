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
// owns CIRC_TW consecutive outputs; its 256 threads are 16 tap groups of 16 lanes,
// group g summing the g-th sixteenth of the k range in chunks of CIRC_KC taps, lane l
// the four outputs 4l .. 4l + 3 from a sliding window held in registers: per tap one
// broadcast read of h and one LDS read of src for four multiply-adds.  Each chunk's
// window of src is staged in LDS decimated by four (element j at [j & 3][j >> 2]), so
// the sixteen lanes of a group read consecutive words.  The groups' partial sums are
// added in a fixed order, so the result is reproducible.
// The sums run in T, loads convert from TS and the result is stored as TD.  COMP
// (the FP32 map): compensated summation, blocks of eight products summed plainly
// and the block sums added with Kahan's correction, the groups' sums and
// corrections combined in TD; the sum's rounding error then no longer grows with
// the number of taps (a plain FP32 sum over N taps loses up to ~N eps of the
// largest partial sum), at 1.5 extra operations per tap.
constexpr int CIRC_TW = 64;
constexpr int CIRC_KC = 64;
constexpr int CIRC_G = 16;
constexpr int CIRC_LQ = (CIRC_TW + CIRC_KC + 8) / 4;     // window row length (4 rows)

template <typename T>
__device__ __forceinline__ void kahan_add(T &s, T &c, T x)
{
    const T y = x - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
}

// s_r (+ compensation c_r with COMP) += sum_{k < kcn} hs[k] W(j0 + r - k), r < 4, taps
// in order, W decimated by four with rows of Lq; every index read is >= j0 - kcn.
template <typename T, bool COMP>
__device__ __forceinline__ void circ_chunk4(const T *W, int Lq, const T *hs, int kcn, int j0, T (&sr)[4], T (&cr)[4])
{
    auto U = [&](int j) __attribute__((always_inline)) { return W[(j & 3) * Lq + (j >> 2)]; };
    T w0 = U(j0), w1 = U(j0 + 1), w2 = U(j0 + 2), w3 = U(j0 + 3);
    // W(j0 - 1 - s - 4q) for s < 4 sits at p_s[-q]: stepping k by four keeps j & 3
    const T *p0 = W + ((j0 - 1) & 3) * Lq + ((j0 - 1) >> 2);
    const T *p1 = W + ((j0 - 2) & 3) * Lq + ((j0 - 2) >> 2);
    const T *p2 = W + ((j0 - 3) & 3) * Lq + ((j0 - 3) >> 2);
    const T *p3 = W + ((j0 - 4) & 3) * Lq + ((j0 - 4) >> 2);
    T b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    if constexpr (!COMP) {
        b0 = sr[0]; b1 = sr[1]; b2 = sr[2]; b3 = sr[3];
    }
    auto step4 = [&](int k, int q) __attribute__((always_inline)) {
        const T h0 = hs[k], h1 = hs[k + 1], h2 = hs[k + 2], h3 = hs[k + 3];
        b0 += h0 * w0; b1 += h0 * w1; b2 += h0 * w2; b3 += h0 * w3;
        const T x = p0[-q];
        b0 += h1 * x; b1 += h1 * w0; b2 += h1 * w1; b3 += h1 * w2;
        const T y = p1[-q];
        b0 += h2 * y; b1 += h2 * x; b2 += h2 * w0; b3 += h2 * w1;
        const T z = p2[-q];
        b0 += h3 * z; b1 += h3 * y; b2 += h3 * x; b3 += h3 * w0;
        const T u = p3[-q];
        w3 = x; w2 = y; w1 = z; w0 = u;
    };
    auto fold = [&]() __attribute__((always_inline)) {
        kahan_add(sr[0], cr[0], b0); kahan_add(sr[1], cr[1], b1);
        kahan_add(sr[2], cr[2], b2); kahan_add(sr[3], cr[3], b3);
        b0 = b1 = b2 = b3 = 0;
    };
    int k = 0;
    if constexpr (COMP) {
        for (; k + 8 <= kcn; k += 8) {
            step4(k, k >> 2);
            step4(k + 4, (k >> 2) + 1);
            fold();
        }
        if (k + 4 <= kcn) {
            step4(k, k >> 2);
            k += 4;
        }
    } else {
        for (; k + 4 <= kcn; k += 4) step4(k, k >> 2);
    }
    for (; k < kcn; k++) {
        const T hv = hs[k];
        b0 += hv * w0; b1 += hv * w1; b2 += hv * w2; b3 += hv * w3;
        w3 = w2; w2 = w1; w1 = w0;
        w0 = U(j0 - k - 1);
    }
    if constexpr (COMP) fold();
    else {
        sr[0] = b0; sr[1] = b1; sr[2] = b2; sr[3] = b3;
    }
}

template <typename T, typename TS = T, typename TD = T, bool COMP = false>
__global__ __launch_bounds__(256) void k_circ1d(const TS *__restrict__ src, long s_cstride, long s_fstride,
                                                const T *__restrict__ h, TD *__restrict__ dst, long d_cstride,
                                                long d_fstride, int d_fofs, int N, int F, T scale, int vec_sign,
                                                const int *gate)
{
    if (gate && *gate) return;
    constexpr int TW = CIRC_TW, KC = CIRC_KC, G = CIRC_G, LQ = CIRC_LQ;
    __shared__ T ws[G][4 * LQ];
    __shared__ T hs[G][KC];
    __shared__ T part[G][TW];
    __shared__ T partc[COMP ? G : 1][TW];
    const int y = blockIdx.y;
    const int c = y / (F - 1), f = 1 + y % (F - 1);
    const TS *s = src + c * s_cstride + f * s_fstride;
    const T *hk = h + (size_t)(c * F + f) * N;
    TD *d = dst + c * d_cstride + (f - d_fofs) * d_fstride;
    T sc = scale;
    if (vec_sign && f == 3) sc = -scale;   // 1D vector slot is f = 3
    const int i0 = blockIdx.x * TW;
    const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int Kg = (N + G - 1) / G;                     // taps per group
    const int kb = g * Kg, ke = min(N, kb + Kg);
    T sr[4] = {0, 0, 0, 0}, cr[4] = {0, 0, 0, 0};
    for (int kc0 = 0; kc0 < Kg; kc0 += KC) {            // the same trip count in every group
        const int kc = kb + kc0;
        const int kcn = max(0, min(KC, ke - kc));
        __syncthreads();
        if (kcn > 0) {
            for (int t = l; t < kcn; t += 16) hs[g][t] = hk[kc + t];
            // window position j holds src[i0 - kc - kcn - 3 + j]: output 4l + r, tap
            // kc + k reads position 4l + r + kcn + 3 - k, from 4l + 3 up to 66 + kcn
            const int wbase = i0 - kc - kcn - 3;
            for (int t = l; t < 4 * LQ; t += 16) {
                int gi = (wbase + t) % N;
                if (gi < 0) gi += N;
                ws[g][(t & 3) * LQ + (t >> 2)] = T(s[gi]);
            }
        }
        __syncthreads();
        if (kcn > 0) circ_chunk4<T, COMP>(ws[g], LQ, hs[g], kcn, 4 * l + kcn + 3, sr, cr);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        part[g][4 * l + r] = sr[r];
        if (COMP) partc[g][4 * l + r] = cr[r];
    }
    __syncthreads();
    const int o = threadIdx.x, i = i0 + o;
    if (o < TW && i < N) {
        if (COMP) {
            TD hi = 0, lo = 0;
#pragma unroll
            for (int q = 0; q < G; q++) {
                hi += TD(part[q][o]);
                lo += TD(partc[q][o]);
            }
            d[i] = TD(sc) * (hi - lo);
        } else {
            T acc = 0;
#pragma unroll
            for (int q = 0; q < G; q++) acc += part[q][o];
            d[i] = TD(sc * acc);
        }
    }
}

// Dense operators of the radial (spherical / cylindrical) grids:
// dst_(c,f)[i] = scale sum_j Op_(c,f)[i][j] src_(c,f)[j], blockIdx.y = c*(F-1) + f-1,
// Op stored column-major per (c, f) (op[(c F + f) N N + j N + i]); vop (optional) is
// the vector slot's (f = 3) operator [c N N] instead.  A workgroup owns 64 outputs,
// its four waves split the j range (src chunks in LDS, the operator read along i,
// consecutive per lane) and their sums are added in a fixed order.
__global__ __launch_bounds__(256) void k_dense1d(const double *__restrict__ src, long s_cstride, long s_fstride,
                                                 const double *__restrict__ op, const double *__restrict__ vop,
                                                 double *__restrict__ dst, long d_cstride, long d_fstride, int d_fofs,
                                                 int N, int F, double scale, const int *gate)
{
    if (gate && *gate) return;
    constexpr int TW = 64, NG = 4, KC = 256;
    __shared__ double xs[NG][KC];
    __shared__ double part[NG][TW];
    const int y = blockIdx.y;
    const int c = y / (F - 1), f = 1 + y % (F - 1);
    const double *s = src + c * s_cstride + f * s_fstride;
    const double *A = (vop && f == 3) ? vop + (size_t)c * N * N : op + (size_t)(c * F + f) * N * N;
    double *d = dst + c * d_cstride + (f - d_fofs) * d_fstride;
    const int g = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int i = blockIdx.x * TW + l;
    const int Kg = (N + NG - 1) / NG;
    const int kb = g * Kg, ke = min(N, kb + Kg);
    double acc = 0.0;
    for (int kc0 = 0; kc0 < Kg; kc0 += KC) {
        const int kc = kb + kc0;
        const int kcn = max(0, min(KC, ke - kc));
        __syncthreads();
        for (int t = l; t < kcn; t += TW) xs[g][t] = s[kc + t];
        __syncthreads();
        if (i < N)
            for (int kk = 0; kk < kcn; kk++) acc += A[(size_t)(kc + kk) * N + i] * xs[g][kk];
    }
    part[g][l] = acc;
    __syncthreads();
    if (g == 0 && i < N) d[i] = scale * ((part[0][l] + part[1][l]) + (part[2][l] + part[3][l]));
}

// Set-up of the radial operators: out[o N N + j N + i] = sum_m B[m N + i] w[o N + m] A[m N + j]
// (B column-major, A row-major), the convolution Q^-1 diag(w_o) Q of multiplier o.
__global__ void k_radial_op(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ w,
                            int N, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, o = blockIdx.z;
    if (i >= N) return;
    const double *wo = w + (size_t)o * N;
    double acc = 0.0;
    for (int m = 0; m < N; m++) acc += B[(size_t)m * N + i] * wo[m] * A[(size_t)m * N + j];
    out[(size_t)o * N * N + (size_t)j * N + i] = acc;
}

// Phi and dPhi/dn at every grid point.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi(const double *__restrict__ rho, const double *__restrict__ n,
                                             double *__restrict__ dphi, double *__restrict__ fval, int N,
                                             double NA, PhiConst P, const int *gate)
{
    if (gate && *gate) return;
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
// lambda-slot gradient is 0 where FP32 does not resolve lambda.  dphi0 (1D): the
// rho slot is also written in FP64 into the FP64 gradient array.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi32(const double *__restrict__ rho, const float *__restrict__ n,
                                               float *__restrict__ dphi, int N, double NA, PhiConst P,
                                               const int *gate, double *__restrict__ dphi0)
{
    if (gate && *gate) return;
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
    if (dphi0)                                  // the rho slot in FP64, where the exponent assembly reads it
#pragma unroll
        for (int c = 0; c < NC; c++) dphi0[(size_t)(c * F) * N + i] = g[c][0];
}

// Hessian-vector product of Phi at every point (no reference counterpart: the
// second derivative behind the fixed-point adjoint).  Slot 0 moves along N_A v,
// the other slots along their weighted densities ndot = W v (in ndio on entry);
// on exit ndio holds H_Phi(n) (N_A v, ndot) for every slot, the forward-mode
// derivative of phi_grad taken with dual numbers.
template <int NC, int ND, int AS = 0>
__global__ __launch_bounds__(256) void k_phi_hvp(const double *__restrict__ rho, const double *__restrict__ v,
                                                 const double *__restrict__ n, double *__restrict__ ndio, int N,
                                                 double NA, PhiConst P, const int *gate)
{
    if (gate && *gate) return;
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
// Reductions: each block writes one partial per quantity; k_aa_fin sums the
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

// Block partials of lambda_k . (G+ - G-) for k < nobj: G = clamp(cp_c - dF/drho - V,
// -100, 100), the undamped map (get_new_profile!, converge.jl:108-153) of two systems
// at one profile, from their dF_res/drho (dp, dm), fields and chemical potentials.
// part[k nblk + block]; grid-stride over the n = N nc points.
__global__ __launch_bounds__(256) void k_adj_param(const double *__restrict__ dp, const double *__restrict__ vp,
                                                   const double *__restrict__ cpp, const double *__restrict__ dm,
                                                   const double *__restrict__ vm, const double *__restrict__ cpm,
                                                   const double *__restrict__ lam, int N, int nc, int nobj,
                                                   double *part, int nblk)
{
    const long n = (long)N * nc, stride = (long)gridDim.x * blockDim.x;
    for (int k = 0; k < nobj; k++) {
        double v[1] = {0.0};
        for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
            const int c = (int)(t / N);
            double sp = dp[t], sm = dm[t];
            if (vp) sp += vp[t];
            if (vm) sm += vm[t];
            const double gp = clamp_lng(cpp[c] - sp), gm = clamp_lng(cpm[c] - sm);
            v[0] += lam[(size_t)k * n + t] * (gp - gm);
        }
        block_partials<1>(v, part + (size_t)k * nblk, nblk);
        __syncthreads();                       // the partials' LDS slots are reused for k + 1
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

__global__ void k_to_f32(const double *x, float *y, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

__global__ void k_exp(const double *x, double *rho, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rho[i] = exp(x[i]);
}

// Exponent assembly (get_new_profile!, converge.jl:108-153) fused with the
// Anderson book-keeping that follows each map evaluation.  Grid-stride: a thread
// sums its points' partials in order, so a capped grid keeps the partials few.
//  mode 0: out = dF_res/drho (no field, no clamp)                functional only
//  mode 1: gx = betafix(lnG); res = gx - sol; partial |res|^2     first/Picard step
//  mode 2: aa_point! (anderson.jl:470-481); partials |res|^2, |df|^2, df.res
// With ctl set, mode and beta come from the solver state (and nothing runs once
// the solve has ended).
// With adj_g set, the map value is the adjoint's g - rho* (dphi + parts) instead of
// the clamped exponent (dphi, part then hold a Hessian-vector product).
__global__ __launch_bounds__(256) void k_update(const double *__restrict__ dphi, const double *__restrict__ part,
                                                const double *__restrict__ vext, int N, int nc, int F,
                                                const double *__restrict__ cp, double *__restrict__ out,
                                                const double *__restrict__ sol, double *__restrict__ gx,
                                                double *__restrict__ res, double *__restrict__ df,
                                                double *__restrict__ dg, double beta, int mode, double *red,
                                                int nblk, const double *__restrict__ adj_g,
                                                const double *__restrict__ adj_rho, const AACtl *ctl)
{
    if (ctl) {                                 // mode and damping from the device-side solver
        if (ctl->g_done) return;
        mode = ctl->umode;
        beta = ctl->beta_u;
    }
    double v[3] = {0.0, 0.0, 0.0};
    const long nt = (long)N * nc, stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        int c = (int)(t / N), i = (int)(t - (long)c * N);
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
                v[0] += r * r;
            } else {
                double gold = gx[t], rold = res[t];
                dg[t] = -gold + g;
                gx[t] = g;
                double r = g - x;
                res[t] = r;
                double d = -rold + r;
                df[t] = d;
                v[0] += r * r;
                v[1] += d * d;
                v[2] += d * r;
            }
        }
    }
    if (mode != 0) block_partials<3>(v, red, nblk);
}

// ----- Anderson on the device (aasol, src/utils/anderson.jl:301-583) --------
// The host enqueues bodies of one map evaluation each (eagerly or as a captured
// hipGraph); the single-thread control kernels below advance the solver state in
// AACtl between the N-wide kernels, which read their work from it and return at
// once when their gate is set, so bodies after the end of a solve do nothing.

// Householder QR of the m x (m-1) block R[:, 2:m] with LAPACK's conventions
// (dgeqrf/dlarfg, then the thin Q of dorg2r); R column-major, leading dim m.
// The Anderson downdate (downdate_aaqr!, src/utils/anderson.jl:552-583).
__host__ __device__ inline void householder_downdate(double *R, int m, double *Qx)
{
#pragma clang fp contract(off)
    // every loop unrolled to CDFT_MAXM with the live entries selected, so the arrays
    // stay in registers; the live operations and their order are those of the plain
    // loops over j < m - 1, i < m
    constexpr int MX = CDFT_MAXM;
    const int nr = m, ncol = m - 1;
    double A[MX][MX], tau[MX], v[MX], Q[MX][MX];
#pragma unroll
    for (int j = 0; j < MX; j++)
#pragma unroll
        for (int i = 0; i < MX; i++) {
            const int t = (j + 1) * m + i;
            A[j][i] = (j < ncol && i < nr) ? R[t < MX * MX ? t : 0] : 0.0;
        }
#pragma unroll
    for (int j = 0; j < MX; j++) {
        tau[j] = 0.0;
        if (j >= ncol) continue;
        const double alpha = A[j][j];
        double xn = 0.0;
#pragma unroll
        for (int i = j + 1; i < MX; i++)
            if (i < nr) xn += A[j][i] * A[j][i];
        xn = sqrt(xn);
        if (xn == 0.0) continue;
        const double beta = -copysign(hypot(alpha, xn), alpha);
        tau[j] = (beta - alpha) / beta;
        const double sc = 1.0 / (alpha - beta);
#pragma unroll
        for (int i = j + 1; i < MX; i++)
            if (i < nr) A[j][i] *= sc;
        A[j][j] = beta;
#pragma unroll
        for (int jj = j + 1; jj < MX; jj++) {
            if (jj >= ncol) continue;
            double w = A[jj][j];
#pragma unroll
            for (int i = j + 1; i < MX; i++)
                if (i < nr) w += A[j][i] * A[jj][i];
            w *= tau[j];
            A[jj][j] -= w;
#pragma unroll
            for (int i = j + 1; i < MX; i++)
                if (i < nr) A[jj][i] -= w * A[j][i];
        }
    }
#pragma unroll
    for (int j = 0; j < MX; j++)
#pragma unroll
        for (int i = 0; i < MX; i++) Q[j][i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int h = MX - 1; h >= 0; h--) {
        if (h >= ncol || tau[h] == 0.0) continue;
        v[h] = 1.0;
#pragma unroll
        for (int i = h + 1; i < MX; i++) v[i] = A[h][i];
#pragma unroll
        for (int j = 0; j < MX; j++) {
            if (j >= ncol) continue;
            double w = 0.0;
#pragma unroll
            for (int i = h; i < MX; i++)
                if (i < nr) w += v[i] * Q[j][i];
            w *= tau[h];
#pragma unroll
            for (int i = h; i < MX; i++)
                if (i < nr) Q[j][i] -= w * v[i];
        }
    }
#pragma unroll
    for (int j = 0; j < MX; j++)
#pragma unroll
        for (int i = 0; i < MX; i++)
            if (j < ncol && i < nr) Qx[j * nr + i] = Q[j][i];
    // R[1:m-1, 1:m-1] = Rd; R[:, m] = 0
#pragma unroll
    for (int j = 0; j < MX; j++)
#pragma unroll
        for (int i = 0; i < MX; i++)
            if (j < ncol && i < ncol) R[j * m + i] = (i <= j) ? A[j][i] : 0.0;
#pragma unroll
    for (int i = 0; i < MX; i++)
        if (i < m) R[(m - 1) * m + i] = 0.0;
}

__device__ inline void aa_hist(AACtl &c, double *hist, double v)
{
    if (c.nh < c.hist_cap) hist[c.nh] = v;
    c.nh++;
}

// AndersonOK's errcode (anderson.jl:727-737)
__device__ inline void aa_finish(AACtl &c)
{
    if (c.toosoon) { c.idid = 1; c.errcode = -1; }
    else if (c.r >= c.up_bd) { c.idid = 0; c.errcode = -2; }
    else if (c.r <= c.tol && !c.fp32) { c.idid = 1; c.errcode = 0; }
    else { c.idid = 0; c.errcode = 10; }
    c.phase = AA_DONE;
}

// The while condition of aasol's Anderson loop (anderson.jl:393).  In mixed
// precision the FP32 phase continues below tol until the switch to the FP64 map.
__device__ inline void aa_loop_check(AACtl &c)
{
    if (c.k < c.maxit + c.picard_maxit && (c.r > c.tol || c.fp32) && !c.toosoon && c.r < c.up_bd) c.phase = AA_LOOP;
    else aa_finish(c);
}

// After a map evaluation: the residual norm and the Picard/Anderson book-keeping
// of the evaluation site that ran.  A non-finite FP32 residual ends the FP32 phase:
// the point goes back to the last one with a finite FP32 value (xsafe).
__device__ inline void aa_post(AACtl &c, const double *fin, double *hist)
{
#pragma clang fp contract(off)
    const double r = sqrt(fin[0]);
    c.nevals++;
    if (c.evfp32) c.nevals32++;
    c.keep = 0;
    switch (c.phase) {
    case AA_FIRST:
        if (c.fp32 && !isfinite(r)) {          // the seed itself: FP64 from the start
            c.fp32 = 0;
            c.r = r;
            return;
        }
        c.r = r;
        c.keep = c.fp32;
        c.up_bd = 1.e4 * r;
        c.tol = c.rtol * r + c.atol;
        c.picard_tol = c.picard_rtol * r + c.picard_atol;
        aa_hist(c, hist, r);
        c.toosoon = (r <= c.tol);
        if (c.toosoon) { aa_finish(c); return; }
        if (c.picard_maxit > 1) { c.phase = AA_PICARD; c.it = 0; }
        else c.phase = AA_START;
        return;
    case AA_PICARD:
        if (c.fp32 && !isfinite(r)) {          // retry this Picard step with the FP64 map at xsafe
            aa_hist(c, hist, r);
            c.r = r;
            c.restore = 1;
            c.fp32 = 0;
            return;
        }
        c.r = r;
        c.it++;
        c.keep = c.fp32;
        aa_hist(c, hist, r);
        if (r <= c.picard_tol || r >= c.up_bd || c.it >= c.picard_maxit - 1) {
            c.toosoon = (r <= c.tol);
            if (c.toosoon) aa_finish(c);
            else c.phase = AA_START;
        }
        return;
    case AA_SWITCH:
        c.r = r;
        aa_hist(c, hist, r);
        if (r <= c.tol) aa_finish(c);
        else c.phase = AA_START;
        return;
    case AA_START:
    case AA_STEP: {
        c.k += 1;
        c.ka = (c.phase == AA_START) ? 1 : c.ka + 1;
        c.r = r;
        c.dfn2 = fin[1];
        c.dfres = fin[2];
        aa_hist(c, hist, r);
        const bool failed = c.fp32 && !isfinite(r);
        if (failed) {
            c.restore = 1;
            c.fp32_failed = 1;
            c.r = 0.0;                         // the switch at the top of the loop takes over
        }
        c.keep = c.fp32 && !failed;
        if (c.phase == AA_START) c.best = c.r;
        aa_loop_check(c);
        return;
    }
    default:
        return;
    }
}

__device__ inline void aa_mid(AACtl &c, const double *fin);

// Before a map evaluation: the work of the next body (which point, which map,
// which update), with BuildDG!'s ring and the QR downdate of aa_qr_update!.  With
// DEFER the downdate is left to the caller (a whole wave, wave_downdate); the
// return value says whether one is due.
template <bool DEFER = false>
__device__ inline bool aa_pre(AACtl &c)
{
#pragma clang fp contract(off)
    bool dd = false;
    c.g_proj = 1;
    c.pmode = 0;
    c.umode = 1;
    c.beta_u = c.beta;
    switch (c.phase) {
    case AA_FIRST:
        break;
    case AA_PICARD:
        c.pmode = c.restore ? 2 : 1;
        c.restore = 0;
        c.beta_u = c.picard_beta;
        break;
    case AA_START:
        c.pmode = 1;
        c.umode = 2;
        for (int j = 0; j < CDFT_MAXM * CDFT_MAXM; j++) c.R[j] = 0.0;
        for (int j = 0; j < CDFT_MAXM; j++) c.ring[j] = j;
        break;
    case AA_LOOP: {
        if (c.fp32) {
            if (c.r < 0.5 * c.best) { c.best = c.r; c.since_best = 0; }
            else c.since_best++;
            if (c.r <= 4.0 * c.tol || c.since_best >= 40 || c.fp32_failed) {
                c.fp32 = 0;                    // the FP64 map from here, at the (restored) point
                c.phase = AA_SWITCH;
                c.pmode = c.restore ? 2 : 0;
                c.restore = 0;
                break;
            }
        }
        c.phase = AA_STEP;
        c.umode = 2;
        const int m = c.m;
        if (m == 0) { c.pmode = 1; break; }
        const int kk = c.ka + 1;               // BuildDG!(DG, m, k+1, dg)
        if (m == 1) {
            c.slot = c.ring[0];
        } else if (kk > m + 1) {
            const int first = c.ring[0];
            for (int j = 0; j < m - 1; j++) c.ring[j] = c.ring[j + 1];
            c.ring[m - 1] = first;
            c.slot = c.ring[m - 1];
        } else {
            c.slot = c.ring[kk - 2];
        }
        const int kq = c.ka - 1;              // aa_qr_update!(Q, R, df, m, k-1, Qd)
        c.down = 0;
        if (kq == 0) {
            c.nrm = sqrt(c.dfn2);
            c.R[0] = c.nrm;
            c.qcol = 0;
            c.tres[0] = c.dfres / c.nrm;
            c.kqq = 0;
            c.pmode = 3;
            aa_mid(c, nullptr);
        } else {
            for (int j = 0; j < CDFT_MAXM * CDFT_MAXM; j++) c.qx.a[j] = 0.0;
            c.down = (kq > m - 1);
            if constexpr (DEFER) dd = c.down;
            else if (c.down) householder_downdate(c.R, m, c.qx.a);
            c.kqq = min(kq, m - 1);
            c.g_proj = 0;
        }
        c.pmode = 3;
        break;
    }
    default:                                   // AA_DONE: a restore left pending by the last step
        c.pmode = c.restore ? 2 : 0;
        c.restore = 0;
        c.keep = 0;
        break;
    }
    const int done = (c.phase == AA_DONE);
    c.g_done = done;
    if (done) c.g_proj = 1;
    c.g_f64 = done || c.fp32;
    c.g_f32 = done || !c.fp32;
    c.evfp32 = !done && c.fp32;
    return dd;
}

// Stop after a control step, with the next body's work already set by aa_pre:
// STOP_PROGRESS after every pause_every-th evaluation (converge!'s per-evaluation
// logger, converge.jl:256-267); STOP_MAP when the next body needs the FP64 map while
// the host's bodies carry the FP32 map only (the switch, or an FP32 failure).  Every
// gate closes, so the bodies still queued do nothing; the next body's gates wait in
// sg for k_aa_resume.  The arithmetic of the solve stays where it is without a stop,
// so a stop changes no bit of it.
enum { STOP_PROGRESS = 1, STOP_MAP = 2 };
__device__ inline void aa_stop(AACtl &c)
{
    int why = (c.pause_every > 0 && c.nevals % c.pause_every == 0) ? STOP_PROGRESS : 0;
    if (c.graph_map == 1 && !c.g_done && !c.g_f64) why |= STOP_MAP;
    if (!why) return;
    c.sg[0] = c.g_done;
    c.sg[1] = c.g_proj;
    c.sg[2] = c.g_f64;
    c.sg[3] = c.g_f32;
    c.paused = why;
    c.g_done = 1;
    c.g_proj = 1;
    c.g_f64 = 1;
    c.g_f32 = 1;
}

// After the projections (or in aa_pre when there were none): the new R column
// and Q^T res, then the triangular solve theta = R \ (Q^T res).
__device__ inline void aa_mid(AACtl &c, const double *fin)
{
#pragma clang fp contract(off)
    if (c.pmode != 3 || c.g_done) return;
    const int m = c.m;
    if (!c.g_proj) {
        const int kqq = c.kqq;
        const double *h1 = fin, *tq = fin + CDFT_MAXM, *h2 = fin + CDFT_NQ;
        for (int j = 0; j < kqq; j++) c.R[kqq * m + j] = h1[j] + h2[j];
        c.nrm = sqrt(fin[2 * CDFT_NQ]);
        c.R[kqq * m + kqq] = c.nrm;
        c.qcol = kqq;
        for (int j = 0; j < kqq; j++) c.tres[j] = tq[j];
        c.tres[kqq] = fin[2 * CDFT_NQ + 1] / c.nrm;
        c.qsel ^= c.down;
    }
    const int mk = min(m, c.ka);
    c.mk = mk;
    for (int j = 0; j < CDFT_MAXM; j++) c.th.a[j] = 0.0;
    for (int j = mk - 1; j >= 0; j--) {
        double acc = c.tres[j];
        for (int l = j + 1; l < mk; l++) acc -= c.R[l * m + j] * c.th.a[l];
        c.th.a[j] = acc / c.R[j * m + j];
    }
    for (int j = 0; j < mk; j++) c.th.col[j] = c.ring[j];
}

// Sums of the partials of quantity q = blockIdx.x over nblk blocks in a fixed
// order (a strided sum per thread, a wave butterfly, the waves in order), then,
// in thread 0 of the last of the nq blocks to finish (a device-scope counter,
// reset for the next launch; the fences order each block's result before the
// count and the others' results before the step), the solver step that needs
// them: STEP 1 aa_mid after the projections, STEP 2 aa_post and aa_pre after a
// map evaluation.  Gated by g_proj (STEP 0, 1) or g_done (STEP 2).
template <int STEP>
__global__ __launch_bounds__(256) void k_aa_fin(const double *part, int nblk, double *out, AACtl *ctl, int site,
                                                double *fin, double *hist)
{
    if (STEP == 2 ? ctl->g_done : ctl->g_proj) return;
    __shared__ double sm[16];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[q * nblk + b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (int w = 0; w < nw; w++) t += sm[w];
    out[q] = t;
    if (STEP == 0) return;
    __threadfence();
    if (atomicAdd(&ctl->cnt[site], 1) != (int)gridDim.x - 1) return;
    atomicExch(&ctl->cnt[site], 0);
    __threadfence();
    if (STEP == 1) aa_mid(*ctl, fin);
    else {
        aa_post(*ctl, fin, hist);
        aa_pre(*ctl);
        aa_stop(*ctl);
    }
}

// The Anderson QR downdate by one whole wave (batched.hip).
template <int MX>
__device__ inline void wave_downdate(double *R, int m, double *Qx);

// Resumes a stopped solve: the next body's gates back; after STOP_MAP the host's
// bodies carry the FP64 map (the FP32 one is not used again).
__global__ void k_aa_resume(AACtl *ctl)
{
    if (threadIdx.x != 0 || !ctl->paused) return;
    if (ctl->paused & STOP_MAP) ctl->graph_map = 0;
    ctl->g_done = ctl->sg[0];
    ctl->g_proj = ctl->sg[1];
    ctl->g_f64 = ctl->sg[2];
    ctl->g_f32 = ctl->sg[3];
    ctl->paused = 0;
}

// pass 1: optional downdate Q' = Q Qx, then h1 = Q'^T df and t = Q'^T res.
__global__ __launch_bounds__(256) void k_aa_proj1(int n, const double *__restrict__ Q0, double *__restrict__ Q1,
                                                  const AACtl *ctl, const double *__restrict__ df,
                                                  const double *__restrict__ res, double *red, int nblk)
{
    if (ctl->g_proj) return;
    const int downdate = ctl->down, m = ctl->m, kq = ctl->kqq;
    const double *Q = ctl->qsel ? Q1 : Q0;
    double *Qout = ctl->qsel ? const_cast<double *>(Q0) : Q1;
    double v[2 * CDFT_MAXM];
    for (int q = 0; q < 2 * CDFT_MAXM; q++) v[q] = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double qv[CDFT_MAXM];
        if (downdate) {
            double qo[CDFT_MAXM];
            for (int l = 0; l < m; l++) qo[l] = Q[(size_t)l * n + i];
            for (int j = 0; j < m - 1; j++) {
                double a = 0.0;
                for (int l = 0; l < m; l++) a += qo[l] * ctl->qx.a[j * m + l];
                qv[j] = a;
                Qout[(size_t)j * n + i] = a;
            }
            Qout[(size_t)(m - 1) * n + i] = 0.0;
        } else {
            for (int j = 0; j < kq; j++) qv[j] = Q[(size_t)j * n + i];
        }
        double dv = df[i], rv = res[i];
        for (int j = 0; j < kq; j++) {
            v[j] += qv[j] * dv;
            v[CDFT_MAXM + j] += qv[j] * rv;
        }
    }
    block_partials<2 * CDFT_MAXM>(v, red, nblk);
}

// pass 2/3: df -= Q' h (h from the finalized reductions); then partials of
// Q'^T df (pass 2) or |df|^2 and df.res (pass 3).  Q' is the basis after the
// downdate of pass 1.
__global__ __launch_bounds__(256) void k_aa_proj2(int n, const double *__restrict__ Q0, const double *__restrict__ Q1,
                                                  const AACtl *ctl, double *__restrict__ df,
                                                  const double *__restrict__ res, const double *__restrict__ hcoef,
                                                  int last, double *red, int nblk)
{
    if (ctl->g_proj) return;
    const int kq = ctl->kqq;
    const double *Q = (ctl->qsel ^ ctl->down) ? Q1 : Q0;
    double v[CDFT_MAXM];
    for (int q = 0; q < CDFT_MAXM; q++) v[q] = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double qv[CDFT_MAXM];
        double a = 0.0;
        for (int j = 0; j < kq; j++) {
            qv[j] = Q[(size_t)j * n + i];
            a += qv[j] * hcoef[j];
        }
        double d = df[i] - a;
        df[i] = d;
        if (!last) {
            for (int j = 0; j < kq; j++) v[j] += qv[j] * d;
        } else {
            v[0] += d * d;
            v[1] += d * res[i];
        }
    }
    block_partials<CDFT_MAXM>(v, red, nblk);
}

// The next point of the body: sol = gx (Picard), sol = xsafe (back to the last
// point with a finite FP32 map value), or the Anderson point sol = gx - DG theta
// with the new Q column (vnew ./= hv[k+1], anderson.jl:538) and BuildDG!'s copy of
// dg into its ring slot; then rho = exp(sol), or rho = mask sol for the adjoint.
__global__ __launch_bounds__(256) void k_aa_point(int n, const AACtl *__restrict__ ctl, const double *__restrict__ df,
                                                  double *__restrict__ Q0, double *__restrict__ Q1,
                                                  double *__restrict__ sol, const double *__restrict__ gx,
                                                  const double *__restrict__ dg, double *__restrict__ DG,
                                                  double *__restrict__ xsafe, double *__restrict__ rho,
                                                  const double *__restrict__ mask)
{
    const int pm = ctl->pmode, keep = ctl->keep;
    if (ctl->paused || (pm == 0 && !keep)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keep) xsafe[i] = sol[i];
    if (pm == 0) return;
    double x;
    if (pm == 1) {
        x = gx[i];
    } else if (pm == 2) {
        x = xsafe[i];
    } else {
        DG[(size_t)ctl->slot * n + i] = dg[i];
        double *Qc = ctl->qsel ? Q1 : Q0;
        Qc[(size_t)ctl->qcol * n + i] = df[i] / ctl->nrm;
        const int mk = ctl->mk;
        double a = 0.0;
        for (int j = 0; j < mk; j++) a += DG[(size_t)ctl->th.col[j] * n + i] * ctl->th.a[j];
        x = gx[i] - a;
    }
    sol[i] = x;
    rho[i] = mask ? mask[i] * x : exp(x);
}

// ---------------------------------------------------------------------------
// Bodies of the plain path in a single workgroup, for small 1D Cartesian FP64
// systems (n = N nc <= SB_MAXN, nc <= 2): the Gram-Schmidt passes, the new point,
// the weighted densities, Phi, the transposed convolutions, the update and the
// control step of k_aa_proj1 / k_aa_proj2 / k_aa_point / k_circ1d / k_phi /
// k_update / k_aa_fin, in that order and with the same arithmetic per point, for
// up to nbody bodies in one launch instead of twelve launches per body (on grids
// this small each launch costs more than its work).  The control state is copied
// into LDS, advanced by the same aa_mid / aa_post / aa_pre in thread 0 (the QR
// downdate by wave 0, wave_downdate) and written back at the end.  Sums run in a
// fixed order (per point over taps in order; over the workgroup by waves in order),
// so a body is reproducible; they are not the multi-kernel path's orders, so the two
// paths agree to rounding.
constexpr int SB_T = 512;          // threads
constexpr int SB_MAXN = 512;       // n = N nc at most
constexpr int SB_F = 7;            // 1D field slots

struct SBArgs {
    AACtl *ctl;
    const double *h;               // circulant kernels h[(c F + f) N + k]
    const double *vext, *cp;
    double *sol, *gx, *res, *df, *dg, *Q0, *Q1, *DG, *xsafe, *rho, *hist;
    int N, nc;
    double NA;
    const float *h32;              // the mixed-precision kernels' FP32 copy of h (or null)
};

// Extended periodic source ue[j] = u[j mod N], j < Le, stored decimated by four:
// ue[j] at ud[(j & 3) Lq + (j >> 2)], Lq = Le / 4, so that the sliding windows of
// four consecutive outputs per lane read consecutive addresses across a wave.
__host__ __device__ inline int sb_le(int N) { return 4 * ((2 * N + 4 + 3) / 4); }

template <typename T>
__device__ __forceinline__ void sb_put(T *ud, int Lq, int N, int i, T v)
{
    for (int j = i; j < 4 * Lq; j += N) ud[(j & 3) * Lq + (j >> 2)] = v;
}

// out_r = sum_{k < N} h[k] ue[4b + r - k + N], r < 4, taps in order (a circular
// convolution at the outputs 4b .. 4b + 3).
__device__ __forceinline__ void sb_conv4(const double *__restrict__ ud, int Lq, const double *__restrict__ hk, int N,
                                         int b, double (&a)[4])
{
    auto U = [&](int j) __attribute__((always_inline)) { return ud[(j & 3) * Lq + (j >> 2)]; };
    const int j0 = 4 * b + N;
    double w0 = U(j0), w1 = U(j0 + 1), w2 = U(j0 + 2), w3 = U(j0 + 3);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    // ue[j0 - 1 - s - 4q] for s < 4 sits at p_s[-q]: stepping k by four keeps j & 3
    const double *p0 = ud + ((j0 - 1) & 3) * Lq + ((j0 - 1) >> 2);
    const double *p1 = ud + ((j0 - 2) & 3) * Lq + ((j0 - 2) >> 2);
    const double *p2 = ud + ((j0 - 3) & 3) * Lq + ((j0 - 3) >> 2);
    const double *p3 = ud + ((j0 - 4) & 3) * Lq + ((j0 - 4) >> 2);
    int k = 0;
    for (int q = 0; k + 4 <= N; k += 4, q++) {
        const double h0 = hk[k], h1 = hk[k + 1], h2 = hk[k + 2], h3 = hk[k + 3];
        a0 += h0 * w0; a1 += h0 * w1; a2 += h0 * w2; a3 += h0 * w3;
        const double x = p0[-q];                       // ue[j0 - k - 1]
        a0 += h1 * x; a1 += h1 * w0; a2 += h1 * w1; a3 += h1 * w2;
        const double y = p1[-q];                       // ue[j0 - k - 2]
        a0 += h2 * y; a1 += h2 * x; a2 += h2 * w0; a3 += h2 * w1;
        const double z = p2[-q];                       // ue[j0 - k - 3]
        a0 += h3 * z; a1 += h3 * y; a2 += h3 * x; a3 += h3 * w0;
        const double u = p3[-q];                       // ue[j0 - k - 4], j0 - k - 4 >= 4 b >= 0
        w3 = x; w2 = y; w1 = z; w0 = u;
    }
    for (; k < N; k++) {
        const double hv = hk[k];
        a0 += hv * w0; a1 += hv * w1; a2 += hv * w2; a3 += hv * w3;
        w3 = w2; w2 = w1; w1 = w0;
        w0 = U(j0 - k - 1);                  // >= 4 b >= 0
    }
    a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3;
}

// The FP32 map's form (k_circ1d's compensated sums): blocks of eight products summed
// plainly, each block sum added with Kahan's correction; the sum is a_r - cc_r.
__device__ __forceinline__ void sb_conv4c(const float *__restrict__ ud, int Lq, const float *__restrict__ hk, int N,
                                          int b, float (&a)[4], float (&cc)[4])
{
    auto U = [&](int j) __attribute__((always_inline)) { return ud[(j & 3) * Lq + (j >> 2)]; };
    const int j0 = 4 * b + N;
    float w0 = U(j0), w1 = U(j0 + 1), w2 = U(j0 + 2), w3 = U(j0 + 3);
    const float *p0 = ud + ((j0 - 1) & 3) * Lq + ((j0 - 1) >> 2);
    const float *p1 = ud + ((j0 - 2) & 3) * Lq + ((j0 - 2) >> 2);
    const float *p2 = ud + ((j0 - 3) & 3) * Lq + ((j0 - 3) >> 2);
    const float *p3 = ud + ((j0 - 4) & 3) * Lq + ((j0 - 4) >> 2);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
    float b0, b1, b2, b3;
    // four taps k .. k + 3 (window index q = k / 4) into the block sums
    auto step4 = [&](int k, int q) __attribute__((always_inline)) {
        const float h0 = hk[k], h1 = hk[k + 1], h2 = hk[k + 2], h3 = hk[k + 3];
        b0 += h0 * w0; b1 += h0 * w1; b2 += h0 * w2; b3 += h0 * w3;
        const float x = p0[-q];
        b0 += h1 * x; b1 += h1 * w0; b2 += h1 * w1; b3 += h1 * w2;
        const float y = p1[-q];
        b0 += h2 * y; b1 += h2 * x; b2 += h2 * w0; b3 += h2 * w1;
        const float z = p2[-q];
        b0 += h3 * z; b1 += h3 * y; b2 += h3 * x; b3 += h3 * w0;
        const float u = p3[-q];                        // >= 4 b >= 0
        w3 = x; w2 = y; w1 = z; w0 = u;
    };
    auto fold = [&]() __attribute__((always_inline)) {
        kahan_add(s0, c0, b0); kahan_add(s1, c1, b1); kahan_add(s2, c2, b2); kahan_add(s3, c3, b3);
    };
    int k = 0;
    for (int q = 0; k + 8 <= N; k += 8, q += 2) {
        b0 = b1 = b2 = b3 = 0.f;
        step4(k, q);
        step4(k + 4, q + 1);
        fold();
    }
    if (k < N) {
        b0 = b1 = b2 = b3 = 0.f;
        if (k + 4 <= N) {
            step4(k, k >> 2);
            k += 4;
        }
        for (; k < N; k++) {
            const float hv = hk[k];
            b0 += hv * w0; b1 += hv * w1; b2 += hv * w2; b3 += hv * w3;
            w3 = w2; w2 = w1; w1 = w0;
            w0 = U(j0 - k - 1);                        // >= 4 b >= 0
        }
        fold();
    }
    a[0] = s0; a[1] = s1; a[2] = s2; a[3] = s3;
    cc[0] = c0; cc[1] = c1; cc[2] = c2; cc[3] = c3;
}

// Workgroup sums of NQ values into out[0..NQ) (LDS), waves added in order.
template <int NQ>
__device__ __forceinline__ void sb_sum(double (&v)[NQ], double *wred, double *out)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) wred[wid * 16 + q] = v[q];
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        double t = 0.0;
        for (int w = 0; w < SB_T / 64; w++) t += wred[w * 16 + threadIdx.x];
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

template <int NC, int AS, bool MIXED>
__device__ __forceinline__ void sb_run(const SBArgs &A, const PhiConst &P, int nbody, double *sbm)
{
    constexpr int F = SB_F;
    __shared__ AACtl C;
    __shared__ double wred[(SB_T / 64) * 16];
    __shared__ double fin[2 * CDFT_NQ + 4];
    const int tid = threadIdx.x;
    const int N = A.N, n = N * NC;
    {
        const int *src = reinterpret_cast<const int *>(A.ctl);
        int *dst = reinterpret_cast<int *>(&C);
        for (int i = tid; i < (int)(sizeof(AACtl) / sizeof(int)); i += SB_T) dst[i] = src[i];
    }
    __syncthreads();
    if (C.g_done) return;                      // the solve has ended (or is paused)
    const int Le = sb_le(N), Lq = Le / 4;
    double *fld = sbm;                         // [NC F N]
    double *hk = fld + NC * F * N;             // [NC (F-1) N]
    double *ud = hk + NC * (F - 1) * N;        // [NC (F-1) Le]
    // the FP32 map (MIXED): its fields, sources and kernels in the FP64 map's regions
    // (as floats) and in hk32; the transposed sums, combined in FP64, in part64
    float *fld32 = reinterpret_cast<float *>(fld);
    float *ud32 = reinterpret_cast<float *>(ud);
    float *hk32 = reinterpret_cast<float *>(ud + NC * (F - 1) * Le);          // [NC (F-1) N]
    double *part64 = reinterpret_cast<double *>(hk32 + ((NC * (F - 1) * N + 1) & ~1));   // [NC (F-1) N]
    for (int t = tid; t < NC * (F - 1) * N; t += SB_T) {
        const int cf = t / N, k = t - cf * N, c = cf / (F - 1), f = 1 + cf % (F - 1);
        hk[t] = A.h[(size_t)(c * F + f) * N + k];
        if constexpr (MIXED) hk32[t] = A.h32[(size_t)(c * F + f) * N + k];
    }

    // up to nbody bodies, until the solve ends or pauses (C stays in LDS between them)
    for (int body = 0; body < nbody; body++) {
        if (body > 0 && C.g_done) break;
        // ---- Gram-Schmidt passes (k_aa_proj1, k_aa_proj2 x 2) and aa_mid
        // (loops over history columns unrolled to CDFT_MAXM with the live ones selected,
        // so every per-thread array stays in registers)
        if (!C.g_proj) {
            constexpr int MX = CDFT_MAXM;
            const int m = C.m, kq = C.kqq, down = C.down;
            const double *Q = C.qsel ? A.Q1 : A.Q0;
            double *Qout = C.qsel ? A.Q0 : A.Q1;
            double v1[2 * MX];
#pragma unroll
            for (int q = 0; q < 2 * MX; q++) v1[q] = 0.0;
            for (int i = tid; i < n; i += SB_T) {
                double qv[MX];
                if (down) {
                    double qo[MX];
#pragma unroll
                    for (int l = 0; l < MX; l++) qo[l] = (l < m) ? Q[(size_t)l * n + i] : 0.0;
#pragma unroll
                    for (int j = 0; j < MX - 1; j++) {
                        double a = 0.0;
#pragma unroll
                        for (int l = 0; l < MX; l++)
                            if (l < m) a += qo[l] * C.qx.a[j * m + l];
                        qv[j] = a;
                        if (j < m - 1) Qout[(size_t)j * n + i] = a;
                    }
                    qv[MX - 1] = 0.0;
                    Qout[(size_t)(m - 1) * n + i] = 0.0;
                } else {
#pragma unroll
                    for (int j = 0; j < MX; j++) qv[j] = (j < kq) ? Q[(size_t)j * n + i] : 0.0;
                }
                const double dv = A.df[i], rv = A.res[i];
#pragma unroll
                for (int j = 0; j < MX; j++)
                    if (j < kq) {
                        v1[j] += qv[j] * dv;
                        v1[MX + j] += qv[j] * rv;
                    }
            }
            sb_sum<2 * MX>(v1, wred, fin);
            const double *Q2 = (C.qsel ^ C.down) ? A.Q1 : A.Q0;
            for (int pass = 0; pass < 2; pass++) {
                const double *hc = pass == 0 ? fin : fin + CDFT_NQ;
                double v2[MX];
#pragma unroll
                for (int q = 0; q < MX; q++) v2[q] = 0.0;
                for (int i = tid; i < n; i += SB_T) {
                    double qv[MX];
                    double a = 0.0;
#pragma unroll
                    for (int j = 0; j < MX; j++) {
                        qv[j] = (j < kq) ? Q2[(size_t)j * n + i] : 0.0;
                        if (j < kq) a += qv[j] * hc[j];
                    }
                    const double d = A.df[i] - a;
                    A.df[i] = d;
                    if (pass == 0) {
#pragma unroll
                        for (int j = 0; j < MX; j++)
                            if (j < kq) v2[j] += qv[j] * d;
                    } else {
                        v2[0] += d * d;
                        v2[1] += d * A.res[i];
                    }
                }
                if (pass == 0) sb_sum<MX>(v2, wred, fin + CDFT_NQ);
                else {
                    double v3[2] = {v2[0], v2[1]};
                    sb_sum<2>(v3, wred, fin + 2 * CDFT_NQ);
                }
            }
            if (tid == 0) aa_mid(C, fin);
            __syncthreads();
        }

        // ---- the new point (k_aa_point) and rho; rho N_A into slot 0, rho as the
        // weighted densities' source (slot 0 of each component's source block)
        {
            const int pm = C.pmode, keep = C.keep;
            double *Qc = C.qsel ? A.Q1 : A.Q0;
            for (int t = tid; t < n; t += SB_T) {
                if (keep) A.xsafe[t] = A.sol[t];
                double r;
                if (pm != 0) {
                    double x;
                    if (pm == 1) {
                        x = A.gx[t];
                    } else if (pm == 2) {
                        x = A.xsafe[t];
                    } else {
                        A.DG[(size_t)C.slot * n + t] = A.dg[t];
                        Qc[(size_t)C.qcol * n + t] = A.df[t] / C.nrm;
                        double a = 0.0;
                        for (int j = 0; j < C.mk; j++) a += A.DG[(size_t)C.th.col[j] * n + t] * C.th.a[j];
                        x = A.gx[t] - a;
                    }
                    A.sol[t] = x;
                    r = exp(x);
                    A.rho[t] = r;
                } else {
                    r = A.rho[t];
                }
                const int c = t / N, i = t - c * N;
                if (MIXED && C.evfp32) {                   // k_circ1d<float, double>, k_phi32
                    fld32[(c * F) * N + i] = float(r * A.NA);
                    sb_put(ud32 + (size_t)c * (F - 1) * Le, Lq, N, i, float(r));
                } else {
                    fld[(c * F) * N + i] = r * A.NA;
                    sb_put(ud + (size_t)c * (F - 1) * Le, Lq, N, i, r);
                }
            }
        }
        __syncthreads();
        const bool f32 = MIXED && C.evfp32;            // this body's map (uniform)

        // ---- weighted densities n_f = N_A h_f (*) rho (k_circ1d)
        const int NB = (N + 3) / 4, NI = NC * (F - 1) * NB;
        for (int it = tid; it < NI; it += SB_T) {
            const int cf = it / NB, b = it - cf * NB, c = cf / (F - 1), f = 1 + cf % (F - 1);
            if (f32) {
                float a[4], cc[4];
                sb_conv4c(ud32 + (size_t)c * (F - 1) * Le, Lq, hk32 + (size_t)cf * N, N, b, a, cc);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (4 * b + r < N) fld32[(c * F + f) * N + 4 * b + r] = float(A.NA) * (a[r] - cc[r]);
            } else {
                double a[4];
                sb_conv4(ud + (size_t)c * (F - 1) * Le, Lq, hk + (size_t)cf * N, N, b, a);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (4 * b + r < N) fld[(c * F + f) * N + 4 * b + r] = A.NA * a[r];
            }
        }
        __syncthreads();

        // ---- Phi and dPhi/dn in place (k_phi)
        for (int i = tid; i < N; i += SB_T) {
            if constexpr (MIXED) {
                if (f32) {
                    float nl[NC][F], g[NC][F];
#pragma unroll
                    for (int c = 0; c < NC; c++)
#pragma unroll
                        for (int f = 0; f < F; f++) nl[c][f] = fld32[(c * F + f) * N + i];
                    phi_grad<float, double, NC, 1, true, AS, true>(P, nl, g);
#pragma unroll
                    for (int c = 0; c < NC; c++)
#pragma unroll
                        for (int f = 0; f < F; f++) fld32[(c * F + f) * N + i] = g[c][f];
                    continue;
                }
            }
            double nl[NC][F], g[NC][F];
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int f = 0; f < F; f++) nl[c][f] = fld[(c * F + f) * N + i];
            (void)phi_grad<double, double, NC, 1, true, AS>(P, nl, g);
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int f = 0; f < F; f++) fld[(c * F + f) * N + i] = g[c][f];
        }
        __syncthreads();
        for (int t = tid; t < NC * (F - 1) * N; t += SB_T) {
            const int cf = t / N, i = t - cf * N, c = cf / (F - 1), f = 1 + cf % (F - 1);
            if (f32) sb_put(ud32 + (size_t)cf * Le, Lq, N, i, fld32[(c * F + f) * N + i]);
            else sb_put(ud + (size_t)cf * Le, Lq, N, i, fld[(c * F + f) * N + i]);
        }
        __syncthreads();

        // ---- transposed convolutions p_f = +-h_f (*) dPhi/dn_f (k_circ1d, minus on the
        // vector slot) into the slots f >= 1
        for (int it = tid; it < NI; it += SB_T) {
            const int cf = it / NB, b = it - cf * NB, c = cf / (F - 1), f = 1 + cf % (F - 1);
            const double sg = (f == 3) ? -1.0 : 1.0;
            if (f32) {                                     // k_circ1d<float, float, double>
                float a[4], cc[4];
                sb_conv4c(ud32 + (size_t)cf * Le, Lq, hk32 + (size_t)cf * N, N, b, a, cc);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (4 * b + r < N) part64[(size_t)cf * N + 4 * b + r] = sg * (double(a[r]) - double(cc[r]));
            } else {
                double a[4];
                sb_conv4(ud + (size_t)cf * Le, Lq, hk + (size_t)cf * N, N, b, a);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (4 * b + r < N) fld[(c * F + f) * N + 4 * b + r] = sg * a[r];
            }
        }
        __syncthreads();

        // ---- the update (k_update with the solver's mode and damping)
        {
            const int mode = C.umode;
            const double beta = C.beta_u;
            double v[3] = {0.0, 0.0, 0.0};
            for (int t = tid; t < n; t += SB_T) {
                const int c = t / N, i = t - c * N;
                double s;
                if (f32) {                                 // dPhi/dn_rho in FP64, then the transposed sums
                    s = double(fld32[(c * F) * N + i]);
                    for (int f = 1; f < F; f++) s += part64[(size_t)(c * (F - 1) + f - 1) * N + i];
                } else {
                    s = fld[(c * F) * N + i];
                    for (int f = 1; f < F; f++) s += fld[(c * F + f) * N + i];
                }
                if (A.vext) s += A.vext[t];
                const double lg = clamp_lng(A.cp[c] - s);
                const double x = A.sol[t];
                const double g = (beta == 1.0) ? lg : (1.0 - beta) * x + beta * lg;
                if (mode == 1) {
                    A.gx[t] = g;
                    const double r = g - x;
                    A.res[t] = r;
                    v[0] += r * r;
                } else {
                    const double gold = A.gx[t], rold = A.res[t];
                    A.dg[t] = -gold + g;
                    A.gx[t] = g;
                    const double r = g - x;
                    A.res[t] = r;
                    const double d = -rold + r;
                    A.df[t] = d;
                    v[0] += r * r;
                    v[1] += d * d;
                    v[2] += d * r;
                }
            }
            sb_sum<3>(v, wred, fin);
        }
        __shared__ int dd;
        if (tid == 0) {
            aa_post(C, fin, A.hist);
            dd = aa_pre<true>(C);
        }
        __syncthreads();
        if (dd && tid < 64) wave_downdate<CDFT_MAXM>(C.R, C.m, C.qx.a);
        if (tid == 0) aa_stop(C);              // after the downdate, in wave 0's order
        __syncthreads();
    }
    {
        const int *src = reinterpret_cast<const int *>(&C);
        int *dst = reinterpret_cast<int *>(A.ctl);
        for (int i = tid; i < (int)(sizeof(AACtl) / sizeof(int)); i += SB_T) dst[i] = src[i];
    }
}

// Up to nbody bodies of one solve (cdft_converge's plain path).  MIXED: the system's
// map is FP32 until the solver's switch, then FP64 (the FP64 map alone otherwise).
template <int NC, int AS, bool MIXED>
__global__ __launch_bounds__(SB_T) void k_small_body(SBArgs A, PhiConst P, int nbody)
{
    extern __shared__ __align__(16) double sbm[];
    sb_run<NC, AS, MIXED>(A, P, nbody, sbm);
}

// One system of a batch (cdft_converge_batch of FP64 systems): its buffers and constants.
struct SBSys {
    SBArgs a;
    PhiConst pc;
};

// The batched form: workgroup b solves system b from its seed (x0 = log rho, the
// start of converge!, converge.jl:244) to the end, at most maxbodies bodies, and
// leaves rho = exp(x) and its control state (errcode, counts) in its own buffers;
// ticks[b] is its time in the workgroup (wall clock).
template <int NC, int AS, bool MIXED>
__global__ __launch_bounds__(SB_T) void k_small_batch(const SBSys *__restrict__ sys, int maxbodies,
                                                      unsigned long long *ticks)
{
    extern __shared__ __align__(16) double sbm[];
    __shared__ PhiConst P;
    const unsigned long long t0 = wall_clock64();
    const SBSys *S = sys + blockIdx.x;
    {
        const int *src = reinterpret_cast<const int *>(&S->pc);
        int *dst = reinterpret_cast<int *>(&P);
        for (int i = threadIdx.x; i < (int)(sizeof(PhiConst) / sizeof(int)); i += SB_T) dst[i] = src[i];
    }
    const SBArgs A = S->a;
    const int n = A.N * NC;
    for (int t = threadIdx.x; t < n; t += SB_T) {
        const double x = log(A.rho[t]);
        A.sol[t] = x;
        A.rho[t] = exp(x);
    }
    __syncthreads();
    sb_run<NC, AS, MIXED>(A, P, maxbodies, sbm);
    __syncthreads();
    // a restore left pending by the last step (back to the last point with a finite
    // FP32 map value, as aasol_run's closing k_aa_point), then rho = exp(x)
    const bool restore = MIXED && A.ctl->pmode == 2;
    for (int t = threadIdx.x; t < n; t += SB_T) {
        if (restore) A.sol[t] = A.xsafe[t];
        A.rho[t] = exp(A.sol[t]);
    }
    if (threadIdx.x == 0) ticks[blockIdx.x] = wall_clock64() - t0;
}

// Dynamic LDS bytes of k_small_body / k_small_batch (mixed: plus the FP32 kernels and
// the FP64 transposed sums).
__host__ __device__ inline size_t sb_lds_bytes(int N, int nc, bool mixed = false)
{
    const size_t m = (size_t)nc * (SB_F - 1) * N;
    return sizeof(double) * ((size_t)nc * SB_F * N + m + (size_t)nc * (SB_F - 1) * sb_le(N)) +
           (mixed ? sizeof(float) * ((m + 1) & ~(size_t)1) + sizeof(double) * m : 0);
}

}  // namespace cdft
