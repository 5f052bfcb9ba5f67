// phi.h -- pointwise PC-SAFT Helmholtz density Phi(n) and its gradient dPhi/dn
// at one grid point, for the HIP kernels.
//
// Restates f_res = f_hs + f_hc + f_disp + f_assoc (PCSAFT.jl:110-117):
//   f_hs    src/models/DFT/FMT.jl:17-47 (White-Bear n3 dependence)
//   f_hc    src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:123-144
//   f_disp  src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:150-211
//   f_assoc src/models/DFT/assoc.jl:702-831 (Wertheim association)
// The gradient is hand-derived (the reference differentiates f_res with Enzyme,
// src/models/models.jl:355-411); it is what the "backward kernel" computes.
//
// Field slots per component (0-based): 0 rho, 1 S (int rho dz, d/2),
// 2 V (int rho z^2 dz, d/2), 3..2+ND vector (int rho z dz, d/2),
// 3+ND zeta slot (int rho z^2 dz, d), 4+ND lambda slot (int rho dz, d),
// 5+ND dispersion slot (int rho z^2 dz, psi d).
//
// Template parameters: T is the arithmetic type of the bulk of the work; TS is
// the type used for the n3-singular hard-sphere factors ln|1-n3|, 1/(1-n3)^k and
// their small-n3 cancellation (FP64 in both precision modes).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#ifndef CDFT_MAXC
#define CDFT_MAXC 8
#endif
#define CDFT_MAXSITE 4      // association sites on the device path
#define CDFT_MAXPAIR 8      // association bonding pairs on the device path

// Per-system constants in reduced units (lengths / L), set up on the host.
struct PhiConst {
    int nc;
    double T;
    double psi;                           // dispersion width factor (PCSAFT.jl:53,160)
    double m[CDFT_MAXC];
    double dd[CDFT_MAXC];                 // d_i / L
    double cm[CDFT_MAXC * CDFT_MAXC];     // m_i m_j (sigma_ij / L)^3
    double eT[CDFT_MAXC * CDFT_MAXC];     // eps_ij / T
    // Wertheim association (pack_assoc_params, assoc.jl:18-57); npair = 0: none
    int nsite, npair;
    int site_comp[CDFT_MAXSITE];          // component of each site
    double site_count[CDFT_MAXSITE];      // multiplicity of each site type
    int pair_a[CDFT_MAXPAIR], pair_b[CDFT_MAXPAIR];   // the two sites of each pair
    double pair_k[CDFT_MAXPAIR];          // (sigma_ab / L)^3 (exp(eps_AB / T) - 1) kappa_AB
    double pair_dij[CDFT_MAXPAIR];        // d_i d_j / (d_i + d_j) / L
    // phi_grad<..., F32IN = true>: per component, the lambda-slot magnitude below
    // which an FP32 convolution's value is rounding noise (1e-7 of its bulk value)
    double lsl_noise[CDFT_MAXC];
};

namespace cdft {

constexpr double kPi = 3.14159265358979323846;

// The a_qk / b_qk tables of PCSAFT_CORR1/2 (PCSAFT.jl:86-103, Gross & Sadowski
// 2001), in constant memory and read through an opaque pointer in i_lite, so the
// compiler re-reads them at each evaluation instead of holding 42 constants in
// registers across a whole solve.
__device__ __constant__ double kCorrDev[2][7][3] = {
    {{0.9105631445, -0.3084016918, -0.0906148351},
     {0.6361281449, 0.1860531159, 0.4527842806},
     {2.6861347891, -2.5030047259, 0.5962700728},
     {-26.547362491, 21.419793629, -1.7241829131},
     {97.759208784, -65.25588533, -4.1302112531},
     {-159.59154087, 83.318680481, 13.77663187},
     {91.297774084, -33.74692293, -8.6728470368}},
    {{0.7240946941, -0.5755498075, 0.0976883116},
     {2.2382791861, 0.6995095521, -0.2557574982},
     {-4.0025849485, 3.892567339, -9.155856153},
     {-21.003576815, -17.215471648, 20.642075974},
     {26.855641363, 192.67226447, -38.804430052},
     {206.55133841, -161.82646165, 93.626774077},
     {-355.60235612, -165.20769346, -29.666905585}}};

// I(mbar, eta) of PCSAFT.jl:199-211 and its partials in mbar and eta (table 0 or 1).
template <typename T>
__device__ __forceinline__ void i_lite(int which, T mb, T eta, T &I, T &Im, T &Ie)
{
    const double *tab = &kCorrDev[which][0][0];
    asm volatile("" : "+s"(tab));
    T m2 = (mb - T(1)) / mb;
    T m3 = m2 * (mb - T(2)) / mb;
    T inv = T(1) / mb;
    T dm2 = inv * inv;
    T dm3 = (T(3) * mb - T(4)) * inv * inv * inv;
    T v = 0, vm = 0, ve = 0, ep = 1, epm1 = 0;
#pragma unroll
    for (int q = 0; q < 7; q++) {
        T c0 = T(tab[3 * q]), c1 = T(tab[3 * q + 1]), c2 = T(tab[3 * q + 2]);
        T a = c0 + m2 * c1 + m3 * c2;
        T am = dm2 * c1 + dm3 * c2;
        v += a * ep;
        vm += am * ep;
        ve += T(q) * a * epm1;
        epm1 = ep;
        ep *= eta;
    }
    I = v;
    Im = vm;
    Ie = ve;
}

// x[idx] for a runtime index into a register array, as a chain of selects (keeps the
// array in registers).
template <typename T, int K>
__device__ __forceinline__ T pick(const T (&x)[K], int idx)
{
    T r = x[0];
#pragma unroll
    for (int k = 1; k < K; k++) r = (idx == k) ? x[k] : r;
    return r;
}

// Association term (f_assoc, assoc.jl:702-831) at one point; adds its gradient to g
// and returns Phi_assoc.  n2m, n3m, nv2m are the mixture densities of the hard-sphere
// term (the same sums, assoc.jl:713-733).  Per component: w_i = n0_i xi_i with
// n0_i = S_i / d_i, xi_i = 1 - |nv2_i|^2 / n2_i^2, n2_i = pi S_i d_i, nv2_i = -2 pi v_i.
// Per pair: Delta_p = g_hs(d_ab; n2m, n3m, xi_m) (sigma_ab^3 (exp(eps/T) - 1) kappa)
// (assoc.jl:159-168).  Site fractions solve X_s (1 + C_s) = 1 with
// C_s = sum_t K_st X_t, K_st = sum of Delta_p v_t over the pairs joining s and t,
// v_t = n_t w_c(t).  AS = 1: one pair of two distinct sites, in closed form
// (assoc.jl:759-764); AS = 2: up to CDFT_MAXSITE sites, X0 = 0.9, 5 NS blended
// substitutions, then 5 NS safeguarded Newton steps (assoc.jl:238-411).
// Gradient: at the solution dPhi/dw_i = sum_{s in i} n_s ln X_s and
// dPhi/dDelta_p = -v_a X_a v_b X_b (stationarity of the mass-action Q function; the
// reference gets the same derivative from an implicit-function rule, assoc.jl:833-916).
template <typename T, int NC, int ND, int AS>
__device__ __forceinline__ T assoc_grad(const PhiConst &P, const T (&nl)[NC][6 + ND], T (&g)[NC][6 + ND], T n2m,
                                        T n3m, const T (&nv2m)[ND])
{
    constexpr int NS = CDFT_MAXSITE;
    const T pi = T(kPi), twopi = T(2.0 * kPi);
    T w[NC], dwS[NC], dwv[NC][ND];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const T di = T(P.dd[i]), S = nl[i][1];
        const T n0 = S / di, n2 = pi * S * di, in2 = T(1) / (n2 * n2);
        T sq = 0;
#pragma unroll
        for (int j = 0; j < ND; j++) sq += (twopi * nl[i][3 + j]) * (twopi * nl[i][3 + j]);
        T r = sq * in2;
        if constexpr (std::is_same<T, float>::value) r = fmin(r, 1.0f);   // FP32 underflow in walls
        const T xi = T(1) - r;
        w[i] = n0 * xi;
        dwS[i] = xi / di + n0 * T(2) * r / n2 * pi * di;
#pragma unroll
        for (int j = 0; j < ND; j++) dwv[i][j] = -n0 * T(2) * twopi * nl[i][3 + j] * in2 * twopi;
    }
    T nv2sq = 0;
#pragma unroll
    for (int j = 0; j < ND; j++) nv2sq += nv2m[j] * nv2m[j];
    const T xim = T(1) - nv2sq / (n2m * n2m);
    const T a1 = T(1) / (T(1) - n3m);
    constexpr int NP = AS == 1 ? 1 : CDFT_MAXPAIR;
    T delta[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const T dd = T(P.pair_dij[p]);
        const T gh = a1 + T(0.5) * dd * xim * n2m * a1 * a1 + dd * dd * n2m * n2m * xim * a1 * a1 * a1 / T(18);
        delta[p] = (p < P.npair) ? gh * T(P.pair_k[p]) : T(0);
    }
    T v[NS], X[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) v[t] = (t < P.nsite) ? T(P.site_count[t]) * pick(w, P.site_comp[t]) : T(0);

    if constexpr (AS == 1) {
        const int a = P.pair_a[0], b = P.pair_b[0];
        const T va = pick(v, a), vb = pick(v, b);
        const T kia = va * delta[0], kjb = vb * delta[0];
        const T bb = T(1) - kia + kjb;
        const T Xa = T(2) / (bb + sqrt(bb * bb + T(4) * kia));
        const T Xb = T(1) / (T(1) + kia * Xa);
#pragma unroll
        for (int t = 0; t < NS; t++) X[t] = (t == a) ? Xa : ((t == b) ? Xb : T(1));
    } else {
        T K[NS][NS];
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int t = 0; t < NS; t++) {
                T k = 0;
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const bool st = (P.pair_a[p] == s && P.pair_b[p] == t), ts = (P.pair_a[p] == t && P.pair_b[p] == s);
                    k += (st ? delta[p] : T(0)) + (ts ? delta[p] : T(0));
                }
                K[s][t] = k * v[t];
            }
#pragma unroll
        for (int t = 0; t < NS; t++) X[t] = T(0.9);
        const int nit = 5 * P.nsite;
        for (int it = 0; it < nit; it++) {
            T Xn[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                T c = 0;
#pragma unroll
                for (int t = 0; t < NS; t++) c += K[s][t] * X[t];
                Xn[s] = T(0.5) * (X[s] + T(1) / (T(1) + c));
            }
#pragma unroll
            for (int s = 0; s < NS; s++) X[s] = Xn[s];
        }
        for (int it = 0; it < nit; it++) {
            T C[NS], J[NS][NS + 1];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                T c = 0;
#pragma unroll
                for (int t = 0; t < NS; t++) c += K[s][t] * X[t];
                C[s] = c;
#pragma unroll
                for (int t = 0; t < NS; t++) J[s][t] = (s == t ? T(1) + c : T(0)) + X[s] * K[s][t];
                J[s][NS] = X[s] * (T(1) + c) - T(1);
            }
#pragma unroll
            for (int k = 0; k < NS; k++)
#pragma unroll
                for (int i = k + 1; i < NS; i++) {
                    const T f = J[i][k] / J[k][k];
#pragma unroll
                    for (int j = k + 1; j <= NS; j++) J[i][j] -= f * J[k][j];
                }
            T dX[NS];
#pragma unroll
            for (int k = NS - 1; k >= 0; k--) {
                T r = J[k][NS];
#pragma unroll
                for (int j = k + 1; j < NS; j++) r -= J[k][j] * dX[j];
                dX[k] = r / J[k][k];
            }
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const T cand = X[s] - dX[s];
                X[s] = (cand > T(0) && cand < T(1)) ? cand : T(1) / (T(1) + C[s]);
            }
        }
    }

    T phi = 0, lx[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) {
        lx[t] = log(fabs(X[t]));
        phi += v[t] * (lx[t] - T(0.5) * X[t] + T(0.5));
    }
    T dw[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) dw[i] = 0;
#pragma unroll
    for (int t = 0; t < NS; t++)
#pragma unroll
        for (int i = 0; i < NC; i++)
            dw[i] += (t < P.nsite && P.site_comp[t] == i) ? T(P.site_count[t]) * lx[t] : T(0);
    T dn3 = 0, dn2 = 0, dxi = 0;
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const int a = P.pair_a[p], b = P.pair_b[p];
        const T vx = pick(v, a) * pick(X, a) * pick(v, b) * pick(X, b);
        const T dg = (p < P.npair) ? -vx * T(P.pair_k[p]) : T(0);
        const T dd = T(P.pair_dij[p]);
        dn3 += dg * (a1 * a1 + dd * xim * n2m * a1 * a1 * a1 + dd * dd * n2m * n2m * xim * a1 * a1 * a1 * a1 / T(6));
        dn2 += dg * (T(0.5) * dd * xim * a1 * a1 + dd * dd * n2m * xim * a1 * a1 * a1 / T(9));
        dxi += dg * (T(0.5) * dd * n2m * a1 * a1 + dd * dd * n2m * n2m * a1 * a1 * a1 / T(18));
    }
    dn2 += dxi * T(2) * nv2sq / (n2m * n2m * n2m);
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const T mi = T(P.m[i]), di = T(P.dd[i]);
        g[i][1] += dn2 * pi * mi * di + dw[i] * dwS[i];
        g[i][2] += dn3 * mi;
#pragma unroll
        for (int j = 0; j < ND; j++) g[i][3 + j] += dxi * (-T(2) * nv2m[j] / (n2m * n2m)) * (-twopi * mi) + dw[i] * dwv[i][j];
    }
    return phi;
}

// Phi at one point; g receives dPhi/dn for every (component, slot).  CHAIN = false
// is for callers that know every segment number is 1 (no chain term): the body is
// then free of branches, so two inlined calls can be scheduled side by side.
// AS selects the association term: 0 none, 1 one pair in closed form, 2 general.
// F32IN = true is for weighted densities from FP32 convolutions of a profile whose
// density is negligible somewhere: where the chain's lambda slot is below its FP32
// resolution (P.lsl_noise), -rho (m - 1) / lambda is rounding noise over rounding
// noise, and that slot's gradient is taken as 0, as the FP64 map's is to rounding
// there (its lambda, though small, is resolved against a negligible rho).  A
// spuriously large value would reach every grid point through the transposed FFT.
template <typename T, typename TS, int NC, int ND, bool CHAIN = true, int AS = 0, bool F32IN = false>
__device__ __forceinline__ T phi_grad(const PhiConst &P, const T (&nl)[NC][6 + ND], T (&g)[NC][6 + ND])
{
    constexpr int F = 6 + ND;
    constexpr int IZ = 3 + ND, IL = 4 + ND, IDP = 5 + ND;
    const T pi = T(kPi), twopi = T(2.0 * kPi);
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
        for (int f = 0; f < F; f++) g[i][f] = T(0);

    // ---- hard sphere ----
    T n0 = 0, n1 = 0, n2 = 0, nv1[ND], nv2[ND];
    TS n3 = 0;
#pragma unroll
    for (int j = 0; j < ND; j++) { nv1[j] = 0; nv2[j] = 0; }
#pragma unroll
    for (int i = 0; i < NC; i++) {
        T mi = T(P.m[i]), di = T(P.dd[i]);
        T nim = nl[i][1] * mi;
        n0 += nim / di;
        n1 += T(0.5) * nim;
        n2 += pi * nim * di;
#pragma unroll
        for (int j = 0; j < ND; j++) {
            T nvi = nl[i][3 + j];
            nv1[j] += nvi * mi / di;
            nv2[j] -= twopi * nvi * mi;
        }
        n3 += TS(nl[i][2]) * TS(P.m[i]);
    }
    T nv1v2 = 0, nv2sq = 0;
#pragma unroll
    for (int j = 0; j < ND; j++) { nv1v2 += nv1[j] * nv2[j]; nv2sq += nv2[j] * nv2[j]; }
    const TS twelvepi = TS(12.0 * kPi);
    TS dn = TS(1) - n3;
    TS log1dn = log(fabs(dn));
    TS inv1dn = TS(1) / dn;
    TS bkt = log1dn / (twelvepi * n3 * n3) + TS(1) / (twelvepi * n3 * dn * dn);
    TS dB = (TS(-1) / (n3 * n3 * dn) - TS(2) * log1dn / (n3 * n3 * n3) +
             (TS(3) * n3 - TS(1)) / (n3 * n3 * dn * dn * dn)) / twelvepi;
    {
        // The closed form cancels two O(1/n3) terms: its rounding error grows as
        // eps/n3^2 (eps/n3^3 for dB), which next to sharp walls, where a convolution
        // rings around n3 = 0 while n2 does not, reaches the gradient through
        // n2^3 dB.  Below |n3| = 1e-3 the same function comes from its Taylor series,
        // sum_k (k+1)(k+3)/(k+2) n3^k / (12 pi), exact to FP64 rounding; selected
        // rather than branched to, so the body stays one block.
        const bool small = fabs(n3) < TS(1e-3);
        const TS bs = (TS(1.5) + n3 * (TS(8.0 / 3.0) + n3 * (TS(15.0 / 4.0) + n3 * (TS(24.0 / 5.0) +
                       n3 * (TS(35.0 / 6.0) + n3 * (TS(48.0 / 7.0) + n3 * TS(63.0 / 8.0))))))) / twelvepi;
        const TS ds = (TS(8.0 / 3.0) + n3 * (TS(2 * 15.0 / 4.0) + n3 * (TS(3 * 24.0 / 5.0) + n3 * (TS(4 * 35.0 / 6.0) +
                       n3 * (TS(5 * 48.0 / 7.0) + n3 * TS(6 * 63.0 / 8.0)))))) / twelvepi;
        bkt = small ? bs : bkt;
        dB = small ? ds : dB;
    }
    T lg = T(log1dn), iv = T(inv1dn), bk = T(bkt);
    T n2c = n2 * n2 * n2 / T(3) - n2 * nv2sq;
    T phi = -n0 * lg + (n1 * n2 - nv1v2) * iv + n2c * bk;
    {
        T d_n0 = -lg;
        T d_n1 = n2 * iv;
        T d_n2 = n1 * iv + (n2 * n2 - nv2sq) * bk;
        T d_n3 = T(TS(n0) * inv1dn + TS(n1 * n2 - nv1v2) * inv1dn * inv1dn + TS(n2c) * dB);
#pragma unroll
        for (int i = 0; i < NC; i++) {
            T mi = T(P.m[i]), di = T(P.dd[i]);
            g[i][1] += mi * (d_n0 / di + T(0.5) * d_n1 + pi * di * d_n2);
            g[i][2] += mi * d_n3;
#pragma unroll
            for (int j = 0; j < ND; j++) {
                T d_v1 = -nv2[j] * iv;
                T d_v2 = -nv1[j] * iv - T(2) * n2 * nv2[j] * bk;
                g[i][3 + j] += mi * d_v1 / di - twopi * mi * d_v2;
            }
        }
    }

    if constexpr (AS > 0) phi += assoc_grad<T, NC, ND, AS>(P, nl, g, n2, T(n3), nv2);

    // ---- hard chain ----
    if (CHAIN) {
        T z3 = 0, z2 = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            T rh = nl[i][IZ];
            z3 += T(P.m[i]) * rh;
            z2 += T(P.m[i]) * rh / T(P.dd[i]);
        }
        z3 /= T(8);
        z2 /= T(8);
        T inv = T(1) / (T(1) - z3);
        T G3 = 0, G2 = 0;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            T mi1 = T(P.m[i]) - T(1);
            if (P.m[i] == 1.0) continue;     // the chain term and its gradient vanish exactly for m = 1
            T di = T(P.dd[i]);
            T rhoi = nl[i][0];
            T lsl = nl[i][IL];
            if (std::is_same<T, float>::value) {   // FP32: keep underflowed densities off 0/0
                rhoi = fmax(rhoi, T(1e-30));
                lsl = copysign(fmax(fabs(lsl), T(1e-30)), lsl);
            }
            T lam = lsl / (T(2) * di);
            T y = inv + T(1.5) * di * z2 * inv * inv + T(0.5) * di * di * z2 * z2 * inv * inv * inv;
            T l = T(log(fabs(TS(y * lam / rhoi))));
            phi += -rhoi * mi1 * l;
            g[i][0] += -mi1 * (l - T(1));
            g[i][IL] += -rhoi * mi1 / lsl;
            T gy = -rhoi * mi1 / y;
            T dy3 = inv * inv + T(3) * di * z2 * inv * inv * inv + T(1.5) * di * di * z2 * z2 * inv * inv * inv * inv;
            T dy2 = T(1.5) * di * inv * inv + di * di * z2 * inv * inv * inv;
            G3 += gy * dy3;
            G2 += gy * dy2;
        }
#pragma unroll
        for (int k = 0; k < NC; k++) g[k][IZ] += T(P.m[k]) / T(8) * G3 + T(P.m[k]) / (T(8) * T(P.dd[k])) * G2;
    }

    // ---- dispersion ----
    const T psi = T(P.psi);
    const T factor = T(3) / (T(4) * psi * psi * psi * pi);
    T rb[NC], S = 0, mt = 0, es = 0;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        T di = T(P.dd[i]);
        rb[i] = nl[i][IDP] * factor / (di * di * di);
        S += rb[i];
        mt += rb[i] * T(P.m[i]);
        es += T(P.m[i]) * rb[i] * di * di * di;
    }
    // FP32: keep an underflowed S off 0/0, and keep its sign (a convolution may ring
    // below zero where the density is negligible, as in FP64)
    if (std::is_same<T, float>::value) S = copysign(fmax(fabs(S), T(1e-30)), S);
    T mb = (std::is_same<T, float>::value && NC == 1) ? T(P.m[0]) : mt / S;
    T eta = es * pi / T(6);
    T A1 = 0, A2 = 0, dA1[NC], dA2[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) { dA1[i] = 0; dA2[i] = 0; }
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
        for (int j = 0; j < NC; j++) {
            T cm = T(P.cm[i * NC + j]);
            T eT = T(P.eT[i * NC + j]);
            T cij = rb[i] * rb[j] * cm;
            A1 += cij * eT;
            A2 += cij * eT * eT;
            dA1[i] += T(2) * rb[j] * cm * eT;
            dA2[i] += T(2) * rb[j] * cm * eT * eT;
        }
    T e2 = eta * eta, om = T(1) - eta, om4 = om * om * om * om;
    T i1e = T(1) / (T(1) - eta), i2e = T(1) / (T(2) - eta);
    T Pn = T(8) * eta - T(2) * e2;
    T Qn = T(20) * eta - T(27) * e2 + T(12) * (eta * e2) - T(2) * (e2 * e2);
    T q12 = i1e * i1e * i2e * i2e;
    T C1 = T(1) + mb * Pn / om4 + (T(1) - mb) * Qn * q12;
    T I1, I1m, I1e, I2, I2m, I2e;
    i_lite<T>(0, mb, eta, I1, I1m, I1e);
    i_lite<T>(1, mb, eta, I2, I2m, I2e);
    T iC1 = T(1) / C1;
    phi += -T(2) * (pi * I1 * A1) - pi * mb * I2 * A2 * iC1;
    {
        T Pv = Pn / om4, Qv = Qn * q12;
        T Pe = (T(8) - T(4) * eta) / om4 + T(4) * Pn / (om4 * om);
        T Qe = (T(20) - T(54) * eta + T(36) * e2 - T(8) * eta * e2) * q12 +
               Qn * (T(2) * i1e * q12 + T(2) * i2e * q12);
        T C1m = Pv - Qv, C1e = mb * Pe + (T(1) - mb) * Qe;
        T iS = T(1) / S;
#pragma unroll
        for (int k = 0; k < NC; k++) {
            T di = T(P.dd[k]);
            T dm = (T(P.m[k]) - mb) * iS;
            T de = pi / T(6) * T(P.m[k]) * di * di * di;
            T gk = -T(2) * pi * ((I1m * dm + I1e * de) * A1 + I1 * dA1[k]);
            gk -= pi * (dm * I2 * A2 * iC1 + mb * (I2m * dm + I2e * de) * A2 * iC1 + mb * I2 * dA2[k] * iC1 -
                        mb * I2 * A2 * (C1m * dm + C1e * de) * iC1 * iC1);
            g[k][IDP] += gk * factor / (di * di * di);
        }
    }
    if constexpr (F32IN) {
#pragma unroll
        for (int i = 0; i < NC; i++) g[i][IL] = (fabs(nl[i][IL]) < T(P.lsl_noise[i])) ? T(0) : g[i][IL];
    }
    return phi;
}

}  // namespace cdft
