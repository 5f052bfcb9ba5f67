// phi.h -- pointwise PC-SAFT Helmholtz density Phi(n) and its gradient dPhi/dn
// at one grid point, for the HIP kernels.
//
// Restates f_res = f_hs + f_hc + f_disp (+ association = 0):
//   f_hs   src/models/DFT/FMT.jl:17-47 (White-Bear n3 dependence)
//   f_hc   src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:123-144
//   f_disp src/models/DFT/SAFT/PCSAFT/PCSAFT.jl:150-211
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

// Per-system constants in reduced units (lengths / L), set up on the host.
struct PhiConst {
    int nc;
    double T;
    double psi;                           // dispersion width factor (PCSAFT.jl:53,160)
    double m[CDFT_MAXC];
    double dd[CDFT_MAXC];                 // d_i / L
    double cm[CDFT_MAXC * CDFT_MAXC];     // m_i m_j (sigma_ij / L)^3
    double eT[CDFT_MAXC * CDFT_MAXC];     // eps_ij / T
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

// Phi at one point; g receives dPhi/dn for every (component, slot).  CHAIN = false
// is for callers that know every segment number is 1 (no chain term): the body is
// then free of branches, so two inlined calls can be scheduled side by side.
template <typename T, typename TS, int NC, int ND, bool CHAIN = true>
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
    if (std::is_same<T, float>::value) {
        // FP32 inputs: n3 is at rounding-noise level inside walls and may be 0.
        // Series of the same bracket, sum_k ((k+2)^2 - 1)/(k+2) n3^k / (12 pi);
        // selected rather than branched to, so the body stays one block.
        const bool small = fabs(n3) < TS(1e-5);
        const TS bs = (TS(1.5) + n3 * (TS(8.0 / 3.0) + n3 * (TS(3.75) + n3 * TS(4.8)))) / twelvepi;
        const TS ds = (TS(8.0 / 3.0) + n3 * (TS(7.5) + n3 * TS(14.4))) / twelvepi;
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
    if (std::is_same<T, float>::value) S = fmax(S, T(1e-30));
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
    return phi;
}

}  // namespace cdft
