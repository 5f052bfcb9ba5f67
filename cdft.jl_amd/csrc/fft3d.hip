// fft3d.hip -- weighted densities and transposed convolutions on 3D Cartesian
// grids (power-of-two axes) with line FFTs in LDS.
//
// The reference convolves every (field, component) with its analytic kernel by an
// FFT, a multiply and an inverse FFT (src/utils/integrals.jl:88-96,
// src/fields/weighted_densities.jl:118-163,345-383).  Here, per component c:
//   P   = FFT3(rho_c)                                         3 passes
//   for each field pair (f1, f2):
//     n_f1 + i n_f2 = IFFT3(P Z) N_A L^3 / N                  3 passes; the multiply is
//                                                             fused into the first, the
//                                                             field split into the last
//   for each pair of gradients (u1, u2) = dPhi/dn_(f1, f2):
//     S += conj(Z) FFT3(u1 + i u2)                            3 passes; accumulate fused
//   dF/drho_c = dPhi/dn_rho + Re IFFT3(S) / N                 3 passes; sum fused
// with Z = H_f1 + i H_f2 and H_f the Hermitian part of the reference's weight (the
// part real() keeps): H = G (scalar weights) or i G (vector weights), G real and
// tabulated once per system.  For real kernels h_f,
// Re IFFT(conj(Z) FFT(u1 + i u2)) = h1^T u1 + h2^T u2, so no conjugate-symmetric
// separation is needed; conj(H) is the transposed operator (the vector weights'
// transpose is their negative, weighted_densities.jl:381).
#pragma once
#include "common.h"

namespace cdft {

template <typename T>
struct cx {
    T x, y;
};

// Complex product with its multiply-adds written out, so that every kernel that
// inlines it rounds alike (the compiler may otherwise fuse either product).
template <typename T>
__device__ __attribute__((always_inline)) inline cx<T> cxmul(cx<T> a, cx<T> b)
{
    return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
}


enum { IN_CPLX = 0, IN_RHO = 1, IN_PAIR = 2, IN_SPEC = 3, IN_MULT = 4 };
enum { OUT_CPLX = 0, OUT_FIELDS = 1, OUT_ACCUM = 2, OUT_REAL = 3, OUT_DDFT0 = 4, OUT_DDFT1 = 5 };

struct Pass3 {
    const void *in;           // IN_CPLX / IN_SPEC: complex T; IN_RHO: double rho; IN_PAIR: T fields
    const void *in2;          // IN_PAIR: second field
    void *out;                // OUT_CPLX / OUT_ACCUM: complex T; OUT_FIELDS: T field 1; OUT_REAL: double
    void *out2;               // OUT_FIELDS: T field 2
    const void *base;         // OUT_REAL: T values added to the result (dPhi/dn_rho), or null
    const cx<double> *tw;     // W_L^t for this axis, t < L
    const void *g1, *g2;      // IN_SPEC / OUT_ACCUM: real spectra G (type T) of the pair's fields
    int v1, v2;               // 1 where the field is a vector weight (H = i G)
    int first;                // OUT_ACCUM: overwrite instead of accumulate
    int inverse;
    double scale;             // OUT_FIELDS / OUT_REAL / OUT_DDFT*
    int mkind;                // IN_MULT: 0 the Laplacian map, 1 + a the gradient map along axis a
    const int *gate;          // the pass returns at once when *gate is nonzero (device-side solver)
    // Several components in one launch (blockIdx.y = component c, ncomp of them):
    // component c's in, in2, out, out2, base, g1, g2 lie cs[0..6] bytes times c
    // further on.  Zero strides and ncomp = 0 or 1: one component.
    int ncomp;
    long cs[7];
};

// Pass p of component c (null pointers stay null).
__device__ __attribute__((always_inline)) inline void pass_component(Pass3 &p, long c)
{
    auto adv = [c](auto *q, long st) { return q ? reinterpret_cast<decltype(q)>(
                                                      reinterpret_cast<const char *>(q) + c * st)
                                                : q; };
    p.in = adv(p.in, p.cs[0]);
    p.in2 = adv(p.in2, p.cs[1]);
    p.out = const_cast<void *>(adv(static_cast<const void *>(p.out), p.cs[2]));
    p.out2 = const_cast<void *>(adv(static_cast<const void *>(p.out2), p.cs[3]));
    p.base = adv(p.base, p.cs[4]);
    p.g1 = adv(p.g1, p.cs[5]);
    p.g2 = adv(p.g2, p.cs[6]);
}

// The DDFT derivative maps at grid index g (ext/SciMLBasecDFTExt.jl:16-19): the
// Laplacian -(2 pi)^2 |k|^2 (real) or the gradient 2 pi i k_a, k the reduced
// frequencies; the gradient's Nyquist term along a is not Hermitian and real()
// drops it, so it is 0 here.
__device__ __attribute__((always_inline)) inline cx<double> ddft_map(const Grid3 &G, long g, int kind)
{
    const int n0 = G.n[0], n1 = G.n[1];
    const int ii[3] = {(int)(g % n0), (int)((g / n0) % n1), (int)(g / ((long)n0 * n1))};
    const double twopi = 2.0 * kPi;
    if (kind == 0) {
        double k2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int ks = (ii[a] < (G.n[a] + 1) / 2) ? ii[a] : ii[a] - G.n[a];
            const double w = (double)ks * G.mult[a];
            k2 += w * w;
        }
        return {-twopi * twopi * k2, 0.0};
    }
    const int a = kind - 1;
    if (2 * ii[a] == G.n[a]) return {0.0, 0.0};
    const int ks = (ii[a] < (G.n[a] + 1) / 2) ? ii[a] : ii[a] - G.n[a];
    return {0.0, twopi * (double)ks * G.mult[a]};
}

// S + conj(Z) A (S ignored when first), with the multiply-adds written out so that
// every pass that accumulates (k_fft3's OUT_ACCUM, k_fft3z_accum) rounds alike.
template <typename T>
__device__ __attribute__((always_inline)) inline cx<T> accum_conj(cx<T> s, cx<T> v, cx<T> z, bool first)
{
    const T tx = fma(v.y, z.y, v.x * z.x);
    const T ty = fma(-v.x, z.y, v.y * z.x);
    return first ? cx<T>{tx, ty} : cx<T>{s.x + tx, s.y + ty};
}

// Z = H1 + i H2 at one frequency from the tabulated G (H = G or i G).
template <typename T>
__device__ __attribute__((always_inline)) inline cx<T> pair_z(const Pass3 &p, long g)
{
    const T a = reinterpret_cast<const T *>(p.g1)[g], b = reinterpret_cast<const T *>(p.g2)[g];
    const cx<T> h1 = p.v1 ? cx<T>{T(0), a} : cx<T>{a, T(0)};
    const cx<T> h2 = p.v2 ? cx<T>{T(0), b} : cx<T>{b, T(0)};
    return {h1.x - h2.y, h1.y + h2.x};
}

// Field table of PC-SAFT in 3D (PCSAFT.jl:55-61): kind 1 int rho dz, 2 int rho z^2 dz,
// 3+j vector component j; width index 0: d/2, 1: d, 2: psi d.
__device__ __attribute__((always_inline)) inline void field3(int f, int &kind, int &w)
{
    switch (f) {
    case 1: kind = 1; w = 0; break;
    case 2: kind = 2; w = 0; break;
    case 3: kind = 3; w = 0; break;
    case 4: kind = 4; w = 0; break;
    case 5: kind = 5; w = 0; break;
    case 6: kind = 2; w = 1; break;
    case 7: kind = 1; w = 1; break;
    default: kind = 2; w = 2; break;
    }
}

// Hermitian part of Omega_f at grid index (i0, i1, i2), reduced units
// (weighted_densities.jl:48-69, :210-227; frequencies structure.jl:153-185).
template <typename T>
__device__ inline cx<T> omega3(const Grid3 &G, int c, int f, int i0, int i1, int i2)
{
    int kind, wi;
    field3(f, kind, wi);
    const int ii[3] = {i0, i1, i2};
    double w[3], wb2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int nn = G.n[a];
        int ks = (ii[a] < (nn + 1) / 2) ? ii[a] : ii[a] - nn;
        w[a] = (double)ks * G.mult[a];
        wb2 += w[a] * w[a];
    }
    const double wb = sqrt(wb2);
    const double R = G.R[c][wi];
    const double twopi = 2.0 * kPi, fourpi = 4.0 * kPi, tp3 = twopi * twopi * twopi;
    if (kind == 1) {
        double om = (wb == 0.0 ? 2.0 * R : 2.0 * sin(wb * R) / wb) / twopi;
        return {T(om), T(0)};
    }
    if (kind == 2) {
        double om = (wb == 0.0 ? fourpi * (R * R * R) / 3.0
                               : fourpi / (wb * wb * wb) * (sin(wb * R) - R * wb * cos(wb * R))) / tp3;
        return {T(om), T(0)};
    }
    // vector component j: i W; no Hermitian part at k = 0 or on the Nyquist plane of axis j
    const int j = kind - 3;
    if (wb == 0.0 || 2 * ii[j] == G.n[j]) return {T(0), T(0)};
    double W = fourpi * w[j] / (wb * wb * wb) * (sin(wb * R) - R * wb * cos(wb * R)) / tp3;
    return {T(0), T(W)};
}

// Lines per LDS tile of the line-FFT passes (the default; the autotuner picks 4, 8
// or 16 per system): 32 KB per workgroup in FP64 at 256 points, so several
// workgroups per CU keep loads in flight.  The arithmetic of a line does not depend
// on it, so neither do the results.
constexpr int TB3 = 8;

// Radix-4 Stockham passes (+ a final radix-2 when log2 L is odd) along j over a tile
// of TB lines in LDS laid out as in k_fft3 (axis AX), in place.  All 256 threads;
// the tile must be complete before the call (barrier) and is complete after it.
// EPT = radix-4 butterflies per thread and pass (L TB / 1024, at least 1).  tw: the
// line's twiddles W_L^t in T (staged in LDS by the caller).  The first pass's twiddles are
// all W^0 = 1 and are skipped (the explicit-fma product by (1, 0) returns its operand).
template <typename T, int AX, int EPT, int TB = TB3>
__device__ __attribute__((always_inline)) inline void tile_fft(cx<T> *a, const cx<T> *tw, int L, int lg,
                                                               bool inverse)
{
    constexpr int EPR = EPT > 0 ? EPT : 1;
    const int tid = threadIdx.x;
    auto lds = [&](int j, int b) -> int { return (AX == 0) ? b * L + j : j * TB + b; };
    // element e of a (2^llen) x TB tile; lengths are powers of two, so shifts, not
    // integer division
    auto split = [&](int e, int llen, int &j, int &b) {
        if (AX == 0) { j = e & ((1 << llen) - 1); b = e >> llen; } else { b = e % TB; j = e / TB; }
    };
    const T sg = inverse ? T(1) : T(-1);
    int Ns = 1, lNs = 0;
    const int nr4 = lg >> 1;
    for (int pass = 0; pass < nr4; pass++) {
        const int nb = L >> 2;
        const int step = L / (Ns * 4);
        cx<T> v[EPR][4];
#pragma unroll
        for (int q = 0; q < EPR; q++) {
            const int e = tid + q * 256;
            if (e < nb * TB) {
                int jj, b;
                split(e, lg - 2, jj, b);
                const int k = jj & (Ns - 1);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    cx<T> x = a[lds(jj + r * nb, b)];
                    if (r && pass > 0) {
                        const cx<T> wd = tw[(r * k * step) & (L - 1)];
                        x = cxmul(x, cx<T>{wd.x, inverse ? -wd.y : wd.y});
                    }
                    v[q][r] = x;
                }
                const cx<T> s02 = {v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y};
                const cx<T> d02 = {v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y};
                const cx<T> s13 = {v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y};
                const cx<T> d13 = {v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y};
                const cx<T> id13 = {-sg * d13.y, sg * d13.x};
                v[q][0] = {s02.x + s13.x, s02.y + s13.y};
                v[q][1] = {d02.x + id13.x, d02.y + id13.y};
                v[q][2] = {s02.x - s13.x, s02.y - s13.y};
                v[q][3] = {d02.x - id13.x, d02.y - id13.y};
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < EPR; q++) {
            const int e = tid + q * 256;
            if (e < nb * TB) {
                int jj, b;
                split(e, lg - 2, jj, b);
                const int k = jj & (Ns - 1);
                const int d = ((jj >> lNs) << (lNs + 2)) + k;
#pragma unroll
                for (int r = 0; r < 4; r++) a[lds(d + r * Ns, b)] = v[q][r];
            }
        }
        __syncthreads();
        Ns <<= 2;
        lNs += 2;
    }
    if (lg & 1) {
        const int nb = L >> 1;
        cx<T> v[2 * EPR][2];
#pragma unroll
        for (int q = 0; q < 2 * EPR; q++) {
            const int e = tid + q * 256;
            if (e < nb * TB) {
                int jj, b;
                split(e, lg - 1, jj, b);
                const int k = jj & (Ns - 1);
                const cx<T> wd = tw[(k * (L / (Ns * 2))) & (L - 1)];
                const cx<T> w = {wd.x, inverse ? -wd.y : wd.y};
                const cx<T> xa = a[lds(jj, b)], xb = cxmul(a[lds(jj + nb, b)], w);
                v[q][0] = {xa.x + xb.x, xa.y + xb.y};
                v[q][1] = {xa.x - xb.x, xa.y - xb.y};
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2 * EPR; q++) {
            const int e = tid + q * 256;
            if (e < nb * TB) {
                int jj, b;
                split(e, lg - 1, jj, b);
                const int k = jj & (Ns - 1);
                const int d = ((jj >> lNs) << (lNs + 1)) + k;
                a[lds(d, b)] = v[q][0];
                a[lds(d + Ns, b)] = v[q][1];
            }
        }
        __syncthreads();
    }
}

// One FFT pass along axis AX over a tile of TB lines held in LDS.  Lines along x
// are contiguous (LDS [b][j], j fastest); along y and z a tile is TB consecutive x
// values (LDS [j][b], b fastest), so global accesses are coalesced and LDS accesses
// of consecutive lanes are consecutive.  EPT = radix-4 butterflies per thread.
template <typename T, int AX, int IN, int OUT, int EPT, int TB = TB3>
__global__ __launch_bounds__(256) void k_fft3(Grid3 G, Pass3 p)
{
    if (p.gate && *p.gate) return;
    if (blockIdx.y) pass_component(p, blockIdx.y);
    constexpr int EPR = EPT > 0 ? EPT : 1;     // radix-4 butterflies per thread and pass
    extern __shared__ __align__(16) unsigned char smem3[];
    cx<T> *a = reinterpret_cast<cx<T> *>(smem3);
    const int tid = threadIdx.x;
    const int L = G.n[AX], lg = G.lg[AX];
    const int n0 = G.n[0], n1 = G.n[1];
    long base, lstride, bstride;
    int x0 = 0, y0 = 0, z0 = 0;
    long line0 = 0;
    if (AX == 0) {
        line0 = (long)blockIdx.x * TB;
        base = line0 * n0;
        lstride = 1;
        bstride = n0;
    } else if (AX == 1) {
        const int tiles_x = n0 / TB;
        x0 = (blockIdx.x % tiles_x) * TB;
        z0 = blockIdx.x / tiles_x;
        base = x0 + (long)n0 * n1 * z0;
        lstride = n0;
        bstride = 1;
    } else {
        const int tiles_x = n0 / TB;
        x0 = (blockIdx.x % tiles_x) * TB;
        y0 = blockIdx.x / tiles_x;
        base = x0 + (long)n0 * y0;
        lstride = (long)n0 * n1;
        bstride = 1;
    }
    auto lds = [&](int j, int b) -> int { return (AX == 0) ? b * L + j : j * TB + b; };
    auto split = [&](int e, int llen, int &j, int &b) {    // element e of a (2^llen) x TB tile
        if (AX == 0) { j = e & ((1 << llen) - 1); b = e >> llen; } else { b = e % TB; j = e / TB; }
    };
    const int total = L * TB;
    // ---- load (unrolled: the tile's global reads are all in flight together).
    // EPT > 0: L TB = 1024 EPT, EL = L TB / 256 = 4 EPT elements per thread, no guard;
    // EPT = 0: short lines (L TB <= 512) with a guard.
    constexpr bool FULL = EPT > 0;
    constexpr int EL = FULL ? 4 * EPT : 2;
#pragma unroll
    for (int q = 0; q < EL; q++) {
        const int e = tid + q * 256;
        if (FULL || e < total) {
            int j, b;
            split(e, lg, j, b);
            const long g = base + (long)j * lstride + (long)b * bstride;
            cx<T> v;
            if (IN == IN_CPLX) {
                v = reinterpret_cast<const cx<T> *>(p.in)[g];
            } else if (IN == IN_RHO) {
                v = {T(reinterpret_cast<const double *>(p.in)[g]), T(0)};
            } else if (IN == IN_PAIR) {
                v = {reinterpret_cast<const T *>(p.in)[g], reinterpret_cast<const T *>(p.in2)[g]};
            } else if (IN == IN_MULT) {
                const cx<T> sv = reinterpret_cast<const cx<T> *>(p.in)[g];
                const cx<double> m = ddft_map(G, g, p.mkind);
                v = cxmul(sv, cx<T>{T(m.x), T(m.y)});
            } else {  // IN_SPEC: P Z
                const cx<T> s = reinterpret_cast<const cx<T> *>(p.in)[g];
                v = cxmul(s, pair_z<T>(p, g));
            }
            a[lds(j, b)] = v;
        }
    }
    cx<T> *twl = a + (size_t)L * TB;                   // W_L^t in T (rounded once, as at each use)
    for (int t = tid; t < L; t += 256) twl[t] = cx<T>{T(p.tw[t].x), T(p.tw[t].y)};
    __syncthreads();
    tile_fft<T, AX, EPT, TB>(a, twl, L, lg, p.inverse != 0);
    // ---- store
#pragma unroll
    for (int q = 0; q < EL; q++) {
        const int e = tid + q * 256;
        if (!FULL && e >= total) continue;
        int j, b;
        split(e, lg, j, b);
        const long g = base + (long)j * lstride + (long)b * bstride;
        const cx<T> v = a[lds(j, b)];
        if (OUT == OUT_CPLX) {
            reinterpret_cast<cx<T> *>(p.out)[g] = v;
        } else if (OUT == OUT_FIELDS) {
            reinterpret_cast<T *>(p.out)[g] = T(p.scale) * v.x;
            reinterpret_cast<T *>(p.out2)[g] = T(p.scale) * v.y;
        } else if (OUT == OUT_REAL) {
            reinterpret_cast<double *>(p.out)[g] =
                p.scale * double(v.x) + (p.base ? double(reinterpret_cast<const T *>(p.base)[g]) : 0.0);
        } else if (OUT == OUT_DDFT0) {          // lap(eta) + lap(mu)
            reinterpret_cast<double *>(p.out)[g] = p.scale * double(v.x) + p.scale * double(v.y);
        } else if (OUT == OUT_DDFT1) {          // += grad(eta)^2 + grad(mu) grad(eta)
            const double ge = p.scale * double(v.x), gm = p.scale * double(v.y);
            reinterpret_cast<double *>(p.out)[g] += ge * ge + gm * ge;
        } else {  // OUT_ACCUM: S += conj(Z) A
            cx<T> *S = reinterpret_cast<cx<T> *>(p.out);
            S[g] = accum_conj<T>(p.first ? cx<T>{T(0), T(0)} : S[g], v, pair_z<T>(p, g), p.first != 0);
        }
    }
}

// The transposed side's z passes of one component in one launch (z lines of 128 or
// 256 points, two LDS tiles of TBZ lines: 64 KB in FP64 at 256 points, so two
// workgroups fit a CU): for each field pair q the tile of B_q (gradient pair
// transformed along x and y) is transformed along z and S += conj(Z_q) FFT_z(B_q)
// accumulates in LDS; S is then transformed back along z and written out.  This
// replaces four accumulating k_fft3 passes, each re-reading and re-writing S, and
// the first inverse pass; arithmetic and order are the same.  (The same fusion on
// the weighted side, P transformed once and four B_q = IFFT_z(P Z_q) written,
// measured slower than the four k_fft3 passes and is not used.)
struct ZPairs3 {
    void *out;                // S (complex T)
    const void *b[4];         // B_q
    const void *g1[4], *g2[4];  // spectra G (type T) of each pair's fields
    int v1[4], v2[4];         // 1 where the field is a vector weight (H = i G)
    int np;
    const cx<double> *tw;     // W_L^t along z
    const int *gate;          // as Pass3::gate
};

template <typename T>
__device__ __attribute__((always_inline)) inline cx<T> zpair_z(const ZPairs3 &p, int q, long g)
{
    const T a = reinterpret_cast<const T *>(p.g1[q])[g], b = reinterpret_cast<const T *>(p.g2[q])[g];
    const cx<T> h1 = p.v1[q] ? cx<T>{T(0), a} : cx<T>{a, T(0)};
    const cx<T> h2 = p.v2[q] ? cx<T>{T(0), b} : cx<T>{b, T(0)};
    return {h1.x - h2.y, h1.y + h2.x};
}

constexpr int TBZ = 8;

template <typename T, int EPT>
__global__ __launch_bounds__(256) void k_fft3z_accum(Grid3 G, ZPairs3 p)
{
    if (p.gate && *p.gate) return;
    static_assert(EPT > 0, "full tiles only");
    constexpr int TB = TBZ, EL = 4 * EPT;             // L = 128 EPT
    extern __shared__ __align__(16) unsigned char smem3[];
    const int tid = threadIdx.x;
    const int L = G.n[2], lg = G.lg[2];
    cx<T> *a0 = reinterpret_cast<cx<T> *>(smem3);     // S
    cx<T> *a1 = a0 + (size_t)L * TB;                    // one pair's line tile
    cx<T> *twl = a1 + (size_t)L * TB;                   // W_L^t in T
    for (int t = tid; t < L; t += 256) twl[t] = cx<T>{T(p.tw[t].x), T(p.tw[t].y)};
    const int n0 = G.n[0], tiles_x = n0 / TB;
    const int x0 = (blockIdx.x % tiles_x) * TB, y0 = blockIdx.x / tiles_x;
    const long base = x0 + (long)n0 * y0, lstride = (long)n0 * G.n[1];
    auto at = [&](int e, int &l, long &g) __attribute__((always_inline)) {   // LDS index and global index
        const int b = e % TB, j = e / TB;
        l = j * TB + b;
        g = base + (long)j * lstride + b;
    };
    for (int k = 0; k < p.np; k++) {
#pragma unroll
        for (int q = 0; q < EL; q++) {
            int l;
            long g;
            at(tid + q * 256, l, g);
            a1[l] = reinterpret_cast<const cx<T> *>(p.b[k])[g];
        }
        __syncthreads();
        tile_fft<T, 2, EPT, TBZ>(a1, twl, L, lg, false);
#pragma unroll
        for (int q = 0; q < EL; q++) {
            int l;
            long g;
            at(tid + q * 256, l, g);
            a0[l] = accum_conj<T>(k == 0 ? cx<T>{T(0), T(0)} : a0[l], a1[l], zpair_z<T>(p, k, g), k == 0);
        }
        __syncthreads();                                 // a1 is rewritten by the next pair; a0 complete
    }
    tile_fft<T, 2, EPT, TBZ>(a0, twl, L, lg, true);
#pragma unroll
    for (int q = 0; q < EL; q++) {
        int l;
        long g;
        at(tid + q * 256, l, g);
        reinterpret_cast<cx<T> *>(p.out)[g] = a0[l];
    }
}

// Tabulated real spectra G of field f of component c (Omega_herm = G or i G),
// evaluated in FP64; out32 (optional) receives the FP32 copy.
__global__ void k_setup_omega3(Grid3 G, int c, int f, double *out, float *out32)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G.N) return;
    const int i0 = (int)(g % G.n[0]);
    const int i1 = (int)((g / G.n[0]) % G.n[1]);
    const int i2 = (int)(g / ((long)G.n[0] * G.n[1]));
    const cx<double> o = omega3<double>(G, c, f, i0, i1, i2);
    int kind, w;
    field3(f, kind, w);
    out[g] = (kind >= 3) ? o.y : o.x;
    if (out32) out32[g] = (float)out[g];
}

}  // namespace cdft
