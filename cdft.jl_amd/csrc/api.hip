// api.hip -- C ABI of the cDFT HIP library (include/cdft_hip.h): contexts, system
// set-up, functional evaluation and the plain Euler-Lagrange / Anderson solve.
// This is the single translation unit of the library; the kernel files are
// included below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "plain.hip"
#include "batched.hip"
#include "fft3d.hip"

using namespace cdft;

namespace {

thread_local std::string g_err;   // errors raised before a context exists

// Scratch of the allocation-free batched entry point
// (cdft_converge_batch_device_noalloc): the system descriptors and the results of
// one launch, in a static device array.  Each GPU holds its own copy for the life
// of the process; its address on a device comes from hipGetSymbolAddress, taken
// once per device at the first context created on it.
constexpr int kNoAllocMaxSys = CDFT_NOALLOC_MAX_SYS;
constexpr size_t kScratchDescBytes = ((sizeof(BSys) * kNoAllocMaxSys + 255) / 256) * 256;
constexpr size_t kScratchBytes = kScratchDescBytes + sizeof(BResult) * kNoAllocMaxSys;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

int fail(cdft_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg; else g_err = msg;
    return code;
}

inline int nblocks(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// Composite Simpson exactly as _int (src/utils/integrals.jl:28-40), 1D.
double simpson1d(const double *f, int n, double dz)
{
    double s = 0.0;
    for (int k = 1; k <= n; k++) {
        double c = (k == 1 || k == n) ? 1.0 : ((k % 2 == 0) ? 4.0 : 2.0);
        s += c * f[k - 1];
    }
    return s * (dz / 3.0);
}

// Captured solver bodies (one graph per map variant).  They hold the line-FFT
// variant they were captured with, so a new variant drops them too.
static void drop_body_graphs(cdft_sys *s)
{
    for (hipGraphExec_t &e : s->aa_exec)
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
}

void free_sys_buffers(cdft_sys *s)
{
    double **bufs[] = {&s->h, &s->hT, &s->vext, &s->cpd, &s->rho, &s->n, &s->dphi, &s->part, &s->fval, &s->out, &s->sol,
                       &s->gx, &s->df, &s->dg, &s->res, &s->Q, &s->Qalt, &s->DG, &s->red, &s->fin, &s->hddft,
                       &s->xsafe};
    for (double **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (s->H64) { (void)hipFree(s->H64); s->H64 = nullptr; }
    drop_body_graphs(s);
    if (s->ctl) { (void)hipFree(s->ctl); s->ctl = nullptr; }
    if (s->dhist) { (void)hipFree(s->dhist); s->dhist = nullptr; s->dhist_cap = 0; }
    double **b3[] = {&s->G3, &s->A3, &s->B3, &s->S3, &s->B3x[0], &s->B3x[1], &s->B3x[2]};
    for (double **b : b3)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    float **f3[] = {&s->G3f, &s->n32, &s->dphi32, &s->h32};
    for (float **b : f3)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
}

// Composite Simpson of _int over nd axes (integrals.jl:28-40): per-axis weights
// 1, 4, 2, ..., 1 times dz_a / 3, f column-major over ngrid.
double simpson_nd(const double *f, int nd, const int *ngrid, const double *dz)
{
    if (nd == 1) return simpson1d(f, ngrid[0], dz[0]);
    auto wgt = [](int k, int n) { return (k == 0 || k == n - 1) ? 1.0 : ((k % 2 == 1) ? 4.0 : 2.0); };
    const int n2 = nd == 3 ? ngrid[2] : 1;
    double s = 0.0;
    for (int i2 = 0; i2 < n2; i2++)
        for (int i1 = 0; i1 < ngrid[1]; i1++) {
            const double w12 = wgt(i1, ngrid[1]) * (nd == 3 ? wgt(i2, n2) : 1.0);
            const double *row = f + ((size_t)i2 * ngrid[1] + i1) * ngrid[0];
            double r = 0.0;
            for (int i0 = 0; i0 < ngrid[0]; i0++) r += wgt(i0, ngrid[0]) * row[i0];
            s += w12 * r;
        }
    double sc = 1.0;
    for (int a = 0; a < nd; a++) sc *= dz[a] / 3.0;
    return s * sc;
}

// 2D and 3D grids take the spectral path (a 2D grid is a 3D grid with one z plane).
inline bool spectral(const cdft_sys *s) { return s->nd >= 2; }

// int f dV over the structure (integrals.jl:20-40): Simpson on Cartesian grids, the
// radial transform's weights (4 pi / 2 pi integrateR) on radial ones.
double integ_sys(const cdft_sys *s, const double *f)
{
    if (s->coord) {
        double acc = 0.0;
        for (int i = 0; i < s->N; i++) acc += s->rw[i] * f[i];
        return acc;
    }
    double dz[3];
    for (int a = 0; a < s->nd; a++) dz[a] = (s->ub[a] - s->lb[a]) / s->ngrid[a];   // structure_dz
    return simpson_nd(f, s->nd, s->ngrid, dz);
}

// Arguments of k_update: per-field parts (1D) or the finished transposed sum (3D).
struct UpdSrc {
    const double *dphi, *part;
    int F;
};
UpdSrc upd_src(const cdft_sys *s)
{
    if (spectral(s)) return {s->part, nullptr, 1};
    return {s->dphi, s->part, s->F};
}

// ---- 3D passes: one line FFT along axis AX over tiles of TB lines (TB: the
// system's tuned lines per tile for precision T)
template <typename T, int AX, int IN, int OUT, int EPT, int TB>
void fft3_launch(const cdft_sys *s, const Pass3 &p)
{
    const Grid3 &G = s->g3;
    const int L = G.n[AX];
    const size_t lds = (size_t)L * (TB + 1) * sizeof(cx<T>);   // tile, twiddles
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_fft3<T, AX, IN, OUT, EPT, TB>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    long nblk;
    if (AX == 0) nblk = (G.N / G.n[0]) / TB;
    else if (AX == 1) nblk = (long)(G.n[0] / TB) * G.n[2];
    else nblk = (long)(G.n[0] / TB) * G.n[1];
    hipLaunchKernelGGL((k_fft3<T, AX, IN, OUT, EPT, TB>), dim3((unsigned)nblk, std::max(p.ncomp, 1)), dim3(256), lds,
                       s->ctx->stream, G, p);
}

// EPT = L TB / 1024 radix-4 butterflies per thread (0: short lines, L TB <= 512,
// with guards); the tuner keeps L TB <= 4096.
template <typename T, int AX, int IN, int OUT, int TB>
void fft3_launch_tb(const cdft_sys *s, const Pass3 &p)
{
    switch (s->g3.n[AX] * TB / 1024) {
    case 4: fft3_launch<T, AX, IN, OUT, 4, TB>(s, p); break;
    case 2: fft3_launch<T, AX, IN, OUT, 2, TB>(s, p); break;
    case 1: fft3_launch<T, AX, IN, OUT, 1, TB>(s, p); break;
    default: fft3_launch<T, AX, IN, OUT, 0, TB>(s, p); break;
    }
}

template <typename T>
inline int tuned_tb(const cdft_sys *s) { return s->tune.lines_per_tile[sizeof(T) == sizeof(float) ? 1 : 0]; }

template <typename T, int AX, int IN, int OUT>
void fft3_pass(const cdft_sys *s, Pass3 p, bool inverse)
{
    p.inverse = inverse ? 1 : 0;
    p.gate = s->gate;
    p.tw = reinterpret_cast<const cx<double> *>(s->ctx->twiddles.at(s->g3.n[AX]));
    switch (tuned_tb<T>(s)) {
    case 4: fft3_launch_tb<T, AX, IN, OUT, 4>(s, p); break;
    case 16: fft3_launch_tb<T, AX, IN, OUT, 16>(s, p); break;
    default: fft3_launch_tb<T, AX, IN, OUT, 8>(s, p); break;
    }
}

// Field pairs of the 3D slots (PCSAFT.jl:55-61): (S, V), (v_x, v_y), (v_z, zeta), (lambda, disp).
constexpr int kPair3[4][2] = {{1, 2}, {3, 4}, {5, 6}, {7, 8}};
inline bool is_vec3(int f) { return f >= 3 && f <= 5; }

template <typename T>
Pass3 pair_pass(const cdft_sys *s, int c, int q)
{
    Pass3 p;
    std::memset(&p, 0, sizeof(p));
    const int f1 = kPair3[q][0], f2 = kPair3[q][1];
    const size_t o1 = (size_t)(c * 8 + f1 - 1) * s->N, o2 = (size_t)(c * 8 + f2 - 1) * s->N;
    if (sizeof(T) == sizeof(double)) {
        p.g1 = s->G3 + o1;
        p.g2 = s->G3 + o2;
    } else {
        p.g1 = s->G3f + o1;
        p.g2 = s->G3f + o2;
    }
    p.v1 = is_vec3(f1);
    p.v2 = is_vec3(f2);
    return p;
}

// The transposed all-pairs z pass (k_fft3z_accum) of component c: spectra and the
// pairs' work arrays.
template <typename T>
ZPairs3 zpairs_for(const cdft_sys *s, int c)
{
    ZPairs3 p;
    std::memset(&p, 0, sizeof(p));
    double *bq[4] = {s->B3, s->B3x[0], s->B3x[1], s->B3x[2]};
    p.np = 4;
    for (int q = 0; q < 4; q++) {
        const Pass3 pq = pair_pass<T>(s, c, q);
        p.g1[q] = pq.g1;
        p.g2[q] = pq.g2;
        p.v1[q] = pq.v1;
        p.v2[q] = pq.v2;
        p.b[q] = bq[q];
    }
    p.tw = reinterpret_cast<const cx<double> *>(s->ctx->twiddles.at(s->g3.n[2]));
    return p;
}

template <typename T, int EPT>
void zpairs_launch_ept(const cdft_sys *s, const ZPairs3 &p)
{
    const Grid3 &G = s->g3;
    const size_t lds = (size_t)G.n[2] * (2 * TBZ + 1) * sizeof(cx<T>);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_fft3z_accum<T, EPT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    const long nblk = (long)(G.n[0] / TBZ) * G.n[1];
    ZPairs3 pg = p;
    pg.gate = s->gate;
    hipLaunchKernelGGL((k_fft3z_accum<T, EPT>), dim3((unsigned)nblk), dim3(256), lds, s->ctx->stream, G, pg);
}

template <typename T>
void zpairs_launch(const cdft_sys *s, const ZPairs3 &p)
{
    if (s->g3.n[2] == 256) zpairs_launch_ept<T, 2>(s, p);
    else zpairs_launch_ept<T, 1>(s, p);                           // 128
}

// Weighted densities of src (FP64, layout [c N + k]) into nout[(c F + f) N + k],
// f >= 1, computed in T.  Every pass takes all components in one launch
// (blockIdx.y = component; A3 and B3 hold one slice per component).
template <typename T>
void launch_weighted3d(cdft_sys *s, const double *src, T *nout)
{
    const int N = s->N, F = s->F, nc = s->nc;
    const bool plane = (s->g3.n[2] == 1);            // a 2D grid: no z passes
    const long cplx = (long)N * sizeof(cx<T>), field = (long)F * N * sizeof(T);
    Pass3 p;
    std::memset(&p, 0, sizeof(p));
    p.ncomp = nc;
    p.in = src;
    p.cs[0] = (long)N * sizeof(double);
    p.out = s->A3;
    p.cs[2] = cplx;
    fft3_pass<T, 0, IN_RHO, OUT_CPLX>(s, p, false);
    p.in = s->A3;
    p.cs[0] = cplx;
    fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, p, false);
    if (!plane) fft3_pass<T, 2, IN_CPLX, OUT_CPLX>(s, p, false);
    for (int q = 0; q < 4; q++) {
        Pass3 pq = pair_pass<T>(s, 0, q);
        pq.ncomp = nc;
        pq.cs[5] = pq.cs[6] = (long)8 * N * sizeof(T);   // spectra: 8 slots per component
        pq.in = s->A3;
        pq.cs[0] = cplx;
        pq.out = s->B3;
        pq.cs[2] = cplx;
        if (plane) {
            fft3_pass<T, 1, IN_SPEC, OUT_CPLX>(s, pq, true);
        } else {
            fft3_pass<T, 2, IN_SPEC, OUT_CPLX>(s, pq, true);
            pq.in = s->B3;
            fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, pq, true);
        }
        pq.in = s->B3;
        pq.out = nout + (size_t)kPair3[q][0] * N;
        pq.out2 = nout + (size_t)kPair3[q][1] * N;
        pq.cs[2] = pq.cs[3] = field;
        pq.scale = s->NA / (double)N;
        fft3_pass<T, 0, IN_CPLX, OUT_FIELDS>(s, pq, true);
    }
}

// dtot[c N + k] = dphi[(c F) N + k] + sum_f (transposed convolution of dphi_f), in T.
// Every pass takes all components in one launch (blockIdx.y = component; B3 and S3
// hold one slice per component).  With the extra work arrays (the tuner's fused_z)
// the four accumulating z passes and the first inverse z pass of a component are one
// all-pairs launch instead, component by component.
template <typename T>
void launch_transposed3d(cdft_sys *s, const T *dphi, double *dtot)
{
    const int N = s->N, F = s->F, nc = s->nc;
    const bool plane = (s->g3.n[2] == 1);            // a 2D grid: no z passes
    if (s->B3x[2] && s->tune.fused_z[sizeof(T) == sizeof(float) ? 1 : 0]) {
        for (int c = 0; c < nc; c++) {
            ZPairs3 zp = zpairs_for<T>(s, c);
            double *bq[4] = {s->B3, s->B3x[0], s->B3x[1], s->B3x[2]};
            for (int q = 0; q < 4; q++) {
                Pass3 pq = pair_pass<T>(s, c, q);
                pq.in = dphi + (size_t)(c * F + kPair3[q][0]) * N;
                pq.in2 = dphi + (size_t)(c * F + kPair3[q][1]) * N;
                pq.out = bq[q];
                fft3_pass<T, 0, IN_PAIR, OUT_CPLX>(s, pq, false);
                pq.in = bq[q];
                fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, pq, false);
            }
            zp.out = s->S3;
            zpairs_launch<T>(s, zp);
            Pass3 p;
            std::memset(&p, 0, sizeof(p));
            p.in = s->S3;
            p.out = s->S3;
            fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, p, true);
            p.out = dtot + (size_t)c * N;
            p.base = dphi + (size_t)(c * F) * N;
            p.scale = 1.0 / (double)N;
            fft3_pass<T, 0, IN_CPLX, OUT_REAL>(s, p, true);
        }
        return;
    }
    const long cplx = (long)N * sizeof(cx<T>), field = (long)F * N * sizeof(T);
    for (int q = 0; q < 4; q++) {
        Pass3 pq = pair_pass<T>(s, 0, q);
        pq.ncomp = nc;
        pq.cs[5] = pq.cs[6] = (long)8 * N * sizeof(T);   // spectra: 8 slots per component
        pq.in = dphi + (size_t)kPair3[q][0] * N;
        pq.in2 = dphi + (size_t)kPair3[q][1] * N;
        pq.cs[0] = pq.cs[1] = field;
        pq.out = s->B3;
        pq.cs[2] = cplx;
        fft3_pass<T, 0, IN_PAIR, OUT_CPLX>(s, pq, false);
        pq.in = s->B3;
        pq.in2 = nullptr;
        pq.cs[0] = cplx;
        pq.cs[1] = 0;
        pq.first = (q == 0);
        if (plane) {
            pq.out = s->S3;
            fft3_pass<T, 1, IN_CPLX, OUT_ACCUM>(s, pq, false);
        } else {
            fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, pq, false);
            pq.out = s->S3;
            fft3_pass<T, 2, IN_CPLX, OUT_ACCUM>(s, pq, false);
        }
    }
    Pass3 p;
    std::memset(&p, 0, sizeof(p));
    p.ncomp = nc;
    p.in = s->S3;
    p.out = s->S3;
    p.cs[0] = p.cs[2] = cplx;
    if (!plane) fft3_pass<T, 2, IN_CPLX, OUT_CPLX>(s, p, true);
    fft3_pass<T, 1, IN_CPLX, OUT_CPLX>(s, p, true);
    p.out = dtot;
    p.cs[2] = (long)N * sizeof(double);
    p.base = dphi;
    p.cs[4] = field;
    p.scale = 1.0 / (double)N;
    fft3_pass<T, 0, IN_CPLX, OUT_REAL>(s, p, true);
}

template <int NC, int AS>
void launch_phi32_as(cdft_sys *s)
{
    if (spectral(s))
        hipLaunchKernelGGL((k_phi32<NC, 3, AS>), dim3(nblocks(s->N)), dim3(256), 0, s->ctx->stream, s->rho, s->n32,
                           s->dphi32, s->N, s->NA, s->pc, s->gate, (double *)nullptr);
    else
        hipLaunchKernelGGL((k_phi32<NC, 1, AS>), dim3(nblocks(s->N)), dim3(256), 0, s->ctx->stream, s->rho, s->n32,
                           s->dphi32, s->N, s->NA, s->pc, s->gate, s->dphi);
}

template <int NC>
void launch_phi32(cdft_sys *s)
{
    switch (s->assoc) {
    case 0: launch_phi32_as<NC, 0>(s); break;
    case 1: launch_phi32_as<NC, 1>(s); break;
    default: launch_phi32_as<NC, 2>(s); break;
    }
}

// The FP32 map of a mixed-precision system: weighted densities, Phi's gradient
// and transposed convolutions in FP32 (n3-singular factors in FP64); the sums
// land in FP64 (s->part) for the exponent assembly.
void launch_functional32(cdft_sys *s)
{
    if (s->nd == 1) {
        // 1D: compensated FP32 circulant sums; the transposed sums land in FP64 in
        // s->part and the rho slot of the gradient in s->dphi, where k_update reads them
        const int N = s->N, F = s->F;
        const dim3 grid((N + CIRC_TW - 1) / CIRC_TW, s->nc * (F - 1));
        hipLaunchKernelGGL((k_circ1d<float, double, float, true>), grid, dim3(256), 0, s->ctx->stream,
                           (const double *)s->rho, (long)N, 0L, (const float *)s->h32, s->n32, (long)F * N, (long)N, 0,
                           N, F, (float)s->NA, 0, s->gate);
        switch (s->nc) {
        case 1: launch_phi32<1>(s); break;
        case 2: launch_phi32<2>(s); break;
        case 3: launch_phi32<3>(s); break;
        default: launch_phi32<4>(s); break;
        }
        hipLaunchKernelGGL((k_circ1d<float, float, double, true>), grid, dim3(256), 0, s->ctx->stream,
                           (const float *)s->dphi32, (long)F * N, (long)N, (const float *)s->h32, s->part,
                           (long)(F - 1) * N, (long)N, 1, N, F, 1.0f, 1, s->gate);
        return;
    }
    launch_weighted3d<float>(s, s->rho, s->n32);
    switch (s->nc) {
    case 1: launch_phi32<1>(s); break;
    case 2: launch_phi32<2>(s); break;
    case 3: launch_phi32<3>(s); break;
    default: launch_phi32<4>(s); break;
    }
    launch_transposed3d<float>(s, s->dphi32, s->part);
}

template <int NC, int AS>
void launch_phi_as(cdft_sys *s, hipStream_t st)
{
    if (spectral(s))
        hipLaunchKernelGGL((k_phi<NC, 3, AS>), dim3(nblocks(s->N)), dim3(256), 0, st, s->rho, s->n, s->dphi, s->fval,
                           s->N, s->NA, s->pc, s->gate);
    else
        hipLaunchKernelGGL((k_phi<NC, 1, AS>), dim3(nblocks(s->N)), dim3(256), 0, st, s->rho, s->n, s->dphi, s->fval,
                           s->N, s->NA, s->pc, s->gate);
}

// s->assoc selects the association variant: 0 none, 1 one pair (closed form), 2 general.
template <int NC>
void launch_phi(cdft_sys *s, hipStream_t st)
{
    switch (s->assoc) {
    case 0: launch_phi_as<NC, 0>(s, st); break;
    case 1: launch_phi_as<NC, 1>(s, st); break;
    default: launch_phi_as<NC, 2>(s, st); break;
    }
}

void launch_phi_any(cdft_sys *s)
{
    hipStream_t st = s->ctx->stream;
    switch (s->nc) {
    case 1: launch_phi<1>(s, st); break;
    case 2: launch_phi<2>(s, st); break;
    case 3: launch_phi<3>(s, st); break;
    default: launch_phi<4>(s, st); break;
    }
}

// Transposed convolutions: src = dphi[(c*F+f)*N], dst = part[(c*(F-1)+f-1)*N], minus
// on the vector slot.
void launch_transposed(cdft_sys *s)
{
    if (spectral(s)) {
        launch_transposed3d<double>(s, s->dphi, s->part);
        return;
    }
    int N = s->N, F = s->F;
    if (s->coord) {        // the scalar slots' operators themselves, the vector slot's transpose
        hipLaunchKernelGGL(k_dense1d, dim3((N + 63) / 64, s->nc * (F - 1)), dim3(256), 0, s->ctx->stream,
                           (const double *)s->dphi, (long)F * N, (long)N, (const double *)s->h, (const double *)s->hT,
                           s->part, (long)(F - 1) * N, (long)N, 1, N, F, 1.0, s->gate);
        return;
    }
    hipLaunchKernelGGL(k_circ1d<double>, dim3((N + CIRC_TW - 1) / CIRC_TW, s->nc * (F - 1)), dim3(256), 0, s->ctx->stream, s->dphi,
                       (long)F * N, (long)N, s->h, s->part, (long)(F - 1) * N, (long)N, 1, N, F, 1.0, 1, s->gate);
}

// Weighted densities of rho (s->rho) into s->n.
void launch_weighted(cdft_sys *s)
{
    if (spectral(s)) {
        launch_weighted3d<double>(s, s->rho, s->n);
        return;
    }
    int N = s->N, F = s->F;
    if (s->coord) {
        hipLaunchKernelGGL(k_dense1d, dim3((N + 63) / 64, s->nc * (F - 1)), dim3(256), 0, s->ctx->stream,
                           (const double *)s->rho, (long)N, 0L, (const double *)s->h, (const double *)nullptr, s->n,
                           (long)F * N, (long)N, 0, N, F, s->NA, s->gate);
        return;
    }
    hipLaunchKernelGGL(k_circ1d<double>, dim3((N + CIRC_TW - 1) / CIRC_TW, s->nc * (F - 1)), dim3(256), 0, s->ctx->stream, s->rho,
                       (long)N, 0L, s->h, s->n, (long)F * N, (long)N, 0, N, F, s->NA, 0, s->gate);
}

// Weighted densities, Phi + gradient and transposed convolutions into s->dphi / s->part.
void launch_functional(cdft_sys *s)
{
    launch_weighted(s);
    launch_phi_any(s);
    launch_transposed(s);
}

template <int NC, int AS>
void launch_phi_hvp_as(cdft_sys *s, const double *v, hipStream_t st)
{
    if (spectral(s))
        hipLaunchKernelGGL((k_phi_hvp<NC, 3, AS>), dim3(nblocks(s->N)), dim3(256), 0, st, s->rho, v, s->n, s->dphi,
                           s->N, s->NA, s->pc, s->gate);
    else
        hipLaunchKernelGGL((k_phi_hvp<NC, 1, AS>), dim3(nblocks(s->N)), dim3(256), 0, st, s->rho, v, s->n, s->dphi,
                           s->N, s->NA, s->pc, s->gate);
}

template <int NC>
void launch_phi_hvp(cdft_sys *s, const double *v, hipStream_t st)
{
    switch (s->assoc) {
    case 0: launch_phi_hvp_as<NC, 0>(s, v, st); break;
    case 1: launch_phi_hvp_as<NC, 1>(s, v, st); break;
    default: launch_phi_hvp_as<NC, 2>(s, v, st); break;
    }
}

// Hessian-vector product of F_res at s->rho (weighted densities already in s->n)
// along v: dphi <- H_Phi (N_A v, W v), part <- transposed convolutions of it.
void launch_hvp(cdft_sys *s, const double *v)
{
    hipStream_t st = s->ctx->stream;
    int N = s->N, F = s->F, nc = s->nc;
    if (spectral(s)) {
        launch_weighted3d<double>(s, v, s->dphi);
    } else if (s->coord) {
        hipLaunchKernelGGL(k_dense1d, dim3((N + 63) / 64, nc * (F - 1)), dim3(256), 0, st, v, (long)N, 0L,
                           (const double *)s->h, (const double *)nullptr, s->dphi, (long)F * N, (long)N, 0, N, F,
                           s->NA, s->gate);
    } else {
        dim3 g((N + CIRC_TW - 1) / CIRC_TW, nc * (F - 1));
        hipLaunchKernelGGL(k_circ1d<double>, g, dim3(256), 0, st, v, (long)N, 0L, s->h, s->dphi, (long)F * N,
                           (long)N, 0, N, F, s->NA, 0, s->gate);
    }
    switch (nc) {
    case 1: launch_phi_hvp<1>(s, v, st); break;
    case 2: launch_phi_hvp<2>(s, v, st); break;
    case 3: launch_phi_hvp<3>(s, v, st); break;
    default: launch_phi_hvp<4>(s, v, st); break;
    }
    launch_transposed(s);
}

// ---- radial grids (Uniform1DSphr / Uniform1DCyl, base/structure.jl:363-433)
// The quasi-discrete Hankel transform of Hankel.jl (Guizar-Sicairos and
// Gutierrez-Vega 2004), order 0: cylindrical = the 2D radial Fourier transform on
// the zeros j_n of J_0, spherical = the 3D one as the order-1/2 transform of
// sqrt(r) f on j_n = n pi.  r_n = j_n R / j_(N+1), k_n = j_n / R, and
// C_mn = 2 / j_(N+1) J_p(j_m j_n / j_(N+1)) / (|J_(p+1)(j_m)| |J_(p+1)(j_n)|).
static double radial_jp(bool sph, double x) { return sph ? std::sqrt(2.0 / (kPi * x)) * std::sin(x) : ::j0(x); }
static double radial_jp1(bool sph, double x)
{
    return sph ? std::sqrt(2.0 / (kPi * x)) * (std::sin(x) / x - std::cos(x)) : ::j1(x);
}
static double radial_zero(bool sph, int s)
{
    if (sph) return s * kPi;
    const double b = (s - 0.25) * kPi;                  // McMahon's expansion, then Newton
    double x = b + 1.0 / (8.0 * b) - 124.0 / (3.0 * std::pow(8.0 * b, 3));
    for (int it = 0; it < 50; it++) {
        const double dx = ::j0(x) / ::j1(x);
        x += dx;
        if (std::fabs(dx) < 1e-15 * x) break;
    }
    return x;
}

// The dense operators of a radial system: for every component and field slot the
// convolution Q^-1 diag(Omega_f) Q (the reference's convolve! with a QDHT,
// integrals.jl:127-137; scalar kernels at omega = k / 2 pi in reduced units without
// a k = 0 sample, weighted_densities.jl:92-116), column-major into s->h; the vector
// slot's T = D Q^-1 diag(H) Q with H in unscaled units and D the radial derivative
// matrix (weighted_densities.jl:290-316, transforms.jl:43-63) into slot 3 and its
// transpose into s->hT (the reference adds transpose(T) dPhi/dn_v, :330-343); the
// integration weights (integrals.jl:20-22) into s->rw.
static hipError_t radial_setup(cdft_sys *s)
{
    const int N = s->N, nc = s->nc, F = s->F;
    const bool sph = (s->coord == CDFT_SPHERICAL);
    const double R = s->ub[0];
    std::vector<double> z(N + 1), r(N), k(N), jp1(N);
    for (int q = 0; q <= N; q++) z[q] = radial_zero(sph, q + 1);
    const double S = z[N], K = S / R;
    for (int q = 0; q < N; q++) {
        r[q] = z[q] * R / S;
        k[q] = z[q] / R;
        jp1[q] = std::fabs(radial_jp1(sph, z[q]));
    }
    // forward A[m N + j] (row-major) and backward B[m N + i] = Bop[i][m] (column-major)
    std::vector<double> A((size_t)N * N), B((size_t)N * N);
    for (int m = 0; m < N; m++) {
        const double km = sph ? std::sqrt(kPi / (2.0 * k[m])) : 1.0;
        for (int j = 0; j < N; j++) {
            const double C = 2.0 / S * radial_jp(sph, z[m] * z[j] / S) / (jp1[m] * jp1[j]);
            A[(size_t)m * N + j] = C * (sph ? std::sqrt(r[j]) : 1.0) * R / jp1[j] * jp1[m] / K * km;
            B[(size_t)m * N + j] = C / km * K / jp1[m] * jp1[j] / R / (sph ? std::sqrt(r[j]) : 1.0);
        }
    }
    // multipliers: w[(c F + f) N + m]; slot 0 unused, slot 3 the unscaled H of the vector weight
    const double twopi = 2.0 * kPi, fourpi = 4.0 * kPi, tp3 = twopi * twopi * twopi;
    std::vector<double> w((size_t)nc * F * N, 0.0);
    for (int c = 0; c < nc; c++) {
        const double dred = s->pc.dd[c];
        for (int f = 1; f < F; f++) {
            const bool sl = (f == 1 || f == 5);                     // int rho dz (d/2), (d)
            const double width = f == 6 ? dred * s->psi : (f <= 3 ? 0.5 : 1.0) * dred;
            const double Rw = twopi * width;
            for (int m = 0; m < N; m++) {
                double v;
                if (f == 3) {
                    const double wb = k[m] / twopi, Rk = twopi * 0.5 * dred * s->L;
                    v = fourpi / (wb * wb * wb) * (std::sin(wb * Rk) - Rk * wb * std::cos(wb * Rk)) / tp3;
                } else {
                    const double wb = k[m] / twopi * s->L, wbR = wb * Rw;
                    v = sl ? 2.0 * std::sin(wbR) / wb / twopi
                           : fourpi / (wb * wb * wb) * (std::sin(wbR) - Rw * wb * std::cos(wbR)) / tp3;
                }
                w[((size_t)c * F + f) * N + m] = v;
            }
        }
    }
    double *dA = nullptr, *dB = nullptr, *dw = nullptr;
    hipError_t e;
    struct Free3 { double **p[3]; ~Free3() { for (auto q : p) if (*q) (void)hipFree(*q); } } fr{{&dA, &dB, &dw}};
    if ((e = hipMalloc((void **)&dA, sizeof(double) * A.size())) != hipSuccess) return e;
    if ((e = hipMalloc((void **)&dB, sizeof(double) * B.size())) != hipSuccess) return e;
    if ((e = hipMalloc((void **)&dw, sizeof(double) * w.size())) != hipSuccess) return e;
    if ((e = hipMemcpy(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(dB, B.data(), sizeof(double) * B.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(dw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_radial_op, dim3((N + 255) / 256, N, nc * F), dim3(256), 0, s->ctx->stream,
                       (const double *)dA, (const double *)dB, (const double *)dw, N, s->h);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s->ctx->stream)) != hipSuccess) return e;
    // T = D M, M the slot-3 operator; D has three entries per row
    std::vector<double> M((size_t)N * N), T((size_t)N * N), TT((size_t)N * N);
    std::vector<double> D((size_t)N * 3), Dc((size_t)N);       // row i: columns Dc[i] .. Dc[i] + 2
    for (int i = 0; i < N; i++) {
        const int i0 = i == 0 ? 0 : (i == N - 1 ? N - 3 : i - 1);
        Dc[i] = i0;
        double *d = &D[(size_t)i * 3];
        if (i == 0) {
            const double h1 = r[1] - r[0], h2 = r[2] - r[1];
            d[0] = -(2 * h1 + h2) / (h1 * (h1 + h2));
            d[1] = (h1 + h2) / (h1 * h2);
            d[2] = -h1 / (h2 * (h1 + h2));
        } else if (i == N - 1) {
            const double h1 = r[N - 2] - r[N - 3], h2 = r[N - 1] - r[N - 2];
            d[0] = h2 / (h1 * (h1 + h2));
            d[1] = -(h1 + h2) / (h1 * h2);
            d[2] = (2 * h2 + h1) / (h2 * (h1 + h2));
        } else {
            const double h1 = r[i] - r[i - 1], h2 = r[i + 1] - r[i];
            d[0] = -h2 / (h1 * (h1 + h2));
            d[1] = (h2 - h1) / (h1 * h2);
            d[2] = h1 / (h2 * (h1 + h2));
        }
    }
    for (int c = 0; c < nc; c++) {
        double *slot = s->h + ((size_t)c * F + 3) * N * N;
        if ((e = hipMemcpy(M.data(), slot, sizeof(double) * M.size(), hipMemcpyDeviceToHost)) != hipSuccess) return e;
        for (int i = 0; i < N; i++)                     // M column-major: M[j N + l] = M_lj
            for (int j = 0; j < N; j++) {
                const double *d = &D[(size_t)i * 3];
                const int l0 = (int)Dc[i];
                double acc = 0.0;
                for (int q = 0; q < 3; q++) acc += d[q] * M[(size_t)j * N + l0 + q];
                T[(size_t)j * N + i] = acc;              // column-major T
                TT[(size_t)i * N + j] = acc;             // column-major T^T
            }
        if ((e = hipMemcpy(slot, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
        if ((e = hipMemcpy(s->hT + (size_t)c * N * N, TT.data(), sizeof(double) * TT.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return e;
    }
    s->rw.resize(N);
    for (int q = 0; q < N; q++) {
        const double wq = 2.0 / (K * K * jp1[q] * jp1[q]);       // int_0^R g r dr quadrature
        s->rw[q] = sph ? 4.0 * kPi * wq * r[q] : 2.0 * kPi * wq;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

const char *cdft_last_error(const cdft_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int cdft_ctx_create(int device, cdft_ctx **out)
{
    if (!out) return fail(nullptr, CDFT_ERR_ARG, "cdft_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(nullptr, CDFT_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, CDFT_ERR_ARG, "device index out of range");
    if (device >= kMaxDevices) return fail(nullptr, CDFT_ERR_UNSUPPORTED, "device index beyond the library's table");
    auto *c = new cdft_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        cdft_ctx_destroy(c);
        return fail(nullptr, CDFT_ERR_HIP, "stream/event creation failed");
    }
    {
        // this device's copy of the batched scratch array (device is current)
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_dev[device]) {
            void *p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                cdft_ctx_destroy(c);
                return fail(nullptr, CDFT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            }
            g_scratch_dev[device] = p;
        }
        c->scratch = g_scratch_dev[device];
    }
    c->bstage.resize(sizeof(BSys) * kNoAllocMaxSys);
    c->rstage.resize(sizeof(BResult) * kNoAllocMaxSys);
    c->border.resize(kNoAllocMaxSys);
    *out = c;
    return CDFT_OK;
}

int cdft_ctx_destroy(cdft_ctx *ctx)
{
    if (!ctx) return CDFT_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (auto &kv : ctx->twiddles) (void)hipFree(kv.second);
    delete ctx;
    return CDFT_OK;
}

double cdft_last_kernel_ms(const cdft_ctx *ctx) { return ctx ? ctx->last_kernel_ms : 0.0; }

int cdft_system_npoints(const cdft_sys *sys) { return sys ? sys->N : 0; }

static int batched_prepare(cdft_ctx *ctx, cdft_sys *s);

int cdft_system_create(cdft_ctx *ctx, const cdft_system_desc *d, cdft_sys **out)
{
    if (!ctx || !d || !out) return fail(ctx, CDFT_ERR_ARG, "cdft_system_create: NULL argument");
    *out = nullptr;
    if (d->nc < 1 || d->nc > 4) return fail(ctx, CDFT_ERR_UNSUPPORTED, "nc must be in 1..4");
    if (d->ndim < 1 || d->ndim > 3) return fail(ctx, CDFT_ERR_UNSUPPORTED, "only 1D, 2D and 3D Cartesian grids are supported");
    if (d->ndim == 1 && d->ngrid[0] < 3) return fail(ctx, CDFT_ERR_ARG, "ngrid must be >= 3");
    if (d->coord != CDFT_CARTESIAN) {
        if (d->coord != CDFT_SPHERICAL && d->coord != CDFT_CYLINDRICAL)
            return fail(ctx, CDFT_ERR_ARG, "unknown coordinate system");
        if (d->ndim != 1) return fail(ctx, CDFT_ERR_ARG, "radial grids are one-dimensional");
        if (d->ngrid[0] > CDFT_RADIAL_MAXN)
            return fail(ctx, CDFT_ERR_UNSUPPORTED, "radial grids of at most CDFT_RADIAL_MAXN points");
        if (!(d->lb[0] >= 0.0)) return fail(ctx, CDFT_ERR_ARG, "radial bounds need 0 <= lb < ub");
        if (d->precision != CDFT_FP64)
            return fail(ctx, CDFT_ERR_UNSUPPORTED, "radial grids run in FP64 only");
    }
    if (d->ndim >= 2) {
        long N3 = 1;
        for (int a = 0; a < d->ndim; a++) {
            const int g = d->ngrid[a];
            if (g < 16 || g > 512 || (g & (g - 1)))
                return fail(ctx, CDFT_ERR_UNSUPPORTED, "2D and 3D grids need power-of-two axes of 16..512 points");
            N3 *= g;
        }
        if (N3 > (1L << 27)) return fail(ctx, CDFT_ERR_UNSUPPORTED, "grid larger than 2^27 points");
    }
    for (int a = 0; a < d->ndim; a++)
        if (!(d->ub[a] > d->lb[a])) return fail(ctx, CDFT_ERR_ARG, "bounds must satisfy ub > lb");
    if (!(d->T > 0) || !(d->L > 0)) return fail(ctx, CDFT_ERR_ARG, "T and L must be positive");
    if (!d->m || !d->d || !d->sigma || !d->eps || !d->rho_bulk || !d->mu_res)
        return fail(ctx, CDFT_ERR_ARG, "parameter array is NULL");
    if (d->assoc_npair < 0 || d->assoc_nsite < 0) return fail(ctx, CDFT_ERR_ARG, "negative association counts");
    if (d->assoc_npair > 0) {
        if (d->assoc_nsite > CDFT_MAXSITE || d->assoc_npair > CDFT_MAXPAIR)
            return fail(ctx, CDFT_ERR_UNSUPPORTED, "association: at most 4 sites and 8 pairs on the device");
        if (!d->assoc_site_comp || !d->assoc_site_count || !d->assoc_pair_site || !d->assoc_pair_eps ||
            !d->assoc_pair_kappa)
            return fail(ctx, CDFT_ERR_ARG, "association array is NULL");
        for (int t = 0; t < d->assoc_nsite; t++)
            if (d->assoc_site_comp[t] < 0 || d->assoc_site_comp[t] >= d->nc || !(d->assoc_site_count[t] > 0))
                return fail(ctx, CDFT_ERR_ARG, "association site with a bad component or count");
        for (int q = 0; q < 2 * d->assoc_npair; q++)
            if (d->assoc_pair_site[q] < 0 || d->assoc_pair_site[q] >= d->assoc_nsite)
                return fail(ctx, CDFT_ERR_ARG, "association pair names a site out of range");
    }
    if (d->precision != CDFT_FP64 && d->precision != CDFT_MIXED)
        return fail(ctx, CDFT_ERR_ARG, "unknown precision mode");
    for (int i = 0; i < d->nc; i++)
        if (!(d->rho_bulk[i] > 0) || !(d->d[i] > 0) || !(d->m[i] > 0))
            return fail(ctx, CDFT_ERR_ARG, "rho_bulk, d and m must be positive");

    auto *s = new cdft_sys;
    s->ctx = ctx;
    s->nc = d->nc;
    s->nd = d->ndim;
    s->N = 1;
    for (int a = 0; a < d->ndim; a++) {
        s->ngrid[a] = d->ngrid[a];
        s->lb[a] = d->lb[a];
        s->ub[a] = d->ub[a];
        s->N *= d->ngrid[a];
    }
    // a 2D grid runs as a 3D grid with one z plane: its third vector slot has a zero
    // weight (the z frequencies are all 0), so its weighted density is zero and Phi
    // is that of the two-component vector weight (PCSAFT.jl:55-61 with D = 2)
    if (d->ndim == 2) {
        s->ngrid[2] = 1;
        s->lb[2] = 0.0;
        s->ub[2] = 1.0;
    }
    s->F = d->ndim == 1 ? 7 : 9;
    s->T = d->T;
    s->p = d->p;
    s->L = d->L;
    s->psi = d->psi > 0 ? d->psi : 1.3862;
    s->NA = 6.02214076e23 * (s->L * s->L * s->L);   // N_A * density_scale^3 (weighted_densities.jl:123)
    s->precision = d->precision;
    s->coord = d->coord;
    int nc = s->nc;
    for (int i = 0; i < nc; i++) {
        s->m[i] = d->m[i];
        s->d[i] = d->d[i];
        s->rho_bulk[i] = d->rho_bulk[i];
        s->mu_res[i] = d->mu_res[i];
        s->cp[i] = d->mu_res[i] + std::log(d->rho_bulk[i]);
    }
    for (int i = 0; i < nc * nc; i++) { s->sigma[i] = d->sigma[i]; s->eps[i] = d->eps[i]; }
    PhiConst &P = s->pc;
    std::memset(&P, 0, sizeof(P));
    P.nc = nc;
    P.T = s->T;
    P.psi = s->psi;
    for (int i = 0; i < nc; i++) { P.m[i] = s->m[i]; P.dd[i] = s->d[i] / s->L; }
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            double sr = s->sigma[i * nc + j] / s->L;
            P.cm[i * nc + j] = s->m[i] * s->m[j] * sr * sr * sr;
            P.eT[i * nc + j] = s->eps[i * nc + j] / s->T;
        }
    // FP32 resolution of the lambda slot (2 d rho_b for a uniform profile):
    // convolution rounding is ~1e-7 of the field's bulk scale
    for (int i = 0; i < nc; i++) P.lsl_noise[i] = 1e-7 * 2.0 * P.dd[i] * d->rho_bulk[i] * s->NA;
    if (d->assoc_npair > 0) {
        // pack_assoc_params (assoc.jl:18-57) in reduced units (PCSAFT.jl:244-246)
        P.nsite = d->assoc_nsite;
        P.npair = d->assoc_npair;
        for (int t = 0; t < P.nsite; t++) {
            P.site_comp[t] = d->assoc_site_comp[t];
            P.site_count[t] = d->assoc_site_count[t];
        }
        for (int q = 0; q < P.npair; q++) {
            const int a = d->assoc_pair_site[2 * q], b = d->assoc_pair_site[2 * q + 1];
            const int ic = P.site_comp[a], jc = P.site_comp[b];
            const double sr = s->sigma[ic * nc + jc] / s->L;
            P.pair_a[q] = a;
            P.pair_b[q] = b;
            P.pair_k[q] = sr * sr * sr * (std::exp(d->assoc_pair_eps[q] / s->T) - 1.0) * d->assoc_pair_kappa[q];
            P.pair_dij[q] = P.dd[ic] * P.dd[jc] / (P.dd[ic] + P.dd[jc]);
        }
        // one pair between two distinct sites has a closed form (assoc.jl:702-769)
        s->assoc = (P.npair == 1 && P.pair_a[0] != P.pair_b[0]) ? 1 : 2;
    }

    int N = s->N, F = s->F;
    long n = (long)N * nc;
    s->nblk = std::min(nblocks(n), 2048);   // reducing kernels: grid-stride over at most 2048 workgroups
    (void)hipSetDevice(ctx->device);
    auto alloc = [&](double **p, size_t count) { return hipMalloc((void **)p, sizeof(double) * count); };
    hipError_t e = hipSuccess;
    double *ddev = nullptr;
    const bool d3 = spectral(s);
    const size_t hsize = s->coord ? (size_t)nc * F * N * N : (size_t)nc * F * N;
    if ((!d3 && (e = alloc(&s->h, hsize)) != hipSuccess) || (s->coord && (e = alloc(&s->hT, (size_t)nc * N * N)) != hipSuccess) ||
        (e = alloc(&s->rho, n)) != hipSuccess ||
        (e = alloc(&s->n, (size_t)nc * F * N)) != hipSuccess || (e = alloc(&s->dphi, (size_t)nc * F * N)) != hipSuccess ||
        (e = alloc(&s->part, d3 ? (size_t)n : (size_t)nc * (F - 1) * N)) != hipSuccess ||
        (e = alloc(&s->fval, N)) != hipSuccess ||
        (d3 && ((e = alloc(&s->G3, (size_t)8 * nc * N)) != hipSuccess ||
                (e = alloc(&s->A3, (size_t)2 * nc * N)) != hipSuccess ||       // one slice per component
                (e = alloc(&s->B3, (size_t)2 * nc * N)) != hipSuccess ||
                (e = alloc(&s->S3, (size_t)2 * nc * N)) != hipSuccess)) ||
        (!d3 && d->precision == CDFT_MIXED &&
         ((e = hipMalloc((void **)&s->h32, sizeof(float) * (size_t)F * nc * N)) != hipSuccess ||
          (e = hipMalloc((void **)&s->n32, sizeof(float) * (size_t)F * nc * N)) != hipSuccess ||
          (e = hipMalloc((void **)&s->dphi32, sizeof(float) * (size_t)F * nc * N)) != hipSuccess ||
          (e = alloc(&s->xsafe, n)) != hipSuccess)) ||
        (d3 && d->precision == CDFT_MIXED &&
         ((e = hipMalloc((void **)&s->G3f, sizeof(float) * 8 * nc * (size_t)N)) != hipSuccess ||
          (e = hipMalloc((void **)&s->n32, sizeof(float) * (size_t)F * nc * N)) != hipSuccess ||
          (e = hipMalloc((void **)&s->dphi32, sizeof(float) * (size_t)F * nc * N)) != hipSuccess ||
          (e = alloc(&s->xsafe, n)) != hipSuccess)) ||
        (e = alloc(&s->out, n)) != hipSuccess || (e = alloc(&s->sol, n)) != hipSuccess ||
        (e = alloc(&s->gx, n)) != hipSuccess || (e = alloc(&s->df, n)) != hipSuccess ||
        (e = alloc(&s->dg, n)) != hipSuccess || (e = alloc(&s->res, n)) != hipSuccess ||
        (e = alloc(&s->Q, (size_t)CDFT_MAXM * n)) != hipSuccess ||
        (e = alloc(&s->Qalt, (size_t)CDFT_MAXM * n)) != hipSuccess ||
        (e = alloc(&s->DG, (size_t)CDFT_MAXM * n)) != hipSuccess ||
        (e = alloc(&s->red, (size_t)2 * CDFT_MAXM * s->nblk)) != hipSuccess ||
        (e = alloc(&s->fin, (size_t)4 * CDFT_NQ)) != hipSuccess || (e = alloc(&s->cpd, nc)) != hipSuccess ||
        (e = alloc(&ddev, nc)) != hipSuccess ||
        (e = hipMemcpy(s->cpd, s->cp, sizeof(double) * nc, hipMemcpyHostToDevice)) != hipSuccess) {
        free_sys_buffers(s);
        if (ddev) (void)hipFree(ddev);
        delete s;
        return fail(ctx, e == hipErrorOutOfMemory ? CDFT_ERR_NOMEM : CDFT_ERR_HIP,
                    std::string("device allocation failed: ") + hipGetErrorString(e));
    }
    if (d3 && s->ngrid[2] >= 128 && s->ngrid[2] <= 256) {
        // optional: without them the per-pair z passes are used
        for (int q = 0; q < 3; q++)
            if (alloc(&s->B3x[q], (size_t)2 * N) != hipSuccess) {
                for (int r = 0; r < 3; r++)
                    if (s->B3x[r]) { (void)hipFree(s->B3x[r]); s->B3x[r] = nullptr; }
                (void)hipGetLastError();
                break;
            }
    }
    for (int pr = 0; pr < 2; pr++) s->tune.fused_z[pr] = s->B3x[2] ? 1 : 0;
    if (d->vext) {
        s->has_vext = true;
        if ((e = alloc(&s->vext, n)) != hipSuccess ||
            (e = hipMemcpy(s->vext, d->vext, sizeof(double) * n, hipMemcpyHostToDevice)) != hipSuccess) {
            free_sys_buffers(s);
            (void)hipFree(ddev);
            delete s;
            return fail(ctx, CDFT_ERR_HIP, std::string("vext upload failed: ") + hipGetErrorString(e));
        }
    }
    // real-space kernels (1D) or tabulated spectra (3D) from the analytic weights
    double dd[CDFT_MAXC];
    for (int i = 0; i < nc; i++) dd[i] = P.dd[i];
    if (d3) {
        Grid3 &G = s->g3;
        std::memset(&G, 0, sizeof(G));
        G.N = N;
        for (int a = 0; a < 3; a++) {
            G.n[a] = s->ngrid[a];
            while ((1 << G.lg[a]) < G.n[a]) G.lg[a]++;
            G.mult[a] = s->L / (s->ub[a] - s->lb[a]);             // fftfreq(n, n/(ub-lb)) * L
            if (G.n[a] > 1 && !ctx->twiddles.count(G.n[a])) {
                double *t64 = nullptr;
                if ((e = alloc(&t64, (size_t)2 * G.n[a])) == hipSuccess) {
                    hipLaunchKernelGGL(k_setup_tw, dim3(nblocks(G.n[a])), dim3(256), 0, ctx->stream, G.n[a], t64);
                    ctx->twiddles[G.n[a]] = t64;
                }
            }
        }
        const double twopi = 2.0 * kPi;
        for (int c = 0; c < nc; c++) {
            G.R[c][0] = twopi * 0.5 * dd[c];
            G.R[c][1] = twopi * dd[c];
            G.R[c][2] = twopi * dd[c] * s->psi;
        }
        for (int c = 0; c < nc && e == hipSuccess; c++)
            for (int f = 1; f <= 8; f++)
                hipLaunchKernelGGL(k_setup_omega3, dim3(nblocks(N)), dim3(256), 0, ctx->stream, G, c, f,
                                   s->G3 + (size_t)(c * 8 + f - 1) * N,
                                   s->G3f ? s->G3f + (size_t)(c * 8 + f - 1) * N : (float *)nullptr);
        if (e == hipSuccess) e = hipGetLastError();
    } else if (s->coord) {
        e = radial_setup(s);
    } else {
        double mult = ((double)N / (s->ub[0] - s->lb[0])) / (double)N;   // fftfreq spacing
        e = hipMemcpy(ddev, dd, sizeof(double) * nc, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemsetAsync(s->h, 0, sizeof(double) * nc * F * N, ctx->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_setup_h1d, dim3(nblocks(N), nc * F), dim3(256), 0, ctx->stream, s->h, N, nc, ddev,
                               s->psi, mult, s->L);
            if (s->h32)
                hipLaunchKernelGGL(k_to_f32, dim3(nblocks((long)nc * F * N)), dim3(256), 0, ctx->stream,
                                   (const double *)s->h, s->h32, (long)nc * F * N);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(ddev);
    if (e != hipSuccess) {
        free_sys_buffers(s);
        delete s;
        return fail(ctx, CDFT_ERR_HIP, std::string("kernel set-up failed: ") + hipGetErrorString(e));
    }
    // a system the batched kernel can take gets its ring spectra now, so that the
    // allocation-free batched entry point finds it ready
    if (!d3 && !s->coord && s->nc == 1 && !s->assoc && s->N <= kRegNPT * BT && s->N >= 5) {
        int rc = batched_prepare(ctx, s);
        if (rc == CDFT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = CDFT_ERR_HIP;
        if (rc != CDFT_OK) {
            free_sys_buffers(s);
            delete s;
            return rc == CDFT_ERR_HIP ? fail(ctx, rc, "batched set-up failed: " + ctx->err) : rc;
        }
    }
    *out = s;
    return CDFT_OK;
}

int cdft_system_destroy(cdft_sys *s)
{
    if (!s) return CDFT_ERR_ARG;
    (void)hipSetDevice(s->ctx->device);
    free_sys_buffers(s);
    delete s;
    return CDFT_OK;
}

int cdft_functional_eval(cdft_sys *s, const double *rho, double *dfdrho, double *f_res, double *fval)
{
    if (!s || !rho || !dfdrho) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_functional_eval: NULL argument");
    cdft_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    long n = (long)s->N * s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    launch_functional(s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(dfdrho, s->out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    std::vector<double> fv;
    if (f_res || fval) {
        fv.resize(s->N);
        HIPCHECK_CTX(ctx, hipMemcpyAsync(fv.data(), s->fval, sizeof(double) * s->N, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    if (fval) std::memcpy(fval, fv.data(), sizeof(double) * s->N);
    if (f_res) *f_res = integ_sys(s, fv.data()) / (s->L * s->L * s->L);    // F_res, models.jl:86-91
    return CDFT_OK;
}

// The linear map of the fixed-point adjoint, lambda -> g - rho* (H (mask lambda)),
// H the Hessian of F_res at rho* (device buffers; rho* in s->rho, W rho* in s->n).
struct AdjMap {
    const double *g;
    const double *mask;
    double *vbuf;              // mask lambda, the direction of the next product
};

// Bodies per graph launch, i.e. per host check of the solver state (CDFT_AA_CHUNK,
// default 16).  Bodies enqueued after the end of a solve find every gate set and
// cost only their launches.
static int aa_chunk()
{
    const char *e = std::getenv("CDFT_AA_CHUNK");
    const int v = e ? std::atoi(e) : 16;
    return std::max(1, std::min(v, 256));
}

// CDFT_AA_GRAPH=0 enqueues the same bodies one launch at a time instead of
// replaying a captured graph (same kernels, same results).
static bool aa_use_graph()
{
    const char *e = std::getenv("CDFT_AA_GRAPH");
    return !(e && e[0] == '0');
}

// The device solver's state at the start of aasol (converge!'s keywords,
// converge.jl:3-13): first evaluation at the seed, then Picard, then Anderson.
static void aa_ctl_init(AACtl *hc, const cdft_solver_opts *o, bool mixed, int hcap, bool adjoint, int pause_every)
{
    std::memset(hc, 0, sizeof(AACtl));
    hc->m = o->m;
    hc->maxit = o->maxit;
    hc->picard_maxit = o->maxit;
    hc->hist_cap = hcap;
    hc->adjoint = adjoint ? 1 : 0;
    hc->rtol = hc->atol = o->tol;
    hc->picard_rtol = hc->picard_atol = o->anderson_start;
    hc->beta = hc->picard_beta = o->beta;
    hc->fp32 = mixed ? 1 : 0;
    hc->phase = AA_FIRST;
    hc->umode = 1;
    hc->beta_u = o->beta;
    hc->g_proj = 1;
    hc->g_f64 = hc->fp32;
    hc->g_f32 = !hc->fp32;
    hc->evfp32 = hc->fp32;
    hc->pause_every = pause_every;
    for (int j = 0; j < CDFT_MAXM; j++) hc->ring[j] = j;
}

// The single-workgroup body (k_small_body) serves small 1D Cartesian FP64 solves of
// converge!; CDFT_SMALL_BODY=0 keeps them on the multi-kernel body.
static bool small_body_ok(const cdft_sys *s, bool adjoint, bool mixed)
{
    if (adjoint || s->nd != 1 || s->coord || s->nc > 2 || s->F != SB_F || (long)s->N * s->nc > SB_MAXN)
        return false;
    if (mixed && !s->h32) return false;
    const char *e = std::getenv("CDFT_SMALL_BODY");
    return !(e && e[0] == '0');
}

extern "C++" {
// Dynamic LDS above the default needs the kernel's attribute raised first (to what
// this system needs; the static LDS comes on top of it).
template <int NC, int AS, bool MIXED>
static hipError_t small_body_launch(const cdft_sys *s, const SBArgs &a, int nbody)
{
    static size_t attr = 0;
    const size_t lds = sb_lds_bytes(s->N, s->nc, MIXED);
    if (lds > attr) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_small_body<NC, AS, MIXED>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = lds;
    }
    hipLaunchKernelGGL((k_small_body<NC, AS, MIXED>), dim3(1), dim3(SB_T), lds, s->ctx->stream, a, s->pc, nbody);
    return hipSuccess;
}

template <bool MIXED>
static hipError_t small_body_m(const cdft_sys *s, const SBArgs &a, int nbody)
{
    const int as = s->assoc < 2 ? s->assoc : 2;
    if (s->nc == 1)
        return as == 0 ? small_body_launch<1, 0, MIXED>(s, a, nbody)
                       : (as == 1 ? small_body_launch<1, 1, MIXED>(s, a, nbody)
                                  : small_body_launch<1, 2, MIXED>(s, a, nbody));
    return as == 0 ? small_body_launch<2, 0, MIXED>(s, a, nbody)
                   : (as == 1 ? small_body_launch<2, 1, MIXED>(s, a, nbody)
                              : small_body_launch<2, 2, MIXED>(s, a, nbody));
}
}  // extern "C++"

static SBArgs sb_args(const cdft_sys *s, AACtl *c, double *hist)
{
    return SBArgs{c,      s->h,   s->vext, s->cpd, s->sol, s->gx,    s->res, s->df, s->dg, s->Q,
                  s->Qalt, s->DG, s->xsafe, s->rho, hist,  s->N,     s->nc,  s->NA,
                  s->precision == CDFT_MIXED ? s->h32 : nullptr};
}

static hipError_t small_body(const cdft_sys *s, AACtl *c, int nbody)
{
    const SBArgs a = sb_args(s, c, s->dhist);
    return s->precision == CDFT_MIXED ? small_body_m<true>(s, a, nbody) : small_body_m<false>(s, a, nbody);
}

// aasol (src/utils/anderson.jl:301-443) entirely on the device: every N-length
// operation and the m x m QR downdate and triangular solve run in kernels, the
// solver's control flow in one thread of the kernels that finish the reductions
// (k_aa_fin: aa_mid, aa_post, aa_pre), twelve launches per body in 1D (one, the
// single-workgroup k_small_body, for small 1D FP64 systems).  The host
// enqueues bodies of one map evaluation each — replayed from a hipGraph of
// aa_chunk() bodies captured once per system and kind of map — and reads the
// solver state once per chunk.  adj == nullptr: the Euler-Lagrange map of
// converge! on x = log rho; otherwise the adjoint map on x.
static int aasol_run(cdft_sys *s, double *rho_inout, const cdft_solver_opts *o, cdft_result *r, const AdjMap *adj)
{
    if (!s || !rho_inout || !o || !r) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_converge: NULL argument");
    cdft_ctx *ctx = s->ctx;
    const int m = o->m;
    const double beta = o->beta;
    if (m < 0 || m > CDFT_MAXM - 1) return fail(ctx, CDFT_ERR_ARG, "anderson m must be in 0..7");
    if (!(std::fabs(beta) > 0 && std::fabs(beta) <= 1)) return fail(ctx, CDFT_ERR_ARG, "abs(beta) must be in (0,1]");
    if (o->maxit < 0) return fail(ctx, CDFT_ERR_ARG, "maxit must be >= 0");
    // the progress hook belongs to converge! (the adjoint's linear solve has none)
    const bool prog = (!adj && o->progress);
    if (prog && o->progress_every < 1) return fail(ctx, CDFT_ERR_ARG, "progress_every must be >= 1");
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    const int nb = s->nblk;                      // reducing kernels (grid-stride)
    const int ne = nblocks(n);                   // elementwise kernels
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    if (!s->ctl) HIPCHECK_CTX(ctx, hipMalloc((void **)&s->ctl, sizeof(AACtl)));
    const int hcap = (o->hist && o->hist_cap > 0) ? o->hist_cap : 0;
    if (hcap > s->dhist_cap) {
        if (s->dhist) HIPCHECK_CTX(ctx, hipFree(s->dhist));
        s->dhist = nullptr;
        s->dhist_cap = 0;
        HIPCHECK_CTX(ctx, hipMalloc((void **)&s->dhist, sizeof(double) * hcap));
        s->dhist_cap = hcap;
    }
    AACtl *hc = nullptr;
    HIPCHECK_CTX(ctx, hipHostMalloc((void **)&hc, sizeof(AACtl)));
    struct HostFree { AACtl *p; ~HostFree() { if (p) (void)hipHostFree(p); } } hf{hc};

    // Mixed precision: the FP32 map until the residual reaches the FP32 noise
    // floor (4 tol, or no halving in 40 iterations), then the FP64 map with a fresh
    // Anderson history; the reported residual and termination are the FP64 ones.
    const bool mixed = (!adj && s->precision == CDFT_MIXED);
    aa_ctl_init(hc, o, mixed, hcap, adj != nullptr, prog ? o->progress_every : 0);
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->ctl, hc, sizeof(AACtl), hipMemcpyHostToDevice, st));

    double *ptgt = adj ? adj->vbuf : s->rho;              // where a new point is expanded
    const double *pmask = adj ? adj->mask : nullptr;
    if (adj) {
        HIPCHECK_CTX(ctx, hipMemcpyAsync(s->sol, rho_inout, sizeof(double) * n, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_set_point, dim3(ne), dim3(256), 0, st, s->sol, s->sol, ptgt, (int)n, pmask);
    } else {
        HIPCHECK_CTX(ctx, hipMemcpyAsync(s->out, rho_inout, sizeof(double) * n, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_init_x, dim3(ne), dim3(256), 0, st, s->out, s->sol, s->rho, (int)n);
    }

    AACtl *c = s->ctl;
    const bool small = small_body_ok(s, adj != nullptr, mixed);
    // Mixed precision with multi-kernel bodies: the FP32 phase's bodies carry the FP32
    // map only and those after the switch the FP64 map only (no gated launches of the
    // other map); a body that needs the FP64 map while only FP32 ones are queued stops
    // the chunk (paused = 2) and the host continues with FP64 bodies.
    const bool split = mixed && !small;
    bool only32 = split && hc->fp32;
    if (only32) {
        hc->graph_map = 1;
        HIPCHECK_CTX(ctx, hipMemcpyAsync(s->ctl, hc, sizeof(AACtl), hipMemcpyHostToDevice, st));
    }
    hipError_t sb_err = hipSuccess;
    auto body = [&](bool map64, bool map32) {
        if (small) {                              // one body (the first, eager)
            const hipError_t e = small_body(s, c, 1);
            if (e != hipSuccess) sb_err = e;
            return;
        }
        hipLaunchKernelGGL(k_aa_proj1, dim3(nb), dim3(256), 0, st, (int)n, (const double *)s->Q, s->Qalt,
                           (const AACtl *)c, (const double *)s->df, (const double *)s->res, s->red, nb);
        hipLaunchKernelGGL(k_aa_fin<0>, dim3(2 * CDFT_MAXM), dim3(256), 0, st, (const double *)s->red, nb, s->fin, c, 0,
                           s->fin, s->dhist);
        hipLaunchKernelGGL(k_aa_proj2, dim3(nb), dim3(256), 0, st, (int)n, (const double *)s->Q,
                           (const double *)s->Qalt, (const AACtl *)c, s->df, (const double *)s->res,
                           (const double *)s->fin, 0, s->red, nb);
        hipLaunchKernelGGL(k_aa_fin<0>, dim3(CDFT_MAXM), dim3(256), 0, st, (const double *)s->red, nb,
                           s->fin + CDFT_NQ, c, 0, s->fin, s->dhist);
        hipLaunchKernelGGL(k_aa_proj2, dim3(nb), dim3(256), 0, st, (int)n, (const double *)s->Q,
                           (const double *)s->Qalt, (const AACtl *)c, s->df, (const double *)s->res,
                           (const double *)(s->fin + CDFT_NQ), 1, s->red, nb);
        hipLaunchKernelGGL(k_aa_fin<1>, dim3(2), dim3(256), 0, st, (const double *)s->red, nb, s->fin + 2 * CDFT_NQ, c, 1,
                           s->fin, s->dhist);
        hipLaunchKernelGGL(k_aa_point, dim3(ne), dim3(256), 0, st, (int)n, (const AACtl *)c, (const double *)s->df,
                           s->Q, s->Qalt, s->sol, (const double *)s->gx, (const double *)s->dg, s->DG, s->xsafe, ptgt,
                           pmask);
        if (map64) {
            s->gate = &c->g_f64;
            if (adj) launch_hvp(s, adj->vbuf);
            else launch_functional(s);
        }
        if (map32) {
            s->gate = &c->g_f32;
            launch_functional32(s);
        }
        s->gate = nullptr;
        hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part, s->vext, s->N, s->nc,
                           upd_src(s).F, s->cpd, s->out, s->sol, s->gx, s->res, s->df, s->dg, beta, 2, s->red, nb,
                           adj ? adj->g : nullptr, adj ? (const double *)s->rho : nullptr, (const AACtl *)c);
        hipLaunchKernelGGL(k_aa_fin<2>, dim3(3), dim3(256), 0, st, (const double *)s->red, nb, s->fin, c, 2, s->fin,
                           s->dhist);
    };

    const bool use_graph = aa_use_graph();
    const int K = aa_chunk();
    const void *key[5] = {adj ? (const void *)adj->g : nullptr, adj ? (const void *)adj->mask : nullptr,
                          adj ? (const void *)adj->vbuf : nullptr, (const void *)s->dhist, (const void *)(intptr_t)K};
    // every path through aasol takes at most 2 evaluations per iteration counted in
    // maxit + picard_maxit, plus the first ones at the seed and at the switch
    const long bound = 4L * (2L * o->maxit + 4) + 2L * K;

    std::vector<double> prho(prog ? n : 0);       // the paused evaluation's profile
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev0, st));
    body(true, mixed);                            // eager first body, both maps
    HIPCHECK_CTX(ctx, sb_err);
    HIPCHECK_CTX(ctx, hipGetLastError());
    long bodies = 1, pauses = 0;
    while (true) {
        HIPCHECK_CTX(ctx, hipMemcpyAsync(hc, c, sizeof(AACtl), hipMemcpyDeviceToHost, st));
        HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
        if (hc->paused) {
            if (hc->paused & STOP_PROGRESS) {
                // the evaluation just finished used rho = exp(x) in s->rho, and only
                // the control step has run since (every gate closed at the stop)
                HIPCHECK_CTX(ctx, hipMemcpyAsync(prho.data(), s->rho, sizeof(double) * n, hipMemcpyDeviceToHost, st));
                HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
                if (o->progress(o->progress_user, hc->nevals, prho.data(), n) != 0) {
                    std::memcpy(rho_inout, prho.data(), sizeof(double) * n);
                    return fail(ctx, CDFT_ERR_STOPPED, "cdft_converge: stopped by the progress callback");
                }
            }
            if (hc->paused & STOP_MAP) only32 = false;   // the FP64 map from here
            hipLaunchKernelGGL(k_aa_resume, dim3(1), dim3(1), 0, st, c);
            HIPCHECK_CTX(ctx, hipGetLastError());
            pauses++;
            continue;
        }
        if (hc->phase == AA_DONE) break;
        // a pause leaves at most one chunk's bodies unused
        if (bodies > bound + pauses * K) return fail(ctx, CDFT_ERR_HIP, "aasol: the device-side solver did not terminate");
        if (small) {                              // a chunk of bodies in one launch
            HIPCHECK_CTX(ctx, small_body(s, c, K));
            HIPCHECK_CTX(ctx, hipGetLastError());
        } else if (use_graph) {
            const int v = only32 ? 1 : 0;
            hipGraphExec_t &exec = s->aa_exec[v];
            bool same = exec != nullptr;
            for (int j = 0; j < 5 && same; j++) same = (s->aa_key[v][j] == key[j]);
            if (!same) {
                if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
                hipGraph_t g = nullptr;
                HIPCHECK_CTX(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                for (int b = 0; b < K; b++) body(!only32, only32);
                hipError_t e1 = hipGetLastError();
                hipError_t e2 = hipStreamEndCapture(st, &g);
                if (e1 != hipSuccess || e2 != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    return fail(ctx, CDFT_ERR_HIP, std::string("aasol graph capture: ") +
                                                       hipGetErrorString(e1 != hipSuccess ? e1 : e2));
                }
                hipError_t e3 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e3 != hipSuccess) {
                    exec = nullptr;
                    return fail(ctx, CDFT_ERR_HIP, std::string("aasol graph instantiate: ") + hipGetErrorString(e3));
                }
                for (int j = 0; j < 5; j++) s->aa_key[v][j] = key[j];
            }
            HIPCHECK_CTX(ctx, hipGraphLaunch(exec, st));
        } else {
            for (int b = 0; b < K; b++) body(!only32, only32);
            HIPCHECK_CTX(ctx, hipGetLastError());
        }
        bodies += K;
    }
    // a restore left pending by the last step (back to the last finite FP32 point)
    hipLaunchKernelGGL(k_aa_point, dim3(ne), dim3(256), 0, st, (int)n, (const AACtl *)c, (const double *)s->df, s->Q,
                       s->Qalt, s->sol, (const double *)s->gx, (const double *)s->dg, s->DG, s->xsafe, ptgt, pmask);
    if (!adj) hipLaunchKernelGGL(k_exp, dim3(ne), dim3(256), 0, st, s->sol, s->out, (int)n);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev1, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(rho_inout, adj ? s->sol : s->out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    const int nh = hc->nh;
    if (hcap > 0 && nh > 0)
        HIPCHECK_CTX(ctx, hipMemcpyAsync(o->hist, s->dhist, sizeof(double) * std::min(nh, hcap), hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    r->idid = hc->idid;
    r->errcode = hc->errcode;
    r->iters = hc->k;
    r->nevals = hc->nevals;
    r->nhist = nh;
    r->resnorm = hc->r;
    r->tol = hc->tol;
    r->nevals_fp32 = hc->nevals32;
    r->reserved = 0;
    float ms = 0.0f;
    HIPCHECK_CTX(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    r->device_ms = ms;
    return CDFT_OK;
}

// ---- autotuner of the spectral path's line-FFT variants
// One weighted-density + transposed-convolution sequence (FP32 or FP64) with the
// system's current variant, ms per sequence (mean of reps after one warm-up).
static int time_spectral_seq(cdft_sys *s, bool fp32, int reps, double *ms)
{
    cdft_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHECK_CTX(ctx, hipEventCreate(&e0));
    struct EvFree { hipEvent_t a, b; ~EvFree() { if (a) (void)hipEventDestroy(a); if (b) (void)hipEventDestroy(b); } } ef{e0, nullptr};
    HIPCHECK_CTX(ctx, hipEventCreate(&e1));
    ef.b = e1;
    auto seq = [&]() {
        if (fp32) {
            launch_weighted3d<float>(s, s->rho, s->n32);
            launch_transposed3d<float>(s, s->dphi32, s->part);
        } else {
            launch_weighted3d<double>(s, s->rho, s->n);
            launch_transposed3d<double>(s, s->dphi, s->part);
        }
    };
    seq();
    HIPCHECK_CTX(ctx, hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) seq();
    HIPCHECK_CTX(ctx, hipEventRecord(e1, st));
    HIPCHECK_CTX(ctx, hipEventSynchronize(e1));
    HIPCHECK_CTX(ctx, hipGetLastError());
    float t = 0.0f;
    HIPCHECK_CTX(ctx, hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return CDFT_OK;
}

// A lines-per-tile count every axis of the grid can take: whole tiles, at most 4
// butterflies per thread, and at most 64 KB of LDS (two workgroups per CU).
static bool tb_fits(const cdft_sys *s, int tb, size_t elem)
{
    if (tb != 4 && tb != 8 && tb != 16) return false;
    for (int a = 0; a < 3; a++) {
        const int L = s->g3.n[a];
        if (L == 1) continue;
        if ((long)L * tb > 4096 || (size_t)L * tb * elem > 64 * 1024) return false;
    }
    return s->g3.n[0] % tb == 0 && (s->g3.N / s->g3.n[0]) % tb == 0;
}

static std::vector<long> tune_key(const cdft_sys *s)
{
    return {s->g3.n[0], s->g3.n[1], s->g3.n[2], s->nc, s->precision};
}

static void set_tune(cdft_sys *s, const cdft_tuning &t)
{
    const cdft_tuning &o = s->tune;
    bool same = true;
    for (int pr = 0; pr < 2; pr++)
        same = same && o.lines_per_tile[pr] == t.lines_per_tile[pr] && o.fused_z[pr] == t.fused_z[pr];
    if (!same) drop_body_graphs(s);
    s->tune = t;
}

static int tune_spectral(cdft_sys *s, bool force)
{
    if (!spectral(s)) return CDFT_OK;
    cdft_ctx *ctx = s->ctx;
    const std::vector<long> key = tune_key(s);
    if (!force) {
        if (s->tune.tuned) return CDFT_OK;
        auto it = ctx->tune_cache.find(key);
        if (it != ctx->tune_cache.end()) {
            set_tune(s, it->second);
            return CDFT_OK;
        }
    }
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    const cdft_tuning before = s->tune;
    cdft_tuning best = s->tune;
    const int nprec = (s->precision == CDFT_MIXED) ? 2 : 1;
    const int reps = s->N >= (1L << 22) ? 2 : 4;
    for (int pr = 0; pr < nprec; pr++) {
        const size_t elem = pr ? sizeof(cx<float>) : sizeof(cx<double>);
        double tbest = -1.0;
        for (int tb : {8, 4, 16}) {
            if (!tb_fits(s, tb, elem)) continue;
            for (int fz = (s->B3x[2] ? 1 : 0); fz >= 0; fz--) {
                s->tune.lines_per_tile[pr] = tb;
                s->tune.fused_z[pr] = fz;
                double ms = 0.0;
                int rc = time_spectral_seq(s, pr == 1, reps, &ms);
                if (rc != CDFT_OK) {
                    s->tune = before;
                    return rc;
                }
                if (tbest < 0 || ms < tbest) {
                    tbest = ms;
                    best.lines_per_tile[pr] = tb;
                    best.fused_z[pr] = fz;
                    best.ms[pr] = ms;
                }
            }
        }
        s->tune = best;
    }
    best.tuned = 1;
    s->tune = before;
    set_tune(s, best);
    ctx->tune_cache[key] = best;
    return CDFT_OK;
}

int cdft_system_tune(cdft_sys *s, int force, cdft_tuning *out)
{
    if (!s) return fail(nullptr, CDFT_ERR_ARG, "cdft_system_tune: NULL system");
    int rc = tune_spectral(s, force != 0);
    if (rc != CDFT_OK) return rc;
    if (out) *out = s->tune;
    return CDFT_OK;
}

int cdft_system_get_tuning(const cdft_sys *s, cdft_tuning *out)
{
    if (!s || !out) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_system_get_tuning: NULL argument");
    *out = s->tune;
    return CDFT_OK;
}

int cdft_system_set_tuning(cdft_sys *s, const cdft_tuning *in)
{
    if (!s || !in) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_system_set_tuning: NULL argument");
    cdft_tuning t = *in;
    for (int pr = 0; pr < 2; pr++) {
        if (spectral(s) && !tb_fits(s, t.lines_per_tile[pr], pr ? sizeof(cx<float>) : sizeof(cx<double>)))
            return fail(s->ctx, CDFT_ERR_ARG, "cdft_system_set_tuning: lines per tile the grid cannot take");
        if (t.fused_z[pr] && !s->B3x[2])
            return fail(s->ctx, CDFT_ERR_ARG, "cdft_system_set_tuning: no all-pairs z pass on this grid");
    }
    t.tuned = 1;                               // an explicit choice: cdft_converge keeps it
    set_tune(s, t);
    return CDFT_OK;
}

static bool autotune_enabled()
{
    const char *e = std::getenv("CDFT_AUTOTUNE");
    return !(e && e[0] == '0');
}

int cdft_converge(cdft_sys *s, double *rho_inout, const cdft_solver_opts *o, cdft_result *r)
{
    if (s && spectral(s) && s->N >= (1L << 16) && autotune_enabled()) {
        int rc = tune_spectral(s, false);
        if (rc != CDFT_OK) return rc;
    }
    return aasol_run(s, rho_inout, o, r, nullptr);
}

int cdft_functional_eval_fp32(cdft_sys *s, const double *rho, double *dfdrho)
{
    if (!s || !rho || !dfdrho) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_functional_eval_fp32: NULL argument");
    cdft_ctx *ctx = s->ctx;
    if (s->precision != CDFT_MIXED)
        return fail(ctx, CDFT_ERR_UNSUPPORTED, "cdft_functional_eval_fp32: the system was not created for mixed precision");
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    launch_functional32(s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(dfdrho, s->out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    return CDFT_OK;
}

// Device scratch of one call, freed on every return path.
struct DevBufs {
    std::vector<double *> p;
    ~DevBufs()
    {
        for (double *q : p) (void)hipFree(q);
    }
    hipError_t get(double **out, size_t count)
    {
        *out = nullptr;
        hipError_t e = hipMalloc((void **)out, sizeof(double) * (count ? count : 1));
        if (e == hipSuccess) p.push_back(*out);
        return e;
    }
};

int cdft_phi_vjp(cdft_sys *s, const double *rho, const double *cot, double *out)
{
    if (!s || !rho || !cot || !out) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_phi_vjp: NULL argument");
    cdft_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    DevBufs tmp;
    double *cotd = nullptr;
    HIPCHECK_CTX(ctx, tmp.get(&cotd, s->N));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(cotd, cot, sizeof(double) * s->N, hipMemcpyHostToDevice, st));
    launch_weighted(s);
    launch_phi_any(s);
    hipLaunchKernelGGL(k_scale_slots, dim3(nblocks(s->N)), dim3(256), 0, st, s->dphi, cotd, s->N, s->nc * s->F);
    launch_transposed(s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(out, s->out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    return CDFT_OK;
}

int cdft_hvp(cdft_sys *s, const double *rho, const double *v, double *out)
{
    if (!s || !rho || !v || !out) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_hvp: NULL argument");
    cdft_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    DevBufs tmp;
    double *vd = nullptr;
    HIPCHECK_CTX(ctx, tmp.get(&vd, n));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(vd, v, sizeof(double) * n, hipMemcpyHostToDevice, st));
    launch_weighted(s);
    launch_hvp(s, vd);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(out, s->out, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    return CDFT_OK;
}

int cdft_adjoint_solve(cdft_sys *s, const double *rho, const double *g, double *lambda_inout,
                       const cdft_solver_opts *o, cdft_result *r)
{
    if (!s || !rho || !g || !lambda_inout || !o || !r)
        return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_adjoint_solve: NULL argument");
    cdft_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    DevBufs tmp;
    double *gd = nullptr, *maskd = nullptr, *vbuf = nullptr;
    HIPCHECK_CTX(ctx, tmp.get(&gd, n));
    HIPCHECK_CTX(ctx, tmp.get(&maskd, n));
    HIPCHECK_CTX(ctx, tmp.get(&vbuf, n));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(gd, g, sizeof(double) * n, hipMemcpyHostToDevice, st));
    // dF_res/drho at rho*, then the mask of points where the exponent clamp is inactive
    launch_functional(s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    hipLaunchKernelGGL(k_clamp_mask, dim3(nblocks(n)), dim3(256), 0, st, s->out, s->vext, s->cpd, s->N, s->nc, maskd);
    HIPCHECK_CTX(ctx, hipGetLastError());
    AdjMap am{gd, maskd, vbuf};
    return aasol_run(s, lambda_inout, o, r, &am);     // s->n keeps W rho* for every product
}

int cdft_adjoint_params(cdft_sys *s, const double *rho, int nobj, const double *g, double *lambda_inout,
                        int nparam, cdft_sys *const *sys_plus, cdft_sys *const *sys_minus, const double *h,
                        const cdft_solver_opts *o, double *out, cdft_result *results)
{
    cdft_ctx *ctx = s ? s->ctx : nullptr;
    if (!s || !rho || !lambda_inout || nobj < 1 || nparam < 0 || (g && (!o || !results)) ||
        (nparam > 0 && (!sys_plus || !sys_minus || !h || !out)))
        return fail(ctx, CDFT_ERR_ARG, "cdft_adjoint_params: NULL or invalid argument");
    const long n = (long)s->N * s->nc;
    for (int p = 0; p < nparam; p++) {
        const cdft_sys *q[2] = {sys_plus[p], sys_minus[p]};
        for (const cdft_sys *x : q) {
            if (!x || x == s) return fail(ctx, CDFT_ERR_ARG, "cdft_adjoint_params: a perturbed system is NULL or sys");
            if (x->ctx != ctx || x->nc != s->nc || x->N != s->N || x->nd != s->nd || x->coord != s->coord)
                return fail(ctx, CDFT_ERR_ARG, "cdft_adjoint_params: perturbed systems must share the context and grid");
        }
        if (!(h[p] != 0.0) || !std::isfinite(h[p])) return fail(ctx, CDFT_ERR_ARG, "cdft_adjoint_params: h must be finite and nonzero");
    }
    // lambda_k for every objective (the same solve as cdft_adjoint_solve)
    if (g)
        for (int k = 0; k < nobj; k++) {
            const int rc = cdft_adjoint_solve(s, rho, g + (size_t)k * n, lambda_inout + (size_t)k * n, o, results + k);
            if (rc != CDFT_OK) return rc;
        }
    if (nparam == 0) return CDFT_OK;
    hipStream_t st = ctx->stream;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    const int nb = std::min(nblocks(n), 2048);
    DevBufs tmp;
    double *lam = nullptr, *part = nullptr;
    HIPCHECK_CTX(ctx, tmp.get(&lam, (size_t)nobj * n));
    HIPCHECK_CTX(ctx, tmp.get(&part, (size_t)nobj * nb));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(lam, lambda_inout, sizeof(double) * nobj * n, hipMemcpyHostToDevice, st));
    std::vector<double> hp((size_t)nobj * nb);
    for (int p = 0; p < nparam; p++) {
        // dF_res/drho of both perturbed systems at the common profile
        cdft_sys *q[2] = {sys_plus[p], sys_minus[p]};
        for (cdft_sys *x : q) {
            HIPCHECK_CTX(ctx, hipMemcpyAsync(x->rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
            launch_functional(x);
            hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(x).dphi, upd_src(x).part,
                               (const double *)nullptr, x->N, x->nc, upd_src(x).F, (const double *)nullptr, x->out,
                               (const double *)nullptr, (double *)nullptr, (double *)nullptr, (double *)nullptr,
                               (double *)nullptr, 1.0, 0, (double *)nullptr, 0, (const double *)nullptr,
                               (const double *)nullptr, (const AACtl *)nullptr);
        }
        hipLaunchKernelGGL(k_adj_param, dim3(nb), dim3(256), 0, st, (const double *)q[0]->out,
                           (const double *)q[0]->vext, (const double *)q[0]->cpd, (const double *)q[1]->out,
                           (const double *)q[1]->vext, (const double *)q[1]->cpd, (const double *)lam, s->N, s->nc,
                           nobj, part, nb);
        HIPCHECK_CTX(ctx, hipGetLastError());
        HIPCHECK_CTX(ctx, hipMemcpyAsync(hp.data(), part, sizeof(double) * nobj * nb, hipMemcpyDeviceToHost, st));
        HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
        for (int k = 0; k < nobj; k++) {
            double acc = 0.0;                          // fixed order: reproducible
            for (int b = 0; b < nb; b++) acc += hp[(size_t)k * nb + b];
            out[(size_t)k * nparam + p] = acc / (2.0 * h[p]);
        }
    }
    return CDFT_OK;
}

int cdft_ddft_rhs(cdft_sys *s, const double *eta, double *deta)
{
    if (!s || !eta || !deta) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_ddft_rhs: NULL argument");
    cdft_ctx *ctx = s->ctx;
    if (s->coord) return fail(ctx, CDFT_ERR_UNSUPPORTED, "cdft_ddft_rhs: Cartesian grids only (spectral derivatives)");
    hipStream_t st = ctx->stream;
    const long n = (long)s->N * s->nc;
    const int N = s->N, nc = s->nc;
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    DevBufs tmp;
    double *etad = nullptr, *detad = nullptr;
    HIPCHECK_CTX(ctx, tmp.get(&etad, n));
    HIPCHECK_CTX(ctx, tmp.get(&detad, n));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(etad, eta, sizeof(double) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ddft_rho, dim3(nblocks(n)), dim3(256), 0, st, etad, s->rho, (int)n);
    launch_functional(s);
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(256), 0, st, upd_src(s).dphi, upd_src(s).part,
                       (const double *)nullptr, s->N, s->nc, upd_src(s).F, (const double *)nullptr, s->out, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr, (double *)nullptr, (double *)nullptr, 1.0, 0, (double *)nullptr, 0,
                       (const double *)nullptr, (const double *)nullptr, (const AACtl *)nullptr);
    if (s->nd == 1) {
        if (!s->hddft) {
            HIPCHECK_CTX(ctx, hipMalloc((void **)&s->hddft, sizeof(double) * 4 * n));
            HIPCHECK_CTX(ctx, hipMemsetAsync(s->hddft, 0, sizeof(double) * 4 * n, st));
            const double mult = ((double)N / (s->ub[0] - s->lb[0])) / (double)N;   // fftfreq spacing
            hipLaunchKernelGGL(k_setup_hddft, dim3(nblocks(N), nc * 4), dim3(256), 0, st, s->hddft, N, nc, mult, s->L);
        }
        double *src = nullptr, *d = nullptr;
        HIPCHECK_CTX(ctx, tmp.get(&src, 4 * n));
        HIPCHECK_CTX(ctx, tmp.get(&d, 3 * n));
        hipLaunchKernelGGL(k_ddft_src, dim3(nblocks(n)), dim3(256), 0, st, etad, s->out, s->vext, N, nc, src,
                           (double *)nullptr);
        hipLaunchKernelGGL(k_circ1d<double>, dim3((N + CIRC_TW - 1) / CIRC_TW, nc * 3), dim3(256), 0, st, src,
                           (long)4 * N, (long)N, s->hddft, d, (long)3 * N, (long)N, 1, N, 4, 1.0, 0,
                           (const int *)nullptr);
        hipLaunchKernelGGL(k_ddft_combine, dim3(nblocks(n)), dim3(256), 0, st, d, N, nc, detad);
    } else {
        const bool plane = (s->g3.n[2] == 1);        // a 2D grid: no z passes
        double *mu = nullptr;
        HIPCHECK_CTX(ctx, tmp.get(&mu, n));
        hipLaunchKernelGGL(k_ddft_src, dim3(nblocks(n)), dim3(256), 0, st, etad, s->out, s->vext, N, nc,
                           (double *)nullptr, mu);
        for (int c = 0; c < nc; c++) {
            Pass3 p;
            std::memset(&p, 0, sizeof(p));
            p.in = etad + (size_t)c * N;
            p.in2 = mu + (size_t)c * N;
            p.out = s->A3;
            fft3_pass<double, 0, IN_PAIR, OUT_CPLX>(s, p, false);          // FFT(eta + i mu)
            p.in = s->A3;
            fft3_pass<double, 1, IN_CPLX, OUT_CPLX>(s, p, false);
            if (!plane) fft3_pass<double, 2, IN_CPLX, OUT_CPLX>(s, p, false);
            for (int kind = 0; kind < (plane ? 3 : 4); kind++) {    // 2D: no z gradient
                Pass3 q;
                std::memset(&q, 0, sizeof(q));
                q.in = s->A3;
                q.out = s->B3;
                q.mkind = kind;
                if (plane) {
                    fft3_pass<double, 1, IN_MULT, OUT_CPLX>(s, q, true);
                } else {
                    fft3_pass<double, 2, IN_MULT, OUT_CPLX>(s, q, true);
                    q.in = s->B3;
                    fft3_pass<double, 1, IN_CPLX, OUT_CPLX>(s, q, true);
                }
                q.in = s->B3;
                q.out = detad + (size_t)c * N;
                q.scale = 1.0 / (double)N;
                if (kind == 0) fft3_pass<double, 0, IN_CPLX, OUT_DDFT0>(s, q, true);
                else fft3_pass<double, 0, IN_CPLX, OUT_DDFT1>(s, q, true);
            }
        }
    }
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(deta, detad, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    return CDFT_OK;
}

int cdft_surface_tension(cdft_sys *s, const double *rho, const double *mu, double p, double *out)
{
    if (!s || !rho || !mu || !out) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_surface_tension: NULL argument");
    const long n = (long)s->N * s->nc;
    std::vector<double> dfd(n), fv(s->N);
    double fres = 0.0;
    int rc = cdft_functional_eval(s, rho, dfd.data(), &fres, nullptr);
    if (rc != CDFT_OK) return rc;
    auto integ = [&](const double *f) { return integ_sys(s, f); };
    std::vector<double> phi(s->N, 0.0), one(s->N, 1.0);
    const double NAv = 6.02214076e23, kB = 1.380649e-23;
    for (int c = 0; c < s->nc; c++)
        for (int i = 0; i < s->N; i++) {
            const double r = rho[(size_t)c * s->N + i];
            phi[i] += NAv * r * (std::log(r) - 1.5 * std::log(s->T) - 1.0);   // f_ideal, BasicIdeal.jl:10-39
        }
    double chem = 0.0;
    for (int c = 0; c < s->nc; c++) chem += mu[c] * integ(rho + (size_t)c * s->N);
    const double F = integ(phi.data()) + fres;
    *out = F * kB * s->T - chem + p * integ(one.data());   // surface_tension.jl:31-53
    return CDFT_OK;
}

int cdft_adsorption(cdft_sys *s, const double *rho, double *out)
{
    if (!s || !rho || !out) return fail(s ? s->ctx : nullptr, CDFT_ERR_ARG, "cdft_adsorption: NULL argument");
    double V = 1.0;                                                  // _domain_volume, adsorption.jl:15-25
    const double lo = s->lb[0], hi = s->ub[0];
    if (s->coord == CDFT_SPHERICAL) V = 4.0 * kPi / 3.0 * (hi * hi * hi - lo * lo * lo);
    else if (s->coord == CDFT_CYLINDRICAL) V = kPi * (hi * hi - lo * lo);
    else
        for (int a = 0; a < s->nd; a++) V *= s->ub[a] - s->lb[a];
    for (int c = 0; c < s->nc; c++) out[c] = integ_sys(s, rho + (size_t)c * s->N) / V;   // adsorption.jl:22-32
    return CDFT_OK;
}

// ---------------------------------------------------------------------------
// Batched persistent path
// ---------------------------------------------------------------------------
static int fft_length(int N)
{
    if ((N & (N - 1)) == 0) return N;          // power of two: the period itself
    int M = 1;
    while (M < 2 * N - 1) M <<= 1;
    return M;
}

static size_t batched_lds_bytes(const cdft_sys *s, bool g_lds, bool dg_lds)
{
    return batched_lds_layout(s->N, s->M, s->nc, g_lds, dg_lds);
}

static int batched_prepare(cdft_ctx *ctx, cdft_sys *s)
{
    hipStream_t st = ctx->stream;
    if (s->M == 0) {
        int M = fft_length(s->N);
        if (M < 8) return fail(ctx, CDFT_ERR_UNSUPPORTED, "batched path: grid too small (N < 5); use cdft_converge");
        if (M > 1024)
            return fail(ctx, CDFT_ERR_UNSUPPORTED,
                        "batched path: grid too large (FFT ring > 1024, i.e. N > 512 unless N is 1024); use cdft_converge");
        const int G1 = (M >> 1) + 1;
        size_t cnt = (size_t)s->nc * (s->F - 1) * G1;
        HIPCHECK_CTX(ctx, hipMalloc((void **)&s->H64, sizeof(double) * cnt));
        hipLaunchKernelGGL(k_setup_spec, dim3(nblocks(G1), s->nc * (s->F - 1)), dim3(256), 0, st, s->h, s->N, M, s->F,
                           s->H64);
        HIPCHECK_CTX(ctx, hipGetLastError());
        s->M = M;
    }
    if (!ctx->twiddles.count(s->M)) {
        double *t64 = nullptr;
        HIPCHECK_CTX(ctx, hipMalloc((void **)&t64, sizeof(double) * 2 * s->M));
        hipLaunchKernelGGL(k_setup_tw, dim3(nblocks(s->M)), dim3(256), 0, st, s->M, t64);
        HIPCHECK_CTX(ctx, hipGetLastError());
        ctx->twiddles[s->M] = t64;
    }
    return CDFT_OK;
}

// FP64 batches of small 1D systems (one or two components, N nc <= SB_MAXN, any
// association variant shared by the batch): one workgroup per system running
// k_small_body's bodies to the end of its solve (k_small_batch), each with its own
// control state; the same per-body arithmetic as cdft_converge of that system.
static bool small_batch_ok(const cdft_sys *s)
{
    return small_body_ok(s, false, s->precision == CDFT_MIXED);
}

extern "C++" {
template <int NC, int AS, bool MIXED>
static hipError_t small_batch_launch(int nsys, size_t lds, hipStream_t st, const SBSys *db, int maxbodies,
                                     unsigned long long *ticks)
{
    const hipError_t e = hipFuncSetAttribute((const void *)k_small_batch<NC, AS, MIXED>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_small_batch<NC, AS, MIXED>), dim3(nsys), dim3(SB_T), lds, st, db, maxbodies, ticks);
    return hipSuccess;
}

template <bool MIXED>
static hipError_t small_batch_m(int nc, int as, int nsys, size_t lds, hipStream_t st, const SBSys *db, int maxbodies,
                                unsigned long long *ticks)
{
    if (nc == 1)
        return as == 0 ? small_batch_launch<1, 0, MIXED>(nsys, lds, st, db, maxbodies, ticks)
                       : (as == 1 ? small_batch_launch<1, 1, MIXED>(nsys, lds, st, db, maxbodies, ticks)
                                  : small_batch_launch<1, 2, MIXED>(nsys, lds, st, db, maxbodies, ticks));
    return as == 0 ? small_batch_launch<2, 0, MIXED>(nsys, lds, st, db, maxbodies, ticks)
                   : (as == 1 ? small_batch_launch<2, 1, MIXED>(nsys, lds, st, db, maxbodies, ticks)
                              : small_batch_launch<2, 2, MIXED>(nsys, lds, st, db, maxbodies, ticks));
}
}  // extern "C++"

static int small_batch_run(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho,
                           const cdft_solver_opts *o, cdft_result *results, bool device_ptrs)
{
    hipStream_t st = ctx->stream;
    const int nc = sys[0]->nc, as = std::min(sys[0]->assoc, 2);
    for (int i = 1; i < nsys; i++)
        if (std::min(sys[i]->assoc, 2) != as)
            return fail(ctx, CDFT_ERR_UNSUPPORTED, "batched FP64 path: the systems must share the association variant");
    const int hcap = (o->hist && o->hist_cap > 0) ? o->hist_cap : 0;
    const bool mixed = (sys[0]->precision == CDFT_MIXED);
    size_t lds = 0;
    for (int i = 0; i < nsys; i++) lds = std::max(lds, sb_lds_bytes(sys[i]->N, nc, mixed));
    // the largest systems first, as in the register-state launch
    std::vector<int> order(nsys);
    for (int i = 0; i < nsys; i++) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return sys[a]->N != sys[b]->N ? sys[a]->N > sys[b]->N : a < b; });
    DevBufs tmp;
    double *dctl_d = nullptr, *dhist = nullptr, *dsys_d = nullptr, *dticks_d = nullptr;
    const size_t ctl_words = (sizeof(AACtl) + sizeof(double) - 1) / sizeof(double);
    const size_t sys_words = (sizeof(SBSys) + sizeof(double) - 1) / sizeof(double);
    HIPCHECK_CTX(ctx, tmp.get(&dctl_d, ctl_words * nsys));
    HIPCHECK_CTX(ctx, tmp.get(&dsys_d, sys_words * nsys));
    HIPCHECK_CTX(ctx, tmp.get(&dticks_d, nsys));
    if (hcap > 0) HIPCHECK_CTX(ctx, tmp.get(&dhist, (size_t)hcap * nsys));
    AACtl *dctl = reinterpret_cast<AACtl *>(dctl_d);
    SBSys *dsys = reinterpret_cast<SBSys *>(dsys_d);
    unsigned long long *dticks = reinterpret_cast<unsigned long long *>(dticks_d);
    std::vector<AACtl> hc(nsys);
    std::vector<SBSys> hs(nsys);
    for (int q = 0; q < nsys; q++) {
        cdft_sys *s = sys[order[q]];
        const long n = (long)s->N * s->nc;
        aa_ctl_init(&hc[q], o, mixed, hcap, false, 0);
        if (device_ptrs)
            HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho[order[q]], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        else
            HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho[order[q]], sizeof(double) * n, hipMemcpyHostToDevice, st));
        hs[q].a = sb_args(s, dctl + q, hcap > 0 ? dhist + (size_t)q * hcap : nullptr);
        hs[q].pc = s->pc;
    }
    HIPCHECK_CTX(ctx, hipMemcpyAsync(dctl, hc.data(), sizeof(AACtl) * nsys, hipMemcpyHostToDevice, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(dsys, hs.data(), sizeof(SBSys) * nsys, hipMemcpyHostToDevice, st));
    // every path through aasol takes at most 2 evaluations per iteration counted in
    // maxit + picard_maxit, plus the first ones (as aasol_run's bound)
    const long bound = 4L * (2L * o->maxit + 4) + 32;
    const int maxbodies = (int)std::min<long>(bound, 0x7fffffff);
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev0, st));
    HIPCHECK_CTX(ctx, mixed ? small_batch_m<true>(nc, as, nsys, lds, st, dsys, maxbodies, dticks)
                            : small_batch_m<false>(nc, as, nsys, lds, st, dsys, maxbodies, dticks));
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev1, st));
    std::vector<unsigned long long> ticks(nsys);
    HIPCHECK_CTX(ctx, hipMemcpyAsync(hc.data(), dctl, sizeof(AACtl) * nsys, hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipMemcpyAsync(ticks.data(), dticks, sizeof(unsigned long long) * nsys, hipMemcpyDeviceToHost, st));
    for (int q = 0; q < nsys; q++) {
        cdft_sys *s = sys[order[q]];
        const long n = (long)s->N * s->nc;
        HIPCHECK_CTX(ctx, hipMemcpyAsync(rho[order[q]], s->rho, sizeof(double) * n,
                                         device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        if (hcap > 0)
            HIPCHECK_CTX(ctx, hipMemcpyAsync(o->hist + (size_t)order[q] * hcap, dhist + (size_t)q * hcap,
                                             sizeof(double) * hcap, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    float ms = 0.0f;
    HIPCHECK_CTX(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    int wall_rate_khz = 0;
    HIPCHECK_CTX(ctx, hipDeviceGetAttribute(&wall_rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    for (int q = 0; q < nsys; q++) {
        const AACtl &c = hc[q];
        if (c.phase != AA_DONE) return fail(ctx, CDFT_ERR_HIP, "batched FP64 path: a solve did not terminate");
        cdft_result &r = results[order[q]];
        r.idid = c.idid;
        r.errcode = c.errcode;
        r.iters = c.k;
        r.nevals = c.nevals;
        r.nhist = hcap > 0 ? c.nh : 0;
        r.resnorm = c.r;
        r.tol = c.tol;
        r.nevals_fp32 = c.nevals32;
        r.reserved = 0;
        r.device_ms = wall_rate_khz > 0 ? (double)ticks[q] / (double)wall_rate_khz : 0.0;
    }
    return CDFT_OK;
}

static int batch_impl(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho, const cdft_solver_opts *o,
                      cdft_result *results, bool device_ptrs, bool noalloc)
{
    const char *who = noalloc ? "cdft_converge_batch_device_noalloc" : "cdft_converge_batch";
    if (!ctx || nsys < 0 || (nsys > 0 && (!sys || !rho || !results)) || !o)
        return fail(ctx, CDFT_ERR_ARG, std::string(who) + ": NULL argument");
    if (nsys == 0) return CDFT_OK;
    if (noalloc && nsys > kNoAllocMaxSys)
        return fail(ctx, CDFT_ERR_ARG,
                    std::string(who) + ": at most " + std::to_string(kNoAllocMaxSys) + " systems per call");
    if (o->m < 0 || o->m > CDFT_MAXM - 1) return fail(ctx, CDFT_ERR_ARG, "anderson m must be in 0..7");
    if (!(std::fabs(o->beta) > 0 && std::fabs(o->beta) <= 1)) return fail(ctx, CDFT_ERR_ARG, "abs(beta) must be in (0,1]");
    if (o->maxit < 0) return fail(ctx, CDFT_ERR_ARG, "maxit must be >= 0");
    const int nc = sys[0] ? sys[0]->nc : 0;
    const int prec = sys[0] ? sys[0]->precision : 0;
    // FP64 batches of small 1D systems: the single-workgroup bodies, one workgroup per
    // system (the allocating entry points)
    // Routes: one-component mixed batches the register-state kernel takes go there (the
    // headline kernel); other batches of small 1D systems (FP64, or mixed with two
    // components, association or a deeper history) take the single-workgroup bodies, one
    // workgroup per system (the allocating entry points)
    bool small = !noalloc, reg = true;
    for (int i = 0; i < nsys && (small || reg); i++) {
        if (!sys[i] || !rho[i]) return fail(ctx, CDFT_ERR_ARG, std::string(who) + ": NULL system or profile");
        if (sys[i]->ctx != ctx) return fail(ctx, CDFT_ERR_ARG, "all systems must belong to the context");
        for (int j = 0; j < i; j++)
            if (sys[j] == sys[i]) return fail(ctx, CDFT_ERR_ARG, "a system appears twice in the batch");
        const cdft_sys *x = sys[i];
        small = small && small_batch_ok(x) && x->nc == nc && x->precision == prec;
        reg = reg && x->nc == 1 && x->precision == CDFT_MIXED && x->N <= kRegNPT * BT && o->m <= kRegAM &&
              x->nd == 1 && !x->coord && !x->assoc;
    }
    if (small && !reg) {
        HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
        return small_batch_run(ctx, nsys, sys, rho, o, results, device_ptrs);
    }
    for (int i = 0; i < nsys; i++) {
        if (!sys[i] || !rho[i]) return fail(ctx, CDFT_ERR_ARG, std::string(who) + ": NULL system or profile");
        if (sys[i]->ctx != ctx) return fail(ctx, CDFT_ERR_ARG, "all systems must belong to the context");
        if (sys[i]->nc != nc) return fail(ctx, CDFT_ERR_UNSUPPORTED, "batched systems must share nc");
        // the multi-component register/HBM-state kernel faulted on the GPU and is
        // switched off until that is understood (FP64 batches of up to two components
        // take the single-workgroup bodies above)
        if (sys[i]->nc != 1)
            return fail(ctx, CDFT_ERR_UNSUPPORTED,
                        "batched path: multi-component batches run in FP64 (CDFT_FP64 systems of at most 512 "
                        "points x components, two components at most); use cdft_converge per system otherwise");
        if (sys[i]->precision != prec) return fail(ctx, CDFT_ERR_UNSUPPORTED, "batched systems must share precision");
        // the register-state kernel running the FP64 map from the first evaluation
        // faulted on the GPU; switched off until that is understood (FP64 batches of
        // small systems take the single-workgroup bodies above)
        if (sys[i]->precision != CDFT_MIXED)
            return fail(ctx, CDFT_ERR_UNSUPPORTED,
                        "batched path: FP64 batches take 1D systems of at most 512 points x components "
                        "(two components at most, one association variant); use cdft_converge otherwise");
        // the kernels with the Anderson state in HBM share the code of the faulting
        // multi-component kernel and have not run on a GPU: only the register-state
        // kernel (at most 512 points, m <= 5) is enabled
        if (sys[i]->N > kRegNPT * BT || o->m > kRegAM)
            return fail(ctx, CDFT_ERR_UNSUPPORTED,
                        "batched path: at most 512 grid points and anderson m <= 5 for now; use cdft_converge");
        if (sys[i]->nd != 1 || sys[i]->coord)
            return fail(ctx, CDFT_ERR_UNSUPPORTED, "the batched path takes 1D Cartesian systems");
        if (sys[i]->assoc)
            return fail(ctx, CDFT_ERR_UNSUPPORTED,
                        "batched path: associating systems are solved per system; use cdft_converge");
        // prepared at creation (spectra on the ring, the ring's twiddles); the
        // allocating entry points prepare a system here if needed
        if (noalloc && (!sys[i]->H64 || !ctx->twiddles.count(sys[i]->M)))
            return fail(ctx, CDFT_ERR_ARG, std::string(who) + ": system not prepared for the batched path");
        for (int j = 0; j < i; j++)
            if (sys[j] == sys[i]) return fail(ctx, CDFT_ERR_ARG, "a system appears twice in the batch");
    }
    HIPCHECK_CTX(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    int maxlds = 0;
    HIPCHECK_CTX(ctx, hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    // Workgroup size: 256 threads (two points per thread) by default; 512 threads (one
    // point per thread, the field-pair FFTs spread over the whole workgroup) with
    // CDFT_BATCH_THREADS=512.  Both run an evaluation of a 446-point pore in the same
    // 23.5 us (the solve is bound by the chain of LDS passes and barriers, not by work
    // per thread), so the smaller kernel stays the default.
    int bt = BT;
    if (const char *env = std::getenv("CDFT_BATCH_THREADS")) {
        const int v = std::atoi(env);
        if (v == BT || v == 2 * BT) bt = v;
    }
    const bool dg_lds = (bt == 2 * BT);               // the 512-thread kernel keeps Delta-G in LDS
    size_t lds_g = 0, lds_nog = 0;
    for (int i = 0; i < nsys; i++) {
        if (!noalloc) {
            int rc = batched_prepare(ctx, sys[i]);
            if (rc != CDFT_OK) return rc;
        }
        lds_g = std::max(lds_g, batched_lds_bytes(sys[i], true, dg_lds));
        lds_nog = std::max(lds_nog, batched_lds_bytes(sys[i], false, dg_lds));
    }
    // spectra in LDS when the largest system leaves room for them, else read from HBM
    const size_t lds_limit = std::max<size_t>((size_t)maxlds, 160 * 1024) - sizeof(AState) - sizeof(PhiConst);
    const bool g_lds = lds_g <= lds_limit;
    const size_t lds = g_lds ? lds_g : lds_nog;
    if (lds > lds_limit)
        return fail(ctx, CDFT_ERR_UNSUPPORTED, "batched path: system too large for one workgroup's LDS; use cdft_converge");
    // host staging (sized at context creation; grown here only by the allocating entry points)
    if (ctx->border.size() < (size_t)nsys) {
        ctx->border.resize(nsys);
        ctx->bstage.resize(sizeof(BSys) * nsys);
        ctx->rstage.resize(sizeof(BResult) * nsys);
    }
    // largest systems first: the tail of the launch is then the short ones (ties in
    // input order; std::sort with the index as tie-break needs no buffer)
    int *order = ctx->border.data();
    for (int i = 0; i < nsys; i++) order[i] = i;
    std::sort(order, order + nsys, [&](int a, int b) { return sys[a]->N != sys[b]->N ? sys[a]->N > sys[b]->N : a < b; });
    BSys *hb = reinterpret_cast<BSys *>(ctx->bstage.data());
    for (int q = 0; q < nsys; q++) {
        cdft_sys *s = sys[order[q]];
        BSys &b = hb[q];
        std::memset(&b, 0, sizeof(b));
        b.N = s->N;
        b.nc = s->nc;
        b.M = s->M;
        int lg = 0;
        while ((1 << lg) < s->M) lg++;
        b.log2M = lg;
        b.NA = s->NA;
        for (int c = 0; c < s->nc; c++) b.cp[c] = s->cp[c];
        b.vext = s->vext;
        b.G = s->H64;
        b.tw64 = ctx->twiddles.find(s->M)->second;
        long n = (long)s->N * s->nc;
        if (device_ptrs) {
            b.rho = rho[order[q]];
        } else {
            HIPCHECK_CTX(ctx, hipMemcpyAsync(s->rho, rho[order[q]], sizeof(double) * n, hipMemcpyHostToDevice, st));
            b.rho = s->rho;
        }
        b.sol = s->sol;
        b.gx = s->gx;
        b.df = s->df;
        b.dg = s->dg;
        b.res = s->res;
        b.Q = s->Q;
        b.DG = s->DG;
        b.hist = nullptr;
        b.hist_cap = 0;
        b.pc = s->pc;
    }
    // device copies of the descriptors and the results: this device's static scratch
    // (allocation-free entry point) or buffers of this call
    BSys *db = nullptr;
    BResult *dr = nullptr;
    struct Free {
        void *p = nullptr;
        ~Free() { if (p) (void)hipFree(p); }
    } f1, f2;
    if (noalloc) {
        db = reinterpret_cast<BSys *>(ctx->scratch);
        dr = reinterpret_cast<BResult *>(static_cast<unsigned char *>(ctx->scratch) + kScratchDescBytes);
    } else {
        HIPCHECK_CTX(ctx, hipMalloc((void **)&db, sizeof(BSys) * nsys));
        f1.p = db;
        HIPCHECK_CTX(ctx, hipMalloc((void **)&dr, sizeof(BResult) * nsys));
        f2.p = dr;
    }
    // residual histories (allocating entry points): system i's at o->hist + i hist_cap
    const int hcap = (!noalloc && o->hist && o->hist_cap > 0) ? o->hist_cap : 0;
    Free f3;
    if (hcap > 0) {
        double *dh = nullptr;
        HIPCHECK_CTX(ctx, hipMalloc((void **)&dh, sizeof(double) * (size_t)hcap * nsys));
        f3.p = dh;
        for (int q = 0; q < nsys; q++) {
            hb[q].hist = dh + (size_t)q * hcap;
            hb[q].hist_cap = hcap;
        }
    }
    HIPCHECK_CTX(ctx, hipMemcpyAsync(db, hb, sizeof(BSys) * nsys, hipMemcpyHostToDevice, st));
    BOpts bo;
    bo.maxit = o->maxit;
    bo.m = o->m;
    bo.beta = o->beta;
    bo.tol = o->tol;
    bo.anderson_start = o->anderson_start;
    bo.mixed = (prec == CDFT_MIXED);
    bo.switch_factor = 4.0;
    bo.stall_window = 40;
    bo.g_lds = g_lds ? 1 : 0;
    // Only the register-state kernel is built and launched: the checks above admit
    // one-component systems of at most kRegNPT * BT points with m <= kRegAM.  The
    // HBM-state variants (k_batched<NC, false>) stay in the source but are not
    // instantiated while the multi-component fault is unexplained.
    // chain-free batches (every segment number 1, e.g. methane) use a kernel without
    // the chain term's code paths, which needs fewer registers
    bool chain_free = true;
    for (int i = 0; i < nsys && chain_free; i++)
        for (int c = 0; c < sys[i]->nc; c++) chain_free = chain_free && sys[i]->m[c] == 1.0;
    auto kern = (bt == BT) ? (chain_free ? k_batched<1, true, BT, true> : k_batched<1, true, BT, false>)
                           : (chain_free ? k_batched<1, true, 2 * BT, true> : k_batched<1, true, 2 * BT, false>);
    if (hcap > 0)
        kern = (bt == BT) ? (chain_free ? k_batched<1, true, BT, true, true> : k_batched<1, true, BT, false, true>)
                          : (chain_free ? k_batched<1, true, 2 * BT, true, true>
                                        : k_batched<1, true, 2 * BT, false, true>);
    HIPCHECK_CTX(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(kern, dim3(nsys), dim3(bt), lds, st, db, bo, dr);
    HIPCHECK_CTX(ctx, hipGetLastError());
    HIPCHECK_CTX(ctx, hipEventRecord(ctx->ev1, st));
    BResult *hr = reinterpret_cast<BResult *>(ctx->rstage.data());
    HIPCHECK_CTX(ctx, hipMemcpyAsync(hr, dr, sizeof(BResult) * nsys, hipMemcpyDeviceToHost, st));
    if (!device_ptrs)
        for (int q = 0; q < nsys; q++) {
            cdft_sys *s = sys[order[q]];
            HIPCHECK_CTX(ctx, hipMemcpyAsync(rho[order[q]], s->rho, sizeof(double) * s->N * s->nc,
                                             hipMemcpyDeviceToHost, st));
        }
    if (hcap > 0)
        for (int q = 0; q < nsys; q++)
            HIPCHECK_CTX(ctx, hipMemcpyAsync(o->hist + (size_t)order[q] * hcap, hb[q].hist, sizeof(double) * hcap,
                                             hipMemcpyDeviceToHost, st));
    HIPCHECK_CTX(ctx, hipStreamSynchronize(st));
    float ms = 0.0f;
    HIPCHECK_CTX(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    int wall_rate_khz = 0;
    HIPCHECK_CTX(ctx, hipDeviceGetAttribute(&wall_rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    for (int q = 0; q < nsys; q++) {
        cdft_result &r = results[order[q]];
        r.idid = hr[q].idid;
        r.errcode = hr[q].errcode;
        r.iters = hr[q].iters;
        r.nevals = hr[q].nevals;
        r.nhist = hcap > 0 ? hr[q].nhist : 0;
        r.resnorm = hr[q].resnorm;
        r.tol = hr[q].tol;
        r.nevals_fp32 = hr[q].nevals32;
        r.reserved = 0;
        r.device_ms = wall_rate_khz > 0 ? (double)hr[q].ticks / (double)wall_rate_khz : 0.0;
    }
    return CDFT_OK;
}

int cdft_converge_batch(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_inout,
                        const cdft_solver_opts *opts, cdft_result *results)
{
    return batch_impl(ctx, nsys, sys, rho_inout, opts, results, false, false);
}

int cdft_converge_batch_device(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_dev,
                               const cdft_solver_opts *opts, cdft_result *results)
{
    return batch_impl(ctx, nsys, sys, rho_dev, opts, results, true, false);
}

int cdft_converge_batch_device_noalloc(cdft_ctx *ctx, int nsys, cdft_sys *const *sys, double *const *rho_dev,
                                       const cdft_solver_opts *opts, cdft_result *results)
{
    return batch_impl(ctx, nsys, sys, rho_dev, opts, results, true, true);
}

}  // extern "C"
