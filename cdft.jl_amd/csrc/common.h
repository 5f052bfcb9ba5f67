// common.h -- shared host/device definitions of the cDFT HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "cdft_hip.h"
#include "phi.h"

#define CDFT_MAXM 8          // largest Anderson depth supported
#define CDFT_NQ 16           // reduction slots per vector pass

#include "kernels.h"

namespace cdft {
// clamp!(ln_Gx, -100, 100) (src/methods/converge.jl:150) with Julia's clamp
// semantics: a NaN stays NaN (fmin/fmax would replace it by a bound), so a
// blown-up map evaluation ends aasol with a NaN residual as in the reference.
__host__ __device__ __attribute__((always_inline)) inline double clamp_lng(double v)
{
    return v > 100.0 ? 100.0 : (v < -100.0 ? -100.0 : v);
}

// Per-system constants of the 3D weights (fft3d.hip).
struct Grid3 {
    int n[3];                 // points per axis (powers of two)
    int lg[3];                // log2 of n
    long N;                   // n0 * n1 * n2
    double mult[3];           // fftfreq spacing times L: omega_a = k_signed * mult[a]
    double R[CDFT_MAXC][3];   // 2 pi * width for widths d/2, d, psi d (reduced)
};
}  // namespace cdft

struct cdft_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    double last_kernel_ms = 0.0;
    std::map<int, double *> twiddles;   // FFT ring length M -> W_M^t table (re, im pairs)
    std::map<std::vector<long>, cdft_tuning> tune_cache;   // (grid, nc, precision) -> measured choice
    // batched launches: this device's static scratch (system descriptors and results
    // of the allocation-free entry point) and host staging sized at creation
    void *scratch = nullptr;
    std::vector<unsigned char> bstage;   // host copies of the system descriptors
    std::vector<unsigned char> rstage;   // host copies of the results
    std::vector<int> border;             // launch order of the systems
};

// Per-system state.  Host copies of the descriptor plus device buffers.
struct cdft_sys {
    cdft_ctx *ctx = nullptr;
    int nc = 0, nd = 0, N = 0, F = 0;
    int ngrid[3] = {1, 1, 1};
    double lb[3] = {0, 0, 0}, ub[3] = {0, 0, 0};
    double T = 0, p = 0, L = 0, psi = 1.3862, NA = 0;
    double m[CDFT_MAXC], d[CDFT_MAXC], rho_bulk[CDFT_MAXC], mu_res[CDFT_MAXC];
    double sigma[CDFT_MAXC * CDFT_MAXC], eps[CDFT_MAXC * CDFT_MAXC];
    double cp[CDFT_MAXC];      // mu_res + ln rho_bulk (converge.jl:121)
    int precision = CDFT_FP64;
    int coord = CDFT_CARTESIAN;  // radial grids: dense operators in h (and hT), weights rw
    bool has_vext = false;
    int assoc = 0;             // association variant of the pointwise kernels: 0 none, 1 one pair, 2 general
    PhiConst pc;

    // device buffers (FP64)
    double *h = nullptr;       // real-space kernels [(c*F+f)*N + j], f >= 1 (radial: dense [(c*F+f)*N*N])
    double *hT = nullptr;      // radial: transposed vector-slot operator [c*N*N]
    std::vector<double> rw;    // radial: weights of int f dV (4 pi or 2 pi times the transform's quadrature)
    double *vext = nullptr;    // [c*N + i]
    double *cpd = nullptr;     // device copy of cp [nc]
    double *rho = nullptr;     // current profile
    double *n = nullptr;       // weighted densities [(c*F+f)*N + i]
    double *dphi = nullptr;    // dPhi/dn, same layout
    double *part = nullptr;    // per-field transposed convolutions [(c*(F-1)+f-1)*N + i]
    double *fval = nullptr;    // Phi [N]
    double *out = nullptr;     // dF/drho or lnG [c*N + i]
    // Anderson storage (anderson.jl:665-703)
    double *sol = nullptr, *gx = nullptr, *df = nullptr, *dg = nullptr, *res = nullptr;
    double *Q = nullptr, *Qalt = nullptr, *DG = nullptr;   // m columns each
    // batched path: real half-spectra of the kernels on the FFT ring of length M
    int M = 0;
    double *H64 = nullptr;
    // 3D grids: tabulated spectra G [(c*8 + f-1)*N + k] and three complex work arrays
    cdft::Grid3 g3;
    double *G3 = nullptr, *A3 = nullptr, *B3 = nullptr, *S3 = nullptr;
    // three more complex work arrays when z lines fit the all-pairs z pass (128, 256 points)
    double *B3x[3] = {nullptr, nullptr, nullptr};
    // mixed precision: FP32 spectra (3D), weighted densities and gradients
    float *G3f = nullptr, *n32 = nullptr, *dphi32 = nullptr;
    float *h32 = nullptr;      // 1D mixed precision: FP32 copy of the real-space kernels h
    double *xsafe = nullptr;   // mixed precision: the last point whose FP32 map value was finite
    double *hddft = nullptr;   // 1D DDFT derivative kernels [(c*4 + f)*N + j], f = 1 (Laplacian), 2, 3 (gradient)
    double *red = nullptr;     // block partials
    double *fin = nullptr;     // finalized reductions [CDFT_NQ]
    int nblk = 0;
    // device-side Anderson solve: its state, residual history, and the captured
    // graph of solver bodies with the pointers it was captured for
    AACtl *ctl = nullptr;
    double *dhist = nullptr;
    int dhist_cap = 0;
    // [0]: bodies with the FP64 map (or the adjoint's), [1]: the FP32 map only
    hipGraphExec_t aa_exec[2] = {nullptr, nullptr};
    const void *aa_key[2][5] = {};
    const int *gate = nullptr;  // gate of the kernels being enqueued (null: always run)
    cdft_tuning tune = {{8, 8}, {1, 1}, {0.0, 0.0}, 0};   // line-FFT variant per map precision
};

#define HIPCHECK_CTX(ctx, call)                                                              \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return CDFT_ERR_HIP;                                                              \
        }                                                                                     \
    } while (0)
