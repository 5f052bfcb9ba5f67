// kernels.h -- small by-value argument structs shared by the kernels and the host.
#pragma once

#ifndef CDFT_MAXM
#define CDFT_MAXM 8
#endif

// A small dense matrix or coefficient vector passed to a kernel by value, with
// an optional column map (the ring order of the Anderson difference history).
struct AAMat {
    double a[CDFT_MAXM * CDFT_MAXM];
    int col[CDFT_MAXM];
};

// State of one plain-path Anderson solve kept on the device (aasol,
// src/utils/anderson.jl:301-443, with converge!'s defaults), so that the whole
// iteration runs without the host: one thread of the kernels that finish the
// reductions advances it, and the N-wide kernels read their work description from it
// and return at once when their gate is set.  One map evaluation per body.
enum AAPhase {
    AA_FIRST = 0,      // the first evaluation at the seed
    AA_PICARD = 1,     // damped Picard warm-up
    AA_START = 2,      // first Anderson step of a (re)start
    AA_LOOP = 3,       // a step of the Anderson loop (decided in the control kernel)
    AA_SWITCH = 4,     // mixed precision: the FP64 map's first evaluation at the switch point
    AA_STEP = 5,       // an Anderson step in flight
    AA_DONE = 6
};

struct AACtl {
    // gates (an N-wide kernel returns at once when its gate is nonzero)
    int g_done;        // the solve has ended
    int g_proj;        // no Gram-Schmidt projections in this body
    int g_f64;         // this body's map is not the FP64 one
    int g_f32;         // this body's map is not the FP32 one
    int paused;        // stopped after a control step (gates kept in sg), bits:
                       // 1 for the host's progress callback, 2 before a body whose
                       // map the host's queued bodies do not carry
    // work of the body in flight
    int umode;         // k_update mode (1 first/Picard, 2 Anderson)
    int pmode;         // point kernel: 0 none, 1 sol = gx, 2 sol = xsafe, 3 Anderson point
    int keep;          // copy sol to xsafe before moving the point
    int qsel;          // which of the two Q arrays holds the current basis
    int down, kqq, qcol, mk, slot;
    int evfp32;        // the body's map is the FP32 one
    double beta_u;     // damping of this body's update
    double nrm;        // norm of the new Q column
    AAMat qx;          // downdate rotation (Householder Q of R[:, 2:m])
    AAMat th;          // theta and the DG columns it weights
    // solver state
    int phase, k, ka, it, nh, nevals, nevals32;
    int fp32, fp32_failed, toosoon, since_best, restore, errcode, idid;
    int m, maxit, picard_maxit, hist_cap, adjoint;
    int pause_every;   // pause after every this many evaluations (0: never)
    int graph_map;     // 1: the host's bodies carry the FP32 map only (0: both maps)
    int sg[4];         // g_done, g_proj, g_f64, g_f32 of the body after a stop
    double r, dfn2, dfres, tol, up_bd, picard_tol, best;
    double rtol, atol, picard_rtol, picard_atol, beta, picard_beta;
    double R[CDFT_MAXM * CDFT_MAXM];
    double tres[CDFT_MAXM];
    int ring[CDFT_MAXM];
    int cnt[3];        // blocks done of the k_aa_fin launches that run a solver step
};
