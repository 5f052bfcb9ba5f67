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
