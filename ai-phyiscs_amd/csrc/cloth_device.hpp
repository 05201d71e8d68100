// SPDX-License-Identifier: Apache-2.0
//
// Plain-old-data structures shared by the host engine and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace chysx_amd {

// A tile is the unit of work of one 256-lane workgroup: up to 256
// consecutive rows of one instance, starting on a 64-row slice boundary.
struct Tile {
    int inst;   // instance id
    int row0;   // first row (padded row space, multiple of 64)
    int nrows;  // rows in this tile (<= 256)
    int part0;  // particle index of row0 in the external buffers
};

// PCG status of one instance.
enum PcgStatus : int {
    kRunning = 0,    // iterating
    kCheck = 1,      // recursive residual passed; true residual pending
    kConverged = 2,  // true relative residual <= tol
    kStopped = 3,    // max_iter iterations done (normal end when tol == 0)
};

// Per-instance solver state, FP64 scalars (inner products, alpha, beta and
// the convergence test).  Lives in device memory; counters implement the
// "last workgroup finishes the reduction" pattern.
struct alignas(16) InstState {
    double rho;      // r . z
    double sigma;    // p . A p
    double bb;       // max(b . b, 1e-30)
    double rr;       // r . r (recursive, or true after a check)
    double beta;     // p <- z + beta p, applied inside the next SpMV
    double relres;   // last true relative residual
    int status;
    int iters;
    int restarts;
    int parity;      // which p buffer holds the current direction
    int ntiles;      // tiles of this instance (set at build)
    int tile0;       // first tile of this instance (set at build)
    unsigned cnt;    // arrival counter for cross-workgroup reductions
    int pad;
};

// Per-instance step parameters and solver options.
struct alignas(16) InstParams {
    double tol2;        // tol^2; 0 runs exactly max_iter iterations
    int max_iter;
    float gx, gy, gz;
    float damping;
    float pin_k;
    int force_active;   // 1: external force enters the predictor
    int springs_active; // 1: spring gradient/Hessian enter the step
};

}  // namespace chysx_amd
