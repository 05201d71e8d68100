// SPDX-License-Identifier: Apache-2.0
//
// Plain-old-data structures shared by the host engine and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace chysx_amd {

// A slice is 64 consecutive rows of one instance (rows 64 s .. 64 s + 63 of
// the padded row space), processed by one 64-lane wavefront, one lane per
// block row.  Instances are padded to whole groups of kGroupSlices slices
// (one PCG workgroup each); slices past an instance's rows have nvalid 0.
constexpr int kGroupSlices = 16;
struct SliceDesc {
    int inst;    // instance id
    int part0;   // particle index (external buffers) of the slice's first row
    int nvalid;  // rows of the slice that exist (<= 64)
    int nlower;  // one past the last entry position holding a lower block (column < row) in
                 // any lane: entries k >= nlower are upper or padding in every lane
    int group0;  // the instance's first 16-slice group and its group count (as in its
    int ngroups; // InstState), so a pass can load group partials before the state arrives
};

// PCG status of one instance.
enum PcgStatus : int {
    kRunning = 0,    // iterating
    kConverged = 2,  // true relative residual <= tol
    kStopped = 3,    // max_iter iterations done (normal end when tol == 0)
};

// What the next SpMV pass of a running instance computes.
enum PcgMode : int {
    kModeIter = 0,   // p = z + beta p_prev, A p, p . A p
    kModeCheck = 1,  // the recursive residual passed the test: true residual b - A x
    kModeFlush = 2,  // x += alpha_pend p (the deferred term) before a check or a stop
};

// Reduction phases of pcg_reduce_kernel.  (The per-iteration reduction,
// r . z and |r|^2 of an update pass, or the true |r|^2 and r . z of a check
// pass, is pcg_advance: run by the next SpMV pass itself, or by
// pcg_advance_kernel at the end of a chunk of iterations.)
enum PcgPhase : int {
    kPhaseStart = 0,  // after the initial residual: |r|^2, |b|^2, r . z
    kPhaseFinal = 3,  // after the report residual of a stopped solve: |r|^2
};

// Inner-product partials: one FP64 value per 16-slice group and slot,
// partials[slot * ngroups + group].  Slots 0-2: |r|^2, |b|^2, r . z of a
// residual pass (start, report), or r . z, |r|^2 of an update pass (slots
// 0-1).  A check pass (the SpMV pass in kModeCheck) writes the true |r|^2
// and r . z to kSlotCheckRR and kSlotCheckRZ; slot kSlotSigma: p . A p of an
// SpMV pass.  No pass writes a slot that the same pass reads: the SpMV pass
// reads slots 0-1 (or the check slots) for its advance while it writes the
// check slots (or sigma), and the update pass reads sigma while it writes
// slots 0-1.
constexpr int kSlotCheckRZ = 2;
constexpr int kSlotCheckRR = 3;
constexpr int kSlotSigma = 4;
constexpr int kPartialSlots = 5;

// Per-instance solver state, FP64 scalars (inner products, alpha, beta and
// the convergence test).  Lives in device memory, in two buffers during the
// iterations: each SpMV pass reads one, advances the state by the previous
// iteration's partials (every workgroup computes the same result), and the
// workgroup of the instance's first group writes it to the other, which the
// update pass and the next SpMV pass read.  The update pass writes only
// `sigma` and `alpha`, which no other workgroup of that pass reads.  No kernel
// needs cross-workgroup synchronisation.
struct alignas(16) InstState {
    double rho;      // r . z
    double sigma;    // p . A p
    double bb;       // max(b . b, 1e-30)
    double rr;       // r . r (recursive, or true after a check)
    double beta;     // p <- z + beta p, applied inside the next SpMV
    double relres;   // last true relative residual
    int status;      // PcgStatus
    int mode;        // PcgMode
    int iters;
    int restarts;
    int parity;      // which p buffer holds the current direction
    int slice0;      // first slice of this instance (set at build)
    int nslices;     // slices holding this instance's rows (set at build)
    int group0;      // first 16-slice group of this instance (set at build)
    int ngroups;     // its groups (set at build)
    // Deferred iterate update: an update pass with `defer` set leaves
    // x += alpha p to the next update pass, which adds both terms while both
    // directions are still in the p buffers (x is read and written every
    // other iteration).  pend: a term alpha_pend p is owed; flush_to: what a
    // flush pass leads to (kModeCheck, or kStopped as a status).
    double alpha_pend;
    int pend;
    int defer;
    int flush_to;
    int owed;  // 1: the partials of the latest SpMV/update pair are not folded into the state yet
    // alpha of the latest update pass, written there by the workgroup of the
    // instance's first group (the update pass itself never reads it)
    double alpha;
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
    int tri_stretch;    // 1: stretch triangles installed
    int tri_shear;      // 1: shear triangles installed
    float bend_k;       // bending stiffness (dihedrals only exist when installed)
    // static-shape contact: thickness h, stiffness k, friction mu, eps_u; the
    // instance's planes (1 float4 each) and boxes (4 float4 each) in the shape table
    float contact_h, contact_k, contact_mu, contact_eps;
    int plane0, nplane, box0, nbox;
    // 1: pins are constraints (extension): a pinned particle's displacement
    // is fixed at target - x_n and eliminated from the system
    int pin_proj;
};

}  // namespace chysx_amd
