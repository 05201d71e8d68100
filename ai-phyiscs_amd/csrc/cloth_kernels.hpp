// SPDX-License-Identifier: Apache-2.0
//
// Device views and launch wrappers for the gfx950 cloth kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "cloth_device.hpp"

namespace chysx_amd {

// Element data (device pointers).  Vertex indices are particle indices.
struct DevMesh {
    const int* tri_v;         // 3 per triangle
    const float4* tri_dminv;  // (D00, D01, D10, D11) per triangle
    const float* tri_area;
    const float* tri_k;
    const int* spr_v;         // 2 per spring
    const float* spr_rest;
    const float* spr_k;
    const int* pin_v;
    const float* pin_tgt;     // 3 per pin
    // row-owned incidence (row space)
    const int* tri_inc_off;
    const int2* tri_inc;
    const int* spr_inc_off;
    const int2* spr_inc;
    const int* pin_inc_off;
    const int* pin_inc;
    // shear triangles (same vertices and incidence as tri_*)
    const float4* sh_dminv;
    const float* sh_area;
    const float* sh_k;
    // bending dihedrals and their row incidence
    const int4* dih_v;
    const float* dih_rest;
    const int* bend_inc_off;
    const int4* bend_inc;
    // static shapes: planes (n.xyz, d); boxes as 4 float4
    // (c.xyz h.x) (h.yz ex.xy) (ex.z ey.xyz) (ez.xyz 0)
    const float4* shapes;
};

// Wave-sliced ELL matrix: entry (s, k, lane) of component c lives at
// val[((slice_ptr[s] + k) * bs + c) * 64 + lane]; diag[(s * 9 + c) * 64 + lane];
// symmetric inverse diagonal minv[(s * 6 + c) * 64 + lane] with c over
// (xx, xy, xz, yy, yz, zz).  bs = 9 stores a block row-major; bs = 6 stores a
// symmetric block as (xx, xy, xz, yy, yz, zz) and needs sym; the diagonal
// blocks then use the same 6-float form, diag[(s * 6 + c) * 64 + lane].
struct DevMatrix {
    const int* slice_ptr;
    const int* col;
    const int* row_deg;
    float* val;
    float* diag;
    float* minv;
    // For an entry (i, j) with j < i: entry index (slice_ptr[s'] + k') * 64 +
    // lane' of the block (j, i) stored by row j (its transpose); -1 otherwise.
    const int* mir;
    // 1: lower blocks are read through mir (assembly writes upper blocks only)
    int sym;
    // floats per stored off-diagonal block: 9, or 6 when every off-diagonal
    // block is symmetric (no bending) and sym is set
    int bs;
    // The same columns and mirror indices as 16-bit offsets, when every one
    // fits (null otherwise): col16 = col - row (kNoCol16 for padding), mir16 =
    // mir - own entry index.  Read by the SpMV pass only.
    const int16_t* col16;
    const int16_t* mir16;
};
constexpr int kNoCol16 = -32768;

// Float offset in val of component 0 of entry index e (slot * 64 + lane).
__host__ __device__ inline size_t val_offset(int e, int bs) {
    return size_t(e & ~63) * size_t(bs) + size_t(e & 63);
}

// Solver vectors in padded row space, 3 components per row (AoS).
struct DevVectors {
    float* rhs;
    float* r;
    float* z;
    float* ap;
    float* p0;
    float* p1;
    double* x;         // the iterate dx (FP64)
    double* partials;  // 4 per slice
};

// Caller-owned buffers (particle space).
struct DevExt {
    float* pos;
    float* vel;
    const float* inv_mass;  // may be null (mass 1)
    const float* force;     // may be null
};

struct LaunchGeom {
    const SliceDesc* slices;
    int nslices;  // a multiple of kGroupSlices
    int ngroups;  // nslices / kGroupSlices
    // Batches: 1 for the groups of instances still iterating, else 0 (kept
    // by the start reduction and the advances), so the iteration passes drop
    // a finished cloth's workgroups after one load.  Null for a single instance.
    int* gactive;
};

// Workgroup shapes: `waves` in {1, 2, 4, 8, 16} (64 x waves lanes), SpMV
// `unroll` in {1, 2, 4}.  They change speed only, never results.
bool valid_waves(int waves);
bool valid_unroll(int unroll);

void launch_rest_lengths(const int* spr_v, const float* pos, float* rest, int nspring, hipStream_t s);
void launch_rest_shapes(const int* tri_v, const float* pos, float4* dminv, float* area, int ntri, float cos_t,
                        float sin_t, hipStream_t s);
// maxw: widest slice (blocks per row) of the problem, sizes the LDS accumulator.
void launch_assemble(const LaunchGeom& g, const DevMesh& m, const DevMatrix& A, const DevVectors& v,
                     const DevExt& e, const InstParams* params, float dt, int waves, int maxw, hipStream_t s);
// Rest dihedral angles (bending_rest_angle_kernel).
void launch_rest_angles(const int4* dih, const float* pos, float* rest, int n, hipStream_t s);
// The PCG passes run one workgroup of kGroupSlices wavefronts per 16-slice
// group and write one partial per group and slot.
// Residual pass: report == 0 starts a solve (true residual, r, z, p = z);
// report == 1 only measures the true residual of stopped instances.
void launch_pcg_residual(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                         int report, hipStream_t s);
// Start pass from the residual already in v.r: z, p = z and the same start
// partials as the residual pass with report == 0, without the product with A.
void launch_pcg_start_given_r(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                              hipStream_t s);
// SpMV pass (direction update, A p, p . A p; the true-residual check; or the
// flush of a deferred iterate term).  With st_out (not st): the pass first
// advances the state in st by the previous iteration's partials (when owed)
// and writes it to st_out, which its update pass and the next SpMV pass read.
// Without: the state in st is current (timing, autotuner).
void launch_pcg_spmv(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                     InstState* st_out, const InstParams* params, int unroll, hipStream_t s, int branch_free = 0,
                     int symmetric = 0);
// Update pass (alpha from the SpMV pass's partials; x, r, z; r . z, r . r).
void launch_pcg_update(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, InstState* st,
                       hipStream_t s);
// One wavefront per instance: the start or final reduction (PcgPhase).
// gactive (batches; may be null): per-group 1/0 flags of running instances,
// set at kPhaseStart and cleared when an instance finishes (by the SpMV pass
// or launch_pcg_advance).
void launch_pcg_reduce(InstState* st, const InstParams* params, const double* partials, int ninst, int ngroups,
                       int phase, hipStream_t s, int* gactive = nullptr);
// The advance of the previous SpMV/update pair from st_in into st_out (one
// workgroup per instance), for an SpMV pass launched without its own advance.
void launch_pcg_advance_copy(const InstState* st_in, InstState* st_out, const InstParams* params,
                             const double* partials, int ninst, int ngroups, hipStream_t s, int* gactive);
// One workgroup per instance: the owed advance of the state in st, in place
// (the end of a chunk of iterations, so the polled state is current).
void launch_pcg_advance(InstState* st, const InstParams* params, const double* partials, int ninst, int ngroups,
                        hipStream_t s, int* gactive = nullptr);
void launch_finalize(const LaunchGeom& g, const DevVectors& v, const DevExt& e, const InstParams* params, float dt,
                     const DevMesh& m, int waves, hipStream_t s);
void launch_area_mass(const int* vt_off, const int* vt_ids, const int* tri_v, const float* pos, float density_over_3,
                      float* inv_mass_out, int n, hipStream_t s);
// Plain SpMV y = A x (FP32 in, FP32 out) over the tiles, for benchmarks and tests.
void launch_spmv(const LaunchGeom& g, const DevMatrix& A, const float* x, float* y, int waves, hipStream_t s);

// HBM probe (stream_bench.hip): kind 0 copy, 1 triad, 2 / 3 the same with
// non-temporal accesses; returns mean ms per launch, *moved = bytes per launch.
double stream_bench(int kind, size_t bytes_per_buffer, int reps, hipStream_t s, double* moved);

}  // namespace chysx_amd
