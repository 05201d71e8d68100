// SPDX-License-Identifier: Apache-2.0
//
// gfx950 kernels of the implicit cloth step.
//
// Work decomposition: one 64-lane wavefront owns one 64-row slice of the
// wave-sliced ELL matrix, one lane per block row.  A workgroup is a group of
// WAVES consecutive slices (a launch parameter picked by the autotuner);
// nothing synchronises inside a workgroup, so slices of different instances
// may share one.  Workgroup ids are remapped so that each of the 8 XCDs walks
// a contiguous eighth of the slices, keeping neighbour gathers (rows +-1,
// +-W) inside one XCD's L2.
//
// Inner products: the PCG passes run one workgroup of 16 wavefronts per
// 16-slice group; each wave sums its lanes (butterfly), the group adds its
// 16 waves in order and writes one FP64 partial per slot.  The consumers
// (the update pass for p . Ap; the next SpMV pass, pcg_advance_kernel and
// pcg_reduce_kernel for the rest) add an instance's group partials in one
// fixed order.  The reduction order
// therefore depends neither on the launch shape nor on scheduling, and every
// result is bitwise reproducible.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "cloth_kernels.hpp"
#include "matrix_access.hpp"

namespace chysx_amd {

namespace {

constexpr float kMinLength = 1.0e-6f;  // spring_constraint.cu:35
constexpr float kMinNorm = 1.0e-6f;    // triangle_stretch_constraint.cu:39
constexpr double kGuard = 1.0e-37;     // pcg_solver.cu:421-423,451
constexpr size_t kAssembleLdsBytes = 64 * 1024;

// Slice owned by this wavefront (>= nslices: nothing to do).
template <int WAVES>
__device__ __forceinline__ int my_slice(int nslices) {
    const int nwg = (nslices + WAVES - 1) / WAVES;
    const int per = (nwg + 7) >> 3;
    const int b = blockIdx.x;
    const int wg = (b & 7) * per + (b >> 3);
    // wave-uniform: in a scalar register, so the slice descriptor and the
    // solver state are fetched with scalar loads and branched on uniformly
    return __builtin_amdgcn_readfirstlane(wg * WAVES + int(threadIdx.x >> 6));
}
inline int grid_for(int nslices, int waves) {
    const int nwg = (nslices + waves - 1) / waves;
    return 8 * ((nwg + 7) / 8);
}

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 ld3(const float* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(float* p, int i, F3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// Symmetric 3x3 (xx, xy, xz, yy, yz, zz).
// a + (m0 p.x + m1 p.y + m2 p.z) as one fixed chain of FMAs.  Written out so
// that every kernel variant (unroll, padding mode, storage) rounds the same
// way whatever the compiler schedules.
__device__ __forceinline__ float fma3(float a, float m0, float m1, float m2, F3 p) {
    return fmaf(m2, p.z, fmaf(m1, p.y, fmaf(m0, p.x, a)));
}
__device__ __forceinline__ double fma3d(double a, double m0, double m1, double m2, double x0, double x1, double x2) {
    return fma(m2, x2, fma(m1, x1, fma(m0, x0, a)));
}
// z + beta p_prev, component-wise FMAs
__device__ __forceinline__ F3 dir3(F3 z, F3 pp, float beta) {
    return F3{fmaf(pp.x, beta, z.x), fmaf(pp.y, beta, z.y), fmaf(pp.z, beta, z.z)};
}

struct S6 {
    float m[6];
};

// Workgroup of kGroupSlices wavefronts (one 16-slice group): wave sums, then
// the 16 waves in order; thread j < N writes slot slot[j] of `group`.  Every
// thread of the workgroup must call it.
template <int N>
__device__ __forceinline__ void group_partials(double (&v)[N], const int (&slot)[N], double* __restrict__ partials,
                                               int group, int ngroups) {
    __shared__ double sh[kGroupSlices][N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double a = v[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        v[j] = a;
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) sh[threadIdx.x >> 6][j] = v[j];
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0.0;
        for (int w = 0; w < kGroupSlices; ++w) t += sh[w][threadIdx.x];
        partials[(size_t)slot[threadIdx.x] * ngroups + group] = t;
    }
}

// ---------------------------------------------------------------------------
// Rest data (set-up).
// ---------------------------------------------------------------------------

// spring_rest_length_kernel (spring_constraint.cu:41-52).
__global__ void rest_length_kernel(const int* __restrict__ spr_v, const float* __restrict__ pos,
                                   float* __restrict__ rest, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const F3 d = ld3(pos, spr_v[2 * e]) - ld3(pos, spr_v[2 * e + 1]);
    rest[e] = sqrtf(dot(d, d));
}

// triangle_rest_shape_kernel (triangle_stretch_constraint.cu:50-110): 2-D rest
// frame, Dm^-1 (identity when det <= 1e-6), rest area, times R(theta)^-1.
__global__ void rest_shape_kernel(const int* __restrict__ tri_v, const float* __restrict__ pos,
                                  float4* __restrict__ dminv, float* __restrict__ area, int n, float ct, float st) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const F3 a = ld3(pos, tri_v[3 * t]), b = ld3(pos, tri_v[3 * t + 1]), c = ld3(pos, tri_v[3 * t + 2]);
    const F3 ab = b - a, ac = c - a;
    const float lab = sqrtf(dot(ab, ab));
    const float lac = sqrtf(dot(ac, ac));
    const F3 abh = (lab > kMinNorm) ? ab * (1.0f / lab) : f3(1.0f, 0.0f, 0.0f);
    const float uc = dot(abh, ac);
    const float vc2 = lac * lac - uc * uc;
    const float vc = (vc2 > 0.0f) ? sqrtf(vc2) : 0.0f;
    area[t] = 0.5f * lab * vc;
    const float det = lab * vc;
    float d00 = 1.0f, d01 = 0.0f, d10 = 0.0f, d11 = 1.0f;
    if (det > kMinNorm) {
        const float inv_det = 1.0f / (lab * vc - uc * 0.0f);
        d00 = vc * inv_det;
        d01 = -uc * inv_det;
        d10 = -0.0f * inv_det;
        d11 = lab * inv_det;
    }
    // Dm_inv * R^-1 with R^-1 = [[c, s], [-s, c]].
    dminv[t] = make_float4(d00 * ct + d01 * (-st), d00 * st + d01 * ct, d10 * ct + d11 * (-st), d10 * st + d11 * ct);
}

// Area-lumped mass: m_v = sum over incident triangles of density/3 * area,
// inv_mass = 1/m (0 when m <= 1e-12).  Restates scatter_triangle_mass_kernel +
// mass_to_inv_mass_kernel (cloth_simulator.cu:151-190) as a per-vertex gather.
__global__ void area_mass_kernel(const int* __restrict__ vt_off, const int* __restrict__ vt_ids,
                                 const int* __restrict__ tri_v, const float* __restrict__ pos, float d3,
                                 float* __restrict__ inv_mass, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float m = 0.0f;
    for (int q = vt_off[i]; q < vt_off[i + 1]; ++q) {
        const int t = vt_ids[q];
        const F3 x0 = ld3(pos, tri_v[3 * t]), x1 = ld3(pos, tri_v[3 * t + 1]), x2 = ld3(pos, tri_v[3 * t + 2]);
        const F3 e1 = x1 - x0, e2 = x2 - x0;
        const F3 cr = f3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
        m += d3 * (0.5f * sqrtf(dot(cr, cr)));
    }
    inv_mass[i] = (m > 1.0e-12f) ? (1.0f / m) : 0.0f;
}

// ---------------------------------------------------------------------------
// Assembly: one lane per block row.  The lane computes the row's right-hand
// side and every Hessian block of its row from the elements incident to it
// (pins, springs, stretch and shear triangles, bending dihedrals, inertia),
// writes the diagonal and its symmetric inverse, and accumulates off-diagonal
// blocks for its own ELL entries.  No atomics; the incidence order is fixed,
// so the result is deterministic.
//
// Element formulas restate:
//   pins      pin_constraint.cu:61-92
//   springs   spring_constraint.cu:76-151
//   stretch   triangle_stretch_constraint.cu:153-311 (shear: the same with
//             the pi/4-rotated frame, cloth_simulator.cu:423-449)
//   bending   bending_constraint.cu:50-246
//   inertia   cloth_simulator.cu:199-210
//   rhs       cloth_simulator.cu:83-142, written as m*(v/dt + g + f/m) - grad,
//             the same quantity as (m/dt^2)(x_tilde - x_n) - grad without the
//             FP32 cancellation of x_tilde - x_n.
// ---------------------------------------------------------------------------
// B d for accumulator entry k (block stride bs floats; bs == 6 holds the
// upper triangle xx, xy, xz, yy, yz, zz of a symmetric block).
__device__ __forceinline__ F3 block_mul(const float* __restrict__ acc, int k, int bs, F3 d) {
    const float* a = acc + k * bs * 64;
    if (bs == 6)
        return f3(a[0] * d.x + a[64] * d.y + a[128] * d.z, a[64] * d.x + a[192] * d.y + a[256] * d.z,
                  a[128] * d.x + a[256] * d.y + a[320] * d.z);
    return f3(a[0] * d.x + a[64] * d.y + a[128] * d.z, a[192] * d.x + a[256] * d.y + a[320] * d.z,
              a[384] * d.x + a[448] * d.y + a[512] * d.z);
}

// Adds sgn H into accumulator entry k (block stride bs floats).  bs == 6
// keeps the upper triangle (xx, xy, xz, yy, yz, zz) of a symmetric H.
__device__ __forceinline__ void add_block(float* __restrict__ acc, int k, const float (&H)[9], float sgn, int bs) {
    if (bs == 6) {
        float* a = acc + k * 6 * 64;
        a[0] += sgn * H[0];
        a[64] += sgn * H[1];
        a[128] += sgn * H[2];
        a[192] += sgn * H[4];
        a[256] += sgn * H[5];
        a[320] += sgn * H[8];
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[(k * 9 + c) * 64] += sgn * H[c];
    }
}

// One of three values, chosen by an index known only at run time, as a chain
// of selects over values.  (A per-thread array indexed at run time would be
// placed in scratch memory, which every launch of the kernel then sets up;
// selects over the elements of such an array are folded back into an indexed
// load, so the element sets here are named values, not arrays.)
__device__ __forceinline__ float sel3(int i, float v0, float v1, float v2) { return i == 0 ? v0 : i == 1 ? v1 : v2; }

// One Baraff-Witkin triangle (stretch or shear, by its D_m^-1) seen from its
// vertex a: gradient into `grad`, the three blocks H_ab into D (b == a or a
// repeated vertex) or the row accumulator (triangle_gradient_kernel and
// triangle_hessian_scatter_kernel, triangle_stretch_constraint.cu:153-311).
__device__ __forceinline__ void tri_element(const float4 Dm, float keff, int a, F3 x0, F3 x1, F3 x2, int packed,
                                            F3& grad, float (&D)[9], float* __restrict__ acc, int bs) {
    const F3 ea = x1 - x0, eb = x2 - x0;
    const F3 wu = ea * Dm.x + eb * Dm.z;
    const F3 wv = ea * Dm.y + eb * Dm.w;
    float lu = sqrtf(dot(wu, wu));
    float lv = sqrtf(dot(wv, wv));
    if (lu < kMinNorm) lu = kMinNorm;
    if (lv < kMinNorm) lv = kMinNorm;
    const float inv_lu = 1.0f / lu, inv_lv = 1.0f / lv;
    const float cu = lu - 1.0f, cv = lv - 1.0f;
    const F3 uh = wu * inv_lu, vh = wv * inv_lv;
    const float dwu0 = -Dm.x - Dm.z, dwu1 = Dm.x, dwu2 = Dm.z;
    const float dwv0 = -Dm.y - Dm.w, dwv1 = Dm.y, dwv2 = Dm.w;
    const float dwu_a = sel3(a, dwu0, dwu1, dwu2), dwv_a = sel3(a, dwv0, dwv1, dwv2);
    grad = grad + (uh * (cu * dwu_a) + vh * (cv * dwv_a)) * keff;
    // H_ab = dwu_a dwu_b Mu + dwv_a dwv_b Mv with
    // Mu = keff uh uh^T + [cu > 0] keff cu / lu (I - uh uh^T), Mv likewise.
    const float su = (cu > 0.0f) ? keff * cu * inv_lu : 0.0f;
    const float sv = (cv > 0.0f) ? keff * cv * inv_lv : 0.0f;
    const float qu = keff - su, qv = keff - sv;
    float Mu[6], Mv[6];
    Mu[0] = qu * uh.x * uh.x + su;
    Mu[1] = qu * uh.x * uh.y;
    Mu[2] = qu * uh.x * uh.z;
    Mu[3] = qu * uh.y * uh.y + su;
    Mu[4] = qu * uh.y * uh.z;
    Mu[5] = qu * uh.z * uh.z + su;
    Mv[0] = qv * vh.x * vh.x + sv;
    Mv[1] = qv * vh.x * vh.y;
    Mv[2] = qv * vh.x * vh.z;
    Mv[3] = qv * vh.y * vh.y + sv;
    Mv[4] = qv * vh.y * vh.z;
    Mv[5] = qv * vh.z * vh.z + sv;
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
        const float cu_ab = dwu_a * sel3(bb, dwu0, dwu1, dwu2), cv_ab = dwv_a * sel3(bb, dwv0, dwv1, dwv2);
        float H[9];
        H[0] = cu_ab * Mu[0] + cv_ab * Mv[0];
        H[1] = cu_ab * Mu[1] + cv_ab * Mv[1];
        H[2] = cu_ab * Mu[2] + cv_ab * Mv[2];
        H[4] = cu_ab * Mu[3] + cv_ab * Mv[3];
        H[5] = cu_ab * Mu[4] + cv_ab * Mv[4];
        H[8] = cu_ab * Mu[5] + cv_ab * Mv[5];
        H[3] = H[1];
        H[6] = H[2];
        H[7] = H[5];
        const int kb = (packed >> (10 * bb)) & 1023;
        if (bb == a || kb == 1023) {
#pragma unroll
            for (int c = 0; c < 9; ++c) D[c] += H[c];
        } else {
            add_block(acc, kb, H, 1.0f, bs);
        }
    }
}

// compute_dihedral_frame (bending_constraint.cu:50-127): dihedral angle theta
// and a_i = d theta / d x_i; invalid for a collapsed hinge or a wing with no
// height (floor 1e-12).
struct DihFrame {
    float theta;
    F3 a0, a1, a2, a3;  // gradient of theta with respect to vertex 0..3
    bool valid;
    // a_i for an index known at compile time (after unrolling)
    __device__ __forceinline__ F3 a(int i) const { return i == 0 ? a0 : i == 1 ? a1 : i == 2 ? a2 : a3; }
    // a_i for an index known only at run time, as a sum with one-hot weights
    // (exact for a valid frame, whose vectors are finite): a select here is
    // folded into a load at a run-time offset, which puts the frame in
    // scratch memory.
    __device__ __forceinline__ F3 a_dyn(int i) const {
        const float w0 = i == 0 ? 1.0f : 0.0f, w1 = i == 1 ? 1.0f : 0.0f;
        const float w2 = i == 2 ? 1.0f : 0.0f, w3 = i == 3 ? 1.0f : 0.0f;
        return a0 * w0 + a1 * w1 + a2 * w2 + a3 * w3;
    }
};
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ DihFrame dihedral_frame(F3 v0, F3 v1, F3 v2, F3 v3) {
    constexpr float kFloor = 1.0e-12f;
    DihFrame F;
    F.valid = false;
    F.theta = 0.0f;
    F.a0 = F.a1 = F.a2 = F.a3 = f3(0.0f, 0.0f, 0.0f);
    const F3 e10 = v1 - v0, e20 = v2 - v0, e30 = v3 - v0;
    const float l = sqrtf(dot(e10, e10));
    if (l < kFloor) return F;
    const float inv_l = 1.0f / l;
    const F3 e10h = e10 * inv_l;
    const F3 n1r = cross(e20, e10), n2r = cross(e10, e30);
    const float n1l = sqrtf(dot(n1r, n1r)), n2l = sqrtf(dot(n2r, n2r));
    if (n1l < kFloor || n2l < kFloor) return F;
    const F3 n1 = n1r * (1.0f / n1l), n2 = n2r * (1.0f / n2l);
    float sin_t = dot(cross(n1, n2), e10h);
    sin_t = fminf(fmaxf(sin_t, -1.0f), 1.0f);
    const float cos_t = dot(n1, n2);
    F.theta = atan2f(sin_t, cos_t);
    const float inv_l2 = inv_l * inv_l;
    const float om1 = dot(e10, e20) * inv_l2, om2 = dot(e10, e30) * inv_l2;
    const F3 p1 = e20 - e10 * om1, p2 = e30 - e10 * om2;
    const float h1 = sqrtf(dot(p1, p1)), h2 = sqrtf(dot(p2, p2));
    if (h1 < kFloor || h2 < kFloor) return F;
    const float ih1 = 1.0f / h1, ih2 = 1.0f / h2;
    F.a0 = n1 * ((om1 - 1.0f) * ih1) + n2 * ((om2 - 1.0f) * ih2);
    F.a1 = n1 * (-om1 * ih1) + n2 * (-om2 * ih2);
    F.a2 = n1 * ih1 + n2 * 0.0f;
    F.a3 = n1 * 0.0f + n2 * ih2;
    F.valid = true;
    return F;
}

// box_signed_dist (static_contact.cu:55-99): signed distance of x to an
// oriented box and the unit outward normal (towards the closest face when
// inside).
__device__ __forceinline__ float box_sdf(const float4* __restrict__ b, F3 x, F3& n) {
    const float4 b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
    const F3 c = f3(b0.x, b0.y, b0.z), he = f3(b0.w, b1.x, b1.y);
    const F3 ex = f3(b1.z, b1.w, b2.x), ey = f3(b2.y, b2.z, b2.w), ez = f3(b3.x, b3.y, b3.z);
    const F3 rel = x - c;
    const float qx = dot(rel, ex), qy = dot(rel, ey), qz = dot(rel, ez);
    const float dx = fabsf(qx) - he.x, dy = fabsf(qy) - he.y, dz = fabsf(qz) - he.z;
    if (dx <= 0.0f && dy <= 0.0f && dz <= 0.0f) {
        if (dx >= dy && dx >= dz) {
            n = (qx >= 0.0f) ? ex : ex * -1.0f;
            return dx;
        }
        if (dy >= dz) {
            n = (qy >= 0.0f) ? ey : ey * -1.0f;
            return dy;
        }
        n = (qz >= 0.0f) ? ez : ez * -1.0f;
        return dz;
    }
    const float cx = dx > 0.0f ? dx : 0.0f, cy = dy > 0.0f ? dy : 0.0f, cz = dz > 0.0f ? dz : 0.0f;
    const float dist = sqrtf(cx * cx + cy * cy + cz * cz);
    const float inv = 1.0f / fmaxf(dist, 1.0e-30f);
    const float sx = qx > 0.0f ? cx : -cx, sy = qy > 0.0f ? cy : -cy, sz = qz > 0.0f ? cz : -cz;
    n = ex * (sx * inv) + ey * (sy * inv) + ez * (sz * inv);
    return dist;
}

// ipc_f1_sf_over_x (friction.cuh:28-36).
__device__ __forceinline__ float f1_sf_over_x(float u, float eps) {
    return (u > eps) ? 1.0f / u : (-u / eps + 2.0f) / eps;
}

// bending_rest_angle_kernel (bending_constraint.cu:136-152).
__global__ void rest_angle_kernel(const int4* __restrict__ dih, const float* __restrict__ pos,
                                  float* __restrict__ rest, int n) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const int4 q = dih[d];
    const DihFrame F = dihedral_frame(ld3(pos, q.x), ld3(pos, q.y), ld3(pos, q.z), ld3(pos, q.w));
    rest[d] = F.valid ? F.theta : 0.0f;
}

// Row-owned assembly: the lane of row i evaluates every element touching i
// (pins, springs, stretch and shear triangles, bending dihedrals) in a fixed
// incidence order and adds the row's blocks into an accumulator with the
// CSR-slot layout of its slice, stride 64 floats.  With LDS = true the
// accumulator is the wave's slab of LDS (maxw blocks per lane) and each
// block is written to HBM once at the end; otherwise it is the matrix row
// itself (zeroed first).  No atomics: every slot has one owner lane.
template <int WAVES, bool LDS>
__global__ __launch_bounds__(64 * WAVES) void assemble_kernel(LaunchGeom g, DevMesh m, DevMatrix A, DevVectors v,
                                                             DevExt e, const InstParams* __restrict__ params,
                                                             float dt, int maxw) {
    extern __shared__ float lds_acc[];
    const int s = my_slice<WAVES>(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63;
    if (lane >= sd.nvalid) return;
    const int row = s * 64 + lane;
    const int p = sd.part0 + lane;
    const InstParams pr = params[sd.inst];
    const int sb = A.slice_ptr[s];
    const int deg = A.row_deg[row];
    const int bs = A.bs;
    float* __restrict__ vrow = A.val + (size_t)sb * bs * 64 + lane;
    float* __restrict__ acc = LDS ? lds_acc + (size_t)(threadIdx.x >> 6) * maxw * bs * 64 + lane : vrow;
    for (int k = 0; k < deg; ++k)
        for (int c = 0; c < bs; ++c) acc[(k * bs + c) * 64] = 0.0f;

    const F3 xi = ld3(e.pos, p);
    const F3 vi = ld3(e.vel, p);
    // mass_from_inv_mass_kernel (cloth_simulator.cu:66-77)
    float mass = 1.0f;
    if (e.inv_mass != nullptr) {
        const float w = e.inv_mass[p];
        mass = (w > 1.0e-12f) ? (1.0f / w) : 1.0e8f;
    }
    F3 ac = f3(pr.gx, pr.gy, pr.gz);
    if (pr.force_active && e.force != nullptr) {
        const float im = (mass > 0.0f) ? (1.0f / mass) : 0.0f;
        ac = ac + ld3(e.force, p) * im;
    }
    const float inv_dt = 1.0f / dt;
    const F3 b = (vi * inv_dt + ac) * mass;

    float D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    F3 grad = f3(0.0f, 0.0f, 0.0f);

    // pins (penalties; with pin projection they are constraints, applied below)
    for (int q = m.pin_inc_off[row]; q < (pr.pin_proj ? m.pin_inc_off[row] : m.pin_inc_off[row + 1]); ++q) {
        const int c = m.pin_inc[q];
        const F3 tg = ld3(m.pin_tgt, c);
        grad = grad + (xi - tg) * pr.pin_k;
        D[0] += pr.pin_k;
        D[4] += pr.pin_k;
        D[8] += pr.pin_k;
    }
    // springs
    if (pr.springs_active) {
        for (int q = m.spr_inc_off[row]; q < m.spr_inc_off[row + 1]; ++q) {
            const int2 ent = m.spr_inc[q];
            const int ei = ent.x >> 1, side = ent.x & 1;
            const int va = m.spr_v[2 * ei], vb = m.spr_v[2 * ei + 1];
            const F3 d = ld3(e.pos, va) - ld3(e.pos, vb);
            float l = sqrtf(dot(d, d));
            if (l < kMinLength) l = kMinLength;
            const float inv_l = 1.0f / l;
            const F3 dh = d * inv_l;
            const float ke = m.spr_k[ei], L = m.spr_rest[ei];
            const F3 ge = dh * (ke * (l - L));
            grad = side ? grad - ge : grad + ge;
            const float a = ke * L * inv_l;
            const float iso = (l < L) ? 0.0f : (ke - a);
            float H[9];
            H[0] = a * (dh.x * dh.x) + iso;
            H[1] = a * (dh.x * dh.y);
            H[2] = a * (dh.x * dh.z);
            H[3] = H[1];
            H[4] = a * (dh.y * dh.y) + iso;
            H[5] = a * (dh.y * dh.z);
            H[6] = H[2];
            H[7] = H[5];
            H[8] = a * (dh.z * dh.z) + iso;
#pragma unroll
            for (int c = 0; c < 9; ++c) D[c] += H[c];
            const int ko = ent.y;
            if (ko == 1023) {
#pragma unroll
                for (int c = 0; c < 9; ++c) D[c] -= H[c];
            } else {
                add_block(acc, ko, H, -1.0f, A.bs);
            }
        }
    }
    // stretch, then shear triangles (cloth_simulator.cu:922-923, 1136-1137)
    for (int q = m.tri_inc_off[row]; q < m.tri_inc_off[row + 1]; ++q) {
        const int2 ent = m.tri_inc[q];
        const int ti = ent.x >> 2, a = ent.x & 3;
        const F3 x0 = ld3(e.pos, m.tri_v[3 * ti]);
        const F3 x1 = ld3(e.pos, m.tri_v[3 * ti + 1]);
        const F3 x2 = ld3(e.pos, m.tri_v[3 * ti + 2]);
        if (pr.tri_stretch)
            tri_element(m.tri_dminv[ti], m.tri_area[ti] * m.tri_k[ti], a, x0, x1, x2, ent.y, grad, D, acc, bs);
        if (pr.tri_shear)
            tri_element(m.sh_dminv[ti], m.sh_area[ti] * m.sh_k[ti], a, x0, x1, x2, ent.y, grad, D, acc, bs);
    }
    // bending dihedrals: gradient k (theta - theta_rest) a_i, Gauss-Newton
    // blocks k a_i a_j^T (bending_constraint.cu:175-246)
    for (int q = m.bend_inc_off[row]; q < m.bend_inc_off[row + 1]; ++q) {
        const int4 ent = m.bend_inc[q];
        const int di = ent.x >> 2, a = ent.x & 3;
        const int4 dv = m.dih_v[di];
        const DihFrame F = dihedral_frame(ld3(e.pos, dv.x), ld3(e.pos, dv.y), ld3(e.pos, dv.z), ld3(e.pos, dv.w));
        if (!F.valid) continue;
        const float kb = pr.bend_k;
        const F3 fa = F.a_dyn(a);
        grad = grad + fa * (kb * (F.theta - m.dih_rest[di]));
        const F3 ka = fa * kb;
        const int kk[4] = {ent.y & 0xFFFF, (ent.y >> 16) & 0xFFFF, ent.z & 0xFFFF, (ent.z >> 16) & 0xFFFF};
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const F3 ab = F.a(bb);
            float H[9];
            H[0] = ka.x * ab.x;
            H[1] = ka.x * ab.y;
            H[2] = ka.x * ab.z;
            H[3] = ka.y * ab.x;
            H[4] = ka.y * ab.y;
            H[5] = ka.y * ab.z;
            H[6] = ka.z * ab.x;
            H[7] = ka.z * ab.y;
            H[8] = ka.z * ab.z;
            if (bb == a || kk[bb] == 0xFFFF) {
#pragma unroll
                for (int c = 0; c < 9; ++c) D[c] += H[c];
            } else {
                add_block(acc, kk[bb], H, 1.0f, A.bs);
            }
        }
    }
    // Pin projection (extension; no reference counterpart, whose pins are
    // penalties, pin_constraint.h:5-15): a pinned particle's displacement is
    // d = target - x_n.  A free row moves its blocks to pinned columns into
    // the right-hand side (rhs -= B d) and drops them; a pinned row drops
    // every block and becomes I dx = d, with the iterate started at d, so its
    // residual is 0 throughout.  Both halves of a coupling are dropped (the
    // pinned row's own block and the free row's), so the symmetric storages'
    // mirrors agree.
    bool pinned = false;
    F3 dpin = f3(0.0f, 0.0f, 0.0f), pcorr = f3(0.0f, 0.0f, 0.0f);
    if (pr.pin_proj) {
        const int q0 = m.pin_inc_off[row];
        if (m.pin_inc_off[row + 1] > q0) {
            pinned = true;
            dpin = ld3(m.pin_tgt, m.pin_inc[q0]) - xi;
        }
        const int* colr = A.col + (size_t)sb * 64 + lane;
        for (int k = 0; k < deg; ++k) {
            const int c = colr[k * 64];
            if (c < 0) continue;
            const int qc = m.pin_inc_off[c];
            const bool pc = m.pin_inc_off[c + 1] > qc;
            if (!pinned && !pc) continue;
            if (!pinned) {
                const int pi = m.pin_inc[qc];
                const F3 dc = ld3(m.pin_tgt, pi) - ld3(e.pos, m.pin_v[pi]);
                pcorr = pcorr + block_mul(acc, k, bs, dc);
            }
            for (int cc = 0; cc < bs; ++cc) acc[(k * bs + cc) * 64] = 0.0f;
        }
    }
    if (LDS) {
        const int* colr = A.col + (size_t)sb * 64 + lane;
        for (int k = 0; k < deg; ++k) {
            if (A.sym && colr[k * 64] < row) continue;  // read through the upper mirror
            for (int c = 0; c < bs; ++c) vrow[(k * bs + c) * 64] = acc[(k * bs + c) * 64];
        }
    }
    // static-shape contact at x_n (detect_kernel + scatter_gradient_kernel,
    // static_contact.cu:107-213): the deepest plane / box penetration below
    // the thickness h; penalty force k depth n and, with friction, the
    // lagged tangential term alpha u_t on the gradient.
    float cdepth = 0.0f, calpha = 0.0f;
    F3 cn = f3(0.0f, 0.0f, 0.0f);
    if (pr.contact_k > 0.0f && pr.nplane + pr.nbox > 0) {
        for (int q = 0; q < pr.nplane; ++q) {
            const float4 pl = m.shapes[pr.plane0 + q];
            const F3 nq = f3(pl.x, pl.y, pl.z);
            const float depth = pr.contact_h - (dot(nq, xi) + pl.w);
            if (depth > cdepth) {
                cdepth = depth;
                cn = nq;
            }
        }
        for (int q = 0; q < pr.nbox; ++q) {
            F3 nq;
            const float depth = pr.contact_h - box_sdf(m.shapes + pr.box0 + 4 * q, xi, nq);
            if (depth > cdepth) {
                cdepth = depth;
                cn = nq;
            }
        }
        if (cdepth > 0.0f) {
            grad = grad + cn * (-pr.contact_k * cdepth);
            if (pr.contact_mu > 0.0f) {
                const F3 dxv = vi * dt;
                const F3 ut = dxv - cn * dot(cn, dxv);
                calpha = pr.contact_mu * (pr.contact_k * cdepth) * f1_sf_over_x(sqrtf(dot(ut, ut)), pr.contact_eps);
                grad = grad + ut * (-calpha);
            }
        }
    }
    // inertia (add_inertia_diag_kernel, cloth_simulator.cu:199-210)
    const float mi = mass * (1.0f / (dt * dt));
    D[0] += mi;
    D[4] += mi;
    D[8] += mi;

    F3 rr = b - grad;
    if (cdepth > 0.0f) {
        // bake_diag_kernel (static_contact.cu:232-300): k n n^T + alpha (I - n n^T)
        const float kc = pr.contact_k;
        D[0] += kc * cn.x * cn.x + calpha * (1.0f - cn.x * cn.x);
        D[1] += kc * cn.x * cn.y - calpha * cn.x * cn.y;
        D[2] += kc * cn.x * cn.z - calpha * cn.x * cn.z;
        D[3] += kc * cn.x * cn.y - calpha * cn.x * cn.y;
        D[4] += kc * cn.y * cn.y + calpha * (1.0f - cn.y * cn.y);
        D[5] += kc * cn.y * cn.z - calpha * cn.y * cn.z;
        D[6] += kc * cn.x * cn.z - calpha * cn.x * cn.z;
        D[7] += kc * cn.y * cn.z - calpha * cn.y * cn.z;
        D[8] += kc * cn.z * cn.z + calpha * (1.0f - cn.z * cn.z);
        // apply_coulomb_friction_kernel (static_contact.cu:325-380): the
        // tangential part of F0 = rhs - k depth n onto the cone of radius
        // mu max(0, -F0 . n).
        if (pr.contact_mu > 0.0f) {
            const F3 F0 = rr - cn * (kc * cdepth);
            const float f0n = dot(F0, cn);
            const float fn = fmaxf(0.0f, -f0n);
            if (fn > 0.0f) {
                const F3 F0t = F0 - cn * f0n;
                const float ft2 = dot(F0t, F0t);
                if (ft2 > 0.0f) {
                    const float ft = sqrtf(ft2);
                    const float cone = pr.contact_mu * fn;
                    rr = rr - F0t * ((ft <= cone) ? 1.0f : cone / ft);
                }
            }
        }
    }
    if (pinned) {
        rr = dpin;
#pragma unroll
        for (int c = 0; c < 9; ++c) D[c] = (c == 0 || c == 4 || c == 8) ? 1.0f : 0.0f;
        v.x[3 * row] = double(dpin.x);
        v.x[3 * row + 1] = double(dpin.y);
        v.x[3 * row + 2] = double(dpin.z);
    } else if (pr.pin_proj) {
        rr = rr - pcorr;
    }
    st3(v.rhs, row, rr);
    if (bs == 6) {  // symmetric diagonal blocks (no bending): upper triangle
        float* dg = A.diag + (size_t)s * 6 * 64 + lane;
        dg[0] = D[0];
        dg[64] = D[1];
        dg[128] = D[2];
        dg[192] = D[4];
        dg[256] = D[5];
        dg[320] = D[8];
    } else {
        float* dg = A.diag + (size_t)s * 9 * 64 + lane;
#pragma unroll
        for (int c = 0; c < 9; ++c) dg[c * 64] = D[c];
    }
    // Inverse of the symmetric diagonal block by adjugate / determinant, no
    // guard (math::inverse, matrix.cuh:219-235; residual_m_inv_kernel,
    // pcg_solver.cu:29-51).
    const float a00 = D[4] * D[8] - D[5] * D[7];
    const float a01 = -D[1] * D[8] + D[2] * D[7];
    const float a02 = D[1] * D[5] - D[2] * D[4];
    const float a11 = D[0] * D[8] - D[2] * D[6];
    const float a12 = -D[0] * D[5] + D[2] * D[3];
    const float a22 = D[0] * D[4] - D[1] * D[3];
    const float a10 = -D[3] * D[8] + D[5] * D[6];
    const float a20 = D[3] * D[7] - D[4] * D[6];
    const float det = D[0] * a00 + D[1] * a10 + D[2] * a20;
    const float id = 1.0f / det;
    float* mv = A.minv + (size_t)s * 6 * 64 + lane;
    mv[0 * 64] = a00 * id;
    mv[1 * 64] = a01 * id;
    mv[2 * 64] = a02 * id;
    mv[3 * 64] = a11 * id;
    mv[4 * 64] = a12 * id;
    mv[5 * 64] = a22 * id;
}

__device__ __forceinline__ F3 minv_mul(const DevMatrix& A, int s, int lane, F3 r) {
    const float* mv = A.minv + (size_t)s * 6 * 64 + lane;
    const float m0 = mv[0], m1 = mv[64], m2 = mv[128], m3 = mv[192], m4 = mv[256], m5 = mv[320];
    return f3(fma3(0.0f, m0, m1, m2, r), fma3(0.0f, m1, m3, m4, r), fma3(0.0f, m2, m4, m5, r));
}

constexpr int kReduceThreads = 64 * kGroupSlices;

// Sums slots slot[0..NS) of an instance's group partials (groups g0 ..
// g0 + ng - 1, `ld` groups per slot).  Thread t adds groups t, t + T, t + 2T,
// ... in order; the waves' sums are then combined in wave order, so the
// result is one fixed sum whatever KU is (every workgroup of a pass computes
// the same alpha or state).  All NS x KU loads of a round are issued before
// any of them is added.  The total of slot slot[j] lands in out[j] in every
// thread: after one barrier each thread adds the 16 wave sums itself, in wave
// order.  One call per kernel (the wave sums' LDS is not reused).
//
// block_sum_issue issues the first round's loads on its own and
// block_sum_finish does the rest, so a caller can issue other loads while
// the first round is in flight; block_sum is the two back to back.
template <int NS, int KU>
__device__ __forceinline__ void block_sum_round(const double* __restrict__ partials, int ld, int s0, int ns,
                                                const int (&slot)[NS], int q0, double (&t)[NS][KU]) {
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const double* pj = partials + (size_t)slot[j] * ld + s0;
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            // unconditional, clamped load (no branch): out-of-range
            // entries re-read the last partial and are not added
            const int q = q0 + u * kReduceThreads;
            t[j][u] = pj[q < ns ? q : ns - 1];
        }
    }
}

template <int NS, int KU>
__device__ __forceinline__ void block_sum_issue(const double* __restrict__ partials, int ld, int s0, int ns,
                                                const int (&slot)[NS], double (&t)[NS][KU]) {
    if (ns > 0) block_sum_round<NS, KU>(partials, ld, s0, ns, slot, int(threadIdx.x), t);
}

template <int NS, int KU>
__device__ __forceinline__ void block_sum_finish(const double* __restrict__ partials, int ld, int s0, int ns,
                                                 const int (&slot)[NS], const double (&t0)[NS][KU],
                                                 double (&out)[NS]) {
    double acc[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) acc[j] = 0.0;
    auto add = [&](int q0, const double (&t)[NS][KU]) {
#pragma unroll
        for (int j = 0; j < NS; ++j)
#pragma unroll
            for (int u = 0; u < KU; ++u)
                if (q0 + u * kReduceThreads < ns) acc[j] += t[j][u];
    };
    // The first round is straight-line code: a loop header here would make
    // the compiler wait for every load the caller has in flight.
    if (ns > 0) {
        add(int(threadIdx.x), t0);
        for (int q0 = int(threadIdx.x) + KU * kReduceThreads; q0 < ns; q0 += KU * kReduceThreads) {
            double t[NS][KU];
            block_sum_round<NS, KU>(partials, ld, s0, ns, slot, q0, t);
            add(q0, t);
        }
    }
    __shared__ double sh[kReduceThreads / 64][NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) sh[threadIdx.x >> 6][j] = acc[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kReduceThreads / 64; ++w) t += sh[w][j];
        out[j] = t;
    }
}

template <int NS, int KU>
__device__ __forceinline__ void block_sum(const double* __restrict__ partials, int ld, int s0, int ns,
                                          const int (&slot)[NS], double (&out)[NS]) {
    double t[NS][KU];
    block_sum_issue<NS, KU>(partials, ld, s0, ns, slot, t);
    block_sum_finish<NS, KU>(partials, ld, s0, ns, slot, t, out);
}

// Sum of ng group partials p[0 .. ng) by one wavefront in a fixed order:
// lane l adds groups l, l + 64, l + 128, ..., then the lanes combine by a
// fixed butterfly; every lane gets the total.  The first 1024 groups are
// loaded by straight-line code (clamped, unconditional loads).
__device__ __forceinline__ double wave_group_sum(const double* __restrict__ p, int ng) {
    constexpr int KU = 16;
    const int lane = threadIdx.x & 63;
    double a = 0.0;
    auto round = [&](int q0) {
        double t[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int q = q0 + u * 64;
            t[u] = p[q < ng ? q : ng - 1];
        }
#pragma unroll
        for (int u = 0; u < KU; ++u)
            if (q0 + u * 64 < ng) a += t[u];
    };
    if (ng > 0) {
        round(lane);
        for (int q0 = lane + KU * 64; q0 < ng; q0 += KU * 64) round(q0);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    return a;
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// True residual of one row: b - A x with FP64 products and sums (x is FP64).
// Returns it and, unless `report_only`, stores r, z = M^-1 r and p = z.
// Partials: [0] |r|^2, [1] |b|^2, [2] r . z.
// Entries per load batch of residual_row: two keeps both the residual kernel
// and the SpMV kernel (whose check mode runs residual_row) at 64 VGPRs or
// fewer, 8 waves per SIMD.
constexpr int kResidualBatch = 2;

// nlow: the slice's nlower (entries at positions >= nlow hold no lower block
// in any lane, so their mirror index is not read).
__device__ __forceinline__ void residual_row(const DevMatrix& A, const DevVectors& v, float* __restrict__ pcur,
                                             int row, int nlow, bool report_only, double (&part)[4]) {
    const int s = row >> 6, lane = row & 63;
    const double* x = v.x;
    float dm[9];
    load_diag(A, s, lane, dm);
    const double x0 = x[3 * row], x1 = x[3 * row + 1], x2 = x[3 * row + 2];
    double a0 = fma3d(0.0, dm[0], dm[1], dm[2], x0, x1, x2);
    double a1 = fma3d(0.0, dm[3], dm[4], dm[5], x0, x1, x2);
    double a2 = fma3d(0.0, dm[6], dm[7], dm[8], x0, x1, x2);
    const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
    const int* colp = A.col + (size_t)sb * 64 + lane;
    const int* mirp = A.mir + (size_t)sb * 64 + lane;
    const float* vp = A.val + (size_t)sb * A.bs * 64 + lane;
    // Batches of entries in three phases (the column and mirror indices, then
    // every block and x gather, then the FP64 FMAs in entry order, so the sums
    // are those of one entry at a time), so a batch costs two memory round
    // trips instead of three per entry.
    constexpr int RB = kResidualBatch;
    for (int k0 = 0; k0 < w; k0 += RB) {
        int c[RB], mi[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            c[u] = -1;
            mi[u] = 0;
            if (k0 + u < w) {  // wave-uniform
                c[u] = colp[(k0 + u) * 64];
                if (A.sym && k0 + u < nlow) mi[u] = mirp[(k0 + u) * 64];  // wave-uniform
            }
        }
        float m[RB][9];
        double y[RB][3];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            if (c[u] >= 0) {
                load_block_at(A, vp + (k0 + u) * A.bs * 64, mi[u], c[u], row, m[u]);
                y[u][0] = x[3 * c[u]];
                y[u][1] = x[3 * c[u] + 1];
                y[u][2] = x[3 * c[u] + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            if (c[u] >= 0) {
                a0 = fma3d(a0, m[u][0], m[u][1], m[u][2], y[u][0], y[u][1], y[u][2]);
                a1 = fma3d(a1, m[u][3], m[u][4], m[u][5], y[u][0], y[u][1], y[u][2]);
                a2 = fma3d(a2, m[u][6], m[u][7], m[u][8], y[u][0], y[u][1], y[u][2]);
            }
        }
    }
    const F3 bf = ld3(v.rhs, row);
    const double r0 = double(bf.x) - a0, r1 = double(bf.y) - a1, r2 = double(bf.z) - a2;
    part[0] = fma3d(0.0, r0, r1, r2, r0, r1, r2);
    part[1] = fma3d(0.0, bf.x, bf.y, bf.z, bf.x, bf.y, bf.z);
    if (!report_only) {
        const F3 rf = f3(float(r0), float(r1), float(r2));
        const F3 zf = minv_mul(A, s, lane, rf);
        st3(v.r, row, rf);
        st3(v.z, row, zf);
        st3(pcur, row, zf);
        part[2] = fma3d(0.0, rf.x, rf.y, rf.z, zf.x, zf.y, zf.z);
    }
}

// Residual pass.  report == 0: start a solve (every instance); report == 1:
// record the true residual of stopped instances only.
// Returns and early exits below are taken by whole workgroups: a group's 16
// slices belong to one instance, and nslices is a multiple of 16.
__global__ __launch_bounds__(64 * kGroupSlices) void pcg_residual_kernel(LaunchGeom g, DevMatrix A, DevVectors v,
                                                                        const InstState* __restrict__ st,
                                                                        int report) {
    const int s = my_slice<kGroupSlices>(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const InstState& S = st[sd.inst];
    if (report && S.status != kStopped) return;
    float* pcur = S.parity ? v.p1 : v.p0;
    const int lane = threadIdx.x & 63;
    double part[3] = {0.0, 0.0, 0.0};
    if (lane < sd.nvalid) {
        double q[4] = {0.0, 0.0, 0.0, 0.0};
        residual_row(A, v, pcur, s * 64 + lane, sd.nlower, report != 0, q);
        part[0] = q[0];
        part[1] = q[1];
        part[2] = q[2];
    }
    const int sl[3] = {0, 1, 2};
    group_partials<3>(part, sl, v.partials, s / kGroupSlices, g.ngroups);
}

// Start pass from a residual the caller already holds in v.r (FP32, as the
// residual pass stores it): z = M^-1 r, p = z and the start partials
// [|r|^2, |b|^2, r . z] as the residual pass forms them, without the product
// with A.  The learned warm start's line search has r = r_b - alpha A d in
// hand; the solve's true-residual check before it stops is unchanged.
__global__ __launch_bounds__(64 * kGroupSlices) void pcg_start_given_r_kernel(LaunchGeom g, DevMatrix A,
                                                                             DevVectors v,
                                                                             const InstState* __restrict__ st) {
    const int s = my_slice<kGroupSlices>(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const InstState& S = st[sd.inst];
    float* pcur = S.parity ? v.p1 : v.p0;
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    double part[3] = {0.0, 0.0, 0.0};
    if (lane < sd.nvalid) {
        const F3 rf = ld3(v.r, row), bf = ld3(v.rhs, row);
        const F3 zf = minv_mul(A, s, lane, rf);
        st3(v.z, row, zf);
        st3(pcur, row, zf);
        part[0] = fma3d(0.0, rf.x, rf.y, rf.z, rf.x, rf.y, rf.z);
        part[1] = fma3d(0.0, bf.x, bf.y, bf.z, bf.x, bf.y, bf.z);
        part[2] = fma3d(0.0, rf.x, rf.y, rf.z, zf.x, zf.y, zf.z);
    }
    const int sl[3] = {0, 1, 2};
    group_partials<3>(part, sl, v.partials, s / kGroupSlices, g.ngroups);
}

// ---------------------------------------------------------------------------
// PCG SpMV pass.  It first advances the solver state by the previous
// iteration's partials (advance_in_pass), then works in the mode that state
// gives.  Mode kModeIter: p = z + beta p_prev (beta = 0 right after a
// (re)start), Ap = A p, partial p . Ap.  The direction update of the
// reference's update_p_kernel (pcg_solver.cu:456-464) is fused here: every
// gathered column recomputes z + beta p_prev and the row's own p is written
// to the other p buffer.  The SpMV restates spmv_kernel (block_csr.cu:54-83)
// on the wave-sliced ELL layout, in batches of UNROLL entries whose loads are
// issued together.  Mode kModeCheck: the recursive residual passed the test,
// so this pass computes the true residual instead.  Mode kModeFlush: the
// owed iterate term is added.
// ---------------------------------------------------------------------------
// Entry order of the symmetric pass: 0, w-1, 1, w-2, ...  A row's columns
// ascend, so step t pairs its t-th lower entry with its t-th upper entry
// counted from the end.  On a grid, the mirror that step t reads for lower
// entry t (row i - W, i - W + 1 or i - 1) is the upper entry that its owner
// row, 16 slices earlier at W = 1024, streams at its own step t: both
// wavefronts touch the block at the same point of their lives, a fraction of
// a microsecond apart, while it is still in L2.  (In ascending order the
// mirror is read first and the owner reads it at the end of its loop, after
// the XCD has streamed several L2 capacities.)
__device__ __forceinline__ int sym_order(int k, int w) { return (k & 1) ? w - 1 - (k >> 1) : (k >> 1); }

// ---------------------------------------------------------------------------
// Advance of an instance's solver state by one SpMV/update pair, given the
// sums of the partials that pair wrote: r . z and |r|^2 of the update pass
// (Iter), or the true ones of the check pass (Check); none for a flush.  Scalar logic restates
// reduce_rho_beta_kernel (pcg_solver.cu:436-454) with its 1e-37 guard, plus
// the stopping rule: true relative residual sqrt(|b-Ax|^2 / max(|b|^2,1e-30))
// <= tol (pcg_solver.cu:595-641, 916-923), or exactly max_iter iterations
// when tol == 0 (pcg_solver.cu:542-592).  Runs in the next SpMV pass (every
// workgroup, identical inputs and order) or in pcg_advance_kernel.
// ---------------------------------------------------------------------------
struct AdvanceSums {
    double rz;  // r . z
    double rr;  // |r|^2
};
__device__ __forceinline__ void pcg_advance(InstState& S, const InstParams& pr, const AdvanceSums& t) {
    const double tol2 = pr.tol2;
    if (S.mode == kModeFlush) {  // x is whole again
        S.pend = 0;
        S.alpha_pend = 0.0;
        S.defer = 1;
        if (S.flush_to == kModeCheck) {
            S.mode = kModeCheck;
        } else {
            S.mode = kModeIter;
            S.status = kStopped;
        }
    } else if (S.mode == kModeCheck) {
        const double rr = t.rr;
        S.rr = rr;
        S.relres = sqrt(rr / S.bb);
        if (rr <= tol2 * S.bb) {
            S.status = kConverged;
        } else {  // residual replacement: restart the recursion from b - A x
            S.restarts += 1;
            S.rho = t.rz;
            S.beta = 0.0;
            S.mode = kModeIter;
            S.defer = 1;
            if (S.iters >= pr.max_iter) S.status = kStopped;
        }
    } else {
        const double rho = S.rho, rho_new = t.rz;
        if (S.pend) {  // this update pass added the owed term
            S.pend = 0;
        } else if (S.defer) {  // this update pass left x += alpha p to the next
            S.alpha_pend = S.alpha;  // the alpha that pass used
            S.pend = 1;
        }
        S.defer = S.pend ? 0 : 1;
        S.beta = (rho > kGuard || rho < -kGuard) ? rho_new / rho : 0.0;
        S.rho = rho_new;
        S.rr = t.rr;
        S.parity ^= 1;
        S.iters += 1;
        const bool check = tol2 > 0.0 && t.rr <= tol2 * S.bb;
        const bool stop = !check && S.iters >= pr.max_iter;
        if ((check || stop) && S.pend) {  // complete x first
            S.mode = kModeFlush;
            S.flush_to = check ? int(kModeCheck) : int(kStopped);
        } else if (check) {
            S.mode = kModeCheck;
        } else if (stop) {
            S.status = kStopped;
        }
    }
}

// Sums the partials an owed advance needs (by the whole workgroup, in the
// fixed order of block_sum; every thread gets the totals).
__device__ __forceinline__ AdvanceSums advance_sums(const double* __restrict__ partials, int ngroups,
                                                    const InstState& S) {
    // constant slot pairs: a slot chosen at run time costs the SpMV pass
    // registers (and occupancy)
    double t[2] = {0.0, 0.0};
    if (S.mode == kModeIter) {
        const int sl[2] = {0, 1};
        block_sum<2, 1>(partials, ngroups, S.group0, S.ngroups, sl, t);
    } else if (S.mode == kModeCheck) {
        const int sl[2] = {kSlotCheckRZ, kSlotCheckRR};
        block_sum<2, 1>(partials, ngroups, S.group0, S.ngroups, sl, t);
    }
    return AdvanceSums{t[0], t[1]};
}

// The state of a workgroup's instance after the owed advance: every
// workgroup sums the previous pair's partials in the same fixed order, so all
// compute the same state; the workgroup of the instance's first group writes
// it to st_out, for the update pass and the next SpMV pass (also when the
// instance is not running, so both buffers agree once it has stopped).
__device__ __forceinline__ int uniform_int(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uniform_double(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(b))));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(b >> 32))));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ void advance_in_pass(const LaunchGeom& g, const DevVectors& v, InstState& S,
                                                const InstParams* __restrict__ params, InstState* __restrict__ st_out,
                                                int inst, int grp) {
    const int status_in = S.status;
    if (S.status == kRunning && S.owed) {
        // every thread holds the same sums and applies the same step
        AdvanceSums t = advance_sums(v.partials, g.ngroups, S);
        t.rz = uniform_double(t.rz);
        t.rr = uniform_double(t.rr);
        pcg_advance(S, params[inst], t);
        S.status = uniform_int(S.status);
        S.mode = uniform_int(S.mode);
        S.parity = uniform_int(S.parity);
        S.beta = uniform_double(S.beta);
        S.alpha_pend = uniform_double(S.alpha_pend);
    }
    S.owed = S.status == kRunning ? 1 : 0;  // this pass and its update pass owe the next advance
    if (grp != S.group0) return;
    if (threadIdx.x == 0) st_out[inst] = S;
    // a cloth that just finished: its groups leave the iteration passes
    if (g.gactive != nullptr && status_in == kRunning && S.status != kRunning)
        for (int q = int(threadIdx.x); q < S.ngroups; q += kReduceThreads) g.gactive[S.group0 + q] = 0;
}

// st_out null: no advance (the state in st is current; for timing and the
// autotuner); otherwise the pass advances st into st_out first.
// Scalar registers are capped at 80: measured on MI355X, this 16-wave
// workgroup at 85 scalar registers runs one workgroup per CU instead of two
// (the SpMV pass 33.4 instead of 27.4 us at 1024^2), although the compiler's
// occupancy estimate still reports 8 waves per SIMD.  At 80 the compiler
// spills a few scalars into vector-register lanes.
template <int UNROLL, bool BF, bool SYM, bool PK, bool I16, bool ADV>
__global__ __launch_bounds__(64 * kGroupSlices) __attribute__((amdgpu_num_sgpr(80))) void pcg_spmv_kernel(LaunchGeom g, DevMatrix A, DevVectors v,
                                                                    const InstState* __restrict__ st,
                                                                    InstState* __restrict__ st_out,
                                                                    const InstParams* __restrict__ params) {
    const int s = my_slice<kGroupSlices>(g.nslices);
    if (s >= g.nslices) return;
    const int grp = s / kGroupSlices;
    const SliceDesc sd = g.slices[s];
    // a finished cloth (its first workgroup still copies the state); the
    // cloth's first group comes with the slice descriptor, so a finished
    // group's workgroup exits after one round of loads
    if (g.gactive != nullptr && g.gactive[grp] == 0 && (!ADV || grp != sd.group0)) return;
    const int lane = threadIdx.x & 63;
    const bool valid = lane < sd.nvalid;
    const int row = s * 64 + lane;
    constexpr int BS = PK ? 6 : 9;
    const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
    const int nlow = SYM ? sd.nlower : 0;  // entries >= nlow hold no lower block in any lane
    const int* colp = A.col + (size_t)sb * 64 + lane;
    const int* mirp = A.mir + (size_t)sb * 64 + lane;
    const int16_t* colp16 = A.col16 + (size_t)sb * 64 + lane;
    const int16_t* mirp16 = A.mir16 + (size_t)sb * 64 + lane;
    const int e0 = sb * 64 + lane;  // entry index of the row's entry 0
    const float* vp = A.val + (size_t)sb * BS * 64 + lane;
    // A batch of UNROLL entries runs in three phases: (1) the column and
    // mirror offsets of every entry, (2) every gather and block load,
    // (3) the FMAs.  No load of phase 2 depends on another load of the
    // batch, so a batch costs two memory round trips.  Whether an entry
    // needs its mirror offset (kk < nlow) is wave-uniform and known
    // before any vector load returns.
    int kk[UNROLL], cix[UNROLL], mix[UNROLL];
    F3 zj[UNROLL], qj[UNROLL];
    float blk[UNROLL][BS];
    auto load_offsets = [&](int k0) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int k = k0 + u;
            kk[u] = SYM ? sym_order(k, w) : k;
            cix[u] = I16 ? int(kNoCol16) : -1;
            mix[u] = 0;
            if (k < w) {  // wave-uniform
                cix[u] = I16 ? int(colp16[kk[u] * 64]) : colp[kk[u] * 64];
                if (SYM && kk[u] < nlow) mix[u] = I16 ? int(mirp16[kk[u] * 64]) : mirp[kk[u] * 64];
            }
        }
    };
    // The first batch's offsets, the diagonal block and the row's z depend
    // on no solver state.  Without the in-pass advance they go out before
    // the state is read, so the two latencies overlap; with it they follow
    // the advance (held across its sums they would cost the pass occupancy).
    F3 zi = f3(0.0f, 0.0f, 0.0f);
    float dg[BS];
    auto load_first = [&] {
        if (valid) {
            load_offsets(0);
            zi = ld3(v.z, row);
            const float* dp = A.diag + (size_t)s * BS * 64 + lane;
#pragma unroll
            for (int q = 0; q < BS; ++q) dg[q] = dp[q * 64];
        }
    };
    if constexpr (!ADV) load_first();
    InstState S = st[sd.inst];
    if constexpr (ADV) advance_in_pass(g, v, S, params, st_out, sd.inst, grp);
    if (S.status != kRunning) return;
    if constexpr (ADV) load_first();
    const int parity = S.parity;
    if (S.mode == kModeCheck) {
        // true |r|^2 and r . z, into slots no workgroup of this pass reads
        // (the advance above read the update pass's slots 0-1)
        double part[2] = {0.0, 0.0};
        if (valid) {
            double q[4] = {0.0, 0.0, 0.0, 0.0};
            residual_row(A, v, parity ? v.p1 : v.p0, row, sd.nlower, false, q);
            part[0] = q[0];
            part[1] = q[2];
        }
        const int sl[2] = {kSlotCheckRR, kSlotCheckRZ};
        group_partials<2>(part, sl, v.partials, s / kGroupSlices, g.ngroups);
        return;
    }
    if (S.mode == kModeFlush) {
        // the owed term of a deferred update; the latest direction is the
        // current p_prev buffer
        if (valid) {
            const F3 pk = ld3(parity ? v.p1 : v.p0, row);
            const double ap = S.alpha_pend;
            double* x = v.x + 3 * row;
            x[0] += ap * double(pk.x);
            x[1] += ap * double(pk.y);
            x[2] += ap * double(pk.z);
        }
        return;
    }
    const float beta = float(S.beta);
    const float* __restrict__ pold = parity ? v.p1 : v.p0;
    float* __restrict__ pnew = parity ? v.p0 : v.p1;
    const float* __restrict__ z = v.z;
    double part[1] = {0.0};
    if (valid) {
        auto column = [&](int u) { return I16 ? (cix[u] == kNoCol16 ? -1 : row + cix[u]) : cix[u]; };
        // BF: padding entries hold zero blocks and gather the row's own
        // vectors instead of branching (k0 + u < w is wave-uniform).
        auto load_operands = [&](int k0) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const int c = column(u);
                if (BF ? (k0 + u < w) : (c >= 0)) {
                    const int cc = (c >= 0) ? c : row;
                    zj[u] = ld3(z, cc);
                    qj[u] = ld3(pold, cc);
                    // the lower block (row, cc) is read from its mirror (cc, row),
                    // which row cc's lane streams at the same step of its loop
                    const float* b = (SYM && cc < row)
                                         ? A.val + val_offset(I16 ? e0 + kk[u] * 64 + mix[u] : mix[u], BS)
                                         : vp + kk[u] * BS * 64;
#pragma unroll
                    for (int q = 0; q < BS; ++q) blk[u][q] = b[q * 64];
                }
            }
        };
        float a0, a1, a2;
        auto accumulate = [&](int k0) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const int c = column(u);
                if (BF ? (k0 + u < w) : (c >= 0)) {
                    const int cc = (c >= 0) ? c : row;
                    const F3 pj = dir3(zj[u], qj[u], beta);
                    const float* m = blk[u];
                    if (PK) {  // symmetric block, upper triangle
                        a0 = fma3(a0, m[0], m[1], m[2], pj);
                        a1 = fma3(a1, m[1], m[3], m[4], pj);
                        a2 = fma3(a2, m[2], m[4], m[5], pj);
                    } else if (SYM && cc < row) {  // transpose of the mirror
                        a0 = fma3(a0, m[0], m[3], m[6], pj);
                        a1 = fma3(a1, m[1], m[4], m[7], pj);
                        a2 = fma3(a2, m[2], m[5], m[8], pj);
                    } else {
                        a0 = fma3(a0, m[0], m[1], m[2], pj);
                        a1 = fma3(a1, m[3], m[4], m[5], pj);
                        a2 = fma3(a2, m[6], m[7], m[8], pj);
                    }
                }
            }
        };
        // The first batch's operands go out with the row's p_prev.
        const F3 qi = ld3(pold, row);
        load_operands(0);
        const F3 pi = dir3(zi, qi, beta);
        if (PK) {
            a0 = fma3(0.0f, dg[0], dg[1], dg[2], pi);
            a1 = fma3(0.0f, dg[1], dg[3], dg[4], pi);
            a2 = fma3(0.0f, dg[2], dg[4], dg[5], pi);
        } else {
            a0 = fma3(0.0f, dg[0], dg[1], dg[2], pi);
            a1 = fma3(0.0f, dg[3], dg[4], dg[5], pi);
            a2 = fma3(0.0f, dg[6], dg[7], dg[8], pi);
        }
        accumulate(0);
        for (int k0 = UNROLL; k0 < w; k0 += UNROLL) {
            load_offsets(k0);
            load_operands(k0);
            accumulate(k0);
        }
        st3(v.ap, row, f3(a0, a1, a2));
        st3(pnew, row, pi);
        part[0] = fma3d(0.0, double(pi.x), double(pi.y), double(pi.z), a0, a1, a2);
    }
    const int sl[1] = {kSlotSigma};
    group_partials<1>(part, sl, v.partials, s / kGroupSlices, g.ngroups);
}

// ---------------------------------------------------------------------------
// PCG update pass: alpha = rho / sigma (guarded), x += alpha p (FP64),
// r -= alpha Ap, z = M^-1 r, partials r . z and r . r.  Restates
// alpha_jacobi_dot_partial_kernel (pcg_solver.cu:406-434) with an FP64 iterate.
// Every other pass defers x += alpha p; the next pass adds both terms, with
// the older direction still in the other p buffer.
// ---------------------------------------------------------------------------
// X: 0 defers x += alpha p; 1 adds the owed term and this one; 2 adds this
// one.  Straight-line per variant, so every operand load (including x) is
// issued before the p . Ap sum and waited for once.
// sig_t: the first round of sigma's partial loads (block_sum_issue), issued
// by update_group before the solver state was read.
// The FP64 iterate of a wave's 64 rows is 192 contiguous doubles; it is read
// and written 64 consecutive doubles per instruction and moved between that
// order and the rows' (x0, x1, x2) through the wave's LDS slab (three plain
// 8-byte accesses per row, 24 bytes apart, covered each line a third at a
// time: the 1024^2 update pass that adds one term took 23.3 us that way,
// 22.3 us staged; 108.0 -> 103.9 us adding two terms at 2048^2).
constexpr int kSigmaKU = 2;
// Orders one wave's LDS accesses around a point (LDS operations of a wave
// complete in issue order; only the compiler's reordering is prevented).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int X>
__device__ __forceinline__ void update_rows(const LaunchGeom& g, const DevVectors& v, InstState& S, int s,
                                           const SliceDesc& sd, F3 r0, F3 api, const float (&m)[6],
                                           const double (&sig_t)[1][kSigmaKU], double& sig_cache, int& sig_inst,
                                           double* __restrict__ slab) {
    const float* __restrict__ p = S.parity ? v.p0 : v.p1;      // written by the SpMV pass
    const float* __restrict__ pprev = S.parity ? v.p1 : v.p0;  // the previous direction
    const int lane = threadIdx.x & 63;
    const int row = s * 64 + lane;
    F3 pi = f3(0.0f, 0.0f, 0.0f), pp = pi;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    double* xw = v.x + (size_t)s * 192;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (X != 0) {
        pi = ld3(p, row);
        t0 = xw[lane];
        t1 = xw[lane + 64];
        t2 = xw[lane + 128];
    }
    if (X == 1) pp = ld3(pprev, row);
    // sigma = p . Ap of this instance, in the fixed order of every per-iteration
    // sum; kept for the workgroup's second group when it is the same instance
    if (sig_inst != sd.inst) {
        double sum[1];
        const int sig_slot[1] = {kSlotSigma};
        block_sum_finish<1, kSigmaKU>(v.partials, g.ngroups, sd.group0, sd.ngroups, sig_slot, sig_t, sum);
        sig_cache = uniform_double(sum[0]);
        sig_inst = sd.inst;
    }
    const double sig = sig_cache, rho = S.rho;
    const double alpha = (sig > kGuard || sig < -kGuard) ? rho / sig : 0.0;
    const float af = float(alpha);
    if (threadIdx.x == 0 && s / kGroupSlices == S.group0) {  // one writer per instance
        S.sigma = sig;
        S.alpha = alpha;
    }
    double part[2] = {0.0, 0.0};
    if (X != 0) {
        slab[lane] = t0;
        slab[lane + 64] = t1;
        slab[lane + 128] = t2;
        wave_lds_sync();
        x0 = slab[3 * lane];
        x1 = slab[3 * lane + 1];
        x2 = slab[3 * lane + 2];
        wave_lds_sync();
    }
    double n0 = x0, n1 = x1, n2 = x2;
    if (lane < sd.nvalid) {
        if (X == 1) {  // the owed term and this one
            const double ap = S.alpha_pend;
            n0 = x0 + (ap * double(pp.x) + alpha * double(pi.x));
            n1 = x1 + (ap * double(pp.y) + alpha * double(pi.y));
            n2 = x2 + (ap * double(pp.z) + alpha * double(pi.z));
        } else if (X == 2) {
            n0 = x0 + alpha * double(pi.x);
            n1 = x1 + alpha * double(pi.y);
            n2 = x2 + alpha * double(pi.z);
        }
        const F3 ri = r0 - api * af;
        const F3 zi = f3(m[0] * ri.x + m[1] * ri.y + m[2] * ri.z, m[1] * ri.x + m[3] * ri.y + m[4] * ri.z,
                         m[2] * ri.x + m[4] * ri.y + m[5] * ri.z);
        st3(v.r, row, ri);
        st3(v.z, row, zi);
        part[0] = double(ri.x) * zi.x + double(ri.y) * zi.y + double(ri.z) * zi.z;
        part[1] = double(ri.x) * ri.x + double(ri.y) * ri.y + double(ri.z) * ri.z;
    }
    if (X != 0) {
        // the rows' x go out through the wave's slab, each store covering 512 contiguous bytes
        slab[3 * lane] = n0;
        slab[3 * lane + 1] = n1;
        slab[3 * lane + 2] = n2;
        wave_lds_sync();
        const int nv3 = 3 * sd.nvalid;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (lane + 64 * c < nv3) xw[lane + 64 * c] = slab[lane + 64 * c];
        wave_lds_sync();
    }
    const int sl[2] = {0, 1};
    group_partials<2>(part, sl, v.partials, s / kGroupSlices, g.ngroups);
}

// One group of the update pass (all 16 waves of the workgroup).
__device__ __forceinline__ void update_group(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v,
                                             InstState* __restrict__ st, int grp, double& sig_cache, int& sig_inst) {
    if (g.gactive != nullptr && g.gactive[grp] == 0) return;  // a finished cloth
    const int s = __builtin_amdgcn_readfirstlane(grp * kGroupSlices + int(threadIdx.x >> 6));
    // The operands every variant needs are loaded before the solver state is
    // read, so the two latencies overlap.  Every lane loads (rows past an
    // instance's end are allocated padding); only valid lanes store.
    const int lane = threadIdx.x & 63;
    const int row = s * 64 + lane;
    const F3 r0 = ld3(v.r, row), api = ld3(v.ap, row);
    float m[6];
    {
        const float* mv = A.minv + (size_t)s * 6 * 64 + lane;
#pragma unroll
        for (int c = 0; c < 6; ++c) m[c] = mv[c * 64];
    }
    const SliceDesc sd = g.slices[s];
    // sigma's partial loads go out before the state is read too (the group
    // range comes with the slice descriptor); update_rows issues its
    // state-dependent loads before it waits for them
    double sig_t[1][kSigmaKU];
    if (sig_inst != sd.inst) {
        const int sig_slot[1] = {kSlotSigma};
        block_sum_issue<1, kSigmaKU>(v.partials, g.ngroups, sd.group0, sd.ngroups, sig_slot, sig_t);
    }
    InstState& S = st[sd.inst];
    if (S.status != kRunning || S.mode != kModeIter) return;  // check and flush rounds update nothing
    __shared__ double xslab[kGroupSlices][192];
    double* slab = xslab[threadIdx.x >> 6];
    if (S.pend) update_rows<1>(g, v, S, s, sd, r0, api, m, sig_t, sig_cache, sig_inst, slab);
    else if (S.defer) update_rows<0>(g, v, S, s, sd, r0, api, m, sig_t, sig_cache, sig_inst, slab);
    else update_rows<2>(g, v, S, s, sd, r0, api, m, sig_t, sig_cache, sig_inst, slab);
}

// With `pairs` (one instance), each workgroup updates two groups, i and
// i + ceil(ngroups / 2), one after the other: half the workgroups (all
// resident at once at 1024^2), and sigma is summed once for both.  A batch
// keeps one group per workgroup: there the two groups would mostly belong to
// different cloths, and a finished cloth's group would hold up a live one
// (the 512-cloth batch ran 3.8% slower paired).  Workgroup ids are remapped
// so each XCD walks contiguous ranges of groups.
// At most 64 VGPRs (two 16-wave workgroups per CU).
__global__ __launch_bounds__(64 * kGroupSlices) __attribute__((amdgpu_waves_per_eu(8))) void pcg_update_kernel(
    LaunchGeom g, DevMatrix A, DevVectors v, InstState* __restrict__ st, int pairs) {
    const int nfirst = pairs ? (g.ngroups + 1) / 2 : g.ngroups;
    const int per = (nfirst + 7) >> 3;
    const int b = blockIdx.x;
    const int first = (b & 7) * per + (b >> 3);
    if (first >= nfirst) return;
    double sig_cache = 0.0;
    int sig_inst = -1;
    update_group(g, A, v, st, first, sig_cache, sig_inst);
    if (!pairs || first + nfirst >= g.ngroups) return;
    __syncthreads();  // the first group's LDS reads are done
    update_group(g, A, v, st, first + nfirst, sig_cache, sig_inst);
}

// Stand-alone advance at the end of a chunk of iterations (one workgroup of
// kReduceThreads per instance, the same sums as the SpMV pass), so the state
// the host polls is up to date.  In place; clears `owed`.
__global__ __launch_bounds__(kReduceThreads) void pcg_advance_kernel(InstState* __restrict__ st,
                                                                     const InstParams* __restrict__ params,
                                                                     const double* __restrict__ partials,
                                                                     int ngroups, int* __restrict__ gactive) {
    InstState S = st[blockIdx.x];
    if (S.status != kRunning || !S.owed) return;
    const AdvanceSums t = advance_sums(partials, ngroups, S);
    if (threadIdx.x != 0) return;
    pcg_advance(S, params[blockIdx.x], t);
    S.owed = 0;
    st[blockIdx.x] = S;
    // a cloth that just finished: its groups leave the iteration passes
    if (gactive != nullptr && S.status != kRunning)
        for (int q = 0; q < S.ngroups; ++q) gactive[S.group0 + q] = 0;
}

// The advance of advance_in_pass as a kernel of its own (one workgroup of
// kReduceThreads per instance, the same sums in the same order and the same
// scalar step): reads the state from st_in and writes the advanced state,
// marked owed while the solve runs, to st_out, for an SpMV pass launched
// without an advance and its update pass.
__global__ __launch_bounds__(kReduceThreads) void pcg_advance_copy_kernel(const InstState* __restrict__ st_in,
                                                                          InstState* __restrict__ st_out,
                                                                          const InstParams* __restrict__ params,
                                                                          const double* __restrict__ partials,
                                                                          int ngroups, int* __restrict__ gactive) {
    InstState S = st_in[blockIdx.x];
    const int status_in = S.status;
    if (S.status == kRunning && S.owed) {
        const AdvanceSums t = advance_sums(partials, ngroups, S);
        pcg_advance(S, params[blockIdx.x], t);
    }
    S.owed = S.status == kRunning ? 1 : 0;
    if (threadIdx.x == 0) st_out[blockIdx.x] = S;
    // a cloth that just finished: its groups leave the iteration passes
    if (gactive != nullptr && status_in == kRunning && S.status != kRunning)
        for (int q = int(threadIdx.x); q < S.ngroups; q += kReduceThreads) gactive[S.group0 + q] = 0;
}

// ---------------------------------------------------------------------------
// Start and final reductions: one wavefront per instance sums that
// instance's group partials in a fixed order (deterministic).  Restates
// reduce_partial_kernel (pcg_solver.cu:374-384) for the initial r . z, and
// the true relative residual of the stopping rule.
// ---------------------------------------------------------------------------
// single_ng > 0: one instance owning groups [0, single_ng) (the partial
// loads then need no state read first); 0: bounds from the instance state.
// The phase's partials are loaded before the state is inspected, so the two
// loads overlap; an instance that is not live discards them.
__global__ __launch_bounds__(64) void pcg_reduce_kernel(InstState* __restrict__ st,
                                                        const InstParams* __restrict__ params,
                                                        const double* __restrict__ partials, int ngroups, int phase,
                                                        int single_ng, int* __restrict__ gactive) {
    InstState& S = st[blockIdx.x];
    const InstParams& pr = params[blockIdx.x];
    const int g0 = single_ng > 0 ? 0 : S.group0;
    const int ng = single_ng > 0 ? single_ng : S.ngroups;
    // Slots: start |r|^2, |b|^2, r . z; final |r|^2.
    const double* base = partials + g0;
    double tot[kPartialSlots] = {};
    tot[0] = wave_group_sum(base, ng);
    if (phase == kPhaseStart) {
        tot[1] = wave_group_sum(base + (size_t)ngroups, ng);
        tot[2] = wave_group_sum(base + (size_t)2 * ngroups, ng);
    }
    if (phase == kPhaseFinal && S.status != kStopped) return;
    if (phase == kPhaseStart && gactive != nullptr) {
        // the instance iterates from here on unless it is done already
        const bool run = !(pr.tol2 > 0.0 && tot[0] <= pr.tol2 * fmax(tot[1], 1.0e-30)) && pr.max_iter > 0;
        for (int q = int(threadIdx.x); q < ng; q += 64) gactive[g0 + q] = run ? 1 : 0;
    }
    if (threadIdx.x != 0) return;
    const double tol2 = pr.tol2;
    if (phase == kPhaseStart) {
        S.bb = fmax(tot[1], 1.0e-30);
        S.rr = tot[0];
        S.relres = sqrt(tot[0] / S.bb);
        S.iters = 0;
        S.restarts = 0;
        S.mode = kModeIter;
        S.beta = 0.0;
        S.rho = tot[2];
        S.pend = 0;
        S.alpha_pend = 0.0;
        S.defer = 1;
        S.owed = 0;
        S.status = (tol2 > 0.0 && tot[0] <= tol2 * S.bb) ? kConverged : (pr.max_iter <= 0 ? kStopped : kRunning);
    } else {
        S.rr = tot[0];
        S.relres = sqrt(tot[0] / S.bb);
    }
}

// finalize_step_kernel (cloth_simulator.cu:222-244): x = x_n + dx,
// v = dx / dt (times exp(-damping dt) when damping > 0).
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void finalize_kernel(LaunchGeom g, DevVectors v, DevExt e,
                                                             const InstParams* __restrict__ params, float dt,
                                                             const int* __restrict__ pin_inc_off,
                                                             const int* __restrict__ pin_inc,
                                                             const float* __restrict__ pin_tgt) {
    const int s = my_slice<WAVES>(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63;
    if (lane >= sd.nvalid) return;
    const int row = s * 64 + lane, p = sd.part0 + lane;
    const InstParams pr = params[sd.inst];
    F3 d = f3(float(v.x[3 * row]), float(v.x[3 * row + 1]), float(v.x[3 * row + 2]));
    const F3 xn = ld3(e.pos, p);
    F3 xnew = xn + d;
    if (pr.pin_proj && pin_inc_off[row + 1] > pin_inc_off[row]) {  // a projected pin ends at its target
        xnew = ld3(pin_tgt, pin_inc[pin_inc_off[row]]);
        d = xnew - xn;
    }
    const float inv_dt = 1.0f / dt;
    F3 vn = d * inv_dt;
    if (pr.damping > 0.0f) vn = vn * expf(-pr.damping * dt);
    st3(e.pos, p, xnew);
    st3(e.vel, p, vn);
}

// Plain y = A x on the ELL layout (spmv_kernel, block_csr.cu:54-83, alpha=1, beta=0).
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void spmv_kernel(LaunchGeom g, DevMatrix A, const float* __restrict__ x,
                                                         float* __restrict__ y) {
    const int s = my_slice<WAVES>(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63;
    if (lane >= sd.nvalid) return;
    const int row = s * 64 + lane;
    const F3 xi = ld3(x, row);
    const float* dg = A.diag + (size_t)s * 9 * 64 + lane;
    float a0 = dg[0] * xi.x + dg[64] * xi.y + dg[128] * xi.z;
    float a1 = dg[192] * xi.x + dg[256] * xi.y + dg[320] * xi.z;
    float a2 = dg[384] * xi.x + dg[448] * xi.y + dg[512] * xi.z;
    const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
    const int* colp = A.col + (size_t)sb * 64 + lane;
    const float* vp = A.val + (size_t)sb * 9 * 64 + lane;
    for (int k = 0; k < w; ++k) {
        const int c = colp[k * 64];
        if (c >= 0) {
            const float* bk = vp + k * 9 * 64;
            const F3 xj = ld3(x, c);
            a0 += bk[0] * xj.x + bk[64] * xj.y + bk[128] * xj.z;
            a1 += bk[192] * xj.x + bk[256] * xj.y + bk[320] * xj.z;
            a2 += bk[384] * xj.x + bk[448] * xj.y + bk[512] * xj.z;
        }
    }
    st3(y, row, f3(a0, a1, a2));
}

}  // namespace

#define CHYSX_LAUNCH_CHECK()                                                            \
    do {                                                                                \
        const hipError_t err_ = hipGetLastError();                                      \
        if (err_ != hipSuccess) throw_hip(err_, __func__);                              \
    } while (0)

[[noreturn]] void throw_hip(hipError_t err, const char* what);

// Launch helpers: WAVES (1, 2, 4, 8, 16) selects the workgroup size.
#define CHYSX_WAVES_SWITCH(waves, LAUNCH) \
    switch (waves) {                      \
        case 1: LAUNCH(1); break;         \
        case 2: LAUNCH(2); break;         \
        case 4: LAUNCH(4); break;         \
        case 8: LAUNCH(8); break;         \
        case 16: LAUNCH(16); break;       \
        default: throw std::invalid_argument("unsupported waves per workgroup"); \
    }

bool valid_waves(int w) { return w == 1 || w == 2 || w == 4 || w == 8 || w == 16; }
bool valid_unroll(int u) { return u == 1 || u == 2 || u == 4; }

void launch_rest_lengths(const int* spr_v, const float* pos, float* rest, int n, hipStream_t s) {
    if (n <= 0) return;
    rest_length_kernel<<<(n + 255) / 256, 256, 0, s>>>(spr_v, pos, rest, n);
    CHYSX_LAUNCH_CHECK();
}
void launch_rest_shapes(const int* tri_v, const float* pos, float4* dminv, float* area, int n, float ct, float st,
                        hipStream_t s) {
    if (n <= 0) return;
    rest_shape_kernel<<<(n + 255) / 256, 256, 0, s>>>(tri_v, pos, dminv, area, n, ct, st);
    CHYSX_LAUNCH_CHECK();
}
void launch_area_mass(const int* vt_off, const int* vt_ids, const int* tri_v, const float* pos, float d3,
                      float* inv_mass, int n, hipStream_t s) {
    if (n <= 0) return;
    area_mass_kernel<<<(n + 255) / 256, 256, 0, s>>>(vt_off, vt_ids, tri_v, pos, d3, inv_mass, n);
    CHYSX_LAUNCH_CHECK();
}
void launch_assemble(const LaunchGeom& g, const DevMesh& m, const DevMatrix& A, const DevVectors& v,
                     const DevExt& e, const InstParams* params, float dt, int waves, int maxw, hipStream_t s) {
    if (g.nslices <= 0) return;
    // LDS accumulator: waves * maxw blocks of bs x 64 floats per workgroup,
    // kept within 64 KiB by lowering the waves per workgroup; rows wider
    // than 28 blocks (42 packed) accumulate in the matrix itself.
    const size_t slab = size_t(std::max(maxw, 1)) * size_t(A.bs) * 64 * sizeof(float);
    if (slab * 1 <= kAssembleLdsBytes) {
        while (waves > 1 && slab * size_t(waves) > kAssembleLdsBytes) waves >>= 1;
        const size_t bytes = slab * size_t(waves);
#define L(W) assemble_kernel<W, true><<<grid_for(g.nslices, W), 64 * W, bytes, s>>>(g, m, A, v, e, params, dt, maxw)
        CHYSX_WAVES_SWITCH(waves, L)
#undef L
    } else {
#define L(W) assemble_kernel<W, false><<<grid_for(g.nslices, W), 64 * W, 0, s>>>(g, m, A, v, e, params, dt, maxw)
        CHYSX_WAVES_SWITCH(waves, L)
#undef L
    }
    CHYSX_LAUNCH_CHECK();
}
void launch_rest_angles(const int4* dih, const float* pos, float* rest, int n, hipStream_t s) {
    if (n <= 0) return;
    rest_angle_kernel<<<(n + 255) / 256, 256, 0, s>>>(dih, pos, rest, n);
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_residual(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                         int report, hipStream_t s) {
    if (g.nslices <= 0) return;
    if (g.nslices % kGroupSlices != 0 || g.ngroups * kGroupSlices != g.nslices)
        throw std::invalid_argument("PCG passes: slices must form whole 16-slice groups");
    pcg_residual_kernel<<<grid_for(g.nslices, kGroupSlices), 64 * kGroupSlices, 0, s>>>(g, A, v, st, report);
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_start_given_r(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                              hipStream_t s) {
    if (g.nslices <= 0) return;
    if (g.nslices % kGroupSlices != 0 || g.ngroups * kGroupSlices != g.nslices)
        throw std::invalid_argument("PCG passes: slices must form whole 16-slice groups");
    pcg_start_given_r_kernel<<<grid_for(g.nslices, kGroupSlices), 64 * kGroupSlices, 0, s>>>(g, A, v, st);
    CHYSX_LAUNCH_CHECK();
}
namespace {
template <bool BF, bool SYM, bool PK, bool I16>
void pcg_spmv_launch(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                     InstState* st_out, const InstParams* params, int unroll, hipStream_t s) {
    const int grid = grid_for(g.nslices, kGroupSlices), block = 64 * kGroupSlices;
    // the in-pass advance is its own instantiation, so the pass without it
    // carries none of its registers
#define L(U, ADV) pcg_spmv_kernel<U, BF, SYM, PK, I16, ADV><<<grid, block, 0, s>>>(g, A, v, st, st_out, params)
    if (st_out != nullptr) {
        if (unroll == 1) L(1, true);
        else if (unroll == 2) L(2, true);
        else L(4, true);
    } else {
        if (unroll == 1) L(1, false);
        else if (unroll == 2) L(2, false);
        else L(4, false);
    }
#undef L
}
template <bool BF, bool SYM, bool PK>
void pcg_spmv_dispatch(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                       InstState* st_out, const InstParams* params, int unroll, hipStream_t s) {
    if (A.col16 != nullptr && A.mir16 != nullptr)
        pcg_spmv_launch<BF, SYM, PK, true>(g, A, v, st, st_out, params, unroll, s);
    else
        pcg_spmv_launch<BF, SYM, PK, false>(g, A, v, st, st_out, params, unroll, s);
}
}  // namespace

void launch_pcg_spmv(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const InstState* st,
                     InstState* st_out, const InstParams* params, int unroll, hipStream_t s, int branch_free,
                     int symmetric) {
    if (g.nslices <= 0) return;
    if (!valid_unroll(unroll)) throw std::invalid_argument("unsupported unroll");
    if (g.nslices % kGroupSlices != 0 || g.ngroups * kGroupSlices != g.nslices)
        throw std::invalid_argument("PCG passes: slices must form whole 16-slice groups");
    if (symmetric != A.sym) throw std::invalid_argument("symmetric SpMV: launch flag and matrix mode differ");
    if (A.bs != 9 && !(A.bs == 6 && A.sym)) throw std::invalid_argument("packed blocks need the symmetric SpMV");
    if (st_out != nullptr && (params == nullptr || st_out == st))
        throw std::invalid_argument("SpMV pass: the advanced state needs the parameters and a second buffer");
    if (branch_free) {
        if (A.bs == 6) pcg_spmv_dispatch<true, true, true>(g, A, v, st, st_out, params, unroll, s);
        else if (symmetric) pcg_spmv_dispatch<true, true, false>(g, A, v, st, st_out, params, unroll, s);
        else pcg_spmv_dispatch<true, false, false>(g, A, v, st, st_out, params, unroll, s);
    } else {
        if (A.bs == 6) pcg_spmv_dispatch<false, true, true>(g, A, v, st, st_out, params, unroll, s);
        else if (symmetric) pcg_spmv_dispatch<false, true, false>(g, A, v, st, st_out, params, unroll, s);
        else pcg_spmv_dispatch<false, false, false>(g, A, v, st, st_out, params, unroll, s);
    }
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_update(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, InstState* st,
                       hipStream_t s) {
    if (g.nslices <= 0) return;
    if (g.nslices % kGroupSlices != 0 || g.ngroups * kGroupSlices != g.nslices)
        throw std::invalid_argument("PCG passes: slices must form whole 16-slice groups");
    // a single instance has no group flags (one instance: gactive is null)
    const int pairs = g.gactive == nullptr ? 1 : 0;
    const int nfirst = pairs ? (g.ngroups + 1) / 2 : g.ngroups;
    pcg_update_kernel<<<8 * ((nfirst + 7) / 8), 64 * kGroupSlices, 0, s>>>(g, A, v, st, pairs);
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_advance_copy(const InstState* st_in, InstState* st_out, const InstParams* params,
                             const double* partials, int ninst, int ngroups, hipStream_t s, int* gactive) {
    if (ninst <= 0) return;
    if (st_in == st_out) throw std::invalid_argument("advance copy: input and output states must differ");
    pcg_advance_copy_kernel<<<ninst, kReduceThreads, 0, s>>>(st_in, st_out, params, partials, ngroups, gactive);
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_advance(InstState* st, const InstParams* params, const double* partials, int ninst, int ngroups,
                        hipStream_t s, int* gactive) {
    if (ninst <= 0) return;
    pcg_advance_kernel<<<ninst, kReduceThreads, 0, s>>>(st, params, partials, ngroups, gactive);
    CHYSX_LAUNCH_CHECK();
}
void launch_pcg_reduce(InstState* st, const InstParams* params, const double* partials, int ninst, int ngroups,
                       int phase, hipStream_t s, int* gactive) {
    if (ninst <= 0) return;
    if (phase != kPhaseStart && phase != kPhaseFinal) throw std::invalid_argument("unknown reduction phase");
    pcg_reduce_kernel<<<ninst, 64, 0, s>>>(st, params, partials, ngroups, phase, ninst == 1 ? ngroups : 0,
                                           gactive);
    CHYSX_LAUNCH_CHECK();
}
void launch_finalize(const LaunchGeom& g, const DevVectors& v, const DevExt& e, const InstParams* params, float dt,
                     const DevMesh& m, int waves, hipStream_t s) {
    if (g.nslices <= 0) return;
#define L(W)                                                                                 \
    finalize_kernel<W><<<grid_for(g.nslices, W), 64 * W, 0, s>>>(g, v, e, params, dt, m.pin_inc_off, m.pin_inc, \
                                                                  m.pin_tgt)
    CHYSX_WAVES_SWITCH(waves, L)
#undef L
    CHYSX_LAUNCH_CHECK();
}
void launch_spmv(const LaunchGeom& g, const DevMatrix& A, const float* x, float* y, int waves, hipStream_t s) {
    // spmv_kernel reads every block itself as 9 floats: the symmetric
    // storages hold only upper blocks (6 floats when packed)
    if (A.sym != 0 || A.bs != 9)
        throw std::invalid_argument(
            "plain SpMV reads the general storage (9-float blocks, both halves): set spmv_symmetric=0 and "
            "spmv_packed=0, then step or assemble");
    if (g.nslices <= 0) return;
#define L(W) spmv_kernel<W><<<grid_for(g.nslices, W), 64 * W, 0, s>>>(g, A, x, y)
    CHYSX_WAVES_SWITCH(waves, L)
#undef L
    CHYSX_LAUNCH_CHECK();
}

}  // namespace chysx_amd
