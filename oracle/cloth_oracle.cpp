// SPDX-License-Identifier: Apache-2.0
//
// ORACLE -- test infrastructure only.
//
// A CPU restatement of ChysX's implicit (backward-Euler, one linear solve per
// step) cloth step, used to check the HIP implementation.  Only the test
// suite, the smoke entry point and the benchmark's CPU-baseline leg may load
// this library, and only as the checker or as the timed CPU reference; the
// product path never routes through it.
//
// Every routine cites the reference source it restates (paths relative to the
// reference repository root).  Two precision variants share one template:
//   Cloth<double>  FP64 everything (the oracle proper);
//   Cloth<float>   FP32 assembly, matrix, SpMV and vectors, FP64 inner
//                  products / alpha / beta / convergence test (mirrors the
//                  GPU precision split; used to size FP32 error).
//
// Extensions over the reference step, each off by default so the defaults
// reproduce the reference behaviour:
//   * springs_active   -- include SpringConstraint gradient/Hessian in the
//                         step (the reference skips them,
//                         src/cloth/cloth_simulator.cu:921).
//   * force_active     -- feed the external force into the predictor (the
//                         reference passes nullptr, cloth_simulator.cu:892-894).
//   * tol > 0          -- stop PCG at true relative residual
//                         sqrt(|b-Ax|^2 / max(|b|^2,1e-30)) <= tol
//                         (definition from src/solver/pcg_solver.cu:595-641,
//                         916-923); tol == 0 runs exactly max_iter iterations
//                         like the reference (pcg_solver.cu:542-592).
//   * per-element stiffness arrays for springs and stretch triangles.
//
// Parity status of each function is pinned by tests/test_oracle_*.py:
// closed forms, finite differences, brute-force pattern builds and
// numpy.linalg.solve.  The reference ships no golden vectors for this path.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Small fixed-size algebra.  Mat3 is row-major m[r*3+c] like the reference's
// math::Mat3 (src/math/matrix.cuh:127-162).
// ---------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T len3(V3<T> a) { return std::sqrt(dot3(a, a)); }
template <class T> inline V3<T> cross3(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class T> struct M3 {
    T m[9];
};
template <class T> inline M3<T> m3_zero() { M3<T> r; for (auto& v : r.m) v = T(0); return r; }
template <class T> inline M3<T> m3_ident() { M3<T> r = m3_zero<T>(); r.m[0] = r.m[4] = r.m[8] = T(1); return r; }
template <class T> inline M3<T> m3_outer(V3<T> a, V3<T> b) {
    const T av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i * 3 + j] = av[i] * bv[j];
    return r;
}
template <class T> inline M3<T> m3_scale(M3<T> a, T s) { for (auto& v : a.m) v *= s; return a; }
template <class T> inline M3<T> m3_add(M3<T> a, const M3<T>& b) { for (int k = 0; k < 9; ++k) a.m[k] += b.m[k]; return a; }
template <class T> inline M3<T> m3_sub(M3<T> a, const M3<T>& b) { for (int k = 0; k < 9; ++k) a.m[k] -= b.m[k]; return a; }
template <class T> inline V3<T> m3_mul(const M3<T>& a, V3<T> v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z,
            a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}

// Adjugate / determinant inverse, no pivoting and no singularity guard:
// restates math::inverse(Mat3) (src/math/matrix.cuh:219-235).
template <class T> inline M3<T> m3_inverse(const M3<T>& a) {
    const T* m = a.m;
    M3<T> c;
    c.m[0] = m[4] * m[8] - m[5] * m[7];
    c.m[1] = -m[1] * m[8] + m[2] * m[7];
    c.m[2] = m[1] * m[5] - m[2] * m[4];
    c.m[3] = -m[3] * m[8] + m[5] * m[6];
    c.m[4] = m[0] * m[8] - m[2] * m[6];
    c.m[5] = -m[0] * m[5] + m[2] * m[3];
    c.m[6] = m[3] * m[7] - m[4] * m[6];
    c.m[7] = -m[0] * m[7] + m[1] * m[6];
    c.m[8] = m[0] * m[4] - m[1] * m[3];
    const T det = m[0] * c.m[0] + m[1] * c.m[3] + m[2] * c.m[6];
    return m3_scale(c, T(1) / det);
}

// Constants of the reference path.
constexpr double kMinLength = 1.0e-6;   // spring_constraint.cu:35
constexpr double kMinNorm = 1.0e-6;     // triangle_stretch_constraint.cu:39
constexpr double kTinyMass = 1.0e-12;   // cloth_simulator.cu:76
constexpr double kHugeMass = 1.0e8;     // cloth_simulator.cu:76
constexpr double kGuard = 1.0e-37;      // pcg_solver.cu:421-423,451
constexpr int kChunk = 4096;            // fixed reduction chunk: results do not depend on thread count

// ---------------------------------------------------------------------------
// Host topology.  Restates ClothSimulator::set_mesh edge extraction
// (src/cloth/cloth_simulator.cu:331-349): per triangle the three undirected
// edges (min,max) of (x,y), (y,z), (z,x); sort; unique.
// ---------------------------------------------------------------------------
static std::vector<std::pair<int, int>> build_edges(const std::vector<int>& tris) {
    const size_t T = tris.size() / 3;
    std::vector<std::pair<int, int>> e;
    e.reserve(3 * T);
    for (size_t t = 0; t < T; ++t) {
        const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        e.emplace_back(std::min(a, b), std::max(a, b));
        e.emplace_back(std::min(b, c), std::max(b, c));
        e.emplace_back(std::min(c, a), std::max(c, a));
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    return e;
}

struct Pattern {
    int n = 0;
    std::vector<int> row_offsets;  // n+1
    std::vector<int> cols;         // nnz, ascending per row, no diagonal
};

// Restates BlockCSR3::build_topology (src/sparse/block_csr.cu:91-147): range
// check (std::out_of_range), drop r==c, lexicographic sort, unique, CSR.
static Pattern build_pattern(int n, const std::vector<int>& rows, const std::vector<int>& cols) {
    std::vector<std::pair<int, int>> tmp;
    tmp.reserve(rows.size());
    for (size_t k = 0; k < rows.size(); ++k) {
        const int r = rows[k], c = cols[k];
        if (r < 0 || r >= n || c < 0 || c >= n) throw std::out_of_range("pattern: pair out of range");
        if (r == c) continue;
        tmp.emplace_back(r, c);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    Pattern p;
    p.n = n;
    p.row_offsets.assign(size_t(n) + 1, 0);
    p.cols.resize(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) {
        p.row_offsets[tmp[i].first + 1] += 1;
        p.cols[i] = tmp[i].second;
    }
    for (int i = 0; i < n; ++i) p.row_offsets[i + 1] += p.row_offsets[i];
    return p;
}

// Restates BlockCSR3::resolve_slots (src/sparse/block_csr.cu:168-204):
// diagonal -> -r-1, off-diagonal -> index of c in row r by lower_bound,
// std::out_of_range when missing.
static int resolve_slot(const Pattern& p, int r, int c) {
    if (r < 0 || r >= p.n || c < 0 || c >= p.n) throw std::out_of_range("slots: pair out of range");
    if (r == c) return -r - 1;
    const auto first = p.cols.begin() + p.row_offsets[r];
    const auto last = p.cols.begin() + p.row_offsets[r + 1];
    const auto it = std::lower_bound(first, last, c);
    if (it == last || *it != c) throw std::out_of_range("slots: (row, col) not in topology");
    return int(it - p.cols.begin());
}

template <class T>
struct Cloth {
    // ---- configuration --------------------------------------------------
    int n = 0;
    std::vector<int> tris;                 // 3T
    std::vector<std::pair<int, int>> edges;
    // springs
    std::vector<int> spr_a, spr_b;
    std::vector<T> spr_rest, spr_k;
    // Baraff-Witkin triangle sets (TriangleStretchConstraint): stretch
    // (theta = 0) and shear (theta = pi/4, cloth_simulator.cu:423-449).
    struct TriSet {
        std::vector<int> idx;    // 3 per installed triangle
        std::vector<T> dminv;    // 4 per triangle, row-major 2x2 (rotation applied)
        std::vector<T> area, k;
        std::vector<int> slots;  // 9 per triangle
    };
    TriSet st, sh;
    // bending dihedrals (BendingConstraint, bending_constraint.h:1-129)
    std::vector<int> dih;                  // 4 per dihedral (v0, v1, v2, v3)
    std::vector<T> dih_rest;
    T bend_k = T(1.0e-3);                  // bending_constraint.h:120
    std::vector<int> dih_slots;            // 16 per dihedral
    // pins
    std::vector<int> pin_idx;
    std::vector<T> pin_tgt;                // 3 per pin
    T pin_k = T(1.0e6);                    // pin_constraint.h:82
    // Pin projection (extension, no reference counterpart; the reference's
    // pins are penalties, pin_constraint.h:5-15): each pinned particle's
    // displacement is fixed at d = target - x_n and eliminated from the
    // system, the particle ends the step exactly at its target.
    bool pin_projection = false;
    // material
    double gx = 0, gy = 0, gz = -9.81, damping = 0;  // cloth_material.h:20-43 defaults
    bool springs_active = false, force_active = false;
    double tol = 0.0;
    int max_iter = 50;                     // cloth_simulator.h:713
    int threads = 1;
    bool assemble_only = false;            // test hook: build the step's system, skip solve + finalize
    // static-shape contact (StaticContactSet, collision/static_contact.h:79-312)
    struct Plane { T n[3]; T d; };
    struct Box { T c[3], h[3], ex[3], ey[3], ez[3]; };
    std::vector<Plane> planes;
    std::vector<Box> boxes;
    T contact_h = T(0), contact_k = T(0), contact_mu = T(0), contact_eps = T(1.0e-4);
    bool contact_active() const { return contact_k > T(0) && (!planes.empty() || !boxes.empty()); }
    // per-particle contact cache of the current step
    std::vector<V3<T>> c_n, c_slip;
    std::vector<T> c_depth, c_slipn;

    // ---- derived --------------------------------------------------------
    bool topo_dirty = true;
    Pattern pat;
    std::vector<int> spr_slots, pin_slots;
    std::vector<M3<T>> diag, vals;
    std::vector<V3<T>> rhs, xn, xt;
    std::vector<V3<double>> dx;            // PCG iterate: always FP64 (see pcg())
    std::vector<T> mass;
    // PCG workspace
    std::vector<V3<T>> r, p, z, Ap;
    std::vector<M3<T>> minv;
    int last_iters = 0;
    double last_relres = 0.0;
    bool last_converged = false;
    int last_restarts = 0;

    int nthreads() const { return threads > 0 ? threads : 1; }

    // -- rest data --------------------------------------------------------
    // spring_rest_length_kernel (spring_constraint.cu:41-52).
    void build_springs(const std::vector<V3<T>>& pos, const T* kper, T kscalar) {
        const size_t E = edges.size();
        spr_a.resize(E); spr_b.resize(E); spr_rest.resize(E); spr_k.resize(E);
        for (size_t e = 0; e < E; ++e) {
            spr_a[e] = edges[e].first;
            spr_b[e] = edges[e].second;
            spr_rest[e] = len3(pos[spr_a[e]] - pos[spr_b[e]]);
            spr_k[e] = kper ? kper[e] : kscalar;
        }
        topo_dirty = true;
    }

    // triangle_rest_shape_kernel (triangle_stretch_constraint.cu:50-110):
    // rest frame [[lab, u_c], [0, v_c]], its inverse (identity when
    // det <= 1e-6), area lab v_c / 2, times R(theta)^-1 = [[c, s], [-s, c]]
    // with c, s the FP32 cos / sin of theta (:340-341).
    void build_tri_set(TriSet& S, const std::vector<V3<T>>& pos, const T* kper, T kscalar, float theta) {
        const size_t Tn = tris.size() / 3;
        const T cs = T(std::cos(theta)), sn = T(std::sin(theta));
        S.idx = tris;
        S.dminv.resize(4 * Tn); S.area.resize(Tn); S.k.resize(Tn);
        for (size_t t = 0; t < Tn; ++t) {
            const V3<T> a = pos[tris[3 * t]], b = pos[tris[3 * t + 1]], c = pos[tris[3 * t + 2]];
            const V3<T> ab = b - a, ac = c - a;
            const T lab = len3(ab), lac = len3(ac);
            const V3<T> abh = (lab > T(kMinNorm)) ? ab * (T(1) / lab) : V3<T>{T(1), T(0), T(0)};
            const T uc = dot3(abh, ac);
            const T vc2 = lac * lac - uc * uc;
            const T vc = (vc2 > T(0)) ? std::sqrt(vc2) : T(0);
            S.area[t] = T(0.5) * lab * vc;
            const T det = lab * vc;
            T d00 = 1, d01 = 0, d10 = 0, d11 = 1;
            if (det > T(kMinNorm)) {
                // Mat2 inverse (matrix.cuh:110-115) of [[lab, uc], [0, vc]].
                const T inv_det = T(1) / (lab * vc - uc * T(0));
                d00 = vc * inv_det;
                d01 = -uc * inv_det;
                d10 = -T(0) * inv_det;
                d11 = lab * inv_det;
            }
            S.dminv[4 * t + 0] = d00 * cs + d01 * (-sn);
            S.dminv[4 * t + 1] = d00 * sn + d01 * cs;
            S.dminv[4 * t + 2] = d10 * cs + d11 * (-sn);
            S.dminv[4 * t + 3] = d10 * sn + d11 * cs;
            S.k[t] = kper ? kper[t] : kscalar;
        }
        topo_dirty = true;
    }
    void build_stretch(const std::vector<V3<T>>& pos, const T* kper, T kscalar) {
        build_tri_set(st, pos, kper, kscalar, 0.0f);
    }
    // build_fem_shear_from_current_positions (cloth_simulator.cu:423-449).
    void build_shear(const std::vector<V3<T>>& pos, const T* kper, T kscalar) {
        build_tri_set(sh, pos, kper, kscalar, 0.7853981633974483f);
    }

    // Dihedral detection of build_bending_from_current_positions
    // (cloth_simulator.cu:451-514): each directed triangle edge records its
    // third vertex (a later triangle overwrites an earlier one); an edge
    // v0 < v1 with both directions present gives (v0, v1, third(v0->v1),
    // third(v1->v0)).  The reference emits them in hash-map order; here
    // they are sorted by (v0, v1) so the order is deterministic.
    static std::vector<int> find_dihedrals(const std::vector<int>& tri) {
        std::map<std::pair<int, int>, int> third;
        const size_t Tn = tri.size() / 3;
        for (size_t t = 0; t < Tn; ++t) {
            const int v[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
            third[{v[0], v[1]}] = v[2];
            third[{v[1], v[2]}] = v[0];
            third[{v[2], v[0]}] = v[1];
        }
        std::vector<int> out;
        for (const auto& kv : third) {
            const int v0 = kv.first.first, v1 = kv.first.second;
            if (v0 >= v1) continue;
            auto rev = third.find({v1, v0});
            if (rev == third.end()) continue;
            out.insert(out.end(), {v0, v1, kv.second, rev->second});
        }
        return out;
    }

    // compute_dihedral_frame (bending_constraint.cu:50-127): dihedral angle
    // theta and a_i = d theta / d x_i; invalid for a collapsed hinge or a
    // wing with no height (floor 1e-12).
    struct DihFrame {
        T theta;
        V3<T> a[4];
        bool valid;
    };
    static DihFrame dihedral_frame(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> v3) {
        DihFrame F;
        F.valid = false;
        F.theta = T(0);
        const T kFloor = T(1.0e-12);
        const V3<T> e10 = v1 - v0, e20 = v2 - v0, e30 = v3 - v0;
        const T l = len3(e10);
        if (l < kFloor) return F;
        const T inv_l = T(1) / l;
        const V3<T> e10h = e10 * inv_l;
        const V3<T> n1r = cross3(e20, e10), n2r = cross3(e10, e30);
        const T n1l = len3(n1r), n2l = len3(n2r);
        if (n1l < kFloor || n2l < kFloor) return F;
        const V3<T> n1 = n1r * (T(1) / n1l), n2 = n2r * (T(1) / n2l);
        T sin_t = dot3(cross3(n1, n2), e10h);
        if (sin_t > T(1)) sin_t = T(1);
        if (sin_t < T(-1)) sin_t = T(-1);
        const T cos_t = dot3(n1, n2);
        F.theta = std::atan2(sin_t, cos_t);
        const T inv_l2 = inv_l * inv_l;
        const T om1 = dot3(e10, e20) * inv_l2, om2 = dot3(e10, e30) * inv_l2;
        const T h1 = len3(e20 - e10 * om1), h2 = len3(e30 - e10 * om2);
        if (h1 < kFloor || h2 < kFloor) return F;
        const T ih1 = T(1) / h1, ih2 = T(1) / h2;
        const T t1[4] = {(om1 - T(1)) * ih1, -om1 * ih1, T(1) * ih1, T(0)};
        const T t2[4] = {(om2 - T(1)) * ih2, -om2 * ih2, T(0), T(1) * ih2};
        for (int i = 0; i < 4; ++i) F.a[i] = n1 * t1[i] + n2 * t2[i];
        F.valid = true;
        return F;
    }
    DihFrame dih_frame(size_t d, const std::vector<V3<T>>& x) const {
        return dihedral_frame(x[dih[4 * d]], x[dih[4 * d + 1]], x[dih[4 * d + 2]], x[dih[4 * d + 3]]);
    }
    // set_dihedrals_from_positions + bending_rest_angle_kernel
    // (bending_constraint.cu:136-152, 260-290).
    void build_bending(const std::vector<V3<T>>& pos, T k) {
        bend_k = k;
        dih = find_dihedrals(tris);
        dih_rest.resize(dih.size() / 4);
        for (size_t d = 0; d < dih_rest.size(); ++d) {
            const DihFrame F = dih_frame(d, pos);
            dih_rest[d] = F.valid ? F.theta : T(0);
        }
        topo_dirty = true;
    }

    // ensure_hessian_topology (cloth_simulator.cu:698-793): off-diagonal pairs
    // of every stretch triangle (a != b and v_a != v_b); spring pairs are
    // added too when springs are active (they are already a subset of the
    // triangle pairs for a triangle mesh, so the pattern is unchanged).
    void ensure_topology() {
        if (!topo_dirty && pat.n == n) return;
        std::vector<int> rows, cols;
        auto tri_pairs = [&](const TriSet& S) {
            const size_t Tn = S.idx.size() / 3;
            for (size_t t = 0; t < Tn; ++t) {
                const int v[3] = {S.idx[3 * t], S.idx[3 * t + 1], S.idx[3 * t + 2]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        if (a != b && v[a] != v[b]) { rows.push_back(v[a]); cols.push_back(v[b]); }
            }
        };
        tri_pairs(st);
        tri_pairs(sh);
        // bending: all 12 directed pairs of the 4 dihedral vertices (:741-761)
        for (size_t d = 0; d < dih.size() / 4; ++d) {
            const int* v = &dih[4 * d];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    if (a != b && v[a] != v[b]) { rows.push_back(v[a]); cols.push_back(v[b]); }
        }
        if (springs_active) {
            for (size_t e = 0; e < spr_a.size(); ++e) {
                rows.push_back(spr_a[e]); cols.push_back(spr_b[e]);
                rows.push_back(spr_b[e]); cols.push_back(spr_a[e]);
            }
        }
        pat = build_pattern(n, rows, cols);
        // ConstraintN<N>::bind_hessian_layout (constraint_n.h:115-153):
        // slot[c*N^2 + a*N + b] for (v_a, v_b).
        pin_slots.resize(pin_idx.size());
        for (size_t c = 0; c < pin_idx.size(); ++c) pin_slots[c] = resolve_slot(pat, pin_idx[c], pin_idx[c]);
        for (TriSet* S : {&st, &sh}) {
            const size_t Tn = S->idx.size() / 3;
            S->slots.resize(9 * Tn);
            for (size_t t = 0; t < Tn; ++t)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        S->slots[9 * t + 3 * a + b] = resolve_slot(pat, S->idx[3 * t + a], S->idx[3 * t + b]);
        }
        dih_slots.resize(dih.size() * 4);
        for (size_t d = 0; d < dih.size() / 4; ++d)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    dih_slots[16 * d + 4 * a + b] = resolve_slot(pat, dih[4 * d + a], dih[4 * d + b]);
        spr_slots.clear();
        if (springs_active) {
            spr_slots.resize(4 * spr_a.size());
            for (size_t e = 0; e < spr_a.size(); ++e) {
                const int a = spr_a[e], b = spr_b[e];
                spr_slots[4 * e + 0] = resolve_slot(pat, a, a);
                spr_slots[4 * e + 1] = resolve_slot(pat, a, b);
                spr_slots[4 * e + 2] = resolve_slot(pat, b, a);
                spr_slots[4 * e + 3] = resolve_slot(pat, b, b);
            }
        }
        diag.assign(size_t(n), m3_zero<T>());
        vals.assign(pat.cols.size(), m3_zero<T>());
        topo_dirty = false;
    }

    // scatter_hessian_block (src/sparse/block_csr_atomic.cuh:33-45), serial.
    void scatter(int slot, const M3<T>& blk) {
        M3<T>& dst = (slot < 0) ? diag[size_t(-slot - 1)] : vals[size_t(slot)];
        for (int k = 0; k < 9; ++k) dst.m[k] += blk.m[k];
    }

    // ---- element physics --------------------------------------------------
    // spring_gradient_kernel (spring_constraint.cu:76-103).
    V3<T> spring_grad(size_t e, const std::vector<V3<T>>& x) const {
        const V3<T> d = x[spr_a[e]] - x[spr_b[e]];
        T l = len3(d);
        if (l < T(kMinLength)) l = T(kMinLength);
        const T inv_l = T(1) / l;
        const V3<T> dh = d * inv_l;
        return dh * (spr_k[e] * (l - spr_rest[e]));
    }
    // spring_hessian_scatter_kernel block H (spring_constraint.cu:110-151).
    M3<T> spring_hess(size_t e, const std::vector<V3<T>>& x) const {
        const T k = spr_k[e], L = spr_rest[e];
        const V3<T> d = x[spr_a[e]] - x[spr_b[e]];
        T l = len3(d);
        if (l < T(kMinLength)) l = T(kMinLength);
        const T inv_l = T(1) / l;
        const V3<T> dh = d * inv_l;
        const T a = k * L * inv_l;
        const M3<T> dd = m3_outer(dh, dh);
        if (l < L) return m3_scale(dd, a);
        return m3_add(m3_scale(m3_ident<T>(), k - a), m3_scale(dd, a));
    }

    struct TriLocal {
        V3<T> wuh, wvh;
        T lu, lv, cu, cv, inv_lu, inv_lv, keff;
        T dwu[3], dwv[3];
    };
    // Shared prologue of triangle_gradient_kernel / triangle_hessian_scatter_kernel
    // (triangle_stretch_constraint.cu:153-196, 231-280).
    TriLocal tri_local(const TriSet& S, size_t t, const std::vector<V3<T>>& x) const {
        TriLocal L;
        const V3<T> x0 = x[S.idx[3 * t]], x1 = x[S.idx[3 * t + 1]], x2 = x[S.idx[3 * t + 2]];
        const T D00 = S.dminv[4 * t + 0], D01 = S.dminv[4 * t + 1];
        const T D10 = S.dminv[4 * t + 2], D11 = S.dminv[4 * t + 3];
        const V3<T> a = x1 - x0, b = x2 - x0;
        const V3<T> wu = a * D00 + b * D10;
        const V3<T> wv = a * D01 + b * D11;
        L.lu = len3(wu); L.lv = len3(wv);
        if (L.lu < T(kMinNorm)) L.lu = T(kMinNorm);
        if (L.lv < T(kMinNorm)) L.lv = T(kMinNorm);
        L.inv_lu = T(1) / L.lu;
        L.inv_lv = T(1) / L.lv;
        L.cu = L.lu - T(1);
        L.cv = L.lv - T(1);
        L.wuh = wu * L.inv_lu;
        L.wvh = wv * L.inv_lv;
        L.dwu[0] = -D00 - D10; L.dwu[1] = D00; L.dwu[2] = D10;
        L.dwv[0] = -D01 - D11; L.dwv[1] = D01; L.dwv[2] = D11;
        L.keff = S.area[t] * S.k[t];
        return L;
    }
    V3<T> tri_grad(const TriLocal& L, int a) const {
        return (L.wuh * (L.cu * L.dwu[a]) + L.wvh * (L.cv * L.dwv[a])) * L.keff;
    }
    M3<T> tri_hess(const TriLocal& L, int a, int b) const {
        const V3<T> dua = L.wuh * L.dwu[a], dub = L.wuh * L.dwu[b];
        const V3<T> dva = L.wvh * L.dwv[a], dvb = L.wvh * L.dwv[b];
        M3<T> H = m3_scale(m3_add(m3_outer(dua, dub), m3_outer(dva, dvb)), L.keff);
        if (L.cu > T(0)) {
            const M3<T> Pu = m3_sub(m3_ident<T>(), m3_outer(L.wuh, L.wuh));
            H = m3_add(H, m3_scale(Pu, L.keff * L.cu * L.inv_lu * L.dwu[a] * L.dwu[b]));
        }
        if (L.cv > T(0)) {
            const M3<T> Pv = m3_sub(m3_ident<T>(), m3_outer(L.wvh, L.wvh));
            H = m3_add(H, m3_scale(Pv, L.keff * L.cv * L.inv_lv * L.dwv[a] * L.dwv[b]));
        }
        return H;
    }

    // ---- sparse algebra ---------------------------------------------------
    // spmv_kernel with alpha=1, beta=0 (src/sparse/block_csr.cu:54-83).
    void spmv(const std::vector<V3<T>>& x, std::vector<V3<T>>& y) const {
        const int N = n;
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (int i = 0; i < N; ++i) {
            V3<T> acc = m3_mul(diag[i], x[i]);
            for (int k = pat.row_offsets[i]; k < pat.row_offsets[i + 1]; ++k)
                acc = acc + m3_mul(vals[k], x[pat.cols[k]]);
            y[i] = acc;
        }
    }
    // FP64 dot product with a fixed chunked summation order.
    template <class A, class B>
    double dot(const std::vector<V3<A>>& a, const std::vector<V3<B>>& b) const {
        const int N = n;
        const int nch = (N + kChunk - 1) / kChunk;
        std::vector<double> part(size_t(nch), 0.0);
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (int c = 0; c < nch; ++c) {
            double s = 0.0;
            const int e = std::min(N, (c + 1) * kChunk);
            for (int i = c * kChunk; i < e; ++i)
                s += double(a[i].x) * double(b[i].x) + double(a[i].y) * double(b[i].y) + double(a[i].z) * double(b[i].z);
            part[c] = s;
        }
        double s = 0.0;
        for (double v : part) s += v;
        return s;
    }
    // True residual b - A x with FP64 products and sums (residual_kernel,
    // pcg_solver.cu:53-71).  Writes it to `out` (rounded to T) and returns
    // |b - A x|^2 computed from the FP64 values.
    std::vector<V3<double>> tres;
    double true_residual(const std::vector<V3<T>>& b, const std::vector<V3<double>>& x, std::vector<V3<T>>& out) {
        const int N = n;
        tres.resize(N);
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (int i = 0; i < N; ++i) {
            double acc[3] = {0, 0, 0};
            auto add = [&](const M3<T>& A, const V3<double>& v) {
                for (int rr = 0; rr < 3; ++rr)
                    acc[rr] += double(A.m[3 * rr]) * v.x + double(A.m[3 * rr + 1]) * v.y + double(A.m[3 * rr + 2]) * v.z;
            };
            add(diag[i], x[i]);
            for (int k = pat.row_offsets[i]; k < pat.row_offsets[i + 1]; ++k) add(vals[k], x[pat.cols[k]]);
            tres[i] = {double(b[i].x) - acc[0], double(b[i].y) - acc[1], double(b[i].z) - acc[2]};
            out[i] = {T(tres[i].x), T(tres[i].y), T(tres[i].z)};
        }
        return dot(tres, tres);
    }

    // Block-Jacobi PCG.  Restates emit_pcg (pcg_solver.cu:467-593) including the
    // 1e-37 guards on sigma and rho, with FP64 scalars; adds the tolerance test
    // described in the file header (tol == 0 -> exactly max_iter iterations).
    //
    // The iterate x is FP64 in both variants.  With an FP32 iterate the
    // rounding of x alone leaves a true residual of about eps32*|A||x|, which
    // for stiff cloth (stiffness/inertia ratios of a few hundred) sits near
    // 1e-5 relative, above the 1e-6 target.  r, z, p, Ap and the SpMV stay in
    // T.  Convergence is judged on the true residual: when the recursive
    // residual passes the test, b - A x is recomputed in FP64; if that fails
    // the recursion restarts from it (residual replacement).
    void pcg(const std::vector<V3<T>>& b, std::vector<V3<double>>& x) {
        const int N = n;
        r.resize(N); p.resize(N); z.resize(N); Ap.resize(N); minv.resize(N);
        // residual_m_inv_kernel (pcg_solver.cu:29-51)
        for (int i = 0; i < N; ++i) minv[i] = m3_inverse(diag[i]);
        const double bb = std::max(dot(b, b), 1.0e-30);
        const double tol2 = tol * tol;
        auto precondition = [&]() {
#pragma omp parallel for schedule(static) num_threads(nthreads())
            for (int i = 0; i < N; ++i) z[i] = m3_mul(minv[i], r[i]);
        };
        last_restarts = 0;
        last_converged = false;
        int it = 0;
        double tt = true_residual(b, x, r);
        if (tol > 0.0 && tt <= tol2 * bb) {
            last_iters = 0;
            last_relres = std::sqrt(tt / bb);
            last_converged = true;
            return;
        }
        // jacobi_copy_dot_partial_kernel (pcg_solver.cu:386-404)
        precondition();
        p = z;
        double rho = dot(r, z);
        while (it < max_iter) {
            spmv(p, Ap);
            const double sigma = dot(p, Ap);
            const double alpha = (sigma > kGuard || sigma < -kGuard) ? rho / sigma : 0.0;
            const T al = T(alpha);
            // alpha_jacobi_dot_partial_kernel (pcg_solver.cu:406-434)
#pragma omp parallel for schedule(static) num_threads(nthreads())
            for (int i = 0; i < N; ++i) {
                x[i] = {x[i].x + double(p[i].x) * alpha, x[i].y + double(p[i].y) * alpha,
                        x[i].z + double(p[i].z) * alpha};
                r[i] = r[i] - Ap[i] * al;
            }
            precondition();
            const double rho_new = dot(r, z);
            // reduce_rho_beta_kernel (pcg_solver.cu:436-454)
            const double beta = (rho > kGuard || rho < -kGuard) ? rho_new / rho : 0.0;
            rho = rho_new;
            ++it;
            if (tol > 0.0) {
                const double rr = dot(r, r);
                if (rr <= tol2 * bb) {
                    tt = true_residual(b, x, r);
                    if (tt <= tol2 * bb) {
                        last_iters = it;
                        last_relres = std::sqrt(tt / bb);
                        last_converged = true;
                        return;
                    }
                    // Residual replacement: restart the recursion from b - A x.
                    precondition();
                    p = z;
                    rho = dot(r, z);
                    ++last_restarts;
                    continue;
                }
            }
            // update_p_kernel (pcg_solver.cu:456-464)
            const T be = T(beta);
#pragma omp parallel for schedule(static) num_threads(nthreads())
            for (int i = 0; i < N; ++i) p[i] = z[i] + p[i] * be;
        }
        last_iters = it;
        tt = true_residual(b, x, Ap);
        last_relres = std::sqrt(tt / bb);
        last_converged = (tol > 0.0) && (last_relres <= tol);
    }

    // ---- static-shape contact ---------------------------------------------
    // box_signed_dist (static_contact.cu:55-99): local coordinates q,
    // d_i = |q_i| - h_i; inside: the largest d_i and that face's normal;
    // outside: |max(d, 0)| and the normalised clamped local offset.
    static T box_sdf(const Box& b, V3<T> x, V3<T>& n) {
        const V3<T> ex{b.ex[0], b.ex[1], b.ex[2]}, ey{b.ey[0], b.ey[1], b.ey[2]}, ez{b.ez[0], b.ez[1], b.ez[2]};
        const V3<T> rel = x - V3<T>{b.c[0], b.c[1], b.c[2]};
        const T qx = dot3(rel, ex), qy = dot3(rel, ey), qz = dot3(rel, ez);
        const T dx = std::fabs(qx) - b.h[0], dy = std::fabs(qy) - b.h[1], dz = std::fabs(qz) - b.h[2];
        if (dx <= T(0) && dy <= T(0) && dz <= T(0)) {
            if (dx >= dy && dx >= dz) { n = (qx >= T(0)) ? ex : ex * T(-1); return dx; }
            if (dy >= dz) { n = (qy >= T(0)) ? ey : ey * T(-1); return dy; }
            n = (qz >= T(0)) ? ez : ez * T(-1);
            return dz;
        }
        const T cx = dx > T(0) ? dx : T(0), cy = dy > T(0) ? dy : T(0), cz = dz > T(0) ? dz : T(0);
        const T dist = std::sqrt(cx * cx + cy * cy + cz * cz);
        const T inv = T(1) / std::max(dist, T(1.0e-30));
        const T sx = qx > T(0) ? cx : -cx, sy = qy > T(0) ? cy : -cy, sz = qz > T(0) ? cz : -cz;
        n = ex * (sx * inv) + ey * (sy * inv) + ez * (sz * inv);
        return dist;
    }
    // detect_kernel (static_contact.cu:107-165): the deepest active contact
    // (largest thickness - signed distance > 0) and the lagged tangential
    // slip u_t = (v dt) - n (n . v dt).
    void contact_detect(int i, const V3<T>& x, const V3<T>& v, T dt) {
        T best = T(0);
        V3<T> bn{0, 0, 0};
        for (const Plane& pl : planes) {
            const V3<T> n{pl.n[0], pl.n[1], pl.n[2]};
            const T depth = contact_h - (dot3(n, x) + pl.d);
            if (depth > best) { best = depth; bn = n; }
        }
        for (const Box& b : boxes) {
            V3<T> n;
            const T depth = contact_h - box_sdf(b, x, n);
            if (depth > best) { best = depth; bn = n; }
        }
        c_depth[i] = best;
        c_n[i] = bn;
        V3<T> ut{0, 0, 0};
        T un = T(0);
        if (best > T(0) && dt > T(0)) {
            const V3<T> dxv = v * dt;
            ut = dxv - bn * dot3(bn, dxv);
            un = len3(ut);
        }
        c_slip[i] = ut;
        c_slipn[i] = un;
    }
    // ipc_f1_sf_over_x (friction.cuh:28-36).
    static T f1_sf_over_x(T u, T eps) { return (u > eps) ? T(1) / u : (-u / eps + T(2)) / eps; }

    // ---- the step -------------------------------------------------------
    // ClothSimulator::step (src/cloth/cloth_simulator.cu:814-1261), PCG branch.
    void step(std::vector<V3<T>>& pos, std::vector<V3<T>>& vel, const std::vector<T>* inv_mass,
              const std::vector<V3<T>>* force, T dt) {
        if (n <= 0) return;
        if (!(dt > T(0))) throw std::invalid_argument("dt must be positive");
        const int N = n;
        // resize_work_buffers (:795-812): dx zeroed only when (re)sized.
        if (int(dx.size()) != N) dx.assign(size_t(N), V3<double>{0, 0, 0});
        xn.resize(N); xt.resize(N); rhs.resize(N); mass.resize(N);
        const V3<T> g{T(gx), T(gy), T(gz)};
        // predictor (:886-900): x_n <- pos; mass_from_inv_mass_kernel (:66-77);
        // compute_x_tilde_kernel (:83-106).
        for (int i = 0; i < N; ++i) {
            xn[i] = pos[i];
            if (!inv_mass) mass[i] = T(1);
            else {
                const T w = (*inv_mass)[i];
                mass[i] = (w > T(kTinyMass)) ? T(1) / w : T(kHugeMass);
            }
            V3<T> acc = g;
            if (force_active && force) {
                const T im = (mass[i] > T(0)) ? T(1) / mass[i] : T(0);
                acc = acc + (*force)[i] * im;
            }
            xt[i] = pos[i] + vel[i] * dt + acc * (dt * dt);
        }
        // gradient at x_n (:914-1110): pins, springs (extension), stretch.
        std::fill(rhs.begin(), rhs.end(), V3<T>{0, 0, 0});
        for (size_t c = 0; c < pin_idx.size() && !pin_projection; ++c) {  // pin_gradient_kernel (pin_constraint.cu:61-75)
            const int i = pin_idx[c];
            const V3<T> t{pin_tgt[3 * c], pin_tgt[3 * c + 1], pin_tgt[3 * c + 2]};
            rhs[i] = rhs[i] + (xn[i] - t) * pin_k;
        }
        if (springs_active) {
            for (size_t e = 0; e < spr_a.size(); ++e) {
                const V3<T> ge = spring_grad(e, xn);
                rhs[spr_a[e]] = rhs[spr_a[e]] + ge;
                rhs[spr_b[e]] = rhs[spr_b[e]] - ge;
            }
        }
        for (const TriSet* S : {&st, &sh}) {  // stretch, then shear (:922-923)
            const size_t Tn = S->idx.size() / 3;
            for (size_t t = 0; t < Tn; ++t) {
                const TriLocal L = tri_local(*S, t, xn);
                for (int a = 0; a < 3; ++a) {
                    const int v = S->idx[3 * t + a];
                    rhs[v] = rhs[v] + tri_grad(L, a);
                }
            }
        }
        for (size_t d = 0; d < dih.size() / 4; ++d) {  // bending_gradient_kernel (bending_constraint.cu:175-203)
            const DihFrame F = dih_frame(d, xn);
            if (!F.valid) continue;
            const T coeff = bend_k * (F.theta - dih_rest[d]);
            for (int i = 0; i < 4; ++i) rhs[dih[4 * d + i]] = rhs[dih[4 * d + i]] + F.a[i] * coeff;
        }
        // static contact (:989-1008): detect at x_n with the current velocity;
        // scatter_gradient_kernel (static_contact.cu:182-213)
        const bool contact = contact_active();
        if (contact) {
            c_n.assign(size_t(N), V3<T>{0, 0, 0});
            c_slip.assign(size_t(N), V3<T>{0, 0, 0});
            c_depth.assign(size_t(N), T(0));
            c_slipn.assign(size_t(N), T(0));
            for (int i = 0; i < N; ++i) {
                contact_detect(i, xn[i], vel[i], dt);
                const T depth = c_depth[i];
                if (depth <= T(0)) continue;
                rhs[i] = rhs[i] + c_n[i] * (-contact_k * depth);
                if (contact_mu > T(0)) {
                    const T alpha = contact_mu * (contact_k * depth) * f1_sf_over_x(c_slipn[i], contact_eps);
                    rhs[i] = rhs[i] + c_slip[i] * (-alpha);
                }
            }
        }
        // assemble_rhs_kernel (:127-142)
        const T inv_dt2 = T(1) / (dt * dt);
        for (int i = 0; i < N; ++i) {
            const T mi = mass[i] * inv_dt2;
            rhs[i] = (xt[i] - xn[i]) * mi - rhs[i];
        }
        // apply_coulomb_friction_kernel (static_contact.cu:325-380): project
        // the inertial + elastic force's tangential part onto the cone of
        // radius mu * max(0, -F0 . n).
        if (contact && contact_mu > T(0)) {
            for (int i = 0; i < N; ++i) {
                const T depth = c_depth[i];
                if (depth <= T(0)) continue;
                const V3<T> n = c_n[i];
                const V3<T> F0 = rhs[i] - n * (contact_k * depth);
                const T F0n = dot3(F0, n);
                const T fn = std::max(T(0), -F0n);
                if (fn <= T(0)) continue;
                const V3<T> F0t = F0 - n * F0n;
                const T ft2 = dot3(F0t, F0t);
                if (ft2 <= T(0)) continue;
                const T ft = std::sqrt(ft2);
                const T cone = contact_mu * fn;
                rhs[i] = rhs[i] - F0t * ((ft <= cone) ? T(1) : cone / ft);
            }
        }
        // topology (:1118-1121) and Hessian (:1129-1217)
        ensure_topology();
        std::fill(diag.begin(), diag.end(), m3_zero<T>());
        std::fill(vals.begin(), vals.end(), m3_zero<T>());
        if (!pin_projection)
            for (size_t c = 0; c < pin_idx.size(); ++c)  // pin_hessian_scatter_kernel (pin_constraint.cu:79-92)
                scatter(pin_slots[c], m3_scale(m3_ident<T>(), pin_k));
        if (springs_active) {
            for (size_t e = 0; e < spr_a.size(); ++e) {
                const M3<T> H = spring_hess(e, xn);
                const M3<T> nH = m3_scale(H, T(-1));
                scatter(spr_slots[4 * e + 0], H);
                scatter(spr_slots[4 * e + 1], nH);
                scatter(spr_slots[4 * e + 2], nH);
                scatter(spr_slots[4 * e + 3], H);
            }
        }
        for (const TriSet* S : {&st, &sh}) {
            const size_t Tn = S->idx.size() / 3;
            for (size_t t = 0; t < Tn; ++t) {
                const TriLocal L = tri_local(*S, t, xn);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) scatter(S->slots[9 * t + 3 * a + b], tri_hess(L, a, b));
            }
        }
        for (size_t d = 0; d < dih.size() / 4; ++d) {  // bending_hessian_scatter_kernel (:214-246)
            const DihFrame F = dih_frame(d, xn);
            if (!F.valid) continue;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    scatter(dih_slots[16 * d + 4 * i + j], m3_scale(m3_outer(F.a[i], F.a[j]), bend_k));
        }
        for (int i = 0; i < N; ++i) {  // add_inertia_diag_kernel (:199-210)
            const T d = mass[i] * inv_dt2;
            diag[i].m[0] += d; diag[i].m[4] += d; diag[i].m[8] += d;
        }
        if (contact) {  // bake_diag_kernel (static_contact.cu:232-300), after inertia (:1189-1193)
            for (int i = 0; i < N; ++i) {
                const T depth = c_depth[i];
                if (depth <= T(0)) continue;
                const V3<T> n = c_n[i];
                M3<T> H = m3_scale(m3_outer(n, n), contact_k);
                if (contact_mu > T(0)) {
                    const T alpha = contact_mu * (contact_k * depth) * f1_sf_over_x(c_slipn[i], contact_eps);
                    H = m3_add(H, m3_scale(m3_sub(m3_ident<T>(), m3_outer(n, n)), alpha));
                }
                diag[i] = m3_add(diag[i], H);
            }
        }
        // Pin projection: eliminate the pinned displacements d_p.  A free row
        // j moves its coupling to a pinned column into the right-hand side
        // (rhs_j -= A_jp d_p, A_jp = 0); a pinned row becomes I dx_p = d_p.
        // The iterate starts at d_p, so the pinned residual is 0 throughout.
        std::vector<int> pin_of;
        std::vector<V3<T>> dp;
        if (pin_projection) {
            pin_of.assign(size_t(N), -1);
            for (size_t c = 0; c < pin_idx.size(); ++c)
                if (pin_of[size_t(pin_idx[c])] < 0) pin_of[size_t(pin_idx[c])] = int(c);
            dp.assign(size_t(N), V3<T>{0, 0, 0});
            for (int i = 0; i < N; ++i)
                if (pin_of[i] >= 0) {
                    const size_t c = size_t(pin_of[i]);
                    dp[i] = V3<T>{pin_tgt[3 * c], pin_tgt[3 * c + 1], pin_tgt[3 * c + 2]} - xn[i];
                }
            for (int j = 0; j < N; ++j) {
                const bool pj = pin_of[j] >= 0;
                for (int k = pat.row_offsets[j]; k < pat.row_offsets[j + 1]; ++k) {
                    const int c = pat.cols[k];
                    if (!pj && pin_of[c] >= 0) rhs[j] = rhs[j] - m3_mul(vals[k], dp[c]);
                    if (pj || pin_of[c] >= 0) vals[k] = m3_zero<T>();
                }
                if (pj) {
                    diag[j] = m3_ident<T>();
                    rhs[j] = dp[j];
                }
            }
        }
        if (assemble_only) return;  // system only: no solve, no state change
        if (pin_projection)
            for (int i = 0; i < N; ++i)
                if (pin_of[i] >= 0) dx[i] = V3<double>{double(dp[i].x), double(dp[i].y), double(dp[i].z)};
        // PCG with warm start dx (:1219-1249)
        pcg(rhs, dx);
        // finalize_step_kernel (:222-244)
        const T inv_dt = T(1) / dt;
        const bool damp = damping > 0.0;
        const T decay = damp ? T(std::exp(-T(damping) * dt)) : T(1);
        for (int i = 0; i < N; ++i) {
            // x_{n+1} = x_n + dx, v = dx / dt, in T like the reference kernel.
            V3<T> d{T(dx[i].x), T(dx[i].y), T(dx[i].z)};
            if (pin_projection && pin_of[i] >= 0) {  // exactly at the target
                const size_t c = size_t(pin_of[i]);
                const V3<T> t{pin_tgt[3 * c], pin_tgt[3 * c + 1], pin_tgt[3 * c + 2]};
                d = t - xn[i];
                pos[i] = t;
            } else {
                pos[i] = xn[i] + d;
            }
            V3<T> v = d * inv_dt;
            if (damp) v = v * decay;
            vel[i] = v;
        }
    }
};

}  // namespace orc

// ===========================================================================
// C ABI for the Python test harness (ctypes).
// ===========================================================================
using namespace orc;

namespace {
thread_local std::string g_err;

struct Handle {
    int precision;  // 64 or 32
    Cloth<double> d;
    Cloth<float> f;
};

template <class F> int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::out_of_range& e) {
        g_err = e.what();
        return 3;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

template <class T> std::vector<V3<T>> to_v3(const double* p, int n) {
    std::vector<V3<T>> v(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) v[i] = {T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])};
    return v;
}
template <class T> void from_v3(const std::vector<V3<T>>& v, double* p) {
    for (size_t i = 0; i < v.size(); ++i) {
        p[3 * i] = double(v[i].x); p[3 * i + 1] = double(v[i].y); p[3 * i + 2] = double(v[i].z);
    }
}
template <class T> std::vector<T> to_vec(const double* p, int n) {
    std::vector<T> v(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) v[i] = T(p[i]);
    return v;
}

template <class F> int dispatch(void* h, F&& f) {
    Handle* H = static_cast<Handle*>(h);
    return guarded([&] {
        if (H->precision == 32) f(H->f);
        else f(H->d);
    });
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(int precision) {
    Handle* h = new Handle();
    h->precision = (precision == 32) ? 32 : 64;
    return h;
}
void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_set_mesh(void* h, const int* tris, int ntri, int n) {
    return dispatch(h, [&](auto& c) {
        if (ntri < 0 || n < 0) throw std::invalid_argument("negative size");
        c.n = n;
        c.tris.assign(tris, tris + 3 * size_t(ntri));
        c.edges = build_edges(c.tris);
        c.topo_dirty = true;
    });
}
int orc_num_edges(void* h) {
    Handle* H = static_cast<Handle*>(h);
    return int(H->precision == 32 ? H->f.edges.size() : H->d.edges.size());
}
int orc_get_edges(void* h, int* out) {
    return dispatch(h, [&](auto& c) {
        for (size_t e = 0; e < c.edges.size(); ++e) { out[2 * e] = c.edges[e].first; out[2 * e + 1] = c.edges[e].second; }
    });
}
int orc_build_springs(void* h, const double* pos, const double* k_per_edge, double k) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        std::vector<T> kv;
        if (k_per_edge) kv = to_vec<T>(k_per_edge, int(c.edges.size()));
        c.build_springs(P, k_per_edge ? kv.data() : nullptr, T(k));
    });
}
int orc_build_stretch(void* h, const double* pos, const double* k_per_tri, double k) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        std::vector<T> kv;
        if (k_per_tri) kv = to_vec<T>(k_per_tri, int(c.tris.size() / 3));
        c.build_stretch(P, k_per_tri ? kv.data() : nullptr, T(k));
    });
}
int orc_set_pins(void* h, const int* idx, const double* targets, int npin, double k) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        c.pin_idx.assign(idx, idx + npin);
        c.pin_tgt = to_vec<T>(targets, 3 * npin);
        c.pin_k = T(k);
        c.topo_dirty = true;
    });
}
int orc_set_pin_projection(void* h, int on) {
    return dispatch(h, [&](auto& c) { c.pin_projection = on != 0; });
}
int orc_set_params(void* h, double gx, double gy, double gz, double damping, int springs_active,
                   int force_active, double tol, int max_iter, int threads) {
    return dispatch(h, [&](auto& c) {
        c.gx = gx; c.gy = gy; c.gz = gz; c.damping = damping;
        if (bool(springs_active) != c.springs_active) c.topo_dirty = true;
        c.springs_active = springs_active != 0;
        c.force_active = force_active != 0;
        c.tol = tol;
        c.max_iter = max_iter > 0 ? max_iter : 1;
        c.threads = threads;
    });
}
// Warm start dx (3n doubles) used by the next solve / left by the last one.
int orc_set_dx(void* h, const double* dx) {
    return dispatch(h, [&](auto& c) {
        c.dx.assign(size_t(c.n), V3<double>{0, 0, 0});
        for (int i = 0; i < c.n; ++i) c.dx[i] = {dx[3 * i], dx[3 * i + 1], dx[3 * i + 2]};
    });
}
int orc_get_dx(void* h, double* out) {
    return dispatch(h, [&](auto& c) {
        for (size_t i = 0; i < c.dx.size(); ++i) {
            out[3 * i] = c.dx[i].x; out[3 * i + 1] = c.dx[i].y; out[3 * i + 2] = c.dx[i].z;
        }
    });
}
int orc_reset_warm_start(void* h) {
    return dispatch(h, [&](auto& c) { c.dx.clear(); });
}
// Advance `nsteps` steps.  pos/vel (3n doubles) are read and written.
// iters/relres (length nsteps, may be null) receive per-step PCG stats.
int orc_step(void* h, double* pos, double* vel, const double* inv_mass, const double* force, double dt,
             int nsteps, int* iters, double* relres, int* converged) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        auto V = to_v3<T>(vel, c.n);
        std::vector<T> im;
        if (inv_mass) im = to_vec<T>(inv_mass, c.n);
        std::vector<V3<T>> F;
        if (force) F = to_v3<T>(force, c.n);
        for (int s = 0; s < nsteps; ++s) {
            c.step(P, V, inv_mass ? &im : nullptr, force ? &F : nullptr, T(dt));
            if (iters) iters[s] = c.last_iters;
            if (relres) relres[s] = c.last_relres;
            if (converged) converged[s] = c.last_converged ? 1 : 0;
        }
        from_v3(P, pos);
        from_v3(V, vel);
    });
}
// Assemble the step's linear system at (pos, vel) without solving or
// changing any state; read it back with orc_get_system.
int orc_assemble(void* h, const double* pos, const double* vel, const double* inv_mass, const double* force,
                 double dt) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        auto V = to_v3<T>(vel, c.n);
        std::vector<T> im;
        if (inv_mass) im = to_vec<T>(inv_mass, c.n);
        std::vector<V3<T>> F;
        if (force) F = to_v3<T>(force, c.n);
        c.assemble_only = true;
        try {
            c.step(P, V, inv_mass ? &im : nullptr, force ? &F : nullptr, T(dt));
        } catch (...) {
            c.assemble_only = false;
            throw;
        }
        c.assemble_only = false;
    });
}
int orc_nnz(void* h) {
    Handle* H = static_cast<Handle*>(h);
    return int(H->precision == 32 ? H->f.pat.cols.size() : H->d.pat.cols.size());
}
int orc_build_topology(void* h) {
    return dispatch(h, [&](auto& c) { c.ensure_topology(); });
}
int orc_get_pattern(void* h, int* row_offsets, int* cols) {
    return dispatch(h, [&](auto& c) {
        std::copy(c.pat.row_offsets.begin(), c.pat.row_offsets.end(), row_offsets);
        std::copy(c.pat.cols.begin(), c.pat.cols.end(), cols);
    });
}
int orc_get_slots(void* h, int* tri_slots, int* spr_slots, int* pin_slots) {
    return dispatch(h, [&](auto& c) {
        if (tri_slots) std::copy(c.st.slots.begin(), c.st.slots.end(), tri_slots);
        if (spr_slots) std::copy(c.spr_slots.begin(), c.spr_slots.end(), spr_slots);
        if (pin_slots) std::copy(c.pin_slots.begin(), c.pin_slots.end(), pin_slots);
    });
}
// Last solve's system: diag (9n), values (9 nnz), rhs (3n), dx (3n).
int orc_get_system(void* h, double* diag, double* values, double* rhs, double* dx) {
    return dispatch(h, [&](auto& c) {
        for (size_t i = 0; i < c.diag.size(); ++i)
            for (int k = 0; k < 9; ++k) diag[9 * i + k] = double(c.diag[i].m[k]);
        for (size_t i = 0; i < c.vals.size(); ++i)
            for (int k = 0; k < 9; ++k) values[9 * i + k] = double(c.vals[i].m[k]);
        from_v3(c.rhs, rhs);
        from_v3(c.dx, dx);
    });
}
int orc_get_rest(void* h, double* spr_rest, double* tri_dminv, double* tri_area) {
    return dispatch(h, [&](auto& c) {
        if (spr_rest) for (size_t e = 0; e < c.spr_rest.size(); ++e) spr_rest[e] = double(c.spr_rest[e]);
        if (tri_dminv) for (size_t i = 0; i < c.st.dminv.size(); ++i) tri_dminv[i] = double(c.st.dminv[i]);
        if (tri_area) for (size_t i = 0; i < c.st.area.size(); ++i) tri_area[i] = double(c.st.area[i]);
    });
}
int orc_build_shear(void* h, const double* pos, const double* k_per_tri, double k) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        std::vector<T> kv;
        if (k_per_tri) kv = to_vec<T>(k_per_tri, int(c.tris.size() / 3));
        c.build_shear(P, k_per_tri ? kv.data() : nullptr, T(k));
    });
}
int orc_build_bending(void* h, const double* pos, double k) {
    return dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        c.build_bending(P, T(k));
    });
}
int orc_num_dihedrals(void* h) {
    Handle* H = static_cast<Handle*>(h);
    return int(H->precision == 32 ? H->f.dih.size() / 4 : H->d.dih.size() / 4);
}
// Dihedrals (4 per), rest angles, shear triangle slots (9 per), dihedral
// slots (16 per); null pointers are skipped.
int orc_get_bending(void* h, int* dih, double* rest, int* shear_slots, int* dih_slots) {
    return dispatch(h, [&](auto& c) {
        if (dih) std::copy(c.dih.begin(), c.dih.end(), dih);
        if (rest) for (size_t d = 0; d < c.dih_rest.size(); ++d) rest[d] = double(c.dih_rest[d]);
        if (shear_slots) std::copy(c.sh.slots.begin(), c.sh.slots.end(), shear_slots);
        if (dih_slots) std::copy(c.dih_slots.begin(), c.dih_slots.end(), dih_slots);
    });
}
int orc_last_stats(void* h, int* iters, double* relres, int* converged, int* restarts) {
    return dispatch(h, [&](auto& c) {
        if (iters) *iters = c.last_iters;
        if (relres) *relres = c.last_relres;
        if (converged) *converged = c.last_converged ? 1 : 0;
        if (restarts) *restarts = c.last_restarts;
    });
}

// ---- element-level entry points (FP64) for finite-difference tests ----------
// Spring element (a=0, b=1): positions x (6), rest L, stiffness k ->
// gradient wrt x_a (3) and the block H (9).
void orc_spring_element(const double* x, double L, double k, double* grad_a, double* H) {
    Cloth<double> c;
    c.spr_a = {0}; c.spr_b = {1}; c.spr_rest = {L}; c.spr_k = {k};
    std::vector<V3<double>> X = {{x[0], x[1], x[2]}, {x[3], x[4], x[5]}};
    const V3<double> g = c.spring_grad(0, X);
    grad_a[0] = g.x; grad_a[1] = g.y; grad_a[2] = g.z;
    const M3<double> h = c.spring_hess(0, X);
    std::copy(h.m, h.m + 9, H);
}
// Triangle element: rest positions xr (9) define Dm_inv/area; evaluate at x (9).
// Outputs grad (9 = 3 vertices) and hess (81 = 3x3 blocks of 3x3, block (a,b)
// at offset 27a + 9b), plus dminv (4) and area (1).
void orc_triangle_element(const double* xr, const double* x, double k, float theta, double* grad, double* hess,
                          double* dminv, double* area) {
    Cloth<double> c;
    c.tris = {0, 1, 2};
    std::vector<V3<double>> R = {{xr[0], xr[1], xr[2]}, {xr[3], xr[4], xr[5]}, {xr[6], xr[7], xr[8]}};
    c.build_tri_set(c.st, R, nullptr, k, theta);
    std::vector<V3<double>> X = {{x[0], x[1], x[2]}, {x[3], x[4], x[5]}, {x[6], x[7], x[8]}};
    const auto L = c.tri_local(c.st, 0, X);
    for (int a = 0; a < 3; ++a) {
        const V3<double> g = c.tri_grad(L, a);
        grad[3 * a] = g.x; grad[3 * a + 1] = g.y; grad[3 * a + 2] = g.z;
        for (int b = 0; b < 3; ++b) {
            const M3<double> h = c.tri_hess(L, a, b);
            std::copy(h.m, h.m + 9, hess + 27 * a + 9 * b);
        }
    }
    std::copy(c.st.dminv.begin(), c.st.dminv.end(), dminv);
    *area = c.st.area[0];
}
// Static shapes and contact parameters (add_static_plane / add_static_box,
// cloth_simulator.h:398-450).  box: center, half extents, ex, ey, ez (15).
int orc_add_plane(void* h, const double* n, double d) {
    return dispatch(h, [&](auto& c) {
        using Cl = std::remove_reference_t<decltype(c)>;
        typename Cl::Plane p;
        for (int k = 0; k < 3; ++k) p.n[k] = n[k];
        p.d = d;
        c.planes.push_back(p);
    });
}
int orc_add_box(void* h, const double* b15) {
    return dispatch(h, [&](auto& c) {
        using Cl = std::remove_reference_t<decltype(c)>;
        typename Cl::Box b;
        for (int k = 0; k < 3; ++k) {
            b.c[k] = b15[k]; b.h[k] = b15[3 + k]; b.ex[k] = b15[6 + k]; b.ey[k] = b15[9 + k]; b.ez[k] = b15[12 + k];
        }
        c.boxes.push_back(b);
    });
}
int orc_clear_shapes(void* h) {
    return dispatch(h, [&](auto& c) { c.planes.clear(); c.boxes.clear(); });
}
int orc_set_contact(void* h, double thickness, double stiffness, double friction, double eps) {
    return dispatch(h, [&](auto& c) {
        c.contact_h = thickness; c.contact_k = stiffness; c.contact_mu = friction; c.contact_eps = eps;
    });
}
// Signed distance and outward normal of a box (15 values as above).
double orc_box_sdf(const double* b15, const double* x, double* n) {
    Cloth<double>::Box b;
    for (int k = 0; k < 3; ++k) {
        b.c[k] = b15[k]; b.h[k] = b15[3 + k]; b.ex[k] = b15[6 + k]; b.ey[k] = b15[9 + k]; b.ez[k] = b15[12 + k];
    }
    V3<double> nn;
    const double d = Cloth<double>::box_sdf(b, V3<double>{x[0], x[1], x[2]}, nn);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    return d;
}
// One dihedral (x: 4 x 3): angle theta, a_i = d theta / d x_i (4 x 3),
// validity (compute_dihedral_frame, bending_constraint.cu:50-127).
int orc_dihedral_element(const double* x, double* theta, double* a) {
    std::vector<V3<double>> X(4);
    for (int i = 0; i < 4; ++i) X[size_t(i)] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const auto F = Cloth<double>::dihedral_frame(X[0], X[1], X[2], X[3]);
    *theta = F.theta;
    for (int i = 0; i < 4; ++i) {
        a[3 * i] = F.valid ? F.a[i].x : 0.0;
        a[3 * i + 1] = F.valid ? F.a[i].y : 0.0;
        a[3 * i + 2] = F.valid ? F.a[i].z : 0.0;
    }
    return F.valid ? 1 : 0;
}
// Dihedrals of a triangle list (4 per), sorted by hinge; returns the count
// and fills `out` when it is not null.
int orc_find_dihedrals(const int* tris, int ntri, int* out) {
    const std::vector<int> t(tris, tris + 3 * size_t(ntri));
    const std::vector<int> d = Cloth<double>::find_dihedrals(t);
    if (out) std::copy(d.begin(), d.end(), out);
    return int(d.size() / 4);
}
void orc_mat3_inverse(const double* m, double* out) {
    M3<double> a;
    std::copy(m, m + 9, a.m);
    const M3<double> r = m3_inverse(a);
    std::copy(r.m, r.m + 9, out);
}
// Stand-alone block-Jacobi PCG on a given split BSR system (FP64), for tests
// against a dense library solve.  x holds the initial guess and the result.
int orc_pcg_solve(int n, const double* diag, const int* row_offsets, const int* cols, const double* values,
                  const double* b, double* x, double tol, int max_iter, int* iters, double* relres) {
    return guarded([&] {
        Cloth<double> c;
        c.n = n;
        c.pat.n = n;
        c.pat.row_offsets.assign(row_offsets, row_offsets + n + 1);
        const int nnz = row_offsets[n];
        c.pat.cols.assign(cols, cols + nnz);
        c.diag.resize(n);
        c.vals.resize(nnz);
        for (int i = 0; i < n; ++i) std::copy(diag + 9 * i, diag + 9 * i + 9, c.diag[i].m);
        for (int i = 0; i < nnz; ++i) std::copy(values + 9 * i, values + 9 * i + 9, c.vals[i].m);
        c.tol = tol;
        c.max_iter = max_iter;
        auto B = to_v3<double>(b, n);
        auto X = to_v3<double>(x, n);
        c.pcg(B, X);
        from_v3(X, x);
        if (iters) *iters = c.last_iters;
        if (relres) *relres = c.last_relres;
    });
}
// Wall-clock timing helper for the CPU baseline: run nsteps steps and return
// seconds spent inside step() (excludes array conversion).
double orc_time_steps(void* h, double* pos, double* vel, const double* inv_mass, const double* force, double dt,
                      int nsteps, int* total_iters) {
    double secs = -1.0;
    int rc = dispatch(h, [&](auto& c) {
        using T = std::remove_reference_t<decltype(c.spr_k[0])>;
        auto P = to_v3<T>(pos, c.n);
        auto V = to_v3<T>(vel, c.n);
        std::vector<T> im;
        if (inv_mass) im = to_vec<T>(inv_mass, c.n);
        std::vector<V3<T>> F;
        if (force) F = to_v3<T>(force, c.n);
        int it = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < nsteps; ++s) {
            c.step(P, V, inv_mass ? &im : nullptr, force ? &F : nullptr, T(dt));
            it += c.last_iters;
        }
        const auto t1 = std::chrono::steady_clock::now();
        secs = std::chrono::duration<double>(t1 - t0).count();
        if (total_iters) *total_iters = it;
        from_v3(P, pos);
        from_v3(V, vel);
    });
    return rc == 0 ? secs : -1.0;
}

}  // extern "C"
