// SPDX-License-Identifier: Apache-2.0
//
// Host-side topology for the implicit cloth step: edge list, block-sparse
// pattern, reference slot maps, and the device layout derived from them
// (wave-sliced ELL storage and row-owned incidence lists).
//
// The pattern (row_offsets, col_indices) and slot maps are built to be
// bit-identical to the reference:
//   edges        src/cloth/cloth_simulator.cu:331-349
//   pattern      src/cloth/cloth_simulator.cu:698-793 + src/sparse/block_csr.cu:91-147
//   slot maps    src/constraint/core/constraint_n.h:115-153 + src/sparse/block_csr.cu:168-204
// The reference sorts all (row, col) pairs with std::sort; here pairs are
// bucketed by row (counting sort) and each short row is sorted, which yields
// the same lexicographic order in O(pairs).
#pragma once

#include <cstdint>
#include <vector>

namespace chysx_amd {

constexpr int kSlice = 64;  // rows per slice = lanes per wavefront

// Unique undirected edges (min, max) in lexicographic order.
std::vector<int> build_edges(const int* tris, int ntri);

struct Pattern {
    int n = 0;
    std::vector<int> row_offsets;  // n + 1
    std::vector<int> cols;         // nnz, ascending per row, no diagonal
};

// Bending dihedrals (v0, v1, v2, v3) of a triangle list, 4 ints each
// (build_bending_from_current_positions, cloth_simulator.cu:451-514): every
// directed triangle edge records its triangle's third vertex (a later
// triangle overwrites an earlier one); an edge v0 < v1 present in both
// directions gives (v0, v1, third(v0 -> v1), third(v1 -> v0)).  The
// reference emits them in hash-map order; here they are sorted by (v0, v1).
std::vector<int> find_dihedrals(const int* tris, int ntri);

// Pattern from the ordered off-diagonal pairs of every triangle
// (a != b, v_a != v_b), of every dihedral (4 vertices) when `dih` is
// non-null, plus both directions of every spring edge when `springs` is
// non-null.  Throws std::out_of_range on an index outside [0, n).
Pattern build_pattern(int n, const int* tris, int ntri, const int* springs, int nspring, const int* dih = nullptr,
                      int ndih = 0);

// Reference slot encoding: -r-1 for a diagonal block, else the CSR index of
// (r, c).  Throws std::out_of_range if (r, c) is not in the pattern.
int resolve_slot(const Pattern& p, int r, int c);

// Slot maps in reference order: triangles 9t+3a+b, springs 4e+{aa,ab,ba,bb},
// pins one per pin.
void build_slot_maps(const Pattern& p, const int* tris, int ntri, const int* springs, int nspring,
                     const int* pins, int npin, std::vector<int>& tri_slots, std::vector<int>& spr_slots,
                     std::vector<int>& pin_slots);
// Dihedral slot map 16d + 4a + b (ConstraintN<4>::bind_hessian_layout).
std::vector<int> build_dihedral_slots(const Pattern& p, const int* dih, int ndih);

// Device layout of one instance, row space padded to a multiple of kSlice.
//
// Wave-sliced ELL: slice s covers rows [64 s, 64 s + 64); its width w_s is
// the longest row in the slice; entry (s, k, lane) is block k (CSR order) of
// row 64 s + lane, stored at entry index (slice_ptr[s] + k) * 64 + lane.
// Missing entries have column -1.
//
// Row-owned incidence lists drive assembly: every row lists the elements
// touching it with the local CSR positions of the element's other vertices,
// so the row's owner thread writes all of that row's blocks without atomics
// and in a fixed order.
struct DeviceLayout {
    int n = 0;              // rows (particles)
    int n_pad = 0;          // rows padded to kSlice * slice_group
    int nslices = 0;
    int max_width = 0;
    std::vector<int> slice_ptr;   // nslices + 1 (in entries of 64 lanes)
    std::vector<int> sell_col;    // slice_ptr[nslices] * 64
    std::vector<int> row_deg;     // n_pad
    // triangle incidence: per row, entries {t*4 + a, packed local positions}
    // packed = k_b0 | k_b1 << 10 | k_b2 << 20 for the three vertices b (own
    // vertex slot unused); 10 bits per position (row degree < 1024).
    std::vector<int> tri_inc_off;  // n_pad + 1
    std::vector<int> tri_inc;      // 2 ints per entry
    // spring incidence: {e*2 + side, local position of the other endpoint}
    std::vector<int> spr_inc_off;
    std::vector<int> spr_inc;
    // pin incidence: pin indices per row
    std::vector<int> pin_inc_off;
    std::vector<int> pin_inc;
    // dihedral incidence: 4 ints per entry {d*4 + a, k_b0 | k_b1 << 16,
    // k_b2 | k_b3 << 16, 0} with 0xFFFF for a diagonal block
    std::vector<int> bend_inc_off;
    std::vector<int> bend_inc;
};

DeviceLayout build_device_layout(const Pattern& p, const int* tris, int ntri, const std::vector<int>& tri_slots,
                                 const int* springs, int nspring, const std::vector<int>& spr_slots,
                                 const int* pins, int npin, const int* dih = nullptr, int ndih = 0,
                                 const std::vector<int>* dih_slots = nullptr, int slice_group = 1);

// Per-vertex triangle incidence (for the mass redistribution helper):
// for each vertex, the triangles touching it, in triangle order.
void build_vertex_triangles(int n, const int* tris, int ntri, std::vector<int>& off, std::vector<int>& ids);

}  // namespace chysx_amd
