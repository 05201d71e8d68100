// SPDX-License-Identifier: Apache-2.0
#include "topology.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace chysx_amd {

namespace {

// Counting-sort (row, col) pairs by row, then sort and unique each row.
// Produces exactly the lexicographic sort + unique of the pair list.
void bucket_rows(int n, const std::vector<int>& rows, const std::vector<int>& cols, std::vector<int>& off,
                 std::vector<int>& out_cols) {
    std::vector<int> cnt(size_t(n) + 1, 0);
    for (int r : rows) cnt[size_t(r) + 1]++;
    for (int i = 0; i < n; ++i) cnt[size_t(i) + 1] += cnt[size_t(i)];
    std::vector<int> tmp(rows.size());
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < rows.size(); ++k) tmp[size_t(pos[size_t(rows[k])]++)] = cols[k];
    off.assign(size_t(n) + 1, 0);
    out_cols.clear();
    out_cols.reserve(tmp.size());
    for (int r = 0; r < n; ++r) {
        auto b = tmp.begin() + cnt[size_t(r)];
        auto e = tmp.begin() + cnt[size_t(r) + 1];
        std::sort(b, e);
        auto u = std::unique(b, e);
        out_cols.insert(out_cols.end(), b, u);
        off[size_t(r) + 1] = int(out_cols.size());
    }
}

}  // namespace

std::vector<int> build_edges(const int* tris, int ntri) {
    if (ntri < 0) throw std::invalid_argument("build_edges: negative triangle count");
    int maxv = -1;
    for (int k = 0; k < 3 * ntri; ++k) {
        if (tris[k] < 0) throw std::out_of_range("build_edges: negative vertex index");
        maxv = std::max(maxv, tris[k]);
    }
    const int n = maxv + 1;
    std::vector<int> rows, cols;
    rows.reserve(size_t(3) * ntri);
    cols.reserve(size_t(3) * ntri);
    for (int t = 0; t < ntri; ++t) {
        const int v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int s = 0; s < 3; ++s) {
            const int a = v[s], b = v[(s + 1) % 3];
            rows.push_back(std::min(a, b));
            cols.push_back(std::max(a, b));
        }
    }
    std::vector<int> off, c;
    bucket_rows(std::max(n, 0), rows, cols, off, c);
    std::vector<int> edges;
    edges.reserve(2 * c.size());
    for (int r = 0; r < n; ++r)
        for (int k = off[size_t(r)]; k < off[size_t(r) + 1]; ++k) {
            edges.push_back(r);
            edges.push_back(c[size_t(k)]);
        }
    return edges;
}

std::vector<int> find_dihedrals(const int* tris, int ntri) {
    // Directed edges (from, to, third, triangle order), bucketed by `from`.
    int nv = 0;
    for (int q = 0; q < 3 * ntri; ++q) {
        if (tris[q] < 0) throw std::out_of_range("find_dihedrals: negative vertex index");
        nv = std::max(nv, tris[q] + 1);
    }
    std::vector<int> off(size_t(nv) + 1, 0);
    for (int q = 0; q < 3 * ntri; ++q) off[size_t(tris[q]) + 1]++;
    for (int i = 0; i < nv; ++i) off[size_t(i) + 1] += off[size_t(i)];
    struct DE {
        int to, third, order;
    };
    std::vector<DE> de(size_t(3) * ntri);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int t = 0; t < ntri; ++t) {
            const int v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
            for (int e = 0; e < 3; ++e) {
                const int from = v[e], to = v[(e + 1) % 3], third = v[(e + 2) % 3];
                de[size_t(fill[size_t(from)]++)] = DE{to, third, 3 * t + e};
            }
        }
    }
    // Per bucket: sort by (to, order) and keep the last record of each `to`
    // (the latest write wins).
    std::vector<int> len(size_t(nv), 0);
    for (int i = 0; i < nv; ++i) {
        auto b = de.begin() + off[size_t(i)], e = de.begin() + off[size_t(i) + 1];
        std::sort(b, e, [](const DE& x, const DE& y) { return x.to != y.to ? x.to < y.to : x.order < y.order; });
        int w = 0;
        for (auto it = b; it != e; ++it) {
            if (it + 1 != e && (it + 1)->to == it->to) continue;
            *(b + w++) = *it;
        }
        len[size_t(i)] = w;
    }
    auto lookup = [&](int from, int to) -> const DE* {
        auto b = de.begin() + off[size_t(from)], e = b + len[size_t(from)];
        auto it = std::lower_bound(b, e, to, [](const DE& x, int v) { return x.to < v; });
        return (it != e && it->to == to) ? &*it : nullptr;
    };
    std::vector<int> out;
    for (int v0 = 0; v0 < nv; ++v0) {
        auto b = de.begin() + off[size_t(v0)];
        for (int q = 0; q < len[size_t(v0)]; ++q) {
            const DE& d = *(b + q);
            if (v0 >= d.to) continue;
            const DE* r = lookup(d.to, v0);
            if (!r) continue;
            out.insert(out.end(), {v0, d.to, d.third, r->third});
        }
    }
    return out;
}

Pattern build_pattern(int n, const int* tris, int ntri, const int* springs, int nspring, const int* dih, int ndih) {
    if (n < 0 || ntri < 0 || nspring < 0 || ndih < 0) throw std::invalid_argument("build_pattern: negative size");
    std::vector<int> rows, cols;
    rows.reserve(size_t(6) * ntri + size_t(2) * nspring);
    cols.reserve(rows.capacity());
    auto push = [&](int r, int c) {
        if (r < 0 || r >= n || c < 0 || c >= n) throw std::out_of_range("build_pattern: pair out of range");
        if (r == c) return;
        rows.push_back(r);
        cols.push_back(c);
    };
    for (int t = 0; t < ntri; ++t) {
        const int v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (a != b && v[a] != v[b]) push(v[a], v[b]);
                else if (a != b) push(v[a], v[b]);  // v[a]==v[b]: range-checked, then dropped
    }
    for (int d = 0; dih && d < ndih; ++d) {
        const int* v = dih + 4 * d;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                if (a != b) push(v[a], v[b]);  // v[a] == v[b]: range-checked, then dropped
    }
    for (int e = 0; springs && e < nspring; ++e) {
        push(springs[2 * e], springs[2 * e + 1]);
        push(springs[2 * e + 1], springs[2 * e]);
    }
    Pattern p;
    p.n = n;
    bucket_rows(n, rows, cols, p.row_offsets, p.cols);
    return p;
}

int resolve_slot(const Pattern& p, int r, int c) {
    if (r < 0 || r >= p.n || c < 0 || c >= p.n) throw std::out_of_range("resolve_slots: pair out of range");
    if (r == c) return -r - 1;
    const auto first = p.cols.begin() + p.row_offsets[size_t(r)];
    const auto last = p.cols.begin() + p.row_offsets[size_t(r) + 1];
    const auto it = std::lower_bound(first, last, c);
    if (it == last || *it != c) throw std::out_of_range("resolve_slots: off-diagonal (row, col) not in topology");
    return int(it - p.cols.begin());
}

void build_slot_maps(const Pattern& p, const int* tris, int ntri, const int* springs, int nspring,
                     const int* pins, int npin, std::vector<int>& tri_slots, std::vector<int>& spr_slots,
                     std::vector<int>& pin_slots) {
    tri_slots.resize(size_t(9) * ntri);
    for (int t = 0; t < ntri; ++t)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                tri_slots[size_t(9) * t + 3 * a + b] = resolve_slot(p, tris[3 * t + a], tris[3 * t + b]);
    spr_slots.resize(springs ? size_t(4) * nspring : 0);
    for (int e = 0; springs && e < nspring; ++e) {
        const int a = springs[2 * e], b = springs[2 * e + 1];
        spr_slots[size_t(4) * e + 0] = resolve_slot(p, a, a);
        spr_slots[size_t(4) * e + 1] = resolve_slot(p, a, b);
        spr_slots[size_t(4) * e + 2] = resolve_slot(p, b, a);
        spr_slots[size_t(4) * e + 3] = resolve_slot(p, b, b);
    }
    pin_slots.resize(size_t(npin));
    for (int c = 0; c < npin; ++c) pin_slots[size_t(c)] = resolve_slot(p, pins[c], pins[c]);
}

std::vector<int> build_dihedral_slots(const Pattern& p, const int* dih, int ndih) {
    std::vector<int> out(size_t(16) * ndih);
    for (int d = 0; d < ndih; ++d)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) out[size_t(16) * d + 4 * a + b] = resolve_slot(p, dih[4 * d + a], dih[4 * d + b]);
    return out;
}

DeviceLayout build_device_layout(const Pattern& p, const int* tris, int ntri, const std::vector<int>& tri_slots,
                                 const int* springs, int nspring, const std::vector<int>& spr_slots,
                                 const int* pins, int npin, const int* dih, int ndih,
                                 const std::vector<int>* dih_slots, int slice_group) {
    DeviceLayout L;
    const int n = p.n;
    L.n = n;
    const int unit = kSlice * std::max(slice_group, 1);
    L.n_pad = ((n + unit - 1) / unit) * unit;
    L.nslices = L.n_pad / kSlice;
    L.row_deg.assign(size_t(L.n_pad), 0);
    for (int r = 0; r < n; ++r) {
        const int d = p.row_offsets[size_t(r) + 1] - p.row_offsets[size_t(r)];
        if (d >= 1024) throw std::invalid_argument("device layout: row degree >= 1024 is not supported");
        L.row_deg[size_t(r)] = d;
    }
    L.slice_ptr.assign(size_t(L.nslices) + 1, 0);
    for (int s = 0; s < L.nslices; ++s) {
        int w = 0;
        for (int l = 0; l < kSlice; ++l) w = std::max(w, L.row_deg[size_t(s) * kSlice + l]);
        L.slice_ptr[size_t(s) + 1] = L.slice_ptr[size_t(s)] + w;
        L.max_width = std::max(L.max_width, w);
    }
    L.sell_col.assign(size_t(L.slice_ptr[size_t(L.nslices)]) * kSlice, -1);
    for (int r = 0; r < n; ++r) {
        const int s = r / kSlice, lane = r % kSlice;
        const int b = p.row_offsets[size_t(r)];
        for (int k = 0; k < L.row_deg[size_t(r)]; ++k)
            L.sell_col[(size_t(L.slice_ptr[size_t(s)]) + k) * kSlice + lane] = p.cols[size_t(b + k)];
    }
    // Local position of a slot inside its row (k = csr index - row start).
    auto local = [&](int slot, int row) -> int {
        if (slot < 0) return 1023;  // diagonal: not an off-diagonal entry
        return slot - p.row_offsets[size_t(row)];
    };
    // Triangle incidence, per row in triangle order.
    std::vector<int> cnt(size_t(L.n_pad) + 1, 0);
    for (int k = 0; k < 3 * ntri; ++k) cnt[size_t(tris[k]) + 1]++;
    for (int i = 0; i < L.n_pad; ++i) cnt[size_t(i) + 1] += cnt[size_t(i)];
    L.tri_inc_off = cnt;
    L.tri_inc.assign(size_t(2) * cnt[size_t(L.n_pad)], 0);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int t = 0; t < ntri; ++t)
            for (int a = 0; a < 3; ++a) {
                const int row = tris[3 * t + a];
                int packed = 0;
                for (int b = 0; b < 3; ++b) {
                    const int k = (b == a) ? 1023 : local(tri_slots[size_t(9) * t + 3 * a + b], row);
                    packed |= (k & 1023) << (10 * b);
                }
                const size_t e = size_t(fill[size_t(row)]++);
                L.tri_inc[2 * e] = t * 4 + a;
                L.tri_inc[2 * e + 1] = packed;
            }
    }
    // Spring incidence.
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int k = 0; springs && k < 2 * nspring; ++k) cnt[size_t(springs[k]) + 1]++;
    for (int i = 0; i < L.n_pad; ++i) cnt[size_t(i) + 1] += cnt[size_t(i)];
    L.spr_inc_off = cnt;
    L.spr_inc.assign(size_t(2) * cnt[size_t(L.n_pad)], 0);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int e = 0; springs && e < nspring; ++e)
            for (int side = 0; side < 2; ++side) {
                const int row = springs[2 * e + side];
                // ab slot (index 1) belongs to row a, ba slot (index 2) to row b.
                const int slot = spr_slots[size_t(4) * e + (side == 0 ? 1 : 2)];
                const size_t q = size_t(fill[size_t(row)]++);
                L.spr_inc[2 * q] = e * 2 + side;
                L.spr_inc[2 * q + 1] = local(slot, row);
            }
    }
    // Pin incidence.
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int c = 0; c < npin; ++c) {
        if (pins[c] < 0 || pins[c] >= n) throw std::out_of_range("pins: index out of range");
        cnt[size_t(pins[c]) + 1]++;
    }
    for (int i = 0; i < L.n_pad; ++i) cnt[size_t(i) + 1] += cnt[size_t(i)];
    L.pin_inc_off = cnt;
    L.pin_inc.assign(size_t(cnt[size_t(L.n_pad)]), 0);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int c = 0; c < npin; ++c) L.pin_inc[size_t(fill[size_t(pins[c])]++)] = c;
    }
    // Dihedral incidence, per row in dihedral order.
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int q = 0; dih && q < 4 * ndih; ++q) cnt[size_t(dih[q]) + 1]++;
    for (int i = 0; i < L.n_pad; ++i) cnt[size_t(i) + 1] += cnt[size_t(i)];
    L.bend_inc_off = cnt;
    L.bend_inc.assign(size_t(4) * cnt[size_t(L.n_pad)], 0);
    if (dih && ndih > 0) {
        if (!dih_slots || dih_slots->size() != size_t(16) * ndih)
            throw std::invalid_argument("device layout: dihedral slots missing");
        auto local16 = [&](int slot, int row) -> int {
            if (slot < 0) return 0xFFFF;
            const int k = slot - p.row_offsets[size_t(row)];
            if (k >= 0xFFFF) throw std::invalid_argument("device layout: row degree >= 65535");
            return k;
        };
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int d = 0; d < ndih; ++d)
            for (int a = 0; a < 4; ++a) {
                const int row = dih[4 * d + a];
                int k[4];
                for (int b = 0; b < 4; ++b) k[b] = (b == a) ? 0xFFFF : local16((*dih_slots)[size_t(16) * d + 4 * a + b], row);
                const size_t e = size_t(fill[size_t(row)]++);
                L.bend_inc[4 * e] = d * 4 + a;
                L.bend_inc[4 * e + 1] = int(unsigned(k[0]) | (unsigned(k[1]) << 16));
                L.bend_inc[4 * e + 2] = int(unsigned(k[2]) | (unsigned(k[3]) << 16));
                L.bend_inc[4 * e + 3] = 0;
            }
    }
    return L;
}

void build_vertex_triangles(int n, const int* tris, int ntri, std::vector<int>& off, std::vector<int>& ids) {
    off.assign(size_t(n) + 1, 0);
    for (int k = 0; k < 3 * ntri; ++k) {
        if (tris[k] < 0 || tris[k] >= n) throw std::out_of_range("triangle index out of range");
        off[size_t(tris[k]) + 1]++;
    }
    for (int i = 0; i < n; ++i) off[size_t(i) + 1] += off[size_t(i)];
    ids.assign(size_t(off[size_t(n)]), 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int t = 0; t < ntri; ++t)
        for (int a = 0; a < 3; ++a) ids[size_t(fill[size_t(tris[3 * t + a])]++)] = t;
}

}  // namespace chysx_amd
