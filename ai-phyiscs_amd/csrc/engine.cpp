// SPDX-License-Identifier: Apache-2.0
#include "engine.hpp"
#include "noalloc.hpp"

#include <climits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace chysx_amd {

[[noreturn]] void throw_hip(hipError_t err, const char* what) {
    throw std::runtime_error(std::string("chysx: ") + what + " failed: " + hipGetErrorString(err));
}

void check_hip(hipError_t err, const char* what) {
    if (err != hipSuccess) throw_hip(err, what);
}

// ---------------------------------------------------------------------------
// DeviceArray
// ---------------------------------------------------------------------------
template <class T>
void DeviceArray<T>::resize(size_t n) {
    if (n == n_ && ptr_) return;
    release();
    if (n == 0) return;
    check_hip(hipMalloc(reinterpret_cast<void**>(&ptr_), n * sizeof(T)), "hipMalloc");
    n_ = n;
}
template <class T>
void DeviceArray<T>::assign(const T* host, size_t n, hipStream_t s) {
    resize(n);
    if (n == 0) return;
    check_hip(hipMemcpyAsync(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync(H2D)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}
template <class T>
void DeviceArray<T>::zero(hipStream_t s) {
    if (n_ == 0) return;
    check_hip(hipMemsetAsync(ptr_, 0, n_ * sizeof(T), s), "hipMemsetAsync");
}
template <class T>
void DeviceArray<T>::release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
}
template class DeviceArray<int>;
template class DeviceArray<int16_t>;
template class DeviceArray<int2>;
template class DeviceArray<float>;
template class DeviceArray<float4>;
template class DeviceArray<int4>;
template class DeviceArray<double>;
template class DeviceArray<SliceDesc>;
template class DeviceArray<InstState>;
template class DeviceArray<InstParams>;
template class DeviceArray<WsInst>;

// ---------------------------------------------------------------------------
// Rest data
// ---------------------------------------------------------------------------
void compute_rest_springs(Instance& inst, const float* pos, const float* k_per_edge, float k, hipStream_t s) {
    const int E = int(inst.edges.size() / 2);
    inst.spr_rest.assign(size_t(E), 0.0f);
    inst.spr_k.assign(size_t(E), k);
    if (k_per_edge) std::copy(k_per_edge, k_per_edge + E, inst.spr_k.begin());
    inst.has_springs = true;
    if (E == 0) return;
    std::vector<int> v(inst.edges);
    for (int& i : v) i += inst.part0;
    DeviceArray<int> dv;
    DeviceArray<float> rest;
    dv.assign(v.data(), v.size(), s);
    rest.resize(size_t(E));
    launch_rest_lengths(dv.data(), pos, rest.data(), E, s);
    check_hip(hipMemcpyAsync(inst.spr_rest.data(), rest.data(), sizeof(float) * E, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(rest lengths)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(rest lengths)");
}

void compute_rest_stretch(Instance& inst, const float* pos, const float* k_per_tri, float k, hipStream_t s) {
    const int T = int(inst.tris.size() / 3);
    inst.tri_dminv.assign(size_t(4) * T, 0.0f);
    inst.tri_area.assign(size_t(T), 0.0f);
    inst.tri_k.assign(size_t(T), k);
    if (k_per_tri) std::copy(k_per_tri, k_per_tri + T, inst.tri_k.begin());
    inst.has_stretch = true;
    if (T == 0) return;
    std::vector<int> v(inst.tris);
    for (int& i : v) i += inst.part0;
    DeviceArray<int> dv;
    DeviceArray<float4> dm;
    DeviceArray<float> ar;
    dv.assign(v.data(), v.size(), s);
    dm.resize(size_t(T));
    ar.resize(size_t(T));
    launch_rest_shapes(dv.data(), pos, dm.data(), ar.data(), T, 1.0f, 0.0f, s);
    check_hip(hipMemcpyAsync(inst.tri_dminv.data(), dm.data(), sizeof(float4) * T, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(rest shapes)");
    check_hip(hipMemcpyAsync(inst.tri_area.data(), ar.data(), sizeof(float) * T, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(rest areas)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(rest shapes)");
}

// FEM shear (build_fem_shear_from_current_positions, cloth_simulator.cu:423-449):
// the stretch rest frame with the material axes rotated by pi/4; cos and sin
// are the FP32 values of the FP32 angle (triangle_stretch_constraint.cu:340-341).
void compute_rest_shear(Instance& inst, const float* pos, const float* k_per_tri, float k, hipStream_t s) {
    const int T = int(inst.tris.size() / 3);
    inst.sh_dminv.assign(size_t(4) * T, 0.0f);
    inst.sh_area.assign(size_t(T), 0.0f);
    inst.sh_k.assign(size_t(T), k);
    if (k_per_tri) std::copy(k_per_tri, k_per_tri + T, inst.sh_k.begin());
    inst.has_shear = true;
    if (T == 0) return;
    constexpr float kQuarterPi = 0.7853981633974483f;
    std::vector<int> v(inst.tris);
    for (int& i : v) i += inst.part0;
    DeviceArray<int> dv;
    DeviceArray<float4> dm;
    DeviceArray<float> ar;
    dv.assign(v.data(), v.size(), s);
    dm.resize(size_t(T));
    ar.resize(size_t(T));
    launch_rest_shapes(dv.data(), pos, dm.data(), ar.data(), T, std::cos(kQuarterPi), std::sin(kQuarterPi), s);
    check_hip(hipMemcpyAsync(inst.sh_dminv.data(), dm.data(), sizeof(float4) * T, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(shear rest shapes)");
    check_hip(hipMemcpyAsync(inst.sh_area.data(), ar.data(), sizeof(float) * T, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(shear rest areas)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(shear rest shapes)");
}

// Bending (build_bending_from_current_positions, cloth_simulator.cu:451-514):
// dihedrals from the mesh, rest angles from the current positions.
void compute_rest_bending(Instance& inst, const float* pos, float k, hipStream_t s) {
    const int T = int(inst.tris.size() / 3);
    inst.dih = find_dihedrals(T ? inst.tris.data() : nullptr, T);
    const int D = int(inst.dih.size() / 4);
    inst.dih_rest.assign(size_t(D), 0.0f);
    inst.bend_k = k;
    inst.has_bending = true;
    if (D == 0) return;
    std::vector<int> v(inst.dih);
    for (int& i : v) i += inst.part0;
    DeviceArray<int4> dv;
    DeviceArray<float> rest;
    dv.assign(reinterpret_cast<const int4*>(v.data()), size_t(D), s);
    rest.resize(size_t(D));
    launch_rest_angles(dv.data(), pos, rest.data(), D, s);
    check_hip(hipMemcpyAsync(inst.dih_rest.data(), rest.data(), sizeof(float) * D, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(rest angles)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(rest angles)");
}

// ---------------------------------------------------------------------------
// Problem
// ---------------------------------------------------------------------------
Problem::Problem() = default;

Problem::~Problem() {
    // an enqueued step still uses the buffers and copies into host_states_
    if (stats_pending_ && done_event_) (void)hipEventSynchronize(done_event_);
    destroy_graph();
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
    if (host_states_) (void)hipHostFree(host_states_);
    if (done_event_) (void)hipEventDestroy(done_event_);
}

void Problem::destroy_graph() {
    for (hipGraphExec_t& g : chunk_exec_) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (step_exec_) (void)hipGraphExecDestroy(step_exec_);
    step_exec_ = nullptr;
    step_key_.clear();
}

void Problem::build(const std::vector<const Instance*>& insts) {
    if (stats_pending_) {  // an enqueued step still uses the buffers rebuilt below
        check_hip(hipEventSynchronize(done_event_), "hipEventSynchronize(previous step)");
        stats_pending_ = false;
    }
    destroy_graph();
    built_ = false;
    ninst_ = int(insts.size());
    pats_.assign(size_t(ninst_), Pattern());
    tri_slots_.assign(size_t(ninst_), {});
    dih_slots_.assign(size_t(ninst_), {});
    spr_slots_.assign(size_t(ninst_), {});
    pin_slots_.assign(size_t(ninst_), {});
    inst_row0_.assign(size_t(ninst_), 0);
    inst_n_.assign(size_t(ninst_), 0);
    inst_slice0_.assign(size_t(ninst_), 0);
    inst_nslices_.assign(size_t(ninst_), 0);
    pin_offset_.assign(size_t(ninst_) + 1, 0);

    std::vector<DeviceLayout> lay(static_cast<size_t>(ninst_));
    int row0 = 0, slice0 = 0;
    n_rows_ = 0;
    nnz_total_ = 0;
    size_t ntri_tot = 0, nspr_tot = 0;
    for (int k = 0; k < ninst_; ++k) {
        const Instance& I = *insts[size_t(k)];
        const int T = (I.has_stretch || I.has_shear) ? int(I.tris.size() / 3) : 0;
        const int E = (I.has_springs && I.springs_active) ? int(I.edges.size() / 2) : 0;
        const int P = int(I.pin_idx.size());
        const int D = I.has_bending ? int(I.dih.size() / 4) : 0;
        const int* tris = T ? I.tris.data() : nullptr;
        const int* spr = E ? I.edges.data() : nullptr;
        const int* dih = D ? I.dih.data() : nullptr;
        pats_[size_t(k)] = build_pattern(I.n, tris, T, spr, E, dih, D);
        build_slot_maps(pats_[size_t(k)], tris, T, spr, E, P ? I.pin_idx.data() : nullptr, P,
                        tri_slots_[size_t(k)], spr_slots_[size_t(k)], pin_slots_[size_t(k)]);
        dih_slots_[size_t(k)] = build_dihedral_slots(pats_[size_t(k)], dih, D);
        lay[size_t(k)] = build_device_layout(pats_[size_t(k)], tris, T, tri_slots_[size_t(k)], spr, E,
                                             spr_slots_[size_t(k)], P ? I.pin_idx.data() : nullptr, P, dih, D,
                                             &dih_slots_[size_t(k)], kGroupSlices);
        inst_row0_[size_t(k)] = row0;
        inst_n_[size_t(k)] = I.n;
        inst_slice0_[size_t(k)] = slice0;
        inst_nslices_[size_t(k)] = lay[size_t(k)].nslices;
        row0 += lay[size_t(k)].n_pad;
        slice0 += lay[size_t(k)].nslices;
        n_rows_ += I.n;
        nnz_total_ += (long long)pats_[size_t(k)].cols.size();
        ntri_tot += size_t(T);
        nspr_tot += size_t(E);
        pin_offset_[size_t(k) + 1] = pin_offset_[size_t(k)] + P;
    }
    const int old_pad = n_pad_;
    n_pad_ = row0;
    const int nslices = slice0;

    // ---- concatenate layouts ---------------------------------------------
    slice_ptr_h_.assign(size_t(nslices) + 1, 0);
    row_deg_h_.assign(size_t(n_pad_), 0);
    std::vector<int> tri_v, spr_v, pin_v, tri_inc_off(size_t(n_pad_) + 1, 0), spr_inc_off(size_t(n_pad_) + 1, 0),
        pin_inc_off(size_t(n_pad_) + 1, 0), pin_inc, bend_inc_off(size_t(n_pad_) + 1, 0);
    std::vector<int2> tri_inc, spr_inc;
    std::vector<int4> dih_v, bend_inc;
    std::vector<float> tri_dminv, tri_area, tri_k, sh_dminv, sh_area, sh_k, spr_rest, spr_k, pin_tgt, dih_rest;
    tri_v.reserve(3 * ntri_tot);
    spr_v.reserve(2 * nspr_tot);
    int entries = 0;
    int tri_base = 0, spr_base = 0, dih_base = 0;
    max_width_ = 0;
    for (int k = 0; k < ninst_; ++k) {
        const Instance& I = *insts[size_t(k)];
        const DeviceLayout& L = lay[size_t(k)];
        const int r0 = inst_row0_[size_t(k)], s0 = inst_slice0_[size_t(k)];
        for (int s = 0; s < L.nslices; ++s) slice_ptr_h_[size_t(s0 + s) + 1] = entries + L.slice_ptr[size_t(s) + 1];
        entries += L.slice_ptr[size_t(L.nslices)];
        for (int r = 0; r < L.n_pad; ++r) row_deg_h_[size_t(r0 + r)] = L.row_deg[size_t(r)];
        max_width_ = std::max(max_width_, L.max_width);
        const int T = (I.has_stretch || I.has_shear) ? int(I.tris.size() / 3) : 0;
        const int E = (I.has_springs && I.springs_active) ? int(I.edges.size() / 2) : 0;
        const int P = int(I.pin_idx.size());
        const int D = I.has_bending ? int(I.dih.size() / 4) : 0;
        for (int q = 0; q < 3 * T; ++q) tri_v.push_back(I.tris[size_t(q)] + I.part0);
        // absent sets contribute zero-stiffness elements (the kernel skips them per instance)
        auto append_set = [&](bool has, const std::vector<float>& dm, const std::vector<float>& ar,
                              const std::vector<float>& kk, std::vector<float>& odm, std::vector<float>& oar,
                              std::vector<float>& ok) {
            if (has && (dm.size() < size_t(4) * T || ar.size() < size_t(T) || kk.size() < size_t(T)))
                throw std::logic_error("triangle rest data does not match the mesh: rebuild after set_mesh");
            for (int t = 0; t < T; ++t) {
                for (int c = 0; c < 4; ++c) odm.push_back(has ? dm[size_t(4) * t + c] : (c == 0 || c == 3 ? 1.0f : 0.0f));
                oar.push_back(has ? ar[size_t(t)] : 0.0f);
                ok.push_back(has ? kk[size_t(t)] : 0.0f);
            }
        };
        append_set(I.has_stretch, I.tri_dminv, I.tri_area, I.tri_k, tri_dminv, tri_area, tri_k);
        append_set(I.has_shear, I.sh_dminv, I.sh_area, I.sh_k, sh_dminv, sh_area, sh_k);
        for (int d = 0; d < D; ++d) {
            dih_v.push_back(make_int4(I.dih[size_t(4) * d] + I.part0, I.dih[size_t(4) * d + 1] + I.part0,
                                      I.dih[size_t(4) * d + 2] + I.part0, I.dih[size_t(4) * d + 3] + I.part0));
            dih_rest.push_back(I.dih_rest[size_t(d)]);
        }
        for (int q = 0; q < 2 * E; ++q) spr_v.push_back(I.edges[size_t(q)] + I.part0);
        if (E) {
            spr_rest.insert(spr_rest.end(), I.spr_rest.begin(), I.spr_rest.begin() + E);
            spr_k.insert(spr_k.end(), I.spr_k.begin(), I.spr_k.begin() + E);
        }
        for (int c = 0; c < P; ++c) {
            pin_v.push_back(I.pin_idx[size_t(c)] + I.part0);
            for (int d = 0; d < 3; ++d) pin_tgt.push_back(I.pin_tgt[size_t(3 * c + d)]);
        }
        // incidence, rows offset by r0, element ids offset by bases
        const int tb = int(tri_inc.size()), sbs = int(spr_inc.size()), pb = int(pin_inc.size()),
                  bb = int(bend_inc.size());
        for (int r = 0; r < L.n_pad; ++r) {
            tri_inc_off[size_t(r0 + r) + 1] = tb + L.tri_inc_off[size_t(r) + 1];
            spr_inc_off[size_t(r0 + r) + 1] = sbs + L.spr_inc_off[size_t(r) + 1];
            pin_inc_off[size_t(r0 + r) + 1] = pb + L.pin_inc_off[size_t(r) + 1];
            bend_inc_off[size_t(r0 + r) + 1] = bb + L.bend_inc_off[size_t(r) + 1];
        }
        for (size_t q = 0; q < L.bend_inc.size() / 4; ++q) {
            const int id = L.bend_inc[4 * q];
            bend_inc.push_back(make_int4(((id >> 2) + dih_base) * 4 + (id & 3), L.bend_inc[4 * q + 1],
                                         L.bend_inc[4 * q + 2], 0));
        }
        for (size_t q = 0; q < L.tri_inc.size() / 2; ++q) {
            const int id = L.tri_inc[2 * q];
            tri_inc.push_back(make_int2(((id >> 2) + tri_base) * 4 + (id & 3), L.tri_inc[2 * q + 1]));
        }
        for (size_t q = 0; q < L.spr_inc.size() / 2; ++q) {
            const int id = L.spr_inc[2 * q];
            spr_inc.push_back(make_int2(((id >> 1) + spr_base) * 2 + (id & 1), L.spr_inc[2 * q + 1]));
        }
        for (int c : L.pin_inc) pin_inc.push_back(c + pin_offset_[size_t(k)]);
        tri_base += T;
        spr_base += E;
        dih_base += D;
        if (tri_base > (1 << 29) || dih_base > (1 << 29)) throw std::invalid_argument("too many elements");
    }
    // fill leading offsets of rows with no later entry (monotone already by construction)
    std::vector<int> sell_col(size_t(entries) * kSlice, -1);
    for (int k = 0; k < ninst_; ++k) {
        const DeviceLayout& L = lay[size_t(k)];
        const int r0 = inst_row0_[size_t(k)], s0 = inst_slice0_[size_t(k)];
        const size_t base = size_t(slice_ptr_h_[size_t(s0)]) * kSlice;
        for (size_t q = 0; q < L.sell_col.size(); ++q) {
            const int c = L.sell_col[q];
            sell_col[base + q] = (c >= 0) ? c + r0 : -1;
        }
    }
    // Mirror offsets of the lower entries: block (i, j), j < i, is the
    // transpose of block (j, i) stored by row j.
    std::vector<int> mir(sell_col.size(), -1);
    std::vector<int> nlower(size_t(nslices), 0);
    if (size_t(entries) * 9 * kSlice > size_t(INT32_MAX)) throw std::invalid_argument("matrix too large");
    for (int sl = 0; sl < nslices; ++sl) {
        const int w = slice_ptr_h_[size_t(sl) + 1] - slice_ptr_h_[size_t(sl)];
        for (int lane = 0; lane < kSlice; ++lane) {
            const int i = sl * kSlice + lane;
            for (int k = 0; k < w; ++k) {
                const size_t e = (size_t(slice_ptr_h_[size_t(sl)]) + k) * kSlice + lane;
                const int j = sell_col[e];
                if (j < 0 || j >= i) continue;
                nlower[size_t(sl)] = std::max(nlower[size_t(sl)], k + 1);  // one past the last lower position
                const int sj = j / kSlice, lj = j % kSlice;
                const int wj = slice_ptr_h_[size_t(sj) + 1] - slice_ptr_h_[size_t(sj)];
                int kk = -1;
                for (int q = 0; q < wj; ++q)
                    if (sell_col[(size_t(slice_ptr_h_[size_t(sj)]) + q) * kSlice + lj] == i) {
                        kk = q;
                        break;
                    }
                if (kk < 0) throw std::logic_error("pattern is not symmetric");
                mir[e] = int((size_t(slice_ptr_h_[size_t(sj)]) + kk) * kSlice + lj);
            }
        }
    }

    // ---- slices ------------------------------------------------------------
    slices_h_.clear();
    std::vector<InstState> st(static_cast<size_t>(ninst_));
    std::memset(st.data(), 0, sizeof(InstState) * st.size());
    for (int k = 0; k < ninst_; ++k) {
        const Instance& I = *insts[size_t(k)];
        st[size_t(k)].slice0 = inst_slice0_[size_t(k)];
        st[size_t(k)].nslices = (I.n + kSlice - 1) / kSlice;  // slices holding rows
        st[size_t(k)].group0 = inst_slice0_[size_t(k)] / kGroupSlices;
        st[size_t(k)].ngroups = inst_nslices_[size_t(k)] / kGroupSlices;
        for (int q = 0; q < inst_nslices_[size_t(k)]; ++q)
            slices_h_.push_back(SliceDesc{k, I.part0 + kSlice * q, std::max(0, std::min(kSlice, I.n - kSlice * q)),
                                          nlower[size_t(inst_slice0_[size_t(k)] + q)], st[size_t(k)].group0,
                                          st[size_t(k)].ngroups});
        st[size_t(k)].status = kStopped;
    }
    if (int(slices_h_.size()) != nslices) throw std::logic_error("slice count mismatch");

    // ---- upload --------------------------------------------------------------
    hipStream_t s = nullptr;
    slices_.assign(slices_h_.data(), slices_h_.size(), s);
    state_.assign(st.data(), st.size(), s);
    state_b_.assign(st.data(), st.size(), s);
    tri_v_.assign(tri_v.data(), tri_v.size(), s);
    tri_dminv_.assign(reinterpret_cast<const float4*>(tri_dminv.data()), tri_dminv.size() / 4, s);
    tri_area_.assign(tri_area.data(), tri_area.size(), s);
    tri_k_.assign(tri_k.data(), tri_k.size(), s);
    sh_dminv_.assign(reinterpret_cast<const float4*>(sh_dminv.data()), sh_dminv.size() / 4, s);
    sh_area_.assign(sh_area.data(), sh_area.size(), s);
    sh_k_.assign(sh_k.data(), sh_k.size(), s);
    dih_v_.assign(dih_v.data(), dih_v.size(), s);
    dih_rest_.assign(dih_rest.data(), dih_rest.size(), s);
    bend_inc_off_.assign(bend_inc_off.data(), bend_inc_off.size(), s);
    bend_inc_.assign(bend_inc.data(), bend_inc.size(), s);
    spr_v_.assign(spr_v.data(), spr_v.size(), s);
    spr_rest_.assign(spr_rest.data(), spr_rest.size(), s);
    spr_k_.assign(spr_k.data(), spr_k.size(), s);
    pin_v_.assign(pin_v.data(), pin_v.size(), s);
    pin_tgt_.assign(pin_tgt.data(), pin_tgt.size(), s);
    tri_inc_off_.assign(tri_inc_off.data(), tri_inc_off.size(), s);
    tri_inc_.assign(tri_inc.data(), tri_inc.size(), s);
    spr_inc_off_.assign(spr_inc_off.data(), spr_inc_off.size(), s);
    spr_inc_.assign(spr_inc.data(), spr_inc.size(), s);
    pin_inc_off_.assign(pin_inc_off.data(), pin_inc_off.size(), s);
    pin_inc_.assign(pin_inc.data(), pin_inc.size(), s);
    slice_ptr_.assign(slice_ptr_h_.data(), slice_ptr_h_.size(), s);
    col_.assign(sell_col.data(), sell_col.size(), s);
    mir_.assign(mir.data(), mir.size(), s);
    {
        // 16-bit offsets for the SpMV pass when the mesh's bandwidth allows
        std::vector<int16_t> c16(sell_col.size()), m16(mir.size(), 0);
        bool fits = true;
        for (int sl = 0; sl < nslices && fits; ++sl) {
            const int w = slice_ptr_h_[size_t(sl) + 1] - slice_ptr_h_[size_t(sl)];
            for (int k = 0; k < w && fits; ++k)
                for (int lane = 0; lane < kSlice; ++lane) {
                    const size_t e = (size_t(slice_ptr_h_[size_t(sl)]) + k) * kSlice + lane;
                    const int row = sl * kSlice + lane, c = sell_col[e];
                    const long long dc = (c < 0) ? kNoCol16 : (long long)c - row;
                    const long long dm = (mir[e] < 0) ? 0 : (long long)mir[e] - (long long)e;
                    if (c >= 0 && (dc <= kNoCol16 || dc > 32767)) fits = false;
                    if (dm < -32768 || dm > 32767) fits = false;
                    c16[e] = int16_t(dc);
                    m16[e] = int16_t(dm);
                }
        }
        if (fits) {
            col16_.assign(c16.data(), c16.size(), s);
            mir16_.assign(m16.data(), m16.size(), s);
        } else {
            col16_.release();
            mir16_.release();
        }
    }
    mir_h_ = std::move(mir);
    row_deg_.assign(row_deg_h_.data(), row_deg_h_.size(), s);
    val_.resize(size_t(entries) * kSlice * 9);
    val_.zero(s);
    diag_.resize(size_t(nslices) * kSlice * 9);
    diag_.zero(s);
    minv_.resize(size_t(nslices) * kSlice * 6);
    minv_.zero(s);
    const size_t v3 = size_t(n_pad_) * 3;
    for (DeviceArray<float>* a : {&rhs_, &r_, &z_, &ap_, &p0_, &p1_}) {
        a->resize(v3);
        a->zero(s);
    }
    if (old_pad != n_pad_ || x_.size() != v3) {
        x_.resize(v3);
        x_.zero(s);
    }
    partials_.resize(std::max<size_t>(size_t(kPartialSlots) * (slices_h_.size() / kGroupSlices), kPartialSlots));
    gactive_.resize(std::max<size_t>(slices_h_.size() / kGroupSlices, 1));
    gactive_.zero(s);
    partials_.zero(s);
    // warm-start buffers (allocated on first use)
    hist_.release();
    rb_.release();
    d_.release();
    ad_.release();
    xb_.release();
    ws_part_.release();
    ws_part2_.release();
    {
        int maxs = 1;
        for (int k = 0; k < ninst_; ++k) maxs = std::max(maxs, inst_nslices_[size_t(k)]);
        ws_chunks_ = (maxs + kWsChunkSlices - 1) / kWsChunkSlices;
    }
    rest_pos_.resize(size_t(std::max(n_rows_, 1)) * 3);
    ws_inst_.resize(size_t(std::max(ninst_, 1)));
    ws_inst_.zero(s);
    ws_host_.assign(size_t(std::max(ninst_, 1)), WsInst{});
    rest_captured_ = false;
    hist_n_ = 0;
    hist_head_ = 0;
    params_.resize(size_t(std::max(ninst_, 1)));
    params_h_.clear();
    shapes_h_.clear();
    if (host_states_) (void)hipHostFree(host_states_);
    host_states_ = nullptr;
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&host_states_), sizeof(InstState) * std::max(ninst_, 1)),
              "hipHostMalloc(states)");
    std::memcpy(host_states_, st.data(), sizeof(InstState) * st.size());
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(build)");
    stats_.assign(size_t(ninst_), SolveStats());
    last_iters_ = 0;
    built_ = true;
}

void Problem::upload_params(const std::vector<const Instance*>& insts, hipStream_t s) {
    wait_pending(s);
    std::vector<InstParams> p(insts.size());
    std::vector<float4> shapes;
    for (size_t k = 0; k < insts.size(); ++k) {
        const Instance& I = *insts[k];
        InstParams& q = p[k];
        q.contact_h = I.contact_h;
        q.contact_k = I.contact_k;
        q.contact_mu = I.contact_mu;
        q.contact_eps = I.contact_eps;
        q.plane0 = int(shapes.size());
        q.nplane = int(I.planes.size() / 4);
        for (size_t j = 0; j < I.planes.size(); j += 4)
            shapes.push_back(make_float4(I.planes[j], I.planes[j + 1], I.planes[j + 2], I.planes[j + 3]));
        q.box0 = int(shapes.size());
        q.nbox = int(I.boxes.size() / 16);
        for (size_t j = 0; j < I.boxes.size(); j += 4)
            shapes.push_back(make_float4(I.boxes[j], I.boxes[j + 1], I.boxes[j + 2], I.boxes[j + 3]));
        q.tol2 = I.tol > 0.0 ? I.tol * I.tol : 0.0;
        q.max_iter = I.max_iter;
        q.gx = I.material.gx;
        q.gy = I.material.gy;
        q.gz = I.material.gz;
        q.damping = I.material.damping;
        q.pin_k = I.pin_k;
        q.force_active = I.force_active ? 1 : 0;
        q.springs_active = (I.springs_active && I.has_springs) ? 1 : 0;
        q.tri_stretch = I.has_stretch ? 1 : 0;
        q.tri_shear = I.has_shear ? 1 : 0;
        q.bend_k = I.has_bending ? I.bend_k : 0.0f;
        q.pin_proj = I.pin_projection ? 1 : 0;
    }
    if (p.empty()) return;
    // Unchanged since the last upload (bytewise; p is value-initialised, so its
    // padding is zero): the device already holds these values.
    if (p.size() == params_h_.size() && shapes.size() == shapes_h_.size() &&
        std::memcmp(p.data(), params_h_.data(), sizeof(InstParams) * p.size()) == 0 &&
        (shapes.empty() || std::memcmp(shapes.data(), shapes_h_.data(), sizeof(float4) * shapes.size()) == 0))
        return;
    if (!shapes.empty()) shapes_.assign(shapes.data(), shapes.size(), s);
    check_hip(hipMemcpyAsync(params_.data(), p.data(), sizeof(InstParams) * p.size(), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(params)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(params)");
    params_h_ = std::move(p);
    shapes_h_ = std::move(shapes);
}

void Problem::wait_pending(hipStream_t s) const {
    if (stats_pending_) check_hip(hipStreamWaitEvent(s, done_event_, 0), "hipStreamWaitEvent(enqueued step)");
}

void Problem::update_pin_targets(const Instance& inst, int k, hipStream_t s) {
    const int P = int(inst.pin_idx.size());
    if (P == 0) return;
    wait_pending(s);
    check_hip(hipMemcpyAsync(pin_tgt_.data() + 3 * pin_offset_[size_t(k)], inst.pin_tgt.data(), sizeof(float) * 3 * P,
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(pin targets)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(pin targets)");
}

DevMatrix Problem::matrix() const {
    const bool i16 = cfg_.spmv_index16 && col16_.size() > 0;
    return DevMatrix{slice_ptr_.data(),
                     col_.data(),
                     row_deg_.data(),
                     val_.data(),
                     diag_.data(),
                     minv_.data(),
                     mir_.data(),
                     cfg_.spmv_symmetric,
                     block_floats(),
                     i16 ? col16_.data() : nullptr,
                     i16 ? mir16_.data() : nullptr};
}
int Problem::block_floats() const {
    return (cfg_.spmv_symmetric && cfg_.spmv_packed && dih_v_.size() == 0) ? 6 : 9;
}
DevVectors Problem::vectors() const {
    return DevVectors{rhs_.data(), r_.data(), z_.data(), ap_.data(), p0_.data(), p1_.data(), x_.data(),
                      partials_.data()};
}
DevMesh Problem::mesh() const {
    DevMesh m{};
    m.tri_v = tri_v_.data();
    m.tri_dminv = tri_dminv_.data();
    m.tri_area = tri_area_.data();
    m.tri_k = tri_k_.data();
    m.spr_v = spr_v_.data();
    m.spr_rest = spr_rest_.data();
    m.spr_k = spr_k_.data();
    m.pin_v = pin_v_.data();
    m.pin_tgt = pin_tgt_.data();
    m.tri_inc_off = tri_inc_off_.data();
    m.tri_inc = tri_inc_.data();
    m.spr_inc_off = spr_inc_off_.data();
    m.spr_inc = spr_inc_.data();
    m.pin_inc_off = pin_inc_off_.data();
    m.pin_inc = pin_inc_.data();
    m.sh_dminv = sh_dminv_.data();
    m.sh_area = sh_area_.data();
    m.sh_k = sh_k_.data();
    m.dih_v = dih_v_.data();
    m.dih_rest = dih_rest_.data();
    m.bend_inc_off = bend_inc_off_.data();
    m.bend_inc = bend_inc_.data();
    m.shapes = shapes_.data();
    return m;
}

double Problem::spmv_bytes() const { return 40.0 * double(nnz_total_) + 64.0 * double(n_rows_) + 4.0; }

void Problem::launch_assemble(float dt, const DevExt& ext, hipStream_t s) {
    chysx_amd::launch_assemble(geom(), mesh(), matrix(), vectors(), ext, params_.data(), dt, cfg_.assemble_waves,
                               max_width_, s);
}
void Problem::launch_pcg_start(hipStream_t s) {
    const LaunchGeom g = geom();
    if (start_r_ready_)
        launch_pcg_start_given_r(g, matrix(), vectors(), state_.data(), s);
    else
        launch_pcg_residual(g, matrix(), vectors(), state_.data(), 0, s);
    start_r_ready_ = false;
    launch_pcg_reduce(state_.data(), params_.data(), partials_.data(), ninst_, g.ngroups, kPhaseStart, s, g.gactive);
}
// One iteration: the SpMV pass (it first folds the previous iteration's
// partials into the state) and the update pass (it sums p . Ap itself).
void Problem::launch_pcg_iteration(hipStream_t s, int i) {
    const LaunchGeom g = geom();
    InstState* in = (i & 1) ? state_b_.data() : state_.data();
    InstState* out = (i & 1) ? state_.data() : state_b_.data();
    if (split_advance()) {
        // the advance of the previous pair, from `in` into `out`, then the SpMV
        // pass on the current state in `out`
        launch_pcg_advance_copy(in, out, params_.data(), partials_.data(), ninst_, g.ngroups, s, g.gactive);
        launch_pcg_spmv(g, matrix(), vectors(), out, nullptr, nullptr, cfg_.spmv_unroll, s, cfg_.spmv_branch_free,
                        cfg_.spmv_symmetric);
    } else {
        launch_pcg_spmv(g, matrix(), vectors(), in, out, params_.data(), cfg_.spmv_unroll, s, cfg_.spmv_branch_free,
                        cfg_.spmv_symmetric);
    }
    launch_pcg_update(g, matrix(), vectors(), out, s);
}
void Problem::launch_pcg_chunk_body(hipStream_t s) {
    static_assert(kChunkIters % 2 == 0, "a chunk must end in the first state buffer");
    for (int i = 0; i < kChunkIters; ++i) launch_pcg_iteration(s, i);
    launch_pcg_advance(state_.data(), params_.data(), partials_.data(), ninst_, geom().ngroups, s, geom().gactive);
}

void Problem::set_launch_config(const LaunchConfig& c) {
    if (!valid_waves(c.assemble_waves) || !valid_waves(c.residual_waves) || !valid_waves(c.spmv_waves) ||
        !valid_waves(c.update_waves) || !valid_waves(c.finalize_waves) || !valid_unroll(c.spmv_unroll) ||
        (c.spmv_branch_free != 0 && c.spmv_branch_free != 1) || (c.spmv_symmetric != 0 && c.spmv_symmetric != 1) ||
        (c.spmv_packed != 0 && c.spmv_packed != 1) || (c.spmv_index16 != 0 && c.spmv_index16 != 1) ||
        c.pcg_split_advance < -1 || c.pcg_split_advance > 1)
        throw std::invalid_argument("launch config: waves must be 1, 2, 4, 8 or 16 and unroll 1, 2 or 4");
    if (c.residual_waves != 16 || c.spmv_waves != 16 || c.update_waves != 16)
        throw std::invalid_argument(
            "launch config: the PCG passes run 16 waves per workgroup (one inner-product partial per 16-slice "
            "group); residual_waves, spmv_waves and update_waves must be 16");
    const int bs_old = block_floats(), sym_old = cfg_.spmv_symmetric;
    cfg_ = c;
    destroy_graph();
    if (built_ && (block_floats() != bs_old || cfg_.spmv_symmetric != sym_old)) {
        // A new storage reuses the value and diagonal buffers with another
        // layout.  Padding entries must read as zero blocks (the branch-free
        // SpMV multiplies them), and assembly writes only real entries, so
        // clear what the old layout left behind.
        val_.zero(nullptr);
        diag_.zero(nullptr);
        check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize(storage change)");
    }
}

namespace {
__global__ void prime_states_kernel(InstState* st, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st[i].status = kRunning;
        st[i].mode = kModeIter;
        st[i].sigma = 1.0;
        st[i].rho = 1.0;
        st[i].beta = 0.5;
        st[i].pend = 0;
        st[i].defer = 0;
        st[i].owed = 0;
        st[i].alpha_pend = 0.0;
    }
}
}  // namespace

void Problem::prime_states(InstState* st, hipStream_t s) const {
    hipLaunchKernelGGL(prime_states_kernel, dim3((ninst_ + 63) / 64), dim3(64), 0, s, st, ninst_);
    check_hip(hipGetLastError(), "prime_states_kernel");
    if (ninst_ > 1 && gactive_.size() > 0)  // every cloth iterating
        check_hip(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(gactive_.data()), 1, gactive_.size(), s),
                  "hipMemsetD32Async(active groups)");
}

// Autotuner: on this problem's own matrix layout, times every SpMV-pass
// variant (unroll x branch-free padding) on scratch vectors and a scratch
// solver state; the assembly at each width of workgroup (it writes this
// step's system, as the step's own assembly will again); and, when the
// launch config leaves it open, whole PCG iterations with the state advanced
// in the SpMV pass and in a kernel of its own, on a scratch state that keeps
// iterating.  It keeps the fastest of each.  Launch shapes never change
// results.
void Problem::autotune(hipStream_t s, float dt, const DevExt& ext) {
    if (!built_ || n_pad_ < kAutotuneMinRows) return;
    const size_t v3 = size_t(n_pad_) * 3;
    DeviceArray<float> f[6];
    for (auto& a : f) {
        a.resize(v3);
        a.zero(s);
    }
    DeviceArray<double> xs, ps;
    xs.resize(v3);
    xs.zero(s);
    ps.resize(partials_.size());
    DeviceArray<InstState> sst, sst_b;
    sst.resize(size_t(ninst_));
    sst_b.resize(size_t(ninst_));
    check_hip(hipMemcpyAsync(sst.data(), state_.data(), sizeof(InstState) * ninst_, hipMemcpyDeviceToDevice, s),
              "hipMemcpyAsync(scratch state)");
    prime_states(sst.data(), s);
    const DevVectors v{f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(), xs.data(),
                       ps.data()};
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "hipEventCreate");
    check_hip(hipEventCreate(&e1), "hipEventCreate");
    auto time_it = [&](auto&& launch) {
        for (int i = 0; i < 2; ++i) launch();
        check_hip(hipEventRecord(e0, s), "hipEventRecord");
        for (int i = 0; i < kAutotuneReps; ++i) launch();
        check_hip(hipEventRecord(e1, s), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return double(ms) / kAutotuneReps;
    };
    tune_log_.clear();
    // 1. the SpMV pass
    const int unroll_list[] = {1, 2, 4};
    double best = 1e30;
    for (int bf = 0; bf < 2; ++bf)
        for (int u : unroll_list) {
            const double t = time_it(
                [&] { launch_pcg_spmv(geom(), matrix(), v, sst.data(), nullptr, nullptr, u, s, bf, cfg_.spmv_symmetric); });
            tune_log_.push_back({bf ? 2 : 0, kGroupSlices, u, t});
            if (t < best) {
                best = t;
                cfg_.spmv_unroll = u;
                cfg_.spmv_branch_free = bf;
            }
        }
    // 2. the assembly's waves per workgroup (the LDS slab caps the width)
    if (ext.pos != nullptr && ext.vel != nullptr) {
        best = 1e30;
        for (int w : {1, 2, 4, 8}) {
            const double t = time_it([&] {
                chysx_amd::launch_assemble(geom(), mesh(), matrix(), vectors(), ext, params_.data(), dt, w,
                                           max_width_, s);
            });
            tune_log_.push_back({3, w, 0, t});
            if (t < best) {
                best = t;
                cfg_.assemble_waves = w;
            }
        }
    }
    // 3. where the state is advanced, over whole iterations: scratch
    // parameters that never stop the solve and partials of 1, so every
    // advance does its full work
    split_tuned_ = -1;
    if (cfg_.pcg_split_advance < 0) {
        std::vector<InstParams> ph = params_h_;
        for (InstParams& q : ph) {
            q.tol2 = 0.0;
            q.max_iter = 1 << 30;
        }
        DeviceArray<InstParams> sp;
        sp.assign(ph.data(), ph.size(), s);
        const std::vector<double> ones(partials_.size(), 1.0);
        ps.assign(ones.data(), ones.size(), s);
        const LaunchGeom g = geom();
        int parity = 0;
        auto iteration = [&](bool split) {
            InstState* in = parity ? sst_b.data() : sst.data();
            InstState* out = parity ? sst.data() : sst_b.data();
            parity ^= 1;
            if (split) {
                launch_pcg_advance_copy(in, out, sp.data(), ps.data(), ninst_, g.ngroups, s, g.gactive);
                launch_pcg_spmv(g, matrix(), v, out, nullptr, nullptr, cfg_.spmv_unroll, s, cfg_.spmv_branch_free,
                                cfg_.spmv_symmetric);
            } else {
                launch_pcg_spmv(g, matrix(), v, in, out, sp.data(), cfg_.spmv_unroll, s, cfg_.spmv_branch_free,
                                cfg_.spmv_symmetric);
            }
            launch_pcg_update(g, matrix(), v, out, s);
        };
        double t2[2];
        for (int split = 0; split < 2; ++split) {
            prime_states(sst.data(), s);
            parity = 0;
            t2[split] = time_it([&] { iteration(split != 0); });
            tune_log_.push_back({4 + split, kGroupSlices, cfg_.spmv_unroll, t2[split]});
        }
        split_tuned_ = t2[1] < t2[0] ? 1 : 0;
        if (ninst_ > 1 && gactive_.size() > 0)  // the timings left every group flagged; the start resets them
            check_hip(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(gactive_.data()), 1, gactive_.size(), s),
                      "hipMemsetD32Async(active groups)");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(autotune)");
    destroy_graph();
}
void Problem::capture_chunk_graph(int level) {
    if (!cap_stream_) check_hip(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking), "hipStreamCreate");
    hipGraph_t graph = nullptr;
    check_hip(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (int c = 0; c < (1 << level); ++c) launch_pcg_chunk_body(cap_stream_);
    } catch (...) {
        (void)hipStreamEndCapture(cap_stream_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    check_hip(hipStreamEndCapture(cap_stream_, &graph), "hipStreamEndCapture");
    hipError_t e = hipGraphInstantiate(&chunk_exec_[level], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    check_hip(e, "hipGraphInstantiate");
}
void Problem::launch_pcg_chunks(hipStream_t s, int count) {
    for (int level = kChunkGraphLevels - 1; level >= 0 && count > 0; --level) {
        while (count >= (1 << level)) {
            if (!chunk_exec_[level]) capture_chunk_graph(level);
            check_hip(hipGraphLaunch(chunk_exec_[level], s), "hipGraphLaunch(pcg chunks)");
            count -= 1 << level;
        }
    }
}
std::vector<unsigned char> Problem::fixed_step_key(float dt, const DevExt& ext, int chunks) const {
    // every field is a pointer or a 4-byte integer, so the structs carry no
    // padding and compare bytewise
    std::vector<unsigned char> k;
    auto put = [&k](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        k.insert(k.end(), b, b + n);
    };
    const LaunchGeom g = geom();
    const DevMesh m = mesh();
    const DevMatrix A = matrix();
    const DevVectors v = vectors();
    const void* ptrs[5] = {params_.data(), state_.data(), state_b_.data(), partials_.data(), host_states_};
    const int ints[4] = {chunks, max_width_, split_advance() ? 1 : 0, ninst_};
    put(&dt, sizeof dt);
    put(&ext, sizeof ext);
    put(&g, sizeof g);
    put(&m, sizeof m);
    put(&A, sizeof A);
    put(&v, sizeof v);
    put(ptrs, sizeof ptrs);
    put(ints, sizeof ints);
    put(&cfg_, sizeof cfg_);
    return k;
}

// The fixed-count step (every instance at tol = 0) as one graph: assembly,
// the PCG start, `chunks` chunks of iterations, the true residual of the
// stopped solves, the final reduction, finalize and the copy of the solver
// states for the statistics.  Captured once and replayed while nothing it
// was captured with changes; no memset is recorded.
void Problem::launch_fixed_step_graph(float dt, const DevExt& ext, int chunks, hipStream_t s) {
    std::vector<unsigned char> key = fixed_step_key(dt, ext, chunks);
    if (!step_exec_ || key != step_key_) {
        if (step_exec_) (void)hipGraphExecDestroy(step_exec_);
        step_exec_ = nullptr;
        step_key_.clear();
        if (!cap_stream_) check_hip(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking), "hipStreamCreate");
        hipGraph_t graph = nullptr;
        check_hip(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch_assemble(dt, ext, cap_stream_);
            launch_pcg_start(cap_stream_);
            for (int c = 0; c < chunks; ++c) launch_pcg_chunk_body(cap_stream_);
            launch_pcg_residual(geom(), matrix(), vectors(), state_.data(), 1, cap_stream_);
            launch_pcg_reduce(state_.data(), params_.data(), partials_.data(), ninst_, geom().ngroups, kPhaseFinal,
                              cap_stream_);
            launch_finalize(dt, ext, cap_stream_);
            check_hip(hipMemcpyAsync(host_states_, state_.data(), sizeof(InstState) * ninst_, hipMemcpyDeviceToHost,
                                     cap_stream_),
                      "hipMemcpyAsync(states)");
        } catch (...) {
            (void)hipStreamEndCapture(cap_stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check_hip(hipStreamEndCapture(cap_stream_, &graph), "hipStreamEndCapture");
        hipError_t e = hipGraphInstantiate(&step_exec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check_hip(e, "hipGraphInstantiate(step)");
        step_key_ = std::move(key);
    }
    check_hip(hipGraphLaunch(step_exec_, s), "hipGraphLaunch(step)");
}

void Problem::launch_finalize(float dt, const DevExt& ext, hipStream_t s) {
    chysx_amd::launch_finalize(geom(), vectors(), ext, params_.data(), dt, mesh(), cfg_.finalize_waves, s);
}
void Problem::read_states(hipStream_t s) {
    check_hip(hipMemcpyAsync(host_states_, state_.data(), sizeof(InstState) * ninst_, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(states)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(states)");
}
bool Problem::all_done() const {
    for (int k = 0; k < ninst_; ++k) {
        const int st = host_states_[k].status;
        if (st == kRunning) return false;
    }
    return true;
}

void Problem::set_warm_start_mode(int mode) {
    if (mode != kWarmPrevious && mode != kWarmLearned && mode != kWarmProjection && mode != kWarmProjLearned)
        throw std::invalid_argument(
            "warm start mode: 0 (previous step), 1 (learned), 2 (projection) or 3 (projection + learned)");
    ws_mode_ = mode;
}

void Problem::set_ws_history(int k) {
    if (k < 2 || k > kWsHistMax) throw std::invalid_argument("warm start history: 2 <= k <= 8");
    if (k != hist_k_) {
        hist_k_ = k;
        hist_.release();
        hist_n_ = 0;
        hist_head_ = 0;
    }
}

WsHist Problem::hist_ptrs() const {
    WsHist h{};
    h.k = std::min(hist_n_, hist_k_);
    const size_t v3 = size_t(n_pad_) * 3;
    for (int j = 0; j < h.k; ++j) {
        const int slot = ((hist_head_ - j) % hist_k_ + hist_k_) % hist_k_;
        h.v[j] = hist_.data() + size_t(slot) * v3;
    }
    return h;
}

WsReduce Problem::ws_reduce() const {
    return WsReduce{ws_part_.data(), int(slices_h_.size()), ws_part2_.data(), ws_chunks_};
}

void Problem::push_history(hipStream_t s) {
    const size_t v3 = size_t(n_pad_) * 3;
    if (hist_.size() != size_t(hist_k_) * v3) {
        hist_.resize(size_t(hist_k_) * v3);
        hist_.zero(s);
        hist_n_ = 0;
        hist_head_ = 0;
    }
    hist_head_ = (hist_n_ == 0) ? 0 : (hist_head_ + 1) % hist_k_;
    launch_ws_history(x_.data(), hist_.data() + size_t(hist_head_) * v3, v3, s);
    hist_n_ = std::min(hist_n_ + 1, hist_k_);
}

WarmStartNet& Problem::net() {
    if (!net_) net_ = std::make_unique<WarmStartNet>();
    return *net_;
}

WsGraph Problem::ws_graph() const {
    return WsGraph{slices_.data(), int(slices_h_.size()), slice_ptr_.data(), col_.data(), row_deg_.data()};
}

void Problem::warm_start_stage(float dt, const DevExt& ext, hipStream_t s) {
    const size_t v3 = size_t(n_pad_) * 3;
    const size_t ns = slices_h_.size();
    if (ws_part_.size() < size_t(kWsMaxValues) * ns) ws_part_.resize(size_t(kWsMaxValues) * ns);
    if (ws_part2_.size() < size_t(kWsMaxValues) * size_t(ninst_) * size_t(ws_chunks_))
        ws_part2_.resize(size_t(kWsMaxValues) * size_t(ninst_) * size_t(ws_chunks_));
    const WsHist h = hist_ptrs();
    const WsReduce red = ws_reduce();
    const bool learned = ws_mode_ == kWarmLearned || ws_mode_ == kWarmProjLearned;
    if (ws_mode_ == kWarmProjection || ws_mode_ == kWarmProjLearned) {
        launch_ws_gram(geom(), matrix(), rhs_.data(), h, red, s);
        launch_ws_finish(red, ws_gram_values(h.k), state_.data(), ws_inst_.data(), ninst_, kWsPhaseProj, h.k, s);
        launch_ws_proj_apply(geom(), h, ws_inst_.data(), x_.data(), s);
        proj_ran_ = true;
    }
    if (!learned && !ws_train_) return;
    WarmStartNet& N = net();
    if (!N.has_params()) throw std::runtime_error("learned warm start: no network parameters set");
    N.reserve(n_pad_, ws_train_);
    if (!rest_captured_) {
        check_hip(hipMemcpyAsync(rest_pos_.data(), ext.pos, sizeof(float) * 3 * size_t(n_rows_),
                                 hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync(rest positions)");
        rest_captured_ = true;
    }
    for (DeviceArray<float>* a : {&rb_, &d_, &ad_})
        if (a->size() != v3) a->resize(v3);
    if (xb_.size() != v3) xb_.resize(v3);
    const DevVectors v = vectors();
    // inference reads the base guess from x itself (unchanged until the line
    // apply); training keeps the copy for the energy after the solve
    double* xb = ws_train_ ? xb_.data() : x_.data();
    launch_ws_base(geom(), matrix(), v, x_.data(), ws_train_ ? xb : nullptr, rb_.data(), red, s);
    launch_ws_finish(red, 3, state_.data(), ws_inst_.data(), ninst_, kWsPhaseBase, proj_ran_ ? 1 : 0, s);
    // inference: the features go straight into the encoder, the layers read
    // the compact edge features and the last layer writes the direction d
    // itself; training keeps F, the 32 edge-feature columns, the full output
    // and the activations for backward
    if (ws_train_) {
        N.set_compact_edges(false);
        launch_ws_features(geom(), matrix(), v, ext, mesh(), rest_pos_.data(), h, xb_.data(), ws_inst_.data(), dt,
                           N.features(), N.edge_features(), nullptr, s);
        N.forward(ws_graph(), s, true, false, ws_inst_.data(), d_.data());
    } else {
        N.features_encode(geom(), matrix(), v, ext, mesh(), rest_pos_.data(), h, xb, ws_inst_.data(), dt, s);
        N.forward(ws_graph(), s, false, true, ws_inst_.data(), d_.data(), true);
    }
    launch_ws_direction(geom(), matrix(), ws_train_ ? N.out() : nullptr, ws_inst_.data(), rb_.data(), d_.data(),
                        ad_.data(), red, s);
    launch_ws_finish(red, 4, state_.data(), ws_inst_.data(), ninst_, kWsPhaseLine, ws_accept_residual_ ? 1 : 0, s);
    if (learned) {
        // x <- x_b + alpha d, and the solve starts from r_b - alpha A d
        launch_ws_line_apply(geom(), ws_inst_.data(), d_.data(), x_.data(), rb_.data(), ad_.data(),
                             vectors().r, s);
        start_r_ready_ = true;
    }
    ws_ran_ = true;
}

void Problem::step(float dt, const DevExt& ext, hipStream_t s) {
    if (!built_ || ninst_ == 0) return;
    wait_pending(s);
    stats_pending_ = false;
    ws_ran_ = false;
    proj_ran_ = false;
    start_r_ready_ = false;
    const bool ws_on = ws_mode_ != kWarmPrevious || ws_train_;
    // Every instance at a fixed iteration count (tol == 0): the number of
    // rounds is known, so the whole solve is enqueued without a poll.
    bool fixed = !ws_on && !params_h_.empty();
    int fixed_iters = 0;
    for (const InstParams& q : params_h_) {
        fixed = fixed && q.tol2 == 0.0;
        fixed_iters = std::max(fixed_iters, q.max_iter);
    }
    if (fixed && !blocking_) {
        // max_iter iterations plus a flush round and the advance that stops
        // the solve (rounds after the stop exit at once): the whole step,
        // assembly to finalize and the statistics' copy, is one graph; very
        // long fixed solves replay the chunk graphs instead, so a graph never
        // holds more than kStepGraphMaxChunks chunks
        const int chunks = std::max(1, (fixed_iters + 2 + kChunkIters - 1) / kChunkIters);
        if (chunks <= kStepGraphMaxChunks) {
            launch_fixed_step_graph(dt, ext, chunks, s);
        } else {
            launch_assemble(dt, ext, s);
            launch_pcg_start(s);
            launch_pcg_chunks(s, chunks);
            launch_pcg_residual(geom(), matrix(), vectors(), state_.data(), 1, s);
            launch_pcg_reduce(state_.data(), params_.data(), partials_.data(), ninst_, geom().ngroups, kPhaseFinal,
                              s);
            launch_finalize(dt, ext, s);
            check_hip(
                hipMemcpyAsync(host_states_, state_.data(), sizeof(InstState) * ninst_, hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync(states)");
        }
        if (!done_event_) check_hip(hipEventCreateWithFlags(&done_event_, hipEventDisableTiming), "hipEventCreate");
        check_hip(hipEventRecord(done_event_, s), "hipEventRecord(step)");
        stats_pending_ = true;
        last_iters_ = fixed_iters;
        return;
    }
    launch_assemble(dt, ext, s);
    if (ws_on && hist_n_ >= 2) warm_start_stage(dt, ext, s);
    launch_pcg_start(s);
    // Launch as many rounds as the previous step used (its iterations plus a
    // flush and a check round) before the first poll; rounds after
    // convergence exit at once, and each extra poll costs a host round trip.
    int chunks = std::max(1, (last_iters_ + 2 + kChunkIters - 1) / kChunkIters);
    int guard = 0;
    for (;;) {
        launch_pcg_chunks(s, chunks);
        read_states(s);
        if (all_done()) break;
        chunks = 1;
        if (++guard > (1 << 22)) throw std::runtime_error("PCG did not terminate");
    }
    bool any_stopped = false;
    int max_it = 0;
    for (int k = 0; k < ninst_; ++k) {
        any_stopped |= host_states_[k].status == kStopped;
        max_it = std::max(max_it, host_states_[k].iters);
    }
    if (any_stopped) {
        launch_pcg_residual(geom(), matrix(), vectors(), state_.data(), 1, s);
        launch_pcg_reduce(state_.data(), params_.data(), partials_.data(), ninst_, geom().ngroups, kPhaseFinal, s);
        read_states(s);
    }
    if (ws_ran_ && ws_train_) {
        WarmStartNet& N = net();
        const WsReduce red = ws_reduce();
        launch_ws_energy(geom(), x_.data(), xb_.data(), rb_.data(), red, s);
        launch_ws_finish(red, 1, state_.data(), ws_inst_.data(), ninst_, kWsPhaseEnergy, 0, s);
        launch_ws_dout(geom(), ws_inst_.data(), rb_.data(), ad_.data(), N.dout(), s);
        N.backward(ws_graph(), s);
        N.adam_step(ws_lr_, s);
    }
    if (ws_on) push_history(s);
    launch_finalize(dt, ext, s);
    if (ws_ran_ || proj_ran_) {
        check_hip(hipMemcpyAsync(ws_host_.data(), ws_inst_.data(), sizeof(WsInst) * ninst_, hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(warm start stats)");
        check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(warm start stats)");
    }
    fill_stats();
    if (ws_ran_ && ws_train_) {
        double l = 0.0;
        for (int k = 0; k < ninst_; ++k) l += ws_host_[size_t(k)].loss;
        ws_loss_ = l / double(ninst_);
    }
    last_iters_ = max_it;
}

void Problem::fill_stats() const {
    for (int k = 0; k < ninst_; ++k) {
        SolveStats& q = stats_[size_t(k)];
        q.iters = host_states_[k].iters;
        q.restarts = host_states_[k].restarts;
        q.status = host_states_[k].status;
        q.relres = host_states_[k].relres;
        const WsInst& W = ws_host_[size_t(k)];
        const double wb = W.bb > 0 ? W.bb : 1.0;
        q.ws_proj_used = proj_ran_ ? W.use_proj : 0;
        q.ws_kdim = proj_ran_ ? W.kdim : 0;
        q.ws_used = (ws_ran_ && (ws_mode_ == kWarmLearned || ws_mode_ == kWarmProjLearned)) ? W.use_cand : 0;
        q.ws_r0_prev = (ws_ran_ || proj_ran_) ? std::sqrt(W.rr_prev / wb) : 0.0;
        q.ws_r0_proj = proj_ran_ ? std::sqrt(W.rr_proj / wb) : 0.0;
        q.ws_r0_learned = ws_ran_ ? std::sqrt(W.rr_cand / wb) : 0.0;
    }
}

void Problem::resolve_stats() const {
    if (!stats_pending_) return;
    check_hip(hipEventSynchronize(done_event_), "hipEventSynchronize(step)");
    stats_pending_ = false;
    for (int k = 0; k < ninst_; ++k)
        if (host_states_[k].status == kRunning)
            throw std::runtime_error("PCG: a fixed-iteration solve did not end within its enqueued rounds");
    fill_stats();
}

// ---- reference-layout copies --------------------------------------------
namespace {
template <class T>
std::vector<T> download(const T* dev, size_t n) {
    std::vector<T> h(n);
    if (n) {
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize(debug)");
        check_hip(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy(debug)");
    }
    return h;
}
}  // namespace

void Problem::copy_diag(int k, float* out) const {
    const std::vector<float> d = download(diag_.data(), diag_.size());
    const int r0 = inst_row0_[size_t(k)];
    const int bs = block_floats();  // 6: symmetric diagonal blocks kept as their upper triangle
    static const int kUnpack[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int i = 0; i < inst_n_[size_t(k)]; ++i) {
        const int row = r0 + i, s = row / kSlice, lane = row % kSlice;
        for (int c = 0; c < 9; ++c)
            out[9 * i + c] = d[(size_t(s) * bs + (bs == 6 ? kUnpack[c] : c)) * kSlice + lane];
    }
}
void Problem::copy_values(int k, float* out) const {
    const std::vector<float> v = download(val_.data(), val_.size());
    const Pattern& P = pats_[size_t(k)];
    const int r0 = inst_row0_[size_t(k)];
    const bool sym = cfg_.spmv_symmetric != 0;
    const int bs = block_floats();
    // packed upper triangle (xx, xy, xz, yy, yz, zz) -> row-major 3x3
    static const int kUnpack[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int i = 0; i < P.n; ++i) {
        const int row = r0 + i, s = row / kSlice, lane = row % kSlice;
        const int sb = slice_ptr_h_[size_t(s)];
        for (int q = P.row_offsets[size_t(i)]; q < P.row_offsets[size_t(i) + 1]; ++q) {
            const int kk = q - P.row_offsets[size_t(i)];
            const int e = int((size_t(sb) + kk) * kSlice + lane);
            // the operator's lower block: transpose of the upper mirror
            const bool mirror = sym && P.cols[size_t(q)] < i;
            const size_t base = val_offset(mirror ? mir_h_[size_t(e)] : e, bs);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const int c = mirror ? 3 * b + a : 3 * a + b;
                    out[9 * size_t(q) + 3 * a + b] = v[base + size_t(bs == 6 ? kUnpack[c] : c) * kSlice];
                }
        }
    }
}
void Problem::copy_rhs(int k, float* out) const {
    const std::vector<float> v = download(rhs_.data(), rhs_.size());
    const int r0 = inst_row0_[size_t(k)];
    std::copy(v.begin() + 3 * size_t(r0), v.begin() + 3 * size_t(r0 + inst_n_[size_t(k)]), out);
}
void Problem::copy_dx64(int k, double* out) const {
    const std::vector<double> v = download(x_.data(), x_.size());
    const int r0 = inst_row0_[size_t(k)];
    std::copy(v.begin() + 3 * size_t(r0), v.begin() + 3 * size_t(r0 + inst_n_[size_t(k)]), out);
}
void Problem::copy_dx(int k, float* out) const {
    std::vector<double> d(static_cast<size_t>(3 * inst_n_[size_t(k)]));
    copy_dx64(k, d.data());
    for (size_t i = 0; i < d.size(); ++i) out[i] = float(d[i]);
}
void Problem::set_dx64(int k, const double* in, hipStream_t s) {
    const int r0 = inst_row0_[size_t(k)];
    wait_pending(s);
    check_hip(hipMemcpyAsync(x_.data() + 3 * size_t(r0), in, sizeof(double) * 3 * inst_n_[size_t(k)],
                             hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(dx)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(dx)");
    reset_history();  // the caller's guess is used as given
}
void Problem::zero_dx(hipStream_t s) {
    wait_pending(s);
    x_.zero(s);
    reset_history();
}

// ---------------------------------------------------------------------------
// ClothSim
// ---------------------------------------------------------------------------
ClothSim::ClothSim() = default;

void ClothSim::set_external_buffers(std::uintptr_t pos, std::uintptr_t vel, int n, std::uintptr_t inv_mass,
                                    std::uintptr_t force) noexcept {
    pos_ = pos;
    vel_ = vel;
    inv_mass_ = inv_mass;
    force_ = force;
    n_ = n;
}

void ClothSim::set_mesh(const int* tris, int ntri) {
    if (ntri < 0) throw std::invalid_argument("ClothSimulator::set_mesh: negative triangle count");
    inst_.tris.assign(tris, tris + 3 * size_t(ntri));
    inst_.edges = build_edges(tris, ntri);
    // Element sets built on a previous mesh are dropped (rebuild them).
    inst_.has_springs = inst_.has_stretch = inst_.has_shear = inst_.has_bending = false;
    topo_dirty_ = true;
}

namespace {
// Kernels read positions at these indices: every index must lie in the
// caller's buffer of n particles.
void check_indices(const std::vector<int>& idx, int n, const char* what) {
    for (int v : idx)
        if (v < 0 || v >= n) throw std::out_of_range(std::string(what) + ": vertex index outside [0, particle_count)");
}
}  // namespace

void ClothSim::build_springs(float k, std::uintptr_t stream) {
    const int E = int(inst_.edges.size() / 2);
    if (E > 0 && pos_ == 0)
        throw std::runtime_error(
            "ClothSimulator::build_springs_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    check_indices(inst_.edges, n_, "build_springs_from_current_positions");
    inst_.part0 = 0;
    compute_rest_springs(inst_, reinterpret_cast<const float*>(pos_), nullptr, k,
                         reinterpret_cast<hipStream_t>(stream));
    topo_dirty_ = true;
}

void ClothSim::build_fem_stretch(float k, std::uintptr_t stream) {
    const int T = int(inst_.tris.size() / 3);
    if (T > 0 && pos_ == 0)
        throw std::runtime_error(
            "ClothSimulator::build_fem_stretch_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    check_indices(inst_.tris, n_, "build_fem_stretch_from_current_positions");
    inst_.part0 = 0;
    compute_rest_stretch(inst_, reinterpret_cast<const float*>(pos_), nullptr, k,
                         reinterpret_cast<hipStream_t>(stream));
    topo_dirty_ = true;
}

void ClothSim::build_fem_shear(float k, std::uintptr_t stream) {
    const int T = int(inst_.tris.size() / 3);
    if (T > 0 && pos_ == 0)
        throw std::runtime_error(
            "ClothSimulator::build_fem_shear_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    check_indices(inst_.tris, n_, "build_fem_shear_from_current_positions");
    inst_.part0 = 0;
    compute_rest_shear(inst_, reinterpret_cast<const float*>(pos_), nullptr, k, reinterpret_cast<hipStream_t>(stream));
    topo_dirty_ = true;
}

void ClothSim::build_bending(float k, std::uintptr_t stream) {
    const int T = int(inst_.tris.size() / 3);
    if (T > 0 && pos_ == 0)
        throw std::runtime_error(
            "ClothSimulator::build_bending_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    check_indices(inst_.tris, n_, "build_bending_from_current_positions");
    inst_.part0 = 0;
    compute_rest_bending(inst_, reinterpret_cast<const float*>(pos_), k, reinterpret_cast<hipStream_t>(stream));
    topo_dirty_ = true;
}

void ClothSim::add_static_plane(const float* n, float d) {
    inst_.planes.insert(inst_.planes.end(), {n[0], n[1], n[2], d});
    params_dirty_ = true;
}

void ClothSim::add_static_box(const float* c, const float* h, const float* ex, const float* ey, const float* ez) {
    inst_.boxes.insert(inst_.boxes.end(), {c[0], c[1], c[2], h[0], h[1], h[2], ex[0], ex[1], ex[2], ey[0], ey[1], ey[2],
                                           ez[0], ez[1], ez[2], 0.0f});
    params_dirty_ = true;
}

void ClothSim::clear_static_shapes() {
    inst_.planes.clear();
    inst_.boxes.clear();
    params_dirty_ = true;
}

void ClothSim::set_spring_stiffnesses(const float* k, int n) {
    if (!inst_.has_springs) throw std::runtime_error("set_spring_stiffnesses: build springs first");
    if (n != int(inst_.spr_k.size())) throw std::invalid_argument("set_spring_stiffnesses: need one value per spring");
    std::copy(k, k + n, inst_.spr_k.begin());
    topo_dirty_ = true;
}

void ClothSim::set_fem_stretch_stiffnesses(const float* k, int n) {
    if (!inst_.has_stretch) throw std::runtime_error("set_fem_stretch_stiffnesses: build stretch first");
    if (n != int(inst_.tri_k.size()))
        throw std::invalid_argument("set_fem_stretch_stiffnesses: need one value per triangle");
    std::copy(k, k + n, inst_.tri_k.begin());
    topo_dirty_ = true;
}

void ClothSim::set_pins(const int* idx, const float* targets, int n, float k) {
    if (n < 0) throw std::invalid_argument("PinConstraint::set_pins: negative count");
    inst_.pin_k = k;
    inst_.pin_idx.assign(idx, idx + n);
    inst_.pin_tgt.assign(targets, targets + 3 * size_t(n));
    topo_dirty_ = true;
}

void ClothSim::clear_pins() {
    inst_.pin_idx.clear();
    inst_.pin_tgt.clear();
    topo_dirty_ = true;
}

void ClothSim::update_pin_targets(const float* targets, int n, std::uintptr_t stream) {
    if (targets == nullptr || n <= 0) return;
    if (n != int(inst_.pin_idx.size()))
        throw std::invalid_argument(
            "ClothSimulator::update_pin_targets: n_pins does not match the currently installed pin set; call "
            "set_pins(...) to change the pin set instead.");
    std::copy(targets, targets + 3 * size_t(n), inst_.pin_tgt.begin());
    if (!topo_dirty_ && prob_.built()) prob_.update_pin_targets(inst_, 0, reinterpret_cast<hipStream_t>(stream));
}

void ClothSim::redistribute_mass_area_weighted(float density, std::uintptr_t inv_mass, int n,
                                               std::uintptr_t stream) {
    if (n <= 0) return;
    if (inv_mass == 0)
        throw std::invalid_argument("ClothSimulator::redistribute_mass_area_weighted: inv_mass_ptr must be non-null");
    if (!(density > 0.0f))
        throw std::invalid_argument("ClothSimulator::redistribute_mass_area_weighted: surface_density must be positive");
    const int T = int(inst_.tris.size() / 3);
    if (T == 0)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted: call set_mesh(...) before redistributing mass.");
    if (pos_ == 0 || n_ < n)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted: external positions must be set first via "
            "set_external_buffers().");
    std::vector<int> off, ids;
    build_vertex_triangles(n, inst_.tris.data(), T, off, ids);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DeviceArray<int> doff, dids, dtri;
    doff.assign(off.data(), off.size(), s);
    dids.assign(ids.data(), ids.size(), s);
    dtri.assign(inst_.tris.data(), inst_.tris.size(), s);
    launch_area_mass(doff.data(), dids.data(), dtri.data(), reinterpret_cast<const float*>(pos_), density / 3.0f,
                     reinterpret_cast<float*>(inv_mass), n, s);
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize(area mass)");
}

void ClothSim::redistribute_mass_area_weighted_noalloc(float density, std::uintptr_t inv_mass, int n,
                                                       std::uintptr_t stream) {
    if (n <= 0) return;
    if (inv_mass == 0)
        throw std::invalid_argument(
            "ClothSimulator::redistribute_mass_area_weighted_noalloc: inv_mass_ptr must be non-null");
    if (!(density > 0.0f))
        throw std::invalid_argument(
            "ClothSimulator::redistribute_mass_area_weighted_noalloc: surface_density must be positive");
    const int T = int(inst_.tris.size() / 3);
    if (T == 0)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted_noalloc: call set_mesh(...) before redistributing mass.");
    if (pos_ == 0 || n_ < n)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted_noalloc: external positions must be set first via "
            "set_external_buffers().");
    area_mass_noalloc(inst_.tris.data(), T, n, reinterpret_cast<const float*>(pos_), density,
                      reinterpret_cast<float*>(inv_mass), reinterpret_cast<hipStream_t>(stream));
}

void ClothSim::set_pcg_tolerance(double tol) {
    if (!(tol >= 0.0)) throw std::invalid_argument("set_pcg_tolerance: tolerance must be >= 0");
    inst_.tol = tol;
    params_dirty_ = true;
}

void ClothSim::set_springs_active(bool on) {
    if (on != inst_.springs_active) topo_dirty_ = true;
    inst_.springs_active = on;
    params_dirty_ = true;
}

void ClothSim::ensure_topology() {
    if (!topo_dirty_ && prob_.built() && inst_.n == n_) return;
    inst_.n = n_;
    inst_.part0 = 0;
    prob_.build({&inst_});
    tune_pending_ = autotune_;
    topo_dirty_ = false;
    params_dirty_ = true;
}

void ClothSim::step(float dt, std::uintptr_t stream) {
    if (n_ <= 0) return;
    if (pos_ == 0 || vel_ == 0)
        throw std::runtime_error(
            "chysx::cloth::ClothSimulator::step: pos / vel spans must be set via set_external_buffers() before "
            "stepping.");
    if (!(dt > 0.0f)) throw std::invalid_argument("chysx::cloth::ClothSimulator::step: dt must be positive");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ensure_topology();
    prob_.upload_params({&inst_}, s);
    params_dirty_ = false;
    DevExt e{reinterpret_cast<float*>(pos_), reinterpret_cast<float*>(vel_),
             reinterpret_cast<const float*>(inv_mass_), reinterpret_cast<const float*>(force_)};
    if (tune_pending_) {
        prob_.autotune(s, dt, e);
        tune_pending_ = false;
    }
    prob_.step(dt, e, s);
    dumped_n_ = n_;
}

int ClothSim::num_off_diag_blocks() const {
    if (!prob_.built() || prob_.ninst() == 0) return 0;
    return int(prob_.pattern(0).cols.size());
}

SolveStats ClothSim::last_stats() const {
    if (!prob_.built() || prob_.ninst() == 0) return SolveStats();
    return prob_.stats()[0];
}

void ClothSim::set_warm_start(const double* dx) {
    ensure_topology();
    prob_.set_dx64(0, dx, nullptr);
}

void ClothSim::zero_warm_start() {
    if (prob_.built()) prob_.zero_dx(nullptr);
}

void ClothSim::debug_copy_hessian_diag(float* out) const {
    if (dumped_n_ == 0) return;
    if (!out) throw std::invalid_argument("ClothSimulator::debug_copy_hessian_diag: out must be non-null");
    prob_.copy_diag(0, out);
}
void ClothSim::debug_copy_hessian_csr(int* ro, int* cols, float* vals) const {
    if (dumped_n_ == 0) return;
    if (!ro) throw std::invalid_argument("ClothSimulator::debug_copy_hessian_csr: out_row_offsets null");
    const Pattern& P = prob_.pattern(0);
    std::copy(P.row_offsets.begin(), P.row_offsets.end(), ro);
    if (!P.cols.empty()) {
        if (!cols || !vals)
            throw std::invalid_argument(
                "ClothSimulator::debug_copy_hessian_csr: col_indices/values null but nnz_off > 0");
        std::copy(P.cols.begin(), P.cols.end(), cols);
        prob_.copy_values(0, vals);
    }
}
void ClothSim::debug_copy_last_rhs(float* out) const {
    if (dumped_n_ == 0) return;
    if (!out) throw std::invalid_argument("ClothSimulator::debug_copy_last_rhs: out must be non-null");
    prob_.copy_rhs(0, out);
}
void ClothSim::debug_copy_last_dx(float* out) const {
    if (dumped_n_ == 0) return;
    if (!out) throw std::invalid_argument("ClothSimulator::debug_copy_last_dx: out must be non-null");
    prob_.copy_dx(0, out);
}

// ---------------------------------------------------------------------------
// ClothBatch
// ---------------------------------------------------------------------------
int ClothBatch::add_cloth(const int* tris, int ntri, int n, const float* spring_k, float spring_k_scalar,
                          const float* stretch_k, float stretch_k_scalar, const int* pins, const float* pin_targets,
                          int npin, float pin_k) {
    if (ntri < 0 || n <= 0 || npin < 0) throw std::invalid_argument("ClothBatch::add_cloth: bad sizes");
    auto I = std::make_unique<Instance>();
    I->n = n;
    I->part0 = total_;
    I->tris.assign(tris, tris + 3 * size_t(ntri));
    for (int v : I->tris)
        if (v < 0 || v >= n) throw std::out_of_range("ClothBatch::add_cloth: triangle index out of range");
    I->edges = build_edges(tris, ntri);
    const int E = int(I->edges.size() / 2);
    pending_spring_k_.emplace_back(spring_k ? std::vector<float>(spring_k, spring_k + E)
                                            : std::vector<float>(size_t(E), spring_k_scalar));
    pending_stretch_k_.emplace_back(stretch_k ? std::vector<float>(stretch_k, stretch_k + ntri)
                                              : std::vector<float>(size_t(ntri), stretch_k_scalar));
    I->pin_idx.assign(pins, pins + npin);
    for (int v : I->pin_idx)
        if (v < 0 || v >= n) throw std::out_of_range("ClothBatch::add_cloth: pin index out of range");
    I->pin_tgt.assign(pin_targets, pin_targets + 3 * size_t(npin));
    I->pin_k = pin_k;
    I->springs_active = true;
    total_ += n;
    insts_.push_back(std::move(I));
    rest_done_.push_back(0);
    dirty_ = true;
    return int(insts_.size()) - 1;
}

void ClothBatch::set_external_buffers(std::uintptr_t pos, std::uintptr_t vel, int total, std::uintptr_t inv_mass,
                                      std::uintptr_t force) noexcept {
    pos_ = pos;
    vel_ = vel;
    ext_n_ = total;
    inv_mass_ = inv_mass;
    force_ = force;
}

void ClothBatch::set_material(int k, const Material& m) {
    insts_.at(size_t(k))->material = m;
    params_dirty_ = true;
}

void ClothBatch::set_options(double tol, int max_iter, bool springs_active, bool force_active) {
    for (auto& I : insts_) {
        if (I->springs_active != springs_active) dirty_ = true;
        I->tol = tol;
        I->max_iter = max_iter > 0 ? max_iter : 1;
        I->springs_active = springs_active;
        I->force_active = force_active;
    }
    params_dirty_ = true;
}

void ClothBatch::step(float dt, std::uintptr_t stream) {
    if (insts_.empty()) return;
    if (pos_ == 0 || vel_ == 0) throw std::runtime_error("ClothBatch::step: set_external_buffers() first");
    if (ext_n_ < total_) throw std::invalid_argument("ClothBatch::step: external buffers hold too few particles");
    if (!(dt > 0.0f)) throw std::invalid_argument("ClothBatch::step: dt must be positive");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (size_t k = 0; k < insts_.size(); ++k) {
        if (rest_done_[k]) continue;
        Instance& I = *insts_[k];
        compute_rest_springs(I, reinterpret_cast<const float*>(pos_), pending_spring_k_[k].data(), 0.0f, s);
        compute_rest_stretch(I, reinterpret_cast<const float*>(pos_), pending_stretch_k_[k].data(), 0.0f, s);
        rest_done_[k] = 1;
        dirty_ = true;
    }
    std::vector<const Instance*> v;
    for (auto& I : insts_) v.push_back(I.get());
    if (dirty_ || !prob_.built()) {
        prob_.build(v);
        tune_pending_ = autotune_;
        dirty_ = false;
    }
    prob_.upload_params(v, s);
    params_dirty_ = false;
    DevExt e{reinterpret_cast<float*>(pos_), reinterpret_cast<float*>(vel_),
             reinterpret_cast<const float*>(inv_mass_), reinterpret_cast<const float*>(force_)};
    if (tune_pending_) {
        prob_.autotune(s, dt, e);
        tune_pending_ = false;
    }
    prob_.step(dt, e, s);
}

void ClothBatch::set_pin_projection(bool on) {
    for (auto& I : insts_) I->pin_projection = on;
    params_dirty_ = true;
}

SolveStats ClothBatch::stats(int k) const {
    if (!prob_.built()) return SolveStats();
    return prob_.stats().at(size_t(k));
}

}  // namespace chysx_amd
