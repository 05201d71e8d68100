// SPDX-License-Identifier: Apache-2.0
//
// extern "C" boundary of libchysx_amd.so (declared in include/chysx_cloth.h).
#include "../../include/chysx_cloth.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace chysx_amd;

struct chysx_cloth {
    ClothSim sim;
};
struct chysx_batch {
    ClothBatch batch;
};
struct chysx_wsnet {
    WarmStartNet net;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guard(F&& f) {
    try {
        f();
        return CHYSX_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return CHYSX_EINVAL;
    } catch (const std::out_of_range& e) {
        g_last_error = e.what();
        return CHYSX_ERANGE;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CHYSX_ERUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return CHYSX_ERUNTIME;
    }
}

void need(const void* p, const char* what) {
    if (p == nullptr) throw std::invalid_argument(std::string(what) + " must be non-null");
}

Material to_mat(const chysx_cloth_material* m) {
    Material r;
    r.lame_mu = m->lame_mu;
    r.lame_lambda = m->lame_lambda;
    r.bending = m->bending;
    r.density = m->density;
    r.damping = m->damping;
    r.gx = m->gx;
    r.gy = m->gy;
    r.gz = m->gz;
    return r;
}

}  // namespace

extern "C" {

const char* chysx_last_error(void) { return g_last_error.c_str(); }
const char* chysx_version(void) { return "0.1.0"; }

int chysx_cloth_create(chysx_cloth** out) {
    return guard([&] {
        need(out, "out");
        *out = new chysx_cloth();
    });
}
int chysx_cloth_destroy(chysx_cloth* sim) {
    return guard([&] { delete sim; });
}
int chysx_cloth_set_material(chysx_cloth* sim, const chysx_cloth_material* m) {
    return guard([&] {
        need(sim, "sim");
        need(m, "material");
        sim->sim.set_material(to_mat(m));
    });
}
int chysx_cloth_get_material(const chysx_cloth* sim, chysx_cloth_material* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        const Material& m = sim->sim.material();
        *out = chysx_cloth_material{m.lame_mu, m.lame_lambda, m.bending, m.density, m.damping, m.gx, m.gy, m.gz};
    });
}
int chysx_cloth_set_external_buffers(chysx_cloth* sim, uintptr_t pos, uintptr_t vel, int n, uintptr_t inv_mass,
                                     uintptr_t force) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_external_buffers(pos, vel, n, inv_mass, force);
    });
}
int chysx_cloth_set_mesh(chysx_cloth* sim, const int32_t* tris, int ntri) {
    return guard([&] {
        need(sim, "sim");
        if (ntri > 0) need(tris, "triangles");
        sim->sim.set_mesh(tris, ntri);
    });
}
int chysx_cloth_build_springs(chysx_cloth* sim, float k, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.build_springs(k, stream);
    });
}
int chysx_cloth_num_springs(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_springs();
    });
}
int chysx_cloth_build_fem_stretch(chysx_cloth* sim, float k, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.build_fem_stretch(k, stream);
    });
}
int chysx_cloth_build_fem_shear(chysx_cloth* sim, float k, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.build_fem_shear(k, stream);
    });
}
int chysx_cloth_num_fem_shear(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_fem_shear();
    });
}
int chysx_cloth_build_bending(chysx_cloth* sim, float k, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.build_bending(k, stream);
    });
}
int chysx_cloth_num_bending(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_bending();
    });
}
int chysx_cloth_add_static_plane(chysx_cloth* sim, const float* n, float d) {
    return guard([&] {
        need(sim, "sim");
        need(n, "n");
        sim->sim.add_static_plane(n, d);
    });
}
int chysx_cloth_add_static_box(chysx_cloth* sim, const float* c, const float* h, const float* ex, const float* ey,
                               const float* ez) {
    return guard([&] {
        need(sim, "sim");
        need(c, "center");
        need(h, "half_ext");
        need(ex, "ex");
        need(ey, "ey");
        need(ez, "ez");
        sim->sim.add_static_box(c, h, ex, ey, ez);
    });
}
int chysx_cloth_clear_static_shapes(chysx_cloth* sim) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.clear_static_shapes();
    });
}
int chysx_cloth_static_shape_counts(const chysx_cloth* sim, int* planes, int* boxes) {
    return guard([&] {
        need(sim, "sim");
        if (planes) *planes = int(sim->sim.instance().planes.size() / 4);
        if (boxes) *boxes = int(sim->sim.instance().boxes.size() / 16);
    });
}
int chysx_cloth_set_static_contact(chysx_cloth* sim, int which, float value) {
    return guard([&] {
        need(sim, "sim");
        Instance& I = sim->sim.contact_instance();
        switch (which) {
            case 0: I.contact_h = value; break;
            case 1: I.contact_k = value; break;
            case 2: I.contact_mu = value; break;
            case 3: I.contact_eps = value; break;
            default: throw std::invalid_argument("set_static_contact: which must be 0..3");
        }
    });
}
int chysx_cloth_get_static_contact(const chysx_cloth* sim, int which, float* value) {
    return guard([&] {
        need(sim, "sim");
        need(value, "value");
        const Instance& I = sim->sim.instance();
        switch (which) {
            case 0: *value = I.contact_h; break;
            case 1: *value = I.contact_k; break;
            case 2: *value = I.contact_mu; break;
            case 3: *value = I.contact_eps; break;
            default: throw std::invalid_argument("get_static_contact: which must be 0..3");
        }
    });
}
int chysx_cloth_get_dihedrals(const chysx_cloth* sim, int32_t* dihedrals, float* rest_angles) {
    return guard([&] {
        need(sim, "sim");
        const Instance& I = sim->sim.instance();
        if (!I.has_bending) return;
        if (dihedrals && !I.dih.empty()) std::memcpy(dihedrals, I.dih.data(), I.dih.size() * sizeof(int));
        if (rest_angles && !I.dih_rest.empty())
            std::memcpy(rest_angles, I.dih_rest.data(), I.dih_rest.size() * sizeof(float));
    });
}
int chysx_cloth_num_fem_stretch(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_fem_stretch();
    });
}
int chysx_cloth_set_pins(chysx_cloth* sim, const int32_t* idx, const float* targets, int n, float k) {
    return guard([&] {
        need(sim, "sim");
        if (n > 0) {
            need(idx, "indices");
            need(targets, "targets");
        }
        sim->sim.set_pins(idx, targets, n, k);
    });
}
int chysx_cloth_clear_pins(chysx_cloth* sim) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.clear_pins();
    });
}
int chysx_cloth_num_pins(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_pins();
    });
}
int chysx_cloth_update_pin_targets(chysx_cloth* sim, const float* targets, int n, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.update_pin_targets(targets, n, stream);
    });
}
int chysx_cloth_redistribute_mass_area_weighted(chysx_cloth* sim, float density, uintptr_t inv_mass, int n,
                                                uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.redistribute_mass_area_weighted(density, inv_mass, n, stream);
    });
}
int chysx_cloth_redistribute_mass_area_weighted_noalloc(chysx_cloth* sim, float density, uintptr_t inv_mass, int n,
                                                        uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.redistribute_mass_area_weighted_noalloc(density, inv_mass, n, stream);
    });
}
int chysx_cloth_set_pcg_iterations(chysx_cloth* sim, int max_iter) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_pcg_iterations(max_iter);
    });
}
int chysx_cloth_pcg_iterations(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.pcg_iterations();
    });
}
int chysx_cloth_step(chysx_cloth* sim, float dt, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.step(dt, stream);
    });
}
int chysx_cloth_set_pin_projection(chysx_cloth* sim, int on) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_pin_projection(on != 0);
    });
}
int chysx_cloth_pin_projection(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.pin_projection() ? 1 : 0;
    });
}
int chysx_cloth_set_blocking_step(chysx_cloth* sim, int on) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.problem().set_blocking_step(on != 0);
    });
}
int chysx_cloth_num_particles(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_particles();
    });
}
int chysx_cloth_num_off_diag_blocks(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_off_diag_blocks();
    });
}
int chysx_cloth_debug_copy_hessian_diag(const chysx_cloth* sim, float* out) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.debug_copy_hessian_diag(out);
    });
}
int chysx_cloth_debug_copy_hessian_csr(const chysx_cloth* sim, int32_t* ro, int32_t* cols, float* vals) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.debug_copy_hessian_csr(ro, cols, vals);
    });
}
int chysx_cloth_debug_copy_last_rhs(const chysx_cloth* sim, float* out) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.debug_copy_last_rhs(out);
    });
}
int chysx_cloth_debug_copy_last_dx(const chysx_cloth* sim, float* out) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.debug_copy_last_dx(out);
    });
}
int chysx_cloth_set_pcg_tolerance(chysx_cloth* sim, double tol) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_pcg_tolerance(tol);
    });
}
int chysx_cloth_set_springs_active(chysx_cloth* sim, int on) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_springs_active(on != 0);
    });
}
int chysx_cloth_set_external_force_active(chysx_cloth* sim, int on) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_external_force_active(on != 0);
    });
}
int chysx_cloth_set_spring_stiffnesses(chysx_cloth* sim, const float* k, int n) {
    return guard([&] {
        need(sim, "sim");
        need(k, "k");
        sim->sim.set_spring_stiffnesses(k, n);
    });
}
int chysx_cloth_set_fem_stretch_stiffnesses(chysx_cloth* sim, const float* k, int n) {
    return guard([&] {
        need(sim, "sim");
        need(k, "k");
        sim->sim.set_fem_stretch_stiffnesses(k, n);
    });
}
int chysx_cloth_last_solve_stats(const chysx_cloth* sim, int* iters, int* restarts, int* status, double* relres) {
    return guard([&] {
        need(sim, "sim");
        const SolveStats s = sim->sim.last_stats();
        if (iters) *iters = s.iters;
        if (restarts) *restarts = s.restarts;
        if (status) *status = s.status;
        if (relres) *relres = s.relres;
    });
}
int chysx_cloth_set_warm_start(chysx_cloth* sim, const double* dx, int n) {
    return guard([&] {
        need(sim, "sim");
        need(dx, "dx");
        sim->sim.ensure_topology();
        if (n != sim->sim.instance().n) throw std::invalid_argument("set_warm_start: n does not match particle count");
        sim->sim.set_warm_start(dx);
    });
}
int chysx_cloth_zero_warm_start(chysx_cloth* sim) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.zero_warm_start();
    });
}
int chysx_cloth_get_dx64(const chysx_cloth* sim, double* out, int n) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        if (n != sim->sim.instance().n) throw std::invalid_argument("get_dx64: n does not match particle count");
        const_cast<chysx_cloth*>(sim)->sim.problem().copy_dx64(0, out);
    });
}
int chysx_cloth_num_edges(const chysx_cloth* sim, int* out) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        *out = sim->sim.num_edges();
    });
}
int chysx_cloth_get_edges(const chysx_cloth* sim, int32_t* out) {
    return guard([&] {
        need(sim, "sim");
        const auto& e = sim->sim.edges();
        if (!e.empty()) need(out, "out");
        std::memcpy(out, e.data(), e.size() * sizeof(int));
    });
}

int chysx_topology_query(int n, const int32_t* tris, int ntri, int include_springs, const int32_t* pins, int npin,
                         int* nnz, int* nedges, int32_t* row_offsets, int32_t* col_indices, int32_t* edges,
                         int32_t* tri_slots, int32_t* spring_slots, int32_t* pin_slots) {
    return guard([&] {
        if (n < 0 || ntri < 0 || npin < 0) throw std::invalid_argument("topology_query: negative size");
        if (ntri > 0) need(tris, "triangles");
        if (npin > 0) need(pins, "pins");
        const std::vector<int> e = build_edges(tris, ntri);
        const int E = int(e.size() / 2);
        const int* spr = (include_springs && E) ? e.data() : nullptr;
        const int Es = spr ? E : 0;
        const Pattern p = build_pattern(n, tris, ntri, spr, Es);
        std::vector<int> ts, ss, ps;
        build_slot_maps(p, tris, ntri, spr, Es, pins, npin, ts, ss, ps);
        if (nnz) *nnz = int(p.cols.size());
        if (nedges) *nedges = E;
        if (row_offsets) std::memcpy(row_offsets, p.row_offsets.data(), p.row_offsets.size() * sizeof(int));
        if (col_indices && !p.cols.empty()) std::memcpy(col_indices, p.cols.data(), p.cols.size() * sizeof(int));
        if (edges && !e.empty()) std::memcpy(edges, e.data(), e.size() * sizeof(int));
        if (tri_slots && !ts.empty()) std::memcpy(tri_slots, ts.data(), ts.size() * sizeof(int));
        if (spring_slots && !ss.empty()) std::memcpy(spring_slots, ss.data(), ss.size() * sizeof(int));
        if (pin_slots && !ps.empty()) std::memcpy(pin_slots, ps.data(), ps.size() * sizeof(int));
    });
}

int chysx_topology_query_bending(int n, const int32_t* tris, int ntri, int* nnz, int* ndih, int32_t* row_offsets,
                                 int32_t* col_indices, int32_t* dihedrals, int32_t* dihedral_slots) {
    return guard([&] {
        if (n < 0 || ntri < 0) throw std::invalid_argument("topology_query: negative size");
        if (ntri > 0) need(tris, "triangles");
        for (int q = 0; q < 3 * ntri; ++q)
            if (tris[q] < 0 || tris[q] >= n) throw std::out_of_range("topology_query: vertex index out of range");
        const std::vector<int> d = find_dihedrals(tris, ntri);
        const int D = int(d.size() / 4);
        const Pattern p = build_pattern(n, tris, ntri, nullptr, 0, D ? d.data() : nullptr, D);
        const std::vector<int> ds = build_dihedral_slots(p, D ? d.data() : nullptr, D);
        if (nnz) *nnz = int(p.cols.size());
        if (ndih) *ndih = D;
        if (row_offsets) std::memcpy(row_offsets, p.row_offsets.data(), p.row_offsets.size() * sizeof(int));
        if (col_indices && !p.cols.empty()) std::memcpy(col_indices, p.cols.data(), p.cols.size() * sizeof(int));
        if (dihedrals && !d.empty()) std::memcpy(dihedrals, d.data(), d.size() * sizeof(int));
        if (dihedral_slots && !ds.empty()) std::memcpy(dihedral_slots, ds.data(), ds.size() * sizeof(int));
    });
}

int chysx_batch_create(chysx_batch** out) {
    return guard([&] {
        need(out, "out");
        *out = new chysx_batch();
    });
}
int chysx_batch_destroy(chysx_batch* b) {
    return guard([&] { delete b; });
}
int chysx_batch_add_cloth(chysx_batch* b, const int32_t* tris, int ntri, int n, const float* spring_k,
                          float spring_k_scalar, const float* stretch_k, float stretch_k_scalar, const int32_t* pins,
                          const float* pin_targets, int npin, float pin_k, int* index) {
    return guard([&] {
        need(b, "batch");
        if (ntri > 0) need(tris, "triangles");
        if (npin > 0) {
            need(pins, "pins");
            need(pin_targets, "pin_targets");
        }
        const int k = b->batch.add_cloth(tris, ntri, n, spring_k, spring_k_scalar, stretch_k, stretch_k_scalar, pins,
                                         pin_targets, npin, pin_k);
        if (index) *index = k;
    });
}
int chysx_batch_set_external_buffers(chysx_batch* b, uintptr_t pos, uintptr_t vel, int total, uintptr_t inv_mass,
                                     uintptr_t force) {
    return guard([&] {
        need(b, "batch");
        b->batch.set_external_buffers(pos, vel, total, inv_mass, force);
    });
}
int chysx_batch_set_material(chysx_batch* b, int index, const chysx_cloth_material* m) {
    return guard([&] {
        need(b, "batch");
        need(m, "material");
        if (index < 0 || index >= b->batch.count()) throw std::out_of_range("batch: cloth index out of range");
        b->batch.set_material(index, to_mat(m));
    });
}
int chysx_batch_set_options(chysx_batch* b, double tol, int max_iter, int springs_active, int force_active) {
    return guard([&] {
        need(b, "batch");
        if (!(tol >= 0.0)) throw std::invalid_argument("batch: tolerance must be >= 0");
        b->batch.set_options(tol, max_iter, springs_active != 0, force_active != 0);
    });
}
int chysx_batch_step(chysx_batch* b, float dt, uintptr_t stream) {
    return guard([&] {
        need(b, "batch");
        b->batch.step(dt, stream);
    });
}
int chysx_batch_set_split_advance(chysx_batch* b, int mode) {
    return guard([&] {
        need(b, "batch");
        b->batch.problem().set_split_advance(mode);
    });
}
int chysx_batch_set_pin_projection(chysx_batch* b, int on) {
    return guard([&] {
        need(b, "batch");
        b->batch.set_pin_projection(on != 0);
    });
}
int chysx_batch_set_blocking_step(chysx_batch* b, int on) {
    return guard([&] {
        need(b, "batch");
        b->batch.problem().set_blocking_step(on != 0);
    });
}
int chysx_batch_stats(const chysx_batch* b, int index, int* iters, int* restarts, int* status, double* relres) {
    return guard([&] {
        need(b, "batch");
        if (index < 0 || index >= b->batch.count()) throw std::out_of_range("batch: cloth index out of range");
        const SolveStats s = b->batch.stats(index);
        if (iters) *iters = s.iters;
        if (restarts) *restarts = s.restarts;
        if (status) *status = s.status;
        if (relres) *relres = s.relres;
    });
}

int chysx_batch_num_cloths(const chysx_batch* b, int* n) {
    return guard([&] {
        need(b, "batch");
        need(n, "n");
        *n = b->batch.count();
    });
}

int chysx_batch_stats_all(const chysx_batch* b, int n, int* iters, int* restarts, int* status, double* relres) {
    return guard([&] {
        need(b, "batch");
        if (n != b->batch.count()) throw std::invalid_argument("batch: n must be the number of cloths");
        for (int k = 0; k < n; ++k) {
            const SolveStats s = b->batch.stats(k);
            if (iters) iters[k] = s.iters;
            if (restarts) restarts[k] = s.restarts;
            if (status) status[k] = s.status;
            if (relres) relres[k] = s.relres;
        }
    });
}

int chysx_cloth_spmv(chysx_cloth* sim, uintptr_t x, uintptr_t y, uintptr_t stream) {
    return guard([&] {
        need(sim, "sim");
        if (x == 0 || y == 0) throw std::invalid_argument("spmv: x and y must be non-null device pointers");
        Problem& P = sim->sim.problem();
        if (!P.built() || sim->sim.num_particles() == 0)
            throw std::runtime_error("spmv: step() once to assemble a system first");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        P.wait_pending(s);
        launch_spmv(P.geom(), P.matrix(), reinterpret_cast<const float*>(x), reinterpret_cast<float*>(y),
                    P.launch_config().spmv_waves, s);
    });
}

int chysx_stream_bench(int kind, uint64_t bytes_per_buffer, int reps, uintptr_t stream, double* ms,
                       double* bytes_moved) {
    return guard([&] {
        need(ms, "ms");
        *ms = stream_bench(kind, size_t(bytes_per_buffer), reps, reinterpret_cast<hipStream_t>(stream), bytes_moved);
    });
}

int chysx_cloth_spmv_bytes(const chysx_cloth* sim, double* bytes) {
    return guard([&] {
        need(sim, "sim");
        need(bytes, "bytes");
        *bytes = const_cast<chysx_cloth*>(sim)->sim.problem().spmv_bytes();
    });
}

int chysx_cloth_time_kernel(chysx_cloth* sim, int which, int reps, float dt, uintptr_t stream, double* ms) {
    return guard([&] {
        need(sim, "sim");
        need(ms, "ms");
        if (reps <= 0) throw std::invalid_argument("time_kernel: reps must be positive");
        if (which < 0 || which > 5) throw std::invalid_argument("time_kernel: unknown kernel id");
        ClothSim& S = sim->sim;
        if (!S.problem().built()) throw std::runtime_error("time_kernel: step() once before timing");
        Problem& P = S.problem();
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        // Keep every PCG kernel live: no tolerance test, no iteration cap.
        Instance tmp = S.instance();
        tmp.tol = 0.0;
        tmp.max_iter = INT_MAX;
        P.upload_params({&tmp}, s);
        const DevExt ext = S.ext();
        const DevMatrix A = P.matrix();
        const DevVectors v = P.vectors();
        const LaunchGeom g = P.geom();
        const LaunchConfig& c = P.launch_config();
        P.prime_states(P.states(), s);
        hipEvent_t e0, e1;
        check_hip(hipEventCreate(&e0), "hipEventCreate");
        check_hip(hipEventCreate(&e1), "hipEventCreate");
        // 1: the SpMV pass as the solve runs it (advancing the state between the
        // two buffers); 4: one iteration, timed as a whole chunk (its passes and
        // closing advance) over its iteration count; 5: the SpMV pass without
        // the advance.
        int pass = 0;
        auto run = [&](int n) {
            for (int i = 0; i < n; ++i) {
                switch (which) {
                    case 0: launch_spmv(g, A, v.p0, v.ap, c.spmv_waves, s); break;
                    case 1: {
                        InstState* in = (pass & 1) ? P.states_b() : P.states();
                        InstState* out = (pass & 1) ? P.states() : P.states_b();
                        ++pass;
                        if (P.split_advance()) {
                            launch_pcg_advance_copy(in, out, P.params(), v.partials, P.ninst(), g.ngroups, s,
                                                    g.gactive);
                            launch_pcg_spmv(g, A, v, out, nullptr, nullptr, c.spmv_unroll, s, c.spmv_branch_free,
                                            c.spmv_symmetric);
                        } else {
                            launch_pcg_spmv(g, A, v, in, out, P.params(), c.spmv_unroll, s, c.spmv_branch_free,
                                            c.spmv_symmetric);
                        }
                        break;
                    }
                    case 2: launch_pcg_update(g, A, v, P.states(), s); break;
                    case 3: launch_assemble(g, P.mesh(), A, v, ext, P.params(), dt, c.assemble_waves, P.max_width(), s); break;
                    case 4: P.launch_pcg_chunk_body(s); break;
                    case 5:
                        launch_pcg_spmv(g, A, v, P.states(), nullptr, nullptr, c.spmv_unroll, s, c.spmv_branch_free,
                                        c.spmv_symmetric);
                        break;
                }
            }
        };
        run(3);
        check_hip(hipEventRecord(e0, s), "hipEventRecord");
        run(reps);
        check_hip(hipEventRecord(e1, s), "hipEventRecord");
        check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0.0f;
        check_hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = double(t) / reps / (which == 4 ? Problem::kChunkIters : 1);
        P.upload_params({&S.instance()}, s);
    });
}

int chysx_cloth_set_autotune(chysx_cloth* sim, int on) {
    return guard([&] {
        need(sim, "sim");
        sim->sim.set_autotune(on != 0);
    });
}

int chysx_cloth_launch_config(const chysx_cloth* sim, int* cfg) {
    return guard([&] {
        need(sim, "sim");
        need(cfg, "cfg");
        const LaunchConfig& c = const_cast<chysx_cloth*>(sim)->sim.problem().launch_config();
        const int v[CHYSX_LAUNCH_CONFIG_LEN] = {c.assemble_waves,   c.residual_waves, c.spmv_waves,
                                                c.spmv_unroll,      c.update_waves,   c.finalize_waves,
                                                c.spmv_branch_free, c.spmv_symmetric, c.spmv_packed,
                                                c.spmv_index16,     c.pcg_split_advance};
        std::memcpy(cfg, v, sizeof(v));
    });
}

int chysx_cloth_set_launch_config(chysx_cloth* sim, const int* cfg) {
    return guard([&] {
        need(sim, "sim");
        need(cfg, "cfg");
        sim->sim.ensure_topology();
        LaunchConfig c;
        c.assemble_waves = cfg[0];
        c.residual_waves = cfg[1];
        c.spmv_waves = cfg[2];
        c.spmv_unroll = cfg[3];
        c.update_waves = cfg[4];
        c.finalize_waves = cfg[5];
        c.spmv_branch_free = cfg[6];
        c.spmv_symmetric = cfg[7];
        c.spmv_packed = cfg[8];
        c.spmv_index16 = cfg[9];
        c.pcg_split_advance = cfg[10];
        sim->sim.problem().set_launch_config(c);
    });
}

int chysx_cloth_index_bits(const chysx_cloth* sim, int* bits) {
    return guard([&] {
        need(sim, "sim");
        need(bits, "bits");
        *bits = const_cast<chysx_cloth*>(sim)->sim.problem().index_bits();
    });
}

int chysx_cloth_block_floats(const chysx_cloth* sim, int* floats) {
    return guard([&] {
        need(sim, "sim");
        need(floats, "floats");
        *floats = const_cast<chysx_cloth*>(sim)->sim.problem().block_floats();
    });
}

int chysx_cloth_tune_log(const chysx_cloth* sim, int* count, int* kernel_waves_unroll, double* ms) {
    return guard([&] {
        need(sim, "sim");
        need(count, "count");
        const auto& L = const_cast<chysx_cloth*>(sim)->sim.problem().tune_log();
        if (kernel_waves_unroll && ms) {
            const int n = std::min<int>(*count, int(L.size()));
            for (int i = 0; i < n; ++i) {
                kernel_waves_unroll[3 * i] = L[size_t(i)].kernel;
                kernel_waves_unroll[3 * i + 1] = L[size_t(i)].waves;
                kernel_waves_unroll[3 * i + 2] = L[size_t(i)].unroll;
                ms[i] = L[size_t(i)].ms;
            }
        }
        *count = int(L.size());
    });
}

int chysx_batch_set_autotune(chysx_batch* b, int on) {
    return guard([&] {
        need(b, "batch");
        b->batch.set_autotune(on != 0);
    });
}

// ---- learned warm start ------------------------------------------------------
}  // extern "C"

namespace {
Problem& cloth_prob(chysx_cloth* sim) { return sim->sim.problem(); }
Problem& batch_prob(chysx_batch* b) { return b->batch.problem(); }

void fill_ws_stats(const SolveStats& st, double loss, double* out) {
    out[0] = st.ws_proj_used;
    out[1] = st.ws_used;
    out[2] = st.ws_kdim;
    out[3] = st.ws_r0_prev;
    out[4] = st.ws_r0_proj;
    out[5] = st.ws_r0_learned;
    out[6] = loss;
}
void set_ws_params(Problem& P, const float* theta, int n) {
    need(theta, "params");
    if (n != WarmStartNet::num_params()) throw std::invalid_argument("warm start: wrong parameter count");
    P.net().set_params(theta, nullptr);
}
void get_ws_params(Problem& P, float* theta, int n) {
    need(theta, "params");
    if (n != WarmStartNet::num_params()) throw std::invalid_argument("warm start: wrong parameter count");
    if (!P.net().has_params()) throw std::runtime_error("warm start: no parameters set");
    P.net().get_params(theta);
}
}  // namespace

extern "C" {

int chysx_ws_num_params(int* out) {
    return guard([&] {
        need(out, "out");
        *out = WarmStartNet::num_params();
    });
}
int chysx_cloth_set_warm_start_mode(chysx_cloth* sim, int mode) {
    return guard([&] {
        need(sim, "sim");
        cloth_prob(sim).set_warm_start_mode(mode);
    });
}
int chysx_cloth_ws_set_params(chysx_cloth* sim, const float* theta, int n) {
    return guard([&] {
        need(sim, "sim");
        set_ws_params(cloth_prob(sim), theta, n);
    });
}
int chysx_cloth_ws_get_params(chysx_cloth* sim, float* theta, int n) {
    return guard([&] {
        need(sim, "sim");
        get_ws_params(cloth_prob(sim), theta, n);
    });
}
int chysx_cloth_ws_init_params(chysx_cloth* sim, uint64_t seed) {
    return guard([&] {
        need(sim, "sim");
        cloth_prob(sim).net().init_params(seed, nullptr);
    });
}
int chysx_cloth_ws_set_training(chysx_cloth* sim, int on, float lr) {
    return guard([&] {
        need(sim, "sim");
        cloth_prob(sim).set_ws_training(on != 0, lr);
    });
}
int chysx_cloth_ws_set_acceptance(chysx_cloth* sim, int residual) {
    return guard([&] {
        need(sim, "sim");
        cloth_prob(sim).set_ws_accept_residual(residual != 0);
    });
}
int chysx_cloth_ws_set_history(chysx_cloth* sim, int k) {
    return guard([&] {
        need(sim, "sim");
        cloth_prob(sim).set_ws_history(k);
    });
}
int chysx_cloth_ws_stats(chysx_cloth* sim, double* out, int n) {
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        if (n < CHYSX_WS_STATS_LEN) throw std::invalid_argument("ws_stats: out too short");
        fill_ws_stats(sim->sim.last_stats(), cloth_prob(sim).last_ws_loss(), out);
    });
}
int chysx_batch_set_warm_start_mode(chysx_batch* b, int mode) {
    return guard([&] {
        need(b, "batch");
        batch_prob(b).set_warm_start_mode(mode);
    });
}
int chysx_batch_ws_set_params(chysx_batch* b, const float* theta, int n) {
    return guard([&] {
        need(b, "batch");
        set_ws_params(batch_prob(b), theta, n);
    });
}
int chysx_batch_ws_get_params(chysx_batch* b, float* theta, int n) {
    return guard([&] {
        need(b, "batch");
        get_ws_params(batch_prob(b), theta, n);
    });
}
int chysx_batch_ws_init_params(chysx_batch* b, uint64_t seed) {
    return guard([&] {
        need(b, "batch");
        batch_prob(b).net().init_params(seed, nullptr);
    });
}
int chysx_batch_ws_set_training(chysx_batch* b, int on, float lr) {
    return guard([&] {
        need(b, "batch");
        batch_prob(b).set_ws_training(on != 0, lr);
    });
}
int chysx_batch_ws_set_acceptance(chysx_batch* b, int residual) {
    return guard([&] {
        need(b, "batch");
        batch_prob(b).set_ws_accept_residual(residual != 0);
    });
}
int chysx_batch_ws_set_history(chysx_batch* b, int k) {
    return guard([&] {
        need(b, "batch");
        batch_prob(b).set_ws_history(k);
    });
}
int chysx_batch_ws_stats(chysx_batch* b, int index, double* out, int n) {
    return guard([&] {
        need(b, "batch");
        need(out, "out");
        if (n < CHYSX_WS_STATS_LEN) throw std::invalid_argument("ws_stats: out too short");
        if (index < 0 || index >= b->batch.count()) throw std::out_of_range("batch: cloth index out of range");
        fill_ws_stats(b->batch.stats(index), batch_prob(b).last_ws_loss(), out);
    });
}
int chysx_ws_small_solve(int k, const double* G, const double* g, double* c, int* kept) {
    return guard([&] {
        need(G, "G");
        need(g, "g");
        need(c, "c");
        const int m = ws_small_solve_host(k, G, g, c);
        if (kept) *kept = m;
    });
}

// ---- stand-alone network (tests) ---------------------------------------------
namespace {
struct NetGraph {
    DeviceArray<SliceDesc> slices;
    DeviceArray<int> slice_ptr, col, row_deg;
    WsGraph g{};
    int rows = 0;
    void build(int n, const int32_t* ro, const int32_t* cols) {
        if (n <= 0) throw std::invalid_argument("wsnet: n must be positive");
        Pattern p;
        p.n = n;
        p.row_offsets.assign(ro, ro + n + 1);
        p.cols.assign(cols, cols + ro[n]);
        for (int c : p.cols)
            if (c < 0 || c >= n) throw std::out_of_range("wsnet: column out of range");
        const DeviceLayout L = build_device_layout(p, nullptr, 0, {}, nullptr, 0, {}, nullptr, 0);
        std::vector<SliceDesc> sd(static_cast<size_t>(L.nslices));
        // nlower: the slice width (every entry may be lower), a bound that holds for any pattern
        for (int q = 0; q < L.nslices; ++q)
            sd[size_t(q)] = SliceDesc{0, 64 * q, std::min(64, n - 64 * q), L.slice_ptr[size_t(q) + 1] - L.slice_ptr[size_t(q)],
                                      0, (L.nslices + kGroupSlices - 1) / kGroupSlices};
        slices.assign(sd.data(), sd.size());
        slice_ptr.assign(L.slice_ptr.data(), L.slice_ptr.size());
        col.assign(L.sell_col.data(), L.sell_col.size());
        row_deg.assign(L.row_deg.data(), L.row_deg.size());
        g = WsGraph{slices.data(), L.nslices, slice_ptr.data(), col.data(), row_deg.data()};
        rows = L.n_pad;
    }
};
}  // namespace

int chysx_wsnet_create(chysx_wsnet** out) {
    return guard([&] {
        need(out, "out");
        *out = new chysx_wsnet();
    });
}
int chysx_wsnet_destroy(chysx_wsnet* net) {
    return guard([&] { delete net; });
}
int chysx_wsnet_set_params(chysx_wsnet* net, const float* theta, int n) {
    return guard([&] {
        need(net, "net");
        need(theta, "params");
        if (n != WarmStartNet::num_params()) throw std::invalid_argument("warm start: wrong parameter count");
        net->net.set_params(theta, nullptr);
    });
}
int chysx_wsnet_forward(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols, const float* F,
                        const float* E, float* out) {
    return guard([&] {
        need(net, "net");
        need(row_offsets, "row_offsets");
        need(F, "F");
        need(E, "E");
        need(out, "out");
        NetGraph G;
        G.build(n, row_offsets, cols);
        net->net.reserve(G.rows, false);
        net->net.set_inputs(F, E, n, nullptr);
        net->net.forward(G.g, nullptr, false, false);
        net->net.copy_out(out, n, nullptr);
    });
}
int chysx_wsnet_forward_direction(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols,
                                  const float* F, const float* E, float* out) {
    return guard([&] {
        need(net, "net");
        need(row_offsets, "row_offsets");
        need(F, "F");
        need(E, "E");
        need(out, "out");
        NetGraph G;
        G.build(n, row_offsets, cols);
        net->net.reserve(G.rows, false);
        net->net.set_inputs(F, E, n, nullptr);
        net->net.forward(G.g, nullptr, false, true);
        std::vector<float> full(size_t(n) * WarmStartNet::kOut);
        net->net.copy_out(full.data(), n, nullptr);
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) out[3 * size_t(i) + c] = full[size_t(i) * WarmStartNet::kOut + c];
    });
}
int chysx_wsnet_loss_grad(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols, const float* F,
                          const float* E, const float* target, double* loss, float* grad) {
    return guard([&] {
        need(net, "net");
        need(row_offsets, "row_offsets");
        need(F, "F");
        need(E, "E");
        need(target, "target");
        need(grad, "grad");
        NetGraph G;
        G.build(n, row_offsets, cols);
        net->net.reserve(G.rows, true);
        net->net.set_inputs(F, E, n, nullptr);
        DeviceArray<float> t;
        std::vector<float> th(size_t(G.rows) * 3, 0.0f);
        std::copy(target, target + 3 * size_t(n), th.begin());
        t.assign(th.data(), th.size());
        net->net.forward(G.g, nullptr, true);
        net->net.loss(G.g, t.data(), nullptr);
        net->net.backward(G.g, nullptr);
        if (loss) *loss = net->net.read_loss(nullptr);
        net->net.copy_grad(grad, nullptr);
    });
}

}  // extern "C"
