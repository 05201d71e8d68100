// SPDX-License-Identifier: Apache-2.0
//
// pybind11 module _chysx_native: the reference's Python surface for the cloth
// step (src/bindings.cpp:22-594, 1069-1075), implemented over the C ABI of
// libchysx_amd.so.  Argument names, defaults and exception types follow the
// reference bindings.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/chysx_cloth.h"

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc == CHYSX_OK) return;
    const std::string msg = chysx_last_error();
    if (rc == CHYSX_EINVAL) throw py::value_error(msg);
    if (rc == CHYSX_ERANGE) throw py::index_error(msg);
    throw std::runtime_error(msg);
}

using IntArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using FltArr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using DblArr = py::array_t<double, py::array::c_style | py::array::forcecast>;

struct Sim;

// ClothMaterial: a free-standing POD, or (as returned by
// ClothSimulator.material) a live view whose writes go to the simulator.
struct Material {
    chysx_cloth_material m{0.0f, 0.0f, 0.0f, 0.05f, 0.0f, 0.0f, 0.0f, -9.81f};
    std::shared_ptr<chysx_cloth*> owner;  // set for views

    chysx_cloth_material get() const {
        if (owner && *owner) {
            chysx_cloth_material r;
            check(chysx_cloth_get_material(*owner, &r));
            return r;
        }
        return m;
    }
    void put(const chysx_cloth_material& v) {
        if (owner && *owner) check(chysx_cloth_set_material(*owner, &v));
        else m = v;
    }
};

struct Sim {
    std::shared_ptr<chysx_cloth*> h;
    Sim() : h(std::make_shared<chysx_cloth*>(nullptr)) { check(chysx_cloth_create(h.get())); }
    ~Sim() {
        if (*h) chysx_cloth_destroy(*h);
        *h = nullptr;
    }
    Sim(const Sim&) = delete;
    Sim& operator=(const Sim&) = delete;
    chysx_cloth* p() const { return *h; }
};

struct Batch {
    chysx_batch* h = nullptr;
    Batch() { check(chysx_batch_create(&h)); }
    ~Batch() {
        if (h) chysx_batch_destroy(h);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
};

#define MAT_FIELD(name, doc)                                                                  \
    def_property(                                                                             \
        #name, [](const Material& s) { return s.get().name; },                                \
        [](Material& s, float v) {                                                            \
            chysx_cloth_material c = s.get();                                                 \
            c.name = v;                                                                       \
            s.put(c);                                                                         \
        },                                                                                    \
        doc)

py::dict stats_dict(int it, int rs, int st, double rr) {
    py::dict d;
    d["iters"] = it;
    d["restarts"] = rs;
    d["status"] = st;
    d["converged"] = (st == CHYSX_PCG_CONVERGED);
    d["relres"] = rr;
    return d;
}

int ws_mode(const std::string& m) {
    if (m == "previous") return 0;
    if (m == "learned") return 1;
    if (m == "projection") return 2;
    if (m == "projection_learned") return 3;
    throw py::value_error("warm start mode must be 'previous', 'learned', 'projection' or 'projection_learned'");
}

py::dict ws_stats_dict(const double* o) {
    py::dict d;
    d["projection_used"] = int(o[0]);
    d["used"] = int(o[1]);
    d["projection_dims"] = int(o[2]);
    d["r0_previous"] = o[3];
    d["r0_projection"] = o[4];
    d["r0_learned"] = o[5];
    d["loss"] = o[6];
    return d;
}

py::array_t<float> params_out(int n) { return py::array_t<float>({n}); }

}  // namespace

PYBIND11_MODULE(_chysx_native, m) {
    m.doc() = "ChysX cloth step on AMD Instinct MI355X (gfx950)";

    py::class_<Material>(m, "ClothMaterial", "Plain-old-data material parameters for the cloth simulator.")
        .def(py::init<>())
        .MAT_FIELD(lame_mu, "Lame mu [Pa] (in-plane elasticity, unused by the step).")
        .MAT_FIELD(lame_lambda, "Lame lambda [Pa] (unused by the step).")
        .MAT_FIELD(bending, "Dihedral bending stiffness [N m] (unused by the step).")
        .MAT_FIELD(density, "Surface mass density [kg/m^2].")
        .MAT_FIELD(damping, "Velocity damping [1/s] (v *= exp(-damping * dt)).")
        .MAT_FIELD(gx, "Gravity x-component [m/s^2].")
        .MAT_FIELD(gy, "Gravity y-component [m/s^2].")
        .MAT_FIELD(gz, "Gravity z-component [m/s^2].");

    py::class_<Sim>(m, "ClothSimulator", "Implicit cloth step (block-CSR Hessian + block-Jacobi PCG) on gfx950.")
        .def(py::init<>())
        .def(
            "set_material", [](Sim& s, const Material& mat) {
                const chysx_cloth_material c = mat.get();
                check(chysx_cloth_set_material(s.p(), &c));
            },
            py::arg("material"), "Copy `material` into the simulator.")
        .def(
            "set_external_buffers",
            [](Sim& s, std::uintptr_t pos, std::uintptr_t vel, int n, std::uintptr_t inv_mass, std::uintptr_t force) {
                check(chysx_cloth_set_external_buffers(s.p(), pos, vel, n, inv_mass, force));
            },
            py::arg("pos_ptr"), py::arg("vel_ptr"), py::arg("particle_count"), py::arg("inv_mass_ptr") = 0,
            py::arg("force_ptr") = 0,
            "Stash caller-owned device pointers (float3 AoS positions / velocities / forces, float inverse mass).")
        .def(
            "step", [](Sim& s, float dt, std::uintptr_t stream) { check(chysx_cloth_step(s.p(), dt, stream)); },
            py::arg("dt"), py::arg("cuda_stream") = 0,
            "Advance the simulation by `dt` seconds; pos / vel are updated in place.")
        .def(
            "set_pins",
            [](Sim& s, IntArr indices, FltArr targets, float stiffness) {
                if (indices.ndim() != 1) throw py::value_error("ClothSimulator.set_pins: indices must be 1-D");
                if (targets.ndim() != 2 || targets.shape(1) != 3)
                    throw py::value_error("ClothSimulator.set_pins: targets must have shape (N, 3)");
                if (indices.shape(0) != targets.shape(0))
                    throw py::value_error("ClothSimulator.set_pins: indices and targets must have the same length");
                check(chysx_cloth_set_pins(s.p(), indices.data(), targets.data(), int(indices.shape(0)), stiffness));
            },
            py::arg("indices"), py::arg("targets"), py::arg("stiffness") = 1.0e6f, "Install penalty pins.")
        .def("clear_pins", [](Sim& s) { check(chysx_cloth_clear_pins(s.p())); })
        .def("num_pins",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_pins(s.p(), &n));
                 return n;
             })
        .def(
            "update_pin_targets",
            [](Sim& s, FltArr targets, std::uintptr_t stream) {
                if (targets.ndim() != 2 || targets.shape(1) != 3)
                    throw py::value_error(
                        "ClothSimulator.update_pin_targets: targets must have shape (N, 3) and dtype float32");
                check(chysx_cloth_update_pin_targets(s.p(), targets.data(), int(targets.shape(0)), stream));
            },
            py::arg("targets"), py::arg("cuda_stream") = 0)
        .def(
            "set_mesh",
            [](Sim& s, IntArr tris) {
                if (tris.ndim() != 2 || tris.shape(1) != 3)
                    throw py::value_error("ClothSimulator.set_mesh: triangles must have shape (M, 3) and dtype int32");
                check(chysx_cloth_set_mesh(s.p(), tris.data(), int(tris.shape(0))));
            },
            py::arg("triangles"))
        .def(
            "build_springs_from_current_positions",
            [](Sim& s, float k, std::uintptr_t stream) { check(chysx_cloth_build_springs(s.p(), k, stream)); },
            py::arg("stiffness"), py::arg("cuda_stream") = 0)
        .def("num_springs",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_springs(s.p(), &n));
                 return n;
             })
        .def(
            "redistribute_mass_area_weighted",
            [](Sim& s, float density, std::uintptr_t inv_mass, int n, std::uintptr_t stream) {
                check(chysx_cloth_redistribute_mass_area_weighted(s.p(), density, inv_mass, n, stream));
            },
            py::arg("surface_density"), py::arg("inv_mass_ptr"), py::arg("particle_count"), py::arg("cuda_stream") = 0)
        .def(
            "build_fem_stretch_from_current_positions",
            [](Sim& s, float k, std::uintptr_t stream) { check(chysx_cloth_build_fem_stretch(s.p(), k, stream)); },
            py::arg("stiffness"), py::arg("cuda_stream") = 0)
        .def("num_fem_stretch_triangles",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_fem_stretch(s.p(), &n));
                 return n;
             })
        .def(
            "build_fem_shear_from_current_positions",
            [](Sim& s, float k, std::uintptr_t stream) { check(chysx_cloth_build_fem_shear(s.p(), k, stream)); },
            py::arg("stiffness"), py::arg("cuda_stream") = 0,
            "One Baraff-Witkin shear element per face: the stretch element with its material axes rotated by "
            "pi/4 (per-element weight area * stiffness).")
        .def("num_fem_shear_triangles",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_fem_shear(s.p(), &n));
                 return n;
             })
        .def(
            "build_bending_from_current_positions",
            [](Sim& s, float k, std::uintptr_t stream) { check(chysx_cloth_build_bending(s.p(), k, stream)); },
            py::arg("stiffness"), py::arg("cuda_stream") = 0,
            "One dihedral bending element per interior edge, rest angles from the current positions.")
        .def("num_bending_dihedrals",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_bending(s.p(), &n));
                 return n;
             })
        .def(
            "add_static_plane",
            [](Sim& s, FltArr n, float d) {
                if (n.ndim() != 1 || n.shape(0) != 3)
                    throw py::value_error("ClothSimulator.add_static_plane: n must be a (3,) float32 vector");
                check(chysx_cloth_add_static_plane(s.p(), n.data(), d));
            },
            py::arg("n"), py::arg("d"),
            "Static plane dot(n, x) + d == 0; n points into the half-space the cloth stays in.")
        .def(
            "add_static_box",
            [](Sim& s, FltArr c, FltArr h, FltArr ex, FltArr ey, FltArr ez) {
                for (auto* a : {&c, &h, &ex, &ey, &ez})
                    if (a->ndim() != 1 || a->shape(0) != 3)
                        throw py::value_error("ClothSimulator.add_static_box: every argument must be a (3,) vector");
                check(chysx_cloth_add_static_box(s.p(), c.data(), h.data(), ex.data(), ey.data(), ez.data()));
            },
            py::arg("center"), py::arg("half_ext"), py::arg("ex"), py::arg("ey"), py::arg("ez"),
            "Static oriented box: unit columns (ex, ey, ez), half extents along them.")
        .def("clear_static_shapes", [](Sim& s) { check(chysx_cloth_clear_static_shapes(s.p())); })
        .def("static_plane_count",
             [](const Sim& s) {
                 int p = 0;
                 check(chysx_cloth_static_shape_counts(s.p(), &p, nullptr));
                 return p;
             })
        .def("static_box_count",
             [](const Sim& s) {
                 int b = 0;
                 check(chysx_cloth_static_shape_counts(s.p(), nullptr, &b));
                 return b;
             })
        .def(
            "set_static_contact_thickness", [](Sim& s, float v) { check(chysx_cloth_set_static_contact(s.p(), 0, v)); },
            py::arg("thickness"))
        .def("static_contact_thickness",
             [](const Sim& s) {
                 float v = 0.0f;
                 check(chysx_cloth_get_static_contact(s.p(), 0, &v));
                 return v;
             })
        .def(
            "set_static_contact_stiffness", [](Sim& s, float v) { check(chysx_cloth_set_static_contact(s.p(), 1, v)); },
            py::arg("stiffness"))
        .def("static_contact_stiffness",
             [](const Sim& s) {
                 float v = 0.0f;
                 check(chysx_cloth_get_static_contact(s.p(), 1, &v));
                 return v;
             })
        .def(
            "set_static_contact_friction", [](Sim& s, float v) { check(chysx_cloth_set_static_contact(s.p(), 2, v)); },
            py::arg("mu"))
        .def("static_contact_friction",
             [](const Sim& s) {
                 float v = 0.0f;
                 check(chysx_cloth_get_static_contact(s.p(), 2, &v));
                 return v;
             })
        .def(
            "set_static_contact_friction_epsilon", [](Sim& s, float v) { check(chysx_cloth_set_static_contact(s.p(), 3, v)); },
            py::arg("eps_u"))
        .def("static_contact_friction_epsilon",
             [](const Sim& s) {
                 float v = 0.0f;
                 check(chysx_cloth_get_static_contact(s.p(), 3, &v));
                 return v;
             })
        .def("bending_dihedrals",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_bending(s.p(), &n));
                 py::array_t<int32_t> d({n, 4});
                 py::array_t<float> r({n});
                 check(chysx_cloth_get_dihedrals(s.p(), d.mutable_data(), r.mutable_data()));
                 return py::make_tuple(d, r);
             },
             "(dihedrals (D, 4), rest angles (D,)) of the installed bending elements.")
        .def(
            "set_pcg_iterations", [](Sim& s, int it) { check(chysx_cloth_set_pcg_iterations(s.p(), it)); },
            py::arg("max_iter"), "Set the maximum number of PCG iterations per step.")
        .def("pcg_iterations",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_pcg_iterations(s.p(), &n));
                 return n;
             })
        .def("num_off_diag_blocks",
             [](const Sim& s) {
                 int nnz = 0;
                 check(chysx_cloth_num_off_diag_blocks(s.p(), &nnz));
                 return nnz;
             })
        .def(
            "debug_dump_last_solve",
            [](const Sim& s) {
                int n = 0, nnz = 0;
                check(chysx_cloth_num_particles(s.p(), &n));
                check(chysx_cloth_num_off_diag_blocks(s.p(), &nnz));
                if (n == 0) nnz = 0;
                py::array_t<float> diag({n, 3, 3});
                py::array_t<int32_t> ro({n + 1});
                py::array_t<int32_t> cols({nnz});
                py::array_t<float> vals({nnz, 3, 3});
                py::array_t<float> rhs({n, 3});
                py::array_t<float> dx({n, 3});
                if (n == 0) return py::make_tuple(diag, ro, cols, vals, rhs, dx);
                check(chysx_cloth_debug_copy_hessian_diag(s.p(), diag.mutable_data()));
                check(chysx_cloth_debug_copy_hessian_csr(s.p(), ro.mutable_data(),
                                                         nnz > 0 ? cols.mutable_data() : nullptr,
                                                         nnz > 0 ? vals.mutable_data() : nullptr));
                check(chysx_cloth_debug_copy_last_rhs(s.p(), rhs.mutable_data()));
                check(chysx_cloth_debug_copy_last_dx(s.p(), dx.mutable_data()));
                return py::make_tuple(diag, ro, cols, vals, rhs, dx);
            },
            "Return (diag, row_offsets, col_indices, values, rhs, dx) of the last solve.")
        .def_property_readonly(
            "material",
            [](Sim& s) {
                Material v;
                v.owner = s.h;
                return v;
            },
            "Live view of the simulator's material (writes go to the simulator).")
        // ---- extensions --------------------------------------------------
        .def(
            "set_pcg_tolerance", [](Sim& s, double tol) { check(chysx_cloth_set_pcg_tolerance(s.p(), tol)); },
            py::arg("tol"),
            "Stop PCG at true relative residual <= tol (capped at pcg_iterations()); 0 = fixed iteration count.")
        .def(
            "set_springs_active", [](Sim& s, bool on) { check(chysx_cloth_set_springs_active(s.p(), on ? 1 : 0)); },
            py::arg("active"))
        .def(
            "set_external_force_active",
            [](Sim& s, bool on) { check(chysx_cloth_set_external_force_active(s.p(), on ? 1 : 0)); },
            py::arg("active"))
        .def(
            "set_spring_stiffnesses",
            [](Sim& s, FltArr k) {
                if (k.ndim() != 1) throw py::value_error("set_spring_stiffnesses: k must be 1-D");
                check(chysx_cloth_set_spring_stiffnesses(s.p(), k.data(), int(k.shape(0))));
            },
            py::arg("k"))
        .def(
            "set_fem_stretch_stiffnesses",
            [](Sim& s, FltArr k) {
                if (k.ndim() != 1) throw py::value_error("set_fem_stretch_stiffnesses: k must be 1-D");
                check(chysx_cloth_set_fem_stretch_stiffnesses(s.p(), k.data(), int(k.shape(0))));
            },
            py::arg("k"))
        .def("last_solve_stats",
             [](const Sim& s) {
                 int it = 0, rs = 0, st = 0;
                 double rr = 0.0;
                 check(chysx_cloth_last_solve_stats(s.p(), &it, &rs, &st, &rr));
                 return stats_dict(it, rs, st, rr);
             })
        .def(
            "set_warm_start",
            [](Sim& s, DblArr dx) {
                if (dx.ndim() != 2 || dx.shape(1) != 3) throw py::value_error("set_warm_start: dx must be (N, 3)");
                check(chysx_cloth_set_warm_start(s.p(), dx.data(), int(dx.shape(0))));
            },
            py::arg("dx"))
        .def("zero_warm_start", [](Sim& s) { check(chysx_cloth_zero_warm_start(s.p())); })
        .def("get_dx64",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_particles(s.p(), &n));
                 py::array_t<double> out({n, 3});
                 if (n) check(chysx_cloth_get_dx64(s.p(), out.mutable_data(), n));
                 return out;
             })
        .def("edges",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_num_edges(s.p(), &n));
                 py::array_t<int32_t> out({n, 2});
                 if (n) check(chysx_cloth_get_edges(s.p(), out.mutable_data()));
                 return out;
             })
        .def(
            "set_pin_projection", [](Sim& s, bool on) { check(chysx_cloth_set_pin_projection(s.p(), on ? 1 : 0)); },
            py::arg("on"),
            "True: pins are constraints (each pinned particle ends every step exactly at its target); False "
            "(default): the reference's penalty pins.")
        .def("pin_projection",
             [](const Sim& s) {
                 int on = 0;
                 check(chysx_cloth_pin_projection(s.p(), &on));
                 return on != 0;
             })
        .def(
            "set_blocking_step", [](Sim& s, bool on) { check(chysx_cloth_set_blocking_step(s.p(), on ? 1 : 0)); },
            py::arg("on"),
            "True: a fixed-iteration step (tolerance 0) waits until its solve has ended; False (default): it "
            "returns once its work is enqueued on the stream.")
        .def(
            "spmv",
            [](Sim& s, std::uintptr_t x, std::uintptr_t y, std::uintptr_t stream) {
                check(chysx_cloth_spmv(s.p(), x, y, stream));
            },
            py::arg("x_ptr"), py::arg("y_ptr"), py::arg("cuda_stream") = 0,
            "y = A x on the last assembled system (general storage only); device float3 pointers.")
        .def(
            "time_kernel",
            [](Sim& s, int which, int reps, float dt, std::uintptr_t stream) {
                double ms = 0.0;
                check(chysx_cloth_time_kernel(s.p(), which, reps, dt, stream, &ms));
                return ms;
            },
            py::arg("which"), py::arg("reps"), py::arg("dt"), py::arg("cuda_stream") = 0)
        .def(
            "set_autotune", [](Sim& s, bool on) { check(chysx_cloth_set_autotune(s.p(), on ? 1 : 0)); },
            py::arg("on"))
        .def("launch_config",
             [](const Sim& s) {
                 int c[CHYSX_LAUNCH_CONFIG_LEN];
                 check(chysx_cloth_launch_config(s.p(), c));
                 py::dict d;
                 d["assemble_waves"] = c[0];
                 d["residual_waves"] = c[1];
                 d["spmv_waves"] = c[2];
                 d["spmv_unroll"] = c[3];
                 d["update_waves"] = c[4];
                 d["finalize_waves"] = c[5];
                 d["spmv_branch_free"] = c[6];
                 d["spmv_symmetric"] = c[7];
                 d["spmv_packed"] = c[8];
                 d["spmv_index16"] = c[9];
                 d["pcg_split_advance"] = c[10];
                 return d;
             })
        .def(
            "set_launch_config",
            [](Sim& s, py::dict d) {
                int c[CHYSX_LAUNCH_CONFIG_LEN];
                check(chysx_cloth_launch_config(s.p(), c));
                const char* keys[CHYSX_LAUNCH_CONFIG_LEN] = {
                    "assemble_waves", "residual_waves",   "spmv_waves",     "spmv_unroll", "update_waves",
                    "finalize_waves", "spmv_branch_free", "spmv_symmetric", "spmv_packed", "spmv_index16",
                    "pcg_split_advance"};
                for (int i = 0; i < CHYSX_LAUNCH_CONFIG_LEN; ++i)
                    if (d.contains(keys[i])) c[i] = d[keys[i]].cast<int>();
                check(chysx_cloth_set_launch_config(s.p(), c));
            },
            py::arg("config"))
        .def("index_bits",
             [](const Sim& s) {
                 int b = 0;
                 check(chysx_cloth_index_bits(s.p(), &b));
                 return b;
             })
        .def("block_floats",
             [](const Sim& s) {
                 int f = 0;
                 check(chysx_cloth_block_floats(s.p(), &f));
                 return f;
             })
        .def("tune_log",
             [](const Sim& s) {
                 int n = 0;
                 check(chysx_cloth_tune_log(s.p(), &n, nullptr, nullptr));
                 std::vector<int> kwu(static_cast<size_t>(3 * n));
                 std::vector<double> ms(static_cast<size_t>(n));
                 check(chysx_cloth_tune_log(s.p(), &n, kwu.data(), ms.data()));
                 py::list out;
                 static const char* const kinds[] = {"spmv", "update", "spmv_branch_free", "assemble",
                                                     "iteration_advance_in_pass", "iteration_advance_kernel"};
                 for (int i = 0; i < n; ++i) {
                     const int k = kwu[3 * i];
                     out.append(py::make_tuple(k >= 0 && k < 6 ? kinds[k] : "unknown", kwu[3 * i + 1],
                                               kwu[3 * i + 2], ms[i]));
                 }
                 return out;
             })
        .def(
            "set_warm_start_mode", [](Sim& s, const std::string& mode) {
                check(chysx_cloth_set_warm_start_mode(s.p(), ws_mode(mode)));
            },
            py::arg("mode"),
            "'previous' (reference: previous step's dx), 'learned' (dx_prev plus the network's line-searched "
            "direction), 'projection' (Galerkin projection onto the last K solutions) or 'projection_learned'.")
        .def(
            "ws_set_history", [](Sim& s, int k) { check(chysx_cloth_ws_set_history(s.p(), k)); }, py::arg("k"),
            "Number of previous solutions the projection uses (2..8).")
        .def(
            "ws_set_params",
            [](Sim& s, FltArr theta) { check(chysx_cloth_ws_set_params(s.p(), theta.data(), int(theta.size()))); },
            py::arg("theta"))
        .def("ws_get_params",
             [](Sim& s) {
                 int n = 0;
                 check(chysx_ws_num_params(&n));
                 auto out = params_out(n);
                 check(chysx_cloth_ws_get_params(s.p(), out.mutable_data(), n));
                 return out;
             })
        .def(
            "ws_init_params", [](Sim& s, uint64_t seed) { check(chysx_cloth_ws_init_params(s.p(), seed)); },
            py::arg("seed") = 0)
        .def(
            "ws_set_training",
            [](Sim& s, bool on, float lr) { check(chysx_cloth_ws_set_training(s.p(), on ? 1 : 0, lr)); },
            py::arg("on"), py::arg("lr") = 1e-3f)
        .def(
            "ws_set_acceptance",
            [](Sim& s, const std::string& rule) {
                if (rule != "energy" && rule != "residual")
                    throw py::value_error("ws_set_acceptance: 'energy' or 'residual'");
                check(chysx_cloth_ws_set_acceptance(s.p(), rule == "residual" ? 1 : 0));
            },
            py::arg("rule"))
        .def("ws_stats",
             [](Sim& s) {
                 double o[CHYSX_WS_STATS_LEN];
                 check(chysx_cloth_ws_stats(s.p(), o, CHYSX_WS_STATS_LEN));
                 return ws_stats_dict(o);
             })
        .def("spmv_bytes", [](const Sim& s) {
            double b = 0.0;
            check(chysx_cloth_spmv_bytes(s.p(), &b));
            return b;
        });

    py::class_<Batch>(m, "ClothBatch", "Many cloths advanced together in one launch per kernel.")
        .def(py::init<>())
        .def(
            "add_cloth",
            [](Batch& b, IntArr tris, int n, py::object spring_k, py::object stretch_k, IntArr pins, FltArr targets,
               float pin_k) {
                if (tris.ndim() != 2 || tris.shape(1) != 3) throw py::value_error("add_cloth: triangles must be (M, 3)");
                if (pins.ndim() != 1 || targets.ndim() != 2 || (targets.shape(0) > 0 && targets.shape(1) != 3) ||
                    targets.shape(0) != pins.shape(0))
                    throw py::value_error("add_cloth: pins (P,) and targets (P, 3) required");
                float ks = 0.0f, kt = 0.0f;
                FltArr ka, kb;
                const float* kap = nullptr;
                const float* kbp = nullptr;
                if (py::isinstance<py::float_>(spring_k) || py::isinstance<py::int_>(spring_k)) ks = spring_k.cast<float>();
                else {
                    ka = spring_k.cast<FltArr>();
                    kap = ka.data();
                }
                if (py::isinstance<py::float_>(stretch_k) || py::isinstance<py::int_>(stretch_k))
                    kt = stretch_k.cast<float>();
                else {
                    kb = stretch_k.cast<FltArr>();
                    kbp = kb.data();
                }
                int idx = -1;
                check(chysx_batch_add_cloth(b.h, tris.data(), int(tris.shape(0)), n, kap, ks, kbp, kt, pins.data(),
                                            targets.data(), int(pins.shape(0)), pin_k, &idx));
                return idx;
            },
            py::arg("triangles"), py::arg("particle_count"), py::arg("spring_k"), py::arg("stretch_k"),
            py::arg("pin_indices"), py::arg("pin_targets"), py::arg("pin_stiffness") = 1.0e6f)
        .def(
            "set_external_buffers",
            [](Batch& b, std::uintptr_t pos, std::uintptr_t vel, int total, std::uintptr_t inv_mass,
               std::uintptr_t force) { check(chysx_batch_set_external_buffers(b.h, pos, vel, total, inv_mass, force)); },
            py::arg("pos_ptr"), py::arg("vel_ptr"), py::arg("particle_count"), py::arg("inv_mass_ptr") = 0,
            py::arg("force_ptr") = 0)
        .def(
            "set_material",
            [](Batch& b, int k, const Material& mat) {
                const chysx_cloth_material c = mat.get();
                check(chysx_batch_set_material(b.h, k, &c));
            },
            py::arg("index"), py::arg("material"))
        .def(
            "set_options",
            [](Batch& b, double tol, int max_iter, bool springs, bool force) {
                check(chysx_batch_set_options(b.h, tol, max_iter, springs ? 1 : 0, force ? 1 : 0));
            },
            py::arg("tol"), py::arg("max_iter"), py::arg("springs_active") = true, py::arg("force_active") = false)
        .def(
            "set_autotune", [](Batch& b, bool on) { check(chysx_batch_set_autotune(b.h, on ? 1 : 0)); },
            py::arg("on"))
        .def(
            "set_warm_start_mode", [](Batch& b, const std::string& mode) {
                check(chysx_batch_set_warm_start_mode(b.h, ws_mode(mode)));
            },
            py::arg("mode"))
        .def(
            "ws_set_history", [](Batch& b, int k) { check(chysx_batch_ws_set_history(b.h, k)); }, py::arg("k"))
        .def(
            "ws_set_params",
            [](Batch& b, FltArr theta) { check(chysx_batch_ws_set_params(b.h, theta.data(), int(theta.size()))); },
            py::arg("theta"))
        .def("ws_get_params",
             [](Batch& b) {
                 int n = 0;
                 check(chysx_ws_num_params(&n));
                 auto out = params_out(n);
                 check(chysx_batch_ws_get_params(b.h, out.mutable_data(), n));
                 return out;
             })
        .def(
            "ws_init_params", [](Batch& b, uint64_t seed) { check(chysx_batch_ws_init_params(b.h, seed)); },
            py::arg("seed") = 0)
        .def(
            "ws_set_training",
            [](Batch& b, bool on, float lr) { check(chysx_batch_ws_set_training(b.h, on ? 1 : 0, lr)); },
            py::arg("on"), py::arg("lr") = 1e-3f)
        .def(
            "ws_set_acceptance",
            [](Batch& b, const std::string& rule) {
                if (rule != "energy" && rule != "residual")
                    throw py::value_error("ws_set_acceptance: 'energy' or 'residual'");
                check(chysx_batch_ws_set_acceptance(b.h, rule == "residual" ? 1 : 0));
            },
            py::arg("rule"))
        .def(
            "ws_stats",
            [](Batch& b, int k) {
                double o[CHYSX_WS_STATS_LEN];
                check(chysx_batch_ws_stats(b.h, k, o, CHYSX_WS_STATS_LEN));
                return ws_stats_dict(o);
            },
            py::arg("index"))
        .def(
            "step", [](Batch& b, float dt, std::uintptr_t stream) { check(chysx_batch_step(b.h, dt, stream)); },
            py::arg("dt"), py::arg("cuda_stream") = 0)
        .def(
            "set_blocking_step", [](Batch& b, bool on) { check(chysx_batch_set_blocking_step(b.h, on ? 1 : 0)); },
            py::arg("on"))
        .def(
            "set_pin_projection",
            [](Batch& b, bool on) { check(chysx_batch_set_pin_projection(b.h, on ? 1 : 0)); }, py::arg("on"))
        .def(
            "set_split_advance", [](Batch& b, int mode) { check(chysx_batch_set_split_advance(b.h, mode)); },
            py::arg("mode"))
        .def("num_cloths",
             [](const Batch& b) {
                 int n = 0;
                 check(chysx_batch_num_cloths(b.h, &n));
                 return n;
             })
        .def(
            "stats",
            [](const Batch& b, int k) {
                int it = 0, rs = 0, st = 0;
                double rr = 0.0;
                check(chysx_batch_stats(b.h, k, &it, &rs, &st, &rr));
                return stats_dict(it, rs, st, rr);
            },
            py::arg("index"))
        .def(
            "stats_all",
            [](const Batch& b) {
                int n = 0;
                check(chysx_batch_num_cloths(b.h, &n));
                py::array_t<int> it(n), rs(n), st(n);
                py::array_t<double> rr(n);
                py::array_t<bool> cv(n);
                check(chysx_batch_stats_all(b.h, n, it.mutable_data(), rs.mutable_data(), st.mutable_data(),
                                            rr.mutable_data()));
                for (int k = 0; k < n; ++k) cv.mutable_data()[k] = st.data()[k] == CHYSX_PCG_CONVERGED;
                py::dict d;
                d["iters"] = it;
                d["restarts"] = rs;
                d["status"] = st;
                d["converged"] = cv;
                d["relres"] = rr;
                return d;
            },
            "Per-cloth iterations, restarts, status and relres of the last step (arrays).");

    m.def(
        "topology_query",
        [](int n, IntArr tris, bool include_springs, IntArr pins) {
            if (tris.ndim() != 2 || tris.shape(1) != 3) throw py::value_error("topology_query: triangles must be (M, 3)");
            const int T = int(tris.shape(0));
            const int P = int(pins.size());
            int nnz = 0, ne = 0;
            check(chysx_topology_query(n, tris.data(), T, include_springs ? 1 : 0, pins.data(), P, &nnz, &ne, nullptr,
                                       nullptr, nullptr, nullptr, nullptr, nullptr));
            py::array_t<int32_t> ro({n + 1}), cols({nnz}), edges({ne, 2}), ts({9 * T}),
                ss({include_springs ? 4 * ne : 0}), ps({P});
            check(chysx_topology_query(n, tris.data(), T, include_springs ? 1 : 0, pins.data(), P, &nnz, &ne,
                                       ro.mutable_data(), cols.mutable_data(), edges.mutable_data(), ts.mutable_data(),
                                       ss.mutable_data(), ps.mutable_data()));
            py::dict d;
            d["row_offsets"] = ro;
            d["col_indices"] = cols;
            d["edges"] = edges;
            d["tri_slots"] = ts;
            d["spring_slots"] = ss;
            d["pin_slots"] = ps;
            return d;
        },
        py::arg("n"), py::arg("triangles"), py::arg("include_springs") = false,
        py::arg("pins") = IntArr(0), "Host-only pattern and slot maps of a mesh (no GPU needed).");
    m.def(
        "topology_query_bending",
        [](int n, IntArr tris) {
            if (tris.ndim() != 2 || tris.shape(1) != 3) throw py::value_error("topology_query: triangles must be (M, 3)");
            const int T = int(tris.shape(0));
            int nnz = 0, nd = 0;
            check(chysx_topology_query_bending(n, tris.data(), T, &nnz, &nd, nullptr, nullptr, nullptr, nullptr));
            py::array_t<int32_t> ro({n + 1}), cols({nnz}), dih({nd, 4}), ds({nd, 16});
            check(chysx_topology_query_bending(n, tris.data(), T, &nnz, &nd, ro.mutable_data(), cols.mutable_data(),
                                               dih.mutable_data(), ds.mutable_data()));
            py::dict d;
            d["row_offsets"] = ro;
            d["col_indices"] = cols;
            d["dihedrals"] = dih;
            d["dihedral_slots"] = ds;
            return d;
        },
        py::arg("n"), py::arg("triangles"),
        "Host-only pattern with FEM triangles and bending dihedrals, dihedral list and slots (no GPU needed).");
    m.def(
        "ws_small_solve",
        [](DblArr G, DblArr g) {
            const int k = int(g.size());
            if (G.ndim() != 2 || G.shape(0) != k || G.shape(1) != k) throw py::value_error("G must be (k, k)");
            py::array_t<double> c({k});
            int kept = 0;
            check(chysx_ws_small_solve(k, G.data(), g.data(), c.mutable_data(), &kept));
            return py::make_tuple(c, kept);
        },
        py::arg("G"), py::arg("g"));
    m.def(
        "stream_bench",
        [](int kind, std::uint64_t bytes, int reps, std::uintptr_t stream) {
            double ms = 0.0, moved = 0.0;
            check(chysx_stream_bench(kind, bytes, reps, stream, &ms, &moved));
            return py::make_tuple(ms, moved);
        },
        py::arg("kind"), py::arg("bytes_per_buffer"), py::arg("reps"), py::arg("cuda_stream") = 0,
        "HBM probe: kind 0 copy, 1 triad, 2 / 3 non-temporal; returns (ms per launch, bytes per launch).");
    m.def("ws_num_params", []() {
        int n = 0;
        check(chysx_ws_num_params(&n));
        return n;
    });
    m.def(
        "wsnet_forward",
        [](FltArr theta, IntArr ro, IntArr cols, FltArr F, FltArr E, bool direction_only) {
            const int n = int(ro.size()) - 1;
            if (F.ndim() != 2 || F.shape(0) != n || F.shape(1) != 32 || E.ndim() != 2 || E.shape(0) != n ||
                E.shape(1) != 32)
                throw py::value_error("wsnet_forward: F and E must be (n, 32)");
            chysx_wsnet* net = nullptr;
            check(chysx_wsnet_create(&net));
            py::array_t<float> out({n, direction_only ? 3 : 32});
            const int rc1 = chysx_wsnet_set_params(net, theta.data(), int(theta.size()));
            const int rc2 = rc1 ? rc1
                                : direction_only ? chysx_wsnet_forward_direction(net, n, ro.data(), cols.data(),
                                                                                 F.data(), E.data(), out.mutable_data())
                                                 : chysx_wsnet_forward(net, n, ro.data(), cols.data(), F.data(),
                                                                       E.data(), out.mutable_data());
            chysx_wsnet_destroy(net);
            check(rc2);
            return out;
        },
        py::arg("theta"), py::arg("row_offsets"), py::arg("cols"), py::arg("F"), py::arg("E"),
        py::arg("direction_only") = false);
    m.def(
        "wsnet_loss_grad",
        [](FltArr theta, IntArr ro, IntArr cols, FltArr F, FltArr E, FltArr target) {
            const int n = int(ro.size()) - 1;
            if (F.ndim() != 2 || F.shape(0) != n || E.ndim() != 2 || E.shape(0) != n || target.ndim() != 2 ||
                target.shape(0) != n || target.shape(1) != 3)
                throw py::value_error("wsnet_loss_grad: F, E (n, 32) and target (n, 3) required");
            chysx_wsnet* net = nullptr;
            check(chysx_wsnet_create(&net));
            py::array_t<float> grad({int(theta.size())});
            double loss = 0.0;
            const int rc1 = chysx_wsnet_set_params(net, theta.data(), int(theta.size()));
            const int rc2 = rc1 ? rc1
                                : chysx_wsnet_loss_grad(net, n, ro.data(), cols.data(), F.data(), E.data(),
                                                        target.data(), &loss, grad.mutable_data());
            chysx_wsnet_destroy(net);
            check(rc2);
            return py::make_tuple(loss, grad);
        },
        py::arg("theta"), py::arg("row_offsets"), py::arg("cols"), py::arg("F"), py::arg("E"), py::arg("target"));
    m.attr("__version__") = chysx_version();
}
