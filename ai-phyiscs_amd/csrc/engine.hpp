// SPDX-License-Identifier: Apache-2.0
//
// Host engine of the implicit cloth step on gfx950.
//
// A Problem is one block-diagonal linear system made of one or more cloth
// instances laid end to end in a padded row space; every kernel launch covers
// all instances, each with its own FP64 PCG state and convergence status.
// ClothSim (one cloth, the reference's ClothSimulator) and ClothBatch (many
// cloths in one launch) are thin owners of Instances plus one Problem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "cloth_device.hpp"
#include "cloth_kernels.hpp"
#include "topology.hpp"
#include "warmstart.hpp"

namespace chysx_amd {

struct Material {  // same field order as ClothMaterial (cloth_material.h:20-43)
    float lame_mu = 0.0f;
    float lame_lambda = 0.0f;
    float bending = 0.0f;
    float density = 0.05f;
    float damping = 0.0f;
    float gx = 0.0f;
    float gy = 0.0f;
    float gz = -9.81f;
};

template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() { release(); }
    void resize(size_t n);          // contents undefined
    void assign(const T* host, size_t n, hipStream_t s = nullptr);
    void zero(hipStream_t s = nullptr);
    void release();
    T* data() const { return ptr_; }
    size_t size() const { return n_; }

private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// One cloth: mesh, element rest data, pins and solver options.
struct Instance {
    int n = 0;      // particles
    int part0 = 0;  // first particle in the external buffers
    std::vector<int> tris;   // 3 per triangle (local indices)
    std::vector<int> edges;  // 2 per unique edge (local), from tris
    // springs (one per edge); present once built
    bool has_springs = false;
    std::vector<float> spr_rest, spr_k;
    // stretch triangles (theta = 0), present once built
    bool has_stretch = false;
    std::vector<float> tri_dminv;  // 4 per triangle
    std::vector<float> tri_area, tri_k;
    // shear triangles (theta = pi/4): the same triangles, rotated frame
    bool has_shear = false;
    std::vector<float> sh_dminv;   // 4 per triangle
    std::vector<float> sh_area, sh_k;
    // bending dihedrals (v0, v1, v2, v3), rest angles, one stiffness
    bool has_bending = false;
    std::vector<int> dih;
    std::vector<float> dih_rest;
    float bend_k = 1.0e-3f;
    // static-shape contact (StaticContactSet): planes 4 floats (n, d), boxes
    // 16 floats (center, half extents, ex, ey, ez, 0)
    std::vector<float> planes, boxes;
    float contact_h = 0.0f, contact_k = 0.0f, contact_mu = 0.0f, contact_eps = 1.0e-4f;
    // pins
    std::vector<int> pin_idx;
    std::vector<float> pin_tgt;
    float pin_k = 1.0e6f;
    // pins as constraints (extension): displacement fixed at target - x_n
    bool pin_projection = false;
    // options
    bool springs_active = false;
    bool force_active = false;
    double tol = 0.0;
    int max_iter = 50;
    Material material;
};

// Workgroup shapes of the step's kernels (speed only; results do not depend
// on them).  The PCG passes (residual, SpMV, update) always run 16-wave
// workgroups, one per 16-slice group of inner-product partials: their waves
// fields are fixed at 16.
struct LaunchConfig {
    int assemble_waves = 4;
    int residual_waves = 16;
    int spmv_waves = 16;
    int spmv_unroll = 1;
    int spmv_branch_free = 0;  // 1: padding entries gather the row's own vector (zero blocks) instead of a branch
    // 1: the SpMV pass reads each lower block as the transpose of its upper
    // mirror (streams half the off-diagonal values; changes results at FP32
    // rounding level, so it is a fixed choice, never picked by the autotuner)
    int spmv_symmetric = 1;
    int update_waves = 16;
    int finalize_waves = 4;
    // 1: with spmv_symmetric and no bending elements (every block, diagonal
    // or not, then symmetric bitwise), blocks are stored as their upper
    // triangle (6 floats).  The operator is bitwise the same as with 9.
    int spmv_packed = 1;
    // 1: the SpMV pass reads 16-bit column and mirror offsets when every
    // offset of the mesh fits (speed only: the same indices).
    int spmv_index16 = 1;
    // 1: each PCG iteration advances the solver state in a small kernel of
    // its own (one workgroup per instance) before the SpMV pass, instead of
    // in every SpMV workgroup.  The in-pass advance costs each generation of
    // SpMV workgroups its start-up, so the kernel of its own wins on large
    // problems (speed only: the same sums, in the same order, and the same
    // scalar step).  -1: chosen from the problem's size at build time.
    int pcg_split_advance = -1;
};

struct TuneSample {
    // 0: SpMV pass, 1: update pass, 2: branch-free SpMV pass, 3: assembly
    // (waves per workgroup), 4 / 5: a whole PCG iteration with the state
    // advanced in the SpMV pass / in a kernel of its own
    int kernel;
    int waves;
    int unroll;
    double ms;
};

struct SolveStats {
    int iters = 0;
    int restarts = 0;
    int status = 0;
    double relres = 0.0;
    // warm start of this solve (zeros when no warm-start stage ran)
    int ws_proj_used = 0;     // 1: the history projection replaced dx_prev
    int ws_used = 0;          // 1: the network's direction was applied
    int ws_kdim = 0;          // projection directions kept
    double ws_r0_prev = 0;    // |b - A dx_prev| / |b|
    double ws_r0_proj = 0;    // |b - A x_proj| / |b| (predicted)
    double ws_r0_learned = 0; // |b - A (x_b + alpha d)| / |b| (predicted)
};

// PCG initial guess.
enum WarmStartMode : int {
    kWarmPrevious = 0,        // the previous step's dx (reference behaviour)
    kWarmLearned = 1,         // dx_prev plus the network's line-searched direction
    kWarmProjection = 2,      // Galerkin projection onto the last K solutions
    kWarmProjLearned = 3,     // the projection plus the network's direction
};

class Problem {
public:
    Problem();
    ~Problem();
    Problem(const Problem&) = delete;
    Problem& operator=(const Problem&) = delete;

    // Build topology, layout and device data for `insts` (all must outlive the
    // call).  Keeps the FP64 iterate when the padded row count is unchanged.
    void build(const std::vector<const Instance*>& insts);
    bool built() const { return built_; }
    // Upload per-instance parameters / options (cheap; call before a step
    // when options changed).
    void upload_params(const std::vector<const Instance*>& insts, hipStream_t s);
    void update_pin_targets(const Instance& inst, int inst_index, hipStream_t s);

    // One step over all instances.  With a tolerance on any instance, or a
    // warm-start stage, it polls the solver states between chunks of
    // iterations, so it returns once every solve has ended.  When every
    // instance runs a fixed iteration count (tol == 0, the reference's mode)
    // the count is known: the step enqueues it all on `s` and returns without
    // waiting (unless set_blocking_step(true)); stats() then waits for it.
    void step(float dt, const DevExt& ext, hipStream_t s);
    void set_blocking_step(bool on) { blocking_ = on; }
    bool blocking_step() const { return blocking_; }
    // True when the last step was enqueued without waiting and its results
    // have not been read yet.
    bool step_pending() const { return stats_pending_; }
    // Orders stream `s` after an enqueued step (no-op otherwise); every
    // writer of the problem's device buffers calls it first.
    void wait_pending(hipStream_t s) const;

    // Per-instance results of the last step (waits for an enqueued step); the
    // reference stays valid until the next step.
    const std::vector<SolveStats>& stats() const {
        resolve_stats();
        return stats_;
    }
    int n_rows_padded() const { return n_pad_; }
    int ninst() const { return ninst_; }
    int nslices() const { return int(slices_h_.size()); }

    // Reference-layout views of instance `k` (host copies).
    const Pattern& pattern(int k) const { return pats_[size_t(k)]; }
    const std::vector<int>& tri_slots(int k) const { return tri_slots_[size_t(k)]; }
    const std::vector<int>& spr_slots(int k) const { return spr_slots_[size_t(k)]; }
    const std::vector<int>& pin_slots(int k) const { return pin_slots_[size_t(k)]; }
    void copy_diag(int k, float* out) const;        // (n, 3, 3)
    void copy_values(int k, float* out) const;      // (nnz, 3, 3) in CSR order
    void copy_rhs(int k, float* out) const;         // (n, 3)
    void copy_dx(int k, float* out) const;          // (n, 3) rounded to FP32
    void copy_dx64(int k, double* out) const;       // (n, 3)
    void set_dx64(int k, const double* in, hipStream_t s);
    void zero_dx(hipStream_t s);

    // Kernel-level access for benchmarks.
    LaunchGeom geom() const {
        return LaunchGeom{slices_.data(), int(slices_h_.size()), int(slices_h_.size()) / kGroupSlices,
                          ninst_ > 1 ? gactive_.data() : nullptr};
    }
    DevMatrix matrix() const;
    // Floats per stored off-diagonal block: 6 (symmetric, packed) or 9.
    int block_floats() const;
    // Bits of the column / mirror offsets the SpMV pass reads: 16 or 32.
    int index_bits() const { return (cfg_.spmv_index16 && col16_.size() > 0) ? 16 : 32; }
    DevVectors vectors() const;
    DevMesh mesh() const;
    // The solver state; a second buffer carries it between the passes of a
    // chunk (iteration i reads buffer i % 2 and writes the other; a chunk has
    // an even number of iterations, so it ends in states()).
    InstState* states() const { return state_.data(); }
    InstState* states_b() const { return state_b_.data(); }
    const InstParams* params() const { return params_.data(); }
    // Bytes of the algorithmic SpMV traffic (reference formulation):
    // 40 per off-diagonal block + 64 per row + 4.
    double spmv_bytes() const;
    long long nnz_total() const { return nnz_total_; }
    int max_width() const { return max_width_; }
    int rows_total() const { return n_rows_; }

    // Launch pieces (used by step and by the benchmark).
    void launch_assemble(float dt, const DevExt& ext, hipStream_t s);
    void launch_pcg_start(hipStream_t s);
    // Iteration i of a chunk: SpMV pass (advancing the state from buffer i % 2
    // into the other) and update pass.
    void launch_pcg_iteration(hipStream_t s, int i);
    // kChunkIters iterations and the closing advance, launched directly.
    void launch_pcg_chunk_body(hipStream_t s);
    // `count` chunks as replayed graphs of 8, 4, 2 and 1 chunks (largest
    // first): each graph launch costs a gap of several microseconds, the
    // chunks inside one graph follow each other directly.
    void launch_pcg_chunks(hipStream_t s, int count);
    void launch_finalize(float dt, const DevExt& ext, hipStream_t s);
    void read_states(hipStream_t s);        // blocking
    bool all_done() const;

    // ---- warm start ----------------------------------------------------------
    void set_warm_start_mode(int mode);
    int warm_start_mode() const { return ws_mode_; }
    // Number of previous solutions the projection uses (2..8, default 6).
    void set_ws_history(int k);
    int ws_history() const { return hist_k_; }
    // Online training: every step runs the network; its direction is applied
    // only in the learned mode; after the solve the line-search loss against
    // the converged dx gives one Adam step.
    void set_ws_training(bool on, float lr) { ws_train_ = on; ws_lr_ = lr; }
    // Acceptance of the network's direction: false (default) when it lowers
    // the energy (always, unless alpha == 0); true only when it also lowers
    // the predicted residual.
    void set_ws_accept_residual(bool on) { ws_accept_residual_ = on; }
    bool ws_accept_residual() const { return ws_accept_residual_; }
    WarmStartNet& net();
    double last_ws_loss() const { return ws_loss_; }

    static constexpr int kChunkIters = 8;
    // Longest fixed-count solve captured with its whole step in one graph
    // (64 chunks: 512 iterations, about 1,600 kernel nodes).
    static constexpr int kStepGraphMaxChunks = 64;
    static constexpr int kAutotuneMinRows = 1 << 15;
    static constexpr int kAutotuneReps = 8;
    // Groups from which the solver state is advanced in a kernel of its own
    // (pcg_split_advance = -1).  Each generation of resident SpMV workgroups
    // repeats the in-pass advance's start-up; measured on MI355X the kernel of
    // its own is as fast at 1024 groups (1024^2) and faster beyond.
    static constexpr int kSplitAdvanceGroups = 512;
    bool split_advance() const {
        // batches keep the in-pass advance by default: at 512 cloths (3130
        // groups) it measured 1% faster than the split one, whose per-instance
        // workgroups also run for finished cloths
        if (cfg_.pcg_split_advance >= 0) return cfg_.pcg_split_advance > 0;
        if (split_tuned_ >= 0) return split_tuned_ > 0;
        return ninst_ == 1 && int(slices_h_.size()) / kGroupSlices >= kSplitAdvanceGroups;
    }
    void set_split_advance(int mode) {
        if (mode < -1 || mode > 1) throw std::invalid_argument("split advance: -1 (by size), 0 or 1");
        cfg_.pcg_split_advance = mode;
        split_tuned_ = -1;
        destroy_graph();
    }

    const LaunchConfig& launch_config() const { return cfg_; }
    void set_launch_config(const LaunchConfig& c);
    // Pick the fastest launch shapes for this problem (no-op for small
    // problems): the SpMV pass variant (entries per batch, branch-free
    // padding), the assembly's waves per workgroup, and, when the launch
    // config leaves it to the size rule (-1), where the PCG state is advanced
    // (timed over whole iterations).  None of these changes results.  The
    // PCG timings use scratch vectors and a scratch solver state; the assembly
    // timings assemble this step's system from `ext` (idempotent: the step's
    // own assembly writes the same values), so call it after upload_params().
    void autotune(hipStream_t s, float dt, const DevExt& ext);
    const std::vector<TuneSample>& tune_log() const { return tune_log_; }
    // Mark every instance of `st` as running (timing hooks only).
    void prime_states(InstState* st, hipStream_t s) const;

private:
    void capture_chunk_graph(int level);  // a graph of 2^level chunks
    // Copies host_states_ (and the warm-start records) into stats_.
    void fill_stats() const;
    // Waits for an enqueued step's states, checks that every solve ended, and
    // fills stats_.
    void resolve_stats() const;
    void destroy_graph();
    void warm_start_stage(float dt, const DevExt& ext, hipStream_t s);
    void push_history(hipStream_t s);
    void reset_history() { hist_n_ = 0; }
    WsGraph ws_graph() const;
    WsHist hist_ptrs() const;
    WsReduce ws_reduce() const;

    bool built_ = false;
    int ninst_ = 0;
    int n_rows_ = 0;  // real rows
    int n_pad_ = 0;   // padded rows
    long long nnz_total_ = 0;
    std::vector<Pattern> pats_;
    std::vector<std::vector<int>> tri_slots_, spr_slots_, pin_slots_;
    std::vector<int> inst_row0_, inst_n_, inst_slice0_, inst_nslices_;
    std::vector<SliceDesc> slices_h_;
    std::vector<int> slice_ptr_h_;
    std::vector<int> row_deg_h_;

    DeviceArray<SliceDesc> slices_;
    DeviceArray<InstState> state_, state_b_;
    DeviceArray<InstParams> params_;
    // What params_ and shapes_ hold on the device (empty: unknown), so an
    // unchanged upload is skipped with its stream synchronisation.
    std::vector<InstParams> params_h_;
    std::vector<float4> shapes_h_;
    // mesh
    DeviceArray<int> tri_v_, spr_v_, pin_v_;
    DeviceArray<float4> tri_dminv_, sh_dminv_;
    DeviceArray<float> tri_area_, tri_k_, sh_area_, sh_k_, spr_rest_, spr_k_, pin_tgt_, dih_rest_;
    DeviceArray<int4> dih_v_, bend_inc_;
    DeviceArray<float4> shapes_;
    DeviceArray<int> bend_inc_off_;
    std::vector<std::vector<int>> dih_slots_;
    int max_width_ = 0;
    DeviceArray<int> tri_inc_off_, spr_inc_off_, pin_inc_off_, pin_inc_;
    DeviceArray<int2> tri_inc_, spr_inc_;
    std::vector<int> pin_offset_;  // first global pin of each instance
    // matrix
    DeviceArray<int> slice_ptr_, col_, row_deg_, mir_;
    std::vector<int> mir_h_;
    DeviceArray<int16_t> col16_, mir16_;  // empty when an offset does not fit in 16 bits
    DeviceArray<float> val_, diag_, minv_;
    // vectors
    DeviceArray<float> rhs_, r_, z_, ap_, p0_, p1_;
    DeviceArray<double> x_, partials_;
    DeviceArray<int> gactive_;  // batches: 1 per group of a running instance

    InstState* host_states_ = nullptr;  // pinned
    mutable std::vector<SolveStats> stats_;
    int last_iters_ = 0;
    bool blocking_ = false;
    mutable bool stats_pending_ = false;  // host_states_ is being copied for an enqueued step
    hipEvent_t done_event_ = nullptr;     // recorded after that copy

    LaunchConfig cfg_;
    std::vector<TuneSample> tune_log_;
    int split_tuned_ = -1;  // the autotuner's split-advance choice (-1: none)
    // warm start
    int ws_mode_ = kWarmPrevious;
    bool ws_train_ = false;
    bool ws_accept_residual_ = false;
    float ws_lr_ = 1e-3f;
    double ws_loss_ = 0.0;
    bool ws_ran_ = false;    // network stage ran this step
    bool proj_ran_ = false;  // projection stage ran this step
    bool start_r_ready_ = false;  // the line search left the start residual in r
    int hist_k_ = 6;         // history length of the projection
    int hist_n_ = 0;         // solutions held (<= hist_k_)
    int hist_head_ = 0;      // slot of the newest
    int ws_chunks_ = 1;
    bool rest_captured_ = false;
    std::unique_ptr<WarmStartNet> net_;
    DeviceArray<float> hist_, rest_pos_, rb_, d_, ad_;
    DeviceArray<double> xb_, ws_part_, ws_part2_;
    DeviceArray<WsInst> ws_inst_;
    std::vector<WsInst> ws_host_;
    hipStream_t cap_stream_ = nullptr;
    static constexpr int kChunkGraphLevels = 4;  // graphs of 1, 2, 4 and 8 chunks
    hipGraphExec_t chunk_exec_[kChunkGraphLevels] = {};
    // The whole fixed-count step as one graph, and what it was captured with
    // (dt, caller buffers, chunk count, every device pointer and launch shape
    // its kernels take): a step whose key differs captures it again.
    hipGraphExec_t step_exec_ = nullptr;
    std::vector<unsigned char> step_key_;
    std::vector<unsigned char> fixed_step_key(float dt, const DevExt& ext, int chunks) const;
    void launch_fixed_step_graph(float dt, const DevExt& ext, int chunks, hipStream_t s);
};

// The reference's ClothSimulator behaviour on one instance.
class ClothSim {
public:
    ClothSim();
    void set_material(const Material& m) { inst_.material = m; params_dirty_ = true; }
    Material& material() { params_dirty_ = true; return inst_.material; }
    const Material& material() const { return inst_.material; }
    void set_external_buffers(std::uintptr_t pos, std::uintptr_t vel, int n, std::uintptr_t inv_mass,
                              std::uintptr_t force) noexcept;
    void set_mesh(const int* tris, int ntri);
    void build_springs(float k, std::uintptr_t stream);
    void build_fem_stretch(float k, std::uintptr_t stream);
    void set_spring_stiffnesses(const float* k, int n);
    void set_fem_stretch_stiffnesses(const float* k, int n);
    void set_pins(const int* idx, const float* targets, int n, float k);
    void clear_pins();
    void update_pin_targets(const float* targets, int n, std::uintptr_t stream);
    void redistribute_mass_area_weighted(float density, std::uintptr_t inv_mass, int n, std::uintptr_t stream);
    // The same without allocating device memory: the incidence is staged in
    // the library's fixed scratch array (noalloc.hpp).
    void redistribute_mass_area_weighted_noalloc(float density, std::uintptr_t inv_mass, int n, std::uintptr_t stream);
    void set_pcg_iterations(int it) { inst_.max_iter = it > 0 ? it : 1; params_dirty_ = true; }
    int pcg_iterations() const { return inst_.max_iter; }
    void set_pcg_tolerance(double tol);
    double pcg_tolerance() const { return inst_.tol; }
    void set_springs_active(bool on);
    void set_external_force_active(bool on) { inst_.force_active = on; params_dirty_ = true; }
    void set_pin_projection(bool on) { inst_.pin_projection = on; params_dirty_ = true; }
    bool pin_projection() const { return inst_.pin_projection; }
    void set_autotune(bool on) { autotune_ = on; }
    void step(float dt, std::uintptr_t stream);

    int num_springs() const { return inst_.has_springs ? int(inst_.edges.size() / 2) : 0; }
    int num_fem_stretch() const { return inst_.has_stretch ? int(inst_.tris.size() / 3) : 0; }
    int num_fem_shear() const { return inst_.has_shear ? int(inst_.tris.size() / 3) : 0; }
    int num_bending() const { return inst_.has_bending ? int(inst_.dih.size() / 4) : 0; }
    void build_fem_shear(float k, std::uintptr_t stream);
    void build_bending(float k, std::uintptr_t stream);
    // static-shape contact (cloth_simulator.h:398-450)
    void add_static_plane(const float* n, float d);
    void add_static_box(const float* center, const float* half_ext, const float* ex, const float* ey, const float* ez);
    void clear_static_shapes();
    Instance& contact_instance() { params_dirty_ = true; return inst_; }
    int num_pins() const { return int(inst_.pin_idx.size()); }
    int num_edges() const { return int(inst_.edges.size() / 2); }
    const std::vector<int>& edges() const { return inst_.edges; }
    int num_particles() const { return dumped_n_; }
    int num_off_diag_blocks() const;
    SolveStats last_stats() const;
    Problem& problem() { return prob_; }
    const Instance& instance() const { return inst_; }
    // Build topology now (normally done lazily by step()).
    void ensure_topology();
    void set_warm_start(const double* dx);  // FP64 (n,3)
    DevExt ext() const {
        return DevExt{reinterpret_cast<float*>(pos_), reinterpret_cast<float*>(vel_),
                      reinterpret_cast<const float*>(inv_mass_), reinterpret_cast<const float*>(force_)};
    }
    void zero_warm_start();

    void debug_copy_hessian_diag(float* out) const;
    void debug_copy_hessian_csr(int* ro, int* cols, float* vals) const;
    void debug_copy_last_rhs(float* out) const;
    void debug_copy_last_dx(float* out) const;

private:
    Instance inst_;
    Problem prob_;
    std::uintptr_t pos_ = 0, vel_ = 0, inv_mass_ = 0, force_ = 0;
    int n_ = 0;
    bool topo_dirty_ = true;
    bool params_dirty_ = true;
    bool autotune_ = true;
    bool tune_pending_ = false;  // topology rebuilt: tune at the next step
    int dumped_n_ = 0;
};

// Many cloths stepped together.  Particles of instance k live at
// [part0_k, part0_k + n_k) of one caller-owned position / velocity buffer.
class ClothBatch {
public:
    // Adds a cloth of n particles (the next n of the external buffers).
    // spring_k / stretch_k: one value per edge / triangle, or null with the
    // scalar.  Rest state is taken from the external positions at the first
    // step after the cloth is added.
    int add_cloth(const int* tris, int ntri, int n, const float* spring_k, float spring_k_scalar,
                  const float* stretch_k, float stretch_k_scalar, const int* pins, const float* pin_targets,
                  int npin, float pin_k);
    void set_external_buffers(std::uintptr_t pos, std::uintptr_t vel, int total, std::uintptr_t inv_mass,
                              std::uintptr_t force) noexcept;
    void set_material(int k, const Material& m);
    void set_options(double tol, int max_iter, bool springs_active, bool force_active);
    void set_pin_projection(bool on);
    void set_autotune(bool on) { autotune_ = on; }
    void step(float dt, std::uintptr_t stream);
    int count() const { return int(insts_.size()); }
    int total_particles() const { return total_; }
    SolveStats stats(int k) const;
    Problem& problem() { return prob_; }
    const Instance& instance(int k) const { return *insts_[size_t(k)]; }

private:
    std::vector<std::unique_ptr<Instance>> insts_;
    std::vector<std::vector<float>> pending_spring_k_, pending_stretch_k_;
    std::vector<char> rest_done_;
    Problem prob_;
    std::uintptr_t pos_ = 0, vel_ = 0, inv_mass_ = 0, force_ = 0;
    int total_ = 0;
    int ext_n_ = 0;
    bool dirty_ = true;
    bool autotune_ = true;
    bool tune_pending_ = false;  // rebuilt: tune at this step, after the parameters are uploaded
    bool params_dirty_ = true;
};

void check_hip(hipError_t err, const char* what);

// Rest lengths and rest shapes of `inst` from device positions `pos`
// (particle space; the instance's particles start at inst.part0).
void compute_rest_springs(Instance& inst, const float* pos, const float* k_per_edge, float k, hipStream_t s);
void compute_rest_stretch(Instance& inst, const float* pos, const float* k_per_tri, float k, hipStream_t s);
void compute_rest_shear(Instance& inst, const float* pos, const float* k_per_tri, float k, hipStream_t s);
void compute_rest_bending(Instance& inst, const float* pos, float k, hipStream_t s);

}  // namespace chysx_amd
