/* SPDX-License-Identifier: Apache-2.0
 *
 * C ABI of the gfx950 implicit cloth step (libchysx_amd.so).
 *
 * Plain pointers and sizes only.  Device buffers are passed as uintptr_t
 * addresses of caller-owned HIP allocations (12-byte float3 AoS for
 * positions, velocities and forces; float for inverse mass), exactly like the
 * reference's raw-pointer interface.  Streams are hipStream_t values cast to
 * uintptr_t; 0 is the null stream.
 *
 * Every function returns a status code; on failure chysx_last_error() holds a
 * thread-local message.  The codes mirror the reference's C++ exception types
 * (src/bindings.cpp re-raises them as Python exceptions):
 *   CHYSX_EINVAL  std::invalid_argument -> ValueError
 *   CHYSX_ERUNTIME std::runtime_error   -> RuntimeError
 *   CHYSX_ERANGE  std::out_of_range     -> IndexError
 *
 * Citations name the reference interface each entry point replaces (paths
 * relative to the reference repository root).
 */
#ifndef CHYSX_CLOTH_H
#define CHYSX_CLOTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CHYSX_OK = 0,
    CHYSX_EINVAL = 1,
    CHYSX_ERUNTIME = 2,
    CHYSX_ERANGE = 3
};

/* PCG status values reported by chysx_cloth_last_solve_stats. */
enum {
    CHYSX_PCG_RUNNING = 0,
    CHYSX_PCG_CHECK = 1,
    CHYSX_PCG_CONVERGED = 2,
    CHYSX_PCG_STOPPED = 3
};

/* Same fields and order as chysx::cloth::ClothMaterial
 * (src/cloth/cloth_material.h:20-43). */
typedef struct chysx_cloth_material {
    float lame_mu;
    float lame_lambda;
    float bending;
    float density;
    float damping;
    float gx;
    float gy;
    float gz;
} chysx_cloth_material;

typedef struct chysx_cloth chysx_cloth;

/* Thread-local message of the last failed call. */
const char* chysx_last_error(void);
/* Library version string ("0.1.0", chysx/__init__.py:51). */
const char* chysx_version(void);

/* ClothSimulator() (src/bindings.cpp:62; src/cloth/cloth_simulator.h:178-838). */
int chysx_cloth_create(chysx_cloth** out);
int chysx_cloth_destroy(chysx_cloth* sim);

/* set_material / material (src/bindings.cpp:63-65, 1069-1075). */
int chysx_cloth_set_material(chysx_cloth* sim, const chysx_cloth_material* m);
int chysx_cloth_get_material(const chysx_cloth* sim, chysx_cloth_material* out);

/* set_external_buffers (src/cloth/cloth_simulator.cu:248-258; bindings.cpp:66-77).
 * Never fails (the reference is noexcept); pointers are borrowed. */
int chysx_cloth_set_external_buffers(chysx_cloth* sim, uintptr_t pos, uintptr_t vel, int particle_count,
                                     uintptr_t inv_mass, uintptr_t force);

/* set_mesh (src/cloth/cloth_simulator.cu:315-377; bindings.cpp:167-192). */
int chysx_cloth_set_mesh(chysx_cloth* sim, const int32_t* triangles, int ntri);

/* build_springs_from_current_positions (cloth_simulator.cu:379-398; bindings.cpp:193-206). */
int chysx_cloth_build_springs(chysx_cloth* sim, float stiffness, uintptr_t stream);
/* num_springs (bindings.cpp:207-212). */
int chysx_cloth_num_springs(const chysx_cloth* sim, int* out);

/* build_fem_stretch_from_current_positions (cloth_simulator.cu:400-421; bindings.cpp:244-259). */
int chysx_cloth_build_fem_stretch(chysx_cloth* sim, float stiffness, uintptr_t stream);
/* num_fem_stretch_triangles (bindings.cpp:260-265). */
int chysx_cloth_num_fem_stretch(const chysx_cloth* sim, int* out);
/* build_fem_shear_from_current_positions (cloth_simulator.cu:423-449;
 * bindings.cpp:266-285): the stretch element with the material axes rotated
 * by pi/4, one per face. */
int chysx_cloth_build_fem_shear(chysx_cloth* sim, float stiffness, uintptr_t stream);
/* num_fem_shear_triangles (bindings.cpp:286-290). */
int chysx_cloth_num_fem_shear(const chysx_cloth* sim, int* out);
/* build_bending_from_current_positions (cloth_simulator.cu:451-514;
 * bindings.cpp:291-313): one dihedral per interior edge, rest angles from the
 * current positions.  Dihedrals are ordered by hinge (v0, v1); the
 * reference's order is its hash map's. */
int chysx_cloth_build_bending(chysx_cloth* sim, float stiffness, uintptr_t stream);
/* num_bending_dihedrals (bindings.cpp:314-318). */
int chysx_cloth_num_bending(const chysx_cloth* sim, int* out);
/* ---- static-shape contact (cloth_simulator.h:398-450; bindings.cpp:595-690) ----
 * Penalty contact with planes dot(n, x) + d = 0 (n points to the free side)
 * and oriented boxes (center, half extents along the unit columns ex, ey, ez):
 * force k (h - dist) n on the deepest contact below thickness h, block
 * k n n^T on the diagonal, and with friction mu > 0 the lagged IPC friction
 * block alpha (I - n n^T) plus the Coulomb-cone projection of the force
 * (static_contact.cu:107-380).  Active when stiffness > 0 and a shape exists. */
int chysx_cloth_add_static_plane(chysx_cloth* sim, const float* n, float d);
int chysx_cloth_add_static_box(chysx_cloth* sim, const float* center, const float* half_ext, const float* ex,
                               const float* ey, const float* ez);
int chysx_cloth_clear_static_shapes(chysx_cloth* sim);
int chysx_cloth_static_shape_counts(const chysx_cloth* sim, int* planes, int* boxes);
/* which: 0 thickness, 1 stiffness, 2 friction mu, 3 friction epsilon (default 1e-4). */
int chysx_cloth_set_static_contact(chysx_cloth* sim, int which, float value);
int chysx_cloth_get_static_contact(const chysx_cloth* sim, int which, float* value);
/* Dihedrals (4 ints each: hinge v0 < v1, wings v2, v3) and rest angles of the
 * installed bending elements; either output may be null. */
int chysx_cloth_get_dihedrals(const chysx_cloth* sim, int32_t* dihedrals, float* rest_angles);

/* set_pins / clear_pins / num_pins / update_pin_targets
 * (cloth_simulator.cu:277-313; bindings.cpp:85-166).  targets: n x 3 floats. */
int chysx_cloth_set_pins(chysx_cloth* sim, const int32_t* indices, const float* targets, int n, float stiffness);
int chysx_cloth_clear_pins(chysx_cloth* sim);
int chysx_cloth_num_pins(const chysx_cloth* sim, int* out);
int chysx_cloth_update_pin_targets(chysx_cloth* sim, const float* targets, int n, uintptr_t stream);

/* redistribute_mass_area_weighted (cloth_simulator.cu:516-575; bindings.cpp:213-243). */
int chysx_cloth_redistribute_mass_area_weighted(chysx_cloth* sim, float surface_density, uintptr_t inv_mass,
                                                int particle_count, uintptr_t stream);

/* No reference counterpart: the entry point above without device allocation,
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The vertex-to-triangle incidence
 * and the triangles are staged in the library's fixed 64 MiB device scratch
 * array (one per GPU, held until the process ends) instead of fresh device
 * buffers; (particle_count + 1 + 6 triangles) * 4 bytes must fit in it, or
 * CHYSX_EINVAL.  Results are bitwise those of the allocating entry point.
 * Calls that overlap on one GPU share the array and race.  Synchronises
 * `stream` before returning. */
int chysx_cloth_redistribute_mass_area_weighted_noalloc(chysx_cloth* sim, float surface_density, uintptr_t inv_mass,
                                                        int particle_count, uintptr_t stream);

/* set_pcg_iterations / pcg_iterations (cloth_simulator.h:597-600; bindings.cpp:544-549). */
int chysx_cloth_set_pcg_iterations(chysx_cloth* sim, int max_iter);
int chysx_cloth_pcg_iterations(const chysx_cloth* sim, int* out);

/* step (src/cloth/cloth_simulator.cu:814-1261; bindings.cpp:78-84).
 * With a fixed iteration count (tolerance 0, the default and the reference's
 * mode) every kernel, the iterations and the pos / vel update are enqueued on
 * `stream` and the call returns without waiting, as the reference returns
 * after its graph launch (pcg_solver.cu:882-894); pos / vel are updated when
 * the stream reaches that point.  The solve statistics wait for it.  With a
 * tolerance (chysx_cloth_set_pcg_tolerance > 0) or a warm-start stage the
 * step synchronises `stream` once per poll of the solver state (after the
 * iterations the previous step needed, then every 8) and returns when the
 * solve has ended, with the pos / vel update enqueued behind it on `stream`.
 * Consecutive steps may use different streams: each is ordered after the
 * previous one. */
int chysx_cloth_step(chysx_cloth* sim, float dt, uintptr_t stream);

/* num_particles / num_off_diag_blocks of the last solve (cloth_simulator.h:632-633). */
int chysx_cloth_num_particles(const chysx_cloth* sim, int* out);
int chysx_cloth_num_off_diag_blocks(const chysx_cloth* sim, int* out);

/* debug_copy_* behind debug_dump_last_solve (cloth_simulator.cu:1272-1348;
 * bindings.cpp:551-594): diag n x 9, row_offsets n+1, col_indices nnz,
 * values nnz x 9 (row-major 3x3 blocks, CSR order), rhs n x 3, dx n x 3. */
int chysx_cloth_debug_copy_hessian_diag(const chysx_cloth* sim, float* out);
int chysx_cloth_debug_copy_hessian_csr(const chysx_cloth* sim, int32_t* row_offsets, int32_t* col_indices,
                                       float* values);
int chysx_cloth_debug_copy_last_rhs(const chysx_cloth* sim, float* out);
int chysx_cloth_debug_copy_last_dx(const chysx_cloth* sim, float* out);

/* ---- extensions (no reference counterpart; additive) ------------------- */

/* Stop PCG at true relative residual sqrt(|b-Ax|^2 / max(|b|^2,1e-30)) <= tol
 * (the reference's definition, pcg_solver.cu:595-641, 916-923), capped at
 * pcg_iterations().  tol = 0 (default) runs exactly pcg_iterations()
 * iterations like the reference (pcg_solver.cu:542-592). */
int chysx_cloth_set_pcg_tolerance(chysx_cloth* sim, double tol);
/* Include springs in the step's gradient and Hessian (the reference builds
 * them but leaves them out of the step, cloth_simulator.cu:921). Default 0. */
int chysx_cloth_set_springs_active(chysx_cloth* sim, int on);
/* Feed the external force buffer into the predictor (the reference passes
 * null, cloth_simulator.cu:892-894).  Default 0. */
int chysx_cloth_set_external_force_active(chysx_cloth* sim, int on);
/* Pins as constraints (no reference counterpart: the reference's pins are
 * penalties of stiffness k, pin_constraint.h:5-15, pin_constraint.cu:61-92).
 * on = 1: each pinned particle's displacement is fixed at target - x_n and
 * eliminated from the system (its couplings move to the right-hand side of
 * the free rows, its own row becomes the identity), so it ends every step
 * exactly at its target; the pin stiffness is unused.  Default 0 (penalty). */
int chysx_cloth_set_pin_projection(chysx_cloth* sim, int on);
int chysx_cloth_pin_projection(const chysx_cloth* sim, int* out);
/* Per-element stiffness (one value per spring in edge order / per triangle). */
int chysx_cloth_set_spring_stiffnesses(chysx_cloth* sim, const float* k, int n);
int chysx_cloth_set_fem_stretch_stiffnesses(chysx_cloth* sim, const float* k, int n);
/* 1: a fixed-iteration step also waits until its solve has ended (the
 * polled path); 0 (default): it returns once enqueued.  Results do not
 * depend on it.  No reference counterpart (the reference never waits). */
int chysx_cloth_set_blocking_step(chysx_cloth* sim, int on);
/* Iterations, restarts, status (CHYSX_PCG_*) and true relative residual of
 * the last solve; waits for an enqueued step. */
int chysx_cloth_last_solve_stats(const chysx_cloth* sim, int* iters, int* restarts, int* status, double* relres);
/* The FP64 warm start dx (n x 3) for the next step; zero resets it. */
int chysx_cloth_set_warm_start(chysx_cloth* sim, const double* dx, int n);
int chysx_cloth_zero_warm_start(chysx_cloth* sim);
int chysx_cloth_get_dx64(const chysx_cloth* sim, double* out, int n);
/* Edge list (2 ints per edge, lexicographic) of the current mesh. */
int chysx_cloth_num_edges(const chysx_cloth* sim, int* out);
int chysx_cloth_get_edges(const chysx_cloth* sim, int32_t* out);

/* Host-only topology query (no GPU needed): pattern and slot maps of a mesh,
 * bit-identical to BlockCSR3::build_topology / resolve_slots
 * (src/sparse/block_csr.cu:91-204; constraint_n.h:115-153).
 * include_springs adds both directions of every edge.  Call once with null
 * outputs to get sizes. */
int chysx_topology_query(int n, const int32_t* triangles, int ntri, int include_springs, const int32_t* pins,
                         int npin, int* nnz, int* nedges, int32_t* row_offsets, int32_t* col_indices,
                         int32_t* edges, int32_t* tri_slots, int32_t* spring_slots, int32_t* pin_slots);

/* Host-only topology query with FEM triangles and bending dihedrals: the
 * pattern of ensure_hessian_topology (cloth_simulator.cu:698-793) over
 * triangle pairs and the 12 directed pairs of every dihedral, the dihedral
 * list and its slot map (16 per dihedral).  Null outputs give sizes. */
int chysx_topology_query_bending(int n, const int32_t* triangles, int ntri, int* nnz, int* ndih,
                                 int32_t* row_offsets, int32_t* col_indices, int32_t* dihedrals,
                                 int32_t* dihedral_slots);

/* ---- batched cloths (extension) ---------------------------------------- */
typedef struct chysx_batch chysx_batch;
int chysx_batch_create(chysx_batch** out);
int chysx_batch_destroy(chysx_batch* b);
/* Adds a cloth of n particles occupying the next n particles of the batch's
 * external buffers.  spring_k / stretch_k: per element or null (then the
 * scalar applies). */
int chysx_batch_add_cloth(chysx_batch* b, const int32_t* triangles, int ntri, int n, const float* spring_k,
                          float spring_k_scalar, const float* stretch_k, float stretch_k_scalar,
                          const int32_t* pins, const float* pin_targets, int npin, float pin_k, int* index);
int chysx_batch_set_external_buffers(chysx_batch* b, uintptr_t pos, uintptr_t vel, int total, uintptr_t inv_mass,
                                     uintptr_t force);
int chysx_batch_set_material(chysx_batch* b, int index, const chysx_cloth_material* m);
int chysx_batch_set_options(chysx_batch* b, double tol, int max_iter, int springs_active, int force_active);
/* Same enqueue / poll behaviour as chysx_cloth_step: enqueued without waiting
 * when every cloth runs a fixed iteration count. */
int chysx_batch_step(chysx_batch* b, float dt, uintptr_t stream);
int chysx_batch_set_blocking_step(chysx_batch* b, int on);
/* Where the PCG state is advanced (launch config's split advance): -1 by
 * size (default: inside the SpMV workgroups for a batch), 0 inside every SpMV
 * workgroup, 1 in a kernel of its own.  Speed only; results are unchanged. */
int chysx_batch_set_split_advance(chysx_batch* b, int mode);
/* Pin projection (as chysx_cloth_set_pin_projection) for every cloth of the batch. */
int chysx_batch_set_pin_projection(chysx_batch* b, int on);
int chysx_batch_stats(const chysx_batch* b, int index, int* iters, int* restarts, int* status, double* relres);
/* Number of cloths added to the batch. */
int chysx_batch_num_cloths(const chysx_batch* b, int* n);
/* The same for every cloth at once (arrays of n = the cloth count; any may be
 * null).  No counterpart in the reference, which steps its cloths one by
 * one; for callers that check every cloth after every batch step. */
int chysx_batch_stats_all(const chysx_batch* b, int n, int* iters, int* restarts, int* status, double* relres);

/* ---- kernel-level benchmark hooks -------------------------------------- */
/* Time `reps` launches of one kernel of the last assembled system on `stream`
 * with HIP events recorded on that stream; returns the mean ms per launch.
 * which: 0 = plain SpMV y = A x, 1 = PCG SpMV pass as the solve runs it
 * (advance of the solver state by the previous iteration's partials,
 * direction + A p + p.Ap), 2 = PCG update, 3 = assembly, 4 = one full PCG
 * iteration (SpMV pass and update pass, plus its share of the advance that
 * closes each chunk of iterations; timed over whole chunks), 5 = the PCG
 * SpMV pass without the advance (the state is taken as current).
 * The solver state is left unusable for warm starts; call before a step. */
int chysx_cloth_time_kernel(chysx_cloth* sim, int which, int reps, float dt, uintptr_t stream, double* ms);
/* Kernel launch shapes and matrix storage.
 * cfg[CHYSX_LAUNCH_CONFIG_LEN] = {assemble waves, residual waves, SpMV-pass
 * waves, SpMV unroll, update-pass waves, finalize waves, SpMV branch-free
 * padding (0/1), SpMV symmetric (0/1: lower blocks read as transposes of
 * their upper mirrors; changes results at FP32 rounding level, never chosen
 * by the autotuner), packed blocks (0/1: with the symmetric SpMV and no
 * bending elements, diagonal and off-diagonal blocks stored as their upper
 * triangle; the operator and results are bitwise those of 9-float blocks),
 * 16-bit indices (0/1: the SpMV pass reads 16-bit column and mirror offsets
 * when every offset of the mesh fits; the same indices), split advance
 * (-1/0/1: each PCG iteration advances the solver state in a small kernel of
 * its own instead of inside every SpMV workgroup; the same arithmetic; -1,
 * the default, splits for a single cloth of 512 or more groups of 1024 rows)};
 * waves in
 * {1,2,4,8,16} wavefronts of 64 lanes per workgroup, except the PCG passes
 * (residual, SpMV, update), which always run 16 (one workgroup per 16-slice
 * group of inner-product partials; other values are rejected); unroll in
 * {1,2,4}.  Shapes never change results.  By default an autotuner times the
 * SpMV variants (unroll x branch-free padding) on the problem's own matrix
 * when its topology is built (problems of 32768 rows or more). */
#define CHYSX_LAUNCH_CONFIG_LEN 11
int chysx_cloth_set_autotune(chysx_cloth* sim, int on);
int chysx_cloth_launch_config(const chysx_cloth* sim, int* cfg);
int chysx_cloth_set_launch_config(chysx_cloth* sim, const int* cfg);
/* Floats per stored off-diagonal block in effect: 6 (packed) or 9. */
int chysx_cloth_block_floats(const chysx_cloth* sim, int* floats);
/* Bits of the column / mirror offsets the SpMV pass reads: 16 or 32. */
int chysx_cloth_index_bits(const chysx_cloth* sim, int* bits);
/* Autotuner samples: count in/out; per sample {kernel (0 SpMV pass, 2
 * branch-free SpMV pass, 3 assembly, 4 / 5 a whole PCG iteration with the
 * state advanced in the SpMV pass / in a kernel of its own), waves, unroll}
 * and mean ms.  The tuner runs at the first step after the topology is built
 * (problems of 32768 padded rows or more) and keeps the fastest of each. */
int chysx_cloth_tune_log(const chysx_cloth* sim, int* count, int* kernel_waves_unroll, double* ms);
int chysx_batch_set_autotune(chysx_batch* b, int on);
/* ---- PCG warm start (extension) ------------------------------------------
 * mode 0: the previous step's dx (reference behaviour, default).
 * mode 2 (projection): Galerkin projection of the new system onto the span
 *   of the last K converged dx (K = chysx_*_ws_set_history, 2..8, default 6),
 *   accepted when it lowers the energy x^T A x / 2 - b^T x below that of
 *   dx_prev (the A-norm error PCG minimises).
 * mode 1 (learned): an edge message-passing network over the mesh graph
 *   (BF16 MFMA, FP32 accumulate) predicts a direction d from dx_prev;
 *   x = dx_prev + alpha d with alpha = d.r / d.Ad, kept when it lowers the
 *   energy (the residual in the A^-1 norm); else dx_prev is used.
 * mode 3 (projection + learned): the projection, then the network's
 *   direction in the same way.
 * Both start once two solutions are held; set/zero_warm_start resets the
 * history.  Parameters: chysx_ws_num_params() FP32 values.  Training: every
 * step runs the network; after the solve the loss
 * 1 - (d.r)^2 / ((d.Ad)(e.r)), e = dx_converged - x_base, drives one Adam
 * step (lr). */
#define CHYSX_WS_STATS_LEN 7
int chysx_ws_num_params(int* out);
int chysx_cloth_set_warm_start_mode(chysx_cloth* sim, int mode);
int chysx_cloth_ws_set_history(chysx_cloth* sim, int k);
int chysx_cloth_ws_set_params(chysx_cloth* sim, const float* theta, int n);
int chysx_cloth_ws_get_params(chysx_cloth* sim, float* theta, int n);
int chysx_cloth_ws_init_params(chysx_cloth* sim, uint64_t seed);
int chysx_cloth_ws_set_training(chysx_cloth* sim, int on, float lr);
/* Acceptance of the network's line-searched direction: residual == 0
 * (default) keeps it whenever it lowers the energy x^T A x / 2 - b^T x (the
 * A-norm error PCG minimises; the exact line search always does unless
 * alpha = 0); residual == 1 keeps it only when the predicted residual
 * |b - A x| also drops, else falls back to the base guess. */
int chysx_cloth_ws_set_acceptance(chysx_cloth* sim, int residual);
/* Last step, out[0..6]: projection used (1/0), learned direction used (1/0),
 * projection directions kept, relative residuals |b - A x0| / |b| of dx_prev,
 * of the projection and of the learned guess (0 when that stage did not run),
 * mean training loss of the step.  n >= CHYSX_WS_STATS_LEN. */
int chysx_cloth_ws_stats(chysx_cloth* sim, double* out, int n);
int chysx_batch_set_warm_start_mode(chysx_batch* b, int mode);
int chysx_batch_ws_set_history(chysx_batch* b, int k);
int chysx_batch_ws_set_params(chysx_batch* b, const float* theta, int n);
int chysx_batch_ws_get_params(chysx_batch* b, float* theta, int n);
int chysx_batch_ws_init_params(chysx_batch* b, uint64_t seed);
int chysx_batch_ws_set_training(chysx_batch* b, int on, float lr);
int chysx_batch_ws_set_acceptance(chysx_batch* b, int residual);
int chysx_batch_ws_stats(chysx_batch* b, int index, double* out, int n);
/* Projection solve on the host (tests): c = G^+ g over the eigen-directions
 * of the symmetric k x k G (row-major) above 1e-12 of the largest; *kept =
 * directions used. */
int chysx_ws_small_solve(int k, const double* G, const double* g, double* c, int* kept);
/* Stand-alone network on a graph (n rows, CSR neighbours), for checking the
 * GPU forward / backward: F and E are n x 32 FP32 inputs, out n x 32, target
 * n x 3, loss = mean over rows of |out[:, 0:3] - target|^2, grad FP32. */
typedef struct chysx_wsnet chysx_wsnet;
int chysx_wsnet_create(chysx_wsnet** out);
int chysx_wsnet_destroy(chysx_wsnet* net);
int chysx_wsnet_set_params(chysx_wsnet* net, const float* theta, int n);
int chysx_wsnet_forward(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols, const float* F,
                        const float* E, float* out);
/* The simulator's inference path: the decoder folded into the last layer,
 * only the three output columns the warm-start direction uses; out n x 3. */
int chysx_wsnet_forward_direction(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols,
                                  const float* F, const float* E, float* out);
int chysx_wsnet_loss_grad(chysx_wsnet* net, int n, const int32_t* row_offsets, const int32_t* cols, const float* F,
                          const float* E, const float* target, double* loss, float* grad);

/* y = A x on the last assembled system with the plain one-lane-per-block-row
 * kernel (BlockCSR3 spmv, src/sparse/block_csr.cu:54-83, 210-243; alpha 1,
 * beta 0).  x, y: device float3 AoS of num_particles rows, y written.  It
 * reads the general storage (9-float blocks, both halves stored); with the
 * symmetric or packed storage in effect it fails with CHYSX_EINVAL. */
int chysx_cloth_spmv(chysx_cloth* sim, uintptr_t x, uintptr_t y, uintptr_t stream);
/* HBM bandwidth probe (no reference counterpart): a hand-written copy
 * (kind 0: b = a) or triad (kind 1: a = b + s c) over device buffers of
 * `bytes_per_buffer` each (rounded down to 16 KiB; use >= 1 GiB, far beyond
 * the 256 MB Infinity Cache), 16-byte accesses; kinds 2 and 3 are the same
 * with non-temporal loads and stores.  Allocates and frees its buffers.
 * *ms = mean per launch over `reps` launches timed with HIP events on
 * `stream`; *bytes_moved = bytes read + written per launch. */
int chysx_stream_bench(int kind, uint64_t bytes_per_buffer, int reps, uintptr_t stream, double* ms,
                       double* bytes_moved);
/* Algorithmic bytes of one plain SpMV: 40 per off-diagonal block + 64 per row + 4. */
int chysx_cloth_spmv_bytes(const chysx_cloth* sim, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* CHYSX_CLOTH_H */
