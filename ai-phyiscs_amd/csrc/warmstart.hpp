// SPDX-License-Identifier: Apache-2.0
//
// Learned PCG warm start: a small edge message-passing network over the mesh
// graph, run in BF16 on MFMA with FP32 accumulation, trained on the GPU by
// its own forward and backward kernels.
//
//   h0 = relu(F We + be)                                    F: N x 32 features
//   3 layers: x = [h | mean_{j in N(i)} h_j | eagg];  h += relu(x Wl + bl)
//   out = h Wd + bd                                         first 3 columns used
//
// eagg holds the mean edge features (x_j - x_i)/L0_ij and |x_j - x_i|/L0_ij - 1
// over the mesh neighbours.  The graph is the block pattern of the cloth's
// Hessian (its neighbours are the mesh neighbours), in the wave-sliced layout
// of the solver.  No counterpart exists in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cloth_device.hpp"
#include "cloth_kernels.hpp"

namespace chysx_amd {

struct WsGraph {
    const SliceDesc* slices;
    int nslices;        // rows = 64 * nslices
    const int* slice_ptr;
    const int* col;
    const int* row_deg;
};

class WarmStartNet {
public:
    static constexpr int kFin = 32;
    static constexpr int kHidden = 64;
    static constexpr int kLayerIn = 160;
    static constexpr int kOut = 32;
    static constexpr int kLayers = 3;
    static int num_params();

    WarmStartNet();
    ~WarmStartNet();
    WarmStartNet(const WarmStartNet&) = delete;
    WarmStartNet& operator=(const WarmStartNet&) = delete;

    // FP32 master parameters in the order [We, be, W0, b0, W1, b1, W2, b2,
    // Wd, bd], matrices stored (K, NOUT) row-major.
    void set_params(const float* theta, hipStream_t s);
    void get_params(float* theta) const;
    bool has_params() const { return has_params_; }
    // Random initialisation (He normal, decoder x0.1, zero biases), seeded.
    void init_params(uint64_t seed, hipStream_t s);

    void reserve(int rows, bool training);
    // Inputs written by the caller before forward().
    __bf16* features() { return F_; }
    __bf16* edge_features() { return E_; }
    // Forward: out (rows x 32, FP32).  keep = store activations for backward.
    void forward(const WsGraph& g, hipStream_t s, bool keep);
    const float* out() const { return out_; }
    // Loss against target (rows x 3, FP32; rows with valid[row] == 0 or
    // beyond a slice's valid count are ignored): mean over valid rows of the
    // squared error.  Writes dOUT and the loss partials (read with loss()).
    void loss(const WsGraph& g, const float* target, hipStream_t s);
    double read_loss(hipStream_t s);
    // Backward from dOUT into the gradient vector (FP32, num_params()).
    void backward(const WsGraph& g, hipStream_t s);
    const float* grad() const { return grad_; }
    void copy_grad(float* host, hipStream_t s) const;
    // Adam step on the master parameters; refreshes the BF16 copies.
    void adam_step(float lr, hipStream_t s);
    // Direct access for tests: upstream gradient dOUT (rows x 32, BF16).
    __bf16* dout() { return dout_; }
    // Upload host FP32 features (n x 32) and edge features (n x 32) into the
    // BF16 inputs; rows n .. rows-1 are zeroed.
    void set_inputs(const float* F, const float* E, int n, hipStream_t s);
    void copy_out(float* host, int n, hipStream_t s) const;

private:
    void refresh_bf16(hipStream_t s);
    void free_all();

    int rows_ = 0;
    bool training_ = false;
    bool has_params_ = false;
    int adam_t_ = 0;
    // parameters
    float* theta_ = nullptr;   // FP32 master
    float* grad_ = nullptr;
    float* m_ = nullptr;
    float* v_ = nullptr;
    __bf16* wnk_ = nullptr;    // forward layout per matrix: (NOUT, K)
    __bf16* wkn_ = nullptr;    // input-gradient layout per matrix: (K, NOUT)
    // activations (rows x width, BF16 unless noted)
    __bf16* F_ = nullptr;      // 32
    __bf16* E_ = nullptr;      // 32 (mean edge features)
    __bf16* H_[4] = {};        // 64 each (H_[0] = encoder output)
    __bf16* A_[3] = {};        // 64 each (mean neighbour h per layer)
    __bf16* P_[4] = {};        // pre-activations: P_[0] encoder, P_[1..3] layers
    float* out_ = nullptr;     // 32, FP32
    // backward scratch
    __bf16* dout_ = nullptr;   // 32
    float* dh_ = nullptr;      // 64
    float* dh2_ = nullptr;     // 64
    __bf16* dpre_ = nullptr;   // 64
    float* dx_ = nullptr;      // 160
    float* part_ = nullptr;    // weight-gradient partials
    double* lpart_ = nullptr;  // loss partials (1 per slice)
    int lpart_n_ = 0;
    long long loss_nv_ = 1;
};

// ---- simulator-side kernels ------------------------------------------------

// Per-instance warm-start bookkeeping (device).
struct WsInst {
    double rr_base;   // |b - A x_base|^2
    double rr_cand;   // |b - A x_cand|^2
    double scale;     // RMS of dx_prev over the instance's rows
    int use_cand;     // 1: the network's guess replaces the base guess
    int pad;
};

// Residual of the base guess x (FP64): r (FP32) and z = M^-1 r into v.r / v.z;
// per-slice partials [|r|^2, |x|^2].
void launch_ws_prep(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const double* x,
                    double* partials, hipStream_t s);
// Per-instance reduction of launch_ws_prep / launch_ws_cand partials.
void launch_ws_reduce(WsInst* wi, const InstState* st, const double* partials, int ninst, int phase, hipStream_t s);
// Network inputs for every row: features F and mean edge features E (BF16);
// also rotates the history (hprev <- dx_prev as FP32).
void launch_ws_features(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const DevExt& e,
                        const DevMesh& m, const float* rest_pos, float* hprev, const WsInst* wi,
                        const InstParams* params, float dt, __bf16* F, __bf16* E, hipStream_t s);
// Candidate guess xc = x + scale * out[:, 0:3] (FP64) and its residual partials.
void launch_ws_candidate(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const float* out,
                         const WsInst* wi, double* xc, double* partials, hipStream_t s);
// x <- xc for instances whose candidate residual is smaller.
void launch_ws_select(const LaunchGeom& g, const WsInst* wi, const double* xc, double* x, hipStream_t s);
// hprev <- x (FP32), n values: keeps the previous solution while the
// network's inputs are still being filled.
void launch_ws_history(const double* x, float* hprev, size_t n, hipStream_t s);
// Training target (x_converged - x_base) / scale, x_base = hprev (FP32).
void launch_ws_target(const LaunchGeom& g, const double* x, const float* hprev, const WsInst* wi, float* target,
                      hipStream_t s);

}  // namespace chysx_amd
