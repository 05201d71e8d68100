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

struct WsInst;
struct WsHist;

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
    // Compact edge features (rows x 4, BF16): the four columns the simulator's
    // features fill, the other 28 being zero.  With set_compact_edges(true) the
    // fused inference layers read these instead of edge_features() (8 bytes a
    // row instead of 64; the same products, since the rest are zeros).
    // set_inputs() switches back to the full 32 columns.
    __bf16* edge_features_compact() { return E4_; }
    void set_compact_edges(bool on) { compact_edges_ = on; }
    // Forward: out (rows x 32, FP32).  keep = store activations for backward
    // (one kernel per aggregation and per GEMM).  Without keep each layer is
    // one fused kernel (the neighbour mean staged in LDS, never in memory); with
    // direction_only the decoder is folded into the last layer and only the
    // output columns 0..2 (the warm-start direction) are written.
    // With direction_only, wi and dir given, the last layer writes the
    // warm-start direction d = scale * out[:, 0:3] (0 on padding rows) into
    // dir itself (rows x 3) instead of out.
    // With encoded (inference only) h0 is already in place (features_encode)
    // and the encoder is skipped.
    void forward(const WsGraph& g, hipStream_t s, bool keep, bool direction_only = false, const WsInst* wi = nullptr,
                 float* dir = nullptr, bool encoded = false);
    // Inference from a simulator's state: the features of every row (as
    // launch_ws_features forms them) go straight into the encoder in one
    // kernel, which writes h0 and the compact edge features (F is not
    // written) and switches the layers to the compact edge features.
    void features_encode(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const DevExt& e,
                         const DevMesh& m, const float* rest_pos, const WsHist& h, const double* xb,
                         const WsInst* wi, float dt, hipStream_t s);
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
    __bf16* E4_ = nullptr;     // 4 (the same, compact)
    bool compact_edges_ = false;
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
//
// Warm-start pipeline of one step (all instances in one launch each):
//   projection: Galerkin projection of the system onto the span of the last
//     K converged dx (newest first): G = V^T A V, g = V^T b, x = V c with
//     c = G^+ g (eigen-truncated), accepted when its energy
//     phi(x) = x^T A x / 2 - b^T x is lower than phi(dx_prev);
//   learned: from the base guess x_b (dx_prev, or the projection) the
//     network predicts a direction d; x = x_b + alpha d with the A-norm line
//     search alpha = d^T r_b / d^T A d, accepted when it lowers the energy
//     (the residual in the A^-1 norm), which the exact line search does
//     unless d is A-orthogonal to the error; the 2-norm residual
//     |r_b - alpha A d| is predicted and reported;
//   training: loss 1 - (d^T r_b)^2 / ((d^T A d)(e^T r_b)), e = x_conv - x_b,
//     the fraction of the base guess's A-norm error left after the line search.

constexpr int kWsHistMax = 8;

struct WsHist {
    const float* v[kWsHistMax];  // newest first, 3 * rows floats each
    int k;
};

// Per-instance warm-start bookkeeping (device).
struct WsInst {
    double coef[kWsHistMax];  // projection coefficients (newest first)
    double bb;        // |b|^2
    double rr_prev;   // |b - A dx_prev|^2
    double rr_proj;   // |b - A x_proj|^2 (predicted from the Gram sums)
    double rr_base;   // |b - A x_b|^2 (learned stage, exact)
    double rr_cand;   // |b - A (x_b + alpha d)|^2 (predicted)
    double scale;     // RMS of x_b over the instance's rows
    double u, w;      // d.r_b, d.A d
    double alpha;     // line-search step (0: rejected)
    double energy;    // e.r_b = e^T A e (training)
    double loss;      // training loss of this instance
    double gcoef;     // dLoss/dd = gcoef * (r_b - (u/w) A d)   (scaled by 1/ninst)
    int use_proj;     // 1: the projection replaced dx_prev
    int use_cand;     // 1: the learned direction was applied
    int kdim;         // projection directions kept
    int pad;
};

// Number of Gram sums for k history vectors: two packed triangles (V^T A V,
// (AV)^T (AV)), V^T b, (AV)^T b and b^T b.
constexpr int ws_gram_values(int k) { return k * (k + 1) + 2 * k + 1; }
constexpr int kWsMaxValues = ws_gram_values(kWsHistMax);
constexpr int kWsChunkSlices = 1024;  // slices per reduction workgroup

// Phases of launch_ws_finish.
enum WsPhase : int { kWsPhaseProj = 0, kWsPhaseBase = 1, kWsPhaseLine = 2, kWsPhaseEnergy = 3 };

struct WsReduce {
    double* part;   // SoA [values][ldp]
    int ldp;        // >= nslices
    double* part2;  // [ninst][chunks][values]
    int chunks;     // ceil(max slices of an instance / kWsChunkSlices)
};

// Gram sums of the history against the current system (per-slice partials).
void launch_ws_gram(const LaunchGeom& g, const DevMatrix& A, const float* rhs, const WsHist& h, const WsReduce& red,
                    hipStream_t s);
// Per-instance sums of `nvalues` per-slice partials (deterministic, fixed
// chunking) followed by the phase's per-instance decision.
void launch_ws_finish(const WsReduce& red, int nvalues, const InstState* st, WsInst* wi, int ninst, int phase,
                      int hist_k, hipStream_t s);
// x <- V c for instances whose projection was accepted.
void launch_ws_proj_apply(const LaunchGeom& g, const WsHist& h, const WsInst* wi, double* x, hipStream_t s);
// Base guess of the learned stage: r_b = b - A x (FP64 sums, stored FP32),
// z = M^-1 r_b into v.z, xb <- x (skipped when xb is null: inference reads x
// itself, which stays the base guess until the line search applies);
// partials [|r_b|^2, |x|^2, |b|^2].
void launch_ws_base(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const double* x, double* xb,
                    float* rb, const WsReduce& red, hipStream_t s);
// Network inputs for every row: features F and mean edge features (BF16),
// into E4 (rows x 4) when it is given, else into E (rows x 32, columns 4..31
// zero).
void launch_ws_features(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const DevExt& e,
                        const DevMesh& m, const float* rest_pos, const WsHist& h, const double* xb,
                        const WsInst* wi, float dt, __bf16* F, __bf16* E, __bf16* E4, hipStream_t s);
// d = scale * out[:, 0:3] and A d; partials [d.r_b, d.Ad, Ad.r_b, Ad.Ad].
void launch_ws_direction(const LaunchGeom& g, const DevMatrix& A, const float* out, const WsInst* wi,
                         const float* rb, float* d, float* ad, const WsReduce& red, hipStream_t s);
// x <- x + alpha d for instances whose direction was accepted; with r, also
// r <- r_b - alpha A d (alpha 0 where the direction was not applied: r = r_b),
// the residual of the new guess with the FP64 sums of r_b and A d.
void launch_ws_line_apply(const LaunchGeom& g, const WsInst* wi, const float* d, double* x, const float* rb,
                          const float* ad, float* r, hipStream_t s);
// Training: partials [(x - xb).r_b].
void launch_ws_energy(const LaunchGeom& g, const double* x, const double* xb, const float* rb, const WsReduce& red,
                      hipStream_t s);
// Training: upstream gradient of the network output (BF16, rows x 32).
void launch_ws_dout(const LaunchGeom& g, const WsInst* wi, const float* rb, const float* ad, __bf16* dout,
                    hipStream_t s);
// hist <- x as FP32 (n values).
void launch_ws_history(const double* x, float* hist, size_t n, hipStream_t s);

// Host twin of the per-instance projection solve (tests): coefficients c of
// the eigen-truncated solve G c = g (G symmetric k x k, row-major); returns
// the number of eigen-directions kept.
int ws_small_solve_host(int k, const double* G, const double* g, double* c);

}  // namespace chysx_amd
