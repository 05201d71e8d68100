// SPDX-License-Identifier: Apache-2.0
//
// Learned PCG warm start on gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16)
// forward / input-gradient / weight-gradient kernels, mean aggregation over
// the mesh graph and its adjoint, loss, Adam, and the simulator-side feature,
// candidate and fallback kernels.  See warmstart.hpp for the network.
//
// MFMA 16x16x32 BF16 operand layout used throughout (lane l, li = l % 16,
// lg = l / 16): A[i][k] with i = li, k = 8 lg .. 8 lg + 7; B[k][j] with
// j = li, k = 8 lg .. 8 lg + 7; D[i][j] with i = 4 lg + r (r = 0..3), j = li.

#include "warmstart.hpp"
#include "matrix_access.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

namespace chysx_amd {

[[noreturn]] void throw_hip(hipError_t err, const char* what);
void check_hip(hipError_t err, const char* what);

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int FIN = WarmStartNet::kFin;
constexpr int HID = WarmStartNet::kHidden;
constexpr int XL = WarmStartNet::kLayerIn;
constexpr int NOUTD = WarmStartNet::kOut;
constexpr int NL = WarmStartNet::kLayers;

// Parameter offsets (floats): [We, be, W0, b0, W1, b1, W2, b2, Wd, bd].
constexpr int OFF_WE = 0;
constexpr int OFF_BE = OFF_WE + FIN * HID;
constexpr int OFF_W0 = OFF_BE + HID;
constexpr int LAYER_STRIDE = XL * HID + HID;
constexpr int OFF_WD = OFF_W0 + NL * LAYER_STRIDE;
constexpr int OFF_BD = OFF_WD + HID * NOUTD;
constexpr int NPARAM = OFF_BD + NOUTD;
inline int off_w(int l) { return OFF_W0 + l * LAYER_STRIDE; }
inline int off_b(int l) { return off_w(l) + XL * HID; }

struct MatDesc {
    int off, K, N;
};
constexpr int NMAT = NL + 2;
const MatDesc kMats[NMAT] = {{OFF_WE, FIN, HID}, {OFF_W0, XL, HID}, {OFF_W0 + LAYER_STRIDE, XL, HID},
                             {OFF_W0 + 2 * LAYER_STRIDE, XL, HID}, {OFF_WD, HID, NOUTD}};

__device__ __forceinline__ float bf2f(__bf16 v) { return float(v); }
__device__ __forceinline__ __bf16 f2bf(float v) { return __bf16(v); }

// m[t] += valid ? v[t] : 0 for t = 0..7, one dot2 per column: v[t] * 1 + v[t^1]
// * 0 + m[t] is exact before its one rounding, so the sum is that of m[t] +=
// float(v[t]) (for finite v), without converting v first.  Invalid entries
// leave m unchanged (m is never -0: it starts at +0 and an exact cancellation
// rounds to +0).  The BF16 pair (1, 0) is kept in a register: as a constant
// the compiler encodes it as the inline constant 1.0, whose meaning for a
// packed BF16 operand is not that pair.
__device__ __forceinline__ void add_bf16x8(float* m, const bf16x8& v, bool valid) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    unsigned one_lo = valid ? 0x3F80u : 0u;  // BF16 1.0 in the low half
    asm volatile("" : "+v"(one_lo));
    const bf16x2 lo = __builtin_bit_cast(bf16x2, one_lo), hi = __builtin_bit_cast(bf16x2, one_lo << 16);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bf16x2 p = {v[2 * j], v[2 * j + 1]};
        m[2 * j] = __builtin_amdgcn_fdot2_f32_bf16(p, lo, m[2 * j], false);
        m[2 * j + 1] = __builtin_amdgcn_fdot2_f32_bf16(p, hi, m[2 * j + 1], false);
    }
}

// Slice of this wavefront in a 4-wave workgroup, XCD-aware: workgroup b runs
// on XCD b % 8, and each XCD walks a contiguous eighth of the slices, so a
// row's +-1 and +-W neighbours (W up to a few thousand rows) were loaded into
// the same XCD's L2 by the neighbouring wavefronts.  The grid is ws_grid.
__device__ __forceinline__ int ws_slice(int nslices) {
    const int nwg = (nslices + 3) >> 2;
    const int per = (nwg + 7) >> 3;
    const int b = int(blockIdx.x);
    // wave-uniform: the slice's bounds and loops then live in scalar registers
    return __builtin_amdgcn_readfirstlane(((b & 7) * per + (b >> 3)) * 4 + int(threadIdx.x >> 6));
}
inline unsigned ws_grid(int nslices) {
    const int nwg = (nslices + 3) / 4;
    return unsigned(8 * ((nwg + 7) / 8));
}

// ---------------------------------------------------------------------------
// GEMM: Y = act(sum_s X_s W_s + bias) (+ resid), rows x NOUT, X row-major
// BF16 segments of widths multiple of 32; weights in (NOUT, K) layout.
// One wave per 64 rows: 4 x NT tiles of 16 x 16, FP32 accumulators.
// ---------------------------------------------------------------------------
struct GemmArgs {
    int rows;
    const __bf16* seg[3];
    int segw[3];
    int nseg;
    int K;
    const __bf16* wnk;
    const float* bias;
    int relu;
    const __bf16* resid;
    float* out_f;
    __bf16* out_b;
    __bf16* pre_b;
};

template <int NT>
__global__ __launch_bounds__(256) void ws_gemm_kernel(GemmArgs a) {
    constexpr int NOUT = 16 * NT;
    const int wave = ws_slice(a.rows / 64);
    const int row0 = wave * 64;
    if (row0 >= a.rows) return;
    const int l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    int kg = 0;
    for (int sg = 0; sg < a.nseg; ++sg) {
        const __bf16* X = a.seg[sg];
        const int w = a.segw[sg];
        for (int k0 = 0; k0 < w; k0 += 32) {
            bf16x8 af[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                af[mt] = *reinterpret_cast<const bf16x8*>(X + (size_t)(row0 + mt * 16 + li) * w + k0 + 8 * lg);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const bf16x8 bfr =
                    *reinterpret_cast<const bf16x8*>(a.wnk + (size_t)(nt * 16 + li) * a.K + kg + k0 + 8 * lg);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mt], bfr, acc[mt][nt], 0, 0, 0);
            }
        }
        kg += w;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + mt * 16 + 4 * lg + r;
                const int n = nt * 16 + li;
                const size_t o = (size_t)row * NOUT + n;
                float v = acc[mt][nt][r] + (a.bias ? a.bias[n] : 0.0f);
                if (a.pre_b) a.pre_b[o] = f2bf(v);
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.resid) v += bf2f(a.resid[o]);
                if (a.out_f) a.out_f[o] = v;
                if (a.out_b) a.out_b[o] = f2bf(v);
            }
}

// ---------------------------------------------------------------------------
// Inference layer, fused: h' = relu([h | mean_{j in N(i)} h_j | e] W + b) + h
// for the 64 rows of one slice per wave, 32 rows at a time, in two phases.
//  1. The neighbour mean of the 32 rows goes to the wave's LDS tile
//     (32 x HID BF16).  Eight lanes share a row and each reads 16 B of a
//     neighbour's 128-B row, so one load instruction covers eight whole rows;
//     the rows' column indices come from one coalesced load per neighbour
//     slot and a lane shuffle, and four neighbour rows of a group of eight
//     rows are in flight together.  The mean is bitwise that of ws_agg_kernel
//     (same order, same 1/deg, same BF16 rounding); it never goes to memory.
//  2. The layer GEMM on MFMA 16x16x32 BF16 with the operands swapped
//     (weights as A, rows as B), K = [h (2 chunks of 32) | mean (2, from LDS)
//     | e (1: the 32 edge-feature columns, or the compact 4 with zeros)] as in
//     the training GEMM.  The weight rows of output tile nt are
//     taken in the order ch(nt, i) = 32 (nt >> 1) + 8 (i >> 2) + 4 (nt & 1) +
//     (i & 3), so the lane (li, lg) ends with output columns 8 lg .. 8 lg + 7
//     and 32 + 8 lg .. 32 + 8 lg + 7 of row 16 mt + li: the h columns it
//     loaded for K chunks 0 and 1 (before the gathers), which are the
//     residual.  Permuting the output columns changes no result (each is the
//     same dot product).  The rows go back through the wave's tile, so each
//     store instruction writes eight whole 128-B rows: written from the MFMA
//     layout directly, each store covered half of 16 rows, and the partial
//     lines made the layer 30% slower.
// With LAST the decoder follows in the epilogue: only the output columns the
// warm start uses (0..2) are formed, from the BF16-rounded h' (as the decoder
// GEMM would read it), by FP32 FMAs and a sum over the four lanes of a row;
// h' itself is not stored.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
constexpr int kDirCols = 3;  // output columns the warm start reads

struct LayerArgs {
    int nslices;
    const SliceDesc* slices;
    const int* slice_ptr;
    const int* col;
    const int* row_deg;
    const __bf16* h;     // rows x HID
    const __bf16* e;     // rows x 32
    const __bf16* e4;    // rows x 4 (columns 0..3 of e, the rest zero), read instead of e when given
    const __bf16* w;     // (HID, XL): the layer's weights, (NOUT, K) layout
    const float* bias;   // HID
    __bf16* h_out;       // rows x HID (not LAST)
    const __bf16* wd;    // (NOUTD, HID): the decoder's weights (LAST)
    const float* bd;     // NOUTD (LAST)
    float* out;          // rows x NOUTD, columns 0..kDirCols-1 written (LAST, dir null)
    // LAST with dir: the warm-start direction d = scale * out[:, 0:3] of the
    // row's instance, 0 on padding rows, written instead of out
    const WsInst* wi;
    float* dir;
};

constexpr int kStageLd = HID + 8;  // LDS row stride (BF16) of the mean tile: 144 B, against bank conflicts
static_assert(HID == 64 && XL == 2 * HID + 32, "the layer kernel tiles h as 2 K chunks and e as 1");
static_assert(OFF_W0 % 8 == 0 && LAYER_STRIDE % 8 == 0 && OFF_WD % 8 == 0, "16-byte weight fragments");

// Output column held in register r of lane group lg for output tile nt (and
// the weight row an A-operand lane li = 4 lg + r supplies).
__device__ __forceinline__ int ws_out_col(int nt, int i) { return 32 * (nt >> 1) + 8 * (i >> 2) + 4 * (nt & 1) + (i & 3); }

// Orders one wave's LDS accesses around a point: LDS operations of a wave
// complete in issue order, so only the compiler's reordering is prevented.
__device__ __forceinline__ void ws_wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool LAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void ws_layer_kernel(LayerArgs a) {
    constexpr int NT = HID / 16, KC = XL / 32;
    constexpr int KB = 8;  // neighbours per gather batch
    // The weights as MFMA A fragments: wfrag[(kc NT + nt) 64 + lane] = W row
    // ws_out_col(nt, li), K columns 32 kc + 8 lg .. 32 kc + 8 lg + 7; the
    // bias, and with LAST the decoder rows 0..2 and biases, beside them.
    __shared__ bf16x8 wfrag[KC * NT * 64];
    __shared__ float4 bias_s[HID / 4];
    __shared__ bf16x8 wd_s[kDirCols * HID / 8];
    __shared__ float bd_s[kDirCols];
    // per wave: the neighbour mean of 32 rows
    __shared__ __attribute__((aligned(16))) __bf16 stage[4][32 * kStageLd];
    for (int f = int(threadIdx.x); f < KC * NT * 64; f += 256) {
        const int lane = f & 63, nt = (f >> 6) % NT, kc = (f >> 6) / NT;
        wfrag[f] = *reinterpret_cast<const bf16x8*>(a.w + (size_t)ws_out_col(nt, lane & 15) * XL + 32 * kc +
                                                    8 * (lane >> 4));
    }
    if (threadIdx.x < HID / 4) bias_s[threadIdx.x] = reinterpret_cast<const float4*>(a.bias)[threadIdx.x];
    if constexpr (LAST) {
        if (threadIdx.x < kDirCols * HID / 8) wd_s[threadIdx.x] = reinterpret_cast<const bf16x8*>(a.wd)[threadIdx.x];
        if (threadIdx.x < kDirCols) bd_s[threadIdx.x] = a.bd[threadIdx.x];
    }
    const int s = ws_slice(a.nslices);
    // a wave past the last slice still reaches the barriers
    const bool live = s < a.nslices;
    __bf16* tile = stage[threadIdx.x >> 6];
    const int row0 = s * 64;
    const int l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
    int sb = 0, w = 0;
    if (live) {
        sb = a.slice_ptr[s];
        w = a.slice_ptr[s + 1] - sb;
    }
    for (int half = 0; half < 2; ++half) {
        bf16x8 hres[2][2], xe[2];
        if (live) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const int row = row0 + 32 * half + 16 * mt + li;
                hres[mt][0] = *reinterpret_cast<const bf16x8*>(a.h + (size_t)row * HID + 8 * lg);
                hres[mt][1] = *reinterpret_cast<const bf16x8*>(a.h + (size_t)row * HID + 32 + 8 * lg);
                if (a.e4 != nullptr) {
                    bf16x8 t;
#pragma unroll
                    for (int u = 0; u < 8; ++u) t[u] = f2bf(0.0f);
                    if (lg == 0) {
                        const bf16x4 e4 = *reinterpret_cast<const bf16x4*>(a.e4 + (size_t)row * 4);
#pragma unroll
                        for (int u = 0; u < 4; ++u) t[u] = e4[u];
                    }
                    xe[mt] = t;
                } else {
                    xe[mt] = *reinterpret_cast<const bf16x8*>(a.e + (size_t)row * 32 + 8 * lg);
                }
            }
            // the first KB column indices and the degree of row l of the slice
            // (one coalesced load each); the gather lanes take theirs by a lane
            // shuffle
            int cl[KB];
#pragma unroll
            for (int u = 0; u < KB; ++u) cl[u] = (u < w) ? a.col[(size_t)(sb + u) * 64 + l] : -1;
            const int degl = a.row_deg[row0 + l];
            const int r8 = l >> 3, q = l & 7;  // row r8 of a group of eight, columns 8 q .. 8 q + 7
#pragma unroll 1
            for (int g = 0; g < 4; ++g) {
                const int sl = 32 * half + 8 * g + r8;
                float m[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) m[t] = 0.0f;
                for (int k0 = 0; k0 < w; k0 += KB) {
#pragma unroll
                    for (int u0 = 0; u0 < KB; u0 += KB / 2) {
                        // padding entries (c < 0) read the row itself and add nothing
                        int c[KB / 2];
                        bf16x8 v[KB / 2];
#pragma unroll
                        for (int u = 0; u < KB / 2; ++u)
                            if (k0 + u0 + u < w) {
                                const int cu = (k0 == 0) ? cl[u0 + u] : a.col[(size_t)(sb + k0 + u0 + u) * 64 + l];
                                c[u] = __shfl(cu, sl, 64);
                                v[u] = *reinterpret_cast<const bf16x8*>(
                                    a.h + (size_t)(c[u] >= 0 ? c[u] : row0 + sl) * HID + 8 * q);
                            }
#pragma unroll
                        for (int u = 0; u < KB / 2; ++u)
                            if (k0 + u0 + u < w) add_bf16x8(m, v[u], c[u] >= 0);
                    }
                }
                const int deg = __shfl(degl, sl, 64);
                const float inv = deg > 0 ? 1.0f / float(deg) : 0.0f;
                bf16x8 o;
#pragma unroll
                for (int t = 0; t < 8; ++t) o[t] = f2bf(m[t] * inv);
                *reinterpret_cast<bf16x8*>(tile + (8 * g + r8) * kStageLd + 8 * q) = o;
            }
        }
        __syncthreads();  // the mean tile (and, the first time, the staged parameters)
        if (live) {
            f32x4 acc[2][NT];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                bf16x8 xf[2];
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const int tl = 16 * mt + li;
                    if (kc < 2) {
                        xf[mt] = hres[mt][kc];
                    } else if (kc < 4) {
                        xf[mt] = *reinterpret_cast<const bf16x8*>(tile + tl * kStageLd + 32 * (kc - 2) + 8 * lg);
                    } else {
                        xf[mt] = xe[mt];
                    }
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const bf16x8 wf = wfrag[(kc * NT + nt) * 64 + l];
#pragma unroll
                    for (int mt = 0; mt < 2; ++mt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xf[mt], acc[mt][nt], 0, 0, 0);
                }
            }
            // lane (li, lg): acc[mt][nt][r] = output (row0 + 32 half + 16 mt + li,
            // column ws_out_col(nt, 4 lg + r))
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const int slane = 32 * half + 16 * mt + li, row = row0 + slane;
                bf16x8 o[2];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const float4 bb = bias_s[ws_out_col(nt, 4 * lg) / 4];
                    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float v = fmaxf(acc[mt][nt][r] + bv[r], 0.0f);
                        v += bf2f(hres[mt][nt >> 1][4 * (nt & 1) + r]);
                        o[nt >> 1][4 * (nt & 1) + r] = f2bf(v);
                    }
                }
                if constexpr (LAST) {
                    float dsum[kDirCols] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int q = 0; q < kDirCols; ++q)
#pragma unroll
                        for (int hh = 0; hh < 2; ++hh) {
                            const bf16x8 wq = wd_s[(q * HID + 32 * hh + 8 * lg) / 8];
#pragma unroll
                            for (int t = 0; t < 8; ++t) dsum[q] = fmaf(bf2f(o[hh][t]), bf2f(wq[t]), dsum[q]);
                        }
#pragma unroll
                    for (int q = 0; q < kDirCols; ++q) {
                        float t = dsum[q];
                        t += __shfl_xor(t, 16, 64);
                        t += __shfl_xor(t, 32, 64);
                        dsum[q] = t + bd_s[q];
                    }
                    if (lg == 0) {
                        if (a.dir != nullptr) {
                            const SliceDesc sd = a.slices[s];
                            const bool valid = slane < sd.nvalid;
                            const float sc = float(a.wi[sd.inst].scale);
#pragma unroll
                            for (int q = 0; q < kDirCols; ++q) a.dir[3 * (size_t)row + q] = valid ? sc * dsum[q] : 0.0f;
                        } else {
#pragma unroll
                            for (int q = 0; q < kDirCols; ++q) a.out[(size_t)row * NOUTD + q] = dsum[q];
                        }
                    }
                } else {
                    // staged in the wave's tile (its mean rows are read) for row-contiguous stores
                    if (mt == 0) ws_wave_lds_sync();
                    bf16x8* t = reinterpret_cast<bf16x8*>(tile + (16 * mt + li) * kStageLd + 8 * lg);
                    t[0] = o[0];
                    t[4] = o[1];
                }
            }
            if constexpr (!LAST) {
                ws_wave_lds_sync();
                const int r8 = l >> 3, q = l & 7;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int tl = 8 * g + r8;
                    *reinterpret_cast<bf16x8*>(a.h_out + (size_t)(row0 + 32 * half + tl) * HID + 8 * q) =
                        *reinterpret_cast<const bf16x8*>(tile + tl * kStageLd + 8 * q);
                }
            }
        }
        __syncthreads();  // the next half rewrites the mean tile
    }
}

// Inference encoder: h0 = relu(F We + be) for the 64 rows of a slice per
// wave, MFMA operands swapped like the layers and the weight rows permuted the
// same way (ws_out_col), so that lane (li, lg) holds columns 8 lg .. 8 lg + 7
// and 32 + 8 lg .. 32 + 8 lg + 7 of its row: two 16-byte stores.
__global__ __launch_bounds__(256) void ws_encode_kernel(int nslices, const __bf16* __restrict__ F,
                                                        const __bf16* __restrict__ w, const float* __restrict__ bias,
                                                        __bf16* __restrict__ h) {
    constexpr int NT = HID / 16;
    const int s = ws_slice(nslices);
    if (s >= nslices) return;
    const int row0 = s * 64;
    const int l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
    bf16x8 wf[NT];
    float bv[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        wf[nt] = *reinterpret_cast<const bf16x8*>(w + (size_t)ws_out_col(nt, li) * FIN + 8 * lg);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ws_out_col(nt, 4 * lg));
        bv[nt][0] = bb.x;
        bv[nt][1] = bb.y;
        bv[nt][2] = bb.z;
        bv[nt][3] = bb.w;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int row = row0 + mt * 16 + li;
        const bf16x8 xf = *reinterpret_cast<const bf16x8*>(F + (size_t)row * FIN + 8 * lg);
        bf16x8 o[2];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[nt], xf, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) o[nt >> 1][4 * (nt & 1) + r] = f2bf(fmaxf(acc[r] + bv[nt][r], 0.0f));
        }
        bf16x8* dst = reinterpret_cast<bf16x8*>(h + (size_t)row * HID + 8 * lg);
        dst[0] = o[0];
        dst[4] = o[1];
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: part[wg] = sum over this workgroup's 32-row chunks of
// X^T dY (K x NOUT) and column sums of dY (bias gradient).  Chunks go to the
// 4 waves of a workgroup round-robin; waves combine in order through LDS, so
// the sum is deterministic for a given grid.
// ---------------------------------------------------------------------------
struct WgradArgs {
    int rows;
    const __bf16* seg[3];
    int segw[3];
    int nseg;
    const __bf16* dy;
    float* part;
};

template <int KT, int NT>
__global__ __launch_bounds__(256) void ws_wgrad_kernel(WgradArgs a) {
    constexpr int K = 16 * KT, NOUT = 16 * NT, OUTN = K * NOUT + NOUT;
    __shared__ __attribute__((aligned(16))) __bf16 xs[4][32][K];
    __shared__ __attribute__((aligned(16))) __bf16 ys[4][32][NOUT];
    __shared__ float red[OUTN];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
    f32x4 acc[KT][NT];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[kt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const int nchunks = a.rows / 32;
    for (int cg = blockIdx.x; cg * 4 < nchunks; cg += gridDim.x) {
        const int c = cg * 4 + wv;
        const bool live = c < nchunks;
        const int m0 = c * 32;
        // stage X (32 x K) and dY (32 x NOUT) of this wave's chunk
        for (int u = l; u < 32 * (K / 8); u += 64) {
            const int r = u / (K / 8), kk = (u % (K / 8)) * 8;
            bf16x8 val = {};
            if (live) {
                int base = 0;
                for (int sg = 0; sg < a.nseg; ++sg) {
                    if (kk < base + a.segw[sg]) {
                        val = *reinterpret_cast<const bf16x8*>(a.seg[sg] + (size_t)(m0 + r) * a.segw[sg] + (kk - base));
                        break;
                    }
                    base += a.segw[sg];
                }
            }
            *reinterpret_cast<bf16x8*>(&xs[wv][r][kk]) = val;
        }
        for (int u = l; u < 32 * (NOUT / 8); u += 64) {
            const int r = u / (NOUT / 8), kk = (u % (NOUT / 8)) * 8;
            bf16x8 val = {};
            if (live) val = *reinterpret_cast<const bf16x8*>(a.dy + (size_t)(m0 + r) * NOUT + kk);
            *reinterpret_cast<bf16x8*>(&ys[wv][r][kk]) = val;
        }
        __syncthreads();
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            bf16x8 bfr;
#pragma unroll
            for (int t = 0; t < 8; ++t) bfr[t] = ys[wv][8 * lg + t][nt * 16 + li];
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                bf16x8 afr;
#pragma unroll
                for (int t = 0; t < 8; ++t) afr[t] = xs[wv][8 * lg + t][kt * 16 + li];
                acc[kt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[kt][nt], 0, 0, 0);
            }
        }
        if (l < NOUT)
            for (int r = 0; r < 32; ++r) bsum += bf2f(ys[wv][r][l]);
        __syncthreads();
    }
    // combine the 4 waves in order
    for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
            for (int kt = 0; kt < KT; ++kt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int idx = (kt * 16 + 4 * lg + r) * NOUT + nt * 16 + li;
                        red[idx] = (w == 0 ? 0.0f : red[idx]) + acc[kt][nt][r];
                    }
            if (l < NOUT) red[K * NOUT + l] = (w == 0 ? 0.0f : red[K * NOUT + l]) + bsum;
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < OUTN; q += 256) a.part[(size_t)blockIdx.x * OUTN + q] = red[q];
}

// grad[woff + q] (q < K*NOUT) and grad[boff + n] from the workgroup partials,
// summed in workgroup order.
__global__ void ws_reduce_parts_kernel(const float* __restrict__ part, int nwg, int K, int NOUT, float* __restrict__ grad,
                                       int woff, int boff) {
    const int outn = K * NOUT + NOUT;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= outn) return;
    float s = 0.0f;
    for (int w = 0; w < nwg; ++w) s += part[(size_t)w * outn + q];
    if (q < K * NOUT) grad[woff + q] = s;
    else grad[boff + (q - K * NOUT)] = s;
}

// ---------------------------------------------------------------------------
// Graph kernels (one lane per row, one wave per 64-row slice).
// ---------------------------------------------------------------------------
// out_i = mean_{j in N(i)} h_j  (64 columns, BF16 in/out, FP32 sums)
__global__ __launch_bounds__(256) void ws_agg_kernel(WsGraph g, const __bf16* __restrict__ h, __bf16* __restrict__ out) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    float acc[HID];
#pragma unroll
    for (int q = 0; q < HID; ++q) acc[q] = 0.0f;
    const int sb = g.slice_ptr[s], w = g.slice_ptr[s + 1] - sb;
    const int* colp = g.col + (size_t)sb * 64 + lane;
    for (int k = 0; k < w; ++k) {
        const int c = colp[k * 64];
        if (c < 0) continue;
        const bf16x8* src = reinterpret_cast<const bf16x8*>(h + (size_t)c * HID);
#pragma unroll
        for (int u = 0; u < HID / 8; ++u) add_bf16x8(acc + 8 * u, src[u], true);
    }
    const int deg = g.row_deg[row];
    const float inv = deg > 0 ? 1.0f / float(deg) : 0.0f;
    bf16x8* dst = reinterpret_cast<bf16x8*>(out + (size_t)row * HID);
#pragma unroll
    for (int u = 0; u < HID / 8; ++u) {
        bf16x8 v;
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = f2bf(acc[8 * u + t] * inv);
        dst[u] = v;
    }
}

// Backward through a layer's input: dh_out_i = dh_in_i + dx_i[0:64]
//   + sum_{j in N(i)} dx_j[64:128] / deg_j   (adjoint of the mean aggregation)
__global__ __launch_bounds__(256) void ws_dh_kernel(WsGraph g, const float* __restrict__ dh_in,
                                                    const float* __restrict__ dx, float* __restrict__ dh_out) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    float acc[HID];
    const float4* a4 = reinterpret_cast<const float4*>(dh_in + (size_t)row * HID);
    const float4* x4 = reinterpret_cast<const float4*>(dx + (size_t)row * XL);
#pragma unroll
    for (int u = 0; u < HID / 4; ++u) {
        const float4 p = a4[u], q = x4[u];
        acc[4 * u] = p.x + q.x;
        acc[4 * u + 1] = p.y + q.y;
        acc[4 * u + 2] = p.z + q.z;
        acc[4 * u + 3] = p.w + q.w;
    }
    const int sb = g.slice_ptr[s], w = g.slice_ptr[s + 1] - sb;
    const int* colp = g.col + (size_t)sb * 64 + lane;
    for (int k = 0; k < w; ++k) {
        const int c = colp[k * 64];
        if (c < 0) continue;
        const int dc = g.row_deg[c];
        const float inv = dc > 0 ? 1.0f / float(dc) : 0.0f;
        const float4* y4 = reinterpret_cast<const float4*>(dx + (size_t)c * XL + HID);
#pragma unroll
        for (int u = 0; u < HID / 4; ++u) {
            const float4 q = y4[u];
            acc[4 * u] += q.x * inv;
            acc[4 * u + 1] += q.y * inv;
            acc[4 * u + 2] += q.z * inv;
            acc[4 * u + 3] += q.w * inv;
        }
    }
    float4* o4 = reinterpret_cast<float4*>(dh_out + (size_t)row * HID);
#pragma unroll
    for (int u = 0; u < HID / 4; ++u) o4[u] = make_float4(acc[4 * u], acc[4 * u + 1], acc[4 * u + 2], acc[4 * u + 3]);
}

// dpre = dh * (pre > 0), BF16 out.  n elements.
__global__ void ws_relu_grad_kernel(const float* __restrict__ dh, const __bf16* __restrict__ pre, __bf16* __restrict__ dpre,
                                    size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dpre[i] = f2bf(bf2f(pre[i]) > 0.0f ? dh[i] : 0.0f);
}

// Loss partials (per slice: sum of squared error over valid rows) and
// dOUT = 2 (out - target) / n_valid in the first 3 columns.
__global__ __launch_bounds__(256) void ws_loss_kernel(WsGraph g, const float* __restrict__ out,
                                                      const float* __restrict__ target, float inv_nv,
                                                      __bf16* __restrict__ dout, double* __restrict__ lpart) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    const bool valid = lane < g.slices[s].nvalid;
    double e2 = 0.0;
    for (int c = 0; c < NOUTD; ++c) {
        float d = 0.0f;
        if (valid && c < 3) {
            const float err = out[(size_t)row * NOUTD + c] - target[3 * row + c];
            e2 += double(err) * err;
            d = 2.0f * err * inv_nv;
        }
        dout[(size_t)row * NOUTD + c] = f2bf(d);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
    if (lane == 0) lpart[s] = e2;
}

__global__ void ws_adam_kernel(float* __restrict__ th, const float* __restrict__ gr, float* __restrict__ m,
                               float* __restrict__ v, int n, float lr, float b1c, float b2c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
}

// BF16 copies of every weight matrix: wkn = (K, NOUT) as stored, wnk = transposed.
__global__ void ws_bf16_kernel(const float* __restrict__ th, __bf16* __restrict__ wnk, __bf16* __restrict__ wkn,
                               int off, int K, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * N) return;
    const int k = i / N, n = i % N;
    const __bf16 b = f2bf(th[off + i]);
    wkn[off + i] = b;
    wnk[off + n * K + k] = b;
}

unsigned grid_rows(int rows) { return ws_grid(rows / 64); }

__global__ void ws_to_bf16_kernel(const float* __restrict__ in, __bf16* __restrict__ out, size_t n_in, size_t n_out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    out[i] = f2bf(i < n_in ? in[i] : 0.0f);
}

template <int NT>
void run_gemm(const GemmArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(ws_gemm_kernel<NT>, dim3(grid_rows(a.rows)), dim3(256), 0, s, a);
    check_hip(hipGetLastError(), "ws_gemm_kernel");
}

constexpr int kWgradBlocks = 240;

template <int KT, int NT>
void run_wgrad(const WgradArgs& a, float* grad, int woff, int boff, hipStream_t s) {
    const int nchunk = a.rows / 32;
    const int nwg = std::max(1, std::min(kWgradBlocks, (nchunk + 3) / 4));
    hipLaunchKernelGGL((ws_wgrad_kernel<KT, NT>), dim3(nwg), dim3(256), 0, s, a);
    check_hip(hipGetLastError(), "ws_wgrad_kernel");
    const int outn = 16 * KT * 16 * NT + 16 * NT;
    hipLaunchKernelGGL(ws_reduce_parts_kernel, dim3((outn + 255) / 256), dim3(256), 0, s, a.part, nwg, 16 * KT,
                       16 * NT, grad, woff, boff);
    check_hip(hipGetLastError(), "ws_reduce_parts_kernel");
}

template <class T>
void dalloc(T*& p, size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (n) check_hip(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)), "hipMalloc(warm start)");
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// WarmStartNet
// ---------------------------------------------------------------------------
int WarmStartNet::num_params() { return NPARAM; }

WarmStartNet::WarmStartNet() {
    dalloc(theta_, NPARAM);
    dalloc(grad_, NPARAM);
    dalloc(m_, NPARAM);
    dalloc(v_, NPARAM);
    dalloc(wnk_, NPARAM);
    dalloc(wkn_, NPARAM);
    check_hip(hipMemset(theta_, 0, NPARAM * sizeof(float)), "hipMemset");
    check_hip(hipMemset(grad_, 0, NPARAM * sizeof(float)), "hipMemset");
    check_hip(hipMemset(m_, 0, NPARAM * sizeof(float)), "hipMemset");
    check_hip(hipMemset(v_, 0, NPARAM * sizeof(float)), "hipMemset");
    check_hip(hipMemset(wnk_, 0, NPARAM * sizeof(__bf16)), "hipMemset");
    check_hip(hipMemset(wkn_, 0, NPARAM * sizeof(__bf16)), "hipMemset");
}

WarmStartNet::~WarmStartNet() {
    free_all();
    dfree(theta_);
    dfree(grad_);
    dfree(m_);
    dfree(v_);
    dfree(wnk_);
    dfree(wkn_);
}

void WarmStartNet::free_all() {
    dfree(F_);
    dfree(E_);
    dfree(E4_);
    for (auto& p : H_) dfree(p);
    for (auto& p : A_) dfree(p);
    for (auto& p : P_) dfree(p);
    dfree(out_);
    dfree(dout_);
    dfree(dh_);
    dfree(dh2_);
    dfree(dpre_);
    dfree(dx_);
    dfree(part_);
    dfree(lpart_);
    rows_ = 0;
    training_ = false;
}

void WarmStartNet::reserve(int rows, bool training) {
    if (rows % 64) throw std::invalid_argument("warm start: rows must be a multiple of 64");
    if (rows <= rows_ && (training_ || !training)) return;
    free_all();
    const size_t R = size_t(rows);
    dalloc(F_, R * FIN);
    dalloc(E_, R * 32);
    dalloc(E4_, R * 4);
    for (auto& p : H_) dalloc(p, R * HID);
    for (auto& p : A_) dalloc(p, R * HID);
    for (auto& p : P_) dalloc(p, R * HID);
    dalloc(out_, R * NOUTD);
    check_hip(hipMemset(F_, 0, R * FIN * sizeof(__bf16)), "hipMemset");
    check_hip(hipMemset(E_, 0, R * 32 * sizeof(__bf16)), "hipMemset");
    check_hip(hipMemset(E4_, 0, R * 4 * sizeof(__bf16)), "hipMemset");
    if (training) {
        dalloc(dout_, R * NOUTD);
        dalloc(dh_, R * HID);
        dalloc(dh2_, R * HID);
        dalloc(dpre_, R * HID);
        dalloc(dx_, R * XL);
        dalloc(part_, size_t(kWgradBlocks) * (XL * HID + HID));
        dalloc(lpart_, R / 64);
        lpart_n_ = int(R / 64);
    }
    rows_ = rows;
    training_ = training;
}

void WarmStartNet::refresh_bf16(hipStream_t s) {
    for (const MatDesc& d : kMats) {
        hipLaunchKernelGGL(ws_bf16_kernel, dim3((d.K * d.N + 255) / 256), dim3(256), 0, s, theta_, wnk_, wkn_, d.off,
                           d.K, d.N);
        check_hip(hipGetLastError(), "ws_bf16_kernel");
    }
}

void WarmStartNet::set_params(const float* theta, hipStream_t s) {
    check_hip(hipMemcpyAsync(theta_, theta, NPARAM * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    check_hip(hipMemsetAsync(m_, 0, NPARAM * sizeof(float), s), "hipMemsetAsync");
    check_hip(hipMemsetAsync(v_, 0, NPARAM * sizeof(float), s), "hipMemsetAsync");
    adam_t_ = 0;
    refresh_bf16(s);
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    has_params_ = true;
}

void WarmStartNet::get_params(float* theta) const {
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    check_hip(hipMemcpy(theta, theta_, NPARAM * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

void WarmStartNet::init_params(uint64_t seed, hipStream_t s) {
    std::vector<float> th(NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    for (const MatDesc& d : kMats) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / float(d.K)));
        const float sc = (d.off == OFF_WD) ? 0.1f : 1.0f;
        for (int i = 0; i < d.K * d.N; ++i) th[size_t(d.off + i)] = sc * nd(rng);
    }
    set_params(th.data(), s);
}

void WarmStartNet::set_inputs(const float* F, const float* E, int n, hipStream_t s) {
    if (n > rows_) throw std::invalid_argument("warm start: reserve() rows first");
    compact_edges_ = false;
    float* tmp = nullptr;
    dalloc(tmp, size_t(n) * 32);
    const size_t nin = size_t(n) * 32, nout = size_t(rows_) * 32;
    for (int which = 0; which < 2; ++which) {
        check_hip(hipMemcpyAsync(tmp, which ? E : F, nin * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        hipLaunchKernelGGL(ws_to_bf16_kernel, dim3(unsigned((nout + 255) / 256)), dim3(256), 0, s, tmp,
                           which ? E_ : F_, nin, nout);
        check_hip(hipGetLastError(), "ws_to_bf16_kernel");
    }
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    dfree(tmp);
}

void WarmStartNet::copy_out(float* host, int n, hipStream_t s) const {
    check_hip(hipMemcpyAsync(host, out_, sizeof(float) * size_t(n) * NOUTD, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(out)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void WarmStartNet::copy_grad(float* host, hipStream_t s) const {
    check_hip(hipMemcpyAsync(host, grad_, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void WarmStartNet::forward(const WsGraph& g, hipStream_t s, bool keep, bool direction_only, const WsInst* wi,
                           float* dir, bool encoded) {
    const int rows = g.nslices * 64;
    if (rows > rows_) throw std::invalid_argument("warm start: reserve() rows first");
    if (keep && !training_) throw std::invalid_argument("warm start: reserve(training) before a kept forward");
    if (keep && encoded) throw std::invalid_argument("warm start: a kept forward runs its own encoder");
    // Pre-activations are only needed by backward().
    if (encoded) {
        // h0 written by features_encode
    } else if (!keep) {
        // inference encoder, vector stores
        hipLaunchKernelGGL(ws_encode_kernel, dim3(grid_rows(rows)), dim3(256), 0, s, g.nslices, F_, wnk_ + OFF_WE,
                           theta_ + OFF_BE, H_[0]);
        check_hip(hipGetLastError(), "ws_encode_kernel");
    } else {
        GemmArgs a{};
        a.rows = rows;
        a.seg[0] = F_;
        a.segw[0] = FIN;
        a.nseg = 1;
        a.K = FIN;
        a.wnk = wnk_ + OFF_WE;
        a.bias = theta_ + OFF_BE;
        a.relu = 1;
        a.out_b = H_[0];
        a.pre_b = P_[0];
        run_gemm<HID / 16>(a, s);
    }
    if (!keep) {
        // inference: each layer in one fused kernel; with direction_only the
        // decoder runs in the last one's epilogue (output columns 0..2, the
        // ones the warm start uses), otherwise as a GEMM below
        for (int l = 0; l < NL; ++l) {
            LayerArgs la{};
            la.nslices = g.nslices;
            la.slices = g.slices;
            la.slice_ptr = g.slice_ptr;
            la.col = g.col;
            la.row_deg = g.row_deg;
            la.h = H_[l];
            la.e = E_;
            la.e4 = compact_edges_ ? E4_ : nullptr;
            la.w = wnk_ + off_w(l);
            la.bias = theta_ + off_b(l);
            la.h_out = H_[l + 1];
            la.wd = wnk_ + OFF_WD;
            la.bd = theta_ + OFF_BD;
            la.out = out_;
            la.wi = wi;
            la.dir = (direction_only && wi != nullptr) ? dir : nullptr;
            if (l + 1 < NL || !direction_only)
                hipLaunchKernelGGL(ws_layer_kernel<false>, dim3(grid_rows(rows)), dim3(256), 0, s, la);
            else
                hipLaunchKernelGGL(ws_layer_kernel<true>, dim3(grid_rows(rows)), dim3(256), 0, s, la);
            check_hip(hipGetLastError(), "ws_layer_kernel");
        }
        if (direction_only) return;
    } else {
        for (int l = 0; l < NL; ++l) {
            hipLaunchKernelGGL(ws_agg_kernel, dim3(ws_grid(g.nslices)), dim3(256), 0, s, g, H_[l], A_[l]);
            check_hip(hipGetLastError(), "ws_agg_kernel");
            GemmArgs b{};
            b.rows = rows;
            b.seg[0] = H_[l];
            b.segw[0] = HID;
            b.seg[1] = A_[l];
            b.segw[1] = HID;
            b.seg[2] = E_;
            b.segw[2] = 32;
            b.nseg = 3;
            b.K = XL;
            b.wnk = wnk_ + off_w(l);
            b.bias = theta_ + off_b(l);
            b.relu = 1;
            b.resid = H_[l];
            b.out_b = H_[l + 1];
            b.pre_b = keep ? P_[l + 1] : nullptr;
            run_gemm<HID / 16>(b, s);
        }
    }
    GemmArgs d{};
    d.rows = rows;
    d.seg[0] = H_[NL];
    d.segw[0] = HID;
    d.nseg = 1;
    d.K = HID;
    d.wnk = wnk_ + OFF_WD;
    d.bias = theta_ + OFF_BD;
    d.out_f = out_;
    run_gemm<NOUTD / 16>(d, s);
}

void WarmStartNet::loss(const WsGraph& g, const float* target, hipStream_t s) {
    if (!training_) throw std::invalid_argument("warm start: reserve(training) first");
    // number of valid rows
    std::vector<SliceDesc> sd(static_cast<size_t>(g.nslices));
    check_hip(hipMemcpyAsync(sd.data(), g.slices, sizeof(SliceDesc) * g.nslices, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(slices)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    long long nv = 0;
    for (const auto& q : sd) nv += q.nvalid;
    hipLaunchKernelGGL(ws_loss_kernel, dim3(ws_grid(g.nslices)), dim3(256), 0, s, g, out_, target,
                       1.0f / float(std::max<long long>(nv, 1)), dout_, lpart_);
    check_hip(hipGetLastError(), "ws_loss_kernel");
    lpart_n_ = g.nslices;
    loss_nv_ = std::max<long long>(nv, 1);
}

double WarmStartNet::read_loss(hipStream_t s) {
    std::vector<double> h(static_cast<size_t>(lpart_n_));
    check_hip(hipMemcpyAsync(h.data(), lpart_, sizeof(double) * lpart_n_, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(loss)");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    double t = 0.0;
    for (double v : h) t += v;
    return t / double(loss_nv_);
}

void WarmStartNet::backward(const WsGraph& g, hipStream_t s) {
    if (!training_) throw std::invalid_argument("warm start: reserve(training) first");
    const int rows = g.nslices * 64;
    // decoder: dWd, dbd from H3 and dOUT; dH3 = dOUT Wd^T
    {
        WgradArgs w{};
        w.rows = rows;
        w.seg[0] = H_[NL];
        w.segw[0] = HID;
        w.nseg = 1;
        w.dy = dout_;
        w.part = part_;
        run_wgrad<HID / 16, NOUTD / 16>(w, grad_, OFF_WD, OFF_BD, s);
        GemmArgs a{};
        a.rows = rows;
        a.seg[0] = dout_;
        a.segw[0] = NOUTD;
        a.nseg = 1;
        a.K = NOUTD;
        a.wnk = wkn_ + OFF_WD;  // (HID, NOUTD) = (NOUT of this product, K)
        a.out_f = dh_;
        run_gemm<HID / 16>(a, s);
    }
    float* dh = dh_;
    float* dh_next = dh2_;
    const size_t nel = size_t(rows) * HID;
    for (int l = NL - 1; l >= 0; --l) {
        hipLaunchKernelGGL(ws_relu_grad_kernel, dim3(unsigned((nel + 255) / 256)), dim3(256), 0, s, dh, P_[l + 1],
                           dpre_, nel);
        check_hip(hipGetLastError(), "ws_relu_grad_kernel");
        WgradArgs w{};
        w.rows = rows;
        w.seg[0] = H_[l];
        w.segw[0] = HID;
        w.seg[1] = A_[l];
        w.segw[1] = HID;
        w.seg[2] = E_;
        w.segw[2] = 32;
        w.nseg = 3;
        w.dy = dpre_;
        w.part = part_;
        run_wgrad<XL / 16, HID / 16>(w, grad_, off_w(l), off_b(l), s);
        GemmArgs a{};
        a.rows = rows;
        a.seg[0] = dpre_;
        a.segw[0] = HID;
        a.nseg = 1;
        a.K = HID;
        a.wnk = wkn_ + off_w(l);  // (XL, HID)
        a.out_f = dx_;
        run_gemm<XL / 16>(a, s);
        hipLaunchKernelGGL(ws_dh_kernel, dim3(ws_grid(g.nslices)), dim3(256), 0, s, g, dh, dx_, dh_next);
        check_hip(hipGetLastError(), "ws_dh_kernel");
        std::swap(dh, dh_next);
    }
    hipLaunchKernelGGL(ws_relu_grad_kernel, dim3(unsigned((nel + 255) / 256)), dim3(256), 0, s, dh, P_[0], dpre_, nel);
    check_hip(hipGetLastError(), "ws_relu_grad_kernel");
    WgradArgs w{};
    w.rows = rows;
    w.seg[0] = F_;
    w.segw[0] = FIN;
    w.nseg = 1;
    w.dy = dpre_;
    w.part = part_;
    run_wgrad<FIN / 16, HID / 16>(w, grad_, OFF_WE, OFF_BE, s);
}

void WarmStartNet::adam_step(float lr, hipStream_t s) {
    ++adam_t_;
    const float b1c = 1.0f - std::pow(0.9f, float(adam_t_));
    const float b2c = 1.0f - std::pow(0.999f, float(adam_t_));
    hipLaunchKernelGGL(ws_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, theta_, grad_, m_, v_, NPARAM, lr,
                       b1c, b2c);
    check_hip(hipGetLastError(), "ws_adam_kernel");
    refresh_bf16(s);
}

// ---------------------------------------------------------------------------
// Simulator-side kernels
// ---------------------------------------------------------------------------

// Cyclic Jacobi eigen-decomposition of the symmetric k x k matrix G (k <= 8)
// and the eigen-truncated solve c = sum_{lam > 1e-12 lam_max} (u^T g / lam) u.
// The directions dropped are those where the history vectors are numerically
// collinear.  Returns the number of directions kept.
__host__ __device__ inline int ws_small_solve(int k, const double* G, const double* g, double* c) {
    double a[kWsHistMax][kWsHistMax], v[kWsHistMax][kWsHistMax];
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            a[i][j] = 0.5 * (G[i * k + j] + G[j * k + i]);
            v[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int p = 0; p < k; ++p) {
            dia += a[p][p] * a[p][p];
            for (int q = p + 1; q < k; ++q) off += a[p][q] * a[p][q];
        }
        if (!(off > 1e-30 * dia)) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                                      : (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < k; ++r) {
                    const double arp = a[r][p], arq = a[r][q];
                    a[r][p] = cs * arp - sn * arq;
                    a[r][q] = sn * arp + cs * arq;
                }
                for (int r = 0; r < k; ++r) {
                    const double apr = a[p][r], aqr = a[q][r];
                    a[p][r] = cs * apr - sn * aqr;
                    a[q][r] = sn * apr + cs * aqr;
                }
                for (int r = 0; r < k; ++r) {
                    const double vrp = v[r][p], vrq = v[r][q];
                    v[r][p] = cs * vrp - sn * vrq;
                    v[r][q] = sn * vrp + cs * vrq;
                }
            }
    }
    double lmax = 0.0;
    for (int i = 0; i < k; ++i) {
        c[i] = 0.0;
        lmax = fmax(lmax, a[i][i]);
    }
    int kept = 0;
    if (!(lmax > 0.0)) return 0;
    for (int m = 0; m < k; ++m) {
        const double lam = a[m][m];
        if (!(lam > 1e-12 * lmax)) continue;
        double pr = 0.0;
        for (int i = 0; i < k; ++i) pr += v[i][m] * g[i];
        const double cf = pr / lam;
        for (int i = 0; i < k; ++i) c[i] += cf * v[i][m];
        ++kept;
    }
    return kept;
}

// ws_small_solve with k fixed at compile time (every loop unrolled, the
// matrices in registers): the same operations in the same order.
// Index of entry (i, j), i <= j, of a K x K symmetric matrix packed as its
// upper triangle row by row.
template <int K>
__device__ __forceinline__ constexpr int ws_tri(int i, int j) {
    return i <= j ? i * K - i * (i - 1) / 2 + (j - i) : j * K - j * (j - 1) / 2 + (i - j);
}

int ws_small_solve_host(int k, const double* G, const double* g, double* c) {
    if (k < 1 || k > kWsHistMax) throw std::invalid_argument("ws_small_solve: 1 <= k <= 8");
    return ws_small_solve(k, G, g, c);
}

namespace {

// Entry order of the symmetric storage's products: 0, w-1, 1, w-2, ...  A
// lower entry's block is its owner row's upper entry; the owner (on a grid,
// a row W earlier, in a wavefront that runs a little earlier on the same XCD)
// streams that entry at the same step from its end, so the mirror read hits
// L2 (in ascending order the owner reads it last, long after).
__device__ __forceinline__ int ws_entry(const DevMatrix& A, int k, int w) {
    return A.sym ? ((k & 1) ? w - 1 - (k >> 1) : (k >> 1)) : k;
}

struct Flt3 {
    float a, b, c;
};

__device__ __forceinline__ void ld3f(const float* p, int i, float& a, float& b, float& c) {
    a = p[3 * i];
    b = p[3 * i + 1];
    c = p[3 * i + 2];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Wave-sum `val` and store it as value `o` of slice `s`.
__device__ __forceinline__ void put_partial(const WsReduce& red, int o, int s, int lane, double val) {
    const double t = wave_sum(val);
    if (lane == 0) red.part[(size_t)o * red.ldp + s] = t;
}

// Gram sums of K history vectors: per row y_j = (A v_j)_row in FP64, then
// [V^T A V (packed i <= j), (AV)^T (AV) (packed), V^T b, (AV)^T b, b^T b].
template <int K>
__global__ __launch_bounds__(256) void ws_gram_kernel(LaunchGeom g, DevMatrix A, const float* __restrict__ rhs,
                                                      WsHist h, WsReduce red) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    double y[K][3];
    float vr[K][3];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            y[j][c] = 0.0;
            vr[j][c] = 0.0f;
        }
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    if (lane < sd.nvalid) {
        float dm[9];
        load_diag(A, s, lane, dm);
        double m[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = double(dm[q]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            ld3f(h.v[j], row, vr[j][0], vr[j][1], vr[j][2]);
            const double x0 = vr[j][0], x1 = vr[j][1], x2 = vr[j][2];
            y[j][0] = m[0] * x0 + m[1] * x1 + m[2] * x2;
            y[j][1] = m[3] * x0 + m[4] * x1 + m[5] * x2;
            y[j][2] = m[6] * x0 + m[7] * x1 + m[8] * x2;
        }
        const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
        const int* colp = A.col + (size_t)sb * 64 + lane;
        const int* mirp = A.mir + (size_t)sb * 64 + lane;
        const float* vp = A.val + (size_t)sb * A.bs * 64 + lane;
        for (int t = 0; t < w; ++t) {
            const int k = ws_entry(A, t, w);
            const int c = colp[k * 64];
            if (c < 0) continue;
            float mf[9];
            load_block(A, vp + k * A.bs * 64, mirp + k * 64, c, row, mf);
#pragma unroll
            for (int q = 0; q < 9; ++q) m[q] = double(mf[q]);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                float u0, u1, u2;
                ld3f(h.v[j], c, u0, u1, u2);
                y[j][0] += m[0] * double(u0) + m[1] * double(u1) + m[2] * double(u2);
                y[j][1] += m[3] * double(u0) + m[4] * double(u1) + m[5] * double(u2);
                y[j][2] += m[6] * double(u0) + m[7] * double(u1) + m[8] * double(u2);
            }
        }
        b0 = rhs[3 * row];
        b1 = rhs[3 * row + 1];
        b2 = rhs[3 * row + 2];
    }
    int o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
            put_partial(red, o++, s, lane,
                        double(vr[i][0]) * y[j][0] + double(vr[i][1]) * y[j][1] + double(vr[i][2]) * y[j][2]);
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
            put_partial(red, o++, s, lane, y[i][0] * y[j][0] + y[i][1] * y[j][1] + y[i][2] * y[j][2]);
#pragma unroll
    for (int i = 0; i < K; ++i)
        put_partial(red, o++, s, lane, double(vr[i][0]) * b0 + double(vr[i][1]) * b1 + double(vr[i][2]) * b2);
#pragma unroll
    for (int i = 0; i < K; ++i) put_partial(red, o++, s, lane, y[i][0] * b0 + y[i][1] * b1 + y[i][2] * b2);
    put_partial(red, o++, s, lane, b0 * b0 + b1 * b1 + b2 * b2);
}

// Sums of `nv` per-slice partials over chunk blockIdx.y of instance
// blockIdx.x; fixed order, so the result does not depend on timing.
// One workgroup per (instance, chunk of kWsChunkSlices slices, value): thread
// t adds slices t, t + 256, ... of the chunk in order, then the wave sums and
// the four waves in order, so every value's sum is fixed whatever the launch.
__global__ __launch_bounds__(256) void ws_chunk_kernel(WsReduce red, int nv, const InstState* __restrict__ st) {
    const int inst = blockIdx.x, c = blockIdx.y, v = blockIdx.z;
    const int s0 = st[inst].slice0, ns = st[inst].nslices;
    const int a = c * kWsChunkSlices, e = min(ns, a + kWsChunkSlices);
    const int wv = int(threadIdx.x >> 6), lane = threadIdx.x & 63;
    __shared__ double sh[4];
    double acc = 0.0;
    const double* __restrict__ pv = red.part + (size_t)v * red.ldp + s0;
    for (int q = a + int(threadIdx.x); q < e; q += 256) acc += pv[q];
    acc = wave_sum(acc);
    if (lane == 0) sh[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) red.part2[((size_t)inst * red.chunks + c) * nv + v] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// The eigen-truncated solve of ws_small_solve by one wavefront: lane
// (i, j) = (lane >> 3, lane & 7) holds a[i][j] and v[i][j] (K <= 8), and the
// Jacobi sweeps rotate K' / 2 disjoint pairs at once (K' = K rounded up to
// even; a round-robin tournament pairs every index with every other once per
// sweep of K' - 1 rounds).  Rotations of disjoint pairs commute and each
// takes its angle from its own 2 x 2 block, which the others leave alone, so
// a round equals those rotations applied one after the other; the angle
// formula and the stopping test are those of ws_small_solve.  Every lane of
// the wavefront calls it; c[0..K-1] (shared memory) and the count of kept
// eigen-directions are the result.
template <int K>
__device__ __forceinline__ int ws_partner(int r, int i) {
    constexpr int KE = K + (K & 1);  // players; index K is a bye when K is odd
    if (i == KE - 1) return r;
    if (i == r) return KE - 1;
    return (2 * r - i + 2 * (KE - 1)) % (KE - 1);
}
__device__ __forceinline__ void ws_rotation(double app, double aqq, double apq, double& cs, double& sn) {
    if (apq == 0.0) {
        cs = 1.0;
        sn = 0.0;
        return;
    }
    const double theta = (aqq - app) / (2.0 * apq);
    const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                          : (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    cs = 1.0 / sqrt(t * t + 1.0);
    sn = t * cs;
}
__device__ __forceinline__ double ws_wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
template <int K>
__device__ int ws_small_solve_wave(const double* G, const double* g, double* c, int lane) {
    constexpr int KE = K + (K & 1);
    const int i = lane >> 3, j = lane & 7;
    const bool act = i < K && j < K;
    double a = act ? G[ws_tri<K>(i, j)] : 0.0;
    double v = (act && i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = ws_wave_sum((act && i < j) ? a * a : 0.0);
        const double dia = ws_wave_sum((act && i == j) ? a * a : 0.0);
        if (!(off > 1e-30 * dia)) break;
        for (int r = 0; r < KE - 1; ++r) {
            // this round's partners of the lane's row and column (K: a bye)
            const int pi = act ? ws_partner<K>(r, i) : i, pj = act ? ws_partner<K>(r, j) : j;
            const bool ri = pi < K, rj = pj < K;
            // rotation of the row's pair and of the column's pair
            const int p1 = ri ? min(i, pi) : i, q1 = ri ? max(i, pi) : i;
            const int p2 = rj ? min(j, pj) : j, q2 = rj ? max(j, pj) : j;
            const double a11 = __shfl(a, 9 * p1, 64), a12 = __shfl(a, 9 * q1, 64), a13 = __shfl(a, 8 * p1 + q1, 64);
            const double a21 = __shfl(a, 9 * p2, 64), a22 = __shfl(a, 9 * q2, 64), a23 = __shfl(a, 8 * p2 + q2, 64);
            double c1 = 1.0, s1 = 0.0, c2 = 1.0, s2 = 0.0;
            if (ri) ws_rotation(a11, a12, a13, c1, s1);
            if (rj) ws_rotation(a21, a22, a23, c2, s2);
            // J of a pair (p, q): J_pp = J_qq = c, J_pq = s, J_qp = -s
            const double jii = c1, jpi = ri ? (i == p1 ? -s1 : s1) : 0.0;
            const double jjj = c2, jpj = rj ? (j == p2 ? -s2 : s2) : 0.0;
            const double a_ipj = __shfl(a, 8 * i + pj, 64), a_pij = __shfl(a, 8 * pi + j, 64),
                         a_pipj = __shfl(a, 8 * pi + pj, 64), v_ipj = __shfl(v, 8 * i + pj, 64);
            if (act) {
                // A' = J^T A J, V' = V J
                a = jii * (jjj * a + jpj * a_ipj) + jpi * (jjj * a_pij + jpj * a_pipj);
                v = jjj * v + jpj * v_ipj;
            }
        }
    }
    // eigenvalue of column j, the largest one, and v_j . g
    const double lam = __shfl(a, 9 * j, 64);
    double lmax = act && i == j ? a : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off, 64));
    double pr = act ? v * g[i] : 0.0;
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) pr += __shfl_xor(pr, off, 64);
    const bool keep = act && lmax > 0.0 && lam > 1e-12 * lmax;
    // c_i = sum over kept columns j of (v_j . g / lambda_j) v_ij
    double ci = keep ? (pr / lam) * v : 0.0;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) ci += __shfl_xor(ci, off, 64);
    if (j == 0 && i < K) c[i] = ci;
    const double kept = ws_wave_sum((keep && i == 0) ? 1.0 : 0.0);
    return int(kept);
}

// The projection's decision for K history vectors from the packed sums in
// LDS and the coefficients c of the eigen-solve (kept directions: kept).
template <int K>
__device__ __forceinline__ void ws_finish_projection(const double* __restrict__ tot, const double* csol, int kept,
                                                     WsInst& W) {
    constexpr int NT = K * (K + 1) / 2;
    const double* G = tot;       // V^T A V, packed
    const double* Q = tot + NT;  // (AV)^T (AV), packed
    double gv[K], bav[K], c[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        gv[i] = tot[2 * NT + i];
        bav[i] = tot[2 * NT + K + i];
    }
    const double bb = tot[2 * NT + 2 * K];
#pragma unroll
    for (int i = 0; i < K; ++i) c[i] = csol[i];
    double phi = 0.0, rr = bb;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double gc = 0.0, qc = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            gc += G[ws_tri<K>(i, j)] * c[j];
            qc += Q[ws_tri<K>(i, j)] * c[j];
        }
        phi += c[i] * (0.5 * gc - gv[i]);
        rr += c[i] * (qc - 2.0 * bav[i]);
    }
    const double phi_prev = 0.5 * G[0] - gv[0];
    W.bb = bb;
    W.rr_prev = fmax(bb - 2.0 * bav[0] + Q[0], 0.0);
    W.rr_proj = fmax(rr, 0.0);
    W.kdim = kept;
    W.use_proj = (kept > 0 && phi < phi_prev) ? 1 : 0;
#pragma unroll
    for (int i = 0; i < kWsHistMax; ++i) W.coef[i] = (i < K) ? c[i] : 0.0;
}

__global__ __launch_bounds__(128) void ws_finish_kernel(WsReduce red, int nv, const InstState* __restrict__ st,
                                                        WsInst* __restrict__ wi, int phase, int hk, double inv_ninst) {
    const int inst = blockIdx.x;
    __shared__ double tot[kWsMaxValues];
    const int t = threadIdx.x;
    if (t < nv) {
        double a = 0.0;
        for (int c = 0; c < red.chunks; ++c) a += red.part2[((size_t)inst * red.chunks + c) * nv + t];
        tot[t] = a;
    }
    __syncthreads();
    if (t != 0) return;
    WsInst& W = wi[inst];
    if (phase == kWsPhaseProj) {
        return;  // ws_finish_proj_kernel
    } else if (phase == kWsPhaseBase) {
        W.rr_base = tot[0];
        W.scale = fmax(sqrt(tot[1] / (3.0 * 64.0 * double(max(1, st[inst].nslices)))), 1e-12);
        W.bb = tot[2];
        if (!hk) {  // no projection this step: the base guess is dx_prev
            W.rr_prev = tot[0];
            W.use_proj = 0;
        }
        W.use_cand = 0;
        W.alpha = 0.0;
    } else if (phase == kWsPhaseLine) {
        const double u = tot[0], w = tot[1], q = tot[2], a2 = tot[3];
        W.u = u;
        W.w = w;
        const double al = (w > 0.0 && isfinite(u / w)) ? u / w : 0.0;
        const double rr = W.rr_base - 2.0 * al * q + al * al * a2;
        W.rr_cand = fmax(rr, 0.0);
        // hk == 0 (default): accept when the energy x^T A x / 2 - b^T x
        // drops, i.e. the residual measured in the A^-1 norm (the A-norm of
        // the error, what PCG minimises): the exact line search lowers it by
        // u^2 / (2 w) > 0 whenever d is not A-orthogonal to the error.
        // hk == 1: accept only when the predicted residual |r_b - alpha A d|
        // is below the base guess's (otherwise keep the base guess).
        W.use_cand = (al != 0.0 && (hk == 0 || W.rr_cand < W.rr_base)) ? 1 : 0;
        W.alpha = W.use_cand ? al : 0.0;
    } else {  // kWsPhaseEnergy
        const double E = tot[0], u = W.u, w = W.w;
        W.energy = E;
        if (w > 0.0 && E > 0.0) {
            W.loss = 1.0 - u * u / (w * E);
            W.gcoef = -2.0 * u / (w * E) * W.scale * inv_ninst;
        } else {
            W.loss = 0.0;
            W.gcoef = 0.0;
        }
    }
}

// The projection phase of ws_finish_kernel, one kernel per history length K
// (its eigen-solve held in registers without spilling the others).
template <int K>
__global__ __launch_bounds__(128) void ws_finish_proj_kernel(WsReduce red, int nv, WsInst* __restrict__ wi) {
    const int inst = blockIdx.x;
    __shared__ double tot[kWsMaxValues];
    const int t = threadIdx.x;
    if (t < nv) {
        double a = 0.0;
        for (int c = 0; c < red.chunks; ++c) a += red.part2[((size_t)inst * red.chunks + c) * nv + t];
        tot[t] = a;
    }
    __syncthreads();
    if (t >= 64) return;
    // the eigen-solve by the first wavefront, then the decision by its lane 0
    __shared__ double csol[K];
    constexpr int NT = K * (K + 1) / 2;
    const int kept = ws_small_solve_wave<K>(tot, tot + 2 * NT, csol, t);
    ws_wave_lds_sync();  // the lanes that wrote csol are in this wavefront
    if (t != 0) return;
    ws_finish_projection<K>(tot, csol, kept, wi[inst]);
}

__global__ __launch_bounds__(256) void ws_proj_apply_kernel(LaunchGeom g, WsHist h, const WsInst* __restrict__ wi,
                                                            double* __restrict__ x) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63;
    const WsInst& W = wi[sd.inst];
    if (!W.use_proj) return;
    // element-wise over the slice's 3 x 64 components, 64 consecutive ones per
    // access (x[3 row + c] is component 3 (row - 64 s) + c of the slice)
    const size_t base = (size_t)s * 192;
    const int nv3 = 3 * sd.nvalid;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = lane + 64 * q;
        if (e >= nv3) break;
        double a = 0.0;
        for (int j = 0; j < h.k; ++j) a += W.coef[j] * double(h.v[j][base + e]);
        x[base + e] = a;
    }
}

// r_b = b - A x with FP64 products and sums (stored FP32), z = M^-1 r_b,
// xb <- x (unless xb is null); partials [|r_b|^2, |x|^2, |b|^2].
__global__ __launch_bounds__(256) void ws_base_kernel(LaunchGeom g, DevMatrix A, DevVectors v,
                                                      const double* __restrict__ x, double* __restrict__ xb,
                                                      float* __restrict__ rb, WsReduce red) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    double rr = 0.0, xx = 0.0, bb = 0.0;
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (lane < sd.nvalid) {
        float dm[9];
        load_diag(A, s, lane, dm);
        x0 = x[3 * row];
        x1 = x[3 * row + 1];
        x2 = x[3 * row + 2];
        double a0 = double(dm[0]) * x0 + double(dm[1]) * x1 + double(dm[2]) * x2;
        double a1 = double(dm[3]) * x0 + double(dm[4]) * x1 + double(dm[5]) * x2;
        double a2 = double(dm[6]) * x0 + double(dm[7]) * x1 + double(dm[8]) * x2;
        const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
        const int* colp = A.col + (size_t)sb * 64 + lane;
        const int* mirp = A.mir + (size_t)sb * 64 + lane;
        const float* vp = A.val + (size_t)sb * A.bs * 64 + lane;
        for (int t = 0; t < w; ++t) {
            const int k = ws_entry(A, t, w);
            const int c = colp[k * 64];
            if (c >= 0) {
                float m[9];
                load_block(A, vp + k * A.bs * 64, mirp + k * 64, c, row, m);
                const double y0 = x[3 * c], y1 = x[3 * c + 1], y2 = x[3 * c + 2];
                a0 += double(m[0]) * y0 + double(m[1]) * y1 + double(m[2]) * y2;
                a1 += double(m[3]) * y0 + double(m[4]) * y1 + double(m[5]) * y2;
                a2 += double(m[6]) * y0 + double(m[7]) * y1 + double(m[8]) * y2;
            }
        }
        const double q0 = double(v.rhs[3 * row]), q1 = double(v.rhs[3 * row + 1]), q2 = double(v.rhs[3 * row + 2]);
        const double r0 = q0 - a0, r1 = q1 - a1, r2 = q2 - a2;
        rr = r0 * r0 + r1 * r1 + r2 * r2;
        xx = x0 * x0 + x1 * x1 + x2 * x2;
        bb = q0 * q0 + q1 * q1 + q2 * q2;
        f0 = float(r0);
        f1 = float(r1);
        f2 = float(r2);
    }
    const float* mv = A.minv + (size_t)s * 6 * 64 + lane;
    const float m0 = mv[0], m1 = mv[64], m2 = mv[128], m3 = mv[192], m4 = mv[256], m5 = mv[320];
    const bool live = lane < sd.nvalid;
    // one 12-byte store per row and vector, so a wave's stores cover whole lines
    *reinterpret_cast<Flt3*>(rb + 3 * row) = Flt3{f0, f1, f2};
    *reinterpret_cast<Flt3*>(v.z + 3 * row) =
        live ? Flt3{m0 * f0 + m1 * f1 + m2 * f2, m1 * f0 + m3 * f1 + m4 * f2, m2 * f0 + m4 * f1 + m5 * f2}
             : Flt3{0.0f, 0.0f, 0.0f};
    if (xb != nullptr) {
        xb[3 * row] = x0;
        xb[3 * row + 1] = x1;
        xb[3 * row + 2] = x2;
    }
    put_partial(red, 0, s, lane, rr);
    put_partial(red, 1, s, lane, xx);
    put_partial(red, 2, s, lane, bb);
}

// Features (scaled by 1 / RMS(x_b)):
//   0-2 dx_prev, 3-5 dx_prev - dx_prev2, 6-8 z = M^-1 r_b, 9-11 x_b - dx_prev,
//   12 pinned, 13 inertia share of the diagonal, 14 constant 1.
// Edge features (mean over mesh neighbours): (x_j - x_i)/L0 (3), |x_j - x_i|/L0 - 1.
// Zeros for the rows of a slice beyond its valid count.
__device__ __forceinline__ void ws_row_features(const DevMatrix& A, const DevVectors& v, const DevExt& e,
                                                const DevMesh& m, const float* __restrict__ rest, const WsHist& h,
                                                const double* __restrict__ xb, const WsInst* __restrict__ wi,
                                                float dt, int s, const SliceDesc& sd, int lane, float (&f)[FIN],
                                                float (&ef)[4]) {
    const int row = s * 64 + lane;
#pragma unroll
    for (int q = 0; q < FIN; ++q) f[q] = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) ef[q] = 0.0f;
    if (lane >= sd.nvalid) return;
    const float is = float(1.0 / fmax(wi[sd.inst].scale, 1e-30));
    float h0, h1, h2, p0, p1, p2;
    ld3f(h.v[0], row, h0, h1, h2);
    ld3f(h.v[1], row, p0, p1, p2);
    f[0] = h0 * is;
    f[1] = h1 * is;
    f[2] = h2 * is;
    f[3] = (h0 - p0) * is;
    f[4] = (h1 - p1) * is;
    f[5] = (h2 - p2) * is;
    f[6] = v.z[3 * row] * is;
    f[7] = v.z[3 * row + 1] * is;
    f[8] = v.z[3 * row + 2] * is;
    f[9] = float(xb[3 * row] - double(h0)) * is;
    f[10] = float(xb[3 * row + 1] - double(h1)) * is;
    f[11] = float(xb[3 * row + 2] - double(h2)) * is;
    f[12] = (m.pin_inc_off[row + 1] > m.pin_inc_off[row]) ? 1.0f : 0.0f;
    const int p = sd.part0 + lane;
    float mass = 1.0f;
    if (e.inv_mass != nullptr) {
        const float w = e.inv_mass[p];
        mass = (w > 1.0e-12f) ? (1.0f / w) : 1.0e8f;
    }
    float dm[9];
    load_diag(A, s, lane, dm);
    const float tr = (dm[0] + dm[4] + dm[8]) * (1.0f / 3.0f);
    f[13] = (tr > 0.0f) ? fminf(mass / (dt * dt) / tr, 1.0f) : 0.0f;
    f[14] = 1.0f;
    float px, py, pz, rx, ry, rz;
    ld3f(e.pos, p, px, py, pz);
    ld3f(rest, p, rx, ry, rz);
    const int shift = sd.part0 - 64 * s;  // particle = row + shift within an instance
    const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
    const int* colp = A.col + (size_t)sb * 64 + lane;
    int cnt = 0;
    for (int k = 0; k < w; ++k) {
        const int c = colp[k * 64];
        if (c < 0) continue;
        const int pj = c + shift;
        float qx, qy, qz, sx, sy, sz;
        ld3f(e.pos, pj, qx, qy, qz);
        ld3f(rest, pj, sx, sy, sz);
        const float dx = qx - px, dy = qy - py, dz = qz - pz;
        const float l0 = sqrtf((sx - rx) * (sx - rx) + (sy - ry) * (sy - ry) + (sz - rz) * (sz - rz));
        const float il = l0 > 1e-12f ? 1.0f / l0 : 0.0f;
        ef[0] += dx * il;
        ef[1] += dy * il;
        ef[2] += dz * il;
        ef[3] += sqrtf(dx * dx + dy * dy + dz * dz) * il - 1.0f;
        ++cnt;
    }
    if (cnt) {
        const float ic = 1.0f / float(cnt);
#pragma unroll
        for (int q = 0; q < 4; ++q) ef[q] *= ic;
    }
}

// The features of every row into F (rows x 32) and the edge features into E4
// (rows x 4) when it is given, else into E (rows x 32, columns 4..31 zero).
__global__ __launch_bounds__(256) void ws_features_kernel(LaunchGeom g, DevMatrix A, DevVectors v, DevExt e,
                                                          DevMesh m, const float* __restrict__ rest, WsHist h,
                                                          const double* __restrict__ xb,
                                                          const WsInst* __restrict__ wi, float dt,
                                                          __bf16* __restrict__ F, __bf16* __restrict__ E,
                                                          __bf16* __restrict__ E4) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    float f[FIN], ef[4];
    ws_row_features(A, v, e, m, rest, h, xb, wi, dt, s, sd, lane, f, ef);
    bf16x8* fo = reinterpret_cast<bf16x8*>(F + (size_t)row * FIN);
#pragma unroll
    for (int u = 0; u < FIN / 8; ++u) {
        bf16x8 t;
#pragma unroll
        for (int q = 0; q < 8; ++q) t[q] = f2bf(f[8 * u + q]);
        fo[u] = t;
    }
    if (E4 != nullptr) {
        bf16x4 t;
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = f2bf(ef[q]);
        *reinterpret_cast<bf16x4*>(E4 + (size_t)row * 4) = t;
        return;
    }
    bf16x8* eo = reinterpret_cast<bf16x8*>(E + (size_t)row * 32);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        bf16x8 t;
#pragma unroll
        for (int q = 0; q < 8; ++q) t[q] = f2bf((u == 0 && q < 4) ? ef[q] : 0.0f);
        eo[u] = t;
    }
}

// Inference: the features of the wave's 64 rows go to an LDS tile instead of
// memory, and the encoder h0 = relu(F We + be) runs on them as in
// ws_encode_kernel (same operands, so bitwise its result); the compact edge
// features go to E4.  Each 16-row tile of h0 goes back through LDS, so a store
// instruction writes eight whole 128-B rows.
constexpr int kFeatLd = FIN + 8;  // LDS row stride (BF16) of the feature tile
__global__ __launch_bounds__(256) void ws_features_encode_kernel(LaunchGeom g, DevMatrix A, DevVectors v, DevExt e,
                                                                 DevMesh m, const float* __restrict__ rest,
                                                                 WsHist h, const double* __restrict__ xb,
                                                                 const WsInst* __restrict__ wi, float dt,
                                                                 const __bf16* __restrict__ w,
                                                                 const float* __restrict__ bias,
                                                                 __bf16* __restrict__ E4, __bf16* __restrict__ H0) {
    constexpr int NT = HID / 16;
    __shared__ __attribute__((aligned(16))) __bf16 ftile[4][64 * kFeatLd];
    __shared__ __attribute__((aligned(16))) __bf16 otile[4][16 * kStageLd];
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int l = threadIdx.x & 63, li = l & 15, lg = l >> 4, row0 = s * 64;
    __bf16* ft = ftile[threadIdx.x >> 6];
    __bf16* ot = otile[threadIdx.x >> 6];
    {
        float f[FIN], ef[4];
        ws_row_features(A, v, e, m, rest, h, xb, wi, dt, s, sd, l, f, ef);
#pragma unroll
        for (int u = 0; u < FIN / 8; ++u) {
            bf16x8 t;
#pragma unroll
            for (int q = 0; q < 8; ++q) t[q] = f2bf(f[8 * u + q]);
            *reinterpret_cast<bf16x8*>(ft + l * kFeatLd + 8 * u) = t;
        }
        bf16x4 t4;
#pragma unroll
        for (int q = 0; q < 4; ++q) t4[q] = f2bf(ef[q]);
        *reinterpret_cast<bf16x4*>(E4 + (size_t)(row0 + l) * 4) = t4;
    }
    bf16x8 wf[NT];
    float bv[NT][4];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        wf[nt] = *reinterpret_cast<const bf16x8*>(w + (size_t)ws_out_col(nt, li) * FIN + 8 * lg);
        const float4 bb = *reinterpret_cast<const float4*>(bias + ws_out_col(nt, 4 * lg));
        bv[nt][0] = bb.x;
        bv[nt][1] = bb.y;
        bv[nt][2] = bb.z;
        bv[nt][3] = bb.w;
    }
    ws_wave_lds_sync();
    const int r8 = l >> 3, q8 = l & 7;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const bf16x8 xf = *reinterpret_cast<const bf16x8*>(ft + (16 * mt + li) * kFeatLd + 8 * lg);
        bf16x8 o[2];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[nt], xf, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) o[nt >> 1][4 * (nt & 1) + r] = f2bf(fmaxf(acc[r] + bv[nt][r], 0.0f));
        }
        ws_wave_lds_sync();  // the previous tile's rows are read
        bf16x8* t = reinterpret_cast<bf16x8*>(ot + li * kStageLd + 8 * lg);
        t[0] = o[0];
        t[4] = o[1];
        ws_wave_lds_sync();
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) {
            const int tl = 8 * gq + r8;
            *reinterpret_cast<bf16x8*>(H0 + (size_t)(row0 + 16 * mt + tl) * HID + 8 * q8) =
                *reinterpret_cast<const bf16x8*>(ot + tl * kStageLd + 8 * q8);
        }
    }
}

// d = scale * out[:, 0:3] (FP32; padding rows 0).
__global__ __launch_bounds__(256) void ws_dir_write_kernel(LaunchGeom g, const float* __restrict__ out,
                                                           const WsInst* __restrict__ wi, float* __restrict__ d) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    const bool live = lane < sd.nvalid;
    const float sc = float(wi[sd.inst].scale);
#pragma unroll
    for (int c = 0; c < 3; ++c) d[3 * row + c] = live ? sc * out[(size_t)row * NOUTD + c] : 0.0f;
}

// A d (FP64 sums, stored FP32) and partials [d.r_b, d.Ad, Ad.r_b, Ad.Ad].
__global__ __launch_bounds__(256) void ws_dir_spmv_kernel(LaunchGeom g, DevMatrix A, const float* __restrict__ d,
                                                          const float* __restrict__ rb, float* __restrict__ ad,
                                                          WsReduce red) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (lane < sd.nvalid) {
        float dm[9];
        load_diag(A, s, lane, dm);
        d0 = d[3 * row];
        d1 = d[3 * row + 1];
        d2 = d[3 * row + 2];
        a0 = double(dm[0]) * d0 + double(dm[1]) * d1 + double(dm[2]) * d2;
        a1 = double(dm[3]) * d0 + double(dm[4]) * d1 + double(dm[5]) * d2;
        a2 = double(dm[6]) * d0 + double(dm[7]) * d1 + double(dm[8]) * d2;
        const int sb = A.slice_ptr[s], w = A.slice_ptr[s + 1] - sb;
        const int* colp = A.col + (size_t)sb * 64 + lane;
        const int* mirp = A.mir + (size_t)sb * 64 + lane;
        const float* vp = A.val + (size_t)sb * A.bs * 64 + lane;
        for (int t = 0; t < w; ++t) {
            const int k = ws_entry(A, t, w);
            const int c = colp[k * 64];
            if (c >= 0) {
                float m[9];
                load_block(A, vp + k * A.bs * 64, mirp + k * 64, c, row, m);
                const double y0 = d[3 * c], y1 = d[3 * c + 1], y2 = d[3 * c + 2];
                a0 += double(m[0]) * y0 + double(m[1]) * y1 + double(m[2]) * y2;
                a1 += double(m[3]) * y0 + double(m[4]) * y1 + double(m[5]) * y2;
                a2 += double(m[6]) * y0 + double(m[7]) * y1 + double(m[8]) * y2;
            }
        }
        r0 = rb[3 * row];
        r1 = rb[3 * row + 1];
        r2 = rb[3 * row + 2];
    }
    ad[3 * row] = float(a0);
    ad[3 * row + 1] = float(a1);
    ad[3 * row + 2] = float(a2);
    const double f0 = float(a0), f1 = float(a1), f2 = float(a2);  // the stored A d
    put_partial(red, 0, s, lane, d0 * r0 + d1 * r1 + d2 * r2);
    put_partial(red, 1, s, lane, d0 * a0 + d1 * a1 + d2 * a2);
    put_partial(red, 2, s, lane, f0 * r0 + f1 * r1 + f2 * r2);
    put_partial(red, 3, s, lane, f0 * f0 + f1 * f1 + f2 * f2);
}

__global__ __launch_bounds__(256) void ws_line_apply_kernel(LaunchGeom g, const WsInst* __restrict__ wi,
                                                            const float* __restrict__ d, double* __restrict__ x,
                                                            const float* __restrict__ rb,
                                                            const float* __restrict__ ad, float* __restrict__ r) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63;
    const WsInst& W = wi[sd.inst];
    const bool use = W.use_cand != 0;
    const double al = use ? W.alpha : 0.0;
    // element-wise over the slice's 3 x 64 components, 64 consecutive ones per access
    const size_t base = (size_t)s * 192;
    const int nv3 = 3 * sd.nvalid;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int e = lane + 64 * q;
        if (e < nv3) {
            if (use) x[base + e] += al * double(d[base + e]);
            // the residual of the new guess, for the solve's start
            if (r != nullptr) r[base + e] = float(double(rb[base + e]) - al * double(ad[base + e]));
        }
    }
}

__global__ __launch_bounds__(256) void ws_energy_kernel(LaunchGeom g, const double* __restrict__ x,
                                                        const double* __restrict__ xb, const float* __restrict__ rb,
                                                        WsReduce red) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    double e = 0.0;
    if (lane < sd.nvalid)
        for (int c = 0; c < 3; ++c) e += (x[3 * row + c] - xb[3 * row + c]) * double(rb[3 * row + c]);
    put_partial(red, 0, s, lane, e);
}

__global__ __launch_bounds__(256) void ws_dout_kernel(LaunchGeom g, const WsInst* __restrict__ wi,
                                                      const float* __restrict__ rb, const float* __restrict__ ad,
                                                      __bf16* __restrict__ dout) {
    const int s = ws_slice(g.nslices);
    if (s >= g.nslices) return;
    const SliceDesc sd = g.slices[s];
    const int lane = threadIdx.x & 63, row = s * 64 + lane;
    float gr[3] = {0.0f, 0.0f, 0.0f};
    if (lane < sd.nvalid) {
        const WsInst& W = wi[sd.inst];
        const double gc = W.gcoef, uw = (W.w > 0.0) ? W.u / W.w : 0.0;
        for (int c = 0; c < 3; ++c) gr[c] = float(gc * (double(rb[3 * row + c]) - uw * double(ad[3 * row + c])));
    }
    bf16x8* o = reinterpret_cast<bf16x8*>(dout + (size_t)row * NOUTD);
#pragma unroll
    for (int u = 0; u < NOUTD / 8; ++u) {
        bf16x8 t;
#pragma unroll
        for (int q = 0; q < 8; ++q) t[q] = f2bf((u == 0 && q < 3) ? gr[q] : 0.0f);
        o[u] = t;
    }
}

__global__ void ws_history_kernel(const double* __restrict__ x, float* __restrict__ hist, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) hist[i] = float(x[i]);
}

inline unsigned grid_slices(int nslices) { return ws_grid(nslices); }

template <int K>
void gram_k(const LaunchGeom& g, const DevMatrix& A, const float* rhs, const WsHist& h, const WsReduce& red,
            hipStream_t s) {
    hipLaunchKernelGGL(ws_gram_kernel<K>, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, A, rhs, h, red);
}

}  // namespace

void launch_ws_gram(const LaunchGeom& g, const DevMatrix& A, const float* rhs, const WsHist& h, const WsReduce& red,
                    hipStream_t s) {
    if (g.nslices <= 0) return;
    switch (h.k) {
        case 2: gram_k<2>(g, A, rhs, h, red, s); break;
        case 3: gram_k<3>(g, A, rhs, h, red, s); break;
        case 4: gram_k<4>(g, A, rhs, h, red, s); break;
        case 5: gram_k<5>(g, A, rhs, h, red, s); break;
        case 6: gram_k<6>(g, A, rhs, h, red, s); break;
        case 7: gram_k<7>(g, A, rhs, h, red, s); break;
        case 8: gram_k<8>(g, A, rhs, h, red, s); break;
        default: throw std::invalid_argument("warm start: history length must be 2..8");
    }
    check_hip(hipGetLastError(), "ws_gram_kernel");
}

void launch_ws_finish(const WsReduce& red, int nvalues, const InstState* st, WsInst* wi, int ninst, int phase,
                      int hist_k, hipStream_t s) {
    if (ninst <= 0) return;
    if (nvalues < 1 || nvalues > kWsMaxValues) throw std::invalid_argument("ws_finish: bad value count");
    if (phase == kWsPhaseProj && (hist_k < 1 || hist_k > kWsHistMax))
        throw std::invalid_argument("ws_finish: the projection needs 1 to 8 history vectors");
    hipLaunchKernelGGL(ws_chunk_kernel, dim3(unsigned(ninst), unsigned(red.chunks), unsigned(nvalues)), dim3(256), 0, s,
                       red, nvalues, st);
    check_hip(hipGetLastError(), "ws_chunk_kernel");
    if (phase == kWsPhaseProj) {
        const dim3 grid{unsigned(ninst)}, block{128u};
        switch (hist_k) {
            case 1: hipLaunchKernelGGL(ws_finish_proj_kernel<1>, grid, block, 0, s, red, nvalues, wi); break;
            case 2: hipLaunchKernelGGL(ws_finish_proj_kernel<2>, grid, block, 0, s, red, nvalues, wi); break;
            case 3: hipLaunchKernelGGL(ws_finish_proj_kernel<3>, grid, block, 0, s, red, nvalues, wi); break;
            case 4: hipLaunchKernelGGL(ws_finish_proj_kernel<4>, grid, block, 0, s, red, nvalues, wi); break;
            case 5: hipLaunchKernelGGL(ws_finish_proj_kernel<5>, grid, block, 0, s, red, nvalues, wi); break;
            case 6: hipLaunchKernelGGL(ws_finish_proj_kernel<6>, grid, block, 0, s, red, nvalues, wi); break;
            case 7: hipLaunchKernelGGL(ws_finish_proj_kernel<7>, grid, block, 0, s, red, nvalues, wi); break;
            default: hipLaunchKernelGGL(ws_finish_proj_kernel<8>, grid, block, 0, s, red, nvalues, wi); break;
        }
        check_hip(hipGetLastError(), "ws_finish_proj_kernel");
        return;
    }
    hipLaunchKernelGGL(ws_finish_kernel, dim3(unsigned(ninst)), dim3(128), 0, s, red, nvalues, st, wi, phase, hist_k,
                       1.0 / double(ninst));
    check_hip(hipGetLastError(), "ws_finish_kernel");
}

void launch_ws_proj_apply(const LaunchGeom& g, const WsHist& h, const WsInst* wi, double* x, hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_proj_apply_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, h, wi, x);
    check_hip(hipGetLastError(), "ws_proj_apply_kernel");
}

void launch_ws_base(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const double* x, double* xb,
                    float* rb, const WsReduce& red, hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_base_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, A, v, x, xb, rb, red);
    check_hip(hipGetLastError(), "ws_base_kernel");
}

void launch_ws_features(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const DevExt& e,
                        const DevMesh& m, const float* rest_pos, const WsHist& h, const double* xb,
                        const WsInst* wi, float dt, __bf16* F, __bf16* E, __bf16* E4, hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_features_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, A, v, e, m, rest_pos, h,
                       xb, wi, dt, F, E, E4);
    check_hip(hipGetLastError(), "ws_features_kernel");
}

void WarmStartNet::features_encode(const LaunchGeom& g, const DevMatrix& A, const DevVectors& v, const DevExt& e,
                                   const DevMesh& m, const float* rest_pos, const WsHist& h, const double* xb,
                                   const WsInst* wi, float dt, hipStream_t s) {
    if (g.nslices * 64 > rows_) throw std::invalid_argument("warm start: reserve() rows first");
    compact_edges_ = true;
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_features_encode_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, A, v, e, m,
                       rest_pos, h, xb, wi, dt, wnk_ + OFF_WE, theta_ + OFF_BE, E4_, H_[0]);
    check_hip(hipGetLastError(), "ws_features_encode_kernel");
}

void launch_ws_direction(const LaunchGeom& g, const DevMatrix& A, const float* out, const WsInst* wi,
                         const float* rb, float* d, float* ad, const WsReduce& red, hipStream_t s) {
    if (g.nslices <= 0) return;
    if (out != nullptr) {  // else the forward wrote d itself
        hipLaunchKernelGGL(ws_dir_write_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, out, wi, d);
        check_hip(hipGetLastError(), "ws_dir_write_kernel");
    }
    hipLaunchKernelGGL(ws_dir_spmv_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, A, d, rb, ad, red);
    check_hip(hipGetLastError(), "ws_dir_spmv_kernel");
}

void launch_ws_line_apply(const LaunchGeom& g, const WsInst* wi, const float* d, double* x, const float* rb,
                          const float* ad, float* r, hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_line_apply_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, wi, d, x, rb, ad, r);
    check_hip(hipGetLastError(), "ws_line_apply_kernel");
}

void launch_ws_energy(const LaunchGeom& g, const double* x, const double* xb, const float* rb, const WsReduce& red,
                      hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_energy_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, x, xb, rb, red);
    check_hip(hipGetLastError(), "ws_energy_kernel");
}

void launch_ws_dout(const LaunchGeom& g, const WsInst* wi, const float* rb, const float* ad, __bf16* dout,
                    hipStream_t s) {
    if (g.nslices <= 0) return;
    hipLaunchKernelGGL(ws_dout_kernel, dim3(grid_slices(g.nslices)), dim3(256), 0, s, g, wi, rb, ad, dout);
    check_hip(hipGetLastError(), "ws_dout_kernel");
}

void launch_ws_history(const double* x, float* hist, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(ws_history_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, hist, n);
    check_hip(hipGetLastError(), "ws_history_kernel");
}

}  // namespace chysx_amd
