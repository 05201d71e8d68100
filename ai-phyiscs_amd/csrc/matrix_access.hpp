// SPDX-License-Identifier: Apache-2.0
//
// Block access on the wave-sliced ELL matrix shared by every kernel that
// multiplies by A.  With A.sym set the lower blocks (row, c), c < row, are
// read as transposes of their upper mirrors (c, row); assembly then writes
// only the upper blocks, and every product (iterations, true residual, warm
// start) uses the same symmetric operator.  With A.bs == 6 every stored
// block is symmetric and kept as its upper triangle.
#pragma once

#include <hip/hip_runtime.h>

#include "cloth_kernels.hpp"

namespace chysx_amd {

// The 3x3 block of entry k of `row` (vp_k = that entry's component 0 in its
// slice, stride 64 floats), row-major into m, with the entry's mirror index
// `mir` already loaded (it is used only for a lower entry of the symmetric
// storage), so a caller can issue the index loads of several entries before
// any block load.
__device__ __forceinline__ void load_block_at(const DevMatrix& A, const float* __restrict__ vp_k, int mir, int c,
                                              int row, float (&m)[9]) {
    if (A.bs == 6) {
        // symmetric block: the mirror (c, row) equals its transpose
        const float* b = (A.sym && c < row) ? A.val + val_offset(mir, 6) : vp_k;
        m[0] = b[0];
        m[1] = b[64];
        m[2] = b[128];
        m[3] = m[1];
        m[4] = b[192];
        m[5] = b[256];
        m[6] = m[2];
        m[7] = m[5];
        m[8] = b[320];
    } else if (A.sym && c < row) {
        const float* bt = A.val + val_offset(mir, 9);
        m[0] = bt[0];
        m[1] = bt[192];
        m[2] = bt[384];
        m[3] = bt[64];
        m[4] = bt[256];
        m[5] = bt[448];
        m[6] = bt[128];
        m[7] = bt[320];
        m[8] = bt[512];
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = vp_k[q * 64];
    }
}

// The same, reading the mirror index (mir_k = the entry's slot in A.mir)
// only for a lower entry of the symmetric storage.
__device__ __forceinline__ void load_block(const DevMatrix& A, const float* __restrict__ vp_k,
                                           const int* __restrict__ mir_k, int c, int row, float (&m)[9]) {
    const int mir = (A.sym && c < row) ? *mir_k : 0;
    load_block_at(A, vp_k, mir, c, row, m);
}

// The diagonal block of row (s, lane), row-major into m.  With A.bs == 6
// (no bending) the diagonal blocks are symmetric too and stored as their
// upper triangle at diag[(s * 6 + c) * 64 + lane].
__device__ __forceinline__ void load_diag(const DevMatrix& A, int s, int lane, float (&m)[9]) {
    if (A.bs == 6) {
        const float* d = A.diag + (size_t)s * 6 * 64 + lane;
        m[0] = d[0];
        m[1] = d[64];
        m[2] = d[128];
        m[3] = m[1];
        m[4] = d[192];
        m[5] = d[256];
        m[6] = m[2];
        m[7] = m[5];
        m[8] = d[320];
    } else {
        const float* d = A.diag + (size_t)s * 9 * 64 + lane;
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = d[q * 64];
    }
}

}  // namespace chysx_amd
