// SPDX-License-Identifier: Apache-2.0
//
// Entry-point helpers that allocate no device memory: their scratch is a
// file-scope device array of kScratchBytes per GPU (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace chysx_amd {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU

// Device address of this GPU's scratch array (looked up once per device).
void* scratch_address();

// Scratch bytes the area-lumped mass of n particles and ntri triangles needs.
size_t area_mass_scratch_bytes(int n, int ntri);

// inv_mass = 1 / (density * A_i / 3 summed over incident triangles), as
// launch_area_mass, with the vertex-to-triangle incidence and the triangles
// staged in the scratch array.  Throws invalid_argument beyond kScratchBytes.
// Synchronises `s` before returning.
void area_mass_noalloc(const int* tris, int ntri, int n, const float* pos, float density, float* inv_mass,
                       hipStream_t s);

}  // namespace chysx_amd
