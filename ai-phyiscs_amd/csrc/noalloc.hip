// SPDX-License-Identifier: Apache-2.0
//
// Device scratch for the entry points that allocate no device memory, for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The scratch is one file-scope device array, `scratch`, of kScratchBytes;
// every GPU the library runs on holds its own copy until the process ends.
// In host code the array's name is only a stand-in for the device address, so
// the address is looked up once per device, with that device current, by
// hipGetSymbolAddress, kept per device, and used everywhere host code needs
// the array's memory (kernel arguments, copies).  Every call on one GPU shares
// that GPU's array and its contents stay from the last call: calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cloth_kernels.hpp"
#include "noalloc.hpp"
#include "topology.hpp"

namespace chysx_amd {

[[noreturn]] void throw_hip(hipError_t err, const char* what);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

void* scratch_address() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw_hip(e, "hipGetDevice(scratch)");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) throw_hip(e, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

size_t area_mass_scratch_bytes(int n, int ntri) {
    // vertex-to-triangle offsets (n + 1), incident triangle ids (3 T) and the
    // triangles (3 T), as ints
    return (size_t(n) + 1 + 6 * size_t(ntri)) * sizeof(int);
}

void area_mass_noalloc(const int* tris, int ntri, int n, const float* pos, float density, float* inv_mass,
                       hipStream_t s) {
    const size_t need = area_mass_scratch_bytes(n, ntri);
    if (need > kScratchBytes)
        throw std::invalid_argument("redistribute_mass_area_weighted_noalloc: " + std::to_string(n) + " particles and " +
                                    std::to_string(ntri) + " triangles need " + std::to_string(need) +
                                    " bytes of scratch; the maximum is " + std::to_string(kScratchBytes) +
                                    " bytes ((n + 1 + 6 T) * 4)");
    std::vector<int> off, ids;
    build_vertex_triangles(n, tris, ntri, off, ids);
    int* base = static_cast<int*>(scratch_address());
    int* d_off = base;
    int* d_ids = d_off + off.size();
    int* d_tri = d_ids + ids.size();
    hipError_t e = hipMemcpyAsync(d_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !ids.empty())
        e = hipMemcpyAsync(d_ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && ntri > 0)
        e = hipMemcpyAsync(d_tri, tris, sizeof(int) * 3 * size_t(ntri), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) throw_hip(e, "hipMemcpyAsync(scratch)");
    launch_area_mass(d_off, d_ids, d_tri, pos, density / 3.0f, inv_mass, n, s);
    // the host staging vectors must outlive the copies
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) throw_hip(e, "hipStreamSynchronize(area mass)");
}

}  // namespace chysx_amd
