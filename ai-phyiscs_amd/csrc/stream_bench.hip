// SPDX-License-Identifier: Apache-2.0
//
// HBM bandwidth probe for gfx950: a copy (b = a) and a triad (a = b + s c)
// over buffers far larger than the 256 MB Infinity Cache, so every byte comes
// from and goes to HBM.  16-byte loads and stores, four independent loads in
// flight per lane, one workgroup per 1024 float4 (no grid-stride loop, so
// every workgroup streams exactly 16 KiB per buffer).  The result is the
// bandwidth a streaming kernel reaches on the device, reported beside the
// published 8 TB/s peak.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "cloth_kernels.hpp"

namespace chysx_amd {

[[noreturn]] void throw_hip(hipError_t err, const char* what);

namespace {

constexpr int kStreamThreads = 256;
constexpr int kStreamPerThread = 4;
constexpr size_t kStreamChunk = size_t(kStreamThreads) * kStreamPerThread;  // float4 per workgroup

template <bool NT>
__device__ __forceinline__ float4 load16(const float4* p) {
    if constexpr (NT) {
        return make_float4(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y),
                           __builtin_nontemporal_load(&p->z), __builtin_nontemporal_load(&p->w));
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void store16(float4* p, float4 v) {
    if constexpr (NT) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
        __builtin_nontemporal_store(v.z, &p->z);
        __builtin_nontemporal_store(v.w, &p->w);
    } else {
        *p = v;
    }
}

template <bool NT>
__global__ __launch_bounds__(kStreamThreads) void stream_copy_kernel(const float4* __restrict__ a,
                                                                     float4* __restrict__ b) {
    const size_t i = size_t(blockIdx.x) * kStreamChunk + threadIdx.x;
    float4 v[kStreamPerThread];
#pragma unroll
    for (int k = 0; k < kStreamPerThread; ++k) v[k] = load16<NT>(a + i + size_t(k) * kStreamThreads);
#pragma unroll
    for (int k = 0; k < kStreamPerThread; ++k) store16<NT>(b + i + size_t(k) * kStreamThreads, v[k]);
}

template <bool NT>
__global__ __launch_bounds__(kStreamThreads) void stream_triad_kernel(float4* __restrict__ a,
                                                                      const float4* __restrict__ b,
                                                                      const float4* __restrict__ c, float s) {
    const size_t i = size_t(blockIdx.x) * kStreamChunk + threadIdx.x;
    float4 x[kStreamPerThread], y[kStreamPerThread];
#pragma unroll
    for (int k = 0; k < kStreamPerThread; ++k) {
        x[k] = load16<NT>(b + i + size_t(k) * kStreamThreads);
        y[k] = load16<NT>(c + i + size_t(k) * kStreamThreads);
    }
#pragma unroll
    for (int k = 0; k < kStreamPerThread; ++k)
        store16<NT>(a + i + size_t(k) * kStreamThreads,
                    make_float4(fmaf(s, y[k].x, x[k].x), fmaf(s, y[k].y, x[k].y), fmaf(s, y[k].z, x[k].z),
                                fmaf(s, y[k].w, x[k].w)));
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw_hip(e, what);
}

struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { check(hipMalloc(&p, bytes), "hipMalloc(stream bench)"); }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

}  // namespace

// kind: 0 copy, 1 triad, 2 copy with non-temporal loads and stores, 3 triad
// non-temporal.  bytes per buffer, rounded down to whole 16 KiB chunks.
// Returns mean ms per launch; *moved = bytes read + written per launch.
double stream_bench(int kind, size_t bytes, int reps, hipStream_t s, double* moved) {
    if (kind < 0 || kind > 3) throw std::invalid_argument("stream bench: kind must be 0..3");
    if (reps <= 0) throw std::invalid_argument("stream bench: reps must be positive");
    const size_t n = (bytes / (kStreamChunk * sizeof(float4))) * kStreamChunk;  // float4 per buffer
    if (n == 0) throw std::invalid_argument("stream bench: at least 16 KiB per buffer");
    const size_t blocks = n / kStreamChunk;
    if (blocks > size_t(0x7fffffff)) throw std::invalid_argument("stream bench: buffer too large");
    const bool triad = (kind & 1) != 0;
    const bool nt = kind >= 2;
    DeviceBuffer a(n * sizeof(float4)), b(n * sizeof(float4));
    DeviceBuffer c(triad ? n * sizeof(float4) : sizeof(float4));
    check(hipMemsetAsync(a.p, 0, n * sizeof(float4), s), "hipMemsetAsync(stream bench)");
    check(hipMemsetAsync(b.p, 0, n * sizeof(float4), s), "hipMemsetAsync(stream bench)");
    if (triad) check(hipMemsetAsync(c.p, 0, n * sizeof(float4), s), "hipMemsetAsync(stream bench)");
    auto launch = [&]() {
        float4* A = static_cast<float4*>(a.p);
        float4* B = static_cast<float4*>(b.p);
        const float4* C = static_cast<const float4*>(c.p);
        const dim3 grid{unsigned(blocks), 1, 1}, block{unsigned(kStreamThreads), 1, 1};
        if (!triad) {
            if (nt) stream_copy_kernel<true><<<grid, block, 0, s>>>(A, B);
            else stream_copy_kernel<false><<<grid, block, 0, s>>>(A, B);
        } else {
            if (nt) stream_triad_kernel<true><<<grid, block, 0, s>>>(A, B, C, 0.5f);
            else stream_triad_kernel<false><<<grid, block, 0, s>>>(A, B, C, 0.5f);
        }
        check(hipGetLastError(), "stream bench launch");
    };
    launch();
    launch();
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    check(hipEventRecord(e0, s), "hipEventRecord");
    for (int r = 0; r < reps; ++r) launch();
    check(hipEventRecord(e1, s), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (moved) *moved = double(n * sizeof(float4)) * (triad ? 3.0 : 2.0);
    return double(ms) / reps;
}

}  // namespace chysx_amd
