// cn_calibrate.hip — kernels of known VALU instruction count, the yardsticks
// for reading the SQ instruction counters of the stage kernels (measurement
// only; no counterpart in the reference). Each lane runs `iterations` rounds
// of eight independent operations of one kind — f32 fused multiply-add, u32
// add, or f32 max, the bulk of a traversal's slab tests and bookkeeping — so
// a wave issues 8 * iterations VALU instructions plus a handful of setup
// ones, and with every SIMD full of such waves the chip issues that kind as
// fast as it can. The fastest kind's counted VALU instructions per second is
// the rate a stage kernel's are divided by.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "cn_handles.h"

namespace {

constexpr uint32_t CAL_BLOCK = 256;

__global__ __launch_bounds__(CAL_BLOCK) void k_valu_calibration(uint32_t iterations, float m, float c, float* out) {
    float a0 = float(threadIdx.x), a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
    float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
    for (uint32_t i = 0; i < iterations; ++i) {
        a0 = fmaf(a0, m, c);
        a1 = fmaf(a1, m, c);
        a2 = fmaf(a2, m, c);
        a3 = fmaf(a3, m, c);
        a4 = fmaf(a4, m, c);
        a5 = fmaf(a5, m, c);
        a6 = fmaf(a6, m, c);
        a7 = fmaf(a7, m, c);
    }
    const float s = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
    if (s == 12345.678f) out[blockIdx.x * CAL_BLOCK + threadIdx.x] = s;  // keeps the chains live
}

// Eight independent chains of one VALU instruction each, written out so the
// compiler can neither fuse nor drop them.
#define CAL_ROUND(OP)                                                           \
    asm volatile(OP " %0, %0, %1" : "+v"(a0) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a1) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a2) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a3) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a4) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a5) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a6) : "v"(k));                         \
    asm volatile(OP " %0, %0, %1" : "+v"(a7) : "v"(k))

__global__ __launch_bounds__(CAL_BLOCK) void k_valu_calibration_add(uint32_t iterations, uint32_t k, uint32_t* out) {
    uint32_t a0 = threadIdx.x, a1 = a0 + 1u, a2 = a0 + 2u, a3 = a0 + 3u;
    uint32_t a4 = a0 + 4u, a5 = a0 + 5u, a6 = a0 + 6u, a7 = a0 + 7u;
    for (uint32_t i = 0; i < iterations; ++i) {
        CAL_ROUND("v_add_u32");
    }
    const uint32_t s = ((a0 ^ a1) ^ (a2 ^ a3)) ^ ((a4 ^ a5) ^ (a6 ^ a7));
    if (s == 0x12345678u) out[blockIdx.x * CAL_BLOCK + threadIdx.x] = s;
}

__global__ __launch_bounds__(CAL_BLOCK) void k_valu_calibration_max(uint32_t iterations, float k, float* out) {
    float a0 = float(threadIdx.x), a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
    float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
    for (uint32_t i = 0; i < iterations; ++i) {
        CAL_ROUND("v_max_f32");
    }
    const float s = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
    if (s == 12345.678f) out[blockIdx.x * CAL_BLOCK + threadIdx.x] = s;
}

#undef CAL_ROUND

void launch_calibration(uint32_t kind, uint32_t blocks, uint32_t iterations, float* out) {
    if (kind == CN_CALIBRATE_ADD_U32) {
        hipLaunchKernelGGL(k_valu_calibration_add, dim3(blocks), dim3(CAL_BLOCK), 0, nullptr, iterations, 3u,
                           reinterpret_cast<uint32_t*>(out));
    } else if (kind == CN_CALIBRATE_MAX_F32) {
        hipLaunchKernelGGL(k_valu_calibration_max, dim3(blocks), dim3(CAL_BLOCK), 0, nullptr, iterations, 0.5f, out);
    } else {
        hipLaunchKernelGGL(k_valu_calibration, dim3(blocks), dim3(CAL_BLOCK), 0, nullptr, iterations, 0.999f, 0.001f,
                           out);
    }
}

}  // namespace

extern "C" int cn_calibrate_valu(cn_context* ctx, uint32_t kind, uint32_t iterations, uint32_t blocks, double* ms,
                                 uint64_t* valu_per_wave) {
    if (!ctx || !ms || !valu_per_wave) return cn::fail(CN_E_INVALID, "cn_calibrate_valu: null argument");
    if (iterations == 0 || blocks == 0) return cn::fail(CN_E_INVALID, "cn_calibrate_valu: empty launch");
    if (kind > CN_CALIBRATE_MAX_F32) return cn::fail(CN_E_INVALID, "cn_calibrate_valu: unknown kind %u", kind);
    CN_HIP(hipSetDevice(ctx->device));
    cn::DevBuf<float> out;
    CN_HIP(out.alloc(size_t(blocks) * CAL_BLOCK));
    hipEvent_t start, stop;
    CN_HIP(hipEventCreate(&start));
    CN_HIP(hipEventCreate(&stop));
    launch_calibration(kind, blocks, iterations, out.ptr);  // warm
    CN_HIP(hipEventRecord(start, nullptr));
    launch_calibration(kind, blocks, iterations, out.ptr);
    CN_HIP(hipEventRecord(stop, nullptr));
    CN_HIP(hipEventSynchronize(stop));
    float elapsed = 0.0f;
    CN_HIP(hipEventElapsedTime(&elapsed, start, stop));
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    *ms = elapsed;
    *valu_per_wave = uint64_t(8) * iterations;
    return CN_OK;
}
