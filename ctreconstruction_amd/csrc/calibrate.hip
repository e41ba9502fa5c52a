// calibrate.hip -- instruction-issue microbenchmarks for gfx950, used by bench.py --calibrate-valu.
//
// Not part of the product ABI (include/ctrecon.h): a separate library, libctrcalib.so, whose one
// entry point launches a kernel that issues a known number of one instruction kind.  Every
// thread runs `iters` iterations of 8 independent chains of that instruction, written in inline
// assembly so that the count is exact: each wavefront issues 8 * iters instructions of the
// kind (plus a handful before and after the loop, whose counter is scalar).  Reading the
// hardware counters (SQ_INSTS_VALU, GRBM_GUI_ACTIVE) over these kernels gives the issue rate
// each kind reaches on its own, the ceiling against which a mixed kernel's rate is judged.

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

enum CalibKind {
    kFma = 0,          // v_fma_f32: FP32 fused multiply-add
    kAddU32 = 1,       // v_add_u32: integer add
    kCndmask = 2,      // v_cndmask_b32: select under a scalar lane mask
    kReadfirstlane = 3,  // v_readfirstlane_b32: VGPR -> SGPR broadcast
    kRcp = 4,          // v_rcp_f32: reciprocal (transcendental unit)
    kAddF64 = 5,       // v_add_f64: FP64 add
    kMul24 = 6,        // v_mul_u32_u24: 24-bit integer multiply
    kNumKinds = 7
};

template <int KIND>
__global__ void __launch_bounds__(256) calib_kernel(float* __restrict__ sink, int iters) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a0 = 1.0f + t * 1e-9f, a1 = a0 + 1e-7f, a2 = a0 + 2e-7f, a3 = a0 + 3e-7f;
    float a4 = a0 + 4e-7f, a5 = a0 + 5e-7f, a6 = a0 + 6e-7f, a7 = a0 + 7e-7f;
    const float m = 0.999999f, c = 1e-7f;
    unsigned u0 = t, u1 = t + 1, u2 = t + 2, u3 = t + 3, u4 = t + 4, u5 = t + 5, u6 = t + 6, u7 = t + 7;
    double d0 = a0, d1 = a1, d2 = a2, d3 = a3, d4 = a4, d5 = a5, d6 = a6, d7 = a7;
    unsigned acc_s = 0;
    const uint64_t lane_mask = 0x5555555555555555ull >> (blockIdx.x & 1);  // uniform: scalar registers
    for (int i = 0; i < iters; ++i) {
        if constexpr (KIND == kFma) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n"
                "v_fma_f32 %3, %3, %8, %9\n v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n"
                "v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
                : "v"(m), "v"(c));
        } else if constexpr (KIND == kAddU32) {
            asm volatile(
                "v_add_u32 %0, %0, %8\n v_add_u32 %1, %1, %8\n v_add_u32 %2, %2, %8\n"
                "v_add_u32 %3, %3, %8\n v_add_u32 %4, %4, %8\n v_add_u32 %5, %5, %8\n"
                "v_add_u32 %6, %6, %8\n v_add_u32 %7, %7, %8"
                : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3), "+v"(u4), "+v"(u5), "+v"(u6), "+v"(u7)
                : "v"(3u));
        } else if constexpr (KIND == kCndmask) {
            // eight independent selects under a lane mask held in scalar registers
            asm volatile(
                "v_cndmask_b32_e64 %0, %0, %8, %9\n v_cndmask_b32_e64 %1, %1, %8, %9\n"
                "v_cndmask_b32_e64 %2, %2, %8, %9\n v_cndmask_b32_e64 %3, %3, %8, %9\n"
                "v_cndmask_b32_e64 %4, %4, %8, %9\n v_cndmask_b32_e64 %5, %5, %8, %9\n"
                "v_cndmask_b32_e64 %6, %6, %8, %9\n v_cndmask_b32_e64 %7, %7, %8, %9"
                : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3), "+v"(u4), "+v"(u5), "+v"(u6), "+v"(u7)
                : "v"(7u), "s"(lane_mask));
        } else if constexpr (KIND == kReadfirstlane) {
            unsigned s0, s1, s2, s3, s4, s5, s6, s7;
            asm volatile(
                "v_readfirstlane_b32 %0, %8\n v_readfirstlane_b32 %1, %9\n v_readfirstlane_b32 %2, %10\n"
                "v_readfirstlane_b32 %3, %11\n v_readfirstlane_b32 %4, %12\n v_readfirstlane_b32 %5, %13\n"
                "v_readfirstlane_b32 %6, %14\n v_readfirstlane_b32 %7, %15"
                : "=s"(s0), "=s"(s1), "=s"(s2), "=s"(s3), "=s"(s4), "=s"(s5), "=s"(s6), "=s"(s7)
                : "v"(u0), "v"(u1), "v"(u2), "v"(u3), "v"(u4), "v"(u5), "v"(u6), "v"(u7));
            acc_s ^= s0 ^ s1 ^ s2 ^ s3 ^ s4 ^ s5 ^ s6 ^ s7;
        } else if constexpr (KIND == kRcp) {
            asm volatile(
                "v_rcp_f32 %0, %0\n v_rcp_f32 %1, %1\n v_rcp_f32 %2, %2\n v_rcp_f32 %3, %3\n"
                "v_rcp_f32 %4, %4\n v_rcp_f32 %5, %5\n v_rcp_f32 %6, %6\n v_rcp_f32 %7, %7"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7));
        } else if constexpr (KIND == kAddF64) {
            asm volatile(
                "v_add_f64 %0, %0, %8\n v_add_f64 %1, %1, %8\n v_add_f64 %2, %2, %8\n"
                "v_add_f64 %3, %3, %8\n v_add_f64 %4, %4, %8\n v_add_f64 %5, %5, %8\n"
                "v_add_f64 %6, %6, %8\n v_add_f64 %7, %7, %8"
                : "+v"(d0), "+v"(d1), "+v"(d2), "+v"(d3), "+v"(d4), "+v"(d5), "+v"(d6), "+v"(d7)
                : "v"(1e-9));
        } else {
            static_assert(KIND == kMul24, "kind");
            asm volatile(
                "v_mul_u32_u24 %0, %0, %8\n v_mul_u32_u24 %1, %1, %8\n v_mul_u32_u24 %2, %2, %8\n"
                "v_mul_u32_u24 %3, %3, %8\n v_mul_u32_u24 %4, %4, %8\n v_mul_u32_u24 %5, %5, %8\n"
                "v_mul_u32_u24 %6, %6, %8\n v_mul_u32_u24 %7, %7, %8"
                : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3), "+v"(u4), "+v"(u5), "+v"(u6), "+v"(u7)
                : "v"(5u));
        }
    }
    // one store per thread keeps every chain alive
    const float f = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
    const unsigned u = u0 ^ u1 ^ u2 ^ u3 ^ u4 ^ u5 ^ u6 ^ u7 ^ acc_s;
    const double d = d0 + d1 + d2 + d3 + d4 + d5 + d6 + d7;
    sink[t] = f + (float)u + (float)d;
}

template <int K>
hipError_t launch(int kind, int blocks, int iters, float* sink, hipStream_t s) {
    if (kind == K) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(calib_kernel<K>), dim3(blocks), dim3(256), 0, s, sink, iters);
        return hipGetLastError();
    }
    if constexpr (K + 1 < kNumKinds) return launch<K + 1>(kind, blocks, iters, sink, s);
    return hipErrorInvalidValue;
}

}  // namespace

extern "C" {

// Number of instruction kinds (kind = 0 .. n-1: fma_f32, add_u32, cndmask_b32, readfirstlane_b32,
// rcp_f32, add_f64, mul_u32_u24).
int ctr_calib_kinds(void) { return kNumKinds; }

// Launches `blocks` workgroups of 256 threads; each wavefront issues 8 * iters instructions of
// `kind`.  `sink` holds blocks * 256 floats.  Returns 0 or a HIP error.
int ctr_calib_run(int kind, int blocks, int iters, float* sink, void* stream) {
    if (kind < 0 || kind >= kNumKinds || blocks <= 0 || iters <= 0 || !sink) return 1;
    return launch<0>(kind, blocks, iters, sink, (hipStream_t)stream) == hipSuccess ? 0 : 4;
}

}  // extern "C"
