// Device-side building blocks shared by the wtsurf kernels (gfx950 / CDNA4).
//
// Matrix storage convention on the device: every per-instance complex matrix
// is an NP x NP column-major array of double2 (re, im), NP = 16*Q the order
// rounded up to the MFMA tile edge.  The pad rows/columns hold exact zeros;
// every product of zero-padded operands keeps them zero, so the kernels work
// on the padded shape without masks.
#pragma once

#include <hip/hip_runtime.h>

namespace wtsurf {

constexpr int kWave = 64;
constexpr int kBlock = 512;  // 8 waves, two per SIMD
constexpr int kGJThreads = 256;  // threads holding the Gauss-Jordan matrix
constexpr int kWaves = kBlock / kWave;
constexpr int kChunk = 8;    // K-chunk (complex) staged through LDS per step

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
// 1/z with the scaling of Smith's algorithm (no overflow for large |z|).
__device__ __forceinline__ double2 crecip(double2 z) {
    if (fabs(z.x) >= fabs(z.y)) {
        const double r = z.y / z.x, d = z.x + z.y * r;
        return make_double2(1.0 / d, -r / d);
    }
    const double r = z.x / z.y, d = z.y + z.x * r;
    return make_double2(r / d, -1.0 / d);
}

// threadIdx.x hidden from loop-invariant code motion: address arithmetic that
// derives from it is then recomputed inside each device-function call instead
// of being hoisted out of the instance loop and held in registers throughout.
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Values every thread of the workgroup holds identically (read from LDS
// after a barrier) made provably uniform for the compiler.  Loops and
// branches that contain barriers must test only such values: a condition the
// compiler believes divergent lets its structurizer send different lanes of
// one wave around different cycles of the loop, and the barriers then pair
// up wrongly.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// One 16x16x4 FP64 MFMA step.  Operand layout on gfx950 (checked on the
// device with a known-answer product):
//   A[i][k]: lane i + 16k;  B[k][j]: lane j + 16k;
//   D[i][j]: lane j + 16*(i%4), element i/4 of the v4d.
__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace wtsurf
