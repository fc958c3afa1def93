// Device-side building blocks shared by the wtsurf kernels (gfx950 / CDNA4).
//
// Matrix storage convention on the device: every per-instance complex matrix
// is an NP x NP column-major array of (re, im) pairs, NP = 16*Q the order
// rounded up to the MFMA tile edge.  The pad rows/columns hold exact zeros;
// every product of zero-padded operands keeps them zero, so the kernels work
// on the padded shape without masks.
#pragma once

#include <hip/hip_runtime.h>

namespace wtsurf {

constexpr int kWave = 64;
constexpr int kBlock = 512;  // 8 waves, two per SIMD
constexpr int kWaves = kBlock / kWave;
constexpr int kChunk = 8;    // K-chunk (complex) staged through LDS per step
constexpr int kPanel = 8;    // Gauss-Jordan panel width (rank of each MFMA update)

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Maximum over the 64 lanes of a wave, returned to every lane: DPP row
// shifts inside each row of 16 lanes, then row broadcasts 15 and 31 carry
// the partial maxima up to lane 63.
__device__ __forceinline__ double wave_max_f64(double v) {
#define WT_DPP_MAX_STEP(ctrl, rmask)                                                  \
    {                                                                                  \
        const int hi = __double2hiint(v), lo = __double2loint(v);                      \
        const int h2 = __builtin_amdgcn_update_dpp(hi, hi, ctrl, rmask, 0xf, false);   \
        const int l2 = __builtin_amdgcn_update_dpp(lo, lo, ctrl, rmask, 0xf, false);   \
        v = fmax(v, __hiloint2double(h2, l2));                                         \
    }
    WT_DPP_MAX_STEP(0x111, 0xf)  // row_shr:1
    WT_DPP_MAX_STEP(0x112, 0xf)  // row_shr:2
    WT_DPP_MAX_STEP(0x114, 0xf)  // row_shr:4
    WT_DPP_MAX_STEP(0x118, 0xf)  // row_shr:8
    WT_DPP_MAX_STEP(0x142, 0xa)  // row_bcast:15
    WT_DPP_MAX_STEP(0x143, 0xc)  // row_bcast:31
#undef WT_DPP_MAX_STEP
    return readlane_f64(v, 63);
}

// Complex arithmetic on (re, im) pairs of either precision.
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
    C r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
// 1/z by Smith's scaling (one real division; no overflow for large |z|).
template <typename C>
__device__ __forceinline__ C crecip(C z) {
    C out;
    if (fabs(z.x) >= fabs(z.y)) {
        const auto r = z.y / z.x;
        const auto inv = decltype(r)(1) / (z.x + z.y * r);
        out.x = inv;
        out.y = -r * inv;
    } else {
        const auto r = z.x / z.y;
        const auto inv = decltype(r)(1) / (z.y + z.x * r);
        out.x = r * inv;
        out.y = -inv;
    }
    return out;
}

// crecip for a value every lane holds identically: the scaling branch is
// taken as one uniform branch instead of evaluating both sides.
template <typename C>
__device__ __forceinline__ C crecip_uniform(C z) {
    C out;
    if (uniform(fabs(z.x) >= fabs(z.y) ? 1 : 0)) {
        const auto r = z.y / z.x;
        const auto inv = decltype(r)(1) / (z.x + z.y * r);
        out.x = inv;
        out.y = -r * inv;
    } else {
        const auto r = z.x / z.y;
        const auto inv = decltype(r)(1) / (z.y + z.x * r);
        out.x = r * inv;
        out.y = -inv;
    }
    return out;
}

// Per-precision MFMA: one 16x16x4 step, D += A B, with the operand layout
//   A[i][k]: lane i + 16k;  B[k][j]: lane j + 16k;  D[i][j]: lane j + 16*q,
// element v, where (checked on the device with known-answer products)
//   FP64: i = 4v + q       FP32: i = 4q + v.
template <typename R>
struct Prec;
template <>
struct Prec<double> {
    using C = double2;
    using V4 = v4d;
    __device__ static __forceinline__ V4 mfma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int q, int v) { return 4 * v + q; }
    __device__ static __forceinline__ double readlane(double x, int l) { return readlane_f64(x, l); }
};
template <>
struct Prec<float> {
    using C = float2;
    using V4 = v4f;
    __device__ static __forceinline__ V4 mfma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int q, int v) { return 4 * q + v; }
    __device__ static __forceinline__ float readlane(float x, int l) { return readlane_f32(x, l); }
};

}  // namespace wtsurf
