// CDNA MFMA wrappers (16x16x4 FP32 and FP64) and 4-row vector access.
//
// v_mfma_*_16x16x4: D(16x16) += A(16x4) B(4x16).  Lane l (0..63) supplies
// A[l%16][l/16] and B[l/16][l%16]; accumulator register r (0..3) of lane l
// holds D[Mfma<T>::drow(l/16, r)][l%16]: row 4*(l/16)+r for FP32, row
// l/16+4*r for FP64 (checked on the device by the layout probe test).
#pragma once

#include <hip/hip_runtime.h>

namespace bx {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
    typedef f32x4 acc;
    static __device__ __forceinline__ int drow(int lk, int r) { return 4 * lk + r; }
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<double> {
    typedef f64x4 acc;
    static __device__ __forceinline__ int drow(int lk, int r) { return lk + 4 * r; }
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// four consecutive elements (16-byte aligned for float, 32-byte for double)
__device__ __forceinline__ void load4(const float *p, float v[4]) {
    const float4 x = *reinterpret_cast<const float4 *>(p);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
}
__device__ __forceinline__ void load4(const double *p, double v[4]) {
    const double2 x = *reinterpret_cast<const double2 *>(p);
    const double2 y = *reinterpret_cast<const double2 *>(p + 2);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = y.x;
    v[3] = y.y;
}
__device__ __forceinline__ void store4(float *p, float a, float b, float c, float d) {
    *reinterpret_cast<float4 *>(p) = make_float4(a, b, c, d);
}
__device__ __forceinline__ void store4(double *p, double a, double b, double c, double d) {
    *reinterpret_cast<double2 *>(p) = make_double2(a, b);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(c, d);
}

}  // namespace bx
