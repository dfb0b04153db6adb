// mfma.h -- the 16x16x4 matrix-core products used by the dense factorizations (FP64 and FP32).
//
// One wave computes a 16 x 16 block D += A B with A 16 x 4 and B 4 x 16: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element i of lane l is
// D[row(l, i)][l % 16], with the rows ordered differently by the FP64 and FP32 instructions.
#pragma once

#include <hip/hip_runtime.h>

namespace mba {

template <typename F>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int i) { return (lane >> 4) + 4 * i; }
    static __device__ __forceinline__ void load8(const double* p, double (&v)[8]) {
        const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double2 t = q[i];
            v[2 * i] = t.x;
            v[2 * i + 1] = t.y;
        }
    }
};
template <>
struct Mfma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
    static __device__ __forceinline__ void load8(const float* p, float (&v)[8]) {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float4 t = q[i];
            v[4 * i] = t.x;
            v[4 * i + 1] = t.y;
            v[4 * i + 2] = t.z;
            v[4 * i + 3] = t.w;
        }
    }
};

}  // namespace mba
