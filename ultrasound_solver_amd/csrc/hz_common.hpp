// Shared device types for the Helmholtz library: interleaved complex values,
// the per-precision operator view, and wavefront/workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hz_multi.hpp"

namespace hz {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T re, T im) { return cx<T>{re, im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__host__ __device__ __forceinline__ T norm2(cx<T> a) { return a.re * a.re + a.im * a.im; }
template <typename T>
__host__ __device__ __forceinline__ T cabs(cx<T> a) { return sqrt(a.re * a.re + a.im * a.im); }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cx<T> cdot(cx<T> a, cx<T> b) {
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> crecip(cx<T> b) {
    T d = b.re * b.re + b.im * b.im;
    return {b.re / d, -b.im / d};
}
template <typename To, typename From>
__host__ __device__ __forceinline__ cx<To> cast(cx<From> a) { return {(To)a.re, (To)a.im}; }

// Multiply-adds of the hot loops as FMAs in a fixed order.  The library is
// compiled with expression-level contraction only, so these (and not the
// compiler's scheduling) set the rounding of a stencil row or a Krylov
// update wherever it is evaluated.
// acc + a b
template <typename T>
__device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> acc) {
    return {fma(-a.im, b.im, fma(a.re, b.re, acc.re)), fma(a.im, b.re, fma(a.re, b.im, acc.im))};
}
// acc - a b
template <typename T>
__device__ __forceinline__ cx<T> cfnma(cx<T> a, cx<T> b, cx<T> acc) {
    return {fma(a.im, b.im, fma(-a.re, b.re, acc.re)), fma(-a.im, b.re, fma(-a.re, b.im, acc.im))};
}
// acc + s x (s real)
template <typename T>
__device__ __forceinline__ cx<T> sfma(T s, cx<T> x, cx<T> acc) {
    return {fma(s, x.re, acc.re), fma(s, x.im, acc.im)};
}
// s x - acc (s real)
template <typename T>
__device__ __forceinline__ cx<T> sfms(T s, cx<T> x, cx<T> acc) {
    return {fma(s, x.re, -acc.re), fma(s, x.im, -acc.im)};
}

// Per-precision operator view passed by value to every kernel.
template <typename T>
struct OpView {
    int nx, ny, nz;
    int64_t n;
    T omega2;
    T omega2_im;        // imaginary mass shift beta * omega^2 (multigrid levels), else 0
    const T *rho;       // N
    const T *kappa;     // N
    const cx<T> *a[3];  // w*s per axis, length n_axis
    const cx<T> *q[3];  // 4/((s_i+s_{i+1}) h^2), length n_axis-1
    const cx<T> *face[6];
    int soft;           // bit f set when face f is pressure-release
    const T *rs_f;      // row scale of A
    const cx<T> *m_f;   // Jacobi inverse of R A
    const T *rs_a;      // row scale of A^H
    const cx<T> *m_a;   // Jacobi inverse of R' A^H
};

constexpr int WAVE = 64;

// Two complex64 values moved with one 16-byte access.
typedef float hz_f4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

// Combine the 4 wavefront totals of a 256-thread workgroup in FP64 when every
// lane already holds its wavefront's total; result valid in thread 0.
__device__ __forceinline__ double block_combine_256(double wave_total, double *scratch) {
    const int lane = threadIdx.x % WAVE;
    const int wid = threadIdx.x / WAVE;
    if (lane == 0) scratch[wid] = wave_total;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
    __syncthreads();
    return r;
}

// Sum over a 256-thread workgroup in FP64; result valid in thread 0.
// `scratch` must hold 4 doubles in LDS.
__device__ __forceinline__ double block_sum_256(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x % WAVE;
    const int wid = threadIdx.x / WAVE;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
    __syncthreads();
    return r;
}

// A double moved across lanes by one DPP pattern (as two 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the wave, the same value in every lane, in a fixed order: pairs
// (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows
// (row_half_mirror) and rows of 16 (row_mirror) by DPP -- each step adds two
// values in both lanes that hold them, so all lanes of a row agree -- then
// the four row sums ((r0 + r1) + r2) + r3 read as scalars.
__device__ __forceinline__ double wave_sum_uniform(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return ((read_lane_f64(v, 0) + read_lane_f64(v, 16)) + read_lane_f64(v, 32)) + read_lane_f64(v, 48);
}

// Value of lane l-1 (wave_shr:1) / lane l+1 (wave_shl:1); lanes at the wave
// edge receive 0 and are fixed up by an explicit load.
__device__ __forceinline__ float lane_prev(float v) { return __builtin_amdgcn_update_dpp(0.0f, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_next(float v) { return __builtin_amdgcn_update_dpp(0.0f, v, 0x130, 0xf, 0xf, false); }

}  // namespace hz
