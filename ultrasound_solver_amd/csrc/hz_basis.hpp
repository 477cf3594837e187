// Online response-basis kernels: u = G w, g = G^H c and the fused
// field_match evaluation, on a phase response basis G held on the device as
// nrows x m complex128 rows (row i = node i in Fortran order, m emitters
// contiguous, row stride ldg), the layout of the reference's basis file
// (phase_response_basis.py:177-253).  All three read G once; the fused
// kernel evaluates the loss, its cotangent and G^H c in that single pass
// (phase_adjoint_optimizer.py:163-200, field_match).
#pragma once

#include "hz_common.hpp"

namespace hz {

constexpr int BASIS_MAX_M = 2048;        // emitters per row held in LDS
constexpr int BASIS_FUSED_MAX_M = 512;   // emitters per row of the fused kernel (registers)

// u[i] = sum_m G[i, m] w[m]: one wavefront per row, w staged in LDS.
__global__ __launch_bounds__(256) void k_basis_field(const cx<double> *__restrict__ G, int64_t ldg, int64_t nrows,
                                                     int m, const cx<double> *__restrict__ w,
                                                     cx<double> *__restrict__ u) {
    __shared__ cx<double> ws[BASIS_MAX_M];
    for (int q = threadIdx.x; q < m; q += 256) ws[q] = w[q];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = wave; i < nrows; i += nwaves) {
        const cx<double> *row = G + i * ldg;
        double re = 0, im = 0;
        for (int q = lane; q < m; q += 64) {
            const cx<double> g = row[q];
            const cx<double> x = ws[q];
            re += g.re * x.re - g.im * x.im;
            im += g.re * x.im + g.im * x.re;
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (lane == 0) u[i] = mk<double>(re, im);
    }
}

// partial[blk * m + q] = sum over the workgroup's rows of conj(G[i, q]) c[i];
// each thread owns columns tid, tid + 256, ...  A fixed row partition and a
// fixed-order combine (k_basis_combine) keep the result reproducible.
__global__ __launch_bounds__(256) void k_basis_adjoint(const cx<double> *__restrict__ G, int64_t ldg, int64_t nrows,
                                                       int m, const cx<double> *__restrict__ c, int64_t rows_per_block,
                                                       cx<double> *__restrict__ partial) {
    constexpr int PER = BASIS_MAX_M / 256;
    cx<double> acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = mk<double>(0, 0);
    const int64_t i0 = blockIdx.x * rows_per_block;
    const int64_t i1 = min(nrows, i0 + rows_per_block);
    for (int64_t i = i0; i < i1; ++i) {
        const cx<double> ci = c[i];
        const cx<double> *row = G + i * ldg;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int col = threadIdx.x + 256 * q;
            if (col < m) acc[q] = acc[q] + conj(row[col]) * ci;
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int col = threadIdx.x + 256 * q;
        if (col < m) partial[blockIdx.x * (int64_t)m + col] = acc[q];
    }
}

// out[q] = sum_blk partial[blk * m + q] (+ the loss partials when asked).
__global__ __launch_bounds__(256) void k_basis_combine(const cx<double> *__restrict__ partial, int nblk, int m,
                                                       cx<double> *__restrict__ out,
                                                       const double *__restrict__ loss_partial,
                                                       double *__restrict__ loss) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < m) {
        cx<double> s = mk<double>(0, 0);
        for (int b = 0; b < nblk; ++b) s = s + partial[(int64_t)b * m + q];
        out[q] = s;
    }
    if (loss_partial && blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0;
        for (int b = 0; b < nblk; ++b) s += loss_partial[b];
        *loss = s;
    }
}

// Fused field_match (reference loss: 0.5 |W (|u| - a)|^2, cotangent
// c = W^2 (|u| - a) u / (|u| + eps)): one wavefront per row computes u_i,
// the loss term and c_i, then accumulates conj(G[i, q]) c_i for its lanes'
// columns from the row already in registers.  Per-wave partials of G^H c
// and of the loss; u is written when u_out is not null.
__global__ __launch_bounds__(256) void k_basis_field_match(const cx<double> *__restrict__ G, int64_t ldg,
                                                           int64_t nrows, int m, const cx<double> *__restrict__ w,
                                                           const double *__restrict__ target,
                                                           const double *__restrict__ weight, double eps,
                                                           int64_t rows_per_wave, cx<double> *__restrict__ partial,
                                                           double *__restrict__ loss_partial,
                                                           cx<double> *__restrict__ u_out) {
    constexpr int PER = BASIS_FUSED_MAX_M / 64;
    __shared__ cx<double> ws[BASIS_FUSED_MAX_M];
    for (int q = threadIdx.x; q < m; q += 256) ws[q] = w[q];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t i0 = wave * rows_per_wave;
    const int64_t i1 = min(nrows, i0 + rows_per_wave);
    const int per = (m + 63) / 64;
    cx<double> acc[PER];
    cx<double> g[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = mk<double>(0, 0);
    double lsum = 0;
    for (int64_t i = i0; i < i1; ++i) {
        const cx<double> *row = G + i * ldg;
        double re = 0, im = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int col = lane + 64 * q;
            g[q] = (q < per && col < m) ? row[col] : mk<double>(0, 0);
            if (q < per && col < m) {
                const cx<double> x = ws[col];
                re += g[q].re * x.re - g[q].im * x.im;
                im += g[q].re * x.im + g[q].im * x.re;
            }
        }
        re = wave_sum(re);
        im = wave_sum(im);
        const cx<double> ui = mk<double>(re, im);
        const double amp = sqrt(re * re + im * im);
        const double res = amp - target[i];
        const double wt = weight[i];
        lsum += 0.5 * (wt * res) * (wt * res);
        const cx<double> ci = (wt * wt * res / (amp + eps)) * ui;
        if (u_out && lane == 0) u_out[i] = ui;
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (q < per) acc[q] = acc[q] + conj(g[q]) * ci;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int col = lane + 64 * q;
        if (q < per && col < m) partial[wave * (int64_t)m + col] = acc[q];
    }
    if (lane == 0) loss_partial[wave] = lsum;
}

}  // namespace hz
