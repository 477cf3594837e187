// Online response-basis kernels: u = G w, g = G^H c and the fused loss
// evaluations, on a phase response basis G held on the device as nrows x m
// complex128 rows (row i = node i in Fortran order, m emitters contiguous,
// row stride ldg), the layout of the reference's basis file
// (phase_response_basis.py:177-253).  Each reads G once; the fused kernels
// evaluate a loss, its cotangent and G^H c in that single pass
// (phase_adjoint_optimizer.py:163-364).
#pragma once

#include "hz_common.hpp"

namespace hz {

constexpr int BASIS_MAX_M = 2048;        // emitters per row held in LDS
constexpr int BASIS_FUSED_MAX_M = 512;   // emitters per row of the fused kernel (registers)
// Per-wave partials of one fused loss launch are capped at this size: the
// wave count shrinks as m grows (1024 waves at m = 512, 2048 at m = 256).
constexpr int64_t BASIS_LOSS_PARTIAL_BYTES = 8ll << 20;

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

// Fused loss evaluations: one wavefront per row computes u_i = G_i w, the
// row's share of the loss and its cotangent c_i, then accumulates
// conj(G[i, q]) c_i for its lanes' columns from the row already in
// registers, so G is read once.  Per-wave partials (NACC accumulators of m
// values, then LOSS_SCALARS doubles) are combined by k_basis_loss_combine
// in a fixed order, which keeps the result reproducible.
//   LOSS_FIELD_MATCH  0.5 sum (W (|u| - a))^2, c = W^2 (|u| - a) u/(|u| + eps)
//                     (phase_adjoint_optimizer.py:163-177)
//   LOSS_ENERGY       background of focal_pressure: 0.5 s sum |u|^2, c = s u
//                     with s = beta/N (:244-258)
//   LOSS_SOFTMAX      background of focal_contrast: beta * smooth-max over the
//                     masked rows, T (log sum exp(|u|/T) - log count), with
//                     c = beta p u/(|u| + eps), p the softmax weights
//                     (:312-338); each wave keeps a running maximum and
//                     rescales its sums when it grows, so the weights need
//                     no second pass
//   LOSS_RATIO        mean_A |u|^2 / mean_B |u|^2 over two row classes
//                     (mask 1 = A, 2 = B): A and B sums of |u|^2 and of
//                     conj(G) u; c follows at the combine
enum BasisLoss { LOSS_FIELD_MATCH = 0, LOSS_ENERGY = 1, LOSS_SOFTMAX = 2, LOSS_RATIO = 3 };

constexpr int LOSS_SCALARS = 2;

struct BasisLossArgs {
    const double *target;
    const double *weight;
    const unsigned char *mask;
    double eps, scale, temperature;
    double count_a, count_b;
};

template <int KIND>
__global__ __launch_bounds__(256) void k_basis_loss(const cx<double> *__restrict__ G, int64_t ldg, int64_t nrows,
                                                    int m, const cx<double> *__restrict__ w, BasisLossArgs a,
                                                    int64_t rows_per_wave, cx<double> *__restrict__ partial,
                                                    double *__restrict__ scal, cx<double> *__restrict__ u_out) {
    constexpr int PER = BASIS_FUSED_MAX_M / 64;
    constexpr int NACC = KIND == LOSS_RATIO ? 2 : 1;
    __shared__ cx<double> ws[BASIS_FUSED_MAX_M];
    for (int q = threadIdx.x; q < m; q += 256) ws[q] = w[q];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t i0 = wave * rows_per_wave;
    const int64_t i1 = min(nrows, i0 + rows_per_wave);
    const int per = (m + 63) / 64;
    cx<double> acc[NACC][PER];
    cx<double> g[PER];
#pragma unroll
    for (int h = 0; h < NACC; ++h)
#pragma unroll
        for (int q = 0; q < PER; ++q) acc[h][q] = mk<double>(0, 0);
    double s0 = 0, s1 = 0;  // SOFTMAX: running maximum (s0) and scaled sum (s1)
    if (KIND == LOSS_SOFTMAX) s0 = -__builtin_inf();
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
        if (u_out && lane == 0) u_out[i] = ui;
        const double p2 = re * re + im * im;
        int slot = 0;
        cx<double> ci;
        if (KIND == LOSS_FIELD_MATCH) {
            const double amp = sqrt(p2);
            const double res = amp - a.target[i];
            const double wt = a.weight[i];
            s0 += 0.5 * (wt * res) * (wt * res);
            ci = (wt * wt * res / (amp + a.eps)) * ui;
        } else if (KIND == LOSS_ENERGY) {
            s0 += p2;
            ci = ui;
        } else if (KIND == LOSS_SOFTMAX) {
            if (!a.mask[i]) continue;
            const double amp = sqrt(p2);
            if (amp > s0) {
                const double f = exp((s0 - amp) / a.temperature);
                s1 *= f;
#pragma unroll
                for (int q = 0; q < PER; ++q) acc[0][q] = f * acc[0][q];
                s0 = amp;
            }
            const double e = exp((amp - s0) / a.temperature);
            s1 += e;
            ci = (e / (amp + a.eps)) * ui;
        } else {
            const unsigned char cls = a.mask[i];
            if (cls != 1 && cls != 2) continue;
            slot = cls - 1;
            if (cls == 1) s0 += p2; else s1 += p2;
            ci = ui;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (q < per) {
                if (NACC == 1 || slot == 0) acc[0][q] = acc[0][q] + conj(g[q]) * ci;
                else acc[NACC - 1][q] = acc[NACC - 1][q] + conj(g[q]) * ci;
            }
    }
#pragma unroll
    for (int h = 0; h < NACC; ++h)
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int col = lane + 64 * q;
            if (q < per && col < m) partial[(h * nwaves + wave) * (int64_t)m + col] = acc[h][q];
        }
    if (lane == 0) {
        scal[wave * LOSS_SCALARS] = s0;
        scal[wave * LOSS_SCALARS + 1] = s1;
    }
}

// g[q] = G^H c from the per-wave partials and the loss into *loss; one
// thread per column, every thread walks the wave scalars in the same order.
template <int KIND>
__global__ __launch_bounds__(256) void k_basis_loss_combine(const cx<double> *__restrict__ partial,
                                                            const double *__restrict__ scal, int64_t nwaves, int m,
                                                            BasisLossArgs a, cx<double> *__restrict__ g,
                                                            double *__restrict__ loss) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    double t0 = 0, t1 = 0, top = -__builtin_inf();
    if (KIND == LOSS_SOFTMAX) {
        for (int64_t v = 0; v < nwaves; ++v) top = fmax(top, scal[v * LOSS_SCALARS]);
        for (int64_t v = 0; v < nwaves; ++v) {
            const double mx = scal[v * LOSS_SCALARS];
            if (mx > -__builtin_inf()) t1 += scal[v * LOSS_SCALARS + 1] * exp((mx - top) / a.temperature);
        }
    } else {
        for (int64_t v = 0; v < nwaves; ++v) {
            t0 += scal[v * LOSS_SCALARS];
            t1 += scal[v * LOSS_SCALARS + 1];
        }
    }
    double lval = 0, f0 = 1, f1 = 0;
    if (KIND == LOSS_FIELD_MATCH) {
        lval = t0;
    } else if (KIND == LOSS_ENERGY) {
        lval = 0.5 * a.scale * t0;
        f0 = a.scale;
    } else if (KIND == LOSS_SOFTMAX) {
        lval = a.scale * (top + a.temperature * (log(t1) - log(a.count_a)));
        f0 = a.scale / t1;
    } else {
        const double ma = t0 / a.count_a, mb = t1 / a.count_b;
        lval = ma / mb;
        f0 = 2.0 / (a.count_a * mb);
        f1 = -2.0 * ma / (a.count_b * mb * mb);
    }
    if (q < m) {
        cx<double> s = mk<double>(0, 0), r = mk<double>(0, 0);
        for (int64_t v = 0; v < nwaves; ++v) {
            cx<double> p = partial[v * (int64_t)m + q];
            if (KIND == LOSS_SOFTMAX) {
                const double mx = scal[v * LOSS_SCALARS];
                p = (mx > -__builtin_inf()) ? exp((mx - top) / a.temperature) * p : mk<double>(0, 0);
            }
            s = s + p;
            if (KIND == LOSS_RATIO) r = r + partial[(nwaves + v) * (int64_t)m + q];
        }
        g[q] = KIND == LOSS_RATIO ? f0 * s + f1 * r : f0 * s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *loss = lval;
}

}  // namespace hz
