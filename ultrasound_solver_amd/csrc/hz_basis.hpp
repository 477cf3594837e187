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
// The per-workgroup partials of one fused loss launch and their folded
// slices are capped at this size together: the workgroup count shrinks as m
// grows (448 workgroups at m = 512 with two accumulators, 960 with one, the
// maximum of 1024 below m = 480 with one).
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
        re = wave_sum_uniform(re);
        im = wave_sum_uniform(im);
        if (lane == 0) u[i] = mk<double>(re, im);
    }
}

// Fused loss evaluations: one wavefront per row computes u_i = G_i w, the
// row's share of the loss and its cotangent c_i, then accumulates
// conj(G[i, q]) c_i for its lanes' columns from the row already in
// registers, so G is read once; the next row's loads are issued before the
// current row's reductions, so two rows per wave are in flight.  The four
// waves of a workgroup sum their accumulators in LDS in a fixed order and
// write one partial per workgroup (NACC accumulators of m values, then
// LOSS_SCALARS doubles); k_basis_loss_fold and k_basis_loss_combine reduce
// the partials, also in a fixed order, which keeps the result reproducible.
//   LOSS_FIELD_MATCH  0.5 sum (W (|u| - a))^2, c = W^2 (|u| - a) u/(|u| + eps)
//                     (phase_adjoint_optimizer.py:163-177)
//   LOSS_ENERGY       background of focal_pressure: 0.5 s sum |u|^2, c = s u
//                     with s = beta/N (:244-258)
//   LOSS_SOFTMAX      background of focal_contrast: beta * smooth-max over the
//                     masked rows, T (log sum exp(|u|/T) - log count), with
//                     c = beta p u/(|u| + eps), p the softmax weights
//                     (:312-338); each wave keeps a running maximum and
//                     rescales its sums when it grows, and every reduction
//                     stage rescales to the largest maximum it combines, so
//                     the weights need no second pass
//   LOSS_RATIO        mean_A |u|^2 / mean_B |u|^2 over two row classes
//                     (mask 1 = A, 2 = B): A and B sums of |u|^2 and of
//                     conj(G) u; c follows at the combine
//   LOSS_ADJOINT      no loss: G^H c for a given cotangent c (no forward
//                     product), the phase VJP of phase_response_basis.py:208-253
enum BasisLoss { LOSS_FIELD_MATCH = 0, LOSS_ENERGY = 1, LOSS_SOFTMAX = 2, LOSS_RATIO = 3, LOSS_ADJOINT = 4 };

constexpr int LOSS_SCALARS = 2;
// Slices the fold reduces the workgroup partials to before the final combine.
constexpr int LOSS_FOLD_SLICES = 64;

struct BasisLossArgs {
    const double *target;
    const double *weight;
    const unsigned char *mask;
    const cx<double> *cot;  // LOSS_ADJOINT
    double eps, scale, temperature;
    double count_a, count_b;
};

// exp((mx - top)/T): the factor that moves a softmax partial kept relative to
// its own maximum mx onto the maximum top; 0 for a partial with no rows.
__device__ __forceinline__ double softmax_rescale(double mx, double top, double temperature) {
    return mx > -__builtin_inf() ? exp((mx - top) / temperature) : 0.0;
}

// PER: columns per lane (m <= 64 PER).  Writes partial[(h nblk + blk) m + q]
// and scal[blk LOSS_SCALARS + {0, 1}] (sums, or for LOSS_SOFTMAX the largest
// |u| and the sum of exp((|u| - that)/T)).
template <int KIND, int PER>
__global__ __launch_bounds__(256) void k_basis_loss(const cx<double> *__restrict__ G, int64_t ldg, int64_t nrows,
                                                    int m, const cx<double> *__restrict__ w, BasisLossArgs a,
                                                    int64_t rows_per_wave, cx<double> *__restrict__ partial,
                                                    double *__restrict__ scal, cx<double> *__restrict__ u_out) {
    constexpr int NACC = KIND == LOSS_RATIO ? 2 : 1;
    __shared__ cx<double> ws[PER * 64];
    __shared__ cx<double> red[NACC][PER * 64];
    __shared__ double rs[2][4];
    if (KIND != LOSS_ADJOINT)
        for (int q = threadIdx.x; q < PER * 64; q += 256) ws[q] = q < m ? w[q] : mk<double>(0, 0);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t wave = blockIdx.x * 4ll + wv;
    const int64_t i0 = wave * rows_per_wave;
    const int64_t i1 = min(nrows, i0 + rows_per_wave);
    cx<double> acc[NACC][PER];
    cx<double> g[PER], gn[PER];
#pragma unroll
    for (int h = 0; h < NACC; ++h)
#pragma unroll
        for (int q = 0; q < PER; ++q) acc[h][q] = mk<double>(0, 0);
    auto load_row = [&](cx<double> *dst, int64_t i) {
        const cx<double> *row = G + i * ldg;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int col = lane + 64 * q;
            dst[q] = col < m ? row[col] : mk<double>(0, 0);
        }
    };
    double s0 = 0, s1 = 0;  // SOFTMAX: running maximum (s0) and scaled sum (s1)
    if (KIND == LOSS_SOFTMAX) s0 = -__builtin_inf();
    if (i0 < i1) load_row(g, i0);
    for (int64_t i = i0; i < i1; ++i) {
        if (i + 1 < i1) load_row(gn, i + 1);
        double re = 0, im = 0;
        if (KIND != LOSS_ADJOINT) {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const cx<double> x = ws[lane + 64 * q];
                re += g[q].re * x.re - g[q].im * x.im;
                im += g[q].re * x.im + g[q].im * x.re;
            }
            re = wave_sum_uniform(re);
            im = wave_sum_uniform(im);
        }
        const cx<double> ui = mk<double>(re, im);
        if (KIND != LOSS_ADJOINT && u_out && lane == 0) u_out[i] = ui;
        const double p2 = re * re + im * im;
        int slot = 0;
        bool take = true;
        cx<double> ci;
        if (KIND == LOSS_ADJOINT) {
            ci = a.cot[i];
        } else if (KIND == LOSS_FIELD_MATCH) {
            const double amp = sqrt(p2);
            const double res = amp - a.target[i];
            const double wt = a.weight[i];
            s0 += 0.5 * (wt * res) * (wt * res);
            ci = (wt * wt * res / (amp + a.eps)) * ui;
        } else if (KIND == LOSS_ENERGY) {
            s0 += p2;
            ci = ui;
        } else if (KIND == LOSS_SOFTMAX) {
            take = a.mask[i] != 0;
            if (take) {
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
            }
        } else {
            const unsigned char cls = a.mask[i];
            take = cls == 1 || cls == 2;
            slot = cls == 2 ? 1 : 0;
            if (cls == 1) s0 += p2;
            else if (cls == 2) s1 += p2;
            ci = ui;
        }
        if (take) {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (NACC == 1 || slot == 0) acc[0][q] = acc[0][q] + conj(g[q]) * ci;
                else acc[NACC - 1][q] = acc[NACC - 1][q] + conj(g[q]) * ci;
            }
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) g[q] = gn[q];
    }
    // Workgroup sum, waves 0..3 in order, softmax sums moved onto the
    // workgroup's largest maximum first.
    if (lane == 0) {
        rs[0][wv] = s0;
        rs[1][wv] = s1;
    }
    __syncthreads();
    double b0, b1, f = 1.0;
    if (KIND == LOSS_SOFTMAX) {
        b0 = fmax(fmax(rs[0][0], rs[0][1]), fmax(rs[0][2], rs[0][3]));
        b1 = 0;
        for (int v = 0; v < 4; ++v) b1 += rs[1][v] * softmax_rescale(rs[0][v], b0, a.temperature);
        f = softmax_rescale(s0, b0, a.temperature);
    } else {
        b0 = ((rs[0][0] + rs[0][1]) + rs[0][2]) + rs[0][3];
        b1 = ((rs[1][0] + rs[1][1]) + rs[1][2]) + rs[1][3];
    }
    for (int v = 0; v < 4; ++v) {
        if (wv == v) {
#pragma unroll
            for (int h = 0; h < NACC; ++h)
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const cx<double> mine = KIND == LOSS_SOFTMAX ? f * acc[h][q] : acc[h][q];
                    red[h][lane + 64 * q] = v == 0 ? mine : red[h][lane + 64 * q] + mine;
                }
        }
        __syncthreads();
    }
    const int64_t nblk = gridDim.x;
    for (int q = threadIdx.x; q < m; q += 256)
#pragma unroll
        for (int h = 0; h < NACC; ++h) partial[(h * nblk + blockIdx.x) * (int64_t)m + q] = red[h][q];
    if (threadIdx.x == 0) {
        scal[blockIdx.x * LOSS_SCALARS] = b0;
        scal[blockIdx.x * LOSS_SCALARS + 1] = b1;
    }
}

// Scalars of up to 64 partials, one per lane of the calling wave: the
// softmax maximum `top` and each partial's rescale factor (written to fac),
// and the sums t0, t1 in a fixed butterfly order (identical in every lane).
template <int KIND>
__device__ __forceinline__ void partial_scalars(const double *__restrict__ scal, int64_t p0, int64_t count,
                                                double temperature, double *fac, double &top, double &t0,
                                                double &t1) {
    const int lane = threadIdx.x & 63;
    const bool ok = lane < count;
    const double a0 = ok ? scal[(p0 + lane) * LOSS_SCALARS] : (KIND == LOSS_SOFTMAX ? -__builtin_inf() : 0.0);
    const double a1 = ok ? scal[(p0 + lane) * LOSS_SCALARS + 1] : 0.0;
    if (KIND == LOSS_SOFTMAX) {
        top = wave_max(a0);
        const double f = softmax_rescale(a0, top, temperature);
        fac[lane] = f;
        t0 = 0;
        t1 = wave_sum(a1 * f);
    } else {
        top = 0;
        t0 = wave_sum(a0);
        t1 = wave_sum(a1);
    }
}

// Folds nparts partials (layout of k_basis_loss) into nslices partials of
// the same layout: slice s sums parts [s nparts / nslices, (s + 1) nparts /
// nslices) in order (at most 64 parts a slice).  Grid (ceil(m / 64),
// nslices), 64 threads, one column per thread.
template <int KIND>
__global__ __launch_bounds__(64) void k_basis_loss_fold(const cx<double> *__restrict__ in,
                                                        const double *__restrict__ in_scal, int64_t nparts, int m,
                                                        double temperature, cx<double> *__restrict__ out,
                                                        double *__restrict__ out_scal, int64_t nslices) {
    constexpr int NACC = KIND == LOSS_RATIO ? 2 : 1;
    constexpr int BATCH = 16;
    __shared__ double fac[64];
    const int64_t s = blockIdx.y;
    const int64_t p0 = s * nparts / nslices, p1 = (s + 1) * nparts / nslices;
    const int q = blockIdx.x * 64 + threadIdx.x;
    double top, t0, t1;
    partial_scalars<KIND>(in_scal, p0, p1 - p0, temperature, fac, top, t0, t1);
    __syncthreads();
    cx<double> acc[NACC];
#pragma unroll
    for (int h = 0; h < NACC; ++h) acc[h] = mk<double>(0, 0);
    for (int64_t v = p0; v < p1; v += BATCH) {
        cx<double> x[NACC][BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            const bool ok = v + j < p1 && q < m;
#pragma unroll
            for (int h = 0; h < NACC; ++h) x[h][j] = ok ? in[(h * nparts + v + j) * (int64_t)m + q] : mk<double>(0, 0);
        }
#pragma unroll
        for (int j = 0; j < BATCH; ++j) {
            const double f = KIND == LOSS_SOFTMAX && v + j < p1 ? fac[v + j - p0] : 1.0;
#pragma unroll
            for (int h = 0; h < NACC; ++h) acc[h] = acc[h] + (KIND == LOSS_SOFTMAX ? f * x[h][j] : x[h][j]);
        }
    }
    if (q < m)
#pragma unroll
        for (int h = 0; h < NACC; ++h) out[(h * nslices + s) * (int64_t)m + q] = acc[h];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out_scal[s * LOSS_SCALARS] = KIND == LOSS_SOFTMAX ? top : t0;
        out_scal[s * LOSS_SCALARS + 1] = t1;
    }
}

// g[q] = G^H c from nparts <= 64 partials and the loss into *loss (none
// when loss is null).  Grid
// ceil(m / 64), 256 threads: wave w sums partials w, w + 4, ... of 64
// columns, the four wave sums are added in order in LDS.
template <int KIND>
__global__ __launch_bounds__(256) void k_basis_loss_combine(const cx<double> *__restrict__ partial,
                                                            const double *__restrict__ scal, int64_t nparts, int m,
                                                            BasisLossArgs a, cx<double> *__restrict__ g,
                                                            double *__restrict__ loss) {
    constexpr int NACC = KIND == LOSS_RATIO ? 2 : 1;
    __shared__ double fac[64];
    __shared__ double sc[3];
    __shared__ cx<double> red[NACC][4][64];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + lane;
    if (wv == 0) {
        double top, t0, t1;
        partial_scalars<KIND>(scal, 0, nparts, a.temperature, fac, top, t0, t1);
        if (lane == 0) {
            sc[0] = top;
            sc[1] = t0;
            sc[2] = t1;
        }
    }
    __syncthreads();
    const double top = sc[0], t0 = sc[1], t1 = sc[2];
    double lval = 0, f0 = 1, f1 = 0;
    if (KIND == LOSS_FIELD_MATCH) {
        lval = t0;
    } else if (KIND == LOSS_ENERGY) {
        lval = 0.5 * a.scale * t0;
        f0 = a.scale;
    } else if (KIND == LOSS_SOFTMAX) {
        lval = a.scale * (top + a.temperature * (log(t1) - log(a.count_a)));
        f0 = a.scale / t1;
    } else if (KIND == LOSS_RATIO) {
        const double ma = t0 / a.count_a, mb = t1 / a.count_b;
        lval = ma / mb;
        f0 = 2.0 / (a.count_a * mb);
        f1 = -2.0 * ma / (a.count_b * mb * mb);
    }
    cx<double> acc[NACC];
#pragma unroll
    for (int h = 0; h < NACC; ++h) acc[h] = mk<double>(0, 0);
#pragma unroll 4
    for (int64_t v = wv; v < nparts; v += 4) {
#pragma unroll
        for (int h = 0; h < NACC; ++h) {
            const cx<double> p = q < m ? partial[(h * nparts + v) * (int64_t)m + q] : mk<double>(0, 0);
            acc[h] = acc[h] + (KIND == LOSS_SOFTMAX ? fac[v] * p : p);
        }
    }
#pragma unroll
    for (int h = 0; h < NACC; ++h) red[h][wv][lane] = acc[h];
    __syncthreads();
    if (wv == 0 && q < m) {
        cx<double> sum[NACC];
#pragma unroll
        for (int h = 0; h < NACC; ++h) sum[h] = ((red[h][0][lane] + red[h][1][lane]) + red[h][2][lane]) + red[h][3][lane];
        g[q] = KIND == LOSS_RATIO ? f0 * sum[0] + f1 * sum[NACC - 1] : f0 * sum[0];
    }
    if (loss && blockIdx.x == 0 && threadIdx.x == 0) *loss = lval;
}

// ------------------------------------------------ phase step of an optimiser
// The parts of one phase-optimiser evaluation around the fused pass over G
// (PhaseProblem.evaluate, src/algorithm/baselines/common.py:275-343), each
// one workgroup: the prologue w = e^{i phi} and the exact focal samples
// f = R w of a (p x m, row-major) target response R; the epilogue adds the
// focal term -tw mean_t |f_t| to the grid loss and forms the phase gradient
// Re(conj(g + R^H c) i w) with c_t = -tw/p f_t/(|f_t| + eps)
// (common.py:189-195), g = G^H c of the grid loss.
constexpr int PHASE_MAX_TARGETS = 256;

__global__ __launch_bounds__(256) void k_phase_prologue(const double *__restrict__ phi, int m,
        const cx<double> *__restrict__ R, int p, cx<double> *__restrict__ w, cx<double> *__restrict__ f) {
    __shared__ cx<double> sw[BASIS_FUSED_MAX_M];
    __shared__ double red[4];
    for (int j = threadIdx.x; j < m; j += 256) {
        double s, c;
        sincos(phi[j], &s, &c);
        sw[j] = mk<double>(c, s);
        w[j] = sw[j];
    }
    __syncthreads();
    for (int t = 0; t < p; ++t) {
        cx<double> acc = mk<double>(0, 0);
        for (int j = threadIdx.x; j < m; j += 256) acc = cfma(R[(int64_t)t * m + j], sw[j], acc);
        const double re = block_sum_256(acc.re, red);
        const double im = block_sum_256(acc.im, red);
        if (threadIdx.x == 0) f[t] = mk<double>(re, im);
    }
}

__global__ __launch_bounds__(256) void k_phase_epilogue(const cx<double> *__restrict__ w, int m,
        const cx<double> *__restrict__ g, const cx<double> *__restrict__ R, int p, const cx<double> *__restrict__ f,
        double tw, double eps, const double *__restrict__ loss_in, int with_gradient, double *__restrict__ out) {
    __shared__ cx<double> scot[PHASE_MAX_TARGETS];
    const double scale = p > 0 ? -tw / p : 0.0;
    for (int t = threadIdx.x; t < p; t += 256) scot[t] = (scale / (hypot(f[t].re, f[t].im) + eps)) * f[t];
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int t = 0; t < p; ++t) sum += hypot(f[t].re, f[t].im);
        out[0] = p > 0 ? loss_in[0] + -tw * (sum / p) : loss_in[0];
    }
    __syncthreads();
    if (!with_gradient) return;
    for (int j = threadIdx.x; j < m; j += 256) {
        cx<double> a = g[j];
        for (int t = 0; t < p; ++t) a = cfma(conj(R[(int64_t)t * m + j]), scot[t], a);
        out[1 + j] = a.im * w[j].re - a.re * w[j].im;
    }
}

}  // namespace hz
