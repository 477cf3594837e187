// Homogeneous surface condensation (reference:
// src/algorithm/solvers/homogeneous_surface.py).  In a homogeneous medium
// the finite-volume operator is a Kronecker sum of three 1-D operators, so
// with the x and y eigenvector bases every (x, y) mode decouples into a
// tridiagonal z system.  Its solution below the baffle is u[j] = H[j-1] u[0]
// with the transfer ratios H from a backward continued fraction, and u[0]
// follows from the surface Schur complement.  Device side: the per-mode
// recurrence, the modal transforms (two-sided products with the eigenvector
// matrices, complex FP64), the sparse source products and the trilinear
// sampling of planes.  Planes are nx x ny, x fastest (column-major), like a
// z-slice of a field vector.
#pragma once

#include "hz_common.hpp"

namespace hz {

// a / b by Smith's scaled algorithm (R. L. Smith, CACM 5(8):435, 1962):
// no overflow or underflow of |b|^2 for widely scaled operands.
__device__ __forceinline__ cx<double> cdiv_smith(cx<double> a, cx<double> b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return mk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return mk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}

// One thread per mode (reference CUDA kernel homogeneous_surface.py:41-82
// and the CPU loop :151-163).  Backward pass: ratio_j = -lower[j-1] /
// (shift + diag[j] + upper[j] ratio_{j+1}), stored as H[j-1]; forward pass:
// prefix products H[j] *= H[j-1]; then the Schur complement
// shift + diag[0] + upper[0] H[0] (:164).  H is (nz-1) x modes, mode fastest.
// The forward pass reads 8 ratios ahead so that its loads overlap.
__global__ __launch_bounds__(256) void k_z_transfer(const cx<double> *__restrict__ shift,
                                                    const cx<double> *__restrict__ diag,
                                                    const cx<double> *__restrict__ upper,
                                                    const cx<double> *__restrict__ lower, int nz, int64_t modes,
                                                    cx<double> *__restrict__ h, cx<double> *__restrict__ schur) {
    const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (m >= modes) return;
    const cx<double> s = shift[m];
    cx<double> ratio = mk(0.0, 0.0);
    for (int j = nz - 1; j >= 1; --j) {
        cx<double> d = s + diag[j];
        if (j < nz - 1) d = d + upper[j] * ratio;
        ratio = cdiv_smith(mk(-lower[j - 1].re, -lower[j - 1].im), d);
        h[(int64_t)(j - 1) * modes + m] = ratio;
    }
    const cx<double> h0 = ratio;  // H[0], unchanged by the prefix products
    cx<double> prev = h0;
    constexpr int AHEAD = 8;
    for (int j0 = 1; j0 < nz - 1; j0 += AHEAD) {
        cx<double> r[AHEAD];
#pragma unroll
        for (int u = 0; u < AHEAD; ++u)
            if (j0 + u < nz - 1) r[u] = h[(int64_t)(j0 + u) * modes + m];
#pragma unroll
        for (int u = 0; u < AHEAD; ++u)
            if (j0 + u < nz - 1) {
                prev = r[u] * prev;
                h[(int64_t)(j0 + u) * modes + m] = prev;
            }
    }
    schur[m] = (s + diag[0]) + upper[0] * h0;
}

// Batched complex FP64 product C[p] = A[p] B[p] (M x N, K inner) with
// element strides, so transposed operands need no copy:
//   A(i, k) = A[i*a_si + k*a_sk] (times W(i, k) when W is given, same strides),
//   B(k, j) = B[k*b_sk + j*b_sj],  C(i, j) = C[i + j*ldc] (divided by D(i, j)
//   at the end when D is given).
// Batch p uses A + p*a_sp, B + p*b_sp, C + p*c_sp and W + sel[p]*w_sp (no
// weight for sel[p] < 0; W + p*w_sp without sel).
struct ZgemmArgs {
    int m, n, k;
    const cx<double> *a;
    int64_t a_si, a_sk, a_sp;
    const cx<double> *w;
    int64_t w_sp;
    const int *sel;
    const cx<double> *b;
    int64_t b_sk, b_sj, b_sp;
    cx<double> *c;
    int64_t ldc, c_sp;
    const cx<double> *d;
};

// A 64 x 64 tile of C per 256-thread workgroup; thread (tx, ty) = (t % 16,
// t / 16) owns rows tx + 16u and columns ty + 16v (u, v < 4), so each K step
// reads 8 complex values from LDS for 16 complex multiply-adds.  K advances
// by 8; the next K tile is loaded into registers while the current one is
// multiplied.  Global loads run along whichever operand index has unit
// stride, so transposed operands stay coalesced.  gfx950's FP64 matrix rate
// equals its FP64 vector rate, so vector FMAs are used.
constexpr int ZG_TILE = 64, ZG_KS = 8;

// acc += a b as four fused multiply-adds (two dependent ones per component).
__device__ __forceinline__ void cmac(cx<double> &acc, cx<double> a, cx<double> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

__global__ __launch_bounds__(256) void k_zgemm(ZgemmArgs g) {
    __shared__ cx<double> sa[ZG_KS][ZG_TILE + 1];
    __shared__ cx<double> sb[ZG_KS][ZG_TILE + 1];
    const int p = blockIdx.z;
    const int i0 = blockIdx.x * ZG_TILE, j0 = blockIdx.y * ZG_TILE;
    const cx<double> *A = g.a + p * g.a_sp;
    const cx<double> *B = g.b + p * g.b_sp;
    const cx<double> *W = nullptr;
    if (g.w) {
        const int64_t slot = g.sel ? g.sel[p] : p;
        if (slot >= 0) W = g.w + slot * g.w_sp;
    }
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const bool a_rows = g.a_si == 1;  // A loads along i (else along k)
    const bool b_rows = g.b_sk == 1;  // B loads along k (else along j)
    // Element e (0, 1) of this thread's share of a K tile: (row, k) of A and (k, col) of B.
    int ai[2], ak[2], bk[2], bj[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int idx = tid + e * 256;
        if (a_rows) ai[e] = idx & 63, ak[e] = idx >> 6;
        else ak[e] = idx & 7, ai[e] = idx >> 3;
        if (b_rows) bk[e] = idx & 7, bj[e] = idx >> 3;
        else bj[e] = idx & 63, bk[e] = idx >> 6;
    }
    auto load_a = [&](int k0, int e) {
        const int gi = i0 + ai[e], gk = k0 + ak[e];
        cx<double> v = mk(0.0, 0.0);
        if (gi < g.m && gk < g.k) {
            const int64_t off = gi * g.a_si + gk * g.a_sk;
            v = A[off];
            if (W) v = W[off] * v;
        }
        return v;
    };
    auto load_b = [&](int k0, int e) {
        const int gk = k0 + bk[e], gj = j0 + bj[e];
        return (gk < g.k && gj < g.n) ? B[gk * g.b_sk + gj * g.b_sj] : mk(0.0, 0.0);
    };
    cx<double> acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = mk(0.0, 0.0);
    cx<double> ra[2], rb[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) ra[e] = load_a(0, e), rb[e] = load_b(0, e);
    for (int k0 = 0; k0 < g.k; k0 += ZG_KS) {
        __syncthreads();  // the previous tile has been read
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            sa[ak[e]][ai[e]] = ra[e];
            sb[bk[e]][bj[e]] = rb[e];
        }
        __syncthreads();
        if (k0 + ZG_KS < g.k) {
#pragma unroll
            for (int e = 0; e < 2; ++e) ra[e] = load_a(k0 + ZG_KS, e), rb[e] = load_b(k0 + ZG_KS, e);
        }
#pragma unroll
        for (int kk = 0; kk < ZG_KS; ++kk) {
            cx<double> a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = sa[kk][tx + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; ++v) b[v] = sb[kk][ty + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) cmac(acc[u][v], a[u], b[v]);
        }
    }
    cx<double> *C = g.c + p * g.c_sp;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int gi = i0 + tx + 16 * u, gj = j0 + ty + 16 * v;
            if (gi < g.m && gj < g.n) {
                const int64_t off = gi + gj * g.ldc;
                cx<double> r = acc[u][v];
                if (g.d) r = cdiv_smith(r, g.d[off]);
                C[off] = r;
            }
        }
}

// Y[p][r] = scale * sum_e val[e] X[p][col[e]] over the CSR row r, one wave
// per (row, batch) pair.  As the surface load (rows = plane nodes, X = the
// emitter weights) and as the emitter response (rows = emitters, X = the
// observation kernels) (homogeneous_surface.py:203-204, 220-221).
__global__ __launch_bounds__(256) void k_csr_rows(int nrows, int count, const int64_t *__restrict__ ptr,
                                                  const int *__restrict__ col, const double *__restrict__ val,
                                                  const cx<double> *__restrict__ x, int64_t ldx, double scale,
                                                  cx<double> *__restrict__ y, int64_t ldy) {
    const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= (int64_t)nrows * count) return;  // whole wave
    const int r = (int)(wid % nrows), p = (int)(wid / nrows);
    const cx<double> *xp = x + p * ldx;
    double re = 0.0, im = 0.0;
    for (int64_t e = ptr[r] + lane; e < ptr[r + 1]; e += WAVE) {
        const cx<double> v = xp[col[e]];
        re += val[e] * v.re;
        im += val[e] * v.im;
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) y[p * ldy + r] = mk(scale * re, scale * im);
}

// Observation kernels of the target response (homogeneous_surface.py:198-202):
// O_p(a, b) = vx_p[a] vy_p[b] hz_p(a, b) / schur(a, b) with the rows of the
// eigenvector matrices and the transfer ratios interpolated linearly at the
// point (low corner lo, fractions fr; hz = 1 at the surface).
__global__ __launch_bounds__(256) void k_observation(int nx, int ny, int nz, const cx<double> *__restrict__ vx,
                                                     const cx<double> *__restrict__ vy,
                                                     const cx<double> *__restrict__ h,
                                                     const cx<double> *__restrict__ schur,
                                                     const int *__restrict__ lo, const double *__restrict__ fr,
                                                     cx<double> *__restrict__ out) {
    const int p = blockIdx.y;
    const int64_t nxy = (int64_t)nx * ny;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nxy) return;
    const int a = (int)(q % nx), b = (int)(q / nx);
    const int x = lo[3 * p], y = lo[3 * p + 1], z = lo[3 * p + 2];
    const double fx = fr[3 * p], fy = fr[3 * p + 1], fz = fr[3 * p + 2];
    // Row x of vx is vx[x + nx * a] (column-major).
    const cx<double> va = (1.0 - fx) * vx[x + (int64_t)nx * a] + fx * vx[x + 1 + (int64_t)nx * a];
    const cx<double> vb = (1.0 - fy) * vy[y + (int64_t)ny * b] + fy * vy[y + 1 + (int64_t)ny * b];
    const cx<double> h_lo = z == 0 ? mk(1.0, 0.0) : h[(int64_t)(z - 1) * nxy + q];
    const cx<double> hz = (1.0 - fz) * h_lo + fz * h[(int64_t)z * nxy + q];
    out[p * nxy + q] = cdiv_smith((va * vb) * hz, schur[q]);
}

// Trilinear samples from computed planes (homogeneous_surface.py:250-259):
// planes[slot[2p + dz]] is plane lo_z + dz of point p.
__global__ __launch_bounds__(256) void k_sample_planes(int nx, int ny, int npoints, const cx<double> *__restrict__ planes,
                                                       const int *__restrict__ slot, const int *__restrict__ lo,
                                                       const double *__restrict__ fr, cx<double> *__restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npoints) return;
    const int64_t nxy = (int64_t)nx * ny;
    const int x = lo[3 * p], y = lo[3 * p + 1];
    const double fx = fr[3 * p], fy = fr[3 * p + 1], fz = fr[3 * p + 2];
    cx<double> acc = mk(0.0, 0.0);
    for (int dz = 0; dz < 2; ++dz) {
        const cx<double> *pl = planes + slot[2 * p + dz] * nxy;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const double w = ((a ? fx : 1.0 - fx) * (b ? fy : 1.0 - fy)) * (dz ? fz : 1.0 - fz);
                acc = acc + w * pl[(x + a) + (int64_t)nx * (y + b)];
            }
    }
    out[p] = acc;
}

}  // namespace hz
