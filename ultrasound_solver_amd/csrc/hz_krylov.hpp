// Restarted-GMRES building blocks for k right-hand sides at once.
//
// Basis layout: slot (i, r) of the Krylov basis is the N-vector
// V + (i*k + r)*N.  Basis vectors are stored unnormalised ("raw") with a
// per-(r, i) scale s such that v_i = s_i * V_i; the normalisation is folded
// into the next stencil apply and into the Gram-Schmidt coefficients, which
// saves a full read+write pass per Arnoldi step.
//
// Reductions: each thread accumulates in the vector precision, a wavefront
// reduces by lane shuffles, and workgroup and grid partials are combined in
// FP64 in a fixed order (reproducible).  The small Hessenberg / Givens least
// squares runs in FP64 on the device, one workgroup per right-hand side.
#pragma once

#include "hz_common.hpp"

namespace hz {

constexpr int DOT_GROUP = 16;  // basis vectors per dot-product pass

// partial[((r*maxv + i)*nblk + blk)] = sum over the workgroup's chunk of
// conj(V_i[r]) * w[r], for i in the group [g*16, min(g*16+16, jp1)).
// VW = 2 (complex64 only; n and chunk even): each lane loads two complex
// values per 16-byte access.
template <typename T, int VW>
__global__ __launch_bounds__(256) void k_dots(const cx<T> *__restrict__ V, int k, int64_t n, int jp1,
                                              const cx<T> *__restrict__ w, const int *__restrict__ live,
                                              double2 *__restrict__ partial, int maxv, int64_t chunk) {
    __shared__ double red[4];
    const int r = blockIdx.y;
    if (!live[r]) return;
    const int g0 = blockIdx.z * DOT_GROUP;
    const int cnt = min(DOT_GROUP, jp1 - g0);
    if (cnt <= 0) return;
    const int64_t p0 = blockIdx.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    const cx<T> *wr = w + (int64_t)r * n;
    T are[DOT_GROUP], aim[DOT_GROUP];
#pragma unroll
    for (int q = 0; q < DOT_GROUP; ++q) are[q] = aim[q] = 0;
    if constexpr (VW == 2) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        const float4 *w4 = reinterpret_cast<const float4 *>(wr);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            const float4 wv = w4[p];
#pragma unroll
            for (int q = 0; q < DOT_GROUP; ++q) {
                if (q < cnt) {
                    const float4 v = reinterpret_cast<const float4 *>(V + ((int64_t)(g0 + q) * k + r) * n)[p];
                    are[q] += v.x * wv.x + v.y * wv.y + v.z * wv.z + v.w * wv.w;
                    aim[q] += v.x * wv.y - v.y * wv.x + v.z * wv.w - v.w * wv.z;
                }
            }
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            const cx<T> wv = wr[p];
#pragma unroll
            for (int q = 0; q < DOT_GROUP; ++q) {
                if (q < cnt) {
                    const cx<T> v = V[((int64_t)(g0 + q) * k + r) * n + p];
                    are[q] += v.re * wv.re + v.im * wv.im;
                    aim[q] += v.re * wv.im - v.im * wv.re;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < DOT_GROUP; ++q) {
        if (q < cnt) {
            // FP32 (or FP64) wavefront sums, combined across wavefronts in FP64.
            const double sre = block_combine_256((double)wave_sum(are[q]), red);
            const double sim = block_combine_256((double)wave_sum(aim[q]), red);
            if (threadIdx.x == 0)
                partial[((int64_t)r * maxv + g0 + q) * gridDim.x + blockIdx.x] = make_double2(sre, sim);
        }
    }
}

// Combine dot partials: d_i = sum_blk partial; H[r](i, j) = s_i d_i and the
// update coefficient coef[r][i] = s_i^2 d_i.  Grid (jp1, k).
__global__ __launch_bounds__(256) void k_reduce_dots(const double2 *__restrict__ partial, int nblk, int maxv,
                                                     int j, const double *__restrict__ scales,
                                                     const int *__restrict__ live, cx<double> *__restrict__ H,
                                                     cx<double> *__restrict__ coef, int restart) {
    __shared__ double red[4];
    const int i = blockIdx.x;
    const int r = blockIdx.y;
    if (!live[r]) return;
    const double2 *src = partial + ((int64_t)r * maxv + i) * nblk;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        re += src[b].x;
        im += src[b].y;
    }
    re = block_sum_256(re, red);
    im = block_sum_256(im, red);
    if (threadIdx.x == 0) {
        const double s = scales[(int64_t)i * gridDim.y + r];
        const cx<double> d = mk<double>(re, im);
        H[((int64_t)r * (restart + 1) + i) * restart + j] = s * d;
        coef[r * maxv + i] = (s * s) * d;
    }
}

// target[r] -= sum_{i<cnt} coef[r][i] * V_i[r]; optional FP64 partial of
// |target|^2 per workgroup (norm_partial[r*nblk + blk]).  Grid (nblk, k).
// VW = 2 as in k_dots (target leading dimension even as well).
template <typename T, int VW>
__global__ __launch_bounds__(256) void k_update(const cx<T> *__restrict__ V, int k, int64_t n, int cnt_all,
                                                const int *__restrict__ cnt_rhs, cx<T> *__restrict__ target,
                                                int64_t ldt, const cx<double> *__restrict__ coef, int maxv,
                                                const int *__restrict__ live, double *__restrict__ norm_partial,
                                                int64_t chunk) {
    __shared__ cx<T> cf[64];
    __shared__ double red[4];
    const int r = blockIdx.y;
    if (!live[r]) return;
    const int cnt = cnt_rhs ? cnt_rhs[r] : cnt_all;
    for (int q = threadIdx.x; q < cnt; q += 256) cf[q] = cast<T>(coef[r * maxv + q]);
    __syncthreads();
    const int64_t p0 = blockIdx.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    cx<T> *tr = target + (int64_t)r * ldt;
    T nrm = 0;
    if constexpr (VW == 2) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        float4 *t4 = reinterpret_cast<float4 *>(tr);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            float4 acc = t4[p];
            int q = 0;
            for (; q + 8 <= cnt; q += 8) {
                float4 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    v[u] = reinterpret_cast<const float4 *>(V + ((int64_t)(q + u) * k + r) * n)[p];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const cx<float> c = cf[q + u];
                    acc.x -= c.re * v[u].x - c.im * v[u].y;
                    acc.y -= c.re * v[u].y + c.im * v[u].x;
                    acc.z -= c.re * v[u].z - c.im * v[u].w;
                    acc.w -= c.re * v[u].w + c.im * v[u].z;
                }
            }
            for (; q + 4 <= cnt; q += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    v[u] = reinterpret_cast<const float4 *>(V + ((int64_t)(q + u) * k + r) * n)[p];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const cx<float> c = cf[q + u];
                    acc.x -= c.re * v[u].x - c.im * v[u].y;
                    acc.y -= c.re * v[u].y + c.im * v[u].x;
                    acc.z -= c.re * v[u].z - c.im * v[u].w;
                    acc.w -= c.re * v[u].w + c.im * v[u].z;
                }
            }
            for (; q < cnt; ++q) {
                const float4 v = reinterpret_cast<const float4 *>(V + ((int64_t)q * k + r) * n)[p];
                const cx<float> c = cf[q];
                acc.x -= c.re * v.x - c.im * v.y;
                acc.y -= c.re * v.y + c.im * v.x;
                acc.z -= c.re * v.z - c.im * v.w;
                acc.w -= c.re * v.w + c.im * v.z;
            }
            t4[p] = acc;
            nrm += acc.x * acc.x + acc.y * acc.y + acc.z * acc.z + acc.w * acc.w;
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            cx<T> acc = tr[p];
            int q = 0;
            for (; q + 4 <= cnt; q += 4) {
                const cx<T> v0 = V[((int64_t)(q + 0) * k + r) * n + p];
                const cx<T> v1 = V[((int64_t)(q + 1) * k + r) * n + p];
                const cx<T> v2 = V[((int64_t)(q + 2) * k + r) * n + p];
                const cx<T> v3 = V[((int64_t)(q + 3) * k + r) * n + p];
                acc = acc - cf[q] * v0;
                acc = acc - cf[q + 1] * v1;
                acc = acc - cf[q + 2] * v2;
                acc = acc - cf[q + 3] * v3;
            }
            for (; q < cnt; ++q) acc = acc - cf[q] * V[((int64_t)q * k + r) * n + p];
            tr[p] = acc;
            nrm += norm2(acc);
        }
    }
    if (norm_partial) {
        const double s = block_combine_256((double)wave_sum(nrm), red);
        if (threadIdx.x == 0) norm_partial[(int64_t)r * gridDim.x + blockIdx.x] = s;
    }
}

// Per-RHS GMRES state, FP64, on the device.
struct GmresState {
    cx<double> *H;     // [k][restart+1][restart]
    double *gc;        // Givens cosines [k][restart]
    cx<double> *gs;    // Givens sines   [k][restart]
    cx<double> *g;     // rotated RHS    [k][restart+1]
    double *scales;    // basis scales   [maxv][k] (vector i of RHS r at i*k + r)
    cx<double> *coef;  // update coefficients [k][maxv]
    double *resid;     // residual estimate / true residual [k]
    double *tol;       // early-exit tolerance (absolute) [k], < 0 disables
    int *steps;        // Arnoldi steps done in this cycle [k]
    int *live;         // RHS still iterating in this cycle [k]
};

// After the cycle-top residual: beta0 = ||r||, s_0 = 1/beta0, g = beta0 e_1.
__global__ __launch_bounds__(256) void k_cycle_start(GmresState st, const double *__restrict__ partial, int nblk,
                                                     int maxv, int restart) {
    __shared__ double red[4];
    const int r = blockIdx.x;
    if (!st.live[r]) return;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(int64_t)r * nblk + b];
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) {
        const double beta = sqrt(s);
        st.resid[r] = beta;
        st.scales[r] = beta > 0 ? 1.0 / beta : 0.0;
        st.g[r * (restart + 1)] = mk<double>(beta, 0);
        for (int q = 1; q <= restart; ++q) st.g[r * (restart + 1) + q] = mk<double>(0, 0);
        st.steps[r] = 0;
    }
}

// Finish Arnoldi step j: beta = ||w||, H(j+1, j) = beta, Givens update,
// residual estimate |g_{j+1}|, early exit when it meets st.tol[r].
__global__ __launch_bounds__(256) void k_arnoldi_finish(GmresState st, const double *__restrict__ norm_partial,
                                                        int nblk, int j, int maxv, int restart) {
    __shared__ double red[4];
    const int r = blockIdx.x;
    if (!st.live[r]) return;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += norm_partial[(int64_t)r * nblk + b];
    s = block_sum_256(s, red);
    if (threadIdx.x != 0) return;
    const double beta = sqrt(s);
    cx<double> *col = st.H + (int64_t)r * (restart + 1) * restart;  // H(i, j) at col[i*restart + j]
    col[(j + 1) * restart + j] = mk<double>(beta, 0);
    st.scales[(int64_t)(j + 1) * gridDim.x + r] = beta > 0 ? 1.0 / beta : 0.0;
    // Apply the previous rotations to column j.
    for (int q = 0; q < j; ++q) {
        const double c = st.gc[r * restart + q];
        const cx<double> sn = st.gs[r * restart + q];
        const cx<double> a = col[q * restart + j];
        const cx<double> b = col[(q + 1) * restart + j];
        col[q * restart + j] = c * a + sn * b;
        col[(q + 1) * restart + j] = c * b - conj(sn) * a;
    }
    // New rotation zeroing H(j+1, j) = beta (real, >= 0).
    const cx<double> a = col[j * restart + j];
    const double aa = cabs(a);
    double c;
    cx<double> sn;
    if (beta == 0.0) {
        c = 1.0;
        sn = mk<double>(0, 0);
    } else if (aa == 0.0) {
        c = 0.0;
        sn = mk<double>(1, 0);
    } else {
        const double t = sqrt(aa * aa + beta * beta);
        c = aa / t;
        sn = (beta / t) * ((1.0 / aa) * a);  // (a/|a|) * beta / t
    }
    st.gc[r * restart + j] = c;
    st.gs[r * restart + j] = sn;
    col[j * restart + j] = c * a + sn * mk<double>(beta, 0);
    col[(j + 1) * restart + j] = mk<double>(0, 0);
    cx<double> *g = st.g + r * (restart + 1);
    const cx<double> gj = g[j];
    g[j] = c * gj;
    g[j + 1] = (-1.0) * (conj(sn) * gj);
    const double est = cabs(g[j + 1]);
    st.resid[r] = est;
    st.steps[r] = j + 1;
    if ((st.tol[r] >= 0 && est <= st.tol[r]) || beta == 0.0) st.live[r] = 0;
}

// Back substitution on the rotated Hessenberg matrix; coef[r][i] = -y_i s_i so
// that k_update adds sum_i y_i v_i to the iterate.
__global__ void k_trisolve(GmresState st, int maxv, int restart, const int *__restrict__ member) {
    const int r = blockIdx.x;
    if (threadIdx.x != 0 || !member[r]) return;
    const int m = st.steps[r];
    const cx<double> *col = st.H + (int64_t)r * (restart + 1) * restart;
    const cx<double> *g = st.g + r * (restart + 1);
    cx<double> y[128];
    for (int i = m - 1; i >= 0; --i) {
        cx<double> acc = g[i];
        for (int q = i + 1; q < m; ++q) acc = acc - col[i * restart + q] * y[q];
        const cx<double> d = col[i * restart + i];
        y[i] = (cabs(d) > 0) ? cdiv(acc, d) : mk<double>(0, 0);
    }
    for (int i = 0; i < m; ++i) st.coef[r * maxv + i] = (-st.scales[(int64_t)i * gridDim.x + r]) * y[i];
}

// ---------------------------------------------------------------- elementwise
// out[r] = cast(rs * b[r]) (scaled right-hand side) with FP64 partials of
// |rs b|^2 and |b|^2 (partial[(r*nblk+blk)*2 + 0/1]).
template <typename T>
__global__ __launch_bounds__(256) void k_scale_rhs(const cx<double> *__restrict__ b, int64_t ldb,
                                                   const double *__restrict__ rs, cx<T> *__restrict__ out,
                                                   int64_t ldo, int64_t n, double *__restrict__ partial,
                                                   int64_t chunk) {
    __shared__ double red[4];
    const int r = blockIdx.y;
    const int64_t p0 = blockIdx.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    double s1 = 0, s2 = 0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
        const cx<double> v = b[r * ldb + p];
        const cx<double> sv = rs[p] * v;
        out[r * ldo + p] = cast<T>(sv);
        s1 += norm2(sv);
        s2 += norm2(v);
    }
    s1 = block_sum_256(s1, red);
    s2 = block_sum_256(s2, red);
    if (threadIdx.x == 0) {
        partial[((int64_t)r * gridDim.x + blockIdx.x) * 2 + 0] = s1;
        partial[((int64_t)r * gridDim.x + blockIdx.x) * 2 + 1] = s2;
    }
}

// x[r] (+)= m * cast<double>(y[r]): apply the Jacobi factor to the iterate.
template <typename T>
__global__ __launch_bounds__(256) void k_precond_out(const cx<T> *__restrict__ y, int64_t ldy,
                                                     const cx<double> *__restrict__ m, cx<double> *__restrict__ x,
                                                     int64_t ldx, int64_t n, const int *__restrict__ member,
                                                     int accumulate) {
    const int r = blockIdx.y;
    if (member && !member[r]) return;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
        const cx<double> v = m[p] * cast<double>(y[r * ldy + p]);
        x[r * ldx + p] = accumulate ? x[r * ldx + p] + v : v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(cx<T> *__restrict__ x, int64_t ldx, int64_t n) {
    const int r = blockIdx.y;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
        x[r * ldx + p] = mk<T>(0, 0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(cx<T> *__restrict__ x, int64_t ldx, int64_t n, cx<T> value) {
    const int r = blockIdx.y;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) x[r * ldx + p] = value;
}

// Sum (or max, where bit q of max_mask is set) of nblk partials for each of
// `width` values per RHS: out[r*width + q] = reduce_b partial[(r*nblk + b)*width + q].
// Grid (width, k).
__global__ __launch_bounds__(256) void k_reduce_rows(const double *__restrict__ partial, int nblk, int width,
                                                     int max_mask, double *__restrict__ out) {
    __shared__ double red[4];
    const int q = blockIdx.x;
    const int r = blockIdx.y;
    const bool is_max = (max_mask >> q) & 1;
    double acc = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double v = partial[((int64_t)r * nblk + b) * width + q];
        acc = is_max ? fmax(acc, v) : acc + v;
    }
    if (is_max) {
        double w = wave_max(acc);
        if ((threadIdx.x % WAVE) == 0) red[threadIdx.x / WAVE] = w;
        __syncthreads();
        if (threadIdx.x == 0) out[r * width + q] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    } else {
        acc = block_sum_256(acc, red);
        if (threadIdx.x == 0) out[r * width + q] = acc;
    }
}

// ------------------------------------------------------------ sources / sampling
// b[node, r] = sum over the node's entries of value * w[column, r]: the
// deposition list grouped by node (row_ptr), so no atomics are needed.  The
// caller zeroes b first.
__global__ void k_scatter_sources(const int64_t *__restrict__ row_node, const int64_t *__restrict__ row_ptr,
                                  const int *__restrict__ column, const double *__restrict__ value, int64_t nrows,
                                  const cx<double> *__restrict__ w, int64_t ldw, cx<double> *__restrict__ b,
                                  int64_t ldb) {
    const int r = blockIdx.y;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nrows; q += (int64_t)gridDim.x * blockDim.x) {
        cx<double> acc = mk<double>(0, 0);
        for (int64_t e = row_ptr[q]; e < row_ptr[q + 1]; ++e) acc = acc + value[e] * w[r * ldw + column[e]];
        b[r * ldb + row_node[q]] = acc;
    }
}

// out[c, r] = sum over column c's entries of value * lambda[node, r].  One
// workgroup per (column, rhs); entries of a column are contiguous (col_ptr).
__global__ __launch_bounds__(256) void k_gather_sources(const int64_t *__restrict__ node,
                                                        const double *__restrict__ value,
                                                        const int64_t *__restrict__ col_ptr,
                                                        const cx<double> *__restrict__ lam, int64_t ld,
                                                        cx<double> *__restrict__ out, int64_t ldo) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    const int r = blockIdx.y;
    double re = 0, im = 0;
    for (int64_t e = col_ptr[c] + threadIdx.x; e < col_ptr[c + 1]; e += 256) {
        const cx<double> v = lam[r * ld + node[e]];
        re += value[e] * v.re;
        im += value[e] * v.im;
    }
    re = block_sum_256(re, red);
    im = block_sum_256(im, red);
    if (threadIdx.x == 0) out[r * ldo + c] = mk<double>(re, im);
}

// out[q, r] = sum_c w[q, c] u[rows[q, c], r] over the 8 trilinear corners.
__global__ void k_sample(const int64_t *__restrict__ rows, const double *__restrict__ wts, int npts,
                         const cx<double> *__restrict__ u, int64_t ld, cx<double> *__restrict__ out, int64_t ldo) {
    const int r = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npts) return;
    cx<double> acc = mk<double>(0, 0);
    for (int c = 0; c < 8; ++c) acc = acc + wts[q * 8 + c] * u[r * ld + rows[q * 8 + c]];
    out[r * ldo + q] = acc;
}

}  // namespace hz
