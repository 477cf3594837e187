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

constexpr int DOT_GROUP = 16;  // basis vectors per dot-product pass (complex64)
// Complex128 dot passes take 4 vectors per workgroup: FP64 accumulators of
// 16 vectors (64 VGPRs) leave room for few loads in flight, and the extra
// reads of w, once per group, come from the caches (GMRES(30) at 128^3, one
// column: 269.5, 258.3, 255.6 and 264.7 us per Arnoldi step with groups of
// 16, 8, 4 and 2).  The sums, and so the results, do not depend on the group.
constexpr int DOT_GROUP_F64 = 4;
template <typename T>
constexpr int dot_group() {
    return sizeof(T) == 8 ? DOT_GROUP_F64 : DOT_GROUP;
}

// Vector width VW of the Gram-Schmidt passes: 1 one complex per access, 2
// two (16-byte accesses, complex64 with even N), 3 the same with
// non-temporal accesses, 4 one complex per access, the basis read with
// non-temporal loads.  The host picks 3 where a basis vector (k columns
// of N) is at least STREAM_MIN_BYTES: each element is read or written once
// per launch and the vectors are far larger than the caches (128^3 x 64
// columns: the two-step update pass 853 -> 802 ms and the dot pass 598 ->
// 569 ms per solve of 64 columns).  Smaller vectors keep cached accesses, so
// that what one kernel writes can still be read from the last-level cache by
// the next (the 32-scene batch of 64^3-128^3 scenes at 16 columns ran 5 %
// slower streamed).  The FP64 plain form (one complex128 column, 33.5 MB at
// 128^3) picks 4 where the pass reads at least STREAM_PASS_BYTES of basis:
// its dot and update passes read up to 31 vectors, and w, which the stencil
// has just written and the update pass reads again, stays cached (GMRES(30)
// at 128^3: 294.6 -> 272.5 us per Arnoldi step; streaming every pass,
// short ones included, measured the same there).
constexpr int64_t STREAM_MIN_BYTES = (int64_t)512 << 20;
constexpr int64_t STREAM_PASS_BYTES = (int64_t)256 << 20;  // the last-level cache
template <int VW>
__device__ __forceinline__ float4 ld_vec(const float4 *p) {
    if constexpr (VW == 3) {
        const hz_f4 t = __builtin_nontemporal_load(reinterpret_cast<const hz_f4 *>(p));
        return make_float4(t.x, t.y, t.z, t.w);
    } else {
        return *p;
    }
}
template <int VW, typename T>
__device__ __forceinline__ cx<T> ld_one(const cx<T> *p) {
    if constexpr (VW == 4) {
        typedef T t2 __attribute__((ext_vector_type(2)));
        const t2 t = __builtin_nontemporal_load(reinterpret_cast<const t2 *>(p));
        return mk<T>(t.x, t.y);
    } else {
        return *p;
    }
}
template <int VW>
__device__ __forceinline__ void st_vec(float4 *p, const float4 &v) {
    if constexpr (VW == 3)
        __builtin_nontemporal_store(hz_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<hz_f4 *>(p));
    else
        *p = v;
}

// partial[((r*maxv + i)*nblk + blk)] = sum over the workgroup's chunk of
// conj(V_i[r]) * w[r], for i in the group [g*G, min(g*G+G, jp1)), G = dot_group<T>().
// VW 2 and 3 (complex64 only; n and chunk even): each lane loads two
// complex values per 16-byte access.
template <typename T, int VW>
__device__ __forceinline__ void dots_body(const cx<T> *__restrict__ V, int k, int64_t n, int jp1,
        const cx<T> *__restrict__ w, const int *__restrict__ live, double2 *__restrict__ partial, int maxv,
        int64_t chunk, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    if (!live[r]) return;
    constexpr int GQ = dot_group<T>();
    const int g0 = bk.z * GQ;
    const int cnt = min(GQ, jp1 - g0);
    if (cnt <= 0) return;
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    const cx<T> *wr = w + (int64_t)r * n;
    T are[GQ], aim[GQ];
#pragma unroll
    for (int q = 0; q < GQ; ++q) are[q] = aim[q] = 0;
    if constexpr (VW == 2 || VW == 3) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        const float4 *w4 = reinterpret_cast<const float4 *>(wr);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            const float4 wv = ld_vec<VW>(w4 + p);
#pragma unroll
            for (int q = 0; q < GQ; ++q) {
                if (q < cnt) {
                    const float4 v =
                            ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)(g0 + q) * k + r) * n) + p);
                    are[q] += v.x * wv.x + v.y * wv.y + v.z * wv.z + v.w * wv.w;
                    aim[q] += v.x * wv.y - v.y * wv.x + v.z * wv.w - v.w * wv.z;
                }
            }
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            const cx<T> wv = wr[p];
#pragma unroll
            for (int q = 0; q < GQ; ++q) {
                if (q < cnt) {
                    const cx<T> v = ld_one<VW>(V + ((int64_t)(g0 + q) * k + r) * n + p);
                    are[q] += v.re * wv.re + v.im * wv.im;
                    aim[q] += v.re * wv.im - v.im * wv.re;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < GQ; ++q) {
        if (q < cnt) {
            // FP32 (or FP64) wavefront sums, combined across wavefronts in FP64.
            const double sre = block_combine_256((double)wave_sum(are[q]), red);
            const double sim = block_combine_256((double)wave_sum(aim[q]), red);
            if (threadIdx.x == 0)
                partial[((int64_t)r * maxv + g0 + q) * bk.gx + bk.x] = make_double2(sre, sim);
        }
    }
}

template <typename T, int VW>
__global__ __launch_bounds__(256) void k_dots(const cx<T> *__restrict__ V, int k, int64_t n, int jp1,
        const cx<T> *__restrict__ w, const int *__restrict__ live, double2 *__restrict__ partial, int maxv,
        int64_t chunk) {
    dots_body<T, VW>(V, k, n, jp1, w, live, partial, maxv, chunk, blk_here());
}

template <typename T, int VW>
struct KDots {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *V;
        int k;
        int64_t n;
        int jp1;
        const cx<T> *w;
        const int *live;
        double2 *partial;
        int maxv;
        int64_t chunk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        dots_body<T, VW>(a.V, a.k, a.n, a.jp1, a.w, a.live, a.partial, a.maxv, a.chunk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_dots<T, VW>), g, dim3(THREADS), 0, s, a.V, a.k, a.n, a.jp1, a.w, a.live, a.partial,
                a.maxv, a.chunk);
    }
};

// Combine dot partials: d_i = sum_blk partial; H[r](i, j) = s_i d_i and the
// update coefficient coef[r][i] = s_i^2 d_i.  Grid (jp1, k).
__device__ __forceinline__ void reduce_dots_body(const double2 *__restrict__ partial, int nblk, int maxv, int j,
        const double *__restrict__ scales, const int *__restrict__ live, cx<double> *__restrict__ H,
        cx<double> *__restrict__ coef, int restart, const Blk &bk) {
    __shared__ double red[4];
    const int i = bk.x;
    const int r = bk.y;
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
        const double s = scales[(int64_t)i * bk.gy + r];
        const cx<double> d = mk<double>(re, im);
        H[((int64_t)r * (restart + 1) + i) * restart + j] = s * d;
        coef[r * maxv + i] = (s * s) * d;
    }
}

__global__ __launch_bounds__(256) void k_reduce_dots(const double2 *__restrict__ partial, int nblk, int maxv, int j,
        const double *__restrict__ scales, const int *__restrict__ live, cx<double> *__restrict__ H,
        cx<double> *__restrict__ coef, int restart) {
    reduce_dots_body(partial, nblk, maxv, j, scales, live, H, coef, restart, blk_here());
}

struct KReduceDots {
    static constexpr int THREADS = 256;
    struct Args {
        const double2 *partial;
        int nblk;
        int maxv;
        int j;
        const double *scales;
        const int *live;
        cx<double> *H;
        cx<double> *coef;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        reduce_dots_body(a.partial, a.nblk, a.maxv, a.j, a.scales, a.live, a.H, a.coef, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_reduce_dots), g, dim3(THREADS), 0, s, a.partial, a.nblk, a.maxv, a.j, a.scales, a.live,
                a.H, a.coef, a.restart);
    }
};

// target[r] -= sum_{i<cnt} coef[r][i] * V_i[r]; optional FP64 partial of
// |target|^2 per workgroup (norm_partial[r*nblk + blk]).  Grid (nblk, k).
// VW as in k_dots (paired: target leading dimension even as well).
template <typename T, int VW>
__device__ __forceinline__ void update_body(const cx<T> *__restrict__ V, int k, int64_t n, int cnt_all,
        const int *__restrict__ cnt_rhs, cx<T> *__restrict__ target, int64_t ldt, const cx<double> *__restrict__ coef,
        int maxv, const int *__restrict__ live, double *__restrict__ norm_partial, int64_t chunk, const Blk &bk) {
    __shared__ cx<T> cf[64];
    __shared__ double red[4];
    const int r = bk.y;
    if (!live[r]) return;
    const int cnt = cnt_rhs ? cnt_rhs[r] : cnt_all;
    for (int q = threadIdx.x; q < cnt; q += 256) cf[q] = cast<T>(coef[r * maxv + q]);
    __syncthreads();
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    cx<T> *tr = target + (int64_t)r * ldt;
    T nrm = 0;
    if constexpr (VW == 2 || VW == 3) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        float4 *t4 = reinterpret_cast<float4 *>(tr);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            float4 acc = ld_vec<VW>(t4 + p);
            int q = 0;
            for (; q + 8 <= cnt; q += 8) {
                float4 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    v[u] = ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)(q + u) * k + r) * n) + p);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const cx<float> c = cf[q + u];
                    acc.x = fmaf(c.im, v[u].y, fmaf(-c.re, v[u].x, acc.x));
                    acc.y = fmaf(-c.im, v[u].x, fmaf(-c.re, v[u].y, acc.y));
                    acc.z = fmaf(c.im, v[u].w, fmaf(-c.re, v[u].z, acc.z));
                    acc.w = fmaf(-c.im, v[u].z, fmaf(-c.re, v[u].w, acc.w));
                }
            }
            for (; q + 4 <= cnt; q += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    v[u] = ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)(q + u) * k + r) * n) + p);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const cx<float> c = cf[q + u];
                    acc.x = fmaf(c.im, v[u].y, fmaf(-c.re, v[u].x, acc.x));
                    acc.y = fmaf(-c.im, v[u].x, fmaf(-c.re, v[u].y, acc.y));
                    acc.z = fmaf(c.im, v[u].w, fmaf(-c.re, v[u].z, acc.z));
                    acc.w = fmaf(-c.im, v[u].z, fmaf(-c.re, v[u].w, acc.w));
                }
            }
            for (; q < cnt; ++q) {
                const float4 v = ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)q * k + r) * n) + p);
                const cx<float> c = cf[q];
                acc.x = fmaf(c.im, v.y, fmaf(-c.re, v.x, acc.x));
                acc.y = fmaf(-c.im, v.x, fmaf(-c.re, v.y, acc.y));
                acc.z = fmaf(c.im, v.w, fmaf(-c.re, v.z, acc.z));
                acc.w = fmaf(-c.im, v.z, fmaf(-c.re, v.w, acc.w));
            }
            st_vec<VW>(t4 + p, acc);
            nrm += acc.x * acc.x + acc.y * acc.y + acc.z * acc.z + acc.w * acc.w;
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            cx<T> acc = tr[p];
            int q = 0;
            for (; q + 4 <= cnt; q += 4) {
                const cx<T> v0 = ld_one<VW>(V + ((int64_t)(q + 0) * k + r) * n + p);
                const cx<T> v1 = ld_one<VW>(V + ((int64_t)(q + 1) * k + r) * n + p);
                const cx<T> v2 = ld_one<VW>(V + ((int64_t)(q + 2) * k + r) * n + p);
                const cx<T> v3 = ld_one<VW>(V + ((int64_t)(q + 3) * k + r) * n + p);
                acc = cfnma(cf[q], v0, acc);
                acc = cfnma(cf[q + 1], v1, acc);
                acc = cfnma(cf[q + 2], v2, acc);
                acc = cfnma(cf[q + 3], v3, acc);
            }
            for (; q < cnt; ++q) acc = cfnma(cf[q], ld_one<VW>(V + ((int64_t)q * k + r) * n + p), acc);
            tr[p] = acc;
            nrm += norm2(acc);
        }
    }
    if (norm_partial) {
        const double s = block_combine_256((double)wave_sum(nrm), red);
        if (threadIdx.x == 0) norm_partial[(int64_t)r * bk.gx + bk.x] = s;
    }
}

template <typename T, int VW>
__global__ __launch_bounds__(256) void k_update(const cx<T> *__restrict__ V, int k, int64_t n, int cnt_all,
        const int *__restrict__ cnt_rhs, cx<T> *__restrict__ target, int64_t ldt, const cx<double> *__restrict__ coef,
        int maxv, const int *__restrict__ live, double *__restrict__ norm_partial, int64_t chunk) {
    update_body<T, VW>(V, k, n, cnt_all, cnt_rhs, target, ldt, coef, maxv, live, norm_partial, chunk, blk_here());
}

template <typename T, int VW>
struct KUpdate {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *V;
        int k;
        int64_t n;
        int cnt_all;
        const int *cnt_rhs;
        cx<T> *target;
        int64_t ldt;
        const cx<double> *coef;
        int maxv;
        const int *live;
        double *norm_partial;
        int64_t chunk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        update_body<T, VW>(a.V, a.k, a.n, a.cnt_all, a.cnt_rhs, a.target, a.ldt, a.coef, a.maxv, a.live,
                a.norm_partial, a.chunk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_update<T, VW>), g, dim3(THREADS), 0, s, a.V, a.k, a.n, a.cnt_all, a.cnt_rhs, a.target,
                a.ldt, a.coef, a.maxv, a.live, a.norm_partial, a.chunk);
    }
};

// Per-RHS GMRES state, FP64, on the device.
struct GmresState {
    cx<double> *H;     // [k][restart+1][restart], Givens-rotated in place
    cx<double> *Hu;    // [k][restart+1][restart], unrotated (two-step blocks only)
    cx<double> *coef2; // second-target update coefficients [k][maxv] (two-step blocks)
    cx<double> *alpha; // per-RHS intra-block coefficients [2][k]: predicted, measured (two-step blocks)
    int *fix;          // [2][k]: the block of RHS r takes the measured coefficient; blocks that did so
    double guard;      // predict the intra-block coefficient only while |W1|^2 > guard |w1|^2
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
__device__ __forceinline__ void cycle_start_body(GmresState st, const double *__restrict__ partial, int nblk,
        int maxv, int restart, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.x;
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

__global__ __launch_bounds__(256) void k_cycle_start(GmresState st, const double *__restrict__ partial, int nblk,
        int maxv, int restart) {
    cycle_start_body(st, partial, nblk, maxv, restart, blk_here());
}

struct KCycleStart {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        const double *partial;
        int nblk;
        int maxv;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        cycle_start_body(a.st, a.partial, a.nblk, a.maxv, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_cycle_start), g, dim3(THREADS), 0, s, a.st, a.partial, a.nblk, a.maxv, a.restart);
    }
};

// partial[r * nblk + blk] = FP64 sum of |v[:, r]|^2 over the workgroup's
// chunk (nblk = gridDim.x; FP32 per thread over its few elements): the norm
// of the complex64 true residual at the cycle top, which the slab stencil
// writes without reducing it (a reduction in the stencil would follow the
// tile shape and round differently per shape).  Grid (nblk, k); VW as in
// k_dots.
template <int VW>
__device__ __forceinline__ void col_norms_body(const cx<float> *__restrict__ v, int64_t n, int64_t ld,
        const int *__restrict__ live, double *__restrict__ partial, int64_t chunk, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    if (live && !live[r]) return;
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    const cx<float> *vr = v + (int64_t)r * ld;
    float s = 0;
    if constexpr (VW >= 2) {
        const float4 *v4 = reinterpret_cast<const float4 *>(vr);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            const float4 a = ld_vec<VW>(v4 + p);
            s += a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w;
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) s += norm2(vr[p]);
    }
    const double t = block_sum_256((double)s, red);
    if (threadIdx.x == 0) partial[(int64_t)r * bk.gx + bk.x] = t;
}

template <int VW>
__global__ __launch_bounds__(256) void k_col_norms(const cx<float> *__restrict__ v, int64_t n, int64_t ld,
        const int *__restrict__ live, double *__restrict__ partial, int64_t chunk) {
    col_norms_body<VW>(v, n, ld, live, partial, chunk, blk_here());
}

template <int VW>
struct KColNorms {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<float> *v;
        int64_t n;
        int64_t ld;
        const int *live;
        double *partial;
        int64_t chunk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        col_norms_body<VW>(a.v, a.n, a.ld, a.live, a.partial, a.chunk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_col_norms<VW>), g, dim3(THREADS), 0, s, a.v, a.n, a.ld, a.live, a.partial, a.chunk);
    }
};

// Givens update of column j of H (rotated storage) whose subdiagonal entry
// is beta (real, >= 0): apply the previous rotations, make the new one,
// rotate g, record the residual estimate |g_{j+1}| and the step count, and
// retire the RHS when the estimate meets st.tol[r] or beta vanished.
// Returns whether the RHS is still live.  The column's entries 0..j are read
// from hj (stride hs) and the rotations 0..j-1 from gcr and gsr: H and the
// rotations in global memory, or a copy in LDS.  The entry that one rotation
// hands to the next stays in a register, so the chain of rotations waits on
// no memory round trip.
__device__ bool givens_column(GmresState &st, int r, int j, double beta, int restart, const cx<double> *hj,
                              int64_t hs, const double *gcr, const cx<double> *gsr) {
    cx<double> *col = st.H + (int64_t)r * (restart + 1) * restart;  // H(i, j) at col[i*restart + j]
    col[(j + 1) * restart + j] = mk<double>(beta, 0);
    cx<double> a = hj[0];
    for (int q = 0; q < j; ++q) {
        const double c = gcr[q];
        const cx<double> sn = gsr[q];
        const cx<double> b = hj[(q + 1) * hs];
        col[q * restart + j] = c * a + sn * b;
        a = c * b - conj(sn) * a;
    }
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
    if ((st.tol[r] >= 0 && est <= st.tol[r]) || beta == 0.0) {
        st.live[r] = 0;
        return false;
    }
    return true;
}

__device__ bool givens_column(GmresState &st, int r, int j, double beta, int restart) {
    const cx<double> *col = st.H + (int64_t)r * (restart + 1) * restart;
    return givens_column(st, r, j, beta, restart, col + j, restart, st.gc + r * restart, st.gs + r * restart);
}

// Finish Arnoldi step j: beta = ||w||, H(j+1, j) = beta, Givens update,
// residual estimate |g_{j+1}|, early exit when it meets st.tol[r].
__device__ __forceinline__ void arnoldi_finish_body(GmresState st, const double *__restrict__ norm_partial, int nblk,
        int j, int maxv, int restart, const Blk &bk) {
    __shared__ double red[4];
    __shared__ cx<double> s_h[64], s_s[64];  // restart <= 63: j <= 62
    __shared__ double s_c[64];
    const int r = bk.x;
    if (!st.live[r]) return;
    // Column j of H and the rotations, loaded by many lanes at once for the
    // one lane that runs the Givens chain (GMRES(30) at 128^3: 3 us less per
    // Arnoldi step than one lane loading them one rotation at a time).
    const cx<double> *col = st.H + (int64_t)r * (restart + 1) * restart;
    for (int q = threadIdx.x; q <= j; q += 256) {
        s_h[q] = col[q * restart + j];
        if (q < j) {
            s_c[q] = st.gc[r * restart + q];
            s_s[q] = st.gs[r * restart + q];
        }
    }
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += norm_partial[(int64_t)r * nblk + b];
    s = block_sum_256(s, red);
    __syncthreads();  // s_h, s_c, s_s
    if (threadIdx.x != 0) return;
    const double beta = sqrt(s);
    st.scales[(int64_t)(j + 1) * bk.gx + r] = beta > 0 ? 1.0 / beta : 0.0;
    givens_column(st, r, j, beta, restart, s_h, 1, s_c, s_s);
}

__global__ __launch_bounds__(256) void k_arnoldi_finish(GmresState st, const double *__restrict__ norm_partial,
        int nblk, int j, int maxv, int restart) {
    arnoldi_finish_body(st, norm_partial, nblk, j, maxv, restart, blk_here());
}

struct KArnoldiFinish {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        const double *norm_partial;
        int nblk;
        int j;
        int maxv;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        arnoldi_finish_body(a.st, a.norm_partial, a.nblk, a.j, a.maxv, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_arnoldi_finish), g, dim3(THREADS), 0, s, a.st, a.norm_partial, a.nblk, a.j, a.maxv,
                a.restart);
    }
};

// ------------------------------------------------------- two-step Arnoldi blocks
// A block at basis index j generates w1 = A M q_j (slot j+1) and
// w2 = A M w1 (slot j+2) before any orthogonalisation, so that one dot pass
// and one update pass over V_0..V_j serve both vectors:
//   W1 = w1 - sum_i h1_i q_i,          h1_i = q_i^H w1
//   W2 = w2 - sum_i h2_i q_i,          h2_i = q_i^H w2
//   q_{j+1} = W1 / r11,  r11 = |W1|
//   q_{j+2} = (W2 - alpha W1) / r22, alpha = W1^H W2 / |W1|^2, r12 = alpha r11
// alpha comes from the dot pass (|W1|^2 = |w1|^2 - |h1|^2, W1^H W2 =
// w1^H w2 - h1^H h2), so the update pass writes W2 - alpha W1 directly.
// Column j of H is (h1, r11); column j+1 follows from linearity:
//   A M q_{j+1} = (w2 - sum_{i<=j} h1_i A M q_i) / r11,
// with A M q_i = sum_l Hu(l, i) q_l the unrotated columns already built.

// Dot partials of both targets: partial[((t*k + r)*maxv + i)*nblk + blk];
// the first group of basis vectors also writes partials of |w1|^2 and
// w1^H w2 (wgram[(r*nblk + blk)*4 + 0..2]), from which the intra-block
// coefficient follows before the update pass.
template <typename T, int VW>
__device__ __forceinline__ void dots2_body(const cx<T> *__restrict__ V, int k, int64_t n, int jp1,
        const cx<T> *__restrict__ w1, const cx<T> *__restrict__ w2, const int *__restrict__ live,
        double2 *__restrict__ partial, int maxv, int64_t chunk, double *__restrict__ wgram, const Blk &bk) {
    constexpr int G = DOT_GROUP / 2;
    __shared__ double red[4];
    const int r = bk.y;
    if (!live[r]) return;
    const int g0 = bk.z * G;
    const int cnt = min(G, jp1 - g0);
    if (cnt <= 0) return;
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    const cx<T> *ar = w1 + (int64_t)r * n;
    const cx<T> *br = w2 + (int64_t)r * n;
    T are[G], aim[G], bre[G], bim[G];
#pragma unroll
    for (int q = 0; q < G; ++q) are[q] = aim[q] = bre[q] = bim[q] = 0;
    const bool first = bk.z == 0;
    // |w1|^2 and w1^H w2 feed a difference (|W1|^2 = |w1|^2 - |h1|^2) that
    // can cancel: they are summed in FP64 from the first product on.
    double waa = 0, wabr = 0, wabi = 0;
    if constexpr (VW >= 2) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        const float4 *a4 = reinterpret_cast<const float4 *>(ar);
        const float4 *b4 = reinterpret_cast<const float4 *>(br);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            const float4 av = ld_vec<VW>(a4 + p), bv = ld_vec<VW>(b4 + p);
            if (first) {
                const double ax = av.x, ay = av.y, az = av.z, aw = av.w;
                const double bx = bv.x, by = bv.y, bz = bv.z, bw = bv.w;
                waa += ax * ax + ay * ay + az * az + aw * aw;
                wabr += ax * bx + ay * by + az * bz + aw * bw;
                wabi += ax * by - ay * bx + az * bw - aw * bz;
            }
#pragma unroll
            for (int q = 0; q < G; ++q) {
                if (q < cnt) {
                    const float4 v =
                            ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)(g0 + q) * k + r) * n) + p);
                    are[q] += v.x * av.x + v.y * av.y + v.z * av.z + v.w * av.w;
                    aim[q] += v.x * av.y - v.y * av.x + v.z * av.w - v.w * av.z;
                    bre[q] += v.x * bv.x + v.y * bv.y + v.z * bv.z + v.w * bv.w;
                    bim[q] += v.x * bv.y - v.y * bv.x + v.z * bv.w - v.w * bv.z;
                }
            }
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            const cx<T> av = ar[p], bv = br[p];
            if (first) {
                const cx<double> a = cast<double>(av), b = cast<double>(bv);
                waa += norm2(a);
                wabr += a.re * b.re + a.im * b.im;
                wabi += a.re * b.im - a.im * b.re;
            }
#pragma unroll
            for (int q = 0; q < G; ++q) {
                if (q < cnt) {
                    const cx<T> v = V[((int64_t)(g0 + q) * k + r) * n + p];
                    are[q] += v.re * av.re + v.im * av.im;
                    aim[q] += v.re * av.im - v.im * av.re;
                    bre[q] += v.re * bv.re + v.im * bv.im;
                    bim[q] += v.re * bv.im - v.im * bv.re;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
        if (q < cnt) {
            const double s0 = block_combine_256((double)wave_sum(are[q]), red);
            const double s1 = block_combine_256((double)wave_sum(aim[q]), red);
            const double s2 = block_combine_256((double)wave_sum(bre[q]), red);
            const double s3 = block_combine_256((double)wave_sum(bim[q]), red);
            if (threadIdx.x == 0) {
                partial[(((int64_t)0 * k + r) * maxv + g0 + q) * bk.gx + bk.x] = make_double2(s0, s1);
                partial[(((int64_t)1 * k + r) * maxv + g0 + q) * bk.gx + bk.x] = make_double2(s2, s3);
            }
        }
    }
    if (first) {
        const double s0 = block_combine_256(wave_sum(waa), red);
        const double s1 = block_combine_256(wave_sum(wabr), red);
        const double s2 = block_combine_256(wave_sum(wabi), red);
        if (threadIdx.x == 0) {
            double *o = wgram + ((int64_t)r * bk.gx + bk.x) * 4;
            o[0] = s0;
            o[1] = s1;
            o[2] = s2;
        }
    }
}

template <typename T, int VW>
__global__ __launch_bounds__(256) void k_dots2(const cx<T> *__restrict__ V, int k, int64_t n, int jp1,
        const cx<T> *__restrict__ w1, const cx<T> *__restrict__ w2, const int *__restrict__ live,
        double2 *__restrict__ partial, int maxv, int64_t chunk, double *__restrict__ wgram) {
    dots2_body<T, VW>(V, k, n, jp1, w1, w2, live, partial, maxv, chunk, wgram, blk_here());
}

template <typename T, int VW>
struct KDots2 {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *V;
        int k;
        int64_t n;
        int jp1;
        const cx<T> *w1;
        const cx<T> *w2;
        const int *live;
        double2 *partial;
        int maxv;
        int64_t chunk;
        double *wgram;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        dots2_body<T, VW>(a.V, a.k, a.n, a.jp1, a.w1, a.w2, a.live, a.partial, a.maxv, a.chunk, a.wgram, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_dots2<T, VW>), g, dim3(THREADS), 0, s, a.V, a.k, a.n, a.jp1, a.w1, a.w2, a.live,
                a.partial, a.maxv, a.chunk, a.wgram);
    }
};

// The intra-block coefficient before the update pass (grid (k)): with
// W1 = w1 - sum h1_i q_i and W2 = w2 - sum h2_i q_i over orthonormal q_i,
// |W1|^2 = |w1|^2 - |h1|^2 and W1^H W2 = w1^H w2 - h1^H h2, so
// alpha = W1^H W2 / |W1|^2, and the update pass can write
// W2 - alpha W1 = (w2 - alpha w1) - sum (coef2_i - alpha coef_i) V_i
// directly (coef2 is rewritten here).  The norms that scale the two new
// vectors are measured in that pass, not predicted.
constexpr double BLOCK_GUARD = 1.0e-2;  // default of GmresState::guard

__device__ __forceinline__ void block_alpha_body(GmresState st, const double *__restrict__ wgram, int nblk, int j,
        int maxv, int restart, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.x;
    if (!st.live[r]) return;
    double v[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double *o = wgram + ((int64_t)r * nblk + b) * 4;
        v[0] += o[0];
        v[1] += o[1];
        v[2] += o[2];
    }
    for (int q = 0; q < 3; ++q) v[q] = block_sum_256(v[q], red);
    if (threadIdx.x != 0) return;
    const cx<double> *hu = st.Hu + (int64_t)r * (restart + 1) * restart;
    double w11 = v[0];
    cx<double> w12 = mk<double>(v[1], v[2]);
    for (int i = 0; i <= j; ++i) {
        const cx<double> h1 = hu[i * restart + j], h2 = hu[i * restart + j + 1];
        w11 -= norm2(h1);
        w12 = w12 - conj(h1) * h2;
    }
    // When W1 keeps less than st.guard of |w1|^2 the prediction cancels
    // (the dot products come from complex64 vectors): no intra-block
    // coefficient is predicted, and the fix-up pass orthogonalises W2
    // against W1 with their measured inner product instead.
    const bool guard = !(w11 > st.guard * v[0]);
    const cx<double> al = guard ? mk<double>(0, 0) : (1.0 / w11) * w12;
    st.alpha[r] = al;
    st.fix[r] = guard ? 1 : 0;
    if (guard) st.fix[bk.gx + r] += 1;
    for (int i = 0; i <= j; ++i) st.coef2[r * maxv + i] = st.coef2[r * maxv + i] - al * st.coef[r * maxv + i];
}

__global__ __launch_bounds__(256) void k_block_alpha(GmresState st, const double *__restrict__ wgram, int nblk, int j,
        int maxv, int restart) {
    block_alpha_body(st, wgram, nblk, j, maxv, restart, blk_here());
}

struct KBlockAlpha {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        const double *wgram;
        int nblk;
        int j;
        int maxv;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        block_alpha_body(a.st, a.wgram, a.nblk, a.j, a.maxv, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_block_alpha), g, dim3(THREADS), 0, s, a.st, a.wgram, a.nblk, a.j, a.maxv, a.restart);
    }
};

// Combine both targets' partials (grid (jp1, k)): Hu(i, j) = h1_i = s_i d1_i,
// Hu(i, j+1) = h2_i (held there until the block finishes), and the update
// coefficients coef = s_i^2 d1_i, coef2 = s_i^2 d2_i.
__device__ __forceinline__ void reduce_dots2_body(const double2 *__restrict__ partial, int nblk, int maxv, int j,
        GmresState st, int restart, const Blk &bk) {
    __shared__ double red[4];
    const int i = bk.x;
    const int r = bk.y;
    const int k = bk.gy;
    if (!st.live[r]) return;
    double v[4] = {0, 0, 0, 0};
    for (int t = 0; t < 2; ++t) {
        const double2 *src = partial + (((int64_t)t * k + r) * maxv + i) * nblk;
        for (int b = threadIdx.x; b < nblk; b += 256) {
            v[2 * t] += src[b].x;
            v[2 * t + 1] += src[b].y;
        }
    }
    for (int q = 0; q < 4; ++q) v[q] = block_sum_256(v[q], red);
    if (threadIdx.x == 0) {
        const double s = st.scales[(int64_t)i * k + r];
        const cx<double> d1 = mk<double>(v[0], v[1]), d2 = mk<double>(v[2], v[3]);
        cx<double> *hu = st.Hu + (int64_t)r * (restart + 1) * restart;
        hu[i * restart + j] = s * d1;
        hu[i * restart + j + 1] = s * d2;
        st.coef[r * maxv + i] = (s * s) * d1;
        st.coef2[r * maxv + i] = (s * s) * d2;
    }
}

__global__ __launch_bounds__(256) void k_reduce_dots2(const double2 *__restrict__ partial, int nblk, int maxv, int j,
        GmresState st, int restart) {
    reduce_dots2_body(partial, nblk, maxv, j, st, restart, blk_here());
}

struct KReduceDots2 {
    static constexpr int THREADS = 256;
    struct Args {
        const double2 *partial;
        int nblk;
        int maxv;
        int j;
        GmresState st;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        reduce_dots2_body(a.partial, a.nblk, a.maxv, a.j, a.st, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_reduce_dots2), g, dim3(THREADS), 0, s, a.partial, a.nblk, a.maxv, a.j, a.st, a.restart);
    }
};

// W1 = w1 - sum coef V_i, W2 = w2 - alpha w1 - sum coef2 V_i in place (coef2
// already carries -alpha coef), with FP64 partials of |W1|^2, Re/Im W1^H W2
// and |W2|^2 (gram[(r*nblk + blk)*4 + 0..3]).
template <typename T, int VW>
__device__ __forceinline__ void update2_body(const cx<T> *__restrict__ V, int k, int64_t n, int cnt,
        cx<T> *__restrict__ w1, cx<T> *__restrict__ w2, const cx<double> *__restrict__ coef,
        const cx<double> *__restrict__ coef2, int maxv, const int *__restrict__ live, double *__restrict__ gram,
        int64_t chunk, const cx<double> *__restrict__ alpha, const Blk &bk) {
    __shared__ cx<T> cf[64], cf2[64];
    __shared__ double red[4];
    const int r = bk.y;
    if (!live[r]) return;
    for (int q = threadIdx.x; q < cnt; q += 256) {
        cf[q] = cast<T>(coef[r * maxv + q]);
        cf2[q] = cast<T>(coef2[r * maxv + q]);
    }
    __syncthreads();
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    cx<T> *ar = w1 + (int64_t)r * n;
    cx<T> *br = w2 + (int64_t)r * n;
    const cx<T> al = cast<T>(alpha[r]);
    T g11 = 0, g22 = 0, g12r = 0, g12i = 0;
    if constexpr (VW >= 2) {
        static_assert(sizeof(T) == 4, "paired loads are for complex64");
        float4 *a4 = reinterpret_cast<float4 *>(ar);
        float4 *b4 = reinterpret_cast<float4 *>(br);
        for (int64_t p = (p0 >> 1) + threadIdx.x; p < (p1 >> 1); p += 256) {
            float4 a = ld_vec<VW>(a4 + p), b = ld_vec<VW>(b4 + p);
            b.x = fmaf(al.im, a.y, fmaf(-al.re, a.x, b.x));
            b.y = fmaf(-al.im, a.x, fmaf(-al.re, a.y, b.y));
            b.z = fmaf(al.im, a.w, fmaf(-al.re, a.z, b.z));
            b.w = fmaf(-al.im, a.z, fmaf(-al.re, a.w, b.w));
            int q = 0;
            for (; q + 4 <= cnt; q += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    v[u] = ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)(q + u) * k + r) * n) + p);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const cx<float> c = cf[q + u], d = cf2[q + u];
                    a.x = fmaf(c.im, v[u].y, fmaf(-c.re, v[u].x, a.x));
                    a.y = fmaf(-c.im, v[u].x, fmaf(-c.re, v[u].y, a.y));
                    a.z = fmaf(c.im, v[u].w, fmaf(-c.re, v[u].z, a.z));
                    a.w = fmaf(-c.im, v[u].z, fmaf(-c.re, v[u].w, a.w));
                    b.x = fmaf(d.im, v[u].y, fmaf(-d.re, v[u].x, b.x));
                    b.y = fmaf(-d.im, v[u].x, fmaf(-d.re, v[u].y, b.y));
                    b.z = fmaf(d.im, v[u].w, fmaf(-d.re, v[u].z, b.z));
                    b.w = fmaf(-d.im, v[u].z, fmaf(-d.re, v[u].w, b.w));
                }
            }
            for (; q < cnt; ++q) {
                const float4 v = ld_vec<VW>(reinterpret_cast<const float4 *>(V + ((int64_t)q * k + r) * n) + p);
                const cx<float> c = cf[q], d = cf2[q];
                a.x = fmaf(c.im, v.y, fmaf(-c.re, v.x, a.x));
                a.y = fmaf(-c.im, v.x, fmaf(-c.re, v.y, a.y));
                a.z = fmaf(c.im, v.w, fmaf(-c.re, v.z, a.z));
                a.w = fmaf(-c.im, v.z, fmaf(-c.re, v.w, a.w));
                b.x = fmaf(d.im, v.y, fmaf(-d.re, v.x, b.x));
                b.y = fmaf(-d.im, v.x, fmaf(-d.re, v.y, b.y));
                b.z = fmaf(d.im, v.w, fmaf(-d.re, v.z, b.z));
                b.w = fmaf(-d.im, v.z, fmaf(-d.re, v.w, b.w));
            }
            st_vec<VW>(a4 + p, a);
            st_vec<VW>(b4 + p, b);
            g11 += a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w;
            g22 += b.x * b.x + b.y * b.y + b.z * b.z + b.w * b.w;
            g12r += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
            g12i += a.x * b.y - a.y * b.x + a.z * b.w - a.w * b.z;
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
            cx<T> a = ar[p], b = br[p];
            b = b - al * a;
            for (int q = 0; q < cnt; ++q) {
                const cx<T> v = V[((int64_t)q * k + r) * n + p];
                a = a - cf[q] * v;
                b = b - cf2[q] * v;
            }
            ar[p] = a;
            br[p] = b;
            g11 += norm2(a);
            g22 += norm2(b);
            g12r += a.re * b.re + a.im * b.im;
            g12i += a.re * b.im - a.im * b.re;
        }
    }
    const double s0 = block_combine_256((double)wave_sum(g11), red);
    const double s1 = block_combine_256((double)wave_sum(g12r), red);
    const double s2 = block_combine_256((double)wave_sum(g12i), red);
    const double s3 = block_combine_256((double)wave_sum(g22), red);
    if (threadIdx.x == 0) {
        double *o = gram + ((int64_t)r * bk.gx + bk.x) * 4;
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
        o[3] = s3;
    }
}

template <typename T, int VW>
__global__ __launch_bounds__(256) void k_update2(const cx<T> *__restrict__ V, int k, int64_t n, int cnt,
        cx<T> *__restrict__ w1, cx<T> *__restrict__ w2, const cx<double> *__restrict__ coef,
        const cx<double> *__restrict__ coef2, int maxv, const int *__restrict__ live, double *__restrict__ gram,
        int64_t chunk, const cx<double> *__restrict__ alpha) {
    update2_body<T, VW>(V, k, n, cnt, w1, w2, coef, coef2, maxv, live, gram, chunk, alpha, blk_here());
}

template <typename T, int VW>
struct KUpdate2 {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *V;
        int k;
        int64_t n;
        int cnt;
        cx<T> *w1;
        cx<T> *w2;
        const cx<double> *coef;
        const cx<double> *coef2;
        int maxv;
        const int *live;
        double *gram;
        int64_t chunk;
        const cx<double> *alpha;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        update2_body<T, VW>(a.V, a.k, a.n, a.cnt, a.w1, a.w2, a.coef, a.coef2, a.maxv, a.live, a.gram, a.chunk,
                a.alpha, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_update2<T, VW>), g, dim3(THREADS), 0, s, a.V, a.k, a.n, a.cnt, a.w1, a.w2, a.coef,
                a.coef2, a.maxv, a.live, a.gram, a.chunk, a.alpha);
    }
};

// Blocks whose coefficient was not predicted (k_block_alpha's guard):
// alpha'[r] = W1^H W2 / |W1|^2 from the update pass's Gram partials (grid (k)).
__device__ __forceinline__ void block_fixcoef_body(GmresState st, const double *__restrict__ gram, int nblk,
        const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.x;
    const int k = bk.gx;
    if (!st.live[r] || !st.fix[r]) return;
    double v[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double *o = gram + ((int64_t)r * nblk + b) * 4;
        v[0] += o[0];
        v[1] += o[1];
        v[2] += o[2];
    }
    for (int q = 0; q < 3; ++q) v[q] = block_sum_256(v[q], red);
    if (threadIdx.x == 0) st.alpha[k + r] = v[0] > 0 ? (1.0 / v[0]) * mk<double>(v[1], v[2]) : mk<double>(0, 0);
}

__global__ __launch_bounds__(256) void k_block_fixcoef(GmresState st, const double *__restrict__ gram, int nblk) {
    block_fixcoef_body(st, gram, nblk, blk_here());
}

struct KBlockFixcoef {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        const double *gram;
        int nblk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        block_fixcoef_body(a.st, a.gram, a.nblk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_block_fixcoef), g, dim3(THREADS), 0, s, a.st, a.gram, a.nblk);
    }
};

// W2 -= alpha' W1 in place for those blocks, and |W2|^2 measured again into
// gram[(r*nblk + blk)*4 + 3] (grid (nblk, k)).
template <typename T>
__device__ __forceinline__ void block_fixup_body(GmresState st, int64_t n, const cx<T> *__restrict__ w1,
        cx<T> *__restrict__ w2, double *__restrict__ gram, int64_t chunk, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    const int k = bk.gy;
    if (!st.live[r] || !st.fix[r]) return;
    const cx<T> al = cast<T>(st.alpha[k + r]);
    const cx<T> *a = w1 + (int64_t)r * n;
    cx<T> *b = w2 + (int64_t)r * n;
    const int64_t p0 = bk.x * chunk;
    const int64_t p1 = min(n, p0 + chunk);
    T g22 = 0;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
        const cx<T> v = b[p] - al * a[p];
        b[p] = v;
        g22 += norm2(v);
    }
    const double s = block_combine_256(wave_sum((double)g22), red);
    if (threadIdx.x == 0) gram[((int64_t)r * bk.gx + bk.x) * 4 + 3] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_block_fixup(GmresState st, int64_t n, const cx<T> *__restrict__ w1,
        cx<T> *__restrict__ w2, double *__restrict__ gram, int64_t chunk) {
    block_fixup_body<T>(st, n, w1, w2, gram, chunk, blk_here());
}

template <typename T>
struct KBlockFixup {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        int64_t n;
        const cx<T> *w1;
        cx<T> *w2;
        double *gram;
        int64_t chunk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        block_fixup_body<T>(a.st, a.n, a.w1, a.w2, a.gram, a.chunk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_block_fixup<T>), g, dim3(THREADS), 0, s, a.st, a.n, a.w1, a.w2, a.gram, a.chunk);
    }
};

// Finish the block (grid (k)) from the update pass's Gram partials:
// r11 = |W1| and r22 = |W2 - alpha W1| as measured (alpha the predicted
// coefficient plus, after a fix-up, the measured one), r12 = alpha r11 (w2 =
// sum h2_i q_i + alpha r11 q_{j+1} + r22 q_{j+2} exactly, by construction),
// the scales of slots j+1 and j+2, column j+1 of Hu by linearity, then the
// Givens updates of columns j and j+1 (the second only if the RHS is still
// live after the first).
__device__ __forceinline__ void block_finish_body(GmresState st, const double *__restrict__ gram, int nblk, int j,
        int restart, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.x;
    if (!st.live[r]) return;
    double v[2] = {0, 0};
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double *o = gram + ((int64_t)r * nblk + b) * 4;
        v[0] += o[0];
        v[1] += o[3];
    }
    for (int q = 0; q < 2; ++q) v[q] = block_sum_256(v[q], red);
    if (threadIdx.x != 0) return;
    const double r11 = sqrt(v[0]), r22 = sqrt(v[1]);
    const cx<double> al = st.fix[r] ? st.alpha[r] + st.alpha[bk.gx + r] : st.alpha[r];
    const cx<double> r12 = r11 * al;
    st.scales[(int64_t)(j + 1) * bk.gx + r] = r11 > 0 ? 1.0 / r11 : 0.0;
    st.scales[(int64_t)(j + 2) * bk.gx + r] = r22 > 0 ? 1.0 / r22 : 0.0;
    cx<double> *hu = st.Hu + (int64_t)r * (restart + 1) * restart;
    cx<double> *h = st.H + (int64_t)r * (restart + 1) * restart;
    hu[(j + 1) * restart + j] = mk<double>(r11, 0);
    // Coefficients of w2 in q_0 .. q_{j+2}: (h2_0 .. h2_j, r12, r22); subtract
    // h1_i * Hu(:, i) for i <= j (column j itself is (h1, r11)).
    cx<double> col[66];
    for (int l = 0; l <= j; ++l) col[l] = hu[l * restart + j + 1];
    col[j + 1] = r12;
    col[j + 2] = mk<double>(r22, 0);
    for (int i = 0; i <= j; ++i) {
        const cx<double> h1 = hu[i * restart + j];
        for (int l = 0; l <= i + 1; ++l) col[l] = col[l] - h1 * hu[l * restart + i];
    }
    const double inv = r11 > 0 ? 1.0 / r11 : 0.0;
    for (int l = 0; l <= j + 2; ++l) hu[l * restart + j + 1] = inv * col[l];
    // Rotated storage: copy the unrotated columns, then the Givens updates.
    for (int l = 0; l <= j; ++l) h[l * restart + j] = hu[l * restart + j];
    if (!givens_column(st, r, j, r11, restart)) return;
    for (int l = 0; l <= j + 1; ++l) h[l * restart + j + 1] = hu[l * restart + j + 1];
    givens_column(st, r, j + 1, hu[(j + 2) * restart + j + 1].re, restart);
}

__global__ __launch_bounds__(256) void k_block_finish(GmresState st, const double *__restrict__ gram, int nblk, int j,
        int restart) {
    block_finish_body(st, gram, nblk, j, restart, blk_here());
}

struct KBlockFinish {
    static constexpr int THREADS = 256;
    struct Args {
        GmresState st;
        const double *gram;
        int nblk;
        int j;
        int restart;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        block_finish_body(a.st, a.gram, a.nblk, a.j, a.restart, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_block_finish), g, dim3(THREADS), 0, s, a.st, a.gram, a.nblk, a.j, a.restart);
    }
};

// Back substitution on the rotated Hessenberg matrix; coef[r][i] = -y_i s_i so
// that k_update adds sum_i y_i v_i to the iterate.
// blocks_z: the update runs over the stored preconditioned slots Z_i instead
// (the multigrid form with two-step blocks).  Z_i = M V_i for a block's first
// slot i, M q_i = s_i Z_i; its second slot holds Z_{i} = M w1 of the raw
// block vector w1 = A M q_{i-1}, so that
//   M q_i = (Z_i - sum_{l < i} Hu(l, i-1) M q_l) / Hu(i, i-1)
// (the linearity the block already uses for column i of H).  The
// coefficients of sum_i y_i M q_i on the Z_i follow from the highest slot
// down.
__device__ __forceinline__ void trisolve_body(GmresState st, int maxv, int restart, const int *__restrict__ member,
        int blocks_z, const Blk &bk) {
    const int r = bk.x;
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
    if (!blocks_z) {
        for (int i = 0; i < m; ++i) st.coef[r * maxv + i] = (-st.scales[(int64_t)i * bk.gx + r]) * y[i];
        return;
    }
    const cx<double> *hu = st.Hu + (int64_t)r * (restart + 1) * restart;
    for (int i = m - 1; i >= 0; --i) {
        if (i & 1) {
            const double r11 = hu[i * restart + i - 1].re;
            const cx<double> c = r11 > 0 ? (1.0 / r11) * y[i] : mk<double>(0, 0);
            for (int l = 0; l < i; ++l) y[l] = y[l] - c * hu[l * restart + i - 1];
            st.coef[r * maxv + i] = -1.0 * c;
        } else {
            st.coef[r * maxv + i] = (-st.scales[(int64_t)i * bk.gx + r]) * y[i];
        }
    }
}

__global__ __launch_bounds__(64) void k_trisolve(GmresState st, int maxv, int restart, const int *__restrict__ member,
        int blocks_z) {
    trisolve_body(st, maxv, restart, member, blocks_z, blk_here());
}

struct KTrisolve {
    static constexpr int THREADS = 64;
    struct Args {
        GmresState st;
        int maxv;
        int restart;
        const int *member;
        int blocks_z;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        trisolve_body(a.st, a.maxv, a.restart, a.member, a.blocks_z, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_trisolve), g, dim3(THREADS), 0, s, a.st, a.maxv, a.restart, a.member, a.blocks_z);
    }
};

// ---------------------------------------------------------------- elementwise
// out[r] = cast(rs * b[r]) (scaled right-hand side) with FP64 partials of
// |rs b|^2 and |b|^2 (partial[(r*nblk+blk)*2 + 0/1]).
template <typename T>
__device__ __forceinline__ void scale_rhs_body(const cx<double> *__restrict__ b, int64_t ldb,
        const double *__restrict__ rs, cx<T> *__restrict__ out, int64_t ldo, int64_t n, double *__restrict__ partial,
        int64_t chunk, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    const int64_t p0 = bk.x * chunk;
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
        partial[((int64_t)r * bk.gx + bk.x) * 2 + 0] = s1;
        partial[((int64_t)r * bk.gx + bk.x) * 2 + 1] = s2;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale_rhs(const cx<double> *__restrict__ b, int64_t ldb,
        const double *__restrict__ rs, cx<T> *__restrict__ out, int64_t ldo, int64_t n, double *__restrict__ partial,
        int64_t chunk) {
    scale_rhs_body<T>(b, ldb, rs, out, ldo, n, partial, chunk, blk_here());
}

template <typename T>
struct KScaleRhs {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<double> *b;
        int64_t ldb;
        const double *rs;
        cx<T> *out;
        int64_t ldo;
        int64_t n;
        double *partial;
        int64_t chunk;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        scale_rhs_body<T>(a.b, a.ldb, a.rs, a.out, a.ldo, a.n, a.partial, a.chunk, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_scale_rhs<T>), g, dim3(THREADS), 0, s, a.b, a.ldb, a.rs, a.out, a.ldo, a.n, a.partial,
                a.chunk);
    }
};

// x[r] (+)= m * cast<double>(y[r]): apply the Jacobi factor to the iterate.
template <typename T>
__device__ __forceinline__ void precond_out_body(const cx<T> *__restrict__ y, int64_t ldy,
        const cx<double> *__restrict__ m, cx<double> *__restrict__ x, int64_t ldx, int64_t n,
        const int *__restrict__ member, int accumulate, const Blk &bk) {
    const int r = bk.y;
    if (member && !member[r]) return;
    for (int64_t p = bk.x * 256 + threadIdx.x; p < n; p += (int64_t)bk.gx * 256) {
        const cx<double> v = m[p] * cast<double>(y[r * ldy + p]);
        x[r * ldx + p] = accumulate ? x[r * ldx + p] + v : v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_precond_out(const cx<T> *__restrict__ y, int64_t ldy,
        const cx<double> *__restrict__ m, cx<double> *__restrict__ x, int64_t ldx, int64_t n,
        const int *__restrict__ member, int accumulate) {
    precond_out_body<T>(y, ldy, m, x, ldx, n, member, accumulate, blk_here());
}

template <typename T>
struct KPrecondOut {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *y;
        int64_t ldy;
        const cx<double> *m;
        cx<double> *x;
        int64_t ldx;
        int64_t n;
        const int *member;
        int accumulate;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        precond_out_body<T>(a.y, a.ldy, a.m, a.x, a.ldx, a.n, a.member, a.accumulate, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_precond_out<T>), g, dim3(THREADS), 0, s, a.y, a.ldy, a.m, a.x, a.ldx, a.n, a.member,
                a.accumulate);
    }
};

template <typename T>
__device__ __forceinline__ void zero_body(cx<T> *__restrict__ x, int64_t ldx, int64_t n, const Blk &bk) {
    const int r = bk.y;
    for (int64_t p = bk.x * 256 + threadIdx.x; p < n; p += (int64_t)bk.gx * 256)
        x[r * ldx + p] = mk<T>(0, 0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(cx<T> *__restrict__ x, int64_t ldx, int64_t n) {
    zero_body<T>(x, ldx, n, blk_here());
}

template <typename T>
struct KZero {
    static constexpr int THREADS = 256;
    struct Args {
        cx<T> *x;
        int64_t ldx;
        int64_t n;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        zero_body<T>(a.x, a.ldx, a.n, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_zero<T>), g, dim3(THREADS), 0, s, a.x, a.ldx, a.n);
    }
};

template <typename T>
__global__ __launch_bounds__(256) void k_fill(cx<T> *__restrict__ x, int64_t ldx, int64_t n, cx<T> value) {
    const int r = blockIdx.y;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) x[r * ldx + p] = value;
}

// Sum (or max, where bit q of max_mask is set) of nblk partials for each of
// `width` values per RHS: out[r*width + q] = reduce_b partial[(r*nblk + b)*width + q].
// Grid (width, k).
__device__ __forceinline__ void reduce_rows_body(const double *__restrict__ partial, int nblk, int width,
        int max_mask, double *__restrict__ out, const Blk &bk) {
    __shared__ double red[4];
    const int q = bk.x;
    const int r = bk.y;
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

__global__ __launch_bounds__(256) void k_reduce_rows(const double *__restrict__ partial, int nblk, int width,
        int max_mask, double *__restrict__ out) {
    reduce_rows_body(partial, nblk, width, max_mask, out, blk_here());
}

struct KReduceRows {
    static constexpr int THREADS = 256;
    struct Args {
        const double *partial;
        int nblk;
        int width;
        int max_mask;
        double *out;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        reduce_rows_body(a.partial, a.nblk, a.width, a.max_mask, a.out, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_reduce_rows), g, dim3(THREADS), 0, s, a.partial, a.nblk, a.width, a.max_mask, a.out);
    }
};

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
