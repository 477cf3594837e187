// Complex-shifted-Laplacian multigrid preconditioner kernels.
//
// Each level is the same finite-volume operator rediscretised on a coarser
// node set with the mass term shifted, omega^2 -> omega^2 (1 + i beta)
// (physical damping for the exp(-i omega t) convention used throughout).
// Coarse node sets keep every even fine index plus the last node, so any
// grid size coarsens (the last coarse interval may be one fine cell long);
// the per-axis dual widths and edge lengths carry the non-uniform spacing.
//
// Transfers are trilinear prolongation P and its transpose R = P^T (rows are
// integrated over control volumes with the fine-grid normalisation on every
// level, so plain P^T is the consistent restriction).  The smoother is
// damped Jacobi on the shifted operator.  Every kernel handles k
// right-hand sides, evaluating the stencil coefficients once per node.
#pragma once

#include "hz_stencil.hpp"

namespace hz {

// Node on a pressure-release face (identity row) for face bitmask `soft`.
__device__ __forceinline__ bool on_soft_face(int i, int j, int k, int nx, int ny, int nz, int soft) {
    return ((soft & 1) && i == 0) || ((soft & 2) && i == nx - 1) || ((soft & 4) && j == 0) ||
           ((soft & 8) && j == ny - 1) || ((soft & 16) && k == 0) || ((soft & 32) && k == nz - 1);
}

// Fine index f is a coarse node iff it is even or the last node.
__device__ __forceinline__ bool coarse_node(int f, int n) { return (f & 1) == 0 || f == n - 1; }
// Coarse index of a coarse-node fine index f (m = coarse size).
__device__ __forceinline__ int coarse_of(int f, int n, int m) { return f == n - 1 ? m - 1 : f >> 1; }
// Fine index of coarse index c.
__device__ __forceinline__ int fine_of(int c, int n) { return min(2 * c, n - 1); }

// Inverse diagonal of the shifted level operator (identity rows -> 1); for the
// adjoint smoother the conjugate is used.
template <typename T>
__global__ __launch_bounds__(256) void k_level_dinv(OpView<T> op, cx<T> *__restrict__ dinv) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= op.n) return;
    const int i = (int)(p % op.nx);
    const int j = (int)((p / op.nx) % op.ny);
    const int k = (int)(p / ((int64_t)op.nx * op.ny));
    const Nbr nb = neighbours(i, j, k, op.nx, op.ny, op.nz);
    cx<T> c[6], d;
    edge_coefs<T, true>(op, i, j, k, p, nb, c, &d);
    dinv[p] = node_fixed(op, i, j, k) ? mk<T>(1, 0) : crecip(d);
}

enum LevelEpi { EPI_SMOOTH = 0, EPI_RESID = 1 };

// EPI_SMOOTH: xo = xi + w * dinv * (b - A xi)      (damped Jacobi sweep)
// EPI_RESID : xo = b - A xi
// MODE is MODE_A (forward) or MODE_AH (adjoint; dinv conjugated).  When
// `bdiv` is set the right-hand side is read as b / bdiv: on the finest level
// this undoes the row scale of the Krylov system, so that R A (P R^-1) is a
// similarity transform of A P and its spectrum does not inherit the spread
// of R (identity rows have scale 1, the others ~1/|A_PP|).
template <typename T, int MODE, int EPI>
__global__ __launch_bounds__(256) void k_level_op(OpView<T> op, TileMap tm, const cx<T> *__restrict__ xi,
                                                  const cx<T> *__restrict__ b, cx<T> *__restrict__ xo,
                                                  const cx<T> *__restrict__ dinv, T w, int k,
                                                  const int *__restrict__ live, const T *__restrict__ bdiv) {
    __shared__ int s_act[APPLY_MAX_RHS];
    __shared__ int s_na;
    if (threadIdx.x == 0) {
        int na = 0;
        for (int r = 0; r < k; ++r)
            if (!live || live[r]) s_act[na++] = r;
        s_na = na;
    }
    __syncthreads();
    const int tile = xcd_tile(blockIdx.x, tm.nblocks);
    if (tile >= tm.ntiles) return;
    const int tx = tile % tm.tiles_x;
    const int ty = (tile / tm.tiles_x) % tm.tiles_y;
    const int tz = tile / (tm.tiles_x * tm.tiles_y);
    const int i = tx * 64 + (threadIdx.x & 63);
    const int j = ty * 4 + (threadIdx.x >> 6);
    if (i >= op.nx || j >= op.ny) return;
    const int64_t n = op.n;
    const int na = s_na;
    const int k0 = tz * tm.zb;
    const int k1 = min(op.nz, k0 + tm.zb);
    for (int kk = k0; kk < k1; ++kk) {
        const int64_t p = i + (int64_t)op.nx * (j + (int64_t)op.ny * kk);
        const Nbr nb = neighbours(i, j, kk, op.nx, op.ny, op.nz);
        cx<T> cs, c[6];
        row_coefs<T, MODE>(op, i, j, kk, p, nb, cs, c);
        cx<T> dv = mk<T>(0, 0);
        if (EPI == EPI_SMOOTH) dv = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
        const T bs = bdiv ? T(1) / bdiv[p] : T(1);
        int t = 0;
        for (; t + 2 <= na; t += 2) {
            const int r0 = s_act[t], r1 = s_act[t + 1];
            const cx<T> *x0 = xi + r0 * n + p;
            const cx<T> *x1 = xi + r1 * n + p;
            cx<T> u0[7], u1[7];
            u0[6] = x0[0];
            u1[6] = x1[0];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                u0[q] = x0[nb.off[q]];
                u1[q] = x1[nb.off[q]];
            }
            const cx<T> b0 = b[r0 * n + p], b1 = b[r1 * n + p];
            cx<T> a0 = cs * u0[6], a1 = cs * u1[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                a0 = a0 + c[q] * u0[q];
                a1 = a1 + c[q] * u1[q];
            }
            const cx<T> e0 = bs * b0 - a0, e1 = bs * b1 - a1;
            xo[r0 * n + p] = (EPI == EPI_SMOOTH) ? u0[6] + w * (dv * e0) : e0;
            xo[r1 * n + p] = (EPI == EPI_SMOOTH) ? u1[6] + w * (dv * e1) : e1;
        }
        if (t < na) {
            const int r0 = s_act[t];
            const cx<T> *x0 = xi + r0 * n + p;
            const cx<T> u06 = x0[0];
            cx<T> a0 = cs * u06;
#pragma unroll
            for (int q = 0; q < 6; ++q) a0 = a0 + c[q] * x0[nb.off[q]];
            const cx<T> e0 = bs * b[r0 * n + p] - a0;
            xo[r0 * n + p] = (EPI == EPI_SMOOTH) ? u06 + w * (dv * e0) : e0;
        }
    }
}

// First sweep from a zero guess: xo = w * dinv * b.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_level_jacobi0(const cx<T> *__restrict__ b, cx<T> *__restrict__ xo,
                                                       const cx<T> *__restrict__ dinv, T w, int64_t n,
                                                       const int *__restrict__ live, const T *__restrict__ bdiv) {
    const int r = blockIdx.y;
    if (live && !live[r]) return;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
        const cx<T> dv = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
        const T ws = bdiv ? w / bdiv[p] : w;
        xo[r * n + p] = ws * (dv * b[r * n + p]);
    }
}

constexpr int COARSE_MAX_NODES = 1024;

// All damped-Jacobi sweeps of the coarsest level in one launch: one
// workgroup per right-hand side keeps the iterate in LDS (ping-pong) and
// each thread holds the coefficients of its <= 4 nodes in registers.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_coarse_jacobi(OpView<T> op, const cx<T> *__restrict__ b,
                                                       cx<T> *__restrict__ x, const cx<T> *__restrict__ dinv, T w,
                                                       int sweeps, const int *__restrict__ live) {
    constexpr int PER = COARSE_MAX_NODES / 256;
    __shared__ cx<T> xs[2][COARSE_MAX_NODES];
    const int r = blockIdx.x;
    if (live && !live[r]) return;
    const int n = (int)op.n;
    cx<T> cs[PER], c[PER][6], dv[PER], bv[PER];
    int off[PER][6];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = threadIdx.x + 256 * q;
        if (p < n) {
            const int i = p % op.nx, j = (p / op.nx) % op.ny, k = p / (op.nx * op.ny);
            const Nbr nb = neighbours(i, j, k, op.nx, op.ny, op.nz);
            row_coefs<T, MODE>(op, i, j, k, p, nb, cs[q], c[q]);
#pragma unroll
            for (int e = 0; e < 6; ++e) off[q][e] = (int)nb.off[e];
            dv[q] = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
            bv[q] = b[(int64_t)r * n + p];
            xs[0][p] = w * (dv[q] * bv[q]);
        }
    }
    __syncthreads();
    int cur = 0;
    for (int s = 1; s < sweeps; ++s) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int p = threadIdx.x + 256 * q;
            if (p < n) {
                const cx<T> x0 = xs[cur][p];
                cx<T> acc = cs[q] * x0;
#pragma unroll
                for (int e = 0; e < 6; ++e) acc = acc + c[q][e] * xs[cur][p + off[q][e]];
                xs[cur ^ 1][p] = x0 + w * (dv[q] * (bv[q] - acc));
            }
        }
        cur ^= 1;
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = threadIdx.x + 256 * q;
        if (p < n) x[(int64_t)r * n + p] = xs[cur][p];
    }
}

// Workgroup -> (64 x 4 tile, z-plane, right-hand side) of an nx x ny x nz
// grid with k slots: the slot is the slowest index and consecutive planes of
// one slot land on one XCD, so the rows that neighbouring tiles share stay in
// that XCD's L2.
struct PlaneTile {
    int i, j, k, r;
    bool in_grid;
};
__device__ __forceinline__ PlaneTile plane_tile(int nx, int ny, int nz, int nslots, int nblocks) {
    const int tile = xcd_tile(blockIdx.x, nblocks);
    const int tx = (nx + 63) >> 6, ty = (ny + 3) >> 2;
    const int per_slot = tx * ty * nz;
    PlaneTile t;
    t.r = tile / per_slot;
    const int rest = tile - t.r * per_slot;
    t.k = rest / (tx * ty);
    const int txy = rest - t.k * tx * ty;
    t.i = (txy % tx) * 64 + (threadIdx.x & 63);
    t.j = (txy / tx) * 4 + (threadIdx.x >> 6);
    t.in_grid = t.r < nslots && t.i < nx && t.j < ny;
    return t;
}

// Coarse right-hand side bc = P^T rf, one thread per (coarse node, slot);
// grid: plane tiles of the coarse grid times k slots, padded to a multiple
// of 8 (nblocks).  The 27 fine taps get separable weights; absent taps get
// weight 0 and a valid address, so every thread issues the same 27 loads.
// Identity rows (pressure-release faces) take no part in the coarse
// correction: their fine residuals are not restricted and coarse identity
// rows get zero, so a residual-scaled value never lands in a solution-scaled
// identity row.
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(const cx<T> *__restrict__ rf, int nxf, int nyf, int nzf,
                                                  cx<T> *__restrict__ bc, int nxc, int nyc, int nzc, int k,
                                                  const int *__restrict__ live, int soft, int nblocks) {
    const PlaneTile t = plane_tile(nxc, nyc, nzc, k, nblocks);
    if (!t.in_grid || (live && !live[t.r])) return;
    const int ic = t.i, jc = t.j, kc = t.k, r = t.r;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int64_t q = ic + (int64_t)nxc * (jc + (int64_t)nyc * kc);
    if (soft && on_soft_face(ic, jc, kc, nxc, nyc, nzc, soft)) {
        bc[r * nc + q] = mk<T>(0, 0);
        return;
    }
    const int fi = fine_of(ic, nxf), fj = fine_of(jc, nyf), fk = fine_of(kc, nzf);
    T wx[3], wy[3], wz[3];
    int ox[3], oy[3], oz[3];
    wx[1] = wy[1] = wz[1] = 1;
    wx[0] = (fi - 1 >= 0 && !coarse_node(fi - 1, nxf)) ? T(0.5) : T(0);
    wx[2] = (fi + 1 < nxf && !coarse_node(fi + 1, nxf)) ? T(0.5) : T(0);
    wy[0] = (fj - 1 >= 0 && !coarse_node(fj - 1, nyf)) ? T(0.5) : T(0);
    wy[2] = (fj + 1 < nyf && !coarse_node(fj + 1, nyf)) ? T(0.5) : T(0);
    wz[0] = (fk - 1 >= 0 && !coarse_node(fk - 1, nzf)) ? T(0.5) : T(0);
    wz[2] = (fk + 1 < nzf && !coarse_node(fk + 1, nzf)) ? T(0.5) : T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ox[d] = min(max(fi + d - 1, 0), nxf - 1);
        oy[d] = min(max(fj + d - 1, 0), nyf - 1);
        oz[d] = min(max(fk + d - 1, 0), nzf - 1);
    }
    const cx<T> *src = rf + r * nf;
    cx<T> val[27];
    T wt[27];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int e = (dz * 3 + dy) * 3 + dx;
                val[e] = src[ox[dx] + (int64_t)nxf * (oy[dy] + (int64_t)nyf * oz[dz])];
                T wv = wz[dz] * wy[dy] * wx[dx];
                if (soft && on_soft_face(ox[dx], oy[dy], oz[dz], nxf, nyf, nzf, soft)) wv = T(0);
                wt[e] = wv;
            }
    cx<T> acc = mk<T>(0, 0);
#pragma unroll
    for (int e = 0; e < 27; ++e) acc = acc + wt[e] * val[e];
    bc[r * nc + q] = acc;
}

// Fine iterate after the coarse correction, one thread per (fine node,
// slot); grid: plane tiles of the fine grid times k slots, padded to a
// multiple of 8 (nblocks).
//   b == nullptr: xf += P xc
//   b != nullptr: xf = f b + P xc with f = w dinv / bdiv (dinv conjugated when
//                 conj_d): the first Jacobi sweep from a zero guess is
//                 recomputed from b instead of being stored and read back.
// The 8 coarse taps always load (absent taps weight 0, valid address).  Fine
// identity rows get no correction.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(const cx<T> *__restrict__ xc, int nxc, int nyc, int nzc,
                                                 cx<T> *__restrict__ xf, int nxf, int nyf, int nzf, int k,
                                                 const int *__restrict__ live, int soft,
                                                 const cx<T> *__restrict__ b, const cx<T> *__restrict__ dinv, T w,
                                                 const T *__restrict__ bdiv, int conj_d, int nblocks) {
    const PlaneTile t = plane_tile(nxf, nyf, nzf, k, nblocks);
    if (!t.in_grid || (live && !live[t.r])) return;
    const int i = t.i, j = t.j, kk = t.k, r = t.r;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t p = i + (int64_t)nxf * (j + (int64_t)nyf * kk);
    const bool fixed = soft && on_soft_face(i, j, kk, nxf, nyf, nzf, soft);
    // Per axis: two coarse indices and weights (the second has weight 0 on
    // coarse nodes).  Coarse identity-row nodes carry no correction
    // (homogeneous Dirichlet for the correction): for A^H their values are
    // -A_IF^H x_I, scaled like a residual, and must not leak into interior
    // fine nodes.
    int ci[2], cj[2], ck[2];
    T wi[2], wj[2], wk[2];
    if (coarse_node(i, nxf)) { ci[0] = ci[1] = coarse_of(i, nxf, nxc); wi[0] = 1; wi[1] = 0; }
    else { ci[0] = coarse_of(i - 1, nxf, nxc); ci[1] = coarse_of(i + 1, nxf, nxc); wi[0] = wi[1] = T(0.5); }
    if (coarse_node(j, nyf)) { cj[0] = cj[1] = coarse_of(j, nyf, nyc); wj[0] = 1; wj[1] = 0; }
    else { cj[0] = coarse_of(j - 1, nyf, nyc); cj[1] = coarse_of(j + 1, nyf, nyc); wj[0] = wj[1] = T(0.5); }
    if (coarse_node(kk, nzf)) { ck[0] = ck[1] = coarse_of(kk, nzf, nzc); wk[0] = 1; wk[1] = 0; }
    else { ck[0] = coarse_of(kk - 1, nzf, nzc); ck[1] = coarse_of(kk + 1, nzf, nzc); wk[0] = wk[1] = T(0.5); }
    if (soft) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (on_soft_face(ci[a], 0, 1, nxc, 3, 3, soft & 3)) wi[a] = 0;
            if (on_soft_face(1, cj[a], 1, 3, nyc, 3, soft & 12)) wj[a] = 0;
            if (on_soft_face(1, 1, ck[a], 3, 3, nzc, soft & 48)) wk[a] = 0;
        }
    }
    const cx<T> *src = xc + r * nc;
    cx<T> val[8];
    T wt[8];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int u = (c * 2 + e) * 2 + a;
                val[u] = src[ci[a] + (int64_t)nxc * (cj[e] + (int64_t)nyc * ck[c])];
                wt[u] = fixed ? T(0) : wk[c] * wj[e] * wi[a];
            }
    cx<T> acc;
    if (b) {
        const cx<T> d = conj_d ? conj(dinv[p]) : dinv[p];
        acc = ((bdiv ? w / bdiv[p] : w) * d) * b[r * nf + p];
    } else {
        acc = xf[r * nf + p];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + wt[u] * val[u];
    xf[r * nf + p] = acc;
}

// x64 += cast<double>(z): apply a complex64 correction to the FP64 iterate.
__global__ __launch_bounds__(256) void k_accumulate(const cx<float> *__restrict__ z, int64_t n,
                                                    cx<double> *__restrict__ x, int64_t ldx,
                                                    const int *__restrict__ member) {
    const int r = blockIdx.y;
    if (member && !member[r]) return;
    for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
        x[r * ldx + p] = x[r * ldx + p] + cast<double>(z[r * n + p]);
}

}  // namespace hz
