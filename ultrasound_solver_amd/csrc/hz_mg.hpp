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

// Workgroup -> (256 consecutive nodes of one xy-plane, z-plane, slot) of an
// nx x ny x nz grid with k slots.  Flattening the plane leaves no idle lanes
// on awkward widths (the coarse levels are 65, 33, 17 ... wide).  The slot is
// the slowest index and consecutive planes of one slot land on one XCD, so
// the rows that neighbouring workgroups share stay in that XCD's L2.  Lanes
// past the plane's end are clamped to its last node and flagged.
struct PlaneTile {
    int i, j, k, r;
    bool in_grid;
};
__device__ __forceinline__ PlaneTile plane_tile(int nx, int ny, int nz, int nslots, int nblocks) {
    const int tile = xcd_tile(blockIdx.x, nblocks);
    const int nxy = nx * ny;
    const int per_plane = (nxy + 255) >> 8;
    const int per_slot = per_plane * nz;
    PlaneTile t;
    t.r = tile / per_slot;
    const int rest = tile - t.r * per_slot;
    t.k = rest / per_plane;
    const int pxy_raw = (rest - t.k * per_plane) * 256 + (int)threadIdx.x;
    const int pxy = min(pxy_raw, nxy - 1);
    t.j = pxy / nx;
    t.i = pxy - t.j * nx;
    t.in_grid = t.r < nslots && pxy_raw < nxy;
    return t;
}

// Coarse right-hand side bc = P^T rf, one thread per (coarse node, group of
// rs slots): the node's taps and weights are computed once for the group.
// Grid: flattened plane tiles of the coarse grid times ceil(k / rs) slot
// groups, padded to a multiple of 8 (nblocks).  Along x the taps are fine 2i-1, 2i, 2i+1 with weights 1/2, 1,
// 1/2 (0 where absent or a coarse node); along y and z likewise.  PAIR (even
// fine x-size): each lane loads its (2i, 2i+1) pair with one 16-byte access
// and takes 2i-1 from the previous lane (DPP), lane 0 loading it itself:
// 9 loads per slot instead of 27.  Lanes past the plane's end run on clamped
// indices (they feed the shifts) and do not store.  Identity rows
// (pressure-release faces) take no part in the coarse correction: their
// fine residuals are not restricted and coarse identity rows get zero, so a
// residual-scaled value never lands in a solution-scaled identity row.
template <typename T, int PAIR>
__global__ __launch_bounds__(256) void k_restrict(const cx<T> *__restrict__ rf, int nxf, int nyf, int nzf,
                                                  cx<T> *__restrict__ bc, int nxc, int nyc, int nzc, int k, int rs,
                                                  const int *__restrict__ live, int soft, int nblocks) {
    const PlaneTile t = plane_tile(nxc, nyc, nzc, (k + rs - 1) / rs, nblocks);
    // Slot group and plane are uniform over the workgroup.
    if (t.r * rs >= k) return;
    const bool valid = t.in_grid;
    const int ic = t.i, jc = t.j, kc = t.k;
    const int r0 = t.r * rs, r1 = min(k, r0 + rs);
    const int lane = threadIdx.x & 63;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int64_t q = ic + (int64_t)nxc * (jc + (int64_t)nyc * kc);
    const bool fixed = soft && on_soft_face(ic, jc, kc, nxc, nyc, nzc, soft);
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
    T wt[3][3][3];  // tap weights, zero on fine identity rows
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                T wv = wz[dz] * wy[dy] * wx[dx];
                if (soft && on_soft_face(ox[dx], oy[dy], oz[dz], nxf, nyf, nzf, soft)) wv = T(0);
                wt[dz][dy][dx] = wv;
            }
    for (int r = r0; r < r1; ++r) {
        if (live && !live[r]) continue;  // uniform over the workgroup
        const cx<T> *src = rf + r * nf;
        cx<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int64_t row = (int64_t)nxf * (oy[dy] + (int64_t)nyf * oz[dz]);
                cx<T> v[3];
                if constexpr (PAIR) {
                    static_assert(sizeof(T) == 4, "paired loads are for complex64");
                    // Pair (2i, 2i+1); the last coarse node of an even grid is fine
                    // node nxf-1, the second element of the pair (nxf-2, nxf-1).
                    const int pidx = min(2 * ic, nxf - 2);
                    const float4 pr = *reinterpret_cast<const float4 *>(src + row + pidx);
                    const bool first = 2 * ic <= nxf - 2;
                    v[1] = first ? mk<float>(pr.x, pr.y) : mk<float>(pr.z, pr.w);
                    v[2] = first ? mk<float>(pr.z, pr.w) : mk<float>(0, 0);
                    // The previous lane holds x-1 of the same row except in lane 0
                    // and at a row start (where the left tap has weight 0).
                    v[0] = mk<float>(lane_prev(pr.z), lane_prev(pr.w));
                    if (lane == 0 || ic == 0) v[0] = src[row + ox[0]];
                } else {
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx) v[dx] = src[row + ox[dx]];
                }
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) acc = acc + wt[dz][dy][dx] * v[dx];
            }
        if (valid) bc[r * nc + q] = fixed ? mk<T>(0, 0) : acc;
    }
}

// Fine nodes of coarse cell c along one axis (fine size n, coarse size m):
// cell c < m-1 spans fine [fine_of(c), fine_of(c+1)) (one or two nodes), the
// last cell is the last fine node.  The first node of a cell is the coarse
// node itself (weight 1 from corner c); a second node sits halfway (0.5 from
// corners c and c+1).
struct CellAxis {
    int f0, count, c1;  // first fine node, fine nodes in the cell, second corner (clamped)
};
__device__ __forceinline__ CellAxis cell_axis(int c, int n, int m) {
    CellAxis a;
    a.f0 = fine_of(c, n);
    a.count = c + 1 < m ? fine_of(c + 1, n) - a.f0 : 1;
    a.c1 = min(c + 1, m - 1);
    return a;
}

// Fine iterate after the coarse correction, one thread per (coarse cell,
// group of rs slots): the cell's corner weights and its (up to 2 x 2 x 2)
// fine nodes' first-sweep factors are computed once, then for each live slot
// of the group the 8 coarse corners are loaded and the fine nodes written.
// Grid: flattened plane tiles of the coarse grid times ceil(k / rs) slot
// groups, padded to a multiple of 8 (nblocks).
//   b == nullptr: xf += P xc
//   b != nullptr: xf = f b + P xc with f = w dinv / bdiv (dinv conjugated when
//                 conj_d): the first Jacobi sweep from a zero guess is
//                 recomputed from b instead of being stored and read back.
// Coarse identity-row corners carry no correction (homogeneous Dirichlet for
// the correction: for A^H their values are -A_IF^H x_I, scaled like a
// residual, and must not leak into interior fine nodes), and fine identity
// rows get none.  PAIR (even fine x-size, complex64): the two fine nodes of a
// cell row, 2i and 2i+1, are read (b, or the iterate) and written with one
// 16-byte access each.
template <typename T, int PAIR>
__global__ __launch_bounds__(256) void k_prolong(const cx<T> *__restrict__ xc, int nxc, int nyc, int nzc,
                                                 cx<T> *__restrict__ xf, int nxf, int nyf, int nzf, int k, int rs,
                                                 const int *__restrict__ live, int soft,
                                                 const cx<T> *__restrict__ b, const cx<T> *__restrict__ dinv, T w,
                                                 const T *__restrict__ bdiv, int conj_d, int nblocks) {
    const PlaneTile t = plane_tile(nxc, nyc, nzc, (k + rs - 1) / rs, nblocks);
    if (!t.in_grid) return;
    const int r0 = t.r * rs, r1 = min(k, r0 + rs);
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const CellAxis ax = cell_axis(t.i, nxf, nxc), ay = cell_axis(t.j, nyf, nyc), az = cell_axis(t.k, nzf, nzc);
    // Corner weights per axis: corner 0 is the cell's own coarse node.
    T cw[3][2];  // [axis][corner]: 0 for a coarse identity-row corner
    cw[0][0] = (soft && on_soft_face(t.i, 0, 1, nxc, 3, 3, soft & 3)) ? T(0) : T(1);
    cw[0][1] = (soft && on_soft_face(ax.c1, 0, 1, nxc, 3, 3, soft & 3)) ? T(0) : T(1);
    cw[1][0] = (soft && on_soft_face(1, t.j, 1, 3, nyc, 3, soft & 12)) ? T(0) : T(1);
    cw[1][1] = (soft && on_soft_face(1, ay.c1, 1, 3, nyc, 3, soft & 12)) ? T(0) : T(1);
    cw[2][0] = (soft && on_soft_face(1, 1, t.k, 3, 3, nzc, soft & 48)) ? T(0) : T(1);
    cw[2][1] = (soft && on_soft_face(1, 1, az.c1, 3, 3, nzc, soft & 48)) ? T(0) : T(1);
    const int cx0[2] = {t.i, ax.c1}, cy0[2] = {t.j, ay.c1}, cz0[2] = {t.k, az.c1};
    int64_t corner[2][2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int bq = 0; bq < 2; ++bq)
#pragma unroll
            for (int a = 0; a < 2; ++a) corner[c][bq][a] = cx0[a] + (int64_t)nxc * (cy0[bq] + (int64_t)nyc * cz0[c]);
    // Weights of fine offset d (0: the coarse node, 1: the midpoint) from corner q.
    auto wgt = [](int d, int q) -> T { return d == 0 ? (q == 0 ? T(1) : T(0)) : T(0.5); };
    // Per fine node of the cell (index dz*4 + dy*2 + dx): offset, presence,
    // identity-row flag and first-sweep factor.
    int64_t pf[8];
    bool here[8], fixed[8];
    cx<T> f[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
        here[q] = dz < az.count && dy < ay.count && dx < ax.count;
        const int fi = ax.f0 + (here[q] ? dx : 0), fj = ay.f0 + (here[q] ? dy : 0), fk = az.f0 + (here[q] ? dz : 0);
        pf[q] = fi + (int64_t)nxf * (fj + (int64_t)nyf * fk);
        fixed[q] = soft && on_soft_face(fi, fj, fk, nxf, nyf, nzf, soft);
        f[q] = mk<T>(0, 0);
        if (b && here[q]) {
            const cx<T> d = conj_d ? conj(dinv[pf[q]]) : dinv[pf[q]];
            f[q] = (bdiv ? w / bdiv[pf[q]] : w) * d;
        }
    }
    for (int r = r0; r < r1; ++r) {
        if (live && !live[r]) continue;
        const cx<T> *src = xc + r * nc;
        cx<T> e[2][2][2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int bq = 0; bq < 2; ++bq)
#pragma unroll
                for (int a = 0; a < 2; ++a) e[c][bq][a] = src[corner[c][bq][a]];
        auto correct = [&](int q, cx<T> acc) {
            const int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
            if (!fixed[q]) {
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int bq = 0; bq < 2; ++bq)
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            const T wt = wgt(dz, c) * cw[2][c] * wgt(dy, bq) * cw[1][bq] * wgt(dx, a) * cw[0][a];
                            acc = acc + wt * e[c][bq][a];
                        }
            }
            return acc;
        };
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            const int q0 = 2 * row, q1 = q0 + 1;
            if (PAIR && here[q1]) {  // here[q1] implies here[q0]; pf[q0] is even
                const hz_f4 in = *reinterpret_cast<const hz_f4 *>((b ? b : xf) + r * nf + pf[q0]);
                cx<T> a0 = mk<T>(in.x, in.y), a1 = mk<T>(in.z, in.w);
                if (b) {
                    a0 = f[q0] * a0;
                    a1 = f[q1] * a1;
                }
                a0 = correct(q0, a0);
                a1 = correct(q1, a1);
                hz_f4 out;
                out.x = a0.re;
                out.y = a0.im;
                out.z = a1.re;
                out.w = a1.im;
                *reinterpret_cast<hz_f4 *>(xf + r * nf + pf[q0]) = out;
                continue;
            }
#pragma unroll
            for (int q = q0; q <= q1; ++q) {
                if (!here[q]) continue;
                const cx<T> acc = b ? f[q] * b[r * nf + pf[q]] : xf[r * nf + pf[q]];
                xf[r * nf + pf[q]] = correct(q, acc);
            }
        }
    }
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
