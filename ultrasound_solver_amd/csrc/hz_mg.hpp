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
__device__ __forceinline__ void level_jacobi0_body(const cx<T> *__restrict__ b, cx<T> *__restrict__ xo,
        const cx<T> *__restrict__ dinv, T w, int64_t n, const int *__restrict__ live, const T *__restrict__ bdiv,
        const Blk &bk) {
    const int r = bk.y;
    if (live && !live[r]) return;
    for (int64_t p = bk.x * 256 + threadIdx.x; p < n; p += (int64_t)bk.gx * 256) {
        const cx<T> dv = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
        const T ws = bdiv ? w / bdiv[p] : w;
        xo[r * n + p] = ws * (dv * b[r * n + p]);
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_level_jacobi0(const cx<T> *__restrict__ b, cx<T> *__restrict__ xo,
        const cx<T> *__restrict__ dinv, T w, int64_t n, const int *__restrict__ live, const T *__restrict__ bdiv) {
    level_jacobi0_body<T, MODE>(b, xo, dinv, w, n, live, bdiv, blk_here());
}

template <typename T, int MODE>
struct KLevelJacobi0 {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *b;
        cx<T> *xo;
        const cx<T> *dinv;
        T w;
        int64_t n;
        const int *live;
        const T *bdiv;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        level_jacobi0_body<T, MODE>(a.b, a.xo, a.dinv, a.w, a.n, a.live, a.bdiv, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_level_jacobi0<T, MODE>), g, dim3(THREADS), 0, s, a.b, a.xo, a.dinv, a.w, a.n, a.live,
                a.bdiv);
    }
};

constexpr int COARSE_MAX_NODES = 1024;

// All damped-Jacobi sweeps of the coarsest level in one launch: one
// workgroup per right-hand side keeps the iterate in LDS (ping-pong) and
// each thread holds the coefficients of its <= 4 nodes in registers.
template <typename T, int MODE>
__device__ __forceinline__ void coarse_jacobi_body(const OpView<T> &op, const cx<T> *__restrict__ b,
        cx<T> *__restrict__ x, const cx<T> *__restrict__ dinv, T w, int sweeps, const int *__restrict__ live,
        const Blk &bk) {
    constexpr int PER = COARSE_MAX_NODES / 256;
    __shared__ cx<T> xs[2][COARSE_MAX_NODES];
    const int r = bk.x;
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
                for (int e = 0; e < 6; ++e) acc = cfma(c[q][e], xs[cur][p + off[q][e]], acc);
                xs[cur ^ 1][p] = sfma(w, dv[q] * (bv[q] - acc), x0);
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

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_coarse_jacobi(OpView<T> op, const cx<T> *__restrict__ b,
        cx<T> *__restrict__ x, const cx<T> *__restrict__ dinv, T w, int sweeps, const int *__restrict__ live) {
    coarse_jacobi_body<T, MODE>(op, b, x, dinv, w, sweeps, live, blk_here());
}

template <typename T, int MODE>
struct KCoarseJacobi {
    static constexpr int THREADS = 256;
    struct Args {
        OpView<T> op;
        const cx<T> *b;
        cx<T> *x;
        const cx<T> *dinv;
        T w;
        int sweeps;
        const int *live;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        coarse_jacobi_body<T, MODE>(a.op, a.b, a.x, a.dinv, a.w, a.sweeps, a.live, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_coarse_jacobi<T, MODE>), g, dim3(THREADS), 0, s, a.op, a.b, a.x, a.dinv, a.w, a.sweeps,
                a.live);
    }
};

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
__device__ __forceinline__ PlaneTile plane_tile(int bid, int nx, int ny, int nz, int nslots, int nblocks) {
    const int tile = xcd_tile(bid, nblocks);
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

// Coarse right-hand side bc = P^T rf marching in z (the restriction of the
// multigrid form): a workgroup owns 256 consecutive coarse nodes of an
// xy-plane (flattened, like plane_tile) and a chunk of RZ_CHUNK coarse planes
// of one slot.  Along x the taps of coarse node i are fine 2i-1, 2i, 2i+1
// with weights 1/2, 1, 1/2 (0 where absent or a coarse node), along y
// likewise.  Each fine plane's xy-restricted value
//   s(f) = sum_{dy, dx} wy(dy) wx(dx) rf(2i + dx, 2j + dy, f)
// is computed once per chunk and
// added to the coarse planes it feeds: weight 1 to the plane whose coarse
// node it is, 1/2 to each neighbouring coarse plane otherwise.  Fine planes
// go in pairs; the fine data are read about once (a thread per coarse node
// loading its nine fine rows reads them 2.25 times on average), and the
// 128^3 -> 65^3 restriction of 64 columns takes 323 us against 420 us for
// that per-node form.  Chunks of 4,
// 16 or all planes, plain workgroup order instead of the XCD remap, loads
// issued without the DPP shift, and four planes per step were all measured
// and no faster.  Identity rows (pressure-release faces) take no part in the
// coarse correction: their fine residuals are not restricted and coarse
// identity rows get zero, so a residual-scaled value never lands in a
// solution-scaled identity row.  Grid: (pieces of a coarse plane) x chunks x slots, padded to
// a multiple of 8 (nblocks).
constexpr int RZ_CHUNK = 8;

template <typename T, int PAIR>
__device__ __forceinline__ void restrict_z_body(const cx<T> *__restrict__ rf, int nxf, int nyf, int nzf,
        cx<T> *__restrict__ bc, int nxc, int nyc, int nzc, int k, const int *__restrict__ live, int soft, int nblocks,
        const Blk &bk) {
    const int tile = xcd_tile(bk.x, nblocks);
    const int nxy = nxc * nyc;
    const int pieces = (nxy + 255) >> 8;
    const int chunks = (nzc + RZ_CHUNK - 1) / RZ_CHUNK;
    const int r = tile / (pieces * chunks);
    if (r >= k) return;  // padding: whole workgroup
    if (live && !live[r]) return;
    const int rest = tile - r * pieces * chunks;
    const int chunk = rest / pieces;
    const int pxy_raw = (rest - chunk * pieces) * 256 + (int)threadIdx.x;
    const bool valid = pxy_raw < nxy;
    const int pxy = min(pxy_raw, nxy - 1);
    const int jc = pxy / nxc, ic = pxy - jc * nxc;
    const int lane = threadIdx.x & 63;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int fi = fine_of(ic, nxf), fj = fine_of(jc, nyf);
    T wx[3], wy[3];
    int ox[3], oy[3];
    wx[1] = wy[1] = 1;
    wx[0] = (fi - 1 >= 0 && !coarse_node(fi - 1, nxf)) ? T(0.5) : T(0);
    wx[2] = (fi + 1 < nxf && !coarse_node(fi + 1, nxf)) ? T(0.5) : T(0);
    wy[0] = (fj - 1 >= 0 && !coarse_node(fj - 1, nyf)) ? T(0.5) : T(0);
    wy[2] = (fj + 1 < nyf && !coarse_node(fj + 1, nyf)) ? T(0.5) : T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ox[d] = min(max(fi + d - 1, 0), nxf - 1);
        oy[d] = min(max(fj + d - 1, 0), nyf - 1);
    }
    T wt[3][3];  // xy tap weights, zero on fine identity rows of the x and y faces
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            T wv = wy[dy] * wx[dx];
            if ((soft & 15) && on_soft_face(ox[dx], oy[dy], 1, nxf, nyf, 3, soft & 15)) wv = T(0);
            wt[dy][dx] = wv;
        }
    const cx<T> *src = rf + r * nf;
    // Row values (x taps 2i-1, 2i, 2i+1) of fine row y = oy[dy], plane f:
    // the pair (2i, 2i+1) with one 16-byte access, 2i-1 from the previous
    // lane (DPP) or, in lane 0 and at a row start, loaded.
    auto load_rows = [&](int f, cx<T> (&v)[3][3]) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const int64_t row = (int64_t)nxf * (oy[dy] + (int64_t)nyf * f);
            if constexpr (PAIR) {
                static_assert(sizeof(T) == 4, "paired loads are for complex64");
                const int pidx = min(2 * ic, nxf - 2);
                const float4 pr = *reinterpret_cast<const float4 *>(src + row + pidx);
                const bool first = 2 * ic <= nxf - 2;
                v[dy][1] = first ? mk<float>(pr.x, pr.y) : mk<float>(pr.z, pr.w);
                v[dy][2] = first ? mk<float>(pr.z, pr.w) : mk<float>(0, 0);
                v[dy][0] = mk<float>(lane_prev(pr.z), lane_prev(pr.w));
                if (lane == 0 || ic == 0) v[dy][0] = src[row + ox[0]];
            } else {
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) v[dy][dx] = src[row + ox[dx]];
            }
        }
    };
    auto plane_sum = [&](const cx<T> (&v)[3][3]) {
        cx<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) acc = sfma(wt[dy][dx], v[dy][dx], acc);
        return acc;
    };
    const int kc0 = chunk * RZ_CHUNK, kc1 = min(nzc, kc0 + RZ_CHUNK);
    const int f_lo = max(fine_of(kc0, nzf) - 1, 0);
    const int f_hi = min(fine_of(kc1 - 1, nzf) + 1, nzf - 1);
    // Accumulators of coarse planes ka and ka + 1; a plane is written once no
    // later fine plane reaches it.
    int ka = kc0;
    cx<T> acc_a = mk<T>(0, 0), acc_b = mk<T>(0, 0);
    auto flush = [&]() {
        const bool fixed = soft && on_soft_face(ic, jc, ka, nxc, nyc, nzc, soft);
        if (valid) bc[r * nc + ic + (int64_t)nxc * (jc + (int64_t)nyc * ka)] = fixed ? mk<T>(0, 0) : acc_a;
        acc_a = acc_b;
        acc_b = mk<T>(0, 0);
        ++ka;
    };
    auto feed = [&](int f, cx<T> sv) {
        if ((soft & 48) && on_soft_face(1, 1, f, 3, 3, nzf, soft & 48)) return;  // identity plane
        const bool own = coarse_node(f, nzf);
        const int t0 = own ? coarse_of(f, nzf, nzc) : coarse_of(f - 1, nzf, nzc);
        const int t1 = own ? -1 : coarse_of(f + 1, nzf, nzc);
        const T w = own ? T(1) : T(0.5);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int kt = q ? t1 : t0;
            if (kt < kc0 || kt >= kc1) continue;
            while (kt > ka + 1) flush();
            if (kt == ka)
                acc_a = acc_a + w * sv;
            else
                acc_b = acc_b + w * sv;
        }
    };
    for (int f = f_lo; f <= f_hi; f += 2) {
        cx<T> v0[3][3], v1[3][3];
        const bool two = f + 1 <= f_hi;  // uniform over the workgroup
        load_rows(f, v0);
        if (two) load_rows(f + 1, v1);
        feed(f, plane_sum(v0));
        if (two) feed(f + 1, plane_sum(v1));
    }
    while (ka < kc1) flush();
}

template <typename T, int PAIR>
__global__ __launch_bounds__(256) void k_restrict_z(const cx<T> *__restrict__ rf, int nxf, int nyf, int nzf,
        cx<T> *__restrict__ bc, int nxc, int nyc, int nzc, int k, const int *__restrict__ live, int soft, int nblocks) {
    restrict_z_body<T, PAIR>(rf, nxf, nyf, nzf, bc, nxc, nyc, nzc, k, live, soft, nblocks, blk_here());
}

template <typename T, int PAIR>
struct KRestrictZ {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *rf;
        int nxf;
        int nyf;
        int nzf;
        cx<T> *bc;
        int nxc;
        int nyc;
        int nzc;
        int k;
        const int *live;
        int soft;
        int nblocks;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        restrict_z_body<T, PAIR>(a.rf, a.nxf, a.nyf, a.nzf, a.bc, a.nxc, a.nyc, a.nzc, a.k, a.live, a.soft, a.nblocks,
                b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_restrict_z<T, PAIR>), g, dim3(THREADS), 0, s, a.rf, a.nxf, a.nyf, a.nzf, a.bc, a.nxc,
                a.nyc, a.nzc, a.k, a.live, a.soft, a.nblocks);
    }
};

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
__device__ __forceinline__ void prolong_body(const cx<T> *__restrict__ xc, int nxc, int nyc, int nzc,
        cx<T> *__restrict__ xf, int nxf, int nyf, int nzf, int k, int rs, const int *__restrict__ live, int soft,
        const cx<T> *__restrict__ b, const cx<T> *__restrict__ dinv, T w, const T *__restrict__ bdiv, int conj_d,
        int nblocks, const Blk &bk) {
    const PlaneTile t = plane_tile(bk.x, nxc, nyc, nzc, (k + rs - 1) / rs, nblocks);
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
        // A node at offset 0 along an axis takes weight 0 from that axis's
        // second corner: those terms are left out (for a finite e, adding
        // 0 * e changes no value), the others are added in the same order.
        auto correct = [&](int q, cx<T> acc) {
            const int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
            if (!fixed[q]) {
#pragma unroll
                for (int c = 0; c <= dz; ++c)
#pragma unroll
                    for (int bq = 0; bq <= dy; ++bq)
#pragma unroll
                        for (int a = 0; a <= dx; ++a) {
                            const T wt = wgt(dz, c) * cw[2][c] * wgt(dy, bq) * cw[1][bq] * wgt(dx, a) * cw[0][a];
                            acc = sfma(wt, e[c][bq][a], acc);
                        }
            }
            return acc;
        };
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            const int q0 = 2 * row, q1 = q0 + 1;
            if (PAIR && here[q1]) {  // here[q1] implies here[q0]; pf[q0] is even
                // Each input pair is read once per launch: streamed past the caches.
                const hz_f4 in =
                        __builtin_nontemporal_load(reinterpret_cast<const hz_f4 *>((b ? b : xf) + r * nf + pf[q0]));
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

template <typename T, int PAIR>
__global__ __launch_bounds__(256) void k_prolong(const cx<T> *__restrict__ xc, int nxc, int nyc, int nzc,
        cx<T> *__restrict__ xf, int nxf, int nyf, int nzf, int k, int rs, const int *__restrict__ live, int soft,
        const cx<T> *__restrict__ b, const cx<T> *__restrict__ dinv, T w, const T *__restrict__ bdiv, int conj_d,
        int nblocks) {
    prolong_body<T, PAIR>(xc, nxc, nyc, nzc, xf, nxf, nyf, nzf, k, rs, live, soft, b, dinv, w, bdiv, conj_d, nblocks,
                          blk_here());
}

template <typename T, int PAIR>
struct KProlong {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<T> *xc;
        int nxc;
        int nyc;
        int nzc;
        cx<T> *xf;
        int nxf;
        int nyf;
        int nzf;
        int k;
        int rs;
        const int *live;
        int soft;
        const cx<T> *b;
        const cx<T> *dinv;
        T w;
        const T *bdiv;
        int conj_d;
        int nblocks;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        prolong_body<T, PAIR>(a.xc, a.nxc, a.nyc, a.nzc, a.xf, a.nxf, a.nyf, a.nzf, a.k, a.rs, a.live, a.soft, a.b,
                a.dinv, a.w, a.bdiv, a.conj_d, a.nblocks, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_prolong<T, PAIR>), g, dim3(THREADS), 0, s, a.xc, a.nxc, a.nyc, a.nzc, a.xf, a.nxf,
                a.nyf, a.nzf, a.k, a.rs, a.live, a.soft, a.b, a.dinv, a.w, a.bdiv, a.conj_d, a.nblocks);
    }
};

// x64 += cast<double>(z): apply a complex64 correction to the FP64 iterate.
__device__ __forceinline__ void accumulate_body(const cx<float> *__restrict__ z, int64_t n,
        cx<double> *__restrict__ x, int64_t ldx, const int *__restrict__ member, const Blk &bk) {
    const int r = bk.y;
    if (member && !member[r]) return;
    for (int64_t p = bk.x * 256 + threadIdx.x; p < n; p += (int64_t)bk.gx * 256)
        x[r * ldx + p] = x[r * ldx + p] + cast<double>(z[r * n + p]);
}

__global__ __launch_bounds__(256) void k_accumulate(const cx<float> *__restrict__ z, int64_t n,
        cx<double> *__restrict__ x, int64_t ldx, const int *__restrict__ member) {
    accumulate_body(z, n, x, ldx, member, blk_here());
}

struct KAccumulate {
    static constexpr int THREADS = 256;
    struct Args {
        const cx<float> *z;
        int64_t n;
        cx<double> *x;
        int64_t ldx;
        const int *member;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        accumulate_body(a.z, a.n, a.x, a.ldx, a.member, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_accumulate), g, dim3(THREADS), 0, s, a.z, a.n, a.x, a.ldx, a.member);
    }
};

// ---------------------------------------------------------------- hierarchy
// Coarse-level material fields derived on the device.  Even-node coarsening
// along an axis of n fine nodes keeps F(I) = min(2I, n - 1); a fine node that
// is not kept (odd, below n - 1) gives weight 1/2 to both kept neighbours.
// Normalised full weighting of a nodal field along one axis:
//   out[I] = sum_f (P[f, I] / sum_g P[g, I]) in[f],  f = F-1, F, F+1 in order,
// for two fields at once (1/rho and kappa).  Input grid (nx, ny, nz), output
// the same with `nc` nodes along `axis`.
__device__ __forceinline__ bool mg_kept(int f, int n) { return (f & 1) == 0 || f == n - 1; }

__global__ __launch_bounds__(256) void k_mg_weight_axis(const double *__restrict__ a_in,
                                                        const double *__restrict__ b_in, int nx, int ny, int nz,
                                                        int axis, int nc, double *__restrict__ a_out,
                                                        double *__restrict__ b_out) {
    const int dims[3] = {nx, ny, nz};
    int od[3] = {nx, ny, nz};
    od[axis] = nc;
    const int64_t nout = (int64_t)od[0] * od[1] * od[2];
    const int64_t q = blockIdx.x * 256ll + threadIdx.x;
    if (q >= nout) return;
    int idx[3];
    idx[0] = (int)(q % od[0]);
    idx[1] = (int)((q / od[0]) % od[1]);
    idx[2] = (int)(q / ((int64_t)od[0] * od[1]));
    const int n = dims[axis];
    const int F = min(2 * idx[axis], n - 1);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)nx : (int64_t)nx * ny);
    int src[3] = {idx[0], idx[1], idx[2]};
    src[axis] = F;
    const int64_t pF = src[0] + (int64_t)nx * (src[1] + (int64_t)ny * src[2]);
    const bool lo = F > 0 && !mg_kept(F - 1, n);
    const bool hi = F + 1 < n && !mg_kept(F + 1, n);
    const double total = 1.0 + (lo ? 0.5 : 0.0) + (hi ? 0.5 : 0.0);
    const double wc = 1.0 / total, wh = 0.5 / total;
    double a = 0.0, b = 0.0;
    if (lo) {
        a += wh * a_in[pF - stride];
        b += wh * b_in[pF - stride];
    }
    a += wc * a_in[pF];
    b += wc * b_in[pF];
    if (hi) {
        a += wh * a_in[pF + stride];
        b += wh * b_in[pF + stride];
    }
    a_out[q] = a;
    b_out[q] = b;
}

// inv_rho = 1 / rho and a copy of kappa (the finest level's FP64 fields).
__global__ __launch_bounds__(256) void k_mg_fields0(const double *__restrict__ rho, const double *__restrict__ kappa,
                                                    int64_t n, double *__restrict__ inv_rho,
                                                    double *__restrict__ kap) {
    const int64_t p = blockIdx.x * 256ll + threadIdx.x;
    if (p >= n) return;
    inv_rho[p] = 1.0 / rho[p];
    kap[p] = kappa[p];
}

// The level's single-precision fields: rho = 1 / inv_rho, kappa.
__global__ __launch_bounds__(256) void k_mg_level_fields(const double *__restrict__ inv_rho,
                                                         const double *__restrict__ kappa, int64_t n,
                                                         float *__restrict__ rho32, float *__restrict__ kap32) {
    const int64_t p = blockIdx.x * 256ll + threadIdx.x;
    if (p >= n) return;
    rho32[p] = (float)(1.0 / inv_rho[p]);
    kap32[p] = (float)kappa[p];
}

// Open-face term of a derived level on face f (axis f / 2, first or last
// plane): i omega (w_b w_c s_b s_c / dx) sqrt(kappa / rho) per face node
// e = ib + n_b ic, with the dual widths w and stretches s of the two in-plane
// axes (b < c), associated as the host-built hierarchy does.
__global__ __launch_bounds__(256) void k_mg_open_face(const double *__restrict__ inv_rho,
                                                      const double *__restrict__ kappa, int nx, int ny, int nz,
                                                      int f, const double *__restrict__ wb,
                                                      const cx<double> *__restrict__ sb,
                                                      const double *__restrict__ wc,
                                                      const cx<double> *__restrict__ sc, double omega, double dx,
                                                      cx<float> *__restrict__ out) {
    const int dims[3] = {nx, ny, nz};
    const int ax = f / 2;
    const int b0 = ax == 0 ? 1 : 0, c0 = ax == 2 ? 1 : 2;
    const int64_t cnt = (int64_t)dims[b0] * dims[c0];
    const int64_t e = blockIdx.x * 256ll + threadIdx.x;
    if (e >= cnt) return;
    const int ib = (int)(e % dims[b0]), ic = (int)(e / dims[b0]);
    int idx[3];
    idx[ax] = (f & 1) ? dims[ax] - 1 : 0;
    idx[b0] = ib;
    idx[c0] = ic;
    const int64_t p = idx[0] + (int64_t)nx * (idx[1] + (int64_t)ny * idx[2]);
    const double adm = sqrt(kappa[p] * inv_rho[p]);
    const cx<double> a = ((wb[ib] * wc[ic]) * sb[ib]) * sc[ic];
    const cx<double> area = mk<double>(a.re / dx, a.im / dx);
    const cx<double> t = mk<double>(-omega * area.im, omega * area.re);
    out[e] = cast<float>(mk<double>(t.re * adm, t.im * adm));
}

}  // namespace hz
