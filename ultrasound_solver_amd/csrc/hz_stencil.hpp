// Matrix-free 7-point finite-volume Helmholtz stencil for gfx950.
//
// Row P of A (reference: src/algorithm/physics/conservative_helmholtz.py:119-149):
//   A_PQ = coef_PQ = a_b a_c q_a / (rho_P + rho_Q)        for the 6 axis neighbours,
//   A_PP = MW_P omega^2 kappa_P - sum_Q coef_PQ + face terms,
// with a = w*s per axis, q_a = 4/((s_i + s_{i+1}) h^2) (the midpoint stretch),
// MW = a_x a_y a_z.  Rows of nodes on soft faces are identity rows; their
// neighbours keep the coefficient into them.  Nothing but rho and kappa is
// stored per node; the 1-D tables and the O(n^2) face tables stay in cache.
//
// Thread layout: a 256-thread workgroup owns a 64 x 4 (x, y) tile and marches
// `zb` planes in z, so x neighbours are in the same cache line, y neighbours
// in the tile and z neighbours were touched one plane earlier.  The 1-D grid
// is remapped so that consecutive tiles share an XCD (and its L2): the
// hardware deals workgroups round-robin over the 8 XCDs.
#pragma once

#include "hz_common.hpp"

namespace hz {

enum ApplyMode { MODE_A = 0, MODE_RAM = 1, MODE_AH = 2, MODE_RAMH = 3, MODE_RA = 4, MODE_RAH = 5 };

struct TileMap {
    int tiles_x, tiles_y, tiles_z, zb;
    int ntiles;
    int nblocks;     // padded to a multiple of 8
    int csplit = 1;  // column groups: workgroup b works on group b / nblocks (slab kernel)
    int py = 0, pz = 0;  // XCD blocks of a slab launch (py * pz == 8), or 0 (xcd_tile order)
};

__host__ inline TileMap make_tilemap(int nx, int ny, int nz, int zb) {
    TileMap t;
    t.tiles_x = (nx + 63) / 64;
    t.tiles_y = (ny + 3) / 4;
    t.zb = zb;
    t.tiles_z = (nz + zb - 1) / zb;
    t.ntiles = t.tiles_x * t.tiles_y * t.tiles_z;
    t.nblocks = (t.ntiles + 7) / 8 * 8;
    return t;
}

// Workgroup -> tile with consecutive tiles on one XCD.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    const int per = nblocks >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// XCD blocks of a slab launch.  Workgroup b runs on XCD b % 8, and each XCD
// has its own L2, so the halo rows and planes a tile reads from a tile of
// another XCD are fetched twice.  The 8 XCDs take py x pz blocks of the tile
// grid (split in y and z; a tile walks every column, so all of a block's
// tiles are resident together).  At 128^3 x 32 columns a z-boundary between
// blocks costs about twice the fetched bytes of a y-boundary of the same
// area (z-halo planes are read by both sides), so the split minimises
// (py - 1) nz + 2 (pz - 1) ny: 4 x 2 there, 1.13x -> 1.06x the algorithmic
// bytes.  Tile counts that do not divide keep xcd_tile's contiguous ranges.
__host__ inline void xcd_blocks(TileMap &t, int ny, int nz) {
    t.py = t.pz = 0;
    long best = -1;
    for (int py = 1; py <= 8; py *= 2) {
        const int pz = 8 / py;
        if (t.tiles_y % py != 0 || t.tiles_z % pz != 0) continue;
        const long cost = (long)(py - 1) * nz + 2L * (pz - 1) * ny;
        if (best < 0 || cost < best) {
            best = cost;
            t.py = py;
            t.pz = pz;
        }
    }
}

// Tile (tx, ty, tz) of workgroup b (0 <= b < nblocks) of a slab launch;
// false for the padding workgroups past the last tile.  With XCD blocks the
// tiles of a block go x fastest, then y, then z.
__device__ __forceinline__ bool slab_tile(const TileMap &tm, int b, int &tx, int &ty, int &tz) {
    if (tm.py > 0) {
        const int xcd = b & 7, w = b >> 3;
        const int by = tm.tiles_y / tm.py, bz = tm.tiles_z / tm.pz;
        tx = w % tm.tiles_x;
        ty = (xcd % tm.py) * by + (w / tm.tiles_x) % by;
        tz = (xcd / tm.py) * bz + w / (tm.tiles_x * by);
        return true;
    }
    const int tile = xcd_tile(b, tm.nblocks);
    tx = tile % tm.tiles_x;
    ty = (tile / tm.tiles_x) % tm.tiles_y;
    tz = tile / (tm.tiles_x * tm.tiles_y);
    return tile < tm.ntiles;
}

template <typename T>
__device__ __forceinline__ bool node_fixed(const OpView<T> &op, int i, int j, int k) {
    const int s = op.soft;
    return ((s & 1) && i == 0) || ((s & 2) && i == op.nx - 1) || ((s & 4) && j == 0) ||
           ((s & 8) && j == op.ny - 1) || ((s & 16) && k == 0) || ((s & 32) && k == op.nz - 1);
}

// Neighbour order: x-, x+, y-, y+, z-, z+.
struct Nbr {
    int64_t off[6];
    bool has[6];
};

__device__ __forceinline__ Nbr neighbours(int i, int j, int k, int nx, int ny, int nz) {
    Nbr nb;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
    nb.has[0] = i > 0;
    nb.has[1] = i < nx - 1;
    nb.has[2] = j > 0;
    nb.has[3] = j < ny - 1;
    nb.has[4] = k > 0;
    nb.has[5] = k < nz - 1;
    nb.off[0] = nb.has[0] ? -sx : 0;
    nb.off[1] = nb.has[1] ? sx : 0;
    nb.off[2] = nb.has[2] ? -sy : 0;
    nb.off[3] = nb.has[3] ? sy : 0;
    nb.off[4] = nb.has[4] ? -sz : 0;
    nb.off[5] = nb.has[5] ? sz : 0;
    return nb;
}

// Edge coefficients coef_PQ (zero where the neighbour is absent) and, when
// asked, the unmodified diagonal D_P (mass - sum coef + face terms).  All
// neighbour loads are issued unconditionally (an absent neighbour aliases P
// and is weighted by zero) so that they overlap instead of serialising
// behind per-neighbour branches.
// 1 / (a + b) for the edge factor of two positive densities: exact
// division in FP64 (the plain form reproduces the reference's matrix), the
// hardware reciprocal (1 ulp) in complex64 arithmetic, where six correctly
// rounded divisions per node were the bulk of the coefficient setup.
__device__ __forceinline__ double inv_sum(double a, double b) { return 1.0 / (a + b); }
__device__ __forceinline__ float inv_sum(float a, float b) { return __builtin_amdgcn_rcpf(a + b); }

template <typename T, bool WANT_DIAG>
__device__ __forceinline__ void edge_coefs(const OpView<T> &op, int i, int j, int k, int64_t p,
                                           const Nbr &nb, cx<T> c[6], cx<T> *diag) {
    T rq[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) rq[n] = op.rho[p + nb.off[n]];
    const T rp = op.rho[p];
    const cx<T> ax = op.a[0][i], ay = op.a[1][j], az = op.a[2][k];
    const cx<T> q0 = op.q[0][i > 0 ? i - 1 : 0], q1 = op.q[0][i < op.nx - 1 ? i : op.nx - 2];
    const cx<T> q2 = op.q[1][j > 0 ? j - 1 : 0], q3 = op.q[1][j < op.ny - 1 ? j : op.ny - 2];
    const cx<T> q4 = op.q[2][k > 0 ? k - 1 : 0], q5 = op.q[2][k < op.nz - 1 ? k : op.nz - 2];
    const cx<T> ayz = ay * az, axz = ax * az, axy = ax * ay;
    c[0] = (nb.has[0] ? inv_sum(rp, rq[0]) : T(0)) * (ayz * q0);
    c[1] = (nb.has[1] ? inv_sum(rp, rq[1]) : T(0)) * (ayz * q1);
    c[2] = (nb.has[2] ? inv_sum(rp, rq[2]) : T(0)) * (axz * q2);
    c[3] = (nb.has[3] ? inv_sum(rp, rq[3]) : T(0)) * (axz * q3);
    c[4] = (nb.has[4] ? inv_sum(rp, rq[4]) : T(0)) * (axy * q4);
    c[5] = (nb.has[5] ? inv_sum(rp, rq[5]) : T(0)) * (axy * q5);
    if (WANT_DIAG) {
        const cx<T> mw = axy * az;
        const T kp = op.kappa[p];
        cx<T> d = (op.omega2 * kp) * mw;
        if (op.omega2_im != T(0)) d = d + mk<T>(0, op.omega2_im * kp) * mw;
        // Same subtraction order as the reference: per axis, the +edge then the -edge.
        d = d - c[1];
        d = d - c[0];
        d = d - c[3];
        d = d - c[2];
        d = d - c[5];
        d = d - c[4];
        if (i == 0 && op.face[0]) d = d + op.face[0][j + (int64_t)op.ny * k];
        if (i == op.nx - 1 && op.face[1]) d = d + op.face[1][j + (int64_t)op.ny * k];
        if (j == 0 && op.face[2]) d = d + op.face[2][i + (int64_t)op.nx * k];
        if (j == op.ny - 1 && op.face[3]) d = d + op.face[3][i + (int64_t)op.nx * k];
        if (k == 0 && op.face[4]) d = d + op.face[4][i + (int64_t)op.nx * j];
        if (k == op.nz - 1 && op.face[5]) d = d + op.face[5][i + (int64_t)op.nx * j];
        *diag = d;
    }
}

// Final row coefficients for a mode: y_P = cs * x_P + sum_n c[n] * x_{P+off[n]}.
//   A, AH      : rows of A / A^H
//   RA, RAH    : the same rows times the row scale of A / A^H
//   RAM, RAMH  : R A M with the Jacobi inverse M of R A (likewise for A^H)
template <typename T, int MODE>
__device__ __forceinline__ void row_coefs(const OpView<T> &op, int i, int j, int k, int64_t p,
                                          const Nbr &nb, cx<T> &cs, cx<T> c[6]) {
    constexpr bool ADJ = (MODE == MODE_AH || MODE == MODE_RAMH || MODE == MODE_RAH);
    constexpr bool PRE = (MODE == MODE_RAM || MODE == MODE_RAMH);
    constexpr bool SCALED = (MODE == MODE_RA || MODE == MODE_RAH);
    cx<T> d = mk<T>(0, 0);
    edge_coefs<T, !PRE>(op, i, j, k, p, nb, c, &d);
    const bool fixed_p = node_fixed(op, i, j, k);
    if (!ADJ) {
        if (fixed_p) {
#pragma unroll
            for (int n = 0; n < 6; ++n) c[n] = mk<T>(0, 0);
        }
    } else {
        // Column P of A: neighbour Q contributes only when row Q is not an identity row.
        const int di[6] = {-1, 1, 0, 0, 0, 0};
        const int dj[6] = {0, 0, -1, 1, 0, 0};
        const int dk[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
        for (int n = 0; n < 6; ++n) {
            if (op.soft && nb.has[n] && node_fixed(op, i + di[n], j + dj[n], k + dk[n]))
                c[n] = mk<T>(0, 0);
            c[n] = conj(c[n]);
        }
        d = conj(d);
    }
    if (!PRE) {
        cs = fixed_p ? mk<T>(1, 0) : d;
        if (SCALED) {
            const T r = (MODE == MODE_RA ? op.rs_f : op.rs_a)[p];
            cs = r * cs;
#pragma unroll
            for (int n = 0; n < 6; ++n) c[n] = r * c[n];
        }
    } else {
        // R A M: the diagonal is r_P D_P / (r_P D_P) = 1 exactly.
        const T *rs = (MODE == MODE_RAM) ? op.rs_f : op.rs_a;
        const cx<T> *m = (MODE == MODE_RAM) ? op.m_f : op.m_a;
        cx<T> mq[6];
#pragma unroll
        for (int n = 0; n < 6; ++n) mq[n] = m[p + nb.off[n]];
        const T r = rs[p];
        cs = mk<T>(1, 0);
#pragma unroll
        for (int n = 0; n < 6; ++n) c[n] = r * (c[n] * mq[n]);
    }
}

constexpr int APPLY_MAX_RHS = 256;

// Compacts the live right-hand sides (all of them when live is null) into
// s_act and their scales into s_sc, with one wave testing 64 columns per
// round: one load latency per round instead of one per column.  Called by
// the whole workgroup; the caller's barrier publishes the result.
template <typename S>
__device__ __forceinline__ void stage_live_columns(int nrhs, const int *__restrict__ live,
                                                   const double *__restrict__ scale, int *s_act, S *s_sc,
                                                   int *s_na) {
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int base = 0;
    for (int r0 = 0; r0 < nrhs; r0 += 64) {
        const int r = r0 + lane;
        const bool on = r < nrhs && (!live || live[r] != 0);
        const uint64_t mask = __ballot(on);
        if (on) {
            const int pos = base + __popcll(mask & below);
            s_act[pos] = r;
            s_sc[pos] = scale ? (S)scale[r] : S(1);
        }
        base += __popcll(mask);
    }
    if (lane == 0) *s_na = base;
}

// out[:, r] = b[:, r] - op(x[:, r]) with per-workgroup FP64 partials of
// |out|^2; one right-hand side per grid row (blockIdx.y).
template <typename T, int MODE>
__device__ __forceinline__ void residual_body(const OpView<T> &op, const TileMap &tm, const cx<T> *__restrict__ x,
        int64_t ldx, const cx<T> *__restrict__ b, int64_t ldb, cx<T> *__restrict__ out, int64_t ldo,
        const int *__restrict__ live, double *__restrict__ partial, const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    if (live && !live[r]) return;
    const int tile = xcd_tile(bk.x, tm.nblocks);
    double acc2 = 0.0;
    if (tile < tm.ntiles) {
        const int tx = tile % tm.tiles_x;
        const int ty = (tile / tm.tiles_x) % tm.tiles_y;
        const int tz = tile / (tm.tiles_x * tm.tiles_y);
        const int i = tx * 64 + (threadIdx.x & 63);
        const int j = ty * 4 + (threadIdx.x >> 6);
        if (i < op.nx && j < op.ny) {
            const int k0 = tz * tm.zb;
            const int k1 = min(op.nz, k0 + tm.zb);
            T loc = 0;
            for (int k = k0; k < k1; ++k) {
                const int64_t p = i + (int64_t)op.nx * (j + (int64_t)op.ny * k);
                const Nbr nb = neighbours(i, j, k, op.nx, op.ny, op.nz);
                cx<T> cs, c[6];
                row_coefs<T, MODE>(op, i, j, k, p, nb, cs, c);
                const cx<T> *xr = x + r * ldx + p;
                cx<T> acc = cs * xr[0];
#pragma unroll
                for (int n = 0; n < 6; ++n) acc = cfma(c[n], xr[nb.off[n]], acc);
                const cx<T> e = b[r * ldb + p] - acc;
                out[r * ldo + p] = e;
                loc += norm2(e);
            }
            acc2 = (double)loc;
        }
    }
    const double s = block_sum_256(acc2, red);
    if (threadIdx.x == 0) partial[(int64_t)r * bk.gx + bk.x] = s;
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_residual(OpView<T> op, TileMap tm, const cx<T> *__restrict__ x, int64_t ldx,
        const cx<T> *__restrict__ b, int64_t ldb, cx<T> *__restrict__ out, int64_t ldo, const int *__restrict__ live,
        double *__restrict__ partial) {
    residual_body<T, MODE>(op, tm, x, ldx, b, ldb, out, ldo, live, partial, blk_here());
}

template <typename T, int MODE>
struct KResidual {
    static constexpr int THREADS = 256;
    struct Args {
        OpView<T> op;
        TileMap tm;
        const cx<T> *x;
        int64_t ldx;
        const cx<T> *b;
        int64_t ldb;
        cx<T> *out;
        int64_t ldo;
        const int *live;
        double *partial;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        residual_body<T, MODE>(a.op, a.tm, a.x, a.ldx, a.b, a.ldb, a.out, a.ldo, a.live, a.partial, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_residual<T, MODE>), g, dim3(THREADS), 0, s, a.op, a.tm, a.x, a.ldx, a.b, a.ldb, a.out,
                a.ldo, a.live, a.partial);
    }
};

// Acceptance check of the reference (iterative_solver.py:70-97) in FP64 for
// x (the physical solution): e = op(x) - b, t = |op| |x| + |b|.  Partials per
// workgroup: [ |R e|^2, |R b|^2, |e|^2, |b|^2, max |e|/max(t,1e-30), #nonfinite ].
// When `rhs_out` is set it also writes the scaled refinement residual
// -R e, rounded to complex64.
template <int MODE>
__device__ __forceinline__ void check_body(const OpView<double> &op, const TileMap &tm,
        const cx<double> *__restrict__ x, int64_t ldx, const cx<double> *__restrict__ b, int64_t ldb,
        const int *__restrict__ live, double *__restrict__ partial, cx<float> *__restrict__ rhs_out, int64_t ldr,
        const Blk &bk) {
    __shared__ double red[4];
    const int r = bk.y;
    if (live && !live[r]) return;
    const int tile = xcd_tile(bk.x, tm.nblocks);
    double se = 0, sb = 0, ue = 0, ub = 0, mx = 0, bad = 0;
    if (tile < tm.ntiles) {
        const int tx = tile % tm.tiles_x;
        const int ty = (tile / tm.tiles_x) % tm.tiles_y;
        const int tz = tile / (tm.tiles_x * tm.tiles_y);
        const int i = tx * 64 + (threadIdx.x & 63);
        const int j = ty * 4 + (threadIdx.x >> 6);
        if (i < op.nx && j < op.ny) {
            const int k0 = tz * tm.zb;
            const int k1 = min(op.nz, k0 + tm.zb);
            const double *rs = (MODE == MODE_A) ? op.rs_f : op.rs_a;
            for (int k = k0; k < k1; ++k) {
                const int64_t p = i + (int64_t)op.nx * (j + (int64_t)op.ny * k);
                const Nbr nb = neighbours(i, j, k, op.nx, op.ny, op.nz);
                cx<double> cs, c[6];
                row_coefs<double, MODE>(op, i, j, k, p, nb, cs, c);
                const cx<double> *xr = x + r * ldx + p;
                const cx<double> xp = xr[0];
                cx<double> acc = cs * xp;
                double t = cabs(cs) * cabs(xp);
#pragma unroll
                for (int n = 0; n < 6; ++n) {
                    const cx<double> xq = xr[nb.off[n]];
                    acc = cfma(c[n], xq, acc);
                    t += cabs(c[n]) * cabs(xq);
                }
                const cx<double> bp = b[r * ldb + p];
                const cx<double> e = acc - bp;
                const double scale = rs[p];
                const double ae = cabs(e), ab = cabs(bp);
                t += ab;
                se += scale * scale * norm2(e);
                sb += scale * scale * norm2(bp);
                ue += norm2(e);
                ub += norm2(bp);
                mx = fmax(mx, ae / fmax(t, 1.0e-30));
                if (!(isfinite(xp.re) && isfinite(xp.im))) bad += 1.0;
                if (rhs_out) rhs_out[r * ldr + p] = cast<float>((-scale) * e);
            }
        }
    }
    double *dst = partial + ((int64_t)r * bk.gx + bk.x) * 6;
    double v;
    v = block_sum_256(se, red);
    if (threadIdx.x == 0) dst[0] = v;
    v = block_sum_256(sb, red);
    if (threadIdx.x == 0) dst[1] = v;
    v = block_sum_256(ue, red);
    if (threadIdx.x == 0) dst[2] = v;
    v = block_sum_256(ub, red);
    if (threadIdx.x == 0) dst[3] = v;
    {
        double w = wave_max(mx);
        if ((threadIdx.x % WAVE) == 0) red[threadIdx.x / WAVE] = w;
        __syncthreads();
        if (threadIdx.x == 0) dst[4] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        __syncthreads();
    }
    v = block_sum_256(bad, red);
    if (threadIdx.x == 0) dst[5] = v;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_check(OpView<double> op, TileMap tm, const cx<double> *__restrict__ x,
        int64_t ldx, const cx<double> *__restrict__ b, int64_t ldb, const int *__restrict__ live,
        double *__restrict__ partial, cx<float> *__restrict__ rhs_out, int64_t ldr) {
    check_body<MODE>(op, tm, x, ldx, b, ldb, live, partial, rhs_out, ldr, blk_here());
}

template <int MODE>
struct KCheck {
    static constexpr int THREADS = 256;
    struct Args {
        OpView<double> op;
        TileMap tm;
        const cx<double> *x;
        int64_t ldx;
        const cx<double> *b;
        int64_t ldb;
        const int *live;
        double *partial;
        cx<float> *rhs_out;
        int64_t ldr;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        check_body<MODE>(a.op, a.tm, a.x, a.ldx, a.b, a.ldb, a.live, a.partial, a.rhs_out, a.ldr, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_check<MODE>), g, dim3(THREADS), 0, s, a.op, a.tm, a.x, a.ldx, a.b, a.ldb, a.live,
                a.partial, a.rhs_out, a.ldr);
    }
};

// Row scales and Jacobi inverses of A and A^H (iterative_solver.py:37-56),
// computed once in FP64 and rounded for the FP32 view.
__global__ __launch_bounds__(256) void k_setup(OpView<double> op, double *rs_f, cx<double> *m_f, double *rs_a,
                                               cx<double> *m_a, float *rs_f32, cx<float> *m_f32, float *rs_a32,
                                               cx<float> *m_a32) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= op.n) return;
    const int i = (int)(p % op.nx);
    const int j = (int)((p / op.nx) % op.ny);
    const int k = (int)(p / ((int64_t)op.nx * op.ny));
    const Nbr nb = neighbours(i, j, k, op.nx, op.ny, op.nz);
    cx<double> c[6], d;
    edge_coefs<double, true>(op, i, j, k, p, nb, c, &d);
    const bool fixed_p = node_fixed(op, i, j, k);
    const cx<double> diag = fixed_p ? mk<double>(1, 0) : d;
    // Row P of A.
    double mf = cabs(diag);
    if (!fixed_p)
        for (int n = 0; n < 6; ++n)
            if (nb.has[n]) mf = fmax(mf, cabs(c[n]));
    // Row P of A^H: conj of column P of A.
    const int di[6] = {-1, 1, 0, 0, 0, 0};
    const int dj[6] = {0, 0, -1, 1, 0, 0};
    const int dk[6] = {0, 0, 0, 0, -1, 1};
    double ma = cabs(diag);
    for (int n = 0; n < 6; ++n)
        if (nb.has[n] && !node_fixed(op, i + di[n], j + dj[n], k + dk[n])) ma = fmax(ma, cabs(c[n]));
    const double rf = mf > 0 ? 1.0 / mf : 1.0;
    const double ra = ma > 0 ? 1.0 / ma : 1.0;
    const cx<double> df = rf * diag;
    const cx<double> da = ra * conj(diag);
    const cx<double> jf = cabs(df) > 0 ? crecip(df) : mk<double>(1, 0);
    const cx<double> ja = cabs(da) > 0 ? crecip(da) : mk<double>(1, 0);
    rs_f[p] = rf;
    m_f[p] = jf;
    rs_a[p] = ra;
    m_a[p] = ja;
    rs_f32[p] = (float)rf;
    m_f32[p] = cast<float>(jf);
    rs_a32[p] = (float)ra;
    m_a32[p] = cast<float>(ja);
}

}  // namespace hz
