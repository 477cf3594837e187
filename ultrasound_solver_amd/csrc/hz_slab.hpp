// 2.5-D complex64 stencil: the hot apply of the mixed form and the
// multigrid smoother/residual.
//
// A workgroup of 64 x TY threads owns a 64 x TY (x, y) tile and a slab of ZB
// z-planes; the autotuner picks (ZB, TY) per grid from a few shapes.  The
// row coefficients of the thread's ZB nodes are computed once and kept in
// registers; then, for each right-hand side, the thread issues all loads of
// its column (planes k0-1 .. k0+ZB, in the first and last tile rows the halo
// row beyond the tile and, in lanes 0 and 63 only, the x-neighbour across the
// tile edge), takes the other x-neighbours from the adjacent lanes with DPP
// wave shifts and the other y-neighbours from the block's planes in LDS, and
// marches z with the (z-1, z, z+1) values in registers.  That is about
// (ZB + 2) / ZB + 2 / TY vector loads per node and right-hand side instead of 7
// (flat tiles of nx < 64 TY nodes: 2 nx / (64 TY) halo loads per node).
#pragma once

#include <algorithm>

#include "hz_stencil.hpp"

namespace hz {

// Tile shapes the autotuner chooses from: z-planes per thread, tile rows,
// right-hand sides per pass (loads of all of them issued together),
// non-temporal output stores, and flat tiles (64 * ty consecutive nodes of
// an xy-plane instead of a 64 x ty rectangle: no idle lanes on widths just
// above a multiple of 64, such as the 65-wide first coarse level).
// flat = 2: z-fastest tile order (rectangular tiles): the workgroups resident
// together on one XCD hold whole z-columns of slabs, so each slab's z-halo
// planes are the neighbouring slabs' own planes, loaded at about the same time.
// flat = 4: paired tiles (k_slab2): each lane owns two x-adjacent nodes of a
// 128 x TY tile and moves them with one 16-byte access per plane and row.
// Even widths and 16-byte aligned vectors only; launch_slab substitutes
// PAIR_FALLBACK elsewhere.
struct SlabShape {
    int zb, ty, rp, nt, flat;
};
constexpr int SLAB_SHAPES = 17;
constexpr SlabShape SLAB_SHAPE[SLAB_SHAPES] = {
    {4, 4, 1, 0, 0}, {2, 4, 1, 0, 0}, {4, 8, 1, 0, 0}, {2, 8, 1, 0, 0}, {4, 2, 1, 0, 0}, {2, 4, 2, 0, 0},
    {4, 4, 1, 1, 0}, {2, 4, 2, 1, 0}, {4, 4, 1, 1, 1}, {2, 4, 1, 1, 1}, {2, 4, 2, 1, 1}, {4, 4, 1, 1, 2},
    {2, 4, 2, 1, 2}, {4, 8, 1, 1, 2}, {2, 4, 1, 1, 4}, {2, 8, 1, 1, 4}, {4, 4, 1, 1, 4}};
constexpr int PAIR_FALLBACK = 6;

// Workgroups a slab launch aims for.  Grids with fewer tiles (the coarse
// multigrid levels, small scenes) split the right-hand sides into groups,
// one workgroup per tile and group, instead of leaving most CUs idle while
// each workgroup walks every column.
constexpr int SLAB_TARGET_BLOCKS = 1024;

__host__ inline TileMap slab_map(int nx, int ny, int nz, int shape, int k) {
    const SlabShape sh = SLAB_SHAPE[shape];
    TileMap t;
    // Flat tiles: tiles_x = 1 and tiles_y counts 64*ty-node pieces of a plane.
    const bool flat = sh.flat == 1;
    t.tiles_x = flat ? 1 : (sh.flat == 4 ? (nx + 127) / 128 : (nx + 63) / 64);
    t.tiles_y = flat ? (nx * ny + 64 * sh.ty - 1) / (64 * sh.ty) : (ny + sh.ty - 1) / sh.ty;
    t.zb = sh.zb;
    t.tiles_z = (nz + sh.zb - 1) / sh.zb;
    t.ntiles = t.tiles_x * t.tiles_y * t.tiles_z;
    t.nblocks = (t.ntiles + 7) / 8 * 8;
    const int passes = (k + sh.rp - 1) / sh.rp;
    t.csplit = std::max(1, std::min(passes, (SLAB_TARGET_BLOCKS + t.ntiles - 1) / t.ntiles));
    if (sh.flat == 0 || sh.flat == 4) xcd_blocks(t, ny, nz);
    return t;
}

enum SlabEpi { SLAB_APPLY = 0, SLAB_SMOOTH = 1, SLAB_RESID = 2, SLAB_RESID0 = 3 };


// Multiplier of the first damped-Jacobi sweep from a zero guess,
// x0 = f * b with f = w * dinv / bdiv (dinv conjugated for the adjoint).
template <int MODE>
__device__ __forceinline__ cx<float> first_sweep_factor(const cx<float> *__restrict__ dinv,
                                                        const float *__restrict__ bdiv, float w, int64_t p) {
    const cx<float> d = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
    return (bdiv ? w / bdiv[p] : w) * d;
}

// A coefficient times a first-sweep factor, rounded the same way in every
// kernel (no fused multiply-adds), so that the folded coefficients of the
// RESID0 epilogue agree bit for bit across tile shapes.
__device__ __forceinline__ cx<float> fold(cx<float> c, cx<float> f) {
#pragma clang fp contract(off)
    return mk<float>(c.re * f.re - c.im * f.im, c.re * f.im + c.im * f.re);
}

// APPLY : y = s_r * op(x)
// SMOOTH: y = x + w * dinv * (b / bdiv - op(x))     (damped Jacobi)
// RESID : y = b / bdiv - op(x)
// RESID0: y = b / bdiv - op(f b), f = w dinv / bdiv, with xin = b: the
//         residual after the first Jacobi sweep from a zero guess, without
//         writing that sweep.  f of each neighbour is folded into the row
//         coefficients once per node, so the right-hand-side loop is the
//         RESID loop reading b in place of x.
// Output store; NT = 1 marks it non-temporal (streamed past the caches, so
// that they keep the input halo rows the neighbouring tiles re-read).
typedef float hz_f2 __attribute__((ext_vector_type(2)));
template <int NT>
__device__ __forceinline__ void store_out(cx<float> *dst, cx<float> v) {
    if (NT) {
        hz_f2 t;
        t.x = v.re;
        t.y = v.im;
        __builtin_nontemporal_store(t, reinterpret_cast<hz_f2 *>(dst));
    } else {
        *dst = v;
    }
}

template <int MODE, int EPI, int ZB, int TY, int RP, int NT, int FLAT>
__device__ __forceinline__ void slab_body(const OpView<float> &op, const TileMap &tm,
        const cx<float> *__restrict__ xin, int64_t ldx, cx<float> *__restrict__ yout, int64_t ldy, int nrhs,
        const int *__restrict__ live, const double *__restrict__ scale, const cx<float> *__restrict__ b,
        const cx<float> *__restrict__ dinv, float w, const float *__restrict__ bdiv, const Blk &bk) {
    __shared__ int s_act[APPLY_MAX_RHS];
    __shared__ float s_sc[APPLY_MAX_RHS];
    __shared__ int s_na;
    // The y-neighbours come from the block's own planes in LDS (double-
    // buffered over passes, one barrier per pass) wherever they lie in the
    // tile: one row up or down (64 threads) in a rectangular tile, nx nodes
    // back or ahead in a flat one.  Only the others are loaded: the halo
    // rows of rectangular tiles, the first and last nx nodes of flat ones.
    // Separate global loads of the y-neighbour rows cost a third of the
    // throughput.
    constexpr int NT_ = 64 * TY;
    __shared__ cx<float> s_pl[2][RP][ZB][NT_];
    // nblocks is a multiple of 8, so a workgroup's XCD (its index modulo 8)
    // is the same in every column group.
    const int part = (int)bk.x / tm.nblocks;
    int tx, ty, tz;
    if (FLAT == 2) {  // z-fastest order
        const int tile = xcd_tile((int)bk.x - part * tm.nblocks, tm.nblocks);
        if (tile >= tm.ntiles) return;  // whole workgroup
        tz = tile % tm.tiles_z;
        tx = (tile / tm.tiles_z) % tm.tiles_x;
        ty = tile / (tm.tiles_z * tm.tiles_x);
    } else if (!slab_tile(tm, (int)bk.x - part * tm.nblocks, tx, ty, tz)) {
        return;  // whole workgroup
    }
    stage_live_columns<float>(nrhs, live, scale, s_act, s_sc, &s_na);
    const int lane = threadIdx.x & 63;
    // Lanes past the grid edge keep running (they feed the DPP shifts) on a
    // clamped node and never store.  In a flat tile the previous/next lane is
    // x-1/x+1 of the same row except at a row end, where the coefficient to
    // the absent neighbour is zero and any finite value may stand in.
    bool valid;
    int i, j;
    if (FLAT == 1) {
        const int nxy = op.nx * op.ny;
        const int raw = ty * (64 * TY) + (int)threadIdx.x;
        const int pxy = min(raw, nxy - 1);
        valid = raw < nxy;
        j = pxy / op.nx;
        i = pxy - j * op.nx;
    } else {
        const int i_raw = tx * 64 + lane;
        const int j_raw = ty * TY + (threadIdx.x >> 6);
        valid = i_raw < op.nx && j_raw < op.ny;
        i = min(i_raw, op.nx - 1);
        j = min(j_raw, op.ny - 1);
    }
    const int k0 = tz * tm.zb;
    const int nzl = min(op.nz - k0, ZB);
    const int64_t sxy = (int64_t)op.nx * op.ny;
    const int64_t p0 = i + (int64_t)op.nx * (j + (int64_t)op.ny * k0);

    cx<float> cs[ZB], c[ZB][6], dv[ZB];
    float bs[ZB];
#pragma unroll
    for (int q = 0; q < ZB; ++q) {
        const int kq = min(k0 + q, op.nz - 1);  // planes past the slab end repeat the last one
        const int64_t p = i + (int64_t)op.nx * (j + (int64_t)op.ny * kq);
        const Nbr nb = neighbours(i, j, kq, op.nx, op.ny, op.nz);
        row_coefs<float, MODE>(op, i, j, kq, p, nb, cs[q], c[q]);
        dv[q] = mk<float>(0, 0);
        bs[q] = 1.0f;
        if (EPI == SLAB_SMOOTH) dv[q] = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
        if (EPI != SLAB_APPLY && bdiv) bs[q] = 1.0f / bdiv[p];
        if (EPI == SLAB_RESID0) {
            cs[q] = fold(cs[q], first_sweep_factor<MODE>(dinv, bdiv, w, p));
#pragma unroll
            for (int e = 0; e < 6; ++e) c[q][e] = fold(c[q][e], first_sweep_factor<MODE>(dinv, bdiv, w, p + nb.off[e]));
        }
    }
    __syncthreads();  // s_act: staged while the coefficients were computed
    const bool has_xm = i > 0, has_xp = i < op.nx - 1;
    const int64_t oym = j > 0 ? -(int64_t)op.nx : 0;
    const int64_t oyp = j < op.ny - 1 ? (int64_t)op.nx : 0;
    const int tid = (int)threadIdx.x;
    const int ystep = FLAT == 1 ? op.nx : 64;  // LDS index distance of a y-step
    const bool lds_ym = tid >= ystep, lds_yp = tid + ystep < NT_;
    // This workgroup's live columns: group `part` of csplit groups of whole
    // passes (the last groups may be short or empty).
    const int na = s_na;
    const int per = ((na + tm.csplit - 1) / tm.csplit + RP - 1) / RP * RP;
    const int t_begin = part * per, t_end = min(na, t_begin + per);
    int buf = 0;
    for (int t = t_begin; t < t_end; t += RP) {
        // RP right-hand sides per pass.  Every load of their columns is
        // issued before any arithmetic: planes k0-1 .. k0+ZB (absent planes
        // read a valid node and are zeroed), the halo rows, the x-edge
        // values of lanes 0 and 63, and b for the SMOOTH/RESID epilogues.  A
        // pass past the group's last column repeats it and does not store.
        int rr[RP];
        cx<float> v[RP][ZB + 2], ym[RP][ZB], yp[RP][ZB], xe[RP][ZB], bb[RP][ZB];
#pragma unroll
        for (int u = 0; u < RP; ++u) {
            rr[u] = s_act[min(t + u, t_end - 1)];
            const cx<float> *xr = xin + rr[u] * ldx;
#pragma unroll
            for (int q = 0; q < ZB + 2; ++q) {
                const int kq = k0 + q - 1;
                const int kc = min(max(kq, 0), op.nz - 1);
                v[u][q] = xr[p0 + (int64_t)(kc - k0) * sxy];
                if (kq < 0 || kq >= op.nz) v[u][q] = mk<float>(0, 0);
            }
#pragma unroll
            for (int q = 0; q < ZB; ++q) {
                const int kc = min(k0 + q, op.nz - 1);
                const int64_t p = p0 + (int64_t)(kc - k0) * sxy;
                ym[u][q] = mk<float>(0, 0);
                yp[u][q] = mk<float>(0, 0);
                if (!lds_ym) ym[u][q] = xr[p + oym];
                if (!lds_yp) yp[u][q] = xr[p + oyp];
                xe[u][q] = mk<float>(0, 0);
                if (lane == 0 && has_xm) xe[u][q] = xr[p - 1];
                if (lane == 63 && has_xp) xe[u][q] = xr[p + 1];
                bb[u][q] = mk<float>(0, 0);
                if (EPI == SLAB_SMOOTH || EPI == SLAB_RESID) bb[u][q] = b[rr[u] * ldy + p];
            }
        }
        {
#pragma unroll
            for (int u = 0; u < RP; ++u)
#pragma unroll
                for (int q = 0; q < ZB; ++q) s_pl[buf][u][q][tid] = v[u][q + 1];
            __syncthreads();
#pragma unroll
            for (int u = 0; u < RP; ++u)
#pragma unroll
                for (int q = 0; q < ZB; ++q) {
                    if (lds_ym) ym[u][q] = s_pl[buf][u][q][tid - ystep];
                    if (lds_yp) yp[u][q] = s_pl[buf][u][q][tid + ystep];
                }
            buf ^= 1;
        }
#pragma unroll
        for (int u = 0; u < RP; ++u) {
            if (u > 0 && t + u >= t_end) break;
            const int r = rr[u];
#pragma unroll
            for (int q = 0; q < ZB; ++q) {
                const cx<float> xc = v[u][q + 1];
                cx<float> xl = mk<float>(lane_prev(xc.re), lane_prev(xc.im));
                cx<float> xn = mk<float>(lane_next(xc.re), lane_next(xc.im));
                if (lane == 0) xl = xe[u][q];
                if (lane == 63) xn = xe[u][q];
                cx<float> acc = cs[q] * xc;
                acc = cfma(c[q][0], xl, acc);
                acc = cfma(c[q][1], xn, acc);
                acc = cfma(c[q][2], ym[u][q], acc);
                acc = cfma(c[q][3], yp[u][q], acc);
                acc = cfma(c[q][4], v[u][q], acc);
                acc = cfma(c[q][5], v[u][q + 2], acc);
                if (valid && q < nzl) {
                    const int64_t p = p0 + (int64_t)q * sxy;
                    cx<float> out;
                    if (EPI == SLAB_APPLY) {
                        out = s_sc[t + u] * acc;
                    } else if (EPI == SLAB_RESID0) {
                        out = sfms(bs[q], xc, acc);
                    } else {
                        const cx<float> e = sfms(bs[q], bb[u][q], acc);
                        out = (EPI == SLAB_SMOOTH) ? sfma(w, dv[q] * e, xc) : e;
                    }
                    store_out<NT>(yout + r * ldy + p, out);
                }
            }
        }
    }
}

template <int MODE, int EPI, int ZB, int TY, int RP, int NT, int FLAT>
__global__ __launch_bounds__(64 * TY) void k_slab(OpView<float> op, TileMap tm, const cx<float> *__restrict__ xin,
        int64_t ldx, cx<float> *__restrict__ yout, int64_t ldy, int nrhs, const int *__restrict__ live,
        const double *__restrict__ scale, const cx<float> *__restrict__ b, const cx<float> *__restrict__ dinv,
        float w, const float *__restrict__ bdiv) {
    slab_body<MODE, EPI, ZB, TY, RP, NT, FLAT>(op, tm, xin, ldx, yout, ldy, nrhs, live, scale, b, dinv, w, bdiv,
                                               blk_here());
}

template <int MODE, int EPI, int ZB, int TY, int RP, int NT, int FLAT>
struct KSlab {
    static constexpr int THREADS = 64 * TY;
    struct Args {
        OpView<float> op;
        TileMap tm;
        const cx<float> *xin;
        int64_t ldx;
        cx<float> *yout;
        int64_t ldy;
        int nrhs;
        const int *live;
        const double *scale;
        const cx<float> *b;
        const cx<float> *dinv;
        float w;
        const float *bdiv;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        slab_body<MODE, EPI, ZB, TY, RP, NT, FLAT>(a.op, a.tm, a.xin, a.ldx, a.yout, a.ldy, a.nrhs, a.live, a.scale,
                a.b, a.dinv, a.w, a.bdiv, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_slab<MODE, EPI, ZB, TY, RP, NT, FLAT>), g, dim3(THREADS), 0, s, a.op, a.tm, a.xin,
                a.ldx, a.yout, a.ldy, a.nrhs, a.live, a.scale, a.b, a.dinv, a.w, a.bdiv);
    }
};

// Paired slab stencil (flat = 4): the arithmetic of k_slab per node, in the
// same order (bit-identical results), with each lane holding the nodes
// (2i', 2i'+1), i' = its lane in a 128-wide tile row.  x-neighbours: 2i'-1
// is the previous lane's second node and 2i'+2 the next lane's first (DPP);
// lanes 0 and 63 load the one across the tile edge.  One right-hand side per
// pass.  Requires an even nx and 16-byte aligned columns (even ldx, ldy).
template <int MODE, int EPI, int ZB, int TY, int NT>
__device__ __forceinline__ void slab2_body(const OpView<float> &op, const TileMap &tm,
        const cx<float> *__restrict__ xin, int64_t ldx, cx<float> *__restrict__ yout, int64_t ldy, int nrhs,
        const int *__restrict__ live, const double *__restrict__ scale, const cx<float> *__restrict__ b,
        const cx<float> *__restrict__ dinv, float w, const float *__restrict__ bdiv, const Blk &bk) {
    __shared__ int s_act[APPLY_MAX_RHS];
    __shared__ float s_sc[APPLY_MAX_RHS];
    __shared__ int s_na;
    constexpr int NT_ = 64 * TY;
    __shared__ hz_f4 s_pl[2][ZB][NT_];
    const int part = (int)bk.x / tm.nblocks;
    int tx, ty, tz;
    if (!slab_tile(tm, (int)bk.x - part * tm.nblocks, tx, ty, tz)) return;  // whole workgroup
    stage_live_columns<float>(nrhs, live, scale, s_act, s_sc, &s_na);
    const int lane = threadIdx.x & 63;
    // Lanes past the grid edge run on the last pair (feeding the shifts)
    // and never store.
    const int i_raw = tx * 128 + 2 * lane;
    const int j_raw = ty * TY + (threadIdx.x >> 6);
    const bool valid = i_raw < op.nx && j_raw < op.ny;
    const int i = min(i_raw, op.nx - 2);
    const int j = min(j_raw, op.ny - 1);
    const int k0 = tz * tm.zb;
    const int nzl = min(op.nz - k0, ZB);
    const int64_t sxy = (int64_t)op.nx * op.ny;
    const int64_t p0 = i + (int64_t)op.nx * (j + (int64_t)op.ny * k0);

    cx<float> cs[ZB][2], c[ZB][2][6], dv[ZB][2];
    float bs[ZB][2];
#pragma unroll
    for (int q = 0; q < ZB; ++q)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int kq = min(k0 + q, op.nz - 1);
            const int64_t p = i + e + (int64_t)op.nx * (j + (int64_t)op.ny * kq);
            const Nbr nb = neighbours(i + e, j, kq, op.nx, op.ny, op.nz);
            row_coefs<float, MODE>(op, i + e, j, kq, p, nb, cs[q][e], c[q][e]);
            dv[q][e] = mk<float>(0, 0);
            bs[q][e] = 1.0f;
            if (EPI == SLAB_SMOOTH) dv[q][e] = (MODE == MODE_AH) ? conj(dinv[p]) : dinv[p];
            if (EPI != SLAB_APPLY && bdiv) bs[q][e] = 1.0f / bdiv[p];
        }
    if constexpr (EPI == SLAB_RESID0) {
        // First-sweep factors, one per distinct node around the pair's
        // column: the window planes k0-1 .. k0+ZB of both nodes, the rows
        // j -+ 1 and the x-neighbours across the pair, each folded into every
        // coefficient that reaches it (as k_slab does per neighbour; an absent
        // neighbour's coefficient is zero, so its clamped stand-in is moot).
        cx<float> fw[ZB + 2][2], fym[ZB][2], fyp[ZB][2], fxm[ZB], fxp[ZB];
        const int jm = max(j - 1, 0), jp = min(j + 1, op.ny - 1);
        const int im = max(i - 1, 0), ip = min(i + 2, op.nx - 1);
        auto at = [&](int ii, int jj, int kk) { return ii + (int64_t)op.nx * (jj + (int64_t)op.ny * kk); };
#pragma unroll
        for (int q = 0; q < ZB + 2; ++q) {
            const int kc = min(max(k0 + q - 1, 0), op.nz - 1);
#pragma unroll
            for (int e = 0; e < 2; ++e) fw[q][e] = first_sweep_factor<MODE>(dinv, bdiv, w, at(i + e, j, kc));
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            const int kc = min(k0 + q, op.nz - 1);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                fym[q][e] = first_sweep_factor<MODE>(dinv, bdiv, w, at(i + e, jm, kc));
                fyp[q][e] = first_sweep_factor<MODE>(dinv, bdiv, w, at(i + e, jp, kc));
            }
            fxm[q] = first_sweep_factor<MODE>(dinv, bdiv, w, at(im, j, kc));
            fxp[q] = first_sweep_factor<MODE>(dinv, bdiv, w, at(ip, j, kc));
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                cs[q][e] = fold(cs[q][e], fw[q + 1][e]);
                c[q][e][0] = fold(c[q][e][0], e ? fw[q + 1][0] : fxm[q]);
                c[q][e][1] = fold(c[q][e][1], e ? fxp[q] : fw[q + 1][1]);
                c[q][e][2] = fold(c[q][e][2], fym[q][e]);
                c[q][e][3] = fold(c[q][e][3], fyp[q][e]);
                c[q][e][4] = fold(c[q][e][4], fw[q][e]);
                c[q][e][5] = fold(c[q][e][5], fw[q + 2][e]);
            }
    }
    __syncthreads();  // s_act
    const bool has_xm = i > 0, has_xp = i + 2 < op.nx;
    const int64_t oym = j > 0 ? -(int64_t)op.nx : 0;
    const int64_t oyp = j < op.ny - 1 ? (int64_t)op.nx : 0;
    const int tid = (int)threadIdx.x;
    const bool lds_ym = tid >= 64, lds_yp = tid + 64 < NT_;
    const int na = s_na;
    const int per = (na + tm.csplit - 1) / tm.csplit;
    const int t_begin = part * per, t_end = min(na, t_begin + per);
    const hz_f4 zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
    int buf = 0;
    for (int t = t_begin; t < t_end; ++t) {
        const int r = s_act[t];
        const cx<float> *xr = xin + r * ldx;
        hz_f4 v[ZB + 2], ym[ZB], yp[ZB], bb[ZB];
        cx<float> xe[ZB];
#pragma unroll
        for (int q = 0; q < ZB + 2; ++q) {
            const int kq = k0 + q - 1;
            const int kc = min(max(kq, 0), op.nz - 1);
            v[q] = *reinterpret_cast<const hz_f4 *>(xr + p0 + (int64_t)(kc - k0) * sxy);
            if (kq < 0 || kq >= op.nz) v[q] = zero4;
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            const int kc = min(k0 + q, op.nz - 1);
            const int64_t p = p0 + (int64_t)(kc - k0) * sxy;
            ym[q] = zero4;
            yp[q] = zero4;
            if (!lds_ym) ym[q] = *reinterpret_cast<const hz_f4 *>(xr + p + oym);
            if (!lds_yp) yp[q] = *reinterpret_cast<const hz_f4 *>(xr + p + oyp);
            xe[q] = mk<float>(0, 0);
            if (lane == 0 && has_xm) xe[q] = xr[p - 1];
            if (lane == 63 && has_xp) xe[q] = xr[p + 2];
            bb[q] = zero4;
            if (EPI == SLAB_SMOOTH || EPI == SLAB_RESID) bb[q] = *reinterpret_cast<const hz_f4 *>(b + r * ldy + p);
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) s_pl[buf][q][tid] = v[q + 1];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            if (lds_ym) ym[q] = s_pl[buf][q][tid - 64];
            if (lds_yp) yp[q] = s_pl[buf][q][tid + 64];
        }
        buf ^= 1;
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            const hz_f4 c4 = v[q + 1];
            cx<float> xl = mk<float>(lane_prev(c4.z), lane_prev(c4.w));
            cx<float> xn = mk<float>(lane_next(c4.x), lane_next(c4.y));
            if (lane == 0) xl = xe[q];
            if (lane == 63) xn = xe[q];
            cx<float> out[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const cx<float> xc = e ? mk<float>(c4.z, c4.w) : mk<float>(c4.x, c4.y);
                const cx<float> xm = e ? mk<float>(c4.x, c4.y) : xl;
                const cx<float> xp = e ? xn : mk<float>(c4.z, c4.w);
                const cx<float> ymv = e ? mk<float>(ym[q].z, ym[q].w) : mk<float>(ym[q].x, ym[q].y);
                const cx<float> ypv = e ? mk<float>(yp[q].z, yp[q].w) : mk<float>(yp[q].x, yp[q].y);
                const cx<float> zm = e ? mk<float>(v[q].z, v[q].w) : mk<float>(v[q].x, v[q].y);
                const cx<float> zp = e ? mk<float>(v[q + 2].z, v[q + 2].w) : mk<float>(v[q + 2].x, v[q + 2].y);
                cx<float> acc = cs[q][e] * xc;
                acc = cfma(c[q][e][0], xm, acc);
                acc = cfma(c[q][e][1], xp, acc);
                acc = cfma(c[q][e][2], ymv, acc);
                acc = cfma(c[q][e][3], ypv, acc);
                acc = cfma(c[q][e][4], zm, acc);
                acc = cfma(c[q][e][5], zp, acc);
                if (EPI == SLAB_APPLY) {
                    out[e] = s_sc[t] * acc;
                } else if (EPI == SLAB_RESID0) {
                    out[e] = sfms(bs[q][e], xc, acc);
                } else {
                    const cx<float> bv = e ? mk<float>(bb[q].z, bb[q].w) : mk<float>(bb[q].x, bb[q].y);
                    const cx<float> ev = sfms(bs[q][e], bv, acc);
                    out[e] = (EPI == SLAB_SMOOTH) ? sfma(w, dv[q][e] * ev, xc) : ev;
                }
            }
            if (valid && q < nzl) {
                hz_f4 o;
                o.x = out[0].re;
                o.y = out[0].im;
                o.z = out[1].re;
                o.w = out[1].im;
                hz_f4 *dst = reinterpret_cast<hz_f4 *>(yout + r * ldy + p0 + (int64_t)q * sxy);
                if (NT)
                    __builtin_nontemporal_store(o, dst);
                else
                    *dst = o;
            }
        }
    }
}

template <int MODE, int EPI, int ZB, int TY, int NT>
__global__ __launch_bounds__(64 * TY) void k_slab2(OpView<float> op, TileMap tm, const cx<float> *__restrict__ xin,
        int64_t ldx, cx<float> *__restrict__ yout, int64_t ldy, int nrhs, const int *__restrict__ live,
        const double *__restrict__ scale, const cx<float> *__restrict__ b, const cx<float> *__restrict__ dinv,
        float w, const float *__restrict__ bdiv) {
    slab2_body<MODE, EPI, ZB, TY, NT>(op, tm, xin, ldx, yout, ldy, nrhs, live, scale, b, dinv, w, bdiv, blk_here());
}

template <int MODE, int EPI, int ZB, int TY, int NT>
struct KSlab2 {
    static constexpr int THREADS = 64 * TY;
    struct Args {
        OpView<float> op;
        TileMap tm;
        const cx<float> *xin;
        int64_t ldx;
        cx<float> *yout;
        int64_t ldy;
        int nrhs;
        const int *live;
        const double *scale;
        const cx<float> *b;
        const cx<float> *dinv;
        float w;
        const float *bdiv;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        slab2_body<MODE, EPI, ZB, TY, NT>(a.op, a.tm, a.xin, a.ldx, a.yout, a.ldy, a.nrhs, a.live, a.scale, a.b,
                a.dinv, a.w, a.bdiv, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_slab2<MODE, EPI, ZB, TY, NT>), g, dim3(THREADS), 0, s, a.op, a.tm, a.xin, a.ldx, a.yout,
                a.ldy, a.nrhs, a.live, a.scale, a.b, a.dinv, a.w, a.bdiv);
    }
};

// FP64 apply of the plain form on the slab layout: y[:, r] = s_r op(x[:, r])
// with the row coefficients of row_coefs and the neighbour terms added in
// the order x-, x+, y-, y+, z-, z+ after the centre, the order of
// k_residual.  A workgroup of 64 x TY threads owns a 64 x TY (x, y) tile and
// ZB z-planes; the row coefficients of the thread's ZB nodes are computed
// once; per right-hand side the thread loads planes k0-1 .. k0+ZB of its own
// column (the z-neighbours stay in registers), takes the y-neighbours from
// the block's planes in LDS (halo rows loaded by the first and last tile
// rows) and the x-neighbours from the adjacent lanes by DPP (lanes 0 and 63
// load the one across the tile edge): about (ZB + 2) / ZB + 2 / TY loads of
// x per node and column instead of 7.
__device__ __forceinline__ double lane_prev_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_next_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int SLAB64_ZB = 2, SLAB64_TY = 4;

__host__ inline TileMap slab64_map(int nx, int ny, int nz, int k) {
    TileMap t;
    t.tiles_x = (nx + 63) / 64;
    t.tiles_y = (ny + SLAB64_TY - 1) / SLAB64_TY;
    t.zb = SLAB64_ZB;
    t.tiles_z = (nz + SLAB64_ZB - 1) / SLAB64_ZB;
    t.ntiles = t.tiles_x * t.tiles_y * t.tiles_z;
    t.nblocks = (t.ntiles + 7) / 8 * 8;
    t.csplit = std::max(1, std::min(k, (SLAB_TARGET_BLOCKS + t.ntiles - 1) / t.ntiles));
    xcd_blocks(t, ny, nz);
    return t;
}

template <int MODE, int ZB, int TY>
__device__ __forceinline__ void slab64_body(const OpView<double> &op, const TileMap &tm,
        const cx<double> *__restrict__ xin, int64_t ldx, cx<double> *__restrict__ yout, int64_t ldy, int nrhs,
        const int *__restrict__ live, const double *__restrict__ scale, const Blk &bk) {
    __shared__ int s_act[APPLY_MAX_RHS];
    __shared__ double s_sc[APPLY_MAX_RHS];
    __shared__ int s_na;
    constexpr int NT_ = 64 * TY;
    __shared__ cx<double> s_pl[2][ZB][NT_];
    const int part = (int)bk.x / tm.nblocks;
    int tx, ty, tz;
    if (!slab_tile(tm, (int)bk.x - part * tm.nblocks, tx, ty, tz)) return;  // whole workgroup
    stage_live_columns<double>(nrhs, live, scale, s_act, s_sc, &s_na);
    const int lane = threadIdx.x & 63;
    // Lanes past the grid edge run on a clamped node (feeding the shifts)
    // and never store.
    const int i_raw = tx * 64 + lane;
    const int j_raw = ty * TY + (int)(threadIdx.x >> 6);
    const bool valid = i_raw < op.nx && j_raw < op.ny;
    const int i = min(i_raw, op.nx - 1);
    const int j = min(j_raw, op.ny - 1);
    const int k0 = tz * tm.zb;
    const int nzl = min(op.nz - k0, ZB);
    const int64_t sxy = (int64_t)op.nx * op.ny;
    const int64_t p0 = i + (int64_t)op.nx * (j + (int64_t)op.ny * k0);

    cx<double> cs[ZB], c[ZB][6];
#pragma unroll
    for (int q = 0; q < ZB; ++q) {
        const int kq = min(k0 + q, op.nz - 1);
        const int64_t p = i + (int64_t)op.nx * (j + (int64_t)op.ny * kq);
        const Nbr nb = neighbours(i, j, kq, op.nx, op.ny, op.nz);
        row_coefs<double, MODE>(op, i, j, kq, p, nb, cs[q], c[q]);
    }
    __syncthreads();  // s_act
    const bool has_xm = i > 0, has_xp = i < op.nx - 1;
    const int64_t oym = j > 0 ? -(int64_t)op.nx : 0;
    const int64_t oyp = j < op.ny - 1 ? (int64_t)op.nx : 0;
    const int tid = (int)threadIdx.x;
    const bool lds_ym = tid >= 64, lds_yp = tid + 64 < NT_;
    const int na = s_na;
    const int per = (na + tm.csplit - 1) / tm.csplit;
    const int t_begin = part * per, t_end = min(na, t_begin + per);
    int buf = 0;
    for (int t = t_begin; t < t_end; ++t) {
        const int r = s_act[t];
        const cx<double> *xr = xin + r * ldx;
        cx<double> v[ZB + 2], ym[ZB], yp[ZB], xe[ZB];
#pragma unroll
        for (int q = 0; q < ZB + 2; ++q) {
            const int kq = k0 + q - 1;
            const int kc = min(max(kq, 0), op.nz - 1);
            v[q] = xr[p0 + (int64_t)(kc - k0) * sxy];
            if (kq < 0 || kq >= op.nz) v[q] = mk<double>(0, 0);
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            const int kc = min(k0 + q, op.nz - 1);
            const int64_t p = p0 + (int64_t)(kc - k0) * sxy;
            ym[q] = mk<double>(0, 0);
            yp[q] = mk<double>(0, 0);
            if (!lds_ym) ym[q] = xr[p + oym];
            if (!lds_yp) yp[q] = xr[p + oyp];
            xe[q] = mk<double>(0, 0);
            if (lane == 0 && has_xm) xe[q] = xr[p - 1];
            if (lane == 63 && has_xp) xe[q] = xr[p + 1];
        }
#pragma unroll
        for (int q = 0; q < ZB; ++q) s_pl[buf][q][tid] = v[q + 1];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            if (lds_ym) ym[q] = s_pl[buf][q][tid - 64];
            if (lds_yp) yp[q] = s_pl[buf][q][tid + 64];
        }
        buf ^= 1;
#pragma unroll
        for (int q = 0; q < ZB; ++q) {
            const cx<double> xc = v[q + 1];
            cx<double> xl = mk<double>(lane_prev_f64(xc.re), lane_prev_f64(xc.im));
            cx<double> xn = mk<double>(lane_next_f64(xc.re), lane_next_f64(xc.im));
            if (lane == 0) xl = xe[q];
            if (lane == 63) xn = xe[q];
            cx<double> acc = cs[q] * xc;
            acc = cfma(c[q][0], xl, acc);
            acc = cfma(c[q][1], xn, acc);
            acc = cfma(c[q][2], ym[q], acc);
            acc = cfma(c[q][3], yp[q], acc);
            acc = cfma(c[q][4], v[q], acc);
            acc = cfma(c[q][5], v[q + 2], acc);
            if (valid && q < nzl) yout[r * ldy + p0 + (int64_t)q * sxy] = s_sc[t] * acc;
        }
    }
}

template <int MODE, int ZB, int TY>
__global__ __launch_bounds__(64 * TY) void k_slab64(OpView<double> op, TileMap tm, const cx<double> *__restrict__ xin,
        int64_t ldx, cx<double> *__restrict__ yout, int64_t ldy, int nrhs, const int *__restrict__ live,
        const double *__restrict__ scale) {
    slab64_body<MODE, ZB, TY>(op, tm, xin, ldx, yout, ldy, nrhs, live, scale, blk_here());
}

template <int MODE, int ZB, int TY>
struct KSlab64 {
    static constexpr int THREADS = 64 * TY;
    struct Args {
        OpView<double> op;
        TileMap tm;
        const cx<double> *xin;
        int64_t ldx;
        cx<double> *yout;
        int64_t ldy;
        int nrhs;
        const int *live;
        const double *scale;
    };
    __device__ __forceinline__ static void run(const Args &a, const Blk &b) {
        slab64_body<MODE, ZB, TY>(a.op, a.tm, a.xin, a.ldx, a.yout, a.ldy, a.nrhs, a.live, a.scale, b);
    }
    static void single(dim3 g, hipStream_t s, const Args &a) {
        hipLaunchKernelGGL((k_slab64<MODE, ZB, TY>), g, dim3(THREADS), 0, s, a.op, a.tm, a.xin, a.ldx, a.yout, a.ldy,
                a.nrhs, a.live, a.scale);
    }
};

}  // namespace hz
