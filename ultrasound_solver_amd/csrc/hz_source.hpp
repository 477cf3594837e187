// Sinc-aperture source deposition on the device (reference:
// src/algorithm/physics/sinc_aperture.py:36-72).  Every quadrature point of
// an emitter's aperture spreads its weight bilinearly over the four nodes
// around it on the z = 0 plane.  The masses must be bit-identical to the
// reference's, so each node's mass is one sequential FP64 sum in the
// reference's order: corners (0,0), (0,1), (1,0), (1,1) in turn, quadrature
// index ascending within each corner, every term w * fx * fy rounded as the
// reference rounds it (no fused multiply-adds).
#pragma once

#include "hz_common.hpp"

namespace hz {

constexpr int DEPOSIT_CHUNK = 256;

// One workgroup per emitter m.  Thread t owns node t (then t + 256, ...) of
// the emitter's box: boxes[4m..4m+3] = (i0, j0, width, height), node
// (i0 + t % width, j0 + t / width), so the box index increases with the
// node number.  The quadrature points are staged through LDS 256 at a time:
// grid coordinates g = (centre + offset) / dx, their floor and fraction.
// mass[m * cap + t] receives the node's mass (0 where no point reaches it).
__global__ __launch_bounds__(256) void k_deposit(double dx, int q, const double *__restrict__ centers,
                                                 const double *__restrict__ offsets,
                                                 const double *__restrict__ weights, const int *__restrict__ boxes,
                                                 int cap, double *__restrict__ mass) {
#pragma clang fp contract(off)
    __shared__ int s_bx[DEPOSIT_CHUNK], s_by[DEPOSIT_CHUNK];
    __shared__ double s_tx[DEPOSIT_CHUNK], s_ty[DEPOSIT_CHUNK], s_w[DEPOSIT_CHUNK];
    const int m = blockIdx.x;
    const int i0 = boxes[4 * m], j0 = boxes[4 * m + 1], bw = boxes[4 * m + 2], bh = boxes[4 * m + 3];
    const int nbox = bw * bh;
    const double cx = centers[2 * m], cy = centers[2 * m + 1];
    for (int group = 0; group < nbox; group += DEPOSIT_CHUNK) {  // uniform across the workgroup
        const int t = group + (int)threadIdx.x;
        const bool mine = t < nbox;
        const int i = i0 + (mine ? t % bw : 0), j = j0 + (mine ? t / bw : 0);
        double acc = 0.0;
        for (int corner = 0; corner < 4; ++corner) {
            const int a = corner >> 1, b = corner & 1;
            for (int q0 = 0; q0 < q; q0 += DEPOSIT_CHUNK) {
                __syncthreads();  // the previous chunk has been read
                const int qq = q0 + (int)threadIdx.x;
                if (qq < q) {
                    const double gx = (cx + offsets[2 * qq]) / dx;
                    const double gy = (cy + offsets[2 * qq + 1]) / dx;
                    const double fx = floor(gx), fy = floor(gy);
                    s_bx[threadIdx.x] = (int)fx;
                    s_by[threadIdx.x] = (int)fy;
                    s_tx[threadIdx.x] = gx - fx;
                    s_ty[threadIdx.x] = gy - fy;
                    s_w[threadIdx.x] = weights[qq];
                }
                __syncthreads();
                const int cnt = min(DEPOSIT_CHUNK, q - q0);
                if (mine) {
                    for (int u = 0; u < cnt; ++u) {
                        if (s_bx[u] + a == i && s_by[u] + b == j) {
                            const double wx = a ? s_tx[u] : 1.0 - s_tx[u];
                            const double wy = b ? s_ty[u] : 1.0 - s_ty[u];
                            const double term = (s_w[u] * wx) * wy;
                            acc = acc + term;
                        }
                    }
                }
            }
        }
        if (mine) mass[(int64_t)m * cap + t] = acc;
    }
}

}  // namespace hz
