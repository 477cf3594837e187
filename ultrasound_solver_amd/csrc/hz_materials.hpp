// Scene rasterisation on the device (reference:
// src/algorithm/physics/warp_utils.py:11-23 blend, :52-77 sphere SDF and
// smoothing).  Per node: the signed distance to a union of primitives
// (spheres, boxes, first-order ellipsoids), the smoothed Heaviside
// h = (1 + t + sin(pi t)/pi)/2 with t = clip(sdf / width, -1, 1) (a sharp step
// for width 0), and the reference's blend, linear in h for 1/rho and
// 1/(rho c^2), into a constant or per-node medium.  Operations are evaluated
// in the reference's order without fused multiply-adds, so that only sin
// (to its last bit) can differ from the host evaluation.
#pragma once

#include "hz_common.hpp"

namespace hz {

enum PrimitiveKind { PRIM_SPHERE = 0, PRIM_BOX = 1, PRIM_ELLIPSOID = 2 };

struct Primitive {
    int kind;
    double c[3];  // centre (m)
    double s[3];  // radius (s[0]) / half extents / semi-axes (m)
};

constexpr int MAX_PRIMITIVES = 16;

struct MaterialArgs {
    int nx, ny, nz;
    double dx, width;
    int nprim;
    Primitive prim[MAX_PRIMITIVES];
    double rho_obs, c_obs, rho_med, c_med;
    const double *rho_field, *c_field;  // per-node medium, or null for the constants
    double *rho, *c, *sdf;              // outputs; any may be null
};

__device__ __forceinline__ double primitive_sdf(const Primitive &p, double x, double y, double z) {
#pragma clang fp contract(off)
    const double d0 = x - p.c[0], d1 = y - p.c[1], d2 = z - p.c[2];
    if (p.kind == PRIM_SPHERE) return sqrt((d0 * d0 + d1 * d1) + d2 * d2) - p.s[0];
    if (p.kind == PRIM_BOX) {
        const double q0 = fabs(d0) - p.s[0], q1 = fabs(d1) - p.s[1], q2 = fabs(d2) - p.s[2];
        const double m0 = fmax(q0, 0.0), m1 = fmax(q1, 0.0), m2 = fmax(q2, 0.0);
        const double outside = sqrt((m0 * m0 + m1 * m1) + m2 * m2);
        const double inside = fmin(fmax(fmax(q0, q1), q2), 0.0);
        return outside + inside;
    }
    // First-order ellipsoid distance k0 (k0 - 1) / k1, k0 = |x / r|, k1 = |x / r^2|.
    const double a0 = d0 / p.s[0], a1 = d1 / p.s[1], a2 = d2 / p.s[2];
    const double b0 = d0 / (p.s[0] * p.s[0]), b1 = d1 / (p.s[1] * p.s[1]), b2 = d2 / (p.s[2] * p.s[2]);
    const double k0 = sqrt((a0 * a0 + a1 * a1) + a2 * a2);
    const double k1 = sqrt((b0 * b0 + b1 * b1) + b2 * b2);
    if (k1 == 0.0) return -fmin(fmin(p.s[0], p.s[1]), p.s[2]);
    return k0 * (k0 - 1.0) / k1;
}

__global__ __launch_bounds__(256) void k_materials(MaterialArgs a) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(p % a.nx);
        const int j = (int)((p / a.nx) % a.ny);
        const int k = (int)(p / ((int64_t)a.nx * a.ny));
        const double x = i * a.dx, y = j * a.dx, z = k * a.dx;
        double sdf = INFINITY;
        for (int q = 0; q < a.nprim; ++q) sdf = fmin(sdf, primitive_sdf(a.prim[q], x, y, z));
        if (a.sdf) a.sdf[p] = sdf;
        if (!a.rho && !a.c) continue;
        double h;
        if (a.width != 0.0) {
            const double t = fmin(fmax(sdf / a.width, -1.0), 1.0);
            h = 0.5 * ((1.0 + t) + sin(M_PI * t) / M_PI);
        } else {
            h = sdf > 0.0 ? 1.0 : 0.0;
        }
        const double rm = a.rho_field ? a.rho_field[p] : a.rho_med;
        const double cm = a.c_field ? a.c_field[p] : a.c_med;
        if (a.nprim == 0) {  // no obstacle: the medium itself
            if (a.rho) a.rho[p] = rm;
            if (a.c) a.c[p] = cm;
            continue;
        }
        const double inv_obs = 1.0 / a.rho_obs;
        const double inv_rho = inv_obs + (1.0 / rm - inv_obs) * h;
        const double k_obs = 1.0 / (a.rho_obs * (a.c_obs * a.c_obs));
        const double kappa = k_obs + (1.0 / (rm * (cm * cm)) - k_obs) * h;
        if (a.rho) a.rho[p] = 1.0 / inv_rho;
        if (a.c) a.c[p] = sqrt(inv_rho / kappa);
    }
}

// One pass of a periodic ("wrap") correlation of a Fortran-order FP64 grid
// along `axis` with 2r+1 symmetric weights w[0..2r] (centre w[r]):
// out = w[r] x_i + sum_{j=r..1} w[r-j] (x_{i-j} + x_{i+j}), indices modulo n
// (the Gaussian filter of a random medium, applied axis by axis).
__global__ __launch_bounds__(256) void k_wrap_filter(int nx, int ny, int nz, int axis, int r,
                                                     const double *__restrict__ w, const double *__restrict__ in,
                                                     double *__restrict__ out) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)nx * ny * nz;
    const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)nx : (int64_t)nx * ny);
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)((axis == 0 ? p : (axis == 1 ? p / nx : p / ((int64_t)nx * ny))) % len);
        const int64_t base = p - (int64_t)i * stride;
        double acc = in[p] * w[r];
        for (int j = r; j >= 1; --j) {
            int lo = (i - j) % len, hi = (i + j) % len;
            if (lo < 0) lo += len;
            acc = acc + (in[base + lo * stride] + in[base + hi * stride]) * w[r - j];
        }
        out[p] = acc;
    }
}

}  // namespace hz
