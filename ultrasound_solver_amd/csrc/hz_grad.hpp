// Material gradient of the discrete operator (conservative_helmholtz.py:151-175):
// for a loss with dL = -Re(lambda^H dA u), dL/drho and dL/dc at every node,
// with lambda zeroed on identity rows.  A's dependence on the materials:
//   mass     D_P += MW_P omega^2 kappa_P, kappa = 1/(rho c^2):
//            d/drho = -mass/rho, d/dc = -2 mass/c
//   edges    A_PQ = F_PQ / (rho_P + rho_Q) with F = a_b a_c q_a (both rows):
//            d/drho_P = d/drho_Q = -F / (rho_P + rho_Q)^2
//   open     D_P += i omega area / (rho c): d/drho = -term/rho, d/dc = -term/c
// Impedance terms do not depend on the materials.
#pragma once

#include "hz_stencil.hpp"

namespace hz {

// One thread per node, summing its terms over k (u, lambda) column pairs;
// every edge is evaluated from both ends, so no atomics are needed.
// open_mask: bit f set when face f is an open (radiating) face.
__global__ __launch_bounds__(256) void k_material_vjp(OpView<double> op, int open_mask,
                                                      const cx<double> *__restrict__ u,
                                                      const cx<double> *__restrict__ lam, int64_t ld, int k,
                                                      double *__restrict__ g_rho, double *__restrict__ g_c) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < op.n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(p % op.nx);
        const int j = (int)((p / op.nx) % op.ny);
        const int kk = (int)(p / ((int64_t)op.nx * op.ny));
        const Nbr nb = neighbours(i, j, kk, op.nx, op.ny, op.nz);
        const double rp = op.rho[p];
        const double cp = 1.0 / sqrt(rp * op.kappa[p]);
        const cx<double> ax = op.a[0][i], ay = op.a[1][j], az = op.a[2][kk];
        const cx<double> mass = (op.omega2 * op.kappa[p]) * (ax * ay * az);
        // d A_PQ / d rho (either end) per neighbour, zero where absent.
        const cx<double> qf[6] = {op.q[0][i > 0 ? i - 1 : 0], op.q[0][i < op.nx - 1 ? i : op.nx - 2],
                                  op.q[1][j > 0 ? j - 1 : 0], op.q[1][j < op.ny - 1 ? j : op.ny - 2],
                                  op.q[2][kk > 0 ? kk - 1 : 0], op.q[2][kk < op.nz - 1 ? kk : op.nz - 2]};
        const cx<double> ab[3] = {ay * az, ax * az, ax * ay};
        const int di[6] = {-1, 1, 0, 0, 0, 0};
        const int dj[6] = {0, 0, -1, 1, 0, 0};
        const int dk[6] = {0, 0, 0, 0, -1, 1};
        cx<double> dcoef[6];
        bool qfixed[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const double s = rp + op.rho[p + nb.off[e]];
            dcoef[e] = nb.has[e] ? (-1.0 / (s * s)) * (ab[e >> 1] * qf[e]) : mk<double>(0, 0);
            qfixed[e] = nb.has[e] && node_fixed(op, i + di[e], j + dj[e], kk + dk[e]);
        }
        // Open-face terms at this node.
        cx<double> face = mk<double>(0, 0);
        if ((open_mask & 1) && i == 0) face = face + op.face[0][j + (int64_t)op.ny * kk];
        if ((open_mask & 2) && i == op.nx - 1) face = face + op.face[1][j + (int64_t)op.ny * kk];
        if ((open_mask & 4) && j == 0) face = face + op.face[2][i + (int64_t)op.nx * kk];
        if ((open_mask & 8) && j == op.ny - 1) face = face + op.face[3][i + (int64_t)op.nx * kk];
        if ((open_mask & 16) && kk == 0) face = face + op.face[4][i + (int64_t)op.nx * j];
        if ((open_mask & 32) && kk == op.nz - 1) face = face + op.face[5][i + (int64_t)op.nx * j];
        const bool pfixed = node_fixed(op, i, j, kk);
        double gr = 0, gc = 0;
        for (int r = 0; r < k; ++r) {
            const cx<double> up = u[r * ld + p];
            const cx<double> lp = pfixed ? mk<double>(0, 0) : lam[r * ld + p];
            const double pm = (conj(lp) * up * mass).re;
            gr += pm / rp;
            gc += 2.0 * pm / cp;
#pragma unroll
            for (int e = 0; e < 6; ++e) {
                if (!nb.has[e]) continue;
                const cx<double> uq = u[r * ld + p + nb.off[e]];
                const cx<double> lq = qfixed[e] ? mk<double>(0, 0) : lam[r * ld + p + nb.off[e]];
                gr -= ((conj(lp) - conj(lq)) * (uq - up) * dcoef[e]).re;
            }
            const double pf = (conj(lp) * up * face).re;
            gr += pf / rp;
            gc += pf / cp;
        }
        g_rho[p] = gr;
        g_c[p] = gc;
    }
}

}  // namespace hz
