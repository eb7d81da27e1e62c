// Galerkin operators of the coarse levels of the p-multigrid preconditioner, formed
// from the point records the fine assembly leaves in its workspace (included by
// fl_assemble.hip, which owns the record layout).
//
// A coarse level is a nodal Lagrange space on the same elements (element vertices, or
// the p=2 lattice of a p=4 hex) whose basis functions lie in the fine space, so with
// the prolongation P (fine nodal values of a coarse function) the Galerkin operator
// P^T K P is the coarse space's stiffness integrated with the fine level's quadrature
// and per-point tangent:
//   Kc^{ik}_{ab} = sum_g sum_{m,n} dNc_a/dxi_m(g) C^{ik}_{mn}(g) dNc_b/dxi_n(g),
//   C^{ik}_{mn} = sum_{j,l} (dxi_m/dx_j) D^{ik}_{jl} (dxi_n/dx_l),
// D^{ik}_{jl} = detJ (c_ijkl + delta_ik sigma_jl) being what every fine stiffness kernel
// contracts.  The records hold C directly (the factorised FP64 path) or dxi/dx and D
// (the dense path); C^{ki}_{mn} = C^{ik}_{nm} gives the three blocks not stored.
#pragma once

namespace fl {

// Kc [nelem][9][NC][NC] (block 3i+k, row a, column b, coarse local order).
// JmCt [G][NC][3]: dNc_a/dxi_m at the fine points.  One workgroup per element; per
// point the record's C and the coarse gradients go to LDS, then
// W^{ik}_{a,n} = sum_m G_am C^{ik}_mn, then each thread adds sum_n W^{ik}_{a,n} G_bn
// to the nine blocks of its (a, b) pairs.
template <int NC>
__global__ __launch_bounds__(256) void mg_galerkin_kernel(const double *__restrict__ qp, int G, int form_c,
                                                          const double *__restrict__ JmCt, int64_t nelem,
                                                          double *__restrict__ Kc) {
    constexpr int NT = 256;
    constexpr int NP2 = NC * NC;
    constexpr int PER = (NP2 + NT - 1) / NT;
    __shared__ double sC[54];
    __shared__ double sG[NC * 3];
    __shared__ double sW[9 * NC * 3];
    const int64_t e = blockIdx.x;
    if (e >= nelem) return;
    const int tid = threadIdx.x;
    const double *rec = qp + e * (int64_t)QS * G;   // field f of point g at rec[f * G + g]
    double acc[PER][9];
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
        for (int ik = 0; ik < 9; ++ik) acc[q][ik] = 0.0;
    for (int g = 0; g < G; ++g) {
        if (tid < 54) {
            if (form_c) {
                sC[tid] = rec[(Q_D + tid) * G + g];
            } else {
                const int p = tid / 9, m = (tid % 9) / 3, n = tid % 3;
                double c = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double ixjm = rec[(Q_IX + 3 * j + m) * G + g];
                    double t = 0.0;
#pragma unroll
                    for (int l = 0; l < 3; ++l) t += rec[(Q_D + 9 * p + 3 * j + l) * G + g] * rec[(Q_IX + 3 * l + n) * G + g];
                    c += ixjm * t;
                }
                sC[tid] = c;
            }
        }
        for (int t = tid; t < NC * 3; t += NT) sG[t] = JmCt[(int64_t)g * NC * 3 + t];
        __syncthreads();
        for (int t = tid; t < 9 * NC * 3; t += NT) {
            const int ik = t / (NC * 3), a = (t / 3) % NC, n = t % 3;
            const int i = ik / 3, k = ik % 3;
            const int lo = i < k ? i : k, hi = i < k ? k : i;
            const int p = lo == hi ? lo : (lo == 0 ? (hi == 1 ? 3 : 4) : 5);
            double w = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) w += sG[3 * a + m] * (i <= k ? sC[9 * p + 3 * m + n] : sC[9 * p + 3 * n + m]);
            sW[t] = w;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int pr = tid + q * NT;
            if (pr < NP2) {
                const int a = pr / NC, b = pr % NC;
                const double g0 = sG[3 * b], g1 = sG[3 * b + 1], g2 = sG[3 * b + 2];
#pragma unroll
                for (int ik = 0; ik < 9; ++ik) {
                    const double *w = sW + (ik * NC + a) * 3;
                    acc[q][ik] += w[0] * g0 + w[1] * g1 + w[2] * g2;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int pr = tid + q * NT;
        if (pr < NP2) {
            const int a = pr / NC, b = pr % NC;
#pragma unroll
            for (int ik = 0; ik < 9; ++ik) Kc[((e * 9 + ik) * NC + a) * NC + b] = acc[q][ik];
        }
    }
}

}  // namespace fl
