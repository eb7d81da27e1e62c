// Phase 2 for tensor-product hexahedra: element stiffness blocks by sum factorisation
// (included by fl_assemble.hip).  No reference counterpart in form; the result is the
// same K as _ConstitutiveStiffnessDF_.h:83-134 + _GeometricStiffness_.h:61-100.
//
// With the parent gradients of a Lagrange hex factorising per direction,
// dN_a/dxi_m = prod_d T_d^{[d == m]}[a_d][q_d] (T^0 = l, T^1 = l', d = zeta, eta, beta),
// every block is
//
//     K^{ik}_{ab} = sum_q sum_{m,n} dN_a/dxi_m (q) C^{ik}_{mn}(q) dN_b/dxi_n (q),
//     C^{ik}_{mn} = sum_{j,l} (dxi_m/dx_j) D^{ik}_{jl} (dxi_n/dx_l),
//
// with D the per-point constitutive + geometric tensor of the point record (detJ
// included).  The point sum is done one direction at a time:
//
//     T1[mn][az,bz][qe,qb]      = sum_qz Fz[az,bz][qz] C[mn][qz,qe,qb]
//     T2[type][ae,be][az,bz][qb] = sum_{mn in type} sum_qe Fe[ae,be][qe] T1[mn][az,bz][qe,qb]
//     K[ab,bb][ae,be][az,bz]    = sum_type sum_qb Fb[ab,bb][qb] T2[type][ae,be][az,bz][qb]
//
// where F_d[a,b][q] = T_d^{s}[a][q] T_d^{t}[b][q] with (s, t) = ([m = d], [n = d]) and
// "type" groups the nine (m, n) by their beta factor.  At p=4 with 6^3 points this is
// ~1.1 MFLOP per block instead of 20 MFLOP for the dense contraction.  The FP64 vector
// rate of MI355X equals its FP64 matrix rate, so the work runs on the VALU from LDS.
//
// One workgroup per (element, block i,k) -- all nine blocks, so every workgroup writes
// its own rows contiguously and no block is stored transposed.  The az slices are
// processed one after another to keep T1 and T2 small (two workgroups per CU).
//
// Tensor basis table (fl_tensor_basis layout, built once on the host):
//   int32  P1, Q, pad, pad
//   int32  g_of_t[Q^3]      flattened point index of tensor point (qb*Q + qe)*Q + qz
//   int32  local_of_t[P1^3] local node of tensor node (ab*P1 + ae)*P1 + az
//   (padded to 8 bytes)
//   double f1d[3][2][P1][Q] l and l' of direction d (zeta, eta, beta) at its points
#pragma once

namespace fl {

template <int P1, int Q>
struct SfLayout {
    static constexpr int Q3 = Q * Q * Q, N3 = P1 * P1 * P1, PP = P1 * P1;
    static constexpr int INTS = 4 + Q3 + N3 + ((Q3 + N3) % 2);   // 8-byte aligned
};

template <int P1, int Q>
__global__ __launch_bounds__(256, 2) void stiffness_sf_kernel(const double *__restrict__ qp,
                                                              const int32_t *__restrict__ basis, int64_t nelem,
                                                              double *__restrict__ Ke) {
    typedef SfLayout<P1, Q> L;
    constexpr int Q3 = L::Q3, N3 = L::N3, PP = L::PP;
    constexpr int NT = 256;
    // LDS: pair factors of zeta and eta for the four (s,t), C, one az slice of T1 and T2
    __shared__ double sFz[4][PP][Q], sFe[4][PP][Q], sFb[2][P1][Q];
    __shared__ double sC[9][Q3];
    __shared__ double sT1[9][P1][Q * Q];
    __shared__ double sT2[4][PP][P1][Q];
    __shared__ int32_t sLoc[N3];

    const int64_t wid = blockIdx.x;
    const int xcd = (int)(wid & 7);
    const int64_t slot = wid >> 3;
    const int blk = (int)(slot % 9);
    const int64_t e = (slot / 9) * 8 + xcd;
    if (e >= nelem) return;
    const int pi = blk / 3, pk = blk % 3;
    // stored pair (i <= k) and whether D must be read transposed (D^{ki}_{jl} = D^{ik}_{lj})
    const int sp = pi == pk ? pi : (pi + pk == 1 ? 3 : (pi + pk == 2 ? 4 : 5));
    const bool tr = pi > pk;
    const int tid = threadIdx.x;
    const int32_t *g_of_t = basis + 4;
    const int32_t *loc = basis + 4 + Q3;
    const double *f1d = reinterpret_cast<const double *>(basis + L::INTS);   // [3][2][P1][Q]

    for (int t = tid; t < 4 * PP * Q; t += NT) {
        const int st = t / (PP * Q), r = t - st * PP * Q;
        const int a = r / (P1 * Q), b = (r / Q) % P1, q = r % Q;
        const int s = st >> 1, u = st & 1;
        sFz[st][a * P1 + b][q] = f1d[((0 * 2 + s) * P1 + a) * Q + q] * f1d[((0 * 2 + u) * P1 + b) * Q + q];
        sFe[st][a * P1 + b][q] = f1d[((1 * 2 + s) * P1 + a) * Q + q] * f1d[((1 * 2 + u) * P1 + b) * Q + q];
    }
    for (int t = tid; t < 2 * P1 * Q; t += NT) sFb[t / (P1 * Q)][(t / Q) % P1][t % Q] = f1d[(4 * P1) * Q + t];
    // table entries are clamped into range: a table that does not match the launch
    // shape gives wrong blocks, never an access outside them
    for (int t = tid; t < N3; t += NT) sLoc[t] = min(max(loc[t], 0), N3 - 1);
    // C[mn][t] = sum_jl ix[j][m] D[j][l] ix[l][n] at tensor point t
    const double *qe_ = qp + e * (int64_t)(Q3)*QS;
    for (int t = tid; t < Q3; t += NT) {
        const double *rec = qe_ + (int64_t)min(max(g_of_t[t], 0), Q3 - 1) * QS;
        double ix[9], D[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) ix[i] = rec[Q_IX + i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int l = 0; l < 3; ++l) D[3 * j + l] = rec[Q_D + 9 * sp + (tr ? 3 * l + j : 3 * j + l)];
        double W[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int n = 0; n < 3; ++n)
                W[3 * j + n] = D[3 * j] * ix[n] + D[3 * j + 1] * ix[3 + n] + D[3 * j + 2] * ix[6 + n];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n)
                sC[3 * m + n][t] = ix[m] * W[n] + ix[3 + m] * W[3 + n] + ix[6 + m] * W[6 + n];
    }
    __syncthreads();

    double *Kd = Ke + (e * 9 + pi * 3 + pk) * (int64_t)N3 * N3;
    for (int az = 0; az < P1; ++az) {
        // T1[mn][bz][qe*Q + qb] = sum_qz Fz^{st(mn)}[az,bz][qz] C[mn][(qb*Q + qe)*Q + qz]
        for (int t = tid; t < 9 * P1 * Q * Q; t += NT) {
            const int mn = t / (P1 * Q * Q), r = t - mn * P1 * Q * Q;
            const int bz = r / (Q * Q), qe = (r / Q) % Q, qb = r % Q;
            const int m = mn / 3, n = mn % 3;
            const int st = ((m == 0) << 1) | (n == 0);
            const double *f = sFz[st][az * P1 + bz];
            const double *c = sC[mn] + (qb * Q + qe) * Q;
            double s = 0.0;
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) s += f[qz] * c[qz];
            sT1[mn][bz][qe * Q + qb] = s;
        }
        __syncthreads();
        // T2[type][ae,be][bz][qb] = sum_{mn in type} sum_qe Fe^{st(mn)}[ae,be][qe] T1[mn][bz][qe*Q + qb]
        for (int t = tid; t < 4 * PP * P1 * Q; t += NT) {
            const int type = t / (PP * P1 * Q), r = t - type * PP * P1 * Q;
            const int aebe = r / (P1 * Q), bz = (r / Q) % P1, qb = r % Q;
            const int sb = type >> 1, tb = type & 1;
            double s = 0.0;
            // (m, n) with [m == 2] = sb, [n == 2] = tb
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                if ((m == 2) != (sb == 1)) continue;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    if ((n == 2) != (tb == 1)) continue;
                    const int st = ((m == 1) << 1) | (n == 1);
                    const double *f = sFe[st][aebe];
                    const double *t1 = &sT1[3 * m + n][bz][qb];
#pragma unroll
                    for (int qe = 0; qe < Q; ++qe) s += f[qe] * t1[qe * Q];
                }
            }
            sT2[type][aebe][bz][qb] = s;
        }
        __syncthreads();
        // rows a = (ab, ae, az), ab = 0..P1-1 per task; columns b = (bb, be, bz)
        for (int task = tid; task < P1 * N3; task += NT) {
            const int ae = task / N3, tb = task - ae * N3;
            const int bb = tb / PP, be = (tb / P1) % P1, bz = tb % P1;
            double u[4][Q];
#pragma unroll
            for (int type = 0; type < 4; ++type)
#pragma unroll
                for (int qb = 0; qb < Q; ++qb)
                    u[type][qb] = sFb[type & 1][bb][qb] * sT2[type][ae * P1 + be][bz][qb];
            const int lb = sLoc[tb];
#pragma unroll
            for (int ab = 0; ab < P1; ++ab) {
                double s = 0.0;
#pragma unroll
                for (int type = 0; type < 4; ++type)
#pragma unroll
                    for (int qb = 0; qb < Q; ++qb) s += sFb[type >> 1][ab][qb] * u[type][qb];
                const int la = sLoc[(ab * P1 + ae) * P1 + az];
                Kd[(int64_t)la * N3 + lb] = s;
            }
        }
        __syncthreads();
    }
}

}  // namespace fl
