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
// ~1.1 MFLOP per block instead of 20 MFLOP for the dense contraction.  Three forms:
// stiffness_sf_kernel (all on the FP64 VALU, whose rate on MI355X equals the FP64
// matrix rate), stiffness_sf_mfma_kernel (eta and beta on 16x16x4 FP64 MFMA, one type
// per reduction) and stiffness_sf_mfma2_kernel (the default: two types per reduction).
//
// One workgroup per (element, block i,k) -- all nine blocks, so every workgroup writes
// its own rows and no block is stored transposed.  Rows and columns of the stored
// blocks are in tensor node order (the gather maps local nodes through tensor_of_l),
// so the stores need no lookups.  The az slices are processed one after another to
// keep T1 and T2 small.
//
// Tensor basis table (fl_tensor_basis layout, built once on the host):
//   int32  P1, Q, pad, pad
//   int32  g_of_t[Q^3]      flattened point index of tensor point (qb*Q + qe)*Q + qz
//   int32  local_of_t[P1^3] local node of tensor node (ab*P1 + ae)*P1 + az
//   int32  tensor_of_l[P1^3] its inverse
//   (padded to 8 bytes)
//   double f1d[3][2][P1][Q] l and l' of direction d (zeta, eta, beta) at its points
#pragma once

namespace fl {

template <int P1, int Q>
struct SfLayout {
    static constexpr int Q3 = Q * Q * Q, N3 = P1 * P1 * P1, PP = P1 * P1;
    static constexpr int INTS = 4 + Q3 + 2 * N3 + (Q3 % 2);   // 8-byte aligned
};

// Element tractions t_(a,i) = sum_q sum_m dN_a/dxi_m (q) P[q][m][i] from the point
// tensors P[3 m + i][t] (tensor point t) in LDS: contracted zeta, eta, beta in turn for the
// nine (m, i), one workgroup per element.  U1 [9][P1][Q^2] and U2 [9][PP][Q] are LDS
// work areas; U2 may share P's space (P's last reads precede the barrier before U2 is
// written).  te in local node order; with T, a node of this element only gets its T
// entry directly (as 0.0 + value, the accumulated form).
template <int P1, int Q>
__device__ __forceinline__ void traction_sf_contract(const double (&sF)[3][2][P1][Q], const double *sP,
                                                     double *sU1, double *sU2, const int32_t *__restrict__ basis,
                                                     int64_t e, double *__restrict__ te,
                                                     const int32_t *__restrict__ elements,
                                                     const int32_t *__restrict__ node_elem_ptr,
                                                     double *__restrict__ T) {
    typedef SfLayout<P1, Q> L;
    constexpr int Q3 = L::Q3, N3 = L::N3, PP = L::PP;
    constexpr int NT = 256;
    const int tid = threadIdx.x;
    const int32_t *loc = basis + 4 + Q3;
    for (int t = tid; t < 9 * P1 * Q * Q; t += NT) {
        const int mi = t / (P1 * Q * Q), r = t - mi * P1 * Q * Q;
        const int az = r / (Q * Q), qe = (r / Q) % Q, qb = r % Q;
        const double *f = sF[0][mi / 3 == 0][az];
        const double *p = sP + mi * Q3 + (qb * Q + qe) * Q;
        double s = 0.0;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) s += f[qz] * p[qz];
        sU1[(mi * P1 + az) * Q * Q + qe * Q + qb] = s;
    }
    __syncthreads();
    for (int t = tid; t < 9 * PP * Q; t += NT) {
        const int mi = t / (PP * Q), r = t - mi * PP * Q;
        const int ae = r / (P1 * Q), az = (r / Q) % P1, qb = r % Q;
        const double *f = sF[1][mi / 3 == 1][ae];
        double s = 0.0;
#pragma unroll
        for (int qe = 0; qe < Q; ++qe) s += f[qe] * sU1[(mi * P1 + az) * Q * Q + qe * Q + qb];
        sU2[(mi * PP + ae * P1 + az) * Q + qb] = s;
    }
    __syncthreads();
    for (int t = tid; t < 3 * N3; t += NT) {
        const int tn = t / 3, i = t - 3 * tn;
        const int ab = tn / PP, aeaz = tn - ab * PP;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const double *f = sF[2][m == 2][ab];
#pragma unroll
            for (int qb = 0; qb < Q; ++qb) s += f[qb] * sU2[((3 * m + i) * PP + aeaz) * Q + qb];
        }
        const int lt = min(max(loc[tn], 0), N3 - 1);
        te[e * 3 * N3 + 3 * lt + i] = s;
        if (T) {
            const int32_t a = elements[e * N3 + lt];
            if (node_elem_ptr[a + 1] - node_elem_ptr[a] == 1) T[3 * (int64_t)a + i] = 0.0 + s;
        }
    }
}

// P[3 m + i] = sum_j (dxi_m/dx_j) (detJ sigma)_ij of one point from dxi/dx (9) and detJ
// sigma (s00 s11 s22 s01 s02 s12)
__device__ __forceinline__ void traction_point_tensor(const double *ix, const double *S, double (&P)[9]) {
    const double Si[3][3] = {{S[0], S[3], S[4]}, {S[3], S[1], S[5]}, {S[4], S[5], S[2]}};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) P[3 * m + i] = ix[m] * Si[i][0] + ix[3 + m] * Si[i][1] + ix[6 + m] * Si[i][2];
}

// Phase 1 for tensor-product hexes: one workgroup per element.  The element's X and x
// are staged in LDS in tensor node order and the parent Jacobians at all points,
// dX_c/dxi_m = sum_a dN_a/dxi_m X_a,c, are contracted one direction at a time:
//
//     A[f][s0][ab,ae][qz]   = sum_az T0^{s0}[az][qz] X_f[ab,ae,az]
//     B[f][m][ab][qe,qz]    = sum_ae T1^{[m=1]}[ae][qe] A[f][[m=0]][ab,ae][qz]
//     P[f][m][qb,qe,qz]     = sum_ab T2^{[m=2]}[ab][qb] B[f][m][ab][qe,qz]
//
// (f = X_0..2, x_0..2; X then x, one pass each), then each point runs qp_point.  Same
// records as qp_kernel, to rounding (the node sum is taken in another order).  LDS:
// the staging shares its space with B, A with P.  With te given, the element tractions
// follow in the same workgroup (traction_sf_contract: the point tensors in the P area,
// the zeta results in the B area, the eta results in the P area again), and the
// records -- if qp is given at all -- omit detJ sigma and detJ.
template <int P1, int Q, bool TANGENT, bool GEN>
__global__ __launch_bounds__(256) void qp_sf_kernel(MatParams mp0, const double *__restrict__ points,
                                                    const int32_t *__restrict__ elements,
                                                    const double *__restrict__ Eulerx,
                                                    const int32_t *__restrict__ basis, const double *__restrict__ w,
                                                    int64_t nelem, int update, double *__restrict__ qp,
                                                    int32_t *__restrict__ status, const int32_t *__restrict__ elem_mat,
                                                    const double *__restrict__ mat_table, int emitC,
                                                    double *__restrict__ te = nullptr,
                                                    const int32_t *__restrict__ node_elem_ptr = nullptr,
                                                    double *__restrict__ T = nullptr) {
    typedef SfLayout<P1, Q> L;
    constexpr int Q3 = L::Q3, N3 = L::N3, PP = L::PP;
    constexpr int NT = 256;
    static_assert(Q3 <= NT, "one point per thread");
    // two passes (X, then x) of three fields each keep the LDS small (~30 KB at p = 4)
    constexpr int NF = 3;
    constexpr int SZ_X = NF * N3, SZ_B = NF * 3 * P1 * Q * Q, SZ_A = NF * 2 * PP * Q, SZ_P = NF * 3 * Q3;
    __shared__ double sF[3][2][P1][Q];
    __shared__ double sXB[SZ_X > SZ_B ? SZ_X : SZ_B];
    __shared__ double sAP[SZ_A > SZ_P ? SZ_A : SZ_P];
    const int64_t e = blockIdx.x;
    if (e >= nelem) return;
    const int tid = threadIdx.x;
    const int32_t *g_of_t = basis + 4;
    const int32_t *loc = basis + 4 + Q3;
    const double *f1d = reinterpret_cast<const double *>(basis + L::INTS);
    const MatParams mp = element_material(mp0, e, elem_mat, mat_table);
    for (int t = tid; t < 3 * 2 * P1 * Q; t += NT) (&sF[0][0][0][0])[t] = f1d[t];
    const int32_t *conn = elements + e * (int64_t)N3;
    const bool valid = tid < Q3;
    const int tq = valid ? tid : 0;
    double PJ[2][9];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        // the previous pass's last reads of sXB and sAP precede the barriers below
        const double *src = pass ? Eulerx : points;
        for (int t = tid; t < N3; t += NT) {
            const int64_t nd = conn[min(max(loc[t], 0), N3 - 1)];
#pragma unroll
            for (int c = 0; c < 3; ++c) sXB[c * N3 + t] = src[3 * nd + c];
        }
        __syncthreads();
        // A[f][s0][(ab,ae)][qz]
        for (int t = tid; t < SZ_A; t += NT) {
            const int qz = t % Q, r = t / Q;
            const int abe = r % PP, r2 = r / PP;
            const int s0 = r2 % 2, f = r2 / 2;
            const double *xf = sXB + f * N3 + abe * P1;
            double acc = 0.0;
#pragma unroll
            for (int az = 0; az < P1; ++az) acc += sF[0][s0][az][qz] * xf[az];
            sAP[t] = acc;
        }
        __syncthreads();
        // B[f][m][ab][(qe,qz)]
        for (int t = tid; t < SZ_B; t += NT) {
            const int qez = t % (Q * Q), r = t / (Q * Q);
            const int ab = r % P1, r2 = r / P1;
            const int m = r2 % 3, f = r2 / 3;
            const int qe = qez / Q, qz = qez % Q;
            const double *a = sAP + ((f * 2 + (m == 0)) * PP + ab * P1) * Q + qz;
            double acc = 0.0;
#pragma unroll
            for (int ae = 0; ae < P1; ++ae) acc += sF[1][m == 1][ae][qe] * a[ae * Q];
            sXB[t] = acc;
        }
        __syncthreads();
        // P[f][m][(qb,qe,qz)]
        for (int t = tid; t < SZ_P; t += NT) {
            const int tqq = t % Q3, r = t / Q3;
            const int m = r % 3, f = r / 3;
            const int qb = tqq / (Q * Q), qez = tqq % (Q * Q);
            const double *bb = sXB + (f * 3 + m) * P1 * Q * Q + qez;
            double acc = 0.0;
#pragma unroll
            for (int ab = 0; ab < P1; ++ab) acc += sF[2][m == 2][ab][qb] * bb[ab * Q * Q];
            sAP[t] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int c = 0; c < 3; ++c) PJ[pass][3 * m + c] = sAP[(c * 3 + m) * Q3 + tq];
    }
    const int g = min(max(g_of_t[tq], 0), Q3 - 1);
    // te given: the element tractions are contracted here from the point values, and
    // the records (qp, when given) are only what the stiffness kernels read
    double pv[15];
    qp_point<TANGENT, GEN>(mp, PJ[0], PJ[1], w[g], update, valid,
                      qp ? qp + e * (int64_t)Q3 * (TANGENT ? QS : QS_RES) + g : nullptr, Q3, status, emitC != 0,
                      te ? pv : nullptr);
    if (te) {
        static_assert(9 * P1 * Q * Q <= (SZ_X > SZ_B ? SZ_X : SZ_B), "U1 in the B area");
        static_assert(9 * PP * Q <= (SZ_A > SZ_P ? SZ_A : SZ_P) && 9 * Q3 <= (SZ_A > SZ_P ? SZ_A : SZ_P),
                      "P and U2 in the P area");
        __syncthreads();   // the last reads of sAP (PJ) precede
        if (valid) {
            double P[9];
            traction_point_tensor(pv, pv + 9, P);
#pragma unroll
            for (int mi = 0; mi < 9; ++mi) sAP[mi * Q3 + tq] = P[mi];
        }
        __syncthreads();
        traction_sf_contract<P1, Q>(sF, sAP, sXB, sAP, basis, e, te, elements, node_elem_ptr, T);
    }
}

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
    // C[mn][t] = sum_jl ix[j][m] D[j][l] ix[l][n] at tensor point t, as the point kernel stored it
    const double *qe_ = qp + e * (int64_t)(Q3)*QS;
    for (int t = tid; t < Q3; t += NT) {
        const double *rec = qe_ + min(max(g_of_t[t], 0), Q3 - 1);   // field f at rec[f * Q3]
        // the record holds C^{ik} of the pair (qp_point with emitC); C^{ki} = (C^{ik})^T
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) sC[3 * m + n][t] = rec[(Q_D + 9 * sp + (tr ? 3 * n + m : 3 * m + n)) * Q3];
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
#pragma unroll
            for (int ab = 0; ab < P1; ++ab) {
                double s = 0.0;
#pragma unroll
                for (int type = 0; type < 4; ++type)
#pragma unroll
                    for (int qb = 0; qb < Q; ++qb) s += sFb[type >> 1][ab][qb] * u[type][qb];
                Kd[(int64_t)((ab * P1 + ae) * P1 + az) * N3 + tb] = s;
            }
        }
        __syncthreads();
    }
}

// The same factorisation with the eta and beta contractions on FP64 MFMA
// (v_mfma_f64_16x16x4f64); the zeta contraction and the point tensor stay on the VALU.
// Per zeta slice and per beta-type (the nine (m, n) grouped by their beta factor):
//   eta:  T2 = sum_{mn in type} Fe^{st}[(ae,be)][qe] x T1[mn][qe][(bz,qb)]
//         (P1^2 x Q x P1 Q, the inner index padded to a multiple of 4)
//   beta: K[(ab,bb)][(ae,be,bz)] += Fb^{sb}Fb^{tb}[(ab,bb)][qb] x T2[qb][(ae,be,bz)]
//         accumulated in the MFMA registers over the four types.
// Pair factors are formed while the operands are loaded, and only one type's T2 is held,
// so the workgroup needs ~37 KB of LDS and four of them share a CU.  Output tiles are
// dealt round-robin to the four waves.
template <int P1, int Q>
__global__ __launch_bounds__(256, 4) void stiffness_sf_mfma_kernel(const double *__restrict__ qp,
                                                                   const int32_t *__restrict__ basis, int64_t nelem,
                                                                   double *__restrict__ Ke) {
    typedef SfLayout<P1, Q> L;
    typedef Mfma16<double> MF;
    constexpr int Q3 = L::Q3, N3 = L::N3, PP = L::PP;
    constexpr int NT = 256;
    constexpr int MT2 = (PP + 15) / 16, NT2 = (P1 * Q + 15) / 16, KQ = (Q + 3) / 4;
    constexpr int MT3 = (PP + 15) / 16, NT3 = (N3 + 15) / 16;
    constexpr int TPW = (MT3 * NT3 + 3) / 4;   // beta tiles per wave
    __shared__ double sF[3][2][P1][Q];
    __shared__ double sC[9][Q3];
    __shared__ double sT1[9][P1][Q * Q];
    __shared__ double sT2[PP][P1][Q];

    const int64_t wid = blockIdx.x;
    const int xcd = (int)(wid & 7);
    const int64_t slot = wid >> 3;
    const int blk = (int)(slot % 9);
    const int64_t e = (slot / 9) * 8 + xcd;
    if (e >= nelem) return;
    const int pi = blk / 3, pk = blk % 3;
    const int sp = pi == pk ? pi : (pi + pk == 1 ? 3 : (pi + pk == 2 ? 4 : 5));
    const bool tr = pi > pk;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int32_t *g_of_t = basis + 4;
    const double *f1d = reinterpret_cast<const double *>(basis + L::INTS);

    for (int t = tid; t < 3 * 2 * P1 * Q; t += NT) (&sF[0][0][0][0])[t] = f1d[t];
    const double *qe_ = qp + e * (int64_t)(Q3)*QS;
    for (int t = tid; t < Q3; t += NT) {
        const double *rec = qe_ + min(max(g_of_t[t], 0), Q3 - 1);   // field f at rec[f * Q3]
        // the record holds C^{ik} of the pair (qp_point with emitC); C^{ki} = (C^{ik})^T
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) sC[3 * m + n][t] = rec[(Q_D + 9 * sp + (tr ? 3 * n + m : 3 * m + n)) * Q3];
    }
    __syncthreads();

    double *Kd = Ke + (e * 9 + pi * 3 + pk) * (int64_t)N3 * N3;
    for (int az = 0; az < P1; ++az) {
        // zeta (VALU): task (mn, bz, qe) forms its factor once and sweeps qb
        for (int t = tid; t < 9 * P1 * Q; t += NT) {
            const int mn = t / (P1 * Q), bz = (t / Q) % P1, qe = t % Q;
            const int m = mn / 3, n = mn % 3;
            double f[Q];
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) f[qz] = sF[0][m == 0][az][qz] * sF[0][n == 0][bz][qz];
#pragma unroll
            for (int qb = 0; qb < Q; ++qb) {
                const double *c = sC[mn] + (qb * Q + qe) * Q;
                double s = 0.0;
#pragma unroll
                for (int qz = 0; qz < Q; ++qz) s += f[qz] * c[qz];
                sT1[mn][bz][qe * Q + qb] = s;
            }
        }
        __syncthreads();
        typename MF::acc_t acc[TPW];
#pragma unroll
        for (int u = 0; u < TPW; ++u) acc[u] = typename MF::acc_t{0.0, 0.0, 0.0, 0.0};
        for (int type = 0; type < 4; ++type) {
            const int sb = type >> 1, tb = type & 1;
            // eta (MFMA) for this type: tiles over rows (ae,be), columns (bz,qb)
            for (int task = wave; task < MT2 * NT2; task += 4) {
                const int tm = task / NT2, tn = task - tm * NT2;
                const int arow = tm * 16 + li, bcol = tn * 16 + li;
                const int ae_ = arow / P1, be_ = arow - ae_ * P1;
                typename MF::acc_t a2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    if ((m == 2) != (sb == 1)) continue;
#pragma unroll
                    for (int n = 0; n < 3; ++n) {
                        if ((n == 2) != (tb == 1)) continue;
#pragma unroll
                        for (int ks = 0; ks < KQ; ++ks) {
                            const int k = ks * 4 + lk;   // qe
                            const double a = (arow < PP && k < Q) ? sF[1][m == 1][ae_][k] * sF[1][n == 1][be_][k] : 0.0;
                            const double b = (bcol < P1 * Q && k < Q) ? sT1[3 * m + n][bcol / Q][k * Q + bcol % Q] : 0.0;
                            a2 = MF::step(a, b, a2);
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = tm * 16 + MF::row(lk, r);
                    if (row < PP && bcol < P1 * Q) sT2[row][bcol / Q][bcol % Q] = a2[r];
                }
            }
            __syncthreads();
            // beta (MFMA), this type's share of the inner index (qb, padded to 4)
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                const int task = wave + 4 * u;
                if (task < MT3 * NT3) {
                    const int tm = task / NT3, tn = task - tm * NT3;
                    const int arow = tm * 16 + li, bcol = tn * 16 + li;
                    const int ab_ = arow / P1, bb_ = arow - ab_ * P1;
                    const int ae_ = bcol / PP, be_ = (bcol / P1) % P1, bz_ = bcol % P1;
#pragma unroll
                    for (int ks = 0; ks < KQ; ++ks) {
                        const int qb = ks * 4 + lk;
                        const double a = (arow < PP && qb < Q) ? sF[2][sb][ab_][qb] * sF[2][tb][bb_][qb] : 0.0;
                        const double b = (bcol < N3 && qb < Q) ? sT2[ae_ * P1 + be_][bz_][qb] : 0.0;
                        acc[u] = MF::step(a, b, acc[u]);
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int task = wave + 4 * u;
            if (task < MT3 * NT3) {
                const int tm = task / NT3, tn = task - tm * NT3;
                const int bcol = tn * 16 + li;
                if (bcol < N3) {
                    const int ae_ = bcol / PP, be_ = (bcol / P1) % P1, bz_ = bcol % P1;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = tm * 16 + MF::row(lk, r);
                        if (row < PP) {
                            const int ab = row / P1, bb = row - ab * P1;
                            const int ta = (ab * P1 + ae_) * P1 + az, tb = (bb * P1 + be_) * P1 + bz_;
                            Kd[(int64_t)ta * N3 + tb] = acc[u][r];
                        }
                    }
                }
            }
        }
    }
}

// Per element and tensor node t: the CSR offsets of rows 3a, 3a+1, 3a+2 when node a lies
// in this element only (else -1) and the position of a among the element's nodes sorted
// by id -- the column block of t in every such row, whose neighbours are exactly the
// element's nodes.  irow[(e N3 + t) 4 + {0,1,2,3}].  One thread per (element, node).
template <int P1, int Q>
__global__ __launch_bounds__(256) void interior_rows_kernel(const int32_t *__restrict__ basis,
                                                            const int32_t *__restrict__ elements, int64_t nelem,
                                                            const int32_t *__restrict__ node_elem_ptr,
                                                            const int32_t *__restrict__ indptr,
                                                            int32_t *__restrict__ irow) {
    typedef SfLayout<P1, Q> L;
    constexpr int Q3 = L::Q3, N3 = L::N3;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nelem * N3) return;
    const int64_t e = g / N3;
    const int t = (int)(g - e * N3);
    const int32_t *loc = basis + 4 + Q3;
    const int32_t *conn = elements + e * N3;
    const int32_t a = conn[min(max(loc[t], 0), N3 - 1)];
    int r = 0;
    for (int u = 0; u < N3; ++u) r += conn[u] < a;
    const bool interior = node_elem_ptr[a + 1] - node_elem_ptr[a] == 1;
    int32_t *o = irow + g * 4;
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = interior ? indptr[3 * (int64_t)a + i] : -1;
    o[3] = r;
}

// The same contraction with the inner indices of two types merged into one MFMA
// reduction: the two types sharing the beta factor of m (a "phase") are contracted
// together, K = (type, qb) of length 2Q for beta and K = (m, n, qe) of length
// pairs x Q for eta, so Q = 6 fills K exactly (3 steps of 4 instead of 2 x 2 padded
// ones) and an az slice needs four barriers instead of nine.  Each wave owns every
// row tile of its column tiles, so an A fragment serves TNW tiles and a B fragment
// MT3.  The products of two 1D basis values that form the zeta factors and the eta A
// fragments are tabulated in LDS once per workgroup, and the beta A fragments, which
// depend only on the lane, stay in registers: LDS traffic, not MFMA issue, bounds this
// kernel.  With V given, rows of element-interior nodes go straight into the CSR
// instead of the block buffer (the gather then skips those nodes).  LDS: T2 of the
// phase's two types and the tables (~51 KB at p = 4, three workgroups per CU).
//
// DIRECT: every entry goes straight into the CSR and no block is stored.  The launch
// covers the elements elist[0, nelem) of one colour (no two of them share a node, so
// no two workgroups of the launch write one entry); scat (sf_scatter_index's order:
// the order this kernel stores in) holds each entry's column block in its row, with
// the sign bit set where an element of an earlier colour holds both nodes: that entry
// is added to, the others are stored.  Launched colour by colour on one stream, each
// CSR entry is the sum of its elements' contributions in colour order.  A slice's
// entries are fetched with one vector load per column tile when the slice starts, and
// the loads of the entries added to precede the slice's stores: stores count in the
// same memory counter as loads, so a load issued after a store waits for it.
template <int P1, int Q, bool DIRECT = false, typename CT = double, typename VT = double>
__global__ __launch_bounds__(256, sizeof(CT) == 8 ? 3 : 4) void stiffness_sf_mfma2_kernel(
    const double *__restrict__ qp, const int32_t *__restrict__ basis, int64_t nelem, CT *__restrict__ Ke,
    const int32_t *__restrict__ irow, VT *__restrict__ V, const int32_t *__restrict__ elist = nullptr,
    const int32_t *__restrict__ elements = nullptr, const int32_t *__restrict__ indptr = nullptr,
    const int16_t *__restrict__ scat = nullptr) {
    typedef SfLayout<P1, Q> L;
    typedef Mfma16<CT> MF;
    // output rows of the beta product (and so the store and scatter-table order): the FP64
    // MFMA's, rows lk + 4 r.  CT = float (the FP32 tangent: FP32 MFMA and accumulation)
    // feeds the A rows of each tile permuted, i -> (i >> 2) + 4 (i & 3), so that the FP32
    // MFMA's rows 4 lk + r come out as the same rows lk + 4 r.
    typedef Mfma16<double> MO;
    constexpr bool F64 = sizeof(CT) == 8;
    static_assert(!F64 || sizeof(VT) == 8, "FP64 blocks go to an FP64 CSR");
    constexpr int Q3 = L::Q3, N3 = L::N3, PP = L::PP;
    constexpr int NT = 256;
    constexpr int MT2 = (PP + 15) / 16, NT2 = (P1 * Q + 15) / 16;
    constexpr int MT3 = (PP + 15) / 16, NT3 = (N3 + 15) / 16, TNW = (NT3 + 3) / 4;
    constexpr int KE_MAX = (4 * Q + 3) / 4, KB = (2 * Q + 3) / 4;
    // pair products of the 1D bases: zeta [s*2+t][az][bz][qz] = T0^s[az][qz] T0^t[bz][qz],
    // eta [s*2+t][ae*P1+be][qe] likewise; the beta ones live in registers
    __shared__ CT sPz[4][P1][P1][Q];
    __shared__ CT sPe[4][PP][Q];
    __shared__ CT sC[9][Q3];
    __shared__ CT sT1[9][P1][Q * Q];
    __shared__ CT sT2[2][PP][P1][Q];
    // interior rows written straight into the CSR (V != NULL): CSR offset of row 3a+pi of
    // tensor node t when node a lies in this element only (else -1), and the position
    // of tensor node t among the element's nodes sorted by id -- the column block of
    // every such row, whose neighbours are exactly the element's nodes
    __shared__ int32_t sRow[N3];
    __shared__ int32_t sRank[N3];

    const int64_t wid = blockIdx.x;
    const int xcd = (int)(wid & 7);
    const int64_t slot = wid >> 3;
    const int blk = (int)(slot % 9);
    const int64_t idx = (slot / 9) * 8 + xcd;
    if (idx >= nelem) return;
    const int64_t e = DIRECT ? (int64_t)elist[idx] : idx;
    const int pi = blk / 3, pk = blk % 3;
    const int sp = pi == pk ? pi : (pi + pk == 1 ? 3 : (pi + pk == 2 ? 4 : 5));
    const bool tr = pi > pk;
    const int tid = threadIdx.x;
    // wave index read through a scalar register, so tests on it are uniform branches
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int32_t *g_of_t = basis + 4;
    const double *f1d = reinterpret_cast<const double *>(basis + L::INTS);

    if (DIRECT) {
        // CSR offset of row 3a+pi of every tensor node t
        const int32_t *loc = basis + 4 + Q3;
        const int32_t *conn = elements + e * N3;
        for (int t = tid; t < N3; t += NT) sRow[t] = indptr[3 * (int64_t)conn[min(max(loc[t], 0), N3 - 1)] + pi];
    } else if (V) {
        // interior_rows_kernel's table: per tensor node three row offsets and the rank
        const int32_t *ir = irow + e * (int64_t)N3 * 4;
        for (int t = tid; t < N3; t += NT) {
            sRow[t] = ir[4 * t + pi];
            sRank[t] = min(max(ir[4 * t + 3], 0), N3 - 1);
        }
    }
    for (int t = tid; t < 4 * PP * Q; t += NT) {
        const int st = t / (PP * Q), r = t - st * (PP * Q);
        const int a = r / (P1 * Q), b = (r / Q) % P1, q = r % Q;
        const int s0 = st >> 1, t0 = st & 1;
        (&sPz[0][0][0][0])[t] = (CT)(f1d[(s0 * P1 + a) * Q + q] * f1d[(t0 * P1 + b) * Q + q]);
        (&sPe[0][0][0])[t] = (CT)(f1d[((2 + s0) * P1 + a) * Q + q] * f1d[((2 + t0) * P1 + b) * Q + q]);
    }
    // this lane's beta A fragments: phase h, step ks, row tile tm -> T2^h[ab][qb] T2^tl[bb][qb]
    CT afr[2][KB][MT3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int ks = 0; ks < KB; ++ks) {
            const int kk = ks * 4 + lk;
            const bool kv = kk < 2 * Q;
            const int tl = kv ? kk / Q : 0, qb = kv ? kk - (kk / Q) * Q : 0;
#pragma unroll
            for (int tm = 0; tm < MT3; ++tm) {
                const int arow = tm * 16 + (F64 ? li : (li >> 2) + 4 * (li & 3));
                const int ab_ = min(arow / P1, P1 - 1), bb_ = arow % P1;
                afr[h][ks][tm] = (kv && arow < PP)
                                     ? (CT)(f1d[((4 + h) * P1 + ab_) * Q + qb] * f1d[((4 + tl) * P1 + bb_) * Q + qb])
                                     : CT(0);
            }
        }
    const double *qe_ = qp + e * (int64_t)(Q3)*QS;
    for (int t = tid; t < Q3; t += NT) {
        const double *rec = qe_ + min(max(g_of_t[t], 0), Q3 - 1);   // field f at rec[f * Q3]
        // the record holds C^{ik} of the pair (qp_point with emitC); C^{ki} = (C^{ik})^T
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) sC[3 * m + n][t] = (CT)rec[(Q_D + 9 * sp + (tr ? 3 * n + m : 3 * m + n)) * Q3];
    }
    __syncthreads();

    CT *Kd = DIRECT ? nullptr : Ke + (e * 9 + pi * 3 + pk) * (int64_t)N3 * N3;
    typedef short TV __attribute__((ext_vector_type(MT3 * 4)));
    for (int az = 0; az < P1; ++az) {
        // this slice's scatter entries: lane's rows (tm, r) of column tile wave + 4 v
        TV tv[TNW];
        if (DIRECT) {
#pragma unroll
            for (int v = 0; v < TNW; ++v)
                tv[v] = wave + 4 * v < NT3
                            ? *reinterpret_cast<const TV *>(scat + sf_scatter_index(P1, e, az, wave + 4 * v, lane, 0, 0))
                            : TV{};
        }
        // zeta (VALU): task (mn, bz, qe) forms its factor once and sweeps qb
        for (int t = tid; t < 9 * P1 * Q; t += NT) {
            const int mn = t / (P1 * Q), bz = (t / Q) % P1, qe = t % Q;
            const int m = mn / 3, n = mn % 3;
            const CT *f = sPz[(m == 0) * 2 + (n == 0)][az][bz];
#pragma unroll
            for (int qb = 0; qb < Q; ++qb) {
                const CT *c = sC[mn] + (qb * Q + qe) * Q;
                CT s = CT(0);
#pragma unroll
                for (int qz = 0; qz < Q; ++qz) s += f[qz] * c[qz];
                sT1[mn][bz][qe * Q + qb] = s;
            }
        }
        __syncthreads();
        typename MF::acc_t acc[MT3][TNW];
#pragma unroll
        for (int tm = 0; tm < MT3; ++tm)
#pragma unroll
            for (int v = 0; v < TNW; ++v) acc[tm][v] = typename MF::acc_t{CT(0), CT(0), CT(0), CT(0)};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // eta (MFMA) for the phase's types (m's beta factor h, n's tl): rows (ae,be),
            // columns (bz,qb), K = (m, n, qe) over the type's (m, n).  A wave takes one tile
            // for both types, whose two accumulation chains interleave.
            for (int tile = wave; tile < MT2 * NT2; tile += 4) {
                const int tm = tile / NT2, tn = tile - tm * NT2;
                const int arow = tm * 16 + li, bcol = tn * 16 + li;
                const bool av = arow < PP, bv = bcol < P1 * Q;
                const int arc = av ? arow : 0;
                const int bz_ = min(bcol / Q, P1 - 1), qb_ = bcol % Q;
                typename MF::acc_t a2[2] = {{CT(0), CT(0), CT(0), CT(0)}, {CT(0), CT(0), CT(0), CT(0)}};
#pragma unroll
                for (int ks = 0; ks < KE_MAX; ++ks) {
                    CT a[2], b[2];
#pragma unroll
                    for (int tl = 0; tl < 2; ++tl) {
                        const int nn = tl ? 1 : 2, ke = (h ? 1 : 2) * nn * Q;
                        a[tl] = CT(0);
                        b[tl] = CT(0);
                        if (ks * 4 < ke) {
                            // clamped loads, zeroed by a select where K or the tile pads
                            const int kk = ks * 4 + lk;
                            const bool kv = ks * 4 + 3 < ke || kk < ke;
                            const int kc = min(kk, ke - 1);
                            const int pr = kc / Q, qe = kc - pr * Q;
                            const int m = h ? 2 : pr / nn, n = tl ? 2 : pr % nn;
                            const CT x = sPe[(m == 1) * 2 + (n == 1)][arc][qe];
                            const CT y = sT1[3 * m + n][bz_][qe * Q + qb_];
                            a[tl] = (kv && av) ? x : CT(0);
                            b[tl] = (kv && bv) ? y : CT(0);
                        }
                    }
#pragma unroll
                    for (int tl = 0; tl < 2; ++tl) {
                        const int ke = (h ? 1 : 2) * (tl ? 1 : 2) * Q;
                        if (ks * 4 < ke) a2[tl] = MF::step(a[tl], b[tl], a2[tl]);
                    }
                }
#pragma unroll
                for (int tl = 0; tl < 2; ++tl)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = tm * 16 + MF::row(lk, r);
                        if (row < PP && bv) sT2[tl][row][bz_][qb_] = a2[tl][r];
                    }
            }
            __syncthreads();
            // beta (MFMA): rows (ab,bb), columns (ae,be,bz), K = (tl, qb)
#pragma unroll
            for (int ks = 0; ks < KB; ++ks) {
                // loads from clamped addresses, zeroed by a select where K or the column
                // pads (no branches around them)
                const int kk = ks * 4 + lk;
                const bool kv = ks * 4 + 3 < 2 * Q || kk < 2 * Q;
                const int kc = min(kk, 2 * Q - 1);
                const int tl = kc / Q, qb = kc - tl * Q;
                CT bf[TNW];
#pragma unroll
                for (int v = 0; v < TNW; ++v) {
                    const int bcol = (wave + 4 * v) * 16 + li;
                    const int bc = min(bcol, N3 - 1);
                    const int ae_ = bc / PP, be_ = (bc / P1) % P1, bz_ = bc % P1;
                    const CT x = sT2[tl][ae_ * P1 + be_][bz_][qb];
                    bf[v] = (kv && bcol < N3) ? x : CT(0);
                }
#pragma unroll
                for (int tm = 0; tm < MT3; ++tm)
#pragma unroll
                    for (int v = 0; v < TNW; ++v)
                        if (wave + 4 * v < NT3) acc[tm][v] = MF::step(afr[h][ks][tm], bf[v], acc[tm][v]);
            }
            // the next phase's eta rewrites sT2; the next slice's zeta only rewrites sT1,
            // whose last reads precede the barrier after this phase's eta
            if (h == 0) __syncthreads();
        }
        if constexpr (DIRECT) {
            // CSR positions and, where an earlier colour has written, the values to add
            // to: every load before the first store
            int32_t off[TNW][MT3][4];
            VT prev[TNW][MT3][4];
#pragma unroll
            for (int v = 0; v < TNW; ++v) {
                const int bcol = (wave + 4 * v) * 16 + li;
                const int ae_ = min(bcol, N3 - 1) / PP;
#pragma unroll
                for (int tm = 0; tm < MT3; ++tm)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = min(tm * 16 + MO::row(lk, r), PP - 1);
                        const int ta = ((row / P1) * P1 + ae_) * P1 + az;
                        const int sv = tv[v][tm * 4 + r];
                        off[v][tm][r] = sRow[ta] + 3 * (sv & 0x7fff) + pk;
                        const bool live = wave + 4 * v < NT3 && bcol < N3 && tm * 16 + MO::row(lk, r) < PP;
                        prev[v][tm][r] = live && sv < 0 ? V[off[v][tm][r]] : VT(0);
                    }
            }
            // vmcnt(0) on every path (expcnt and lgkmcnt left open): no load of this phase
            // is still counted when the next slice starts, so the waits the compiler places
            // there for its registers do not also wait for these stores
            __builtin_amdgcn_s_waitcnt(0x0F70);
            // the first contribution is stored as 0.0 + value, the sum the gather starts
            // from; later colours add to it
#pragma unroll
            for (int v = 0; v < TNW; ++v) {
                const int bcol = (wave + 4 * v) * 16 + li;
                if (wave + 4 * v < NT3 && bcol < N3) {
#pragma unroll
                    for (int tm = 0; tm < MT3; ++tm)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int rlo = tm * 16 + 4 * r;
                            const int row = tm * 16 + MO::row(lk, r);
                            if (rlo < PP && (rlo + 3 < PP || row < PP))
                                V[off[v][tm][r]] = (VT)(prev[v][tm][r] + acc[tm][v][r]);
                        }
                }
            }
            continue;
        }
#pragma unroll
        for (int v = 0; v < TNW; ++v) {
            const int bcol = (wave + 4 * v) * 16 + li;
            if (wave + 4 * v < NT3 && bcol < N3) {
                const int ae_ = bcol / PP, be_ = (bcol / P1) % P1, bz_ = bcol % P1;
#pragma unroll
                for (int tm = 0; tm < MT3; ++tm)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        // rows tm 16 + lk + 4 r, lk in [0, 4): whole groups of four are kept
                        // or dropped at compile time
                        const int rlo = tm * 16 + 4 * r;
                        const int row = tm * 16 + MO::row(lk, r);
                        if (rlo < PP && (rlo + 3 < PP || row < PP)) {
                            const int ab = row / P1, bb = row - ab * P1;
                            const int ta = (ab * P1 + ae_) * P1 + az, tb = (bb * P1 + be_) * P1 + bz_;
                            // one store through a selected address (the 0.0 + value form is
                            // what the gather would have written for an interior row, and
                            // leaves a block entry's sums unchanged)
                            CT *dst = Kd + ((int64_t)ta * N3 + tb);
                            if constexpr (F64) {
                                if (V) {
                                    const int32_t ro = sRow[ta];
                                    const int32_t rk = sRank[tb];
                                    dst = ro >= 0 ? V + ((int64_t)ro + 3 * rk + pk) : dst;
                                }
                            }
                            *dst = CT(0) + acc[tm][v][r];
                        }
                    }
            }
        }
    }
}

}  // namespace fl
