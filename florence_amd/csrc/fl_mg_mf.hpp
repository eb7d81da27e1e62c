// The assembled tangent applied element by element from the point records, for the
// multigrid's fine-level smoother (included by fl_assemble.hip, which owns the record
// layout; the per-node sums are in fl_mg.hip).
//
// For a tensor-product hex with the factorised path's records (the C tensors, detJ and
// the point weight included),
//     y_{a,i} = sum_q sum_m dN_a/dxi_m(q) S_{im}(q),
//     S_{im}  = sum_{k,n} C^{ik}_{mn}(q) G_{kn}(q),   G_{kn} = sum_b dN_b/dxi_n(q) x_{b,k},
// with dN_b/dxi_n = prod_d T_d^{[d == n]}[b_d][q_d] (T^0 = l, T^1 = l'; d = zeta, eta,
// beta; tensor node (ab P1 + ae) P1 + az, tensor point (qb Q + qe) Q + qz).  The gradients
// are three one-direction contractions,
//   U1[k][sb][qb][ae,az]  = sum_ab T_b^sb[ab][qb] x_k[ab,ae,az]                    sb = 0, 1
//   U2[k][n][qb][qe][az]  = sum_ae T_e^se[ae][qe] U1[k][sb][qb][ae,az]   n: (sb, se) = (0,0) (0,1) (1,0)
//   G[k][n][qb,qe,qz]     = sum_az T_z^{[n = 0]}[az][qz] U2[k][n][qb][qe][az]
// and the transposed sum runs them backwards:
//   W[i][m][qb][qe][az]   = sum_qz T_z^{[m = 0]}[az][qz] S[i][m][qb,qe,qz]
//   Z[i][sb][qb][ae,az]   = sum_qe (sb = 0: T_e^0 W[i][0] + T_e^1 W[i][1]; sb = 1: T_e^0 W[i][2])
//   y[i][ab,ae,az]        = sum_qb T_b^0[ab][qb] Z[i][0] + T_b^1[ab][qb] Z[i][1]
// About 124 kflop per p = 4 element, and the 54 C values per point read once (K_e itself
// is 140,625 doubles).  One workgroup per element, each XCD taking a contiguous range of
// elements (its L2 then holds the x of neighbouring elements); two LDS buffers alternate
// between the stages (the per-point stage works in place: each thread reads its point's
// nine gradients before it writes the nine S values over them), and each stage reads an
// operand from LDS once per pencil, not once per multiply-add.
#pragma once

namespace fl {

template <int P1, int Q>
struct MfLayout {
    static constexpr int N3 = P1 * P1 * P1, Q3 = Q * Q * Q, PP = P1 * P1, QQ = Q * Q;
    static constexpr int mx(int a, int b) { return a > b ? a : b; }
    // buffer A: x (3 N3), U2 (9 QQ P1), W (9 QQ P1); buffer B: U1 (6 Q PP), G then S (9 Q3), Z (6 Q PP)
    static constexpr int BUFA = mx(3 * N3, 9 * QQ * P1);
    static constexpr int BUFB = mx(6 * Q * PP, 9 * Q3);
};

// C32: the C tensors rounded to FP32 (c32 [e][54][Q3]) instead of the records; each
// thread's 54 values are loaded into registers when the workgroup starts, so their latency
// overlaps the gradient stages.  The BF16 form fits five workgroups per CU (LDS and
// registers); the others four.
template <int P1, int Q, int CK>
__global__ __launch_bounds__(256, CK == 2 ? 5 : 4) void mf_element_kernel(const double *__restrict__ qp,
                                                         const float *__restrict__ c32,
                                                         const uint32_t *__restrict__ c16,
                                                         const int32_t *__restrict__ basis,
                                                         const int32_t *__restrict__ elements, int64_t nelem,
                                                         const uint8_t *__restrict__ free_mask,
                                                         const double *__restrict__ x, double *__restrict__ ye) {
    typedef SfLayout<P1, Q> L;
    typedef MfLayout<P1, Q> M;
    constexpr int N3 = M::N3, Q3 = M::Q3, PP = M::PP, QQ = M::QQ, NT = 256;
    // CK: 0 the FP64 records, 1 their FP32 copy, 2 their BF16 copy
    constexpr bool C32 = CK != 0;
    static_assert(!C32 || Q3 <= NT, "one point per thread");
    __shared__ double sF[3][2][P1][Q];
    __shared__ double sA[M::BUFA];
    __shared__ double sB[M::BUFB];
    __shared__ int32_t sLoc[N3];   // local node of tensor node t
    __shared__ int16_t sT[Q3];     // tensor point of record point g

    const int64_t chunk = (nelem + 7) / 8;
    const int64_t k8 = blockIdx.x >> 3;
    const int64_t e = (int64_t)(blockIdx.x & 7) * chunk + k8;
    if (k8 >= chunk || e >= nelem) return;
    const int tid = threadIdx.x;
    const int32_t *g_of_t = basis + 4;
    const int32_t *loc = basis + 4 + Q3;
    const double *f1d = reinterpret_cast<const double *>(basis + L::INTS);
    const int32_t *conn = elements + e * N3;
    float cr[C32 ? 54 : 1];
    if constexpr (CK == 1) {
        const float *r32 = c32 + e * (int64_t)54 * Q3 + min(tid, Q3 - 1);
#pragma unroll
        for (int f = 0; f < 54; ++f) cr[f] = r32[f * Q3];
    } else if constexpr (CK == 2) {
        const uint32_t *r16 = c16 + e * (int64_t)27 * Q3 + min(tid, Q3 - 1);
        uint32_t w[27];
#pragma unroll
        for (int j = 0; j < 27; ++j) w[j] = r16[j * Q3];
#pragma unroll
        for (int j = 0; j < 27; ++j) {
            cr[2 * j] = __uint_as_float(w[j] << 16);
            cr[2 * j + 1] = __uint_as_float(w[j] & 0xFFFF0000u);
        }
    }
    for (int t = tid; t < 6 * P1 * Q; t += NT) (&sF[0][0][0][0])[t] = f1d[t];
    for (int t = tid; t < Q3; t += NT) sT[min(max(g_of_t[t], 0), Q3 - 1)] = (int16_t)t;
    for (int t = tid; t < N3; t += NT) {
        const int l = min(max(loc[t], 0), N3 - 1);
        sLoc[t] = l;
        const int64_t a = conn[l];
#pragma unroll
        for (int k = 0; k < 3; ++k) sA[k * N3 + t] = free_mask[3 * a + k] ? x[3 * a + k] : 0.0;
    }
    __syncthreads();
    // Each stage runs over "pencils": a thread loads the P1 (or Q) values along the
    // contracted direction into registers once and writes every output of its pencil.
    // beta: U1 -> B, pencils (k, sb, ae az)
    for (int t = tid; t < 6 * PP; t += NT) {
        const int r = t % PP, kb = t / PP;
        const int k = kb >> 1, sb = kb & 1;
        double xv[P1];
#pragma unroll
        for (int ab = 0; ab < P1; ++ab) xv[ab] = sA[k * N3 + ab * PP + r];
        double *o = sB + kb * Q * PP + r;
#pragma unroll
        for (int qb = 0; qb < Q; ++qb) {
            double s = 0.0;
#pragma unroll
            for (int ab = 0; ab < P1; ++ab) s += sF[2][sb][ab][qb] * xv[ab];
            o[qb * PP] = s;
        }
    }
    __syncthreads();
    // eta: U2 -> A, pencils (k n, qb, az)
    for (int t = tid; t < 9 * Q * P1; t += NT) {
        const int az = t % P1, qb = (t / P1) % Q, kn = t / (P1 * Q);
        const int k = kn / 3, n = kn - 3 * k;
        const int sb = n == 2, se = n == 1;
        const double *u = sB + ((k * 2 + sb) * Q + qb) * PP + az;
        double uv[P1];
#pragma unroll
        for (int ae = 0; ae < P1; ++ae) uv[ae] = u[ae * P1];
        double *o = sA + (kn * Q + qb) * Q * P1 + az;
#pragma unroll
        for (int qe = 0; qe < Q; ++qe) {
            double s = 0.0;
#pragma unroll
            for (int ae = 0; ae < P1; ++ae) s += sF[1][se][ae][qe] * uv[ae];
            o[qe * P1] = s;
        }
    }
    __syncthreads();
    // zeta: G -> B, pencils (k n, qb qe)
    for (int t = tid; t < 9 * QQ; t += NT) {
        const int qbe = t % QQ, kn = t / QQ;
        const int n = kn % 3;
        const double *u = sA + (kn * QQ + qbe) * P1;
        double uv[P1];
#pragma unroll
        for (int az = 0; az < P1; ++az) uv[az] = u[az];
        double *o = sB + kn * Q3 + qbe * Q;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
            double s = 0.0;
#pragma unroll
            for (int az = 0; az < P1; ++az) s += sF[0][n == 0][az][qz] * uv[az];
            o[qz] = s;
        }
    }
    __syncthreads();
    // per point: S_im = sum_kn C^{ik}_mn G_kn, in place in B.  The record holds C^{ik} of the six
    // pairs i <= k (0,0 1,1 2,2 0,1 0,2 1,2); C^{ki}_mn = C^{ik}_nm.
    const double *rec = qp + e * (int64_t)QS * Q3 + (int64_t)Q_D * Q3;
    for (int g = tid; g < Q3; g += NT) {
        const int t = min(max((int)sT[g], 0), Q3 - 1);
        double gk[3][3], S[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                gk[k][n] = sB[(3 * k + n) * Q3 + t];
                S[k][n] = 0.0;
            }
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const int i = p < 3 ? p : (p == 5 ? 1 : 0), k = p < 3 ? p : (p == 3 ? 1 : 2);
            double c[9];
#pragma unroll
            for (int mn = 0; mn < 9; ++mn) {
                if constexpr (C32)
                    c[mn] = (double)cr[9 * p + mn];
                else
                    c[mn] = rec[(int64_t)(9 * p + mn) * Q3 + g];
            }
#pragma unroll
            for (int m = 0; m < 3; ++m)
                S[i][m] += c[3 * m] * gk[k][0] + c[3 * m + 1] * gk[k][1] + c[3 * m + 2] * gk[k][2];
            if (i != k) {
#pragma unroll
                for (int m = 0; m < 3; ++m) S[k][m] += c[m] * gk[i][0] + c[3 + m] * gk[i][1] + c[6 + m] * gk[i][2];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int m = 0; m < 3; ++m) sB[(3 * i + m) * Q3 + t] = S[i][m];
    }
    __syncthreads();
    // zeta transposed: W -> A, pencils (i m, qb qe)
    for (int t = tid; t < 9 * QQ; t += NT) {
        const int qbe = t % QQ, im = t / QQ;
        const int m = im % 3;
        const double *sv = sB + im * Q3 + qbe * Q;
        double v[Q];
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) v[qz] = sv[qz];
        double *o = sA + (im * QQ + qbe) * P1;
#pragma unroll
        for (int az = 0; az < P1; ++az) {
            double s = 0.0;
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) s += sF[0][m == 0][az][qz] * v[qz];
            o[az] = s;
        }
    }
    __syncthreads();
    // eta transposed: Z -> B, pencils (i, sb, qb, az)
    for (int t = tid; t < 6 * Q * P1; t += NT) {
        const int az = t % P1, qb = (t / P1) % Q, is = t / (P1 * Q);
        const int i = is >> 1, sb = is & 1;
        double *o = sB + ((i * 2 + sb) * Q + qb) * PP + az;
        if (sb == 0) {
            const double *w0 = sA + ((i * 3) * QQ + qb * Q) * P1 + az;
            const double *w1 = sA + ((i * 3 + 1) * QQ + qb * Q) * P1 + az;
            double a0[Q], a1[Q];
#pragma unroll
            for (int qe = 0; qe < Q; ++qe) {
                a0[qe] = w0[qe * P1];
                a1[qe] = w1[qe * P1];
            }
#pragma unroll
            for (int ae = 0; ae < P1; ++ae) {
                double s = 0.0;
#pragma unroll
                for (int qe = 0; qe < Q; ++qe) s += sF[1][0][ae][qe] * a0[qe] + sF[1][1][ae][qe] * a1[qe];
                o[ae * P1] = s;
            }
        } else {
            const double *w2 = sA + ((i * 3 + 2) * QQ + qb * Q) * P1 + az;
            double a2[Q];
#pragma unroll
            for (int qe = 0; qe < Q; ++qe) a2[qe] = w2[qe * P1];
#pragma unroll
            for (int ae = 0; ae < P1; ++ae) {
                double s = 0.0;
#pragma unroll
                for (int qe = 0; qe < Q; ++qe) s += sF[1][0][ae][qe] * a2[qe];
                o[ae * P1] = s;
            }
        }
    }
    __syncthreads();
    // beta transposed: y, in local node order, pencils (i, ae az)
    double *yo = ye + e * (int64_t)N3 * 3;
    for (int t = tid; t < 3 * PP; t += NT) {
        const int r = t % PP, i = t / PP;
        double z0[Q], z1[Q];
#pragma unroll
        for (int qb = 0; qb < Q; ++qb) {
            z0[qb] = sB[((i * 2) * Q + qb) * PP + r];
            z1[qb] = sB[((i * 2 + 1) * Q + qb) * PP + r];
        }
#pragma unroll
        for (int ab = 0; ab < P1; ++ab) {
            double s = 0.0;
#pragma unroll
            for (int qb = 0; qb < Q; ++qb) s += sF[2][0][ab][qb] * z0[qb] + sF[2][1][ab][qb] * z1[qb];
            yo[3 * sLoc[ab * PP + r] + i] = s;
        }
    }
}

// c32[e][f][g] = (float) records' C field f at point g (f < 54)
__global__ __launch_bounds__(256) void mf_round_kernel(const double *__restrict__ qp, int G, int64_t n,
                                                       float *__restrict__ c32) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t per = (int64_t)54 * G;
    const int64_t e = i / per, r = i - e * per;
    c32[i] = (float)qp[e * (int64_t)QS * G + (int64_t)Q_D * G + r];
}

// c16[e][j][g] = bf16(C field 2j) | bf16(C field 2j+1) << 16, rounded to nearest even
__device__ inline uint32_t to_bf16_bits(double v) {
    const uint32_t u = __float_as_uint((float)v);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__global__ __launch_bounds__(256) void mf_round16_kernel(const double *__restrict__ qp, int G, int64_t n,
                                                         uint32_t *__restrict__ c16) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t per = (int64_t)27 * G;
    const int64_t e = i / per, r = i - e * per;
    const int64_t j = r / G, g = r - j * G;
    const double *rec = qp + e * (int64_t)QS * G + (int64_t)Q_D * G;
    c16[i] = to_bf16_bits(rec[(2 * j) * G + g]) | (to_bf16_bits(rec[(2 * j + 1) * G + g]) << 16);
}

}  // namespace fl
