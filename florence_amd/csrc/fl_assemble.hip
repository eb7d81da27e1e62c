// Displacement-formulation K + residual assembly for high-order tets/hexes, gfx950.
//
// Restates the element loop of _GlobalAssemblyDF_ (Florence/FiniteElements/Assembly/
// _Assembly_/_LowLevelAssemblyDF_.h:69-161) as four device phases:
//
//  1. qp_kernel       one thread per (element, quadrature point): gather X, x; the
//                     kinematics of _KinematicMeasures_.h:64-119 (parent Jacobians,
//                     adjugate inverses, detJ = w |det dx/dxi|, F); the Neo-Hookean /
//                     Mooney-Rivlin laws of _NeoHookean_.h:27-51 / _MooneyRivlin_.h:30-70
//                     (and the linear elastic, nearly incompressible and anisotropic
//                     Mooney-Rivlin laws, material_law_gen below).
//                     Emits per point: dxi/dx (9), the six coupling blocks
//                     D^{ik}_{jl} = detJ (c_ijkl + delta_ik sigma_jl) (54) and detJ*sigma (6).
//  2. stiffness_kernel one workgroup per (element, block pair i<=k): the constitutive
//                     B.H.B^T of _ConstitutiveStiffnessDF_.h:83-134 plus the geometric term
//                     of _GeometricStiffness_.h:61-100, written per 3x3 displacement block
//                       K^{ik}_{ab} = sum_{g,j} G_{a,(g,j)} W^{ik}_{b,(g,j)},
//                       W^{ik}_{b,(g,j)} = sum_l D^{ik}_{jl}(g) G_{b,(g,l)},
//                     G_{a,(g,j)} = dN_a/dx_j -- a dense GEMM over the 3G inner index run
//                     on FP64 MFMA (v_mfma_f64_16x16x4f64) with the spatial gradients
//                     and W rebuilt per 8-point chunk in LDS.  K^{ki} = (K^{ik})^T.
//  3. traction_kernel one workgroup per element: t_(a,i) = sum_g detJ sigma_ij dN_a/dx_j
//                     (_ConstitutiveStiffnessDF_.h:136-148).
//  4. gather_kernel   one workgroup per node: its three CSR rows accumulated in LDS from
//                     the incident elements in ascending element order (the order of the
//                     reference's serial CSR accumulation, SparseAssemblyNative.h:32-45),
//                     then written once, coalesced; same for T (_LowLevelAssemblyDF_.h:153-161).
//                     No floating-point atomics: the result is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "fl_common.hpp"
#include "fl_profile.hpp"

namespace fl {

typedef double double4_t __attribute__((ext_vector_type(4)));

// per quadrature point record
// Point records are stored field-major per element: field f of point g of element e at
// qp[(e * QS + f) * G + g] (QS_RES for residual-only records), so the threads of an
// element, one per point, write and read each field coalesced.
constexpr int QS = 72;        // fields per point (doubles)
constexpr int Q_IX = 0;       // dxi/dx: ix[j*3+m] = d xi_m / d x_j
constexpr int Q_D = 9;        // 6 pairs x 9: D^{ik}[j*3+l]
constexpr int Q_S = 63;       // detJ*sigma: s00 s11 s22 s01 s02 s12
constexpr int Q_DETJ = 69;
constexpr int QS_RES = 16;    // residual-only record: ix (9), S (6 at offset 9), detJ (15)
constexpr int QR_S = 9;
constexpr int QR_DETJ = 15;

__constant__ int c_pair_i[6] = {0, 1, 2, 0, 0, 1};
__constant__ int c_pair_k[6] = {0, 1, 2, 1, 2, 2};

struct MatParams {
    int material;
    double mu, mu1, mu2, lamb, mu3;
    double n0, n1, n2;   // the element's fibre direction (FL_MAT_ANISOTROPICMOONEYRIVLIN_1)
};

// the laws whose tangent needs the general form below (qp kernels instantiated with GEN)
__host__ __device__ inline bool general_law(int material) { return material > FL_MAT_MOONEYRIVLIN; }

// rows of the per-element material table: {mu, mu1, mu2, lamb} or, for the anisotropic
// law, {mu, mu1, mu2, lamb, mu3, N0, N1, N2}
__host__ __device__ inline int material_table_stride(int material) {
    return material == FL_MAT_ANISOTROPICMOONEYRIVLIN_1 ? 8 : 4;
}

// cofactor of F (row-major): H_ij = dJ/dF_ij = J F^{-T}
__device__ __forceinline__ void cofactor3(const double (&F)[9], double (&H)[9]) {
    H[0] = F[4] * F[8] - F[5] * F[7];
    H[1] = F[5] * F[6] - F[3] * F[8];
    H[2] = F[3] * F[7] - F[4] * F[6];
    H[3] = F[2] * F[7] - F[1] * F[8];
    H[4] = F[0] * F[8] - F[2] * F[6];
    H[5] = F[1] * F[6] - F[0] * F[7];
    H[6] = F[1] * F[5] - F[2] * F[4];
    H[7] = F[2] * F[3] - F[0] * F[5];
    H[8] = F[0] * F[4] - F[1] * F[3];
}

// The other low-level laws: Cauchy stress sig and the coefficients of a spatial tangent
//   c_ijkl = c1 (2 b_ij b_kl - b_ik b_jl - b_il b_jk) + a1 (d_ik d_jl + d_il d_jk) + a2 d_ij d_kl
//          + A_ij d_kl + d_ij A_kl + B_ik d_jl + B_il d_jk + d_ik B_jl + d_il B_jk + e4 g_ij g_kl
// (A, B, g symmetric), and geo = 0 where the low-level assembler has no geometric term.
//  * linear elastic (_LinearElastic_.h:35-49): sigma = 2 mu e + lamb tr(e) I, e = sym(F) - I;
//    c = mu (d_ik d_jl + d_il d_jk) + lamb d_ij d_kl; no geometric stiffness
//    (AOT_Assembler.py:84-93).
//  * nearly incompressible Mooney-Rivlin (_NearlyIncompressibleMooneyRivlin_.h:34-81) with
//    (alpha, beta, kappa) = (mu1, mu2, lamb) as the generated assembler constructs it
//    (AOT_Assembler.py:69-70): with H = cof F, g = H H^T, c0 = J^(-5/3), c1' = sqrt(tr g),
//    c4 = J^-3, sigma = 2 alpha c0 b - 2/3 alpha c0 tr(b) I + beta c4 tr(g) c1' I
//    - 3 beta c4 c1' g + kappa (J-1) I.
//  * anisotropic Mooney-Rivlin 1 (_AnisotropicMooneyRivlin_1_.h:40-89) with n = H N:
//    the Mooney-Rivlin terms with mu1+2mu2+mu3 in the isotropic part, plus
//    2mu3/J (FN FN^T + (n.n) I - n n^T) in sigma.
struct GenTangent {
    double c1, a1, a2, e4, geo;
    double A[9], B[9], g[9];
};

__device__ __forceinline__ void material_law_gen(const MatParams &mp, const double (&F)[9], double J,
                                                 const double (&b)[9], double (&sig)[9], GenTangent &t) {
#pragma unroll
    for (int i = 0; i < 9; ++i) t.A[i] = t.B[i] = t.g[i] = 0.0;
    t.c1 = t.e4 = 0.0;
    t.geo = 1.0;
    if (mp.material == FL_MAT_LINEARELASTIC) {
        const double tre = F[0] + F[4] + F[8] - 3.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                sig[3 * i + j] = mp.mu * (F[3 * i + j] + F[3 * j + i] - (i == j ? 2.0 : 0.0)) +
                                 (i == j ? mp.lamb * tre : 0.0);
        t.a1 = mp.mu;
        t.a2 = mp.lamb;
        t.geo = 0.0;
    } else if (mp.material == FL_MAT_NEARLYINCOMPRESSIBLEMOONEYRIVLIN) {
        const double alpha = mp.mu1, beta = mp.mu2, kappa = mp.lamb;
        double Hc[9];
        cofactor3(F, Hc);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                t.g[3 * i + j] = Hc[3 * i] * Hc[3 * j] + Hc[3 * i + 1] * Hc[3 * j + 1] + Hc[3 * i + 2] * Hc[3 * j + 2];
        const double trb = b[0] + b[4] + b[8], trg = t.g[0] + t.g[4] + t.g[8];
        const double c0 = pow(J, -5.0 / 3.0), q1 = sqrt(trg), q2 = 1.0 / q1, q3 = trg * q1, c4 = 1.0 / (J * J * J);
        const double diag = -2.0 / 3.0 * alpha * c0 * trb + beta * c4 * q3 + kappa * (J - 1.0);
#pragma unroll
        for (int i = 0; i < 9; ++i) sig[i] = 2.0 * alpha * c0 * b[i] - 3.0 * beta * c4 * q1 * t.g[i];
        sig[0] += diag;
        sig[4] += diag;
        sig[8] += diag;
        t.a1 = 2.0 / 3.0 * alpha * c0 * trb - beta * c4 * q3 - kappa * (J - 1.0);
        t.a2 = 4.0 * alpha / 9.0 * c0 * trb + beta * c4 * q3 + kappa * (2.0 * J - 1.0);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            t.A[i] = -4.0 / 3.0 * alpha * c0 * b[i] - 3.0 * beta * c4 * q1 * t.g[i];
            t.B[i] = 3.0 * beta * c4 * q1 * t.g[i];
        }
        t.e4 = 3.0 * beta * c4 * q2;
    } else {   // FL_MAT_ANISOTROPICMOONEYRIVLIN_1
        double Hc[9];
        cofactor3(F, Hc);
        const double N[3] = {mp.n0, mp.n1, mp.n2};
        double FN[3], HN[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            FN[i] = F[3 * i] * N[0] + F[3 * i + 1] * N[1] + F[3 * i + 2] * N[2];
            HN[i] = Hc[3 * i] * N[0] + Hc[3 * i + 1] * N[1] + Hc[3 * i + 2] * N[2];
        }
        const double iHN = HN[0] * HN[0] + HN[1] * HN[1] + HN[2] * HN[2], trb = b[0] + b[4] + b[8];
        const double s3 = 2.0 * mp.mu3 / J;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double bb = b[3 * i] * b[j] + b[3 * i + 1] * b[3 + j] + b[3 * i + 2] * b[6 + j];
                sig[3 * i + j] = 2.0 * mp.mu1 / J * b[3 * i + j] + 2.0 * mp.mu2 / J * (trb * b[3 * i + j] - bb) +
                                 s3 * (FN[i] * FN[j] - HN[i] * HN[j]);
                t.A[3 * i + j] = -2.0 * s3 * HN[i] * HN[j];
                t.B[3 * i + j] = s3 * HN[i] * HN[j];
            }
        const double diag = -2.0 * (mp.mu1 + 2.0 * mp.mu2 + mp.mu3) / J + mp.lamb * (J - 1.0) + s3 * iHN;
        sig[0] += diag;
        sig[4] += diag;
        sig[8] += diag;
        t.c1 = 2.0 * mp.mu2 / J;
        t.a1 = 2.0 * (mp.mu1 + 2.0 * mp.mu2 + mp.mu3) / J - mp.lamb * (J - 1.0) - s3 * iHN;
        t.a2 = mp.lamb * (2.0 * J - 1.0) + 2.0 * s3 * iHN;
    }
}

__device__ inline double invdet3x3(const double *s, double *d) {
    d[0] = s[4] * s[8] - s[5] * s[7];
    d[1] = -s[1] * s[8] + s[2] * s[7];
    d[2] = s[1] * s[5] - s[2] * s[4];
    d[3] = -s[3] * s[8] + s[5] * s[6];
    d[4] = s[0] * s[8] - s[2] * s[6];
    d[5] = -s[0] * s[5] + s[2] * s[3];
    d[6] = s[3] * s[7] - s[4] * s[6];
    d[7] = -s[0] * s[7] + s[1] * s[6];
    d[8] = s[0] * s[4] - s[1] * s[3];
    const double det = s[0] * d[0] + s[1] * d[3] + s[2] * d[6];
    const double r = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] *= r;
    return det;
}

__global__ void clear_status_kernel(int32_t *status) {
    if (threadIdx.x < 2) status[threadIdx.x] = 0;
}

__global__ void zero_kernel(double *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Jm^T padded: JmT[(g*3+m)*NP + a] = dN_a/dxi_m at point g (0 for a >= npe)
__global__ void transpose_jm_kernel(const double *__restrict__ Jm, int npe, int G, int NP, double *__restrict__ JmT) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G * 3 * NP) return;
    const int a = t % NP, m = (t / NP) % 3, g = t / (3 * NP);
    JmT[t] = a < npe ? Jm[((int64_t)m * npe + a) * G + g] : 0.0;
}

// ------------------------------------------------------------------ phase 1
// Per-point kinematics, kinetics and record of one quadrature point from its parent
// Jacobians PX[3m+c] = dX_c/dxi_m, Px = dx_c/dxi_m (_KinematicMeasures_.h:64-119,
// _NeoHookean_.h:27-51, _MooneyRivlin_.h:30-70).  valid = false lanes only join the
// wave's status ballot.
// emitC: store, in place of each pair's D^{ik}, the tensor C^{ik}_{mn} = sum_{jl}
// (dxi_m/dx_j) D^{ik}_{jl} (dxi_n/dx_l) that the sum-factorised stiffness contracts
// (computed once per point instead of in each of the element's block workgroups).
template <bool TANGENT, bool GEN = false>
__device__ __forceinline__ void qp_point(const MatParams &mp, const double (&PX)[9], const double (&Px)[9], double wg,
                                         int update, bool valid, double *__restrict__ rec, int fs,
                                         int32_t *__restrict__ status, bool emitC = false,
                                         double *__restrict__ out = nullptr) {
    double iX[9], ix[9];
    const double detX = invdet3x3(PX, iX);
    const double detx = invdet3x3(Px, ix);
    const double detJ = wg * fabs(update == 1 ? detx : detX);
    // F_lj = sum_m (dx_l/dxi_m)(dxi_m/dX_j) = sum_m Px[m][l] iX[j][m]
    double F[9];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * l + j] = Px[l] * iX[3 * j] + Px[3 + l] * iX[3 * j + 1] + Px[6 + l] * iX[3 * j + 2];
    const double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
    double b[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) b[3 * i + j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
    double sig[9], c1, a1, a2;
    GenTangent gt;
    if (GEN) {
        material_law_gen(mp, F, J, b, sig, gt);
        c1 = gt.c1;
        a1 = gt.a1;
        a2 = gt.a2;
    } else if (mp.material == FL_MAT_NEOHOOKEAN) {
        // sigma = mu/J (b - I) + lamb (J-1) I; c = a1 (d_ik d_jl + d_il d_jk) + a2 d_ij d_kl
#pragma unroll
        for (int i = 0; i < 9; ++i) sig[i] = mp.mu / J * (b[i] - ((i % 4) == 0 ? 1.0 : 0.0));
        sig[0] += mp.lamb * (J - 1.0);
        sig[4] += mp.lamb * (J - 1.0);
        sig[8] += mp.lamb * (J - 1.0);
        c1 = 0.0;
        a1 = mp.mu / J - mp.lamb * (J - 1.0);
        a2 = mp.lamb * (2.0 * J - 1.0);
    } else {
        // sigma = 2mu1/J b + 2mu2/J (tr b b - b b) - 2(mu1+2mu2)/J I + lamb (J-1) I
        const double trb = b[0] + b[4] + b[8];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double bb = b[3 * i] * b[j] + b[3 * i + 1] * b[3 + j] + b[3 * i + 2] * b[6 + j];
                sig[3 * i + j] = 2.0 * mp.mu1 / J * b[3 * i + j] + 2.0 * mp.mu2 / J * (trb * b[3 * i + j] - bb);
            }
        const double diag = -2.0 * (mp.mu1 + 2.0 * mp.mu2) / J + mp.lamb * (J - 1.0);
        sig[0] += diag;
        sig[4] += diag;
        sig[8] += diag;
        c1 = 2.0 * mp.mu2 / J;
        a1 = 2.0 * (mp.mu1 + 2.0 * mp.mu2) / J - mp.lamb * (J - 1.0);
        a2 = mp.lamb * (2.0 * J - 1.0);
    }
    // status: det F <= 0 or non-finite values (wave-reduced, one atomic per wave)
    bool nonpos = valid && !(J > 0.0);
    bool nonfin =
        valid && (!isfinite(J) || !isfinite(detJ) || !isfinite(sig[0] + sig[4] + sig[8] + sig[1] + sig[2] + sig[5]));
    const unsigned long long m_np = __ballot(nonpos), m_nf = __ballot(nonfin);
    if ((m_np | m_nf) && (threadIdx.x & 63) == __ffsll((long long)(m_np | m_nf)) - 1) {
        atomicOr(status, (m_np ? FL_S_NONPOSITIVE_J : 0) | (m_nf ? FL_S_NONFINITE : 0));
        atomicAdd(status + 1, (int)__popcll(m_np));
    }
    if (!valid) return;
    if (out) {
        // dxi/dx and detJ sigma (s00 s11 s22 s01 s02 s12) for a fused traction
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = ix[i];
        out[9] = detJ * sig[0];
        out[10] = detJ * sig[4];
        out[11] = detJ * sig[8];
        out[12] = detJ * sig[1];
        out[13] = detJ * sig[2];
        out[14] = detJ * sig[5];
    }
    if (!rec) return;
    if (TANGENT) {
        // dxi/dx: read by the dense contractions (the C tensors already hold it)
        if (!emitC) {
#pragma unroll
            for (int i = 0; i < 9; ++i) rec[(Q_IX + i) * fs] = ix[i];
        }
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const int i = p < 3 ? p : (p == 5 ? 1 : 0);
            const int k = p < 3 ? p : (p == 3 ? 1 : 2);
            double D[9];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    const double dik = i == k ? 1.0 : 0.0, djl = j == l ? 1.0 : 0.0;
                    const double dil = i == l ? 1.0 : 0.0, djk = j == k ? 1.0 : 0.0;
                    const double dij = i == j ? 1.0 : 0.0, dkl = k == l ? 1.0 : 0.0;
                    double cijkl = c1 * (2.0 * b[3 * i + j] * b[3 * k + l] - b[3 * i + k] * b[3 * j + l] -
                                         b[3 * i + l] * b[3 * j + k]) +
                                   a1 * (dik * djl + dil * djk) + a2 * dij * dkl;
                    if (GEN) {
                        cijkl += gt.A[3 * i + j] * dkl + dij * gt.A[3 * k + l] + gt.B[3 * i + k] * djl +
                                 gt.B[3 * i + l] * djk + dik * gt.B[3 * j + l] + dil * gt.B[3 * j + k] +
                                 gt.e4 * gt.g[3 * i + j] * gt.g[3 * k + l];
                        D[3 * j + l] = detJ * (cijkl + gt.geo * dik * sig[3 * j + l]);
                    } else {
                        D[3 * j + l] = detJ * (cijkl + dik * sig[3 * j + l]);
                    }
                }
            if (emitC) {
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
                        rec[(Q_D + 9 * p + 3 * m + n) * fs] = ix[m] * W[n] + ix[3 + m] * W[3 + n] + ix[6 + m] * W[6 + n];
            } else {
#pragma unroll
                for (int t = 0; t < 9; ++t) rec[(Q_D + 9 * p + t) * fs] = D[t];
            }
        }
        if (!out) {
            // read by the traction kernels (a fused traction has them already)
            rec[(Q_S + 0) * fs] = detJ * sig[0];
            rec[(Q_S + 1) * fs] = detJ * sig[4];
            rec[(Q_S + 2) * fs] = detJ * sig[8];
            rec[(Q_S + 3) * fs] = detJ * sig[1];
            rec[(Q_S + 4) * fs] = detJ * sig[2];
            rec[(Q_S + 5) * fs] = detJ * sig[5];
            rec[(Q_DETJ) * fs] = detJ;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) rec[(i) * fs] = ix[i];
        rec[(QR_S + 0) * fs] = detJ * sig[0];
        rec[(QR_S + 1) * fs] = detJ * sig[4];
        rec[(QR_S + 2) * fs] = detJ * sig[8];
        rec[(QR_S + 3) * fs] = detJ * sig[1];
        rec[(QR_S + 4) * fs] = detJ * sig[2];
        rec[(QR_S + 5) * fs] = detJ * sig[5];
        rec[(QR_DETJ) * fs] = detJ;
    }
}

__device__ __forceinline__ MatParams element_material(const MatParams &mp0, int64_t e,
                                                      const int32_t *__restrict__ elem_mat,
                                                      const double *__restrict__ mat_table) {
    MatParams mp = mp0;
    if (elem_mat) {
        const double *row = mat_table + material_table_stride(mp0.material) * (int64_t)elem_mat[e];
        mp.mu = row[0];
        mp.mu1 = row[1];
        mp.mu2 = row[2];
        mp.lamb = row[3];
        if (mp0.material == FL_MAT_ANISOTROPICMOONEYRIVLIN_1) {
            mp.mu3 = row[4];
            mp.n0 = row[5];
            mp.n1 = row[6];
            mp.n2 = row[7];
        }
    }
    return mp;
}

// One thread per point: the parent Jacobians summed over the element's nodes.
// elem_mat/mat_table (optional): per-element material parameters (material_table_stride)
template <bool TANGENT, bool GEN>
__global__ __launch_bounds__(256) void qp_kernel(MatParams mp0, const double *__restrict__ points,
                                                 const int32_t *__restrict__ elements,
                                                 const double *__restrict__ Eulerx, const double *__restrict__ Jm,
                                                 const double *__restrict__ w, int G, int64_t nelem, int npe,
                                                 int update, double *__restrict__ qp, int32_t *__restrict__ status,
                                                 const int32_t *__restrict__ elem_mat,
                                                 const double *__restrict__ mat_table) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < nelem * G;
    const int64_t e = valid ? t / G : 0;
    const int g = valid ? (int)(t - e * G) : 0;
    const MatParams mp = element_material(mp0, e, elem_mat, mat_table);
    double PX[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Px[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int32_t *conn = elements + e * npe;
    for (int a = 0; valid && a < npe; ++a) {
        const int64_t nd = conn[a];
        const double X0 = points[3 * nd], X1 = points[3 * nd + 1], X2 = points[3 * nd + 2];
        const double x0 = Eulerx[3 * nd], x1 = Eulerx[3 * nd + 1], x2 = Eulerx[3 * nd + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double jm = Jm[((int64_t)i * npe + a) * G + g];
            PX[3 * i + 0] += jm * X0;
            PX[3 * i + 1] += jm * X1;
            PX[3 * i + 2] += jm * X2;
            Px[3 * i + 0] += jm * x0;
            Px[3 * i + 1] += jm * x1;
            Px[3 * i + 2] += jm * x2;
        }
    }
    if (!valid) PX[0] = PX[4] = PX[8] = Px[0] = Px[4] = Px[8] = 1.0;
    qp_point<TANGENT, GEN>(mp, PX, Px, valid ? w[g] : 1.0, update, valid,
                           qp + e * (int64_t)G * (TANGENT ? QS : QS_RES) + g, G, status);
}

// ------------------------------------------------------------------ phase 2
// ------------------------------------------------------------------ phase 2
}  // namespace fl
#include "fl_stiffness.hpp"
#include "fl_stiffness_bf16.hpp"
#include "fl_stiffness_sf.hpp"
#include "fl_mg_galerkin.hpp"
#include "fl_mg_mf.hpp"
namespace fl {

// ------------------------------------------------------------------ phase 3
// t_(a,i) = sum_g sum_m dN_a/dxi_m(g) P_g[m][i],  P_g[m][i] = sum_j ix[j][m] S_g[i][j]
__global__ __launch_bounds__(256) void traction_kernel(const double *__restrict__ JmT, int NP,
                                                       const double *__restrict__ qp, int qs, int soff, int npe,
                                                       int G, double *__restrict__ te) {
    extern __shared__ double P[];   // G*9
    const int64_t e = blockIdx.x;
    const double *qe = qp + e * (int64_t)G * qs;
    for (int idx = threadIdx.x; idx < G * 9; idx += blockDim.x) {
        const int g = idx / 9, m = (idx / 3) % 3, i = idx % 3;
        const double *rec = qe + g;   // field f at rec[f * G]
        const double *S = rec + (int64_t)soff * G;   // s00 s11 s22 s01 s02 s12
        const double Si[3][3] = {{S[0], S[3 * G], S[4 * G]}, {S[3 * G], S[1 * G], S[5 * G]},
                                 {S[4 * G], S[5 * G], S[2 * G]}};
        P[idx] = rec[(0 * 3 + m) * G] * Si[i][0] + rec[(1 * 3 + m) * G] * Si[i][1] + rec[(2 * 3 + m) * G] * Si[i][2];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 3 * npe; idx += blockDim.x) {
        const int a = idx / 3, i = idx % 3;
        double acc = 0.0;
        for (int g = 0; g < G; ++g) {
            const double *jt = JmT + (int64_t)g * 3 * NP + a;
            acc += jt[0] * P[g * 9 + 0 + i] + jt[NP] * P[g * 9 + 3 + i] + jt[2 * NP] * P[g * 9 + 6 + i];
        }
        te[e * 3 * npe + idx] = acc;
    }
}

// ------------------------------------------------------------------ phase 4
__device__ inline int find_rank(const int32_t *s, int n, int32_t v) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v < s[mid]) hi = mid; else lo = mid;
    }
    return s[lo] == v ? lo : -1;
}

// eslot[q] = stored row of node-element entry q's block (entry q = e npe + la)
__global__ void entry_slot_kernel(const int32_t *__restrict__ node_elem, int64_t n, int npe,
                                  const int32_t *__restrict__ kperm, int32_t *__restrict__ eslot) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) eslot[q] = min(max(kperm[node_elem[q] % npe], 0), npe - 1);
}

// Stored row/column of local node l in an element block: blocks of the sum-factorised
// kernels are stored in tensor node order (kperm = tensor_of_l), the others in local
// order (kperm = NULL).
__device__ __forceinline__ int kslot(const int32_t *__restrict__ kperm, int l, int npe) {
    return kperm ? min(max(kperm[l], 0), npe - 1) : l;
}

// A node inside one element: its three rows are that element's rows with the columns
// permuted -- no accumulation.  The nine element rows (9 npe values) are read
// coalesced into LDS (sv), then each output row is written coalesced, its values
// picked from sv through the column map (as 0.0 + value, the accumulated form, so
// signed zeros match the sum path).  LDS: inv (npe ints), snb (deg ints, used only
// without a rank table), sv (9 npe doubles).
template <typename KT, int GT, typename VT>
__device__ __forceinline__ void gather_single_node(int64_t a, int32_t ent, int rslot, int npe, int deg, int32_t p0,
                                                   const int32_t *__restrict__ elements,
                                                   const int32_t *__restrict__ nbrs,
                                                   const int32_t *__restrict__ indptr, const KT *__restrict__ Ke,
                                                   const double *__restrict__ te, const int16_t *__restrict__ ranks,
                                                   const int32_t *__restrict__ kperm, int32_t *inv, int32_t *snb,
                                                   double *sv, VT *__restrict__ V, double *__restrict__ T,
                                                   int32_t *__restrict__ status) {
    const int tid = threadIdx.x;
    const int64_t npe2 = (int64_t)npe * npe;
    const int64_t e = ent / npe;
    const int la = ent - (int)(e * npe);
    const KT *Kb = Ke + e * 9 * npe2 + (int64_t)rslot * npe;
    // issue the row loads first (9 npe <= UF GT for npe <= 128)
    constexpr int UF = (9 * 128 + GT - 1) / GT;
    double v[UF];
#pragma unroll
    for (int u = 0; u < UF; ++u) {
        const int idx = u * GT + tid;
        v[u] = 0.0;
        if (idx < 9 * npe) {
            const int ik = idx / npe, lb = idx - ik * npe;
            v[u] = (double)Kb[ik * npe2 + lb];
        }
    }
    const double tv = tid < 3 ? te[e * 3 * npe + 3 * la + tid] : 0.0;
    bool bad = false;
    for (int r = tid; r < npe; r += GT) inv[r] = 0;   // position -> stored column
    __syncthreads();
    if (ranks) {
        const int16_t *rk = ranks + ((int64_t)ent) * npe;
        for (int lb = tid; lb < npe; lb += GT) inv[rk[lb]] = kslot(kperm, lb, npe);
    } else {
        for (int i = tid; i < deg; i += GT) snb[i] = nbrs[p0 + i];
        __syncthreads();
        for (int lb = tid; lb < npe; lb += GT) {
            const int r = find_rank(snb, deg, elements[e * npe + lb]);
            bad |= r < 0;
            if (r >= 0) inv[r] = kslot(kperm, lb, npe);
        }
    }
#pragma unroll
    for (int u = 0; u < UF; ++u) {
        const int idx = u * GT + tid;
        if (idx < 9 * npe) sv[idx] = v[u];
    }
    __syncthreads();
    const int rowlen = 3 * deg;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        VT *row = V + indptr[3 * a + i];
        const double *si = sv + 3 * i * npe;
        for (int c = tid; c < rowlen; c += GT) {
            const int r = c / 3, k = c - 3 * r;
            row[c] = (VT)(0.0 + si[k * npe + inv[r]]);
        }
    }
    if (tid < 3) T[3 * a + tid] = 0.0 + tv;
    if (bad) atomicOr(status, 4);
}

// The lowest degree class when every node in it lies inside one element: one wave per
// node, LDS inv (npe ints) | snb (npe ints) | sv (9 npe doubles).
template <typename KT, typename VT>
__global__ __launch_bounds__(64) void gather_single_kernel(const int32_t *__restrict__ elements, int npe,
                                                           const int32_t *__restrict__ node_order,
                                                           const int32_t *__restrict__ node_elem_ptr,
                                                           const int32_t *__restrict__ node_elem,
                                                           const int32_t *__restrict__ nbr_ptr,
                                                           const int32_t *__restrict__ nbrs,
                                                           const int32_t *__restrict__ indptr,
                                                           const KT *__restrict__ Ke, const double *__restrict__ te,
                                                           const int16_t *__restrict__ ranks,
                                                           const int32_t *__restrict__ kperm,
                                                           const int32_t *__restrict__ eslot, VT *__restrict__ V,
                                                           double *__restrict__ T, int32_t *__restrict__ status) {
    extern __shared__ double dsm[];
    int32_t *ism = reinterpret_cast<int32_t *>(dsm);
    const int64_t a = node_order[blockIdx.x];
    const int deg = nbr_ptr[a + 1] - nbr_ptr[a];
    const int32_t q0 = node_elem_ptr[a], q1 = node_elem_ptr[a + 1];
    if (q1 - q0 != 1 || deg != npe) {
        if (threadIdx.x == 0) atomicOr(status, FL_S_PATTERN);
        return;
    }
    const int32_t ent = node_elem[q0];
    const int rs = eslot ? min(max(eslot[q0], 0), npe - 1) : ent % npe;
    gather_single_node<KT, 64, VT>(a, ent, rs, npe, deg, nbr_ptr[a], elements, nbrs, indptr, Ke, te, ranks, kperm,
                                   ism, ism + npe, dsm + npe, V, T, status);
}

// One workgroup of GT threads per node.  Nodes are visited through node_order (grouped
// by neighbour count so each launch sizes its LDS to its own degree class); lds_deg is
// that class's max.  The rows of up to EB elements (and their tractions) are loaded
// into registers (UB loads per lane per element, UB GT >= 9 npe) before any is added,
// so a node's loads are in flight together; the adds then run element by element in
// node_elem order with a barrier between elements, so the sums do not depend on timing.
// LDS: rr (column positions by stored column, EB npe ints) | snb (lds_deg ints) | acc
// (9 lds_deg doubles; the staging rows of a single-element node).
template <typename KT, int GT, int EB, int UB, typename VT>
__global__ __launch_bounds__(GT) void gather_kernel(const int32_t *__restrict__ elements, int npe,
                                                    const int32_t *__restrict__ node_order,
                                                    const int32_t *__restrict__ node_elem_ptr,
                                                    const int32_t *__restrict__ node_elem,
                                                    const int32_t *__restrict__ nbr_ptr,
                                                    const int32_t *__restrict__ nbrs, const int32_t *__restrict__ indptr,
                                                    const KT *__restrict__ Ke, const double *__restrict__ te,
                                                    int lds_deg, const int16_t *__restrict__ ranks,
                                                    const int32_t *__restrict__ kperm,
                                                    const int32_t *__restrict__ kinv,
                                                    const int32_t *__restrict__ eslot, VT *__restrict__ V,
                                                    double *__restrict__ T, int32_t *__restrict__ status) {
    extern __shared__ double smem[];
    const int64_t a = node_order[blockIdx.x];
    const int tid = threadIdx.x;
    const int deg = nbr_ptr[a + 1] - nbr_ptr[a];
    const int32_t p0 = nbr_ptr[a];
    int32_t *rr = reinterpret_cast<int32_t *>(smem);
    int32_t *snb = rr + EB * npe;
    double *acc = smem + (EB * npe + lds_deg + 1) / 2;
    if (deg > lds_deg || 9 * npe > UB * GT) {
        if (tid == 0) atomicOr(status, FL_S_PATTERN);
        return;
    }
    const int32_t q0 = node_elem_ptr[a], q1 = node_elem_ptr[a + 1];
    if (q1 - q0 == 1 && deg == npe) {
        const int32_t ent = node_elem[q0];
        const int rs = eslot ? min(max(eslot[q0], 0), npe - 1) : ent % npe;
        gather_single_node<KT, GT, VT>(a, ent, rs, npe, deg, p0, elements, nbrs, indptr, Ke, te, ranks, kperm, rr,
                                       snb, acc, V, T, status);
        return;
    }
    const int rowlen = 3 * deg;
    const int64_t npe2 = (int64_t)npe * npe;
    for (int i = tid; i < 9 * deg; i += GT) acc[i] = 0.0;
    if (!ranks)
        for (int i = tid; i < deg; i += GT) snb[i] = nbrs[p0 + i];
    // this lane's slots of an element's 9 npe stored values (consecutive lanes read
    // consecutive stored columns): source offset, target base, stored column
    int src[UB], tgt[UB], col[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        const int idx = u * GT + tid;
        src[u] = -1;
        tgt[u] = 0;
        col[u] = 0;
        if (idx < 9 * npe) {
            const int ik = idx / npe, ts = idx - ik * npe;
            const int i = ik / 3, k = ik - 3 * i;
            src[u] = ik * (int)npe2 + ts;
            tgt[u] = i * rowlen + k;
            col[u] = ts;
        }
    }
    __syncthreads();
    bool bad = false;
    double tsum = 0.0;
    for (int32_t c0 = q0; c0 < q1; c0 += EB) {
        const int ne = min(EB, (int)(q1 - c0));
        // column positions of the chunk's elements (precomputed per mesh, or searched), by
        // stored column
        for (int t = tid; t < ne * npe; t += GT) {
            const int j = t / npe, lb = kslot(kinv, t - j * npe, npe);
            const int32_t ent = node_elem[c0 + j];
            int r;
            if (ranks) {
                r = ranks[(int64_t)ent * npe + lb];
            } else {
                r = find_rank(snb, deg, elements[(int64_t)(ent / npe) * npe + lb]);
                bad |= r < 0;
                r = r < 0 ? 0 : r;
            }
            rr[t] = r;
        }
        double v[EB][UB], tv[EB];
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            tv[j] = 0.0;
            if (j < ne) {
                const int32_t ent = node_elem[c0 + j];
                const int64_t e = ent / npe;
                const int la = ent - (int)(e * npe);
                const int rs = eslot ? min(max(eslot[c0 + j], 0), npe - 1) : la;
                const KT *Kb = Ke + e * 9 * npe2 + (int64_t)rs * npe;
#pragma unroll
                for (int u = 0; u < UB; ++u) v[j][u] = src[u] >= 0 ? (double)Kb[src[u]] : 0.0;
                if (tid < 3) tv[j] = te[e * 3 * npe + 3 * la + tid];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            if (j < ne) {
#pragma unroll
                for (int u = 0; u < UB; ++u)
                    if (src[u] >= 0) acc[tgt[u] + 3 * rr[j * npe + col[u]]] += v[j][u];
                tsum += tv[j];
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        VT *row = V + indptr[3 * a + i];
        const double *ai = acc + i * rowlen;
        for (int c = tid; c < rowlen; c += GT) row[c] = (VT)ai[c];
    }
    if (tid < 3) T[3 * a + tid] = tsum;
    if (bad) atomicOr(status, 4);
}

__global__ void gather_traction_kernel(int npe, const int32_t *__restrict__ node_elem_ptr,
                                       const int32_t *__restrict__ node_elem, const double *__restrict__ te,
                                       int64_t nnode, double *__restrict__ T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nnode) return;
    const int64_t a = t / 3;
    const int i = (int)(t - 3 * a);
    double s = 0.0;
    for (int32_t q = node_elem_ptr[a]; q < node_elem_ptr[a + 1]; ++q) {
        const int32_t ent = node_elem[q];
        const int64_t e = ent / npe;
        const int la = ent - (int)(e * npe);
        s += te[e * 3 * npe + 3 * la + i];
    }
    T[t] = s;
}

// ------------------------------------------------------------------ host side
// Gather launch shapes (threads per node, elements per chunk, loads per lane per
// element).  Class 0 (the smallest neighbour count) whose degree equals npe holds
// element-interior nodes only (a node of two distinct elements has more than npe
// neighbours): gather_single_kernel.  Elements of at most 64 nodes: one wave per node
// in every class, <64, 8, 2> for npe <= 14 and <64, 4, 9> up to 64.  Larger elements:
// the other classes use <256, EB, 5> with EB = 2, 4 or 8 as the class's degree suggests
// its nodes' element count (deg <= 2 npe, <= 4 npe, more: chunks longer than that only
// cost registers), class 0 one wave per node, <64, 2, 18>.
enum GatherForm { G_SINGLE, G_W8X2, G_W4X9, G_W2X18, G_WIDE1, G_WIDE2, G_WIDE4, G_WIDE8 };

// FL_TUNE_GATHER: bit 0 = the wide classes one element at a time (<256, 1, 5>), bit 1 =
// the single-element class through gather_kernel<64, 2, 18>
static GatherForm gather_form(int b, int deg, int npe) {
    const int t = tuning_value(TUNE_GATHER, 0);
    if (b == 0 && deg == npe && !(t & 2)) return G_SINGLE;
    // up to 64 nodes an element's 9 npe values fill at most 9 loads of one wave
    if (npe <= 14) return G_W8X2;
    if (npe <= 64) return G_W4X9;
    if (b > 0) {
        if (t & 1) return G_WIDE1;
        return deg <= 2 * npe ? G_WIDE2 : (deg <= 4 * npe ? G_WIDE4 : G_WIDE8);
    }
    return G_W2X18;
}

static int gather_eb(GatherForm f) {
    switch (f) {
        case G_W8X2: case G_WIDE8: return 8;
        case G_W4X9: case G_WIDE4: return 4;
        case G_W2X18: case G_WIDE2: return 2;
        default: return 1;
    }
}

// LDS of one launch: gather_single_kernel's inv + snb + sv, or gather_kernel's
// rr (eb npe ints) | snb (deg ints) | acc (9 deg doubles)
static size_t gather_lds_bytes(GatherForm f, int deg, int npe) {
    if (f == G_SINGLE) return (size_t)8 * npe + (size_t)72 * npe + 16;
    return (size_t)(gather_eb(f) * npe + deg + 1) / 2 * 8 + (size_t)deg * 9 * 8 + 16;
}

static int padded_np(int npe) {
    if (npe <= 16) return 16;
    if (npe <= 32) return 32;
    if (npe <= 48) return 48;
    if (npe <= 64) return 64;
    if (npe <= 128) return 128;
    return -1;
}

template <typename CT, typename KT>
static int launch_stiffness_t(int NP, const double *JmT, const double *qp, int npe, int G, int64_t nelem, KT *Ke,
                              hipStream_t st) {
    const int64_t groups = (nelem + 7) / 8 * 8 * 6;
    if (groups > (int64_t)INT32_MAX) {
        set_error("too many elements for one launch");
        return FL_E_LIMIT;
    }
    const dim3 grid((unsigned)groups);
    // tile-shape variants per padded node count (fl_set_tuning key
    // FL_TUNE_STIFF_VARIANT_BASE + class, value = variant + 1; 0 = default variant 0)
    // <MT, NT, WGM, WGN>: MT x NT tiles of 16 x 16 per wave, WGM x WGN waves per block
#define FL_DENSE(NPV, MT_, NT_, WGM_, WGN_) \
    stiffness_kernel<CT, KT, NPV, MT_, NT_, WGM_, WGN_><<<grid, 64 * WGM_ * WGN_, 0, st>>>(JmT, qp, npe, G, nelem, Ke)
    switch (NP) {
        case 16: FL_DENSE(16, 1, 1, 1, 1); break;
        case 32:
            switch (tuning_value(TUNE_STIFF_VARIANT_BASE + 1, 1)) {
                case 2: FL_DENSE(32, 1, 1, 2, 2); break;
                case 3: FL_DENSE(32, 2, 1, 1, 2); break;
                case 4: FL_DENSE(32, 1, 2, 2, 1); break;
                default: FL_DENSE(32, 2, 2, 1, 1); break;
            }
            break;
        case 48:
            switch (tuning_value(TUNE_STIFF_VARIANT_BASE + 2, 1)) {
                case 2: FL_DENSE(48, 1, 3, 3, 1); break;
                case 3: FL_DENSE(48, 3, 1, 1, 3); break;
                default: FL_DENSE(48, 3, 3, 1, 1); break;
            }
            break;
        case 64:
            switch (tuning_value(TUNE_STIFF_VARIANT_BASE + 3, 1)) {
                case 2: FL_DENSE(64, 4, 4, 1, 1); break;
                case 3: FL_DENSE(64, 4, 2, 1, 2); break;
                case 4: FL_DENSE(64, 2, 4, 2, 1); break;
                case 5: FL_DENSE(64, 2, 1, 2, 4); break;
                default: FL_DENSE(64, 2, 2, 2, 2); break;
            }
            break;
#undef FL_DENSE
        case 128:
            // default: registers capped for two waves per SIMD (measured faster on MI355X)
            if (tuning_value(TUNE_STIFF_VARIANT_BASE + 4, 1) == 2 || tuning_value(TUNE_STIFF_MINW, 2) == 1) {
                stiffness_kernel<CT, KT, 128, 4, 4, 2, 2, 1><<<grid, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke);
            } else {
                // off-diagonal pairs as full 128 x 128 blocks (v0: 2 x 2 waves of 4 x 4
                // tiles; v2: 1 x 4 waves of 8 x 2 tiles, fewer W operands per MFMA), then
                // the diagonal pairs as balanced upper triangles
                const dim3 half((unsigned)(groups / 2));
                if (tuning_value(TUNE_STIFF_VARIANT_BASE + 4, 1) == 3)
                    stiffness_kernel<CT, KT, 128, 8, 2, 1, 4, 2, 1><<<half, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke);
                else
                    stiffness_kernel<CT, KT, 128, 4, 4, 2, 2, 2, 1><<<half, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke);
                stiffness_kernel<CT, KT, 128, 4, 4, 2, 2, 2, 2><<<half, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke);
            }
            break;
        default: set_error("unsupported node count"); return FL_E_ARG;
    }
    FL_LAUNCH_CHECK();
    return 0;
}

// Sum-factorised FP64 stiffness of tensor-product hexes (P1^3 nodes, Q^3 points).
// Vdirect (the default form only): rows of element-interior nodes go straight into V
static int launch_stiffness_sf(int npe, int G, const double *qp, const void *basis, int64_t nelem, double *Ke,
                               const int32_t *elements, const int32_t *node_elem_ptr, const int32_t *indptr,
                               int32_t *irow, double *Vdirect, hipStream_t st) {
    const int64_t groups = (nelem + 7) / 8 * 8 * 9;
    if (groups > (int64_t)INT32_MAX) {
        set_error("too many elements for one launch");
        return FL_E_LIMIT;
    }
    const dim3 grid((unsigned)groups);
    const int32_t *b = static_cast<const int32_t *>(basis);
    // FL_TUNE_STIFF_SF value 3: the all-VALU form; 5: eta and beta on MFMA one type at a
    // time; default: two types per MFMA reduction
    const int form = tuning_value(TUNE_STIFF_SF, 0);
#define FL_SF_CASE(P, QQ)                                                                  \
    if (npe == P * P * P && G == QQ * QQ * QQ) {                                           \
        if (form == 3)                                                                     \
            stiffness_sf_kernel<P, QQ><<<grid, 256, 0, st>>>(qp, b, nelem, Ke);            \
        else if (form == 5)                                                                \
            stiffness_sf_mfma_kernel<P, QQ><<<grid, 256, 0, st>>>(qp, b, nelem, Ke);       \
        else {                                                                             \
            if (Vdirect) {                                                                 \
                const int64_t nt = nelem * P * P * P;                                      \
                interior_rows_kernel<P, QQ><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>( \
                    b, elements, nelem, node_elem_ptr, indptr, irow);                      \
                FL_LAUNCH_CHECK();                                                         \
            }                                                                              \
            stiffness_sf_mfma2_kernel<P, QQ><<<grid, 256, 0, st>>>(qp, b, nelem, Ke, irow, \
                                                                   Vdirect);               \
        }                                                                                  \
        FL_LAUNCH_CHECK();                                                                 \
        return 0;                                                                          \
    }
    FL_SF_CASE(2, 2)
    FL_SF_CASE(3, 3)
    FL_SF_CASE(4, 4)
    FL_SF_CASE(5, 5)
    FL_SF_CASE(5, 6)
#undef FL_SF_CASE
    set_error("no sum-factorised kernel for %d nodes and %d points", npe, G);
    return FL_E_ARG;
}

// Galerkin operators of hex coarse levels (the p=2 lattice or the vertices of the fine
// hexes): the factorised FP64 stiffness of the coarse space over the fine rule's C
// records, every block stored in the coarse table's tensor node order.
static bool galerkin_sf_shape(int npe_c, int G) {
    return (npe_c == 27 && (G == 216 || G == 125)) || (npe_c == 8 && (G == 216 || G == 125 || G == 64 || G == 27));
}

static int launch_galerkin_sf(int npe_c, int G, const double *qp, const int32_t *basis_c, int64_t nelem, double *Kc,
                              hipStream_t st) {
    const int64_t groups = (nelem + 7) / 8 * 8 * 9;
    if (groups > (int64_t)INT32_MAX) {
        set_error("too many elements for one launch");
        return FL_E_LIMIT;
    }
    const dim3 grid((unsigned)groups);
#define FL_GSF_CASE(P, QQ)                                                                                   \
    if (npe_c == P * P * P && G == QQ * QQ * QQ) {                                                         \
        stiffness_sf_mfma2_kernel<P, QQ, false, double, double>                                            \
            <<<grid, 256, 0, st>>>(qp, basis_c, nelem, Kc, nullptr, nullptr);                               \
        FL_LAUNCH_CHECK();                                                                                 \
        return 0;                                                                                          \
    }
    FL_GSF_CASE(3, 6)
    FL_GSF_CASE(3, 5)
    FL_GSF_CASE(2, 6)
    FL_GSF_CASE(2, 5)
    FL_GSF_CASE(2, 4)
    FL_GSF_CASE(2, 3)
#undef FL_GSF_CASE
    set_error("no factorised Galerkin kernel for %d coarse nodes and %d points", npe_c, G);
    return FL_E_ARG;
}

// Sum-factorised FP32 stiffness blocks (FP32 MFMA, FP32 accumulation) of tensor-product
// hexes, stored in tensor node order for the gather (no interior-row fusion).
static int launch_stiffness_sf32(int npe, int G, const double *qp, const void *basis, int64_t nelem, float *Ke,
                                 hipStream_t st) {
    const int64_t groups = (nelem + 7) / 8 * 8 * 9;
    if (groups > (int64_t)INT32_MAX) {
        set_error("too many elements for one launch");
        return FL_E_LIMIT;
    }
    const dim3 grid((unsigned)groups);
    const int32_t *b = static_cast<const int32_t *>(basis);
#define FL_SF_CASE(P, QQ)                                                                                      \
    if (npe == P * P * P && G == QQ * QQ * QQ) {                                                               \
        stiffness_sf_mfma2_kernel<P, QQ, false, float, double><<<grid, 256, 0, st>>>(qp, b, nelem, Ke, nullptr, nullptr); \
        FL_LAUNCH_CHECK();                                                                                     \
        return 0;                                                                                              \
    }
    FL_SF_CASE(2, 2)
    FL_SF_CASE(3, 3)
    FL_SF_CASE(4, 4)
    FL_SF_CASE(5, 5)
    FL_SF_CASE(5, 6)
#undef FL_SF_CASE
    set_error("no sum-factorised kernel for %d nodes and %d points", npe, G);
    return FL_E_ARG;
}

// The factorised stiffness straight into the CSR, one launch per colour in colour order
// (stiffness_sf_mfma2_kernel's DIRECT form).  offsets (host): the colour ranges of elems.
// CT: the contraction type (double; float for the FP32 tangent), VT: the CSR values'
// (double; float for FL_PREC_CSR32 with the FP32 tangent).
template <typename CT, typename VT>
static int launch_stiffness_direct(int npe, int G, const double *qp, const void *basis, const int32_t *elements,
                                   const int32_t *indptr, int ncolour, const int32_t *offsets,
                                   const int32_t *elems, const int16_t *scat, VT *V, hipStream_t st) {
    const int32_t *b = static_cast<const int32_t *>(basis);
    for (int c = 0; c < ncolour; ++c) {
        const int64_t n = offsets[c + 1] - offsets[c];
        if (n <= 0) continue;
        if ((n + 7) / 8 * 8 * 9 > (int64_t)INT32_MAX) {
            set_error("too many elements of one colour for one launch");
            return FL_E_LIMIT;
        }
        const dim3 grid((unsigned)((n + 7) / 8 * 8 * 9));
        const int32_t *el = elems + offsets[c];
#define FL_SF_CASE(P, QQ)                                                                              \
    if (npe == P * P * P && G == QQ * QQ * QQ) {                                                       \
        stiffness_sf_mfma2_kernel<P, QQ, true, CT, VT><<<grid, 256, 0, st>>>(qp, b, n, nullptr, nullptr, V, el, \
                                                                             elements, indptr, scat);  \
        FL_LAUNCH_CHECK();                                                                             \
        continue;                                                                                      \
    }
        FL_SF_CASE(2, 2)
        FL_SF_CASE(3, 3)
        FL_SF_CASE(4, 4)
        FL_SF_CASE(5, 5)
        FL_SF_CASE(5, 6)
#undef FL_SF_CASE
        set_error("no sum-factorised kernel for %d nodes and %d points", npe, G);
        return FL_E_ARG;
    }
    return 0;
}

// Tdirect: T entries of element-interior nodes written here (their gather is skipped)
template <bool TANGENT>
static int launch_qp_sf(const MatParams &mp, const double *points, const int32_t *elements, const double *Eulerx,
                        const void *basis, const double *w, int npe, int G, int64_t nelem, int update, double *qp,
                        int32_t *status, const int32_t *elem_mat, const double *mat_table, hipStream_t st,
                        int emitC = 0, double *te = nullptr, const int32_t *node_elem_ptr = nullptr,
                        double *Tdirect = nullptr) {
    const int32_t *b = static_cast<const int32_t *>(basis);
#define FL_QSF_CASE(P, QQ)                                                                                     \
    if (npe == P * P * P && G == QQ * QQ * QQ) {                                                               \
        if (general_law(mp.material))                                                                          \
            qp_sf_kernel<P, QQ, TANGENT, true><<<(unsigned)nelem, 256, 0, st>>>(                               \
                mp, points, elements, Eulerx, b, w, nelem, update, qp, status, elem_mat, mat_table, emitC, te,  \
                node_elem_ptr, Tdirect);                                                                       \
        else                                                                                                   \
            qp_sf_kernel<P, QQ, TANGENT, false><<<(unsigned)nelem, 256, 0, st>>>(                              \
                mp, points, elements, Eulerx, b, w, nelem, update, qp, status, elem_mat, mat_table, emitC, te,  \
                node_elem_ptr, Tdirect);                                                                       \
        FL_LAUNCH_CHECK();                                                                                     \
        return 0;                                                                                              \
    }
    FL_QSF_CASE(2, 2)
    FL_QSF_CASE(3, 3)
    FL_QSF_CASE(4, 4)
    FL_QSF_CASE(5, 5)
    FL_QSF_CASE(5, 6)
#undef FL_QSF_CASE
    set_error("no sum-factorised point kernel for %d nodes and %d points", npe, G);
    return FL_E_ARG;
}

// the FP32 tangent takes the factorised kernel unless a form other than the default
// factorised one is selected (FL_TUNE_STIFF_SF 3 and 5 are FP64-only forms: dense there)
static bool sf32_form() {
    const int f = tuning_value(TUNE_STIFF_SF, 0);
    return f != 1 && f != 3 && f != 5;
}

static bool sf_shape(int npe, int G) {
    return (npe == 8 && G == 8) || (npe == 27 && G == 27) || (npe == 64 && G == 64) || (npe == 125 && G == 125) ||
           (npe == 125 && G == 216);
}

static int launch_stiffness_bf16(int NP, const double *JmT, const double *qp, int npe, int G, int64_t nelem,
                                 float *Ke, hipStream_t st) {
    const int64_t groups = (nelem + 7) / 8 * 8 * 6;
    if (groups > (int64_t)INT32_MAX) {
        set_error("too many elements for one launch");
        return FL_E_LIMIT;
    }
    const dim3 grid((unsigned)groups);
    switch (NP) {
        case 16: stiffness_bf16_kernel<16, 1, 1, 1, 1><<<grid, 64, 0, st>>>(JmT, qp, npe, G, nelem, Ke); break;
        case 32: stiffness_bf16_kernel<32, 2, 2, 1, 1><<<grid, 64, 0, st>>>(JmT, qp, npe, G, nelem, Ke); break;
        case 48: stiffness_bf16_kernel<48, 3, 3, 1, 1><<<grid, 64, 0, st>>>(JmT, qp, npe, G, nelem, Ke); break;
        case 64: stiffness_bf16_kernel<64, 2, 2, 2, 2><<<grid, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke); break;
        case 128: stiffness_bf16_kernel<128, 4, 4, 2, 2><<<grid, 256, 0, st>>>(JmT, qp, npe, G, nelem, Ke); break;
        default: set_error("unsupported node count"); return FL_E_ARG;
    }
    FL_LAUNCH_CHECK();
    return 0;
}

template <typename KT, int GT, int EB, int UB, typename VT>
static int launch_gather_t(int32_t nb, size_t lds, const int32_t *elements, int npe, const int32_t *order,
                           const int32_t *node_elem_ptr, const int32_t *node_elem, const int32_t *nbr_ptr,
                           const int32_t *nbrs, const int32_t *indptr, const KT *Ke, const double *te, int deg,
                           const int16_t *ranks, const int32_t *kperm, const int32_t *kinv, const int32_t *eslot,
                           VT *V, double *T, int32_t *status, hipStream_t st) {
    if (lds > 64 * 1024)
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)gather_kernel<KT, GT, EB, UB, VT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    gather_kernel<KT, GT, EB, UB, VT><<<(unsigned)nb, GT, lds, st>>>(elements, npe, order, node_elem_ptr,
                                                                     node_elem, nbr_ptr, nbrs, indptr, Ke, te, deg,
                                                                     ranks, kperm, kinv, eslot, V, T, status);
    FL_LAUNCH_CHECK();
    return 0;
}

// one degree class's gather launch
template <typename KT, typename VT>
static int launch_gather(int b, int32_t nb, const int32_t *elements, int npe, const int32_t *order,
                         const int32_t *node_elem_ptr, const int32_t *node_elem, const int32_t *nbr_ptr,
                         const int32_t *nbrs, const int32_t *indptr, const KT *Ke, const double *te, int deg,
                         const int16_t *ranks, const int32_t *kperm, const int32_t *kinv, const int32_t *eslot,
                         VT *V, double *T, int32_t *status, hipStream_t st) {
    const GatherForm f = gather_form(b, deg, npe);
    const size_t lds = gather_lds_bytes(f, deg, npe);
    switch (f) {
        case G_SINGLE:
            gather_single_kernel<KT, VT><<<(unsigned)nb, 64, lds, st>>>(elements, npe, order, node_elem_ptr,
                                                                       node_elem, nbr_ptr, nbrs, indptr, Ke, te,
                                                                       ranks, kperm, eslot, V, T, status);
            FL_LAUNCH_CHECK();
            return 0;
        case G_W8X2:
            return launch_gather_t<KT, 64, 8, 2, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem, nbr_ptr,
                                                     nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V, T, status, st);
        case G_W4X9:
            return launch_gather_t<KT, 64, 4, 9, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem, nbr_ptr,
                                                     nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V, T, status, st);
        case G_W2X18:
            return launch_gather_t<KT, 64, 2, 18, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem,
                                                      nbr_ptr, nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V,
                                                      T, status, st);
        case G_WIDE1:
            return launch_gather_t<KT, 256, 1, 5, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem,
                                                      nbr_ptr, nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V,
                                                      T, status, st);
        case G_WIDE2:
            return launch_gather_t<KT, 256, 2, 5, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem,
                                                      nbr_ptr, nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V,
                                                      T, status, st);
        case G_WIDE4:
            return launch_gather_t<KT, 256, 4, 5, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem,
                                                      nbr_ptr, nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V,
                                                      T, status, st);
        default:
            return launch_gather_t<KT, 256, 8, 5, VT>(nb, lds, elements, npe, order, node_elem_ptr, node_elem,
                                                      nbr_ptr, nbrs, indptr, Ke, te, deg, ranks, kperm, kinv, eslot, V,
                                                      T, status, st);
    }
}

struct DFWork {
    double *JmT, *qp, *te;
    void *Ke;
    int32_t *slot;   // stored row of each node-element entry's block (tensor-ordered blocks)
    int32_t *irow;   // interior_rows_kernel's table (4 ints per element node)
};

// bytes per element-block entry for each tangent precision
static size_t ke_bytes(int precision) { return precision == FL_PREC_FP64 ? sizeof(double) : sizeof(float); }

// blocks = false: no element-block buffer and tables (residual-only pass, direct scatter);
// records = false: no point records (the factorised residual pass contracts the
// tractions inside its point kernel)
static DFWork carve_df(void *ws, size_t bytes, int G, int64_t nelem, int npe, bool tangent, size_t *need,
                       int precision = FL_PREC_FP64, bool blocks = true, bool records = true) {
    const int NP = padded_np(npe);
    Carve cv(ws, bytes);
    DFWork w;
    w.JmT = cv.take<double>((size_t)G * 3 * NP);
    w.qp = records ? cv.take<double>((size_t)nelem * G * (tangent ? QS : QS_RES)) : nullptr;
    w.te = cv.take<double>((size_t)nelem * 3 * npe);
    blocks = blocks && tangent;
    w.Ke = blocks ? (void *)cv.take<char>((size_t)nelem * 9 * npe * npe * ke_bytes(precision)) : nullptr;
    w.slot = blocks ? cv.take<int32_t>((size_t)nelem * npe) : nullptr;
    w.irow = blocks ? cv.take<int32_t>((size_t)nelem * npe * 4) : nullptr;
    *need = cv.off + 256;
    if (!cv.ok) w.JmT = nullptr;
    return w;
}

static bool check_common(int material, const double *params, int G, int64_t nelem, int npe, int64_t nnode) {
    if (material < FL_MAT_NEOHOOKEAN || material > FL_MAT_ANISOTROPICMOONEYRIVLIN_1) {
        set_error("unknown material id %d", material);
        return false;
    }
    if (!params || G <= 0 || nelem <= 0 || nnode <= 0 || padded_np(npe) < 0) {
        set_error("invalid sizes (ngauss=%d nelem=%lld npe=%d nnode=%lld); npe must be in [1,128]", G,
                  (long long)nelem, npe, (long long)nnode);
        return false;
    }
    if (nelem * (int64_t)npe >= INT32_MAX) {
        set_error("nelem*npe exceeds int32");
        return false;
    }
    return true;
}

// params: {mu, mu1, mu2, lamb} and, for the anisotropic law, mu3 in params[4] (its
// fibre directions only come per element, through the material table)
static MatParams mat_params(int material, const double *p) {
    MatParams m;
    m.material = material;
    m.mu = p[0];
    m.mu1 = p[1];
    m.mu2 = p[2];
    m.lamb = p[3];
    m.mu3 = material == FL_MAT_ANISOTROPICMOONEYRIVLIN_1 ? p[4] : 0.0;
    m.n0 = m.n1 = m.n2 = 0.0;
    return m;
}

// Workspace of fl_df_residual_noalloc: a file-scope device array of a fixed size.  HIP
// gives every GPU the library runs on its own copy, at its own address, held until the
// process ends.  In host code the name `scratch` is a host-side stand-in, not a device
// address: host code only ever uses the per-device address hipGetSymbolAddress returns
// (scratch_address).  Nothing in it has to start at zero (the residual pass writes every
// element traction before the gather reads it); calls that overlap on one GPU share it
// and race.
static __device__ __align__(256) unsigned char scratch[FL_NOALLOC_SCRATCH_BYTES];

constexpr int FL_NOALLOC_MAX_DEVICES = 64;
static void *g_scratch_addr[FL_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current device, looked up on the first call per
// device (that device current) and kept
static int scratch_address(void **out) {
    int dev = -1;
    FL_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= FL_NOALLOC_MAX_DEVICES) {
        set_error("fl_df_residual_noalloc: device %d is beyond the %d devices the library tracks", dev,
                  FL_NOALLOC_MAX_DEVICES);
        return FL_E_LIMIT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("fl_df_residual_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return FL_E_HIP;
        }
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

static bool check_table(int material, const int32_t *elem_mat, const double *mat_table, const char *who) {
    if ((elem_mat == nullptr) != (mat_table == nullptr)) {
        set_error("%s: elem_material and material_table go together", who);
        return false;
    }
    if (material == FL_MAT_ANISOTROPICMOONEYRIVLIN_1 && !elem_mat) {
        set_error("%s: the anisotropic Mooney-Rivlin law takes its fibre directions from a material table", who);
        return false;
    }
    return true;
}

int mf_prepare(const FlMgMatrixFree *in, MfOp *op) {
    if (!in || !op || !in->df_workspace || !in->tensor_basis || !in->elements || !in->node_elem_ptr ||
        !in->node_elem || !in->element_vectors || in->nelem <= 0 || in->nnode <= 0 ||
        in->nelem * (int64_t)in->npe >= INT32_MAX) {
        set_error("matrix-free operator: incomplete description");
        return FL_E_ARG;
    }
    if (in->records != FL_RECORDS_C) {
        set_error("matrix-free operator: needs the C point records of the factorised hex path");
        return FL_E_ARG;
    }
    if (!sf_shape(in->npe, in->ngauss)) {
        set_error("matrix-free operator: %d nodes and %d points are not a tensor-product hex rule", in->npe,
                  in->ngauss);
        return FL_E_ARG;
    }
    size_t need = 0;
    DFWork fw = carve_df(const_cast<void *>(in->df_workspace), in->df_workspace_bytes, in->ngauss, in->nelem,
                         in->npe, true, &need, FL_PREC_FP64, false);
    if (!fw.JmT) {
        set_error("matrix-free operator: the workspace (%zu bytes) holds no point records for these sizes",
                  in->df_workspace_bytes);
        return FL_E_SPACE;
    }
    int p1 = 1, q = 1;
    while (p1 * p1 * p1 < in->npe) ++p1;
    while (q * q * q < in->ngauss) ++q;
    op->qp = fw.qp;
    op->basis = static_cast<const int32_t *>(in->tensor_basis);
    op->elements = in->elements;
    op->nelem = in->nelem;
    op->nnode = in->nnode;
    op->npe = in->npe;
    op->G = in->ngauss;
    op->p1 = p1;
    op->q = q;
    op->node_elem_ptr = in->node_elem_ptr;
    op->node_elem = in->node_elem;
    op->ye = in->element_vectors;
    op->c32 = nullptr;
    op->c16 = nullptr;
    return 0;
}

int mf_round_records16(const MfOp &op, uint32_t *c16, hipStream_t st) {
    const int64_t n = op.nelem * 27 * (int64_t)op.G;
    mf_round16_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(op.qp, op.G, n, c16);
    FL_LAUNCH_CHECK();
    return 0;
}

int mf_round_records(const MfOp &op, float *c32, hipStream_t st) {
    const int64_t n = op.nelem * 54 * (int64_t)op.G;
    mf_round_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(op.qp, op.G, n, c32);
    FL_LAUNCH_CHECK();
    return 0;
}

int mf_element_apply(const MfOp &op, const uint8_t *free_mask, const double *x, hipStream_t st) {
    const int64_t chunk = (op.nelem + 7) / 8;
    const unsigned grid = (unsigned)(8 * chunk);
#define FL_MF_LAUNCH(P, QQ)                                                                                        \
    do {                                                                                                           \
        if (op.c16)                                                                                                \
            mf_element_kernel<P, QQ, 2><<<grid, 256, 0, st>>>(op.qp, op.c32, op.c16, op.basis, op.elements,         \
                                                              op.nelem, free_mask, x, op.ye);                      \
        else if (op.c32)                                                                                           \
            mf_element_kernel<P, QQ, 1><<<grid, 256, 0, st>>>(op.qp, op.c32, op.c16, op.basis, op.elements,         \
                                                              op.nelem, free_mask, x, op.ye);                      \
        else                                                                                                       \
            mf_element_kernel<P, QQ, 0><<<grid, 256, 0, st>>>(op.qp, op.c32, op.c16, op.basis, op.elements,         \
                                                              op.nelem, free_mask, x, op.ye);                      \
    } while (0)
    if (op.p1 == 2 && op.q == 2)
        FL_MF_LAUNCH(2, 2);
    else if (op.p1 == 3 && op.q == 3)
        FL_MF_LAUNCH(3, 3);
    else if (op.p1 == 4 && op.q == 4)
        FL_MF_LAUNCH(4, 4);
    else if (op.p1 == 5 && op.q == 5)
        FL_MF_LAUNCH(5, 5);
    else if (op.p1 == 5 && op.q == 6)
        FL_MF_LAUNCH(5, 6);
    else {
        set_error("matrix-free operator: no kernel for %d^3 nodes and %d^3 points", op.p1, op.q);
        return FL_E_ARG;
    }
#undef FL_MF_LAUNCH
    FL_LAUNCH_CHECK();
    return 0;
}

}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_df_workspace_size(int32_t ngauss, int64_t nelem, int32_t npe) {
    if (padded_np(npe) < 0) return 0;
    size_t need = 0;
    carve_df(nullptr, 0, ngauss, nelem, npe, true, &need);
    return need;
}

size_t fl_df_residual_workspace_size(int32_t ngauss, int64_t nelem, int32_t npe) {
    if (padded_np(npe) < 0) return 0;
    size_t need = 0;
    carve_df(nullptr, 0, ngauss, nelem, npe, false, &need);
    return need;
}

size_t fl_tensor_basis_bytes(int32_t p1, int32_t q) {
    if (p1 <= 0 || q <= 0) return 0;
    const int64_t q3 = (int64_t)q * q * q, n3 = (int64_t)p1 * p1 * p1;
    const int64_t ints = 4 + q3 + 2 * n3 + (q3 % 2);
    return (size_t)(4 * ints + 8 * 3 * 2 * p1 * q);
}

size_t fl_df_workspace_size_ex(int32_t ngauss, int64_t nelem, int32_t npe, int32_t precision) {
    if (padded_np(npe) < 0 || precision < FL_PREC_FP64 || precision > FL_PREC_BF16) return 0;
    size_t need = 0;
    carve_df(nullptr, 0, ngauss, nelem, npe, true, &need, precision);
    return need;
}

int fl_df_assemble(int32_t material, const double *params, const double *points, const int32_t *elements,
                   const double *Eulerx, const double *Jm, const double *AllGauss, int32_t G, int64_t nelem,
                   int32_t npe, int64_t nnode, int32_t update, const int32_t *node_elem_ptr,
                   const int32_t *node_elem, const int32_t *node_nbr_ptr, const int32_t *node_nbrs,
                   const int32_t *node_order, const int32_t *bucket_offsets, const int32_t *bucket_degree,
                   const int32_t *indptr, double *V, double *T, void *workspace, size_t workspace_bytes,
                   int32_t *status, void *stream) {
    return fl_df_assemble_ex(material, params, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe, nnode, update,
                             node_elem_ptr, node_elem, node_nbr_ptr, node_nbrs, node_order, bucket_offsets,
                             bucket_degree, indptr, V, T, FL_PREC_FP64, workspace, workspace_bytes, status, stream);
}

int fl_df_assemble_ex(int32_t material, const double *params, const double *points, const int32_t *elements,
                      const double *Eulerx, const double *Jm, const double *AllGauss, int32_t G, int64_t nelem,
                      int32_t npe, int64_t nnode, int32_t update, const int32_t *node_elem_ptr,
                      const int32_t *node_elem, const int32_t *node_nbr_ptr, const int32_t *node_nbrs,
                      const int32_t *node_order, const int32_t *bucket_offsets, const int32_t *bucket_degree,
                      const int32_t *indptr, double *V, double *T, int32_t precision, void *workspace,
                      size_t workspace_bytes, int32_t *status, void *stream) {
    return fl_df_assemble_mt(material, params, nullptr, nullptr, points, elements, Eulerx, Jm, AllGauss, G, nelem,
                             npe, nnode, update, node_elem_ptr, node_elem, node_nbr_ptr, node_nbrs, node_order,
                             bucket_offsets, bucket_degree, indptr, V, T, precision, nullptr, nullptr, workspace,
                             workspace_bytes, status, stream);
}

}  // extern "C"

namespace fl {
// the colouring and scatter table of the direct form (offsets on the host)
struct ColourPlan {
    int32_t ncolour;
    const int32_t *offsets;
    const int32_t *elems;
    const int16_t *scat;
};
}  // namespace fl

static int df_assemble(int32_t material, const double *params, const int32_t *elem_mat, const double *mat_table,
                       const double *points, const int32_t *elements, const double *Eulerx, const double *Jm,
                       const double *AllGauss, int32_t G, int64_t nelem, int32_t npe, int64_t nnode, int32_t update,
                       const int32_t *node_elem_ptr, const int32_t *node_elem, const int32_t *node_nbr_ptr,
                       const int32_t *node_nbrs, const int32_t *node_order, const int32_t *bucket_offsets,
                       const int32_t *bucket_degree, const int32_t *indptr, double *V, double *T, int32_t precision,
                       const int16_t *ranks, const void *tensor_basis, const ColourPlan *cp, void *workspace,
                       size_t workspace_bytes, int32_t *status, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!check_common(material, params, G, nelem, npe, nnode)) return FL_E_ARG;
    if (!check_table(material, elem_mat, mat_table, "fl_df_assemble_mt")) return FL_E_ARG;
    // FL_PREC_CSR32 (with an FP32/BF16 contraction): V is written as float[nnz]
    const bool v32 = (precision & FL_PREC_CSR32) != 0;
    precision &= ~FL_PREC_CSR32;
    if (precision < FL_PREC_FP64 || precision > FL_PREC_BF16 || (v32 && precision == FL_PREC_FP64)) {
        set_error("fl_df_assemble_mt: unsupported precision %d%s", precision, v32 ? " with FL_PREC_CSR32" : "");
        return FL_E_ARG;
    }
    if (!points || !elements || !Eulerx || !Jm || !AllGauss || !node_elem_ptr || !node_elem || !node_nbr_ptr ||
        !node_nbrs || !node_order || !bucket_offsets || !bucket_degree || !indptr || !V || !T || !status) {
        set_error("fl_df_assemble: null pointer");
        return FL_E_ARG;
    }
    if (bucket_offsets[0] != 0 || bucket_offsets[FL_GATHER_BUCKETS] != nnode) {
        set_error("fl_df_assemble: gather plan does not cover the %lld nodes", (long long)nnode);
        return FL_E_ARG;
    }
    // the factorised point, traction and (FP64) stiffness kernels whenever the tensor
    // basis is given, so T is the same for every tangent precision
    const bool sf_ok = tensor_basis != nullptr && sf_shape(npe, G) && tuning_value(TUNE_STIFF_SF, 0) != 1;
    const bool use_sf = precision == FL_PREC_FP64 && sf_ok;
    // the FP32 tangent of tensor-product hexes: the default factorised form in FP32
    const bool use_sf32 = precision == FL_PREC_FP32 && sf_ok && sf32_form();
    // the direct scatter takes the default factorised form only (FP64, or FP32 into the
    // FP64 or the FP32-stored CSR)
    const bool direct = cp != nullptr && (use_sf || use_sf32) && tuning_value(TUNE_STIFF_SF, 0) != 3 &&
                        tuning_value(TUNE_STIFF_SF, 0) != 5;
    size_t need = 0;
    DFWork w = carve_df(workspace, workspace_bytes, G, nelem, npe, true, &need, precision, !direct);
    if (!w.JmT) {
        set_error("fl_df_assemble: workspace too small (%zu < %zu)", workspace_bytes, need);
        return FL_E_SPACE;
    }
    for (int b = 0; b < FL_GATHER_BUCKETS && !direct; ++b) {
        if (bucket_offsets[b + 1] > bucket_offsets[b] &&
            gather_lds_bytes(gather_form(b, bucket_degree[b], npe), bucket_degree[b], npe) > 160 * 1024) {
            set_error("fl_df_assemble: node degree %d needs more LDS than a CU has", bucket_degree[b]);
            return FL_E_LIMIT;
        }
    }
    const int NP = padded_np(npe);
    const MatParams mp = mat_params(material, params);
    prof_begin(st, PROF_DF_ASSEMBLE);
    // status cleared by a kernel, not a memset node, so the call can be graph-captured
    clear_status_kernel<<<1, 64, 0, st>>>(status);
    transpose_jm_kernel<<<(G * 3 * NP + 255) / 256, 256, 0, st>>>(Jm, npe, G, NP, w.JmT);
    FL_LAUNCH_CHECK();
    const int64_t nq = nelem * G;
    // element-interior rows straight from the default factorised kernel into V (and their
    // T entries from the fused traction) when the lowest gather class holds exactly
    // those nodes (its degree equals npe); not with update != 1, whose T is zero
    const bool fuse_interior = use_sf && !direct && update == 1 && tuning_value(TUNE_STIFF_SF, 0) != 3 &&
                               tuning_value(TUNE_STIFF_SF, 0) != 5 && bucket_offsets[1] > bucket_offsets[0] &&
                               bucket_degree[0] == npe;
    if (sf_ok) {
        // the factorised FP64 stiffness reads C^{ik} from the records; the dense FP32/BF16
        // contractions read D^{ik}.  The element tractions are contracted in the same kernel.
        const int rq = launch_qp_sf<true>(mp, points, elements, Eulerx, tensor_basis, AllGauss, npe, G, nelem, update,
                                          w.qp, status, elem_mat, mat_table, st, use_sf || use_sf32 ? 1 : 0, w.te,
                                          node_elem_ptr, fuse_interior ? T : nullptr);
        if (rq) return rq;
    } else {
        if (general_law(material))
            qp_kernel<true, true><<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(
                mp, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe, update, w.qp, status, elem_mat, mat_table);
        else
            qp_kernel<true, false><<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(
                mp, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe, update, w.qp, status, elem_mat, mat_table);
        FL_LAUNCH_CHECK();
    }
    prof_mark(st);
    // the factorised blocks are stored in tensor node order: tensor_of_l of the basis table,
    // and each node-element entry's stored row read next to the entry itself
    const bool tensor_blocks = (use_sf || use_sf32) && !direct;
    const int32_t *kperm = tensor_blocks ? static_cast<const int32_t *>(tensor_basis) + 4 + G + npe : nullptr;
    const int32_t *kinv = tensor_blocks ? static_cast<const int32_t *>(tensor_basis) + 4 + G : nullptr;   // local_of_t
    const int32_t *eslot = nullptr;
    if (kperm) {
        const int64_t ne = nelem * npe;
        entry_slot_kernel<<<(unsigned)((ne + 255) / 256), 256, 0, st>>>(node_elem, ne, npe, kperm, w.slot);
        FL_LAUNCH_CHECK();
        eslot = w.slot;
    }
    int rc = direct ? (use_sf ? launch_stiffness_direct<double, double>(npe, G, w.qp, tensor_basis, elements, indptr,
                                                                        cp->ncolour, cp->offsets, cp->elems, cp->scat,
                                                                        V, st)
                       : v32  ? launch_stiffness_direct<float, float>(npe, G, w.qp, tensor_basis, elements, indptr,
                                                                      cp->ncolour, cp->offsets, cp->elems, cp->scat,
                                                                      reinterpret_cast<float *>(V), st)
                              : launch_stiffness_direct<float, double>(npe, G, w.qp, tensor_basis, elements, indptr,
                                                                       cp->ncolour, cp->offsets, cp->elems, cp->scat,
                                                                       V, st))
           : use_sf ? launch_stiffness_sf(npe, G, w.qp, tensor_basis, nelem, (double *)w.Ke, elements, node_elem_ptr,
                                          indptr, w.irow, fuse_interior ? V : nullptr, st)
           : use_sf32 ? launch_stiffness_sf32(npe, G, w.qp, tensor_basis, nelem, (float *)w.Ke, st)
           : precision == FL_PREC_FP64
                 ? launch_stiffness_t<double, double>(NP, w.JmT, w.qp, npe, G, nelem, (double *)w.Ke, st)
                 : precision == FL_PREC_FP32
                       ? launch_stiffness_t<float, float>(NP, w.JmT, w.qp, npe, G, nelem, (float *)w.Ke, st)
                       : launch_stiffness_bf16(NP, w.JmT, w.qp, npe, G, nelem, (float *)w.Ke, st);
    if (rc) return rc;
    prof_mark(st);
    if (!sf_ok) {
        const size_t lds_tr = (size_t)G * 9 * 8;
        if (lds_tr > 64 * 1024)
            FL_HIP_CHECK(hipFuncSetAttribute((const void *)traction_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds_tr));
        traction_kernel<<<(unsigned)nelem, 256, lds_tr, st>>>(w.JmT, NP, w.qp, QS, Q_S, npe, G, w.te);
        FL_LAUNCH_CHECK();
    }
    if (update != 1) {
        const int64_t nte = nelem * 3 * npe;
        zero_kernel<<<(unsigned)((nte + 255) / 256), 256, 0, st>>>(w.te, nte);
        FL_LAUNCH_CHECK();
    }
    prof_mark(st);
    if (direct) {
        // K is complete; T by node from the element vectors
        gather_traction_kernel<<<(unsigned)((3 * nnode + 255) / 256), 256, 0, st>>>(npe, node_elem_ptr, node_elem,
                                                                                     w.te, nnode, T);
        FL_LAUNCH_CHECK();
        prof_mark(st);
        return 0;
    }
    for (int b = fuse_interior ? 1 : 0; b < FL_GATHER_BUCKETS; ++b) {
        const int32_t nb = bucket_offsets[b + 1] - bucket_offsets[b];
        if (nb <= 0) continue;
        int rc2 = 0;
        const int32_t *ord = node_order + bucket_offsets[b];
        const int dg = bucket_degree[b];
        if (precision == FL_PREC_FP64)
            rc2 = launch_gather<double, double>(b, nb, elements, npe, ord, node_elem_ptr, node_elem, node_nbr_ptr,
                                                node_nbrs, indptr, (const double *)w.Ke, w.te, dg, ranks, kperm,
                                                kinv, eslot, V, T, status, st);
        else if (!v32)
            rc2 = launch_gather<float, double>(b, nb, elements, npe, ord, node_elem_ptr, node_elem, node_nbr_ptr,
                                               node_nbrs, indptr, (const float *)w.Ke, w.te, dg, ranks, kperm,
                                               kinv, eslot, V, T, status, st);
        else
            rc2 = launch_gather<float, float>(b, nb, elements, npe, ord, node_elem_ptr, node_elem, node_nbr_ptr,
                                              node_nbrs, indptr, (const float *)w.Ke, w.te, dg, ranks, kperm,
                                              kinv, eslot, reinterpret_cast<float *>(V), T, status, st);
        if (rc2) return rc2;
    }
    prof_mark(st);
    return 0;
}

extern "C" {

int fl_df_assemble_mt(int32_t material, const double *params, const int32_t *elem_mat, const double *mat_table,
                      const double *points, const int32_t *elements, const double *Eulerx, const double *Jm,
                      const double *AllGauss, int32_t G, int64_t nelem, int32_t npe, int64_t nnode, int32_t update,
                      const int32_t *node_elem_ptr, const int32_t *node_elem, const int32_t *node_nbr_ptr,
                      const int32_t *node_nbrs, const int32_t *node_order, const int32_t *bucket_offsets,
                      const int32_t *bucket_degree, const int32_t *indptr, double *V, double *T, int32_t precision,
                      const int16_t *ranks, const void *tensor_basis, void *workspace, size_t workspace_bytes,
                      int32_t *status, void *stream) {
    return df_assemble(material, params, elem_mat, mat_table, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, update, node_elem_ptr, node_elem, node_nbr_ptr, node_nbrs, node_order, bucket_offsets,
                       bucket_degree, indptr, V, T, precision, ranks, tensor_basis, nullptr, workspace,
                       workspace_bytes, status, stream);
}

size_t fl_df_workspace_size_cs(int32_t ngauss, int64_t nelem, int32_t npe) {
    if (padded_np(npe) < 0) return 0;
    size_t need = 0;
    carve_df(nullptr, 0, ngauss, nelem, npe, true, &need, FL_PREC_FP64, false);
    return need;
}

int fl_df_assemble_cs(int32_t material, const double *params, const int32_t *elem_mat, const double *mat_table,
                      const double *points, const int32_t *elements, const double *Eulerx, const double *Jm,
                      const double *AllGauss, int32_t G, int64_t nelem, int32_t npe, int64_t nnode, int32_t update,
                      const int32_t *node_elem_ptr, const int32_t *node_elem, const int32_t *node_nbr_ptr,
                      const int32_t *node_nbrs, const int32_t *node_order, const int32_t *bucket_offsets,
                      const int32_t *bucket_degree, const int32_t *indptr, double *V, double *T, int32_t precision,
                      const int16_t *ranks, const void *tensor_basis, int32_t ncolour, const int32_t *colour_offsets,
                      const int32_t *colour_elems, const int16_t *scatter, void *workspace, size_t workspace_bytes,
                      int32_t *status, void *stream) {
    if (ncolour < 1 || ncolour > 64 || !colour_offsets || !colour_elems || !scatter) {
        set_error("fl_df_assemble_cs: invalid colour plan (ncolour %d)", ncolour);
        return FL_E_ARG;
    }
    if (colour_offsets[0] != 0 || colour_offsets[ncolour] != nelem) {
        set_error("fl_df_assemble_cs: colour offsets do not cover the %lld elements", (long long)nelem);
        return FL_E_ARG;
    }
    for (int c = 0; c < ncolour; ++c)
        if (colour_offsets[c + 1] < colour_offsets[c]) {
            set_error("fl_df_assemble_cs: colour offsets decrease at colour %d", c);
            return FL_E_ARG;
        }
    const ColourPlan cp = {ncolour, colour_offsets, colour_elems, scatter};
    return df_assemble(material, params, elem_mat, mat_table, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, update, node_elem_ptr, node_elem, node_nbr_ptr, node_nbrs, node_order, bucket_offsets,
                       bucket_degree, indptr, V, T, precision, ranks, tensor_basis, &cp, workspace, workspace_bytes,
                       status, stream);
}

// the residual pass takes the factorised point kernel (tractions contracted in it, no
// point records) for tensor-product hexes given their basis table
static bool residual_sf(const void *tensor_basis, int npe, int G) {
    return tensor_basis != nullptr && sf_shape(npe, G) && tuning_value(TUNE_STIFF_SF, 0) != 1;
}

static int df_residual(int32_t material, const double *params, const int32_t *elem_mat, const double *mat_table,
                       const double *points, const int32_t *elements, const double *Eulerx, const double *Jm,
                       const double *AllGauss, int32_t G, int64_t nelem, int32_t npe, int64_t nnode,
                       const int32_t *node_elem_ptr, const int32_t *node_elem, double *T, const void *tensor_basis,
                       void *workspace, size_t workspace_bytes, int32_t *status, void *stream,
                       bool profile = true) {
    hipStream_t st = as_stream(stream);
    if (!check_common(material, params, G, nelem, npe, nnode)) return FL_E_ARG;
    if (!check_table(material, elem_mat, mat_table, "fl_df_residual")) return FL_E_ARG;
    if (!points || !elements || !Eulerx || !Jm || !AllGauss || !node_elem_ptr || !node_elem || !T || !status) {
        set_error("fl_df_residual: null pointer");
        return FL_E_ARG;
    }
    const bool sf_ok = residual_sf(tensor_basis, npe, G);
    size_t need = 0;
    DFWork w = carve_df(workspace, workspace_bytes, G, nelem, npe, false, &need, FL_PREC_FP64, false, !sf_ok);
    if (!w.JmT) {
        set_error("fl_df_residual: workspace too small (%zu < %zu)", workspace_bytes, need);
        return FL_E_SPACE;
    }
    const int NP = padded_np(npe);
    const MatParams mp = mat_params(material, params);
    if (profile) prof_begin(st, PROF_DF_RESIDUAL);
    clear_status_kernel<<<1, 64, 0, st>>>(status);
    if (sf_ok) {
        // no records: the tractions are contracted in the point kernel
        const int rq = launch_qp_sf<false>(mp, points, elements, Eulerx, tensor_basis, AllGauss, npe, G, nelem, 1,
                                           nullptr, status, elem_mat, mat_table, st, 0, w.te);
        if (rq) return rq;
    } else {
        transpose_jm_kernel<<<(G * 3 * NP + 255) / 256, 256, 0, st>>>(Jm, npe, G, NP, w.JmT);
        FL_LAUNCH_CHECK();
        const int64_t nq = nelem * G;
        if (general_law(material))
            qp_kernel<false, true><<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(
                mp, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe, 1, w.qp, status, elem_mat, mat_table);
        else
            qp_kernel<false, false><<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(
                mp, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe, 1, w.qp, status, elem_mat, mat_table);
        FL_LAUNCH_CHECK();
    }
    if (profile) prof_mark(st);
    if (!sf_ok) {
        const size_t lds_tr = (size_t)G * 9 * 8;
        if (lds_tr > 64 * 1024)
            FL_HIP_CHECK(hipFuncSetAttribute((const void *)traction_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds_tr));
        traction_kernel<<<(unsigned)nelem, 256, lds_tr, st>>>(w.JmT, NP, w.qp, QS_RES, QR_S, npe, G, w.te);
        FL_LAUNCH_CHECK();
    }
    if (profile) prof_mark(st);
    gather_traction_kernel<<<(unsigned)((3 * nnode + 255) / 256), 256, 0, st>>>(npe, node_elem_ptr, node_elem, w.te,
                                                                                 nnode, T);
    FL_LAUNCH_CHECK();
    if (profile) prof_mark(st);
    return 0;
}

int fl_df_residual(int32_t material, const double *params, const double *points, const int32_t *elements,
                   const double *Eulerx, const double *Jm, const double *AllGauss, int32_t G, int64_t nelem,
                   int32_t npe, int64_t nnode, const int32_t *node_elem_ptr, const int32_t *node_elem, double *T,
                   void *workspace, size_t workspace_bytes, int32_t *status, void *stream) {
    return df_residual(material, params, nullptr, nullptr, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, node_elem_ptr, node_elem, T, nullptr, workspace, workspace_bytes, status, stream);
}

int fl_df_residual_tb(int32_t material, const double *params, const double *points, const int32_t *elements,
                      const double *Eulerx, const double *Jm, const double *AllGauss, int32_t G, int64_t nelem,
                      int32_t npe, int64_t nnode, const int32_t *node_elem_ptr, const int32_t *node_elem, double *T,
                      const void *tensor_basis, void *workspace, size_t workspace_bytes, int32_t *status,
                      void *stream) {
    return df_residual(material, params, nullptr, nullptr, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, node_elem_ptr, node_elem, T, tensor_basis, workspace, workspace_bytes, status, stream);
}

int fl_df_residual_mt(int32_t material, const double *params, const int32_t *elem_mat, const double *mat_table,
                      const double *points, const int32_t *elements, const double *Eulerx, const double *Jm,
                      const double *AllGauss, int32_t G, int64_t nelem, int32_t npe, int64_t nnode,
                      const int32_t *node_elem_ptr, const int32_t *node_elem, double *T, const void *tensor_basis,
                      void *workspace, size_t workspace_bytes, int32_t *status, void *stream) {
    return df_residual(material, params, elem_mat, mat_table, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, node_elem_ptr, node_elem, T, tensor_basis, workspace, workspace_bytes, status, stream);
}

// fl_df_residual_noalloc: fl_df_residual_mt with its workspace in `scratch` below, for
// callers that cannot allocate (see scratch_address).
int fl_df_residual_noalloc(int32_t material, const double *params, const int32_t *elem_mat,
                           const double *mat_table, const double *points, const int32_t *elements,
                           const double *Eulerx, const double *Jm, const double *AllGauss, int32_t G, int64_t nelem,
                           int32_t npe, int64_t nnode, const int32_t *node_elem_ptr, const int32_t *node_elem,
                           double *T, const void *tensor_basis, int32_t *status, void *stream) {
    if (!check_common(material, params, G, nelem, npe, nnode)) return FL_E_ARG;
    size_t need = 0;
    carve_df(nullptr, 0, G, nelem, npe, false, &need, FL_PREC_FP64, false, !residual_sf(tensor_basis, npe, G));
    if (need > FL_NOALLOC_SCRATCH_BYTES) {
        set_error("fl_df_residual_noalloc: %lld elements of %d nodes at %d points need %zu bytes of scratch, above "
                  "the fixed maximum of %zu (use fl_df_residual_mt with a workspace)",
                  (long long)nelem, npe, G, need, (size_t)FL_NOALLOC_SCRATCH_BYTES);
        return FL_E_SPACE;
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    return df_residual(material, params, elem_mat, mat_table, points, elements, Eulerx, Jm, AllGauss, G, nelem, npe,
                       nnode, node_elem_ptr, node_elem, T, tensor_basis, ws, FL_NOALLOC_SCRATCH_BYTES, status, stream,
                       false);
}

// the form of the point records fl_df_assemble_* leaves in its workspace: the C tensors
// of the factorised FP64 path or dxi/dx and D of the dense path
static int records_form(int32_t precision, const void *tensor_basis, int npe, int G) {
    precision &= ~FL_PREC_CSR32;
    const bool sf = tensor_basis != nullptr && sf_shape(npe, G) && tuning_value(TUNE_STIFF_SF, 0) != 1;
    return sf && (precision == FL_PREC_FP64 || (precision == FL_PREC_FP32 && sf32_form())) ? FL_RECORDS_C
                                                                                          : FL_RECORDS_D;
}

int32_t fl_df_records_form(int32_t precision, const void *tensor_basis, int32_t npe, int32_t ngauss) {
    return records_form(precision, tensor_basis, npe, ngauss);
}

size_t fl_mg_galerkin_workspace_size(int64_t nelem, int32_t npe_c, int64_t nnode_c) {
    if (nelem <= 0 || npe_c <= 0 || nnode_c <= 0) return 0;
    Carve cv(nullptr, 0);
    cv.take<double>((size_t)nelem * 9 * npe_c * npe_c);
    cv.take<double>((size_t)nelem * 3 * npe_c);
    cv.take<double>((size_t)3 * nnode_c);
    cv.take<int32_t>((size_t)nelem * npe_c);
    return cv.off + 256;
}

int fl_mg_galerkin(const void *df_workspace, size_t df_workspace_bytes, int32_t G, int64_t nelem, int32_t npe,
                   int32_t records, const double *JmCt, const void *tensor_basis_c, int32_t npe_c,
                   const int32_t *elements_c, int64_t nnode_c,
                   const int32_t *node_elem_ptr_c, const int32_t *node_elem_c, const int32_t *node_nbr_ptr_c,
                   const int32_t *node_nbrs_c, const int32_t *node_order_c, const int32_t *bucket_offsets_c,
                   const int32_t *bucket_degree_c, const int32_t *indptr_c, const int16_t *ranks_c, double *V_c,
                   void *workspace, size_t workspace_bytes, int32_t *status, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!df_workspace || !JmCt || !elements_c || !node_elem_ptr_c || !node_elem_c || !node_nbr_ptr_c ||
        !node_nbrs_c || !node_order_c || !bucket_offsets_c || !bucket_degree_c || !indptr_c || !V_c || !status ||
        G <= 0 || nelem <= 0 || nnode_c <= 0 || padded_np(npe) < 0 ||
        (records != FL_RECORDS_C && records != FL_RECORDS_D)) {
        set_error("fl_mg_galerkin: invalid argument");
        return FL_E_ARG;
    }
    if (npe_c != 4 && npe_c != 8 && npe_c != 27) {
        set_error("fl_mg_galerkin: coarse elements of %d nodes are not supported (4, 8 or 27)", npe_c);
        return FL_E_ARG;
    }
    if (bucket_offsets_c[0] != 0 || bucket_offsets_c[FL_GATHER_BUCKETS] != nnode_c) {
        set_error("fl_mg_galerkin: gather plan does not cover the %lld coarse nodes", (long long)nnode_c);
        return FL_E_ARG;
    }
    size_t need = 0;
    DFWork fw = carve_df(const_cast<void *>(df_workspace), df_workspace_bytes, G, nelem, npe, true, &need,
                         FL_PREC_FP64, false);
    if (!fw.JmT) {
        set_error("fl_mg_galerkin: the fine workspace (%zu bytes) holds no point records for these sizes",
                  df_workspace_bytes);
        return FL_E_SPACE;
    }
    Carve cv(workspace, workspace_bytes);
    double *Kc = cv.take<double>((size_t)nelem * 9 * npe_c * npe_c);
    double *te = cv.take<double>((size_t)nelem * 3 * npe_c);
    double *Tdum = cv.take<double>((size_t)3 * nnode_c);
    int32_t *slot = cv.take<int32_t>((size_t)nelem * npe_c);
    if (!cv.ok) {
        set_error("fl_mg_galerkin: workspace too small (%zu < %zu)", workspace_bytes, cv.off + 256);
        return FL_E_SPACE;
    }
    for (int b = 0; b < FL_GATHER_BUCKETS; ++b) {
        if (bucket_offsets_c[b + 1] > bucket_offsets_c[b] &&
            gather_lds_bytes(gather_form(b, bucket_degree_c[b], npe_c), bucket_degree_c[b], npe_c) > 160 * 1024) {
            set_error("fl_mg_galerkin: coarse node degree %d needs more LDS than a CU has", bucket_degree_c[b]);
            return FL_E_LIMIT;
        }
    }
    clear_status_kernel<<<1, 64, 0, st>>>(status);
    const int64_t nte = nelem * 3 * npe_c;
    zero_kernel<<<(unsigned)((nte + 255) / 256), 256, 0, st>>>(te, nte);
    FL_LAUNCH_CHECK();
    const int form_c = records == FL_RECORDS_C ? 1 : 0;
    // a tensor-product coarse space over C records: the factorised stiffness kernel with
    // the coarse level's basis table (blocks in its tensor node order, as the fine blocks)
    const int32_t *kperm = nullptr, *kinv = nullptr, *eslot = nullptr;
    if (tensor_basis_c && form_c && galerkin_sf_shape(npe_c, G)) {
        const int32_t *tb = static_cast<const int32_t *>(tensor_basis_c);
        const int rc = launch_galerkin_sf(npe_c, G, fw.qp, tb, nelem, Kc, st);
        if (rc) return rc;
        kperm = tb + 4 + G + npe_c;   // tensor_of_l
        kinv = tb + 4 + G;            // local_of_t
        const int64_t ne = nelem * npe_c;
        entry_slot_kernel<<<(unsigned)((ne + 255) / 256), 256, 0, st>>>(node_elem_c, ne, npe_c, kperm, slot);
        FL_LAUNCH_CHECK();
        eslot = slot;
    } else if (npe_c == 4) {
        mg_galerkin_kernel<4><<<(unsigned)nelem, 256, 0, st>>>(fw.qp, G, form_c, JmCt, nelem, Kc);
    } else if (npe_c == 8) {
        mg_galerkin_kernel<8><<<(unsigned)nelem, 256, 0, st>>>(fw.qp, G, form_c, JmCt, nelem, Kc);
    } else {
        mg_galerkin_kernel<27><<<(unsigned)nelem, 256, 0, st>>>(fw.qp, G, form_c, JmCt, nelem, Kc);
    }
    FL_LAUNCH_CHECK();
    // the coarse CSR: owner-computes rows summed in ascending element order, as the fine gather
    for (int b = 0; b < FL_GATHER_BUCKETS; ++b) {
        const int32_t nb = bucket_offsets_c[b + 1] - bucket_offsets_c[b];
        if (nb <= 0) continue;
        const int rc = launch_gather<double, double>(b, nb, elements_c, npe_c, node_order_c + bucket_offsets_c[b],
                                                     node_elem_ptr_c, node_elem_c, node_nbr_ptr_c, node_nbrs_c,
                                                     indptr_c, Kc, te, bucket_degree_c[b], ranks_c, kperm, kinv,
                                                     eslot, V_c, Tdum, status, st);
        if (rc) return rc;
    }
    return 0;
}

int32_t fl_stiffness_variant_count(int32_t np_class) {
    static const int32_t counts[5] = {1, 4, 3, 5, 3};
    return np_class >= 0 && np_class < 5 ? counts[np_class] : 0;
}

int32_t fl_material_table_stride(int32_t material) {
    if (material < FL_MAT_NEOHOOKEAN || material > FL_MAT_ANISOTROPICMOONEYRIVLIN_1) return 0;
    return material_table_stride(material);
}

}  // extern "C"
