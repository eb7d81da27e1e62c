/* C restatement of the reference's low-level displacement-formulation assembly --
 * TEST INFRASTRUCTURE (see oracle/__init__.py).  Used as the timed CPU baseline
 * ("port") and as a second restatement next to oracle/df_numpy.py.
 *
 * The loop nest is the reference's (Florence/FiniteElements/Assembly/_Assembly_/
 * _LowLevelAssemblyDF_.h:69-161): per element, gather X and x, then per quadrature
 * point the kinematics of _KinematicMeasures_.h:64-119, the Neo-Hookean or
 * Mooney-Rivlin kinetics of _NeoHookean_.h:27-51 / _MooneyRivlin_.h:30-70, the dense
 * B, HBT = H*B^T and BDB = B*HBT of _ConstitutiveStiffnessDF_.h:83-156 (the two
 * cblas_dgemm calls written as plain loops), the traction B*s, the geometric term of
 * _GeometricStiffness_.h:61-100 and finally the CSR accumulation with a binary search
 * of every column in its row (SparseAssemblyNative.h:49-111) and the traction
 * scatter.  Single threaded, like the reference kernel.
 *
 * Laws (material ids of include/florence_amd.h): 1 Neo-Hookean, 2 Mooney-Rivlin,
 * 3 linear elastic (_LinearElastic_.h:26-52; the generated assembler has no geometric
 * term, AOT_Assembler.py:84-93), 4 nearly incompressible Mooney-Rivlin
 * (_NearlyIncompressibleMooneyRivlin_.h:28-87, constructed from mu1, mu2, lamb as
 * alpha, beta, kappa, AOT_Assembler.py:69-70), 5 anisotropic Mooney-Rivlin 1
 * (_AnisotropicMooneyRivlin_1_.h:30-95, fibre direction orient[3 e], AOT_Assembler.py:78-82).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int VI[6] = {0, 1, 2, 0, 0, 1};
static const int VJ[6] = {0, 1, 2, 1, 2, 2};

static double invdet3x3(const double *s, double *d) {   /* Tensor/_det_inv_.h:149-190 */
    d[0] = +s[4] * s[8] - s[5] * s[7];
    d[1] = -s[1] * s[8] + s[2] * s[7];
    d[2] = +s[1] * s[5] - s[2] * s[4];
    d[3] = -s[3] * s[8] + s[5] * s[6];
    d[4] = +s[0] * s[8] - s[2] * s[6];
    d[5] = -s[0] * s[5] + s[2] * s[3];
    d[6] = +s[3] * s[7] - s[4] * s[6];
    d[7] = -s[0] * s[7] + s[1] * s[6];
    d[8] = +s[0] * s[4] - s[1] * s[3];
    double det = s[0] * d[0] + s[1] * d[3] + s[2] * d[6];
    double r = 1.0 / det;
    for (int i = 0; i < 9; ++i) d[i] *= r;
    return det;
}

static double kd(int a, int b) { return a == b ? 1.0 : 0.0; }

/* cofactor of F: H = J F^{-T}, H_ij = dJ/dF_ij */
static void cofactor3(const double *F, double *H) {
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

/* sigma (3x3 row-major) and Voigt H (6x6) at one quadrature point; N: fibre direction
 * (material 5 only). */
static void kinetics(int material, const double *prm, const double *N, const double *F, double *sig, double *H) {
    double b[9], J, bb[9];
    J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) + F[2] * (F[3] * F[7] - F[4] * F[6]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += F[i * 3 + k] * F[j * 3 + k];
            b[i * 3 + j] = s;
        }
    if (material == 3) {   /* linear elastic, _LinearElastic_.h:35-49 */
        double mu = prm[0], lamb = prm[3], e[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) e[i * 3 + j] = 0.5 * (F[i * 3 + j] + F[j * 3 + i]) - kd(i, j);
        double tre = e[0] + e[4] + e[8];
        for (int i = 0; i < 9; ++i) sig[i] = 2.0 * mu * e[i] + lamb * tre * kd(i / 3, i % 3);
        for (int A = 0; A < 6; ++A)
            for (int B = 0; B < 6; ++B) {
                int i = VI[A], j = VJ[A], k = VI[B], l = VJ[B];
                H[A * 6 + B] = mu * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k)) + lamb * kd(i, j) * kd(k, l);
            }
        return;
    }
    if (material == 4) {   /* nearly incompressible Mooney-Rivlin, _NearlyIncompressibleMooneyRivlin_.h:34-81 */
        double alpha = prm[1], beta = prm[2], kappa = prm[3], Hc[9], g[9];
        cofactor3(F, Hc);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += Hc[i * 3 + k] * Hc[j * 3 + k];
                g[i * 3 + j] = s;
            }
        double trb = b[0] + b[4] + b[8], trg = g[0] + g[4] + g[8];
        double c0 = pow(J, -5. / 3.), c1 = sqrt(trg), c2 = 1. / c1, c3 = trg * c1, c4 = 1. / (J * J * J);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                sig[i * 3 + j] = 2. * alpha * c0 * b[i * 3 + j] - 2. / 3. * alpha * c0 * trb * kd(i, j)
                                 + beta * c4 * c3 * kd(i, j) - 3 * beta * c4 * c1 * g[i * 3 + j] + kappa * (J - 1.0) * kd(i, j);
        for (int A = 0; A < 6; ++A)
            for (int B = 0; B < 6; ++B) {
                int i = VI[A], j = VJ[A], k = VI[B], l = VJ[B];
                double Iijkl = kd(i, j) * kd(k, l), Iikjl = kd(i, k) * kd(j, l), Iiljk = kd(i, l) * kd(j, k);
                H[A * 6 + B] = -4 / 3. * alpha * c0 * (b[i * 3 + j] * kd(k, l) + kd(i, j) * b[k * 3 + l])
                               + 4. * alpha / 9. * c0 * trb * Iijkl + 2 / 3. * alpha * c0 * trb * (Iikjl + Iiljk)
                               + beta * c4 * c3 * (Iijkl - Iikjl - Iiljk)
                               - 3. * beta * c4 * c1 * (g[i * 3 + j] * kd(k, l) + kd(i, j) * g[k * 3 + l])
                               + 3. * beta * c4 * c1 * (g[i * 3 + k] * kd(j, l) + g[i * 3 + l] * kd(j, k)
                                                        + kd(i, k) * g[j * 3 + l] + kd(i, l) * g[j * 3 + k])
                               + 3. * beta * c4 * c2 * g[i * 3 + j] * g[k * 3 + l]
                               + kappa * (2.0 * J - 1) * Iijkl - kappa * (J - 1) * (Iikjl + Iiljk);
            }
        return;
    }
    if (material == 5) {   /* anisotropic Mooney-Rivlin 1, _AnisotropicMooneyRivlin_1_.h:40-89 */
        double mu1 = prm[1], mu2 = prm[2], lamb = prm[3], mu3 = prm[4], Hc[9], FN[3], HN[3];
        cofactor3(F, Hc);
        for (int i = 0; i < 3; ++i) {
            FN[i] = F[i * 3] * N[0] + F[i * 3 + 1] * N[1] + F[i * 3 + 2] * N[2];
            HN[i] = Hc[i * 3] * N[0] + Hc[i * 3 + 1] * N[1] + Hc[i * 3 + 2] * N[2];
        }
        double iHN = HN[0] * HN[0] + HN[1] * HN[1] + HN[2] * HN[2], trb = b[0] + b[4] + b[8];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += b[i * 3 + k] * b[k * 3 + j];
                bb[i * 3 + j] = s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                sig[i * 3 + j] = 2. * mu1 / J * b[i * 3 + j] + 2. * mu2 / J * (trb * b[i * 3 + j] - bb[i * 3 + j])
                                 - 2. * (mu1 + 2. * mu2 + mu3) / J * kd(i, j) + lamb * (J - 1.) * kd(i, j)
                                 + 2. * mu3 / J * FN[i] * FN[j] + 2. * mu3 / J * iHN * kd(i, j)
                                 - 2. * mu3 / J * HN[i] * HN[j];
        for (int A = 0; A < 6; ++A)
            for (int B = 0; B < 6; ++B) {
                int i = VI[A], j = VJ[A], k = VI[B], l = VJ[B];
                double Iijkl = kd(i, j) * kd(k, l), Iikjl = kd(i, k) * kd(j, l), Iiljk = kd(i, l) * kd(j, k);
                H[A * 6 + B] = 2.0 * mu2 / J * (2.0 * b[i * 3 + j] * b[k * 3 + l] - b[i * 3 + k] * b[j * 3 + l]
                                                - b[i * 3 + l] * b[j * 3 + k])
                               + (2. * (mu1 + 2 * mu2 + mu3) / J - lamb * (J - 1.)) * (Iikjl + Iiljk)
                               + lamb * (2. * J - 1.) * Iijkl
                               - 4. * mu3 / J * (kd(i, j) * HN[k] * HN[l] + HN[i] * HN[j] * kd(k, l))
                               + 2. * mu3 / J * iHN * (2.0 * Iijkl - Iikjl - Iiljk)
                               + 2. * mu3 / J * (kd(i, k) * HN[j] * HN[l] + kd(i, l) * HN[j] * HN[k]
                                                 + kd(j, l) * HN[i] * HN[k] + kd(j, k) * HN[i] * HN[l]);
            }
        return;
    }
    if (material == 1) {   /* Neo-Hookean, _NeoHookean_.h:35-46 */
        double mu = prm[0], lamb = prm[3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                sig[i * 3 + j] = mu / J * (b[i * 3 + j] - kd(i, j)) + lamb * (J - 1.0) * kd(i, j);
        double a1 = mu / J - lamb * (J - 1.0), a2 = lamb * (2.0 * J - 1.0);
        for (int A = 0; A < 6; ++A)
            for (int B = 0; B < 6; ++B) {
                int i = VI[A], j = VJ[A], k = VI[B], l = VJ[B];
                H[A * 6 + B] = a1 * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k)) + a2 * kd(i, j) * kd(k, l);
            }
    } else {               /* Mooney-Rivlin, _MooneyRivlin_.h:120-147 */
        double mu1 = prm[1], mu2 = prm[2], lamb = prm[3];
        double trb = b[0] + b[4] + b[8];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += b[i * 3 + k] * b[k * 3 + j];
                bb[i * 3 + j] = s;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                sig[i * 3 + j] = 2. * mu1 / J * b[i * 3 + j] + 2. * mu2 / J * (trb * b[i * 3 + j] - bb[i * 3 + j])
                                 - 2. * (mu1 + 2 * mu2) / J * kd(i, j) + lamb * (J - 1) * kd(i, j);
        double c1 = 2.0 * mu2 / J, a1 = 2. * (mu1 + 2 * mu2) / J - lamb * (J - 1.), a2 = lamb * (2. * J - 1.);
        for (int A = 0; A < 6; ++A)
            for (int B = 0; B < 6; ++B) {
                int i = VI[A], j = VJ[A], k = VI[B], l = VJ[B];
                H[A * 6 + B] = c1 * (2.0 * b[i * 3 + j] * b[k * 3 + l] - b[i * 3 + k] * b[j * 3 + l] - b[i * 3 + l] * b[j * 3 + k])
                               + a1 * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k)) + a2 * kd(i, j) * kd(k, l);
            }
    }
}

static int64_t locate(const int32_t *row, int64_t n, int32_t val) {   /* binary search, sorted row */
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int64_t mid = lo + (hi - lo) / 2;
        if (val < row[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

/* Assemble elements [e0, e1) into CSR data V (accumulated) and T (accumulated).
 * material: 1..5 as above; prm = {mu, mu1, mu2, lamb, mu3}; orient: f64[nelem][3] fibre
 * directions (material 5; NULL otherwise).
 * Returns 0, or -1 when a column is missing from the pattern. */
int oracle_df_assemble_ex(const double *points, const int64_t *elements, const double *Eulerx,
                          const double *Jm, const double *AllGauss, int64_t ngauss, int64_t npe,
                          int material, const double *prm, const double *orient, int update,
                          const int32_t *indptr, const int32_t *indices, double *V, double *T,
                          int64_t e0, int64_t e1) {
    if (material == 5 && !orient) return -2;
    const int64_t G = ngauss, n = 3 * npe;
    double *X = malloc(sizeof(double) * npe * 3), *x = malloc(sizeof(double) * npe * 3);
    double *SG = malloc(sizeof(double) * G * npe * 3), *Fq = malloc(sizeof(double) * G * 9);
    double *detJ = malloc(sizeof(double) * G), *S = malloc(sizeof(double) * G * 9), *Hq = malloc(sizeof(double) * G * 36);
    double *B = malloc(sizeof(double) * n * 6), *HBT = malloc(sizeof(double) * 6 * n);
    double *K = malloc(sizeof(double) * n * n), *Kg = malloc(sizeof(double) * n * n), *t = malloc(sizeof(double) * n), *MG = malloc(sizeof(double) * 3 * npe);
    int64_t *dof = malloc(sizeof(int64_t) * n);
    int status = 0;
    for (int64_t e = e0; e < e1; ++e) {
        for (int64_t a = 0; a < npe; ++a) {
            int64_t nd = elements[e * npe + a];
            for (int d = 0; d < 3; ++d) {
                X[a * 3 + d] = points[nd * 3 + d];
                x[a * 3 + d] = Eulerx[nd * 3 + d];
                dof[a * 3 + d] = nd * 3 + d;
            }
        }
        /* kinematics */
        for (int64_t g = 0; g < G; ++g) {
            double PX[9] = {0}, Px[9] = {0}, iX[9], ix[9], Ft[9];
            for (int i = 0; i < 3; ++i)
                for (int64_t a = 0; a < npe; ++a) {
                    double jm = Jm[i * G * npe + a * G + g];
                    for (int l = 0; l < 3; ++l) {
                        PX[i * 3 + l] += jm * X[a * 3 + l];
                        Px[i * 3 + l] += jm * x[a * 3 + l];
                    }
                }
            double detX = invdet3x3(PX, iX), detx = invdet3x3(Px, ix);
            detJ[g] = AllGauss[g] * fabs(update == 1 ? detx : detX);
            for (int j = 0; j < 3; ++j)
                for (int64_t a = 0; a < npe; ++a) {
                    double m = 0, s = 0;
                    for (int k = 0; k < 3; ++k) {
                        double jm = Jm[k * G * npe + a * G + g];
                        m += iX[j * 3 + k] * jm;
                        s += ix[j * 3 + k] * jm;
                    }
                    MG[j * npe + a] = m;
                    SG[g * npe * 3 + a * 3 + j] = s;
                }
            for (int j = 0; j < 3; ++j)
                for (int l = 0; l < 3; ++l) {
                    double s = 0;
                    for (int64_t a = 0; a < npe; ++a) s += MG[j * npe + a] * x[a * 3 + l];
                    Ft[j * 3 + l] = s;
                }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Fq[g * 9 + i * 3 + j] = Ft[j * 3 + i];
            kinetics(material, prm, orient ? orient + 3 * e : NULL, Fq + g * 9, S + g * 9, Hq + g * 36);
        }
        memset(K, 0, sizeof(double) * n * n);
        memset(t, 0, sizeof(double) * n);
        memset(B, 0, sizeof(double) * n * 6);
        for (int64_t g = 0; g < G; ++g) {
            const double *sg = SG + g * npe * 3, *H = Hq + g * 36, *s = S + g * 9;
            for (int64_t a = 0; a < npe; ++a) {   /* FillConstitutiveB_ */
                double a0 = sg[a * 3], a1 = sg[a * 3 + 1], a2 = sg[a * 3 + 2];
                double *r = B + a * 3 * 6;
                r[0] = a0; r[6 + 1] = a1; r[12 + 2] = a2;
                r[6 + 5] = a2; r[12 + 5] = a1;
                r[4] = a2; r[12 + 4] = a0;
                r[3] = a1; r[6 + 3] = a0;
            }
            for (int A = 0; A < 6; ++A)          /* HBT = H * B^T */
                for (int64_t m = 0; m < n; ++m) {
                    double v = 0;
                    for (int C = 0; C < 6; ++C) v += H[A * 6 + C] * B[m * 6 + C];
                    HBT[A * n + m] = v;
                }
            const double dj = detJ[g];
            for (int64_t r = 0; r < n; ++r)      /* K += detJ * B * HBT */
                for (int64_t c = 0; c < n; ++c) {
                    double v = 0;
                    for (int A = 0; A < 6; ++A) v += B[r * 6 + A] * HBT[A * n + c];
                    K[r * n + c] += v * dj;
                }
            if (update == 1) {
                double tv[6] = {s[0], s[4], s[8], s[1], s[2], s[5]};
                for (int64_t r = 0; r < n; ++r) {
                    double v = 0;
                    for (int A = 0; A < 6; ++A) v += B[r * 6 + A] * tv[A];
                    t[r] += v * dj;
                }
            }
        }
        /* geometric stiffness, added after the constitutive term (none for material 3) */
        memset(Kg, 0, sizeof(double) * n * n);
        for (int64_t g = 0; g < (material == 3 ? 0 : G); ++g) {
            const double *sg = SG + g * npe * 3, *s = S + g * 9;
            const double dj = detJ[g];
            for (int64_t a = 0; a < npe; ++a)
                for (int64_t b = 0; b < npe; ++b) {
                    double a0 = sg[a * 3], a1 = sg[a * 3 + 1], a2 = sg[a * 3 + 2];
                    double b0 = sg[b * 3], b1 = sg[b * 3 + 1], b2 = sg[b * 3 + 2];
                    double dum = a0 * (s[0] * b0 + s[1] * b1 + s[2] * b2) + a1 * (s[1] * b0 + s[4] * b1 + s[5] * b2)
                                 + a2 * (s[2] * b0 + s[5] * b1 + s[8] * b2);
                    for (int d = 0; d < 3; ++d) Kg[(a * 3 + d) * n + b * 3 + d] += dum * dj;
                }
        }
        for (int64_t i = 0; i < n * n; ++i) K[i] += Kg[i];
        /* CSR accumulation and traction scatter */
        for (int64_t r = 0; r < n; ++r) {
            int64_t gr = dof[r];
            const int32_t *row = indices + indptr[gr];
            int64_t len = indptr[gr + 1] - indptr[gr];
            for (int64_t c = 0; c < n; ++c) {
                int64_t p = locate(row, len, (int32_t)dof[c]);
                if (row[p] != dof[c]) status = -1;
                V[indptr[gr] + p] += K[r * n + c];
            }
            T[gr] += t[r];
        }
    }
    free(X); free(x); free(SG); free(Fq); free(detJ); free(S); free(Hq);
    free(B); free(HBT); free(K); free(Kg); free(t); free(MG); free(dof);
    return status;
}

/* Neo-Hookean / Mooney-Rivlin / isotropic laws without orientations. */
int oracle_df_assemble(const double *points, const int64_t *elements, const double *Eulerx,
                       const double *Jm, const double *AllGauss, int64_t ngauss, int64_t npe,
                       int material, const double *prm, int update,
                       const int32_t *indptr, const int32_t *indices, double *V, double *T,
                       int64_t e0, int64_t e1) {
    return oracle_df_assemble_ex(points, elements, Eulerx, Jm, AllGauss, ngauss, npe, material, prm, NULL, update,
                                 indptr, indices, V, T, e0, e1);
}
