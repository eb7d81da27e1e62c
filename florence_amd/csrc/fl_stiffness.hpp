// Phase 2 of the DF assembly: element stiffness blocks on MFMA (included by
// fl_assemble.hip).
//
// One workgroup per (element, displacement block pair i <= k).  With G_{a,(g,j)} =
// dN_a/dx_j at point g and D^{ik}_{jl}(g) = detJ_g (c_ijkl + delta_ik sigma_jl), the
// constitutive + geometric stiffness of the reference
// (_ConstitutiveStiffnessDF_.h:83-134 and _GeometricStiffness_.h:61-100) is, per block,
//
//     K^{ik}_{ab} = sum_{(g,j)} G_{a,(g,j)} W^{ik}_{b,(g,j)},
//     W^{ik}_{b,(g,j)} = sum_l D^{ik}_{jl}(g) G_{b,(g,l)},
//
// a dense npe x npe x 3G product.  The inner index runs in chunks of 8 points (24
// indices).  Per chunk the workgroup builds G for all nodes in LDS (from the padded
// parent gradients JmT and the per-point inverse Jacobians, formed in FP64), and each
// lane forms its W operand on the fly from three G values and one row of D while the
// MFMAs run.  LDS is double buffered and the next chunk's global loads are issued
// before the current chunk's MFMAs, so one barrier per chunk remains.
//
// CT is the contraction type: double -> v_mfma_f64_16x16x4f64 (the FP64 path);
// float -> v_mfma_f32_16x16x4f32 (the FP32 tangent of the mixed-precision form, FP32
// accumulation).  KT is the storage type of the element blocks.
#pragma once

namespace fl {

typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename CT>
struct Mfma16;

template <>
struct Mfma16<double> {
    typedef double4_t acc_t;
    __device__ static acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output layout of the f64 16x16 MFMA: lane holds rows lk + 4r of column li
    __device__ static int row(int lk, int r) { return lk + 4 * r; }
};

template <>
struct Mfma16<float> {
    typedef float4_t acc_t;
    __device__ static acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // output layout of the f32 16x16 MFMA: lane holds rows 4 lk + r of column li
    __device__ static int row(int lk, int r) { return 4 * lk + r; }
};

// Triangular split of a 128-node diagonal block: wave W accumulates tile columns W and
// 7 - W over the rows up to the column index (W + 1 and 8 - W tiles: 9), over one chunk
// of KC inner indices.  Splitting by columns keeps the costly operand (W, formed from
// three G values and a D row) at two fragments per step; the cheap one (G rows) is a
// single LDS read per tile row.  acc[0..1][*] hold column W's tiles (row r at
// [r / 4][r % 4]), acc[2..3][*] column 7 - W's.
template <int W, int KC, int LDA, typename MF, typename CT, typename acc_t>
__device__ __forceinline__ void tri_chunk(const CT *__restrict__ A, const CT *__restrict__ Dc, int li, int lk,
                                          acc_t (&acc)[4][4]) {
    constexpr int C0 = W, C1 = 7 - W;
#pragma unroll 1
    for (int kk = 0; kk < KC / 4; ++kk) {
        const int q = kk * 4 + lk;
        const int gq = q / 3, j = q - 3 * gq;
        const CT d0 = Dc[gq * 9 + 3 * j], d1 = Dc[gq * 9 + 3 * j + 1], d2 = Dc[gq * 9 + 3 * j + 2];
        const CT *Gb = A + 3 * gq * LDA;
        const int b0 = C0 * 16 + li, b1 = C1 * 16 + li;
        const CT w0 = d0 * Gb[b0] + d1 * Gb[LDA + b0] + d2 * Gb[2 * LDA + b0];
        const CT w1 = d0 * Gb[b1] + d1 * Gb[LDA + b1] + d2 * Gb[2 * LDA + b1];
#pragma unroll
        for (int r = 0; r <= C1; ++r) {
            const CT a = A[q * LDA + r * 16 + li];
            if (r <= C0) acc[r / 4][r % 4] = MF::step(a, w0, acc[r / 4][r % 4]);
            acc[2 + r / 4][r % 4] = MF::step(a, w1, acc[2 + r / 4][r % 4]);
        }
    }
}

// PS selects the block pairs of the launch: 0 = all six, 1 = the three off-diagonal
// pairs (i < k), 2 = the three diagonal pairs (i = k).
template <typename CT, typename KT, int NP, int MT, int NT, int WGM, int WGN, int MINW = 1, int PS = 0>
__global__ __launch_bounds__(64 * WGM *WGN, MINW) void stiffness_kernel(const double *__restrict__ JmT,
                                                                   const double *__restrict__ qp, int npe, int G,
                                                                   int64_t nelem, KT *__restrict__ Ke) {
    static_assert(NP == 16 * MT * WGM && NP == 16 * NT * WGN, "tile shape must cover NP");
    typedef Mfma16<CT> MF;
    typedef typename MF::acc_t acc_t;
    constexpr int NTH = 64 * WGM * WGN;
    constexpr int QC = 8;          // quadrature points per chunk
    constexpr int KC = 3 * QC;     // inner (g, j) indices per chunk
    constexpr int LDA = NP + 4;    // LDS row stride (elements)
    constexpr int TPG = NTH / QC;  // fill threads per point
    constexpr int IPT = NP / TPG;  // nodes filled per thread
    static_assert(NTH % QC == 0 && NP % TPG == 0, "fill mapping");
    constexpr int PDN = (QC * 9 + NTH - 1) / NTH;   // D values staged per thread
    constexpr int TT = 16 * NT * 17;                // per-wave transpose tile (elements)
    static_assert(WGM * WGN * TT <= 2 * KC * LDA, "transpose tile must fit the staging buffers");
    __shared__ CT sA[2 * KC * LDA];
    __shared__ CT sD[2 * QC * 9];

    // XCD-aware order: workgroup ids are dealt round-robin over the 8 XCDs; the six
    // block pairs of one element land on the same XCD and share its L2.
    constexpr int NPAIR = PS == 0 ? 6 : 3;
    const int64_t wid = blockIdx.x;
    const int xcd = (int)(wid & 7);
    const int64_t slot = wid >> 3;
    const int pair = (int)(slot % NPAIR) + (PS == 1 ? 3 : 0);
    const int64_t e = (slot / NPAIR) * 8 + xcd;
    if (e >= nelem) return;
    const int pi = c_pair_i[pair], pk = c_pair_k[pair];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave / WGN, wn = wave % WGN;
    const int li = lane & 15, lk = lane >> 4;
    const double *qe = qp + e * (int64_t)G * QS;
    // diagonal blocks K^{ii} are symmetric: a wave tile strictly below the diagonal is
    // skipped and written as the transpose of its mirror tile
    const bool diag = pi == pk;
    // 128-node diagonal blocks (PS = 2): the upper triangle of the 8 x 8 grid of 16 x 16
    // tiles (36 tiles) is dealt to the four waves as column pairs {w, 7 - w} (w + 1 and
    // 8 - w tiles: 9 each), so these workgroups do 9/16 of an off-diagonal block's MFMAs
    constexpr bool TRI = PS == 2 && NP == 128 && WGM == 2 && WGN == 2 && MT == 4 && NT == 4;
    constexpr bool tri = TRI;
    const bool skip = !tri && WGM == WGN && diag && wm > wn;
    const bool mirror = !diag || (WGM == WGN && wm < wn);
    const int R0 = wave, R1 = 7 - wave;   // tile columns of the triangular split

    // fill role: point fg of the chunk, nodes fa0 + TPG*s
    const int fg = tid / TPG, fa0 = tid % TPG;
    double pj[IPT][3], pix[9], pd[PDN];
    auto prefetch = [&](int g0) {
        const int g = g0 + fg;
        if (g < G) {
            const double *jt = JmT + (int64_t)g * 3 * NP + fa0;
#pragma unroll
            for (int s = 0; s < IPT; ++s) {
                pj[s][0] = jt[TPG * s];
                pj[s][1] = jt[TPG * s + NP];
                pj[s][2] = jt[TPG * s + 2 * NP];
            }
#pragma unroll
            for (int m = 0; m < 9; ++m) pix[m] = qe[(int64_t)(Q_IX + m) * G + g];
        } else {
#pragma unroll
            for (int s = 0; s < IPT; ++s) pj[s][0] = pj[s][1] = pj[s][2] = 0.0;
#pragma unroll
            for (int m = 0; m < 9; ++m) pix[m] = 0.0;
        }
#pragma unroll
        for (int s = 0; s < PDN; ++s) {
            const int t = tid + s * NTH;
            const int gq = t / 9, f = t - 9 * gq, gg = g0 + gq;
            pd[s] = (t < QC * 9 && gg < G) ? qe[(int64_t)(Q_D + 9 * pair + f) * G + gg] : 0.0;
        }
    };
    auto commit = [&](int buf) {
        CT *A = sA + buf * KC * LDA;
#pragma unroll
        for (int s = 0; s < IPT; ++s) {
            const int a = fa0 + TPG * s;
            A[(3 * fg + 0) * LDA + a] = (CT)(pix[0] * pj[s][0] + pix[1] * pj[s][1] + pix[2] * pj[s][2]);
            A[(3 * fg + 1) * LDA + a] = (CT)(pix[3] * pj[s][0] + pix[4] * pj[s][1] + pix[5] * pj[s][2]);
            A[(3 * fg + 2) * LDA + a] = (CT)(pix[6] * pj[s][0] + pix[7] * pj[s][1] + pix[8] * pj[s][2]);
        }
#pragma unroll
        for (int s = 0; s < PDN; ++s)
            if (tid + s * NTH < QC * 9) sD[buf * QC * 9 + tid + s * NTH] = (CT)pd[s];
    };

    acc_t acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = acc_t{0, 0, 0, 0};

    prefetch(0);
    commit(0);
    __syncthreads();
    int buf = 0;
    for (int g0 = 0; g0 < G; g0 += QC) {
        const bool more = g0 + QC < G;
        if (more) prefetch(g0 + QC);
        const CT *A = sA + buf * KC * LDA;
        const CT *Dc = sD + buf * QC * 9;
        // not unrolled: each step already issues MT*NT long MFMAs, and hoisting the LDS
        // operands of several steps would cost the registers of a second wave per SIMD
        if constexpr (TRI) {
            // one straight-line body per wave (tile rows are compile-time there)
            switch (wave) {
                case 0: tri_chunk<0, KC, LDA, MF>(A, Dc, li, lk, acc); break;
                case 1: tri_chunk<1, KC, LDA, MF>(A, Dc, li, lk, acc); break;
                case 2: tri_chunk<2, KC, LDA, MF>(A, Dc, li, lk, acc); break;
                default: tri_chunk<3, KC, LDA, MF>(A, Dc, li, lk, acc); break;
            }
        } else {
#pragma unroll 1
        for (int kk = 0; kk < (skip ? 0 : KC / 4); ++kk) {
            const int q = kk * 4 + lk;
            const int gq = q / 3, j = q - 3 * gq;
            const CT d0 = Dc[gq * 9 + 3 * j], d1 = Dc[gq * 9 + 3 * j + 1], d2 = Dc[gq * 9 + 3 * j + 2];
            const CT *Gb = A + 3 * gq * LDA;
            CT af[MT], bf[NT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) af[mt] = A[q * LDA + (wm * MT + mt) * 16 + li];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int b = (wn * NT + nt) * 16 + li;
                bf[nt] = d0 * Gb[b] + d1 * Gb[LDA + b] + d2 * Gb[2 * LDA + b];
            }
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = MF::step(af[mt], bf[nt], acc[mt][nt]);
        }
        }
        if (more) commit(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    // epilogue: K^{ik} rows a, cols b.
    const int64_t npe2 = (int64_t)npe * npe;
    KT *Kd = Ke + (e * 9 + pi * 3 + pk) * npe2;
    if (tri) {
        // tile (t, C), t <= C, and its mirror (C, t) inside the same symmetric block
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int C = h ? R1 : R0;
                if (t <= C) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int a = t * 16 + MF::row(lk, r), b = C * 16 + li;
                        const KT v = (KT)acc[2 * h + t / 4][t % 4][r];
                        if (a < npe && b < npe) {
                            Kd[(int64_t)a * npe + b] = v;
                            if (t < C) Kd[(int64_t)b * npe + a] = v;
                        }
                    }
                }
            }
        return;
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = (wm * MT + mt) * 16 + MF::row(lk, r);
                const int b = (wn * NT + nt) * 16 + li;
                if (!skip && a < npe && b < npe) Kd[(int64_t)a * npe + b] = (KT)acc[mt][nt][r];
            }
    if (!diag || WGM * WGN > 1) {
        // transposed tiles (K^{ki}_{ba}, or the skipped half of K^{ii}) through a per-wave
        // LDS transpose for coalesced stores; barriers are workgroup-uniform
        KT *Kt = Ke + (e * 9 + pk * 3 + pi) * npe2;
        CT *tile = sA + wave * TT;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            if (mirror) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) tile[(nt * 16 + li) * 17 + MF::row(lk, r)] = acc[mt][nt][r];
            }
            __syncthreads();
            if (mirror) {
                for (int idx = lane; idx < 16 * NT * 16; idx += 64) {
                    const int bl = idx >> 4, al = idx & 15;
                    const int b = wn * NT * 16 + bl, a = (wm * MT + mt) * 16 + al;
                    if (a < npe && b < npe) Kt[(int64_t)b * npe + a] = (KT)tile[bl * 17 + al];
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace fl
