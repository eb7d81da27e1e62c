// Phase 2, BF16 form: element stiffness blocks on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation (included by fl_assemble.hip).  Same block formulation as
// fl_stiffness.hpp:
//
//     K^{ik}_{ab} = sum_{(g,j)} G_{a,(g,j)} W^{ik}_{b,(g,j)},
//     W^{ik}_{b,(g,j)} = sum_l D^{ik}_{jl}(g) G_{b,(g,l)}.
//
// G and W are formed in FP64 from the per-point records (kinematics stay FP64) and
// rounded once to BF16 into LDS, node-major with the inner index contiguous, so each
// lane's 8-wide A and B operands are one 16-byte LDS read.  32 points (96 inner
// indices = 3 MFMA steps of 32) per chunk.
#pragma once

namespace fl {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <int NP, int MT, int NT, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM *WGN) void stiffness_bf16_kernel(const double *__restrict__ JmT,
                                                                        const double *__restrict__ qp, int npe, int G,
                                                                        int64_t nelem, float *__restrict__ Ke) {
    static_assert(NP == 16 * MT * WGM && NP == 16 * NT * WGN, "tile shape must cover NP");
    constexpr int NTH = 64 * WGM * WGN;
    constexpr int QC = 32;         // quadrature points per chunk
    constexpr int KC = 3 * QC;     // 96 inner indices = 3 MFMA k-steps
    constexpr int LDK = KC + 8;    // bf16 row stride: 208 B, a multiple of 16 B
    constexpr int TT = 16 * NT * 17;
    static_assert(WGM * WGN * TT * 4 <= NP * LDK * 2, "transpose tile must fit the G staging buffer");
    __shared__ __attribute__((aligned(16))) __bf16 sG[NP * LDK];
    __shared__ __attribute__((aligned(16))) __bf16 sW[NP * LDK];
    __shared__ double sQ[QC * 18];   // per point: ix (9), D^{ik} (9)

    const int64_t wid = blockIdx.x;
    const int xcd = (int)(wid & 7);
    const int64_t slot = wid >> 3;
    const int pair = (int)(slot % 6);
    const int64_t e = (slot / 6) * 8 + xcd;
    if (e >= nelem) return;
    const int pi = c_pair_i[pair], pk = c_pair_k[pair];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave / WGN, wn = wave % WGN;
    const int li = lane & 15, lk = lane >> 4;
    const double *qe = qp + e * (int64_t)G * QS;

    float4_t acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = float4_t{0.f, 0.f, 0.f, 0.f};

    for (int g0 = 0; g0 < G; g0 += QC) {
        for (int idx = tid; idx < QC * 18; idx += NTH) {
            const int gq = idx / 18, f = idx - 18 * gq, g = g0 + gq;
            double v = 0.0;
            if (g < G)
                v = f < 9 ? qe[(int64_t)(Q_IX + f) * G + g] : qe[(int64_t)(Q_D + 9 * pair + (f - 9)) * G + g];
            sQ[idx] = v;
        }
        __syncthreads();
        for (int idx = tid; idx < QC * NP; idx += NTH) {
            const int gq = idx / NP, a = idx - gq * NP, g = g0 + gq;
            double j0 = 0.0, j1 = 0.0, j2 = 0.0;
            if (g < G) {
                const double *jt = JmT + (int64_t)g * 3 * NP + a;
                j0 = jt[0];
                j1 = jt[NP];
                j2 = jt[2 * NP];
            }
            const double *ix = sQ + gq * 18, *D = ix + 9;
            const double G0 = ix[0] * j0 + ix[1] * j1 + ix[2] * j2;
            const double G1 = ix[3] * j0 + ix[4] * j1 + ix[5] * j2;
            const double G2 = ix[6] * j0 + ix[7] * j1 + ix[8] * j2;
            __bf16 *gr = sG + a * LDK + 3 * gq;
            __bf16 *wr = sW + a * LDK + 3 * gq;
            gr[0] = (__bf16)(float)G0;
            gr[1] = (__bf16)(float)G1;
            gr[2] = (__bf16)(float)G2;
            wr[0] = (__bf16)(float)(D[0] * G0 + D[1] * G1 + D[2] * G2);
            wr[1] = (__bf16)(float)(D[3] * G0 + D[4] * G1 + D[5] * G2);
            wr[2] = (__bf16)(float)(D[6] * G0 + D[7] * G1 + D[8] * G2);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KC / 32; ++ks) {
            const int q = 32 * ks + 8 * lk;
            bf16x8_t af[MT], bf[NT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                af[mt] = *reinterpret_cast<const bf16x8_t *>(sG + ((wm * MT + mt) * 16 + li) * LDK + q);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bf[nt] = *reinterpret_cast<const bf16x8_t *>(sW + ((wn * NT + nt) * 16 + li) * LDK + q);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mt], bf[nt], acc[mt][nt], 0, 0, 0);
        }
        __syncthreads();
    }

    // epilogue (f32 16x16 output layout: lane holds rows 4 lk + r of column li)
    const int64_t npe2 = (int64_t)npe * npe;
    float *Kd = Ke + (e * 9 + pi * 3 + pk) * npe2;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = (wm * MT + mt) * 16 + 4 * lk + r;
                const int b = (wn * NT + nt) * 16 + li;
                if (a < npe && b < npe) Kd[(int64_t)a * npe + b] = acc[mt][nt][r];
            }
    if (pi != pk) {
        float *Kt = Ke + (e * 9 + pk * 3 + pi) * npe2;
        float *tile = reinterpret_cast<float *>(sG) + wave * TT;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) tile[(nt * 16 + li) * 17 + 4 * lk + r] = acc[mt][nt][r];
            __syncthreads();
            for (int idx = lane; idx < 16 * NT * 16; idx += 64) {
                const int bl = idx >> 4, al = idx & 15;
                const int b = wn * NT * 16 + bl, a = (wm * MT + mt) * 16 + al;
                if (a < npe && b < npe) Kt[(int64_t)b * npe + a] = tile[bl * 17 + al];
            }
            __syncthreads();
        }
    }
}

}  // namespace fl
