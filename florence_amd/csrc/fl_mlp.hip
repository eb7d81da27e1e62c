// Warm-start predictor: a 3-layer MLP (in -> 128 -> 128 -> out, tanh) with FP32 master
// weights and BF16 MFMA GEMMs (v_mfma_f32_16x16x32_bf16, FP32 accumulation), forward,
// backward (MSE loss) and Adam, all on the device.  No reference counterpart.
//
// Parameter block (FP32, contiguous):  W1[IP][H] b1[H] W2[H][H] b2[H] W3[H][OP] b3[OP]
// with IP = 32 (inputs padded), H = 128, OP = 32 (outputs padded).  Activations are
// row-major [N][width] FP32.  Reductions are two-pass in a fixed order: bitwise
// reproducible.
#include <cmath>

#include "fl_common.hpp"

namespace fl {

constexpr int MLP_IP = 32, MLP_H = 128, MLP_OP = 32;
constexpr int64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IP * MLP_H, OFF_W2 = OFF_B1 + MLP_H,
                  OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H * MLP_OP,
                  MLP_NPARAM = OFF_B3 + MLP_OP;

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));

// C[M][N] (row-major, ldc) = sum_k A(m,k) B(k,n) (+ bias[n]) (tanh if act), A(m,k) =
// A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].  K % 32 == 0.  Tile 64x64x32, 4 waves.
constexpr int GB = 64, GK = 32, GLD = GK + 8;
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const float *__restrict__ A, int64_t sam, int64_t sak,
                                                        const float *__restrict__ B, int64_t sbk, int64_t sbn,
                                                        float *__restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                        int64_t K, const float *__restrict__ bias, int act) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[GB * GLD];
    __shared__ __attribute__((aligned(16))) __bf16 sB[GB * GLD];   // [n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1, li = lane & 15, lk = lane >> 4;
    const int64_t m0 = (int64_t)blockIdx.x * GB, n0 = (int64_t)blockIdx.y * GB;
    f4_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = 0; k0 < K; k0 += GK) {
        for (int idx = tid; idx < GB * GK; idx += 256) {
            const int r = idx / GK, c = idx % GK;
            const int64_t m = m0 + r, k = k0 + c;
            sA[r * GLD + c] = (__bf16)(m < M ? A[m * sam + k * sak] : 0.f);
            const int64_t n = n0 + r;
            sB[r * GLD + c] = (__bf16)(n < N ? B[k * sbk + n * sbn] : 0.f);
        }
        __syncthreads();
        b8_t af[2], bf[2];
        for (int t = 0; t < 2; ++t) {
            af[t] = *reinterpret_cast<const b8_t *>(sA + (wm * 32 + t * 16 + li) * GLD + 8 * lk);
            bf[t] = *reinterpret_cast<const b8_t *>(sB + (wn * 32 + t * 16 + li) * GLD + 8 * lk);
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int64_t m = m0 + wm * 32 + i * 16 + 4 * lk + r;
                const int64_t n = n0 + wn * 32 + j * 16 + li;
                if (m < M && n < N) {
                    float v = acc[i][j][r] + (bias ? bias[n] : 0.f);
                    if (act) v = tanhf(v);
                    C[m * ldc + n] = v;
                }
            }
}

// dY = 2 (Y - T) / (nout N) on the first nout columns (rest 0); per-block loss partials
__global__ __launch_bounds__(256) void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T,
                                                       int64_t N, int nout, float *__restrict__ dY,
                                                       double *__restrict__ partial) {
    __shared__ double sh[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < N * MLP_OP) {
        const int c = (int)(i % MLP_OP);
        float g = 0.f;
        if (c < nout) {
            const float d = Y[i] - T[i];
            g = 2.f * d / (float)(nout * N);
            s = (double)d * d;
        }
        dY[i] = g;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void sum_partials_kernel(const double *__restrict__ partial, int64_t n, double scale,
                                    float *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(sh[0] * scale);
}

// dZ = dH * (1 - H^2)  (in place on dH)
__global__ void tanh_grad_kernel(const float *__restrict__ H, float *__restrict__ dH, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dH[i] *= (1.f - H[i] * H[i]);
}

// column sums of dZ [N][W] -> out[W]: per (column, row-chunk) partials, then a fixed-order sum
constexpr int CS_ROWS = 1024;
__global__ void colsum_partial_kernel(const float *__restrict__ dZ, int64_t N, int W, float *__restrict__ part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W) return;
    const int64_t r0 = (int64_t)blockIdx.y * CS_ROWS;
    const int64_t r1 = r0 + CS_ROWS < N ? r0 + CS_ROWS : N;
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += dZ[r * W + c];
    part[(int64_t)blockIdx.y * W + c] = s;
}

__global__ void colsum_final_kernel(const float *__restrict__ part, int64_t nchunk, int W, float *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W) return;
    float s = 0.f;
    for (int64_t q = 0; q < nchunk; ++q) s += part[q * W + c];
    out[c] = s;
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Per-node inputs of the warm-start predictor: reference coordinates, the previous
// displacement, the last increment's step (u1 - u2) and the load factor, each
// standardised with norm = {mean[10], inv_std[10]}; columns 10..31 are zero.  Feeding
// u1 - u2 rather than u2 keeps the step resolvable after BF16 rounding.
constexpr int WS_NIN = 10;
__global__ void ws_features_kernel(const double *__restrict__ X, const double *__restrict__ u1,
                                   const double *__restrict__ u2, double lam, int64_t nnode,
                                   const float *__restrict__ norm, float *__restrict__ F) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnode * MLP_IP) return;
    const int64_t n = i / MLP_IP;
    const int c = (int)(i % MLP_IP);
    double v = 0.0;
    if (c < 3) v = X[3 * n + c];
    else if (c < 6) v = u1[3 * n + c - 3];
    else if (c < 9) v = u1[3 * n + c - 6] - u2[3 * n + c - 6];
    else if (c == 9) v = lam;
    F[i] = c < WS_NIN ? (float)((v - (double)norm[c]) * (double)norm[WS_NIN + c]) : 0.f;
}

// u = free ? 2 u1 - u2 + (scale * Y + shift) : u_bc   (scale/shift = out_norm[0..2]/[3..5])
__global__ void ws_apply_kernel(const double *__restrict__ u1, const double *__restrict__ u2,
                                const float *__restrict__ Y, int64_t nnode, const float *__restrict__ out_norm,
                                const uint8_t *__restrict__ free_mask, const double *__restrict__ u_bc,
                                double *__restrict__ u) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= 3 * nnode) return;
    const int64_t n = d / 3;
    const int c = (int)(d % 3);
    const double corr = (double)out_norm[c] * (double)Y[n * MLP_OP + c] + (double)out_norm[3 + c];
    u[d] = free_mask[d] ? 2.0 * u1[d] - u2[d] + corr : u_bc[d];
}

static int gemm(const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                int64_t ldc, int64_t M, int64_t N, int64_t K, const float *bias, int act, hipStream_t st) {
    if (K % GK != 0 || M <= 0 || N <= 0) {
        set_error("gemm: K must be a multiple of %d", GK);
        return FL_E_ARG;
    }
    dim3 grid((unsigned)((M + GB - 1) / GB), (unsigned)((N + GB - 1) / GB));
    gemm_bf16_kernel<<<grid, 256, 0, st>>>(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, bias, act);
    FL_LAUNCH_CHECK();
    return 0;
}

static int colsum(const float *dZ, int64_t N, int W, float *part, float *out, hipStream_t st) {
    const int64_t nch = (N + CS_ROWS - 1) / CS_ROWS;
    colsum_partial_kernel<<<dim3((W + 127) / 128, (unsigned)nch), 128, 0, st>>>(dZ, N, W, part);
    colsum_final_kernel<<<(W + 127) / 128, 128, 0, st>>>(part, nch, W, out);
    FL_LAUNCH_CHECK();
    return 0;
}

}  // namespace fl

using namespace fl;

extern "C" {

int64_t fl_mlp_param_count(void) { return MLP_NPARAM; }

int fl_mlp_forward(const float *P, const float *X, int64_t N, float *H1, float *H2, float *Y, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!P || !X || !H1 || !H2 || !Y || N <= 0) {
        set_error("fl_mlp_forward: invalid argument");
        return FL_E_ARG;
    }
    int rc = gemm(X, MLP_IP, 1, P + OFF_W1, MLP_H, 1, H1, MLP_H, N, MLP_H, MLP_IP, P + OFF_B1, 1, st);
    if (!rc) rc = gemm(H1, MLP_H, 1, P + OFF_W2, MLP_H, 1, H2, MLP_H, N, MLP_H, MLP_H, P + OFF_B2, 1, st);
    if (!rc) rc = gemm(H2, MLP_H, 1, P + OFF_W3, MLP_OP, 1, Y, MLP_OP, N, MLP_OP, MLP_H, P + OFF_B3, 0, st);
    return rc;
}

size_t fl_mlp_workspace_size(int64_t N) {
    const int64_t nch = (N + CS_ROWS - 1) / CS_ROWS;
    return align_up(sizeof(float) * N * MLP_OP, 256) + 2 * align_up(sizeof(float) * N * MLP_H, 256) +
           align_up(sizeof(float) * nch * MLP_H, 256) + align_up(sizeof(double) * ((N * MLP_OP + 255) / 256), 256) +
           1024;
}

int fl_mlp_backward(const float *P, const float *X, const float *H1, const float *H2, const float *Y,
                    const float *Tgt, int64_t N, int32_t nout, float *Gr, float *loss, void *workspace,
                    size_t workspace_bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!P || !X || !H1 || !H2 || !Y || !Tgt || !Gr || !loss || N <= 0 || nout <= 0 || nout > MLP_OP) {
        set_error("fl_mlp_backward: invalid argument");
        return FL_E_ARG;
    }
    if (N % GK != 0) {
        set_error("fl_mlp_backward: batch size must be a multiple of %d", GK);
        return FL_E_ARG;
    }
    const int64_t nch = (N + CS_ROWS - 1) / CS_ROWS;
    Carve cv(workspace, workspace_bytes);
    float *dY = cv.take<float>(N * MLP_OP);
    float *dH2 = cv.take<float>(N * MLP_H);
    float *dH1 = cv.take<float>(N * MLP_H);
    float *part = cv.take<float>(nch * MLP_H);
    double *lp = cv.take<double>((N * MLP_OP + 255) / 256);
    if (!cv.ok) {
        set_error("fl_mlp_backward: workspace too small");
        return FL_E_SPACE;
    }
    const int64_t nb = (N * MLP_OP + 255) / 256;
    mse_grad_kernel<<<(unsigned)nb, 256, 0, st>>>(Y, Tgt, N, nout, dY, lp);
    sum_partials_kernel<<<1, 256, 0, st>>>(lp, nb, 1.0 / ((double)nout * N), loss);
    FL_LAUNCH_CHECK();
    // layer 3: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T
    int rc = gemm(H2, 1, MLP_H, dY, MLP_OP, 1, Gr + OFF_W3, MLP_OP, MLP_H, MLP_OP, N, nullptr, 0, st);
    if (!rc) rc = colsum(dY, N, MLP_OP, part, Gr + OFF_B3, st);
    if (!rc) rc = gemm(dY, MLP_OP, 1, P + OFF_W3, 1, MLP_OP, dH2, MLP_H, N, MLP_H, MLP_OP, nullptr, 0, st);
    if (rc) return rc;
    tanh_grad_kernel<<<(unsigned)((N * MLP_H + 255) / 256), 256, 0, st>>>(H2, dH2, N * MLP_H);
    FL_LAUNCH_CHECK();
    // layer 2: dW2 = H1^T dZ2, db2, dH1 = dZ2 W2^T
    rc = gemm(H1, 1, MLP_H, dH2, MLP_H, 1, Gr + OFF_W2, MLP_H, MLP_H, MLP_H, N, nullptr, 0, st);
    if (!rc) rc = colsum(dH2, N, MLP_H, part, Gr + OFF_B2, st);
    if (!rc) rc = gemm(dH2, MLP_H, 1, P + OFF_W2, 1, MLP_H, dH1, MLP_H, N, MLP_H, MLP_H, nullptr, 0, st);
    if (rc) return rc;
    tanh_grad_kernel<<<(unsigned)((N * MLP_H + 255) / 256), 256, 0, st>>>(H1, dH1, N * MLP_H);
    FL_LAUNCH_CHECK();
    // layer 1: dW1 = X^T dZ1, db1
    rc = gemm(X, 1, MLP_IP, dH1, MLP_H, 1, Gr + OFF_W1, MLP_H, MLP_IP, MLP_H, N, nullptr, 0, st);
    if (!rc) rc = colsum(dH1, N, MLP_H, part, Gr + OFF_B1, st);
    return rc;
}

int fl_mlp_adam(float *P, const float *Gr, float *m, float *v, float lr, float beta1, float beta2, float eps,
                int32_t step, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!P || !Gr || !m || !v || step <= 0) {
        set_error("fl_mlp_adam: invalid argument");
        return FL_E_ARG;
    }
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    adam_kernel<<<(unsigned)((MLP_NPARAM + 255) / 256), 256, 0, st>>>(P, Gr, m, v, MLP_NPARAM, lr, beta1, beta2,
                                                                       eps, c1, c2);
    FL_LAUNCH_CHECK();
    return 0;
}

int fl_warmstart_features(const double *X, const double *u1, const double *u2, double load_factor, int64_t nnode,
                          const float *norm, float *F, void *stream) {
    if (!X || !u1 || !u2 || !norm || !F || nnode <= 0) {
        set_error("fl_warmstart_features: invalid argument");
        return FL_E_ARG;
    }
    const int64_t n = nnode * MLP_IP;
    ws_features_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(X, u1, u2, load_factor, nnode,
                                                                                   norm, F);
    FL_LAUNCH_CHECK();
    return 0;
}

int fl_warmstart_apply(const double *u1, const double *u2, const float *Y, int64_t nnode, const float *out_norm,
                       const uint8_t *free_mask, const double *u_bc, double *u, void *stream) {
    if (!u1 || !u2 || !Y || !out_norm || !free_mask || !u_bc || !u || nnode <= 0) {
        set_error("fl_warmstart_apply: invalid argument");
        return FL_E_ARG;
    }
    ws_apply_kernel<<<(unsigned)((3 * nnode + 255) / 256), 256, 0, as_stream(stream)>>>(u1, u2, Y, nnode, out_norm,
                                                                                       free_mask, u_bc, u);
    FL_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
