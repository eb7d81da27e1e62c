// Kernels of the device Krylov solvers, shared by fl_solve.hip (block-Jacobi PCG) and
// fl_mg.hip (multigrid-preconditioned flexible CG).  Internal linkage: each translation
// unit gets its own copy.
#pragma once
#include <cmath>

#include "fl_common.hpp"

namespace fl {
namespace {

constexpr int SV_T = 256;   // threads per block for vector kernels
constexpr int SP_W = 4;     // waves (nodes) per SpMV block

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// y = A x (masked when free != nullptr).  One wave per node a, its three rows share the
// neighbour list; lanes stride over neighbours.  Optional: partial of dot(x, y).
// MASK_X = false: the caller guarantees x is +0 on fixed dofs (the Krylov vectors of the
// solvers are), so the neighbours' mask bytes are not read.
template <typename VT, bool MASK_X = true>
__global__ __launch_bounds__(64 * SP_W) void spmv_kernel(const int32_t *__restrict__ indptr,
                                                         const int32_t *__restrict__ nbr_ptr,
                                                         const int32_t *__restrict__ nbrs,
                                                         const VT *__restrict__ V, int64_t nnode,
                                                         const uint8_t *__restrict__ free_mask,
                                                         const double *__restrict__ x, double *__restrict__ y,
                                                         double *__restrict__ dot_partial) {
    __shared__ double sdot[SP_W];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t a = (int64_t)blockIdx.x * SP_W + wave;
    double d = 0.0;
    if (a < nnode) {
        // a node whose three dofs are all fixed keeps x: its rows are not read
        const bool all_fixed = free_mask && !free_mask[3 * a] && !free_mask[3 * a + 1] && !free_mask[3 * a + 2];
        const int32_t p0 = nbr_ptr[a], deg = all_fixed ? 0 : nbr_ptr[a + 1] - p0;
        const VT *r0 = V + indptr[3 * a];
        const VT *r1 = V + indptr[3 * a + 1];
        const VT *r2 = V + indptr[3 * a + 2];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int r = lane; r < deg; r += 64) {
            const int64_t b = nbrs[p0 + r];
            double x0 = x[3 * b], x1 = x[3 * b + 1], x2 = x[3 * b + 2];
            if (MASK_X && free_mask) {
                x0 = free_mask[3 * b] ? x0 : 0.0;
                x1 = free_mask[3 * b + 1] ? x1 : 0.0;
                x2 = free_mask[3 * b + 2] ? x2 : 0.0;
            }
            s0 += (double)r0[3 * r] * x0 + (double)r0[3 * r + 1] * x1 + (double)r0[3 * r + 2] * x2;
            s1 += (double)r1[3 * r] * x0 + (double)r1[3 * r + 1] * x1 + (double)r1[3 * r + 2] * x2;
            s2 += (double)r2[3 * r] * x0 + (double)r2[3 * r + 1] * x1 + (double)r2[3 * r + 2] * x2;
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        if (lane == 0) {
            double out[3] = {s0, s1, s2};
            for (int i = 0; i < 3; ++i) {
                const int64_t row = 3 * a + i;
                const double v = (free_mask && !free_mask[row]) ? x[row] : out[i];
                y[row] = v;
                d += v * x[row];
            }
        }
    }
    if (dot_partial) {
        if (lane == 0) sdot[wave] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < SP_W; ++w) s += sdot[w];
            dot_partial[blockIdx.x] = s;
        }
    }
}

// Per-node inverse of the masked 3x3 diagonal block (identity rows/cols on fixed dofs).
template <typename VT>
__global__ void block_jacobi_kernel(const int32_t *__restrict__ indptr, const int32_t *__restrict__ nbr_ptr,
                                    const int32_t *__restrict__ nbrs, const VT *__restrict__ V, int64_t nnode,
                                    const uint8_t *__restrict__ free_mask, double *__restrict__ P,
                                    int32_t *__restrict__ bad) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnode) return;
    const int32_t p0 = nbr_ptr[a], deg = nbr_ptr[a + 1] - p0;
    int lo = 0, hi = deg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a < nbrs[p0 + mid]) hi = mid; else lo = mid;
    }
    double B[9];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const bool fi = free_mask ? free_mask[3 * a + i] != 0 : true;
            const bool fk = free_mask ? free_mask[3 * a + k] != 0 : true;
            B[3 * i + k] = (fi && fk) ? (double)V[indptr[3 * a + i] + 3 * lo + k] : (i == k ? 1.0 : 0.0);
        }
    double inv[9];
    inv[0] = B[4] * B[8] - B[5] * B[7];
    inv[1] = -B[1] * B[8] + B[2] * B[7];
    inv[2] = B[1] * B[5] - B[2] * B[4];
    inv[3] = -B[3] * B[8] + B[5] * B[6];
    inv[4] = B[0] * B[8] - B[2] * B[6];
    inv[5] = -B[0] * B[5] + B[2] * B[3];
    inv[6] = B[3] * B[7] - B[4] * B[6];
    inv[7] = -B[0] * B[7] + B[1] * B[6];
    inv[8] = B[0] * B[4] - B[1] * B[3];
    const double det = B[0] * inv[0] + B[1] * inv[3] + B[2] * inv[6];
    if (!(fabs(det) > 0.0) || nbrs[p0 + lo] != a) {
        atomicOr(bad, 1);
        for (int i = 0; i < 9; ++i) P[9 * a + i] = (i % 4) == 0 ? 1.0 : 0.0;
        return;
    }
    for (int i = 0; i < 9; ++i) P[9 * a + i] = inv[i] / det;
}

__global__ void reduce2_kernel(const double *__restrict__ partial, int64_t n, double *__restrict__ out0,
                               double *__restrict__ out1) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += SV_T) {
        s0 += partial[2 * i];
        s1 += partial[2 * i + 1];
    }
    sh0[threadIdx.x] = s0;
    sh1[threadIdx.x] = s1;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh0[threadIdx.x] += sh0[threadIdx.x + off];
            sh1[threadIdx.x] += sh1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out0 = sh0[0];
        *out1 = sh1[0];
    }
}

// z = P r with partials of (r.z, r.r); used to start the iteration (x = 0, r = b).
__global__ __launch_bounds__(SV_T) void cg_start_kernel(int64_t nnode, const double *__restrict__ P,
                                                        const double *__restrict__ b, const uint8_t *__restrict__ free_mask,
                                                        double *__restrict__ x, double *__restrict__ r,
                                                        double *__restrict__ z, double *__restrict__ p,
                                                        double *__restrict__ partial) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (a < nnode) {
        double rv[3];
        for (int i = 0; i < 3; ++i) {
            const int64_t k = 3 * a + i;
            rv[i] = (free_mask && !free_mask[k]) ? 0.0 : b[k];
            r[k] = rv[i];
            x[k] = 0.0;
        }
        for (int i = 0; i < 3; ++i) {
            const double zi = P[9 * a + 3 * i] * rv[0] + P[9 * a + 3 * i + 1] * rv[1] + P[9 * a + 3 * i + 2] * rv[2];
            z[3 * a + i] = zi;
            p[3 * a + i] = zi;
            rz += rv[i] * zi;
            rr += rv[i] * rv[i];
        }
    }
    sh0[threadIdx.x] = rz;
    sh1[threadIdx.x] = rr;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh0[threadIdx.x] += sh0[threadIdx.x + off];
            sh1[threadIdx.x] += sh1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sh0[0];
        partial[2 * blockIdx.x + 1] = sh1[0];
    }
}

// scalars s[]: 0 rz, 1 pq, 2 alpha, 3 rz_new, 4 beta, 5 rr, 6 bb

template <int NT>
__device__ inline void block_sum2(const double *__restrict__ part, int64_t n, int stride, double &o0, double &o1,
                                  double *sh0, double *sh1) {
    // four independent partial sums per thread keep four loads in flight; fixed order
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t i = threadIdx.x;
    for (; i + 3 * NT < n; i += 4 * NT) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] += part[stride * (i + u * NT)];
            if (stride == 2) b[u] += part[2 * (i + u * NT) + 1];
        }
    }
    for (; i < n; i += NT) {
        a[0] += part[stride * i];
        if (stride == 2) b[0] += part[2 * i + 1];
    }
    sh0[threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
    sh1[threadIdx.x] = (b[0] + b[1]) + (b[2] + b[3]);
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh0[threadIdx.x] += sh0[threadIdx.x + off];
            sh1[threadIdx.x] += sh1[threadIdx.x + off];
        }
        __syncthreads();
    }
    o0 = sh0[0];
    o1 = sh1[0];
}

// One CG step, part 1: p_new = z + beta p_old (formed on the fly for every neighbour,
// stored for the block's own rows), q = A p_new and per-block partials of p_new . q.
// p_old and p_new are distinct buffers (other blocks still read p_old).
template <typename VT>
__global__ __launch_bounds__(64 * SP_W) void pcg_spmv_kernel(const int32_t *__restrict__ indptr,
                                                             const int32_t *__restrict__ nbr_ptr,
                                                             const int32_t *__restrict__ nbrs,
                                                             const VT *__restrict__ V, int64_t nnode,
                                                             const uint8_t *__restrict__ free_mask,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ p_old,
                                                             double *__restrict__ p_new, double *__restrict__ q,
                                                             double *__restrict__ part,
                                                             const double *__restrict__ s) {
    __shared__ double sh0[SP_W];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t a = (int64_t)blockIdx.x * SP_W + wave;
    const double beta = s[4];
    double d = 0.0;
    if (a < nnode) {
        // a node whose three dofs are all fixed has p = 0 and q = 0: its rows are not read
        const bool all_fixed = free_mask && !free_mask[3 * a] && !free_mask[3 * a + 1] && !free_mask[3 * a + 2];
        const int32_t p0 = nbr_ptr[a], deg = all_fixed ? 0 : nbr_ptr[a + 1] - p0;
        const VT *r0 = V + indptr[3 * a];
        const VT *r1 = V + indptr[3 * a + 1];
        const VT *r2 = V + indptr[3 * a + 2];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int r = lane; r < deg; r += 64) {
            // z and p_old are +0 on fixed dofs (r starts masked, the block inverse has
            // identity rows there, q and p_new are set to 0 there): no mask is needed
            const int64_t b = nbrs[p0 + r];
            const double x0 = z[3 * b] + beta * p_old[3 * b];
            const double x1 = z[3 * b + 1] + beta * p_old[3 * b + 1];
            const double x2 = z[3 * b + 2] + beta * p_old[3 * b + 2];
            s0 += (double)r0[3 * r] * x0 + (double)r0[3 * r + 1] * x1 + (double)r0[3 * r + 2] * x2;
            s1 += (double)r1[3 * r] * x0 + (double)r1[3 * r + 1] * x1 + (double)r1[3 * r + 2] * x2;
            s2 += (double)r2[3 * r] * x0 + (double)r2[3 * r + 1] * x1 + (double)r2[3 * r + 2] * x2;
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        if (lane == 0) {
            const double out[3] = {s0, s1, s2};
            for (int i = 0; i < 3; ++i) {
                const int64_t row = 3 * a + i;
                double pn = z[row] + beta * p_old[row];
                const bool fixed = free_mask && !free_mask[row];
                if (fixed) pn = 0.0;
                p_new[row] = pn;
                const double v = fixed ? pn : out[i];
                q[row] = v;
                d += v * pn;
            }
        }
    }
    if (lane == 0) sh0[wave] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < SP_W; ++w) t += sh0[w];
        part[blockIdx.x] = t;
    }
}

// One workgroup: pq = sum of the partials (fixed order), alpha = rz / pq.
__global__ __launch_bounds__(SV_T) void pcg_alpha_kernel(const double *__restrict__ part, int64_t n,
                                                         double *__restrict__ s) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double pq, unused;
    block_sum2<SV_T>(part, n, 1, pq, unused, sh0, sh1);
    if (threadIdx.x == 0) {
        s[1] = pq;
        s[2] = pq != 0.0 ? s[0] / pq : 0.0;
    }
}

// One workgroup: rz_new, rr from the partials of pcg_update_kernel; beta = rz_new / rz.
__global__ __launch_bounds__(SV_T) void pcg_beta_kernel(const double *__restrict__ part, int64_t n,
                                                        double *__restrict__ s) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double rzn, rrn;
    block_sum2<SV_T>(part, n, 2, rzn, rrn, sh0, sh1);
    if (threadIdx.x == 0) {
        s[4] = s[0] != 0.0 ? rzn / s[0] : 0.0;
        s[0] = rzn;
        s[3] = rzn;
        s[5] = rrn;
    }
}

// One CG step, part 2: x += alpha p; r -= alpha q; z = P r; per-block partials of
// (r.z, r.r).
__global__ __launch_bounds__(SV_T) void pcg_update_kernel(int64_t nnode, const double *__restrict__ s,
                                                          const double *__restrict__ P, const double *__restrict__ p,
                                                          const double *__restrict__ q, double *__restrict__ x,
                                                          double *__restrict__ r, double *__restrict__ z,
                                                          double *__restrict__ part) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (a < nnode) {
        const double alpha = s[2];
        double rv[3];
        for (int i = 0; i < 3; ++i) {
            const int64_t k = 3 * a + i;
            x[k] += alpha * p[k];
            rv[i] = r[k] - alpha * q[k];
            r[k] = rv[i];
        }
        for (int i = 0; i < 3; ++i) {
            const double zi = P[9 * a + 3 * i] * rv[0] + P[9 * a + 3 * i + 1] * rv[1] + P[9 * a + 3 * i + 2] * rv[2];
            z[3 * a + i] = zi;
            rz += rv[i] * zi;
            rr += rv[i] * rv[i];
        }
    }
    sh0[threadIdx.x] = rz;
    sh1[threadIdx.x] = rr;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh0[threadIdx.x] += sh0[threadIdx.x + off];
            sh1[threadIdx.x] += sh1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh0[0];
        part[2 * blockIdx.x + 1] = sh1[0];
    }
}

}  // namespace
}  // namespace fl
