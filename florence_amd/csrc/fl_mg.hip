// p-multigrid preconditioned flexible conjugate gradient for the Dirichlet-reduced
// tangent of the Newton steps.
//
// The reference solves each Newton step with a sparse direct solver
// (LinearSolver.py:208-300, called from FEMSolver.py:954).  Here the masked operator
// A = M K M + (I - M) of the assembled CSR is solved by flexible CG (the
// Polak-Ribiere beta, which tolerates a preconditioner that is not a fixed linear
// map) preconditioned by one V-cycle over nested nodal spaces on the same elements:
//   level 0   the assembled CSR (FP64, or the FP32-stored values of the mixed form);
//   level l   a coarser Lagrange space (p=2 lattice, element vertices) whose Galerkin
//             operator P^T K P fl_mg_galerkin forms from the fine point records;
//   coarsest  a fixed number of block-Jacobi CG steps (no host synchronisation).
// Smoother: Chebyshev polynomial of block-Jacobi (3x3 node blocks) on
// [lmax / range, lmax], lmax of D^-1 A estimated per solve by power iteration (warm
// started from the previous solve's vector when the caller says the operator is the
// previous one's successor); the same polynomial before and after the coarse correction
// keeps the cycle symmetric.  The fine level's smoother applies the tangent either
// matrix-free from the assembly's C point records (fl_mf_apply's operator, optionally
// from an FP32 or BF16 copy of them) or from the CSR values (optionally an FP32 copy);
// the smoothers of the intermediate levels may read FP32 copies of their values; the
// outer iteration always applies level 0's own CSR.
// Transfers are the nodal interpolation P (fine node <- coarse nodes of an element
// containing it) and its transpose, both as per-node CSR lists summed in a fixed order.
// Every reduction is in a fixed order: a solve is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "fl_common.hpp"
#include "fl_solve_kernels.hpp"

namespace fl {
namespace {

constexpr double FL_MG_LMAX_SAFETY = 1.15;   // the estimate approaches lmax from below
constexpr int FL_MG_COLD_POWER_STEPS = 12;   // power steps of a restart from the random vector

__device__ inline void bj_apply(const double *__restrict__ P, int64_t a, const double (&r)[3], double (&z)[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) z[i] = P[9 * a + 3 * i] * r[0] + P[9 * a + 3 * i + 1] * r[1] + P[9 * a + 3 * i + 2] * r[2];
}

// sum over a group of LPN consecutive lanes (LPN a power of two <= 64), in every lane
template <int LPN>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int off = LPN / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// q = A d for the three rows of node a (a group of LPN lanes; d is +0 on fixed dofs,
// fixed rows give q = d = 0); returns in the group's lane 0.
template <typename VT, int LPN>
__device__ inline void node_spmv(const int32_t *__restrict__ indptr, const int32_t *__restrict__ nbr_ptr,
                                 const int32_t *__restrict__ nbrs, const VT *__restrict__ V,
                                 const uint8_t *__restrict__ free_mask, const double *__restrict__ d, int64_t a,
                                 int lane, double (&q)[3]) {
    const bool all_fixed = !free_mask[3 * a] && !free_mask[3 * a + 1] && !free_mask[3 * a + 2];
    const int32_t p0 = nbr_ptr[a], deg = all_fixed ? 0 : nbr_ptr[a + 1] - p0;
    const VT *r0 = V + indptr[3 * a];
    const VT *r1 = V + indptr[3 * a + 1];
    const VT *r2 = V + indptr[3 * a + 2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int r = lane; r < deg; r += LPN) {
        const int64_t b = nbrs[p0 + r];
        const double x0 = d[3 * b], x1 = d[3 * b + 1], x2 = d[3 * b + 2];
        s0 += (double)r0[3 * r] * x0 + (double)r0[3 * r + 1] * x1 + (double)r0[3 * r + 2] * x2;
        s1 += (double)r1[3 * r] * x0 + (double)r1[3 * r + 1] * x1 + (double)r1[3 * r + 2] * x2;
        s2 += (double)r2[3 * r] * x0 + (double)r2[3 * r + 1] * x1 + (double)r2[3 * r + 2] * x2;
    }
    q[0] = group_sum<LPN>(s0);
    q[1] = group_sum<LPN>(s1);
    q[2] = group_sum<LPN>(s2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (!free_mask[3 * a + i]) q[i] = 0.0;
}

// Chebyshev start: r = b (if b given), d = D^-1 r / theta, x = d (fresh) or x += d.
__global__ __launch_bounds__(SV_T) void cheb_init_kernel(int64_t nnode, const double *__restrict__ P,
                                                         const double *__restrict__ b, double *__restrict__ r,
                                                         double *__restrict__ d, double *__restrict__ x,
                                                         double inv_theta, int accumulate) {
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (a >= nnode) return;
    double rv[3], z[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rv[i] = b ? b[3 * a + i] : r[3 * a + i];
        if (b) r[3 * a + i] = rv[i];
    }
    bj_apply(P, a, rv, z);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double di = z[i] * inv_theta;
        d[3 * a + i] = di;
        x[3 * a + i] = accumulate ? x[3 * a + i] + di : di;
    }
}

// One Chebyshev step: r -= A d_old; d_new = c1 d_old + c2 D^-1 r; x += d_new.
// c2 = 0 and x = d_new = nullptr: only the residual update r -= A d.  LPN lanes per node
// (64 on the fine level; 32 on the coarse levels, whose nodes have fewer neighbours).
template <typename VT, int LPN>
__global__ __launch_bounds__(64 * SP_W) void cheb_step_kernel(const int32_t *__restrict__ indptr,
                                                              const int32_t *__restrict__ nbr_ptr,
                                                              const int32_t *__restrict__ nbrs,
                                                              const VT *__restrict__ V, int64_t nnode,
                                                              const uint8_t *__restrict__ free_mask,
                                                              const double *__restrict__ P,
                                                              const double *__restrict__ d_old,
                                                              double *__restrict__ d_new, double *__restrict__ r,
                                                              double *__restrict__ x, double c1, double c2) {
    const int grp = threadIdx.x / LPN, lane = threadIdx.x % LPN;
    const int64_t a = (int64_t)blockIdx.x * (64 * SP_W / LPN) + grp;
    if (a >= nnode) return;
    double q[3];
    node_spmv<VT, LPN>(indptr, nbr_ptr, nbrs, V, free_mask, d_old, a, lane, q);
    if (lane != 0) return;
    double rv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rv[i] = r[3 * a + i] - q[i];
        r[3 * a + i] = rv[i];
    }
    if (!d_new) return;
    double z[3];
    bj_apply(P, a, rv, z);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double dn = c1 * d_old[3 * a + i] + c2 * z[i];
        d_new[3 * a + i] = dn;
        x[3 * a + i] += dn;
    }
}

// The matrix-free product's node sums: q = sum of the element vectors ye of node a's
// elements (node_elem: element * npe + local, ascending element order), 0 on fixed dofs.
// With q_out: q_out = q.  Otherwise as cheb_step_kernel: r -= q and, with d_new, the
// Chebyshev update.
__global__ __launch_bounds__(SV_T) void mf_node_kernel(int64_t nnode, const int32_t *__restrict__ node_elem_ptr,
                                                       const int32_t *__restrict__ node_elem,
                                                       const double *__restrict__ ye,
                                                       const uint8_t *__restrict__ free_mask,
                                                       const double *__restrict__ P, const double *__restrict__ d_old,
                                                       double *__restrict__ d_new, double *__restrict__ r,
                                                       double *__restrict__ x, double *__restrict__ q_out, double c1,
                                                       double c2) {
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (a >= nnode) return;
    double q[3] = {0.0, 0.0, 0.0};
    for (int32_t j = node_elem_ptr[a]; j < node_elem_ptr[a + 1]; ++j) {
        const double *v = ye + 3 * (int64_t)node_elem[j];
        q[0] += v[0];
        q[1] += v[1];
        q[2] += v[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (!free_mask[3 * a + i]) q[i] = 0.0;
    if (q_out) {
#pragma unroll
        for (int i = 0; i < 3; ++i) q_out[3 * a + i] = q[i];
        return;
    }
    double rv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rv[i] = r[3 * a + i] - q[i];
        r[3 * a + i] = rv[i];
    }
    if (!d_new) return;
    double z[3];
    bj_apply(P, a, rv, z);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double dn = c1 * d_old[3 * a + i] + c2 * z[i];
        d_new[3 * a + i] = dn;
        x[3 * a + i] += dn;
    }
}

// Sum of part[stride * i + col], i < n, by one wave: lane l takes i = l, l + 64, ... in
// order, then a butterfly; every wave of every block gets the same value.
__device__ inline double wave_sum_partials(const double *__restrict__ part, int64_t n, int stride, int col,
                                           int lane) {
    double a = 0.0;
    for (int64_t i = lane; i < n; i += 64) a += part[stride * i + col];
    return wave_sum(a);
}

// The coarsest level's block-Jacobi CG in two launches per step.  Update partials U_k =
// per-block (r.z, r.r) after k updates (U_0 from cg_start_kernel) rotate through three
// buffers; each block re-sums the partials it needs in the same fixed order, so every
// block holds the same alpha and beta and no single-block reduction launch is needed.
//   step k, part 1 (cc_spmv_kernel): beta = sum U_k[rz] / sum U_{k-1}[rz] (0 at k = 0),
//     p_new = z + beta p_old on the fly, q = A p_new, partials of p_new.q;
//   step k, part 2 (cc_update_kernel): alpha = sum U_k[rz] / sum(p.q), x += alpha p,
//     r -= alpha q, z = P r, partials U_{k+1}.
// The coarsest level is a vertex space (≈ 27 neighbours per node at most on hexes): half a
// wave per node, CC_NB nodes per workgroup.
constexpr int CC_LPN = 32;
constexpr int CC_NB = 64 * SP_W / CC_LPN;


template <typename VT>
__global__ __launch_bounds__(64 * SP_W) void cc_spmv_kernel(const int32_t *__restrict__ indptr,
                                                            const int32_t *__restrict__ nbr_ptr,
                                                            const int32_t *__restrict__ nbrs,
                                                            const VT *__restrict__ V, int64_t nnode,
                                                            const uint8_t *__restrict__ free_mask,
                                                            const double *__restrict__ z,
                                                            const double *__restrict__ p_old, double *__restrict__ p_new,
                                                            double *__restrict__ q, double *__restrict__ spart,
                                                            const double *__restrict__ ucur,
                                                            const double *__restrict__ uprev, int64_t nu) {
    __shared__ double sh0[CC_NB];
    const int lane = threadIdx.x & 63, grp = threadIdx.x / CC_LPN, gl = threadIdx.x % CC_LPN;
    double beta = 0.0;
    if (uprev) {
        const double rz = wave_sum_partials(ucur, nu, 2, 0, lane), rzo = wave_sum_partials(uprev, nu, 2, 0, lane);
        beta = rzo != 0.0 ? rz / rzo : 0.0;
    }
    const int64_t a = (int64_t)blockIdx.x * CC_NB + grp;
    double d = 0.0;
    if (a < nnode) {
        const bool all_fixed = !free_mask[3 * a] && !free_mask[3 * a + 1] && !free_mask[3 * a + 2];
        const int32_t p0 = nbr_ptr[a], deg = all_fixed ? 0 : nbr_ptr[a + 1] - p0;
        const VT *r0 = V + indptr[3 * a];
        const VT *r1 = V + indptr[3 * a + 1];
        const VT *r2 = V + indptr[3 * a + 2];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int r = gl; r < deg; r += CC_LPN) {
            // z and p_old are +0 on fixed dofs (the block inverse has identity rows there
            // and r starts masked; p_new is set to 0 there)
            const int64_t b = nbrs[p0 + r];
            const double x0 = z[3 * b] + beta * p_old[3 * b];
            const double x1 = z[3 * b + 1] + beta * p_old[3 * b + 1];
            const double x2 = z[3 * b + 2] + beta * p_old[3 * b + 2];
            s0 += (double)r0[3 * r] * x0 + (double)r0[3 * r + 1] * x1 + (double)r0[3 * r + 2] * x2;
            s1 += (double)r1[3 * r] * x0 + (double)r1[3 * r + 1] * x1 + (double)r1[3 * r + 2] * x2;
            s2 += (double)r2[3 * r] * x0 + (double)r2[3 * r + 1] * x1 + (double)r2[3 * r + 2] * x2;
        }
        s0 = group_sum<CC_LPN>(s0);
        s1 = group_sum<CC_LPN>(s1);
        s2 = group_sum<CC_LPN>(s2);
        if (gl == 0) {
            const double out[3] = {s0, s1, s2};
            for (int i = 0; i < 3; ++i) {
                const int64_t row = 3 * a + i;
                const bool fixed = !free_mask[row];
                const double pn = fixed ? 0.0 : z[row] + beta * p_old[row];
                p_new[row] = pn;
                const double v = fixed ? 0.0 : out[i];
                q[row] = v;
                d += v * pn;
            }
        }
    }
    if (gl == 0) sh0[grp] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < CC_NB; ++w) t += sh0[w];
        spart[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(SV_T) void cc_update_kernel(int64_t nnode, const double *__restrict__ P,
                                                         const double *__restrict__ p, const double *__restrict__ q,
                                                         double *__restrict__ x, double *__restrict__ r,
                                                         double *__restrict__ z, const double *__restrict__ spart,
                                                         int64_t ns, const double *__restrict__ ucur, int64_t nu,
                                                         double *__restrict__ unext) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    // p.q over the SpMV blocks (thousands): the whole workgroup, four loads in flight per
    // thread; r.z over the update blocks (tens): one wave
    double pq, unused;
    block_sum2<SV_T>(spart, ns, 1, pq, unused, sh0, sh1);
    __syncthreads();
    const double rzc = wave_sum_partials(ucur, nu, 2, 0, threadIdx.x & 63);
    const double alpha = pq != 0.0 ? rzc / pq : 0.0;
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (a < nnode) {
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
        unext[2 * blockIdx.x] = sh0[0];
        unext[2 * blockIdx.x + 1] = sh1[0];
    }
}

// coarse b[c] = sum_j R_w[j] r_fine[R_idx[j]], masked by the coarse free dofs; one wave
// per coarse node (its list holds tens of fine nodes), lanes striding over the list
__global__ __launch_bounds__(64 * SP_W) void restrict_kernel(int64_t nnode_c, const int32_t *__restrict__ R_ptr,
                                                             const int32_t *__restrict__ R_idx,
                                                             const double *__restrict__ R_w,
                                                             const uint8_t *__restrict__ free_c,
                                                             const double *__restrict__ r, double *__restrict__ bc) {
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * SP_W + (threadIdx.x >> 6);
    if (c >= nnode_c) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int32_t j = R_ptr[c] + lane; j < R_ptr[c + 1]; j += 64) {
        const int64_t f = R_idx[j];
        const double w = R_w[j];
        s0 += w * r[3 * f];
        s1 += w * r[3 * f + 1];
        s2 += w * r[3 * f + 2];
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        bc[3 * c] = free_c[3 * c] ? s0 : 0.0;
        bc[3 * c + 1] = free_c[3 * c + 1] ? s1 : 0.0;
        bc[3 * c + 2] = free_c[3 * c + 2] ? s2 : 0.0;
    }
}

// d[f] = sum_j P_w[j] e_coarse[P_idx[j]] (masked by the fine free dofs); x += d
__global__ __launch_bounds__(SV_T) void prolong_kernel(int64_t nnode, const int32_t *__restrict__ P_ptr,
                                                       const int32_t *__restrict__ P_idx,
                                                       const double *__restrict__ P_w,
                                                       const uint8_t *__restrict__ free_mask,
                                                       const double *__restrict__ ec, double *__restrict__ d,
                                                       double *__restrict__ x) {
    const int64_t f = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (f >= nnode) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int32_t j = P_ptr[f]; j < P_ptr[f + 1]; ++j) {
        const int64_t c = P_idx[j];
        const double w = P_w[j];
        s0 += w * ec[3 * c];
        s1 += w * ec[3 * c + 1];
        s2 += w * ec[3 * c + 2];
    }
    const double v[3] = {free_mask[3 * f] ? s0 : 0.0, free_mask[3 * f + 1] ? s1 : 0.0, free_mask[3 * f + 2] ? s2 : 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        d[3 * f + i] = v[i];
        x[3 * f + i] += v[i];
    }
}

// start of the power iteration, zero on fixed dofs: a deterministic pseudo-random
// vector, or (warm) the vector already in v restricted to this call's free dofs -- on a
// fixed dof the smoother's operator is the identity, so a component left there would
// pull the estimate towards 1
__global__ __launch_bounds__(SV_T) void power_start_kernel(int64_t n, const uint8_t *__restrict__ free_mask,
                                                           double *__restrict__ v, int warm) {
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (i >= n) return;
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    const double start = warm ? v[i] : 0.5 + (double)(h >> 11) * (1.0 / 9007199254740992.0);
    v[i] = free_mask[i] ? start : 0.0;
}

// z = D^-1 y with per-block partials of z.z
__global__ __launch_bounds__(SV_T) void bj_norm_kernel(int64_t nnode, const double *__restrict__ P,
                                                       const double *__restrict__ y, double *__restrict__ z,
                                                       double *__restrict__ part) {
    __shared__ double sh[SV_T];
    const int64_t a = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double s = 0.0;
    if (a < nnode) {
        const double yv[3] = {y[3 * a], y[3 * a + 1], y[3 * a + 2]};
        double zv[3];
        bj_apply(P, a, yv, zv);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            z[3 * a + i] = zv[i];
            s += zv[i] * zv[i];
        }
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(SV_T) void sum1_kernel(const double *__restrict__ part, int64_t n,
                                                    double *__restrict__ out) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double v, unused;
    block_sum2<SV_T>(part, n, 1, v, unused, sh0, sh1);
    if (threadIdx.x == 0) *out = v;
}

// v = z / sqrt(zz)
__global__ __launch_bounds__(SV_T) void scale_kernel(int64_t n, const double *__restrict__ z,
                                                     const double *__restrict__ zz, double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (i >= n) return;
    const double s = *zz > 0.0 ? 1.0 / sqrt(*zz) : 0.0;
    v[i] = z[i] * s;
}

__global__ void zero_small_kernel(double *__restrict__ s, int n) {
    if ((int)threadIdx.x < n) s[threadIdx.x] = 0.0;
}

// FCG start: r = b (masked), x = 0, partial of r.r
__global__ __launch_bounds__(SV_T) void fcg_start_kernel(int64_t n, const double *__restrict__ b,
                                                         const uint8_t *__restrict__ free_mask,
                                                         double *__restrict__ x, double *__restrict__ r,
                                                         double *__restrict__ part) {
    __shared__ double sh[SV_T];
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const double v = free_mask[i] ? b[i] : 0.0;
        r[i] = v;
        x[i] = 0.0;
        s = v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// partials of (z.r, z.r_old)
__global__ __launch_bounds__(SV_T) void fcg_dots_kernel(int64_t n, const double *__restrict__ z,
                                                        const double *__restrict__ r,
                                                        const double *__restrict__ r_old, double *__restrict__ part) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        a = z[i] * r[i];
        b = r_old ? z[i] * r_old[i] : 0.0;
    }
    sh0[threadIdx.x] = a;
    sh1[threadIdx.x] = b;
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

// out[b] = sum of part[b * PART_CHUNK, (b + 1) * PART_CHUNK) in a fixed order: the first
// of two stages summing one partial per SpMV workgroup (a quarter of a million at 24^3)
constexpr int PART_CHUNK = 1024;
__global__ __launch_bounds__(SV_T) void part_sum_kernel(const double *__restrict__ part, int64_t n,
                                                        double *__restrict__ out) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    const int64_t lo = (int64_t)blockIdx.x * PART_CHUNK;
    const int64_t cnt = n - lo < PART_CHUNK ? n - lo : PART_CHUNK;
    double v, unused;
    block_sum2<SV_T>(part + lo, cnt, 1, v, unused, sh0, sh1);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// scalars s[]: 0 zr (current), 1 pq, 2 alpha, 3 beta, 4 rr, 5 zr_old
// beta = (z.r - z.r_old) / zr_old  (first iteration: first = 1, beta = 0)
__global__ __launch_bounds__(SV_T) void fcg_beta_kernel(const double *__restrict__ part, int64_t n,
                                                        double *__restrict__ s, int first) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double zr, zro;
    block_sum2<SV_T>(part, n, 2, zr, zro, sh0, sh1);
    if (threadIdx.x == 0) {
        s[3] = (first || s[0] == 0.0) ? 0.0 : (zr - zro) / s[0];
        s[5] = s[0];
        s[0] = zr;
    }
}

__global__ __launch_bounds__(SV_T) void fcg_alpha_kernel(const double *__restrict__ part, int64_t n,
                                                         double *__restrict__ s) {
    __shared__ double sh0[SV_T], sh1[SV_T];
    double pq, unused;
    block_sum2<SV_T>(part, n, 1, pq, unused, sh0, sh1);
    if (threadIdx.x == 0) {
        s[1] = pq;
        s[2] = pq != 0.0 ? s[0] / pq : 0.0;
    }
}

// p = z + beta p (first: p = z)
__global__ __launch_bounds__(SV_T) void fcg_dir_kernel(int64_t n, const double *__restrict__ s,
                                                       const double *__restrict__ z, double *__restrict__ p,
                                                       int first) {
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    if (i >= n) return;
    p[i] = first ? z[i] : z[i] + s[3] * p[i];
}

// x += alpha p; r_old = r; r -= alpha q; partial of r.r
__global__ __launch_bounds__(SV_T) void fcg_update_kernel(int64_t n, const double *__restrict__ s,
                                                          const double *__restrict__ p, const double *__restrict__ q,
                                                          double *__restrict__ x, double *__restrict__ r,
                                                          double *__restrict__ r_old, double *__restrict__ part) {
    __shared__ double sh[SV_T];
    const int64_t i = (int64_t)blockIdx.x * SV_T + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double alpha = s[2];
        x[i] += alpha * p[i];
        const double ro = r[i];
        r_old[i] = ro;
        const double rn = ro - alpha * q[i];
        r[i] = rn;
        v = rn * rn;
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = SV_T / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Per level: block-Jacobi inverses and the cycle's vectors.  On Chebyshev levels x is
// the cycle's result, r its running residual, d / d2 the Chebyshev corrections and b
// the restricted right-hand side; on the coarsest level d, d2, t1, t2 are CG's z, the
// two search-direction buffers and q.
struct LevelWork {
    double *P;
    double *x, *r, *d, *d2, *b, *t1, *t2;
    double *pv;   // the power iteration's vector, kept from call to call (warm starts)
    double lmax;
};

constexpr int NVEC = 8;

struct Mg {
    const FlMgLevel *lv;
    int nl;
    int v32;
    const float *Vs32[FL_MG_MAX_LEVELS];   // optional FP32 copies of the levels' values for their smoothers
    bool use_mf;        // level 0's smoother applies the matrix-free operator mf
    MfOp mf;
    int degree;
    double range;
    int coarse_its;
    LevelWork w[FL_MG_MAX_LEVELS];
    double *part;
    double *s;   // [0, 8) FCG scalars, [8, 16) coarse CG scalars, [16] power-iteration norm
    hipStream_t st;
};

template <typename VT>
static const VT *lvals(const FlMgLevel &L) {
    return static_cast<const VT *>(L.V);
}

static unsigned nb_sp(int64_t nnode) { return (unsigned)((nnode + SP_W - 1) / SP_W); }
static unsigned nb_v(int64_t n) { return (unsigned)((n + SV_T - 1) / SV_T); }

// the smoother's view of level l's values: level 0 reads the FP32 copy when one is given
static bool smooth_f32(const Mg &m, int l) {
    if (l == 0 && (m.use_mf || m.v32)) return !m.use_mf;
    return m.Vs32[l] != nullptr;
}

// the matrix-free product on level 0: element vectors, then node_kernel's sums (q_out, or
// the Chebyshev step of mf_node_kernel)
static int mf_product(const Mg &m, const double *d, double *q_out, const double *d_old, double *d_new, double *x,
                      double c1, double c2) {
    const FlMgLevel &L = m.lv[0];
    const int rc = mf_element_apply(m.mf, L.free_mask, d, m.st);
    if (rc) return rc;
    mf_node_kernel<<<nb_v(L.nnode), SV_T, 0, m.st>>>(L.nnode, m.mf.node_elem_ptr, m.mf.node_elem, m.mf.ye,
                                                     L.free_mask, m.w[0].P, d_old, d_new, m.w[0].r, x, q_out, c1, c2);
    FL_LAUNCH_CHECK();
    return 0;
}
static const float *smooth_v32(const Mg &m, int l) {
    return l == 0 && m.v32 ? lvals<float>(m.lv[l]) : m.Vs32[l];
}

// q = A_l d, with per-block partials of d.q when part is given; smoother: with the
// values the smoother reads (the spectrum estimate), else the level's own operator
static int level_spmv(const Mg &m, int l, const double *d, double *q, double *part, bool smoother = false) {
    const FlMgLevel &L = m.lv[l];
    if (smoother && l == 0 && m.use_mf && !part) return mf_product(m, d, q, nullptr, nullptr, nullptr, 0.0, 0.0);
    // d is +0 on fixed dofs here (the cycle's and the outer iteration's vectors are)
    if (smoother ? smooth_f32(m, l) : (l == 0 && m.v32))
        spmv_kernel<float, false><<<nb_sp(L.nnode), 64 * SP_W, 0, m.st>>>(
            L.indptr, L.node_nbr_ptr, L.node_nbrs, smoother ? smooth_v32(m, l) : lvals<float>(L), L.nnode,
            L.free_mask, d, q, part);
    else
        spmv_kernel<double, false><<<nb_sp(L.nnode), 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs,
                                                                            lvals<double>(L), L.nnode, L.free_mask,
                                                                            d, q, part);
    FL_LAUNCH_CHECK();
    return 0;
}

// r -= A d_old and, with d_new, the Chebyshev update d_new = c1 d_old + c2 D^-1 r, x += d_new
static int level_cheb_step(const Mg &m, int l, const double *d_old, double *d_new, double *x, double c1, double c2) {
    const FlMgLevel &L = m.lv[l];
    const LevelWork &w = m.w[l];
    if (l == 0 && m.use_mf) return mf_product(m, d_old, nullptr, d_old, d_new, x, c1, c2);
    const float *v32 = smooth_f32(m, l) ? smooth_v32(m, l) : nullptr;
    const double *v64 = lvals<double>(L);
    const unsigned nb64 = nb_sp(L.nnode), nb32 = (unsigned)((L.nnode + 2 * SP_W - 1) / (2 * SP_W));
    if (l == 0 && v32)
        cheb_step_kernel<float, 64><<<nb64, 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs, v32,
                                                                   L.nnode, L.free_mask, w.P, d_old, d_new, w.r, x,
                                                                   c1, c2);
    else if (l == 0)
        cheb_step_kernel<double, 64><<<nb64, 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs, v64,
                                                                    L.nnode, L.free_mask, w.P, d_old, d_new, w.r, x,
                                                                    c1, c2);
    else if (v32)
        cheb_step_kernel<float, 32><<<nb32, 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs, v32,
                                                                   L.nnode, L.free_mask, w.P, d_old, d_new, w.r, x,
                                                                   c1, c2);
    else
        cheb_step_kernel<double, 32><<<nb32, 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs, v64,
                                                                    L.nnode, L.free_mask, w.P, d_old, d_new, w.r, x,
                                                                    c1, c2);
    FL_LAUNCH_CHECK();
    return 0;
}

// Degree-k Chebyshev smoothing of A_l x = b on [lmax / range, lmax]: from x = 0 with
// r = b (b given), or on top of x with r its residual (b = nullptr).  On return r is the
// residual of x before the last correction, which *d_last holds.
static int chebyshev(Mg &m, int l, const double *b, double **d_last) {
    LevelWork &w = m.w[l];
    const FlMgLevel &L = m.lv[l];
    const double lmax = w.lmax, lmin = lmax / m.range;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    cheb_init_kernel<<<nb_v(L.nnode), SV_T, 0, m.st>>>(L.nnode, w.P, b, w.r, w.d, w.x, 1.0 / theta, b ? 0 : 1);
    FL_LAUNCH_CHECK();
    double *dc = w.d, *dn = w.d2;
    for (int k = 1; k < m.degree; ++k) {
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        const int rc = level_cheb_step(m, l, dc, dn, w.x, rho_new * rho, 2.0 * rho_new / delta);
        if (rc) return rc;
        rho = rho_new;
        double *t = dc;
        dc = dn;
        dn = t;
    }
    *d_last = dc;
    return 0;
}

// fixed-step block-Jacobi CG from x = 0 on the coarsest level: x ~ A^-1 b (two
// launches per step: cc_spmv_kernel, cc_update_kernel)
static int coarse_solve(Mg &m, int l) {
    const FlMgLevel &L = m.lv[l];
    LevelWork &w = m.w[l];
    const int64_t nnode = L.nnode;
    const int64_t ns = (nnode + CC_NB - 1) / CC_NB, nu = nb_v(nnode);
    double *z = w.d, *q = w.t2;
    double *pbuf[2] = {w.d2, w.t1};
    double *spart = m.part;
    double *ubuf[3] = {m.part + ns, m.part + ns + 2 * nu, m.part + ns + 4 * nu};
    cg_start_kernel<<<(unsigned)nu, SV_T, 0, m.st>>>(nnode, w.P, w.b, L.free_mask, w.x, w.r, z, pbuf[0], ubuf[0]);
    FL_LAUNCH_CHECK();
    for (int it = 0; it < m.coarse_its; ++it) {
        double *p_old = pbuf[it & 1], *p_new = pbuf[(it + 1) & 1];
        const double *ucur = ubuf[it % 3], *uprev = it ? ubuf[(it + 2) % 3] : nullptr;
        cc_spmv_kernel<double><<<(unsigned)ns, 64 * SP_W, 0, m.st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs,
                                                                     lvals<double>(L), nnode, L.free_mask, z, p_old,
                                                                     p_new, q, spart, ucur, uprev, nu);
        cc_update_kernel<<<(unsigned)nu, SV_T, 0, m.st>>>(nnode, w.P, p_new, q, w.x, w.r, z, spart, ns, ucur, nu,
                                                          ubuf[(it + 1) % 3]);
        FL_LAUNCH_CHECK();
    }
    return 0;
}

// one V-cycle on level l: m.w[l].x ~ A_l^-1 b_l (b_l = rhs, or m.w[l].b below the top)
static int vcycle(Mg &m, int l, const double *rhs) {
    LevelWork &w = m.w[l];
    const FlMgLevel &L = m.lv[l];
    if (l == m.nl - 1) return coarse_solve(m, l);
    double *dl = nullptr;
    // pre-smoothing from x = 0 on this level's right-hand side
    int rc = chebyshev(m, l, rhs ? rhs : w.b, &dl);
    if (rc) return rc;
    rc = level_cheb_step(m, l, dl, nullptr, nullptr, 0.0, 0.0);   // r = b - A x
    if (rc) return rc;
    const FlMgLevel &C = m.lv[l + 1];
    LevelWork &wc = m.w[l + 1];
    restrict_kernel<<<nb_sp(C.nnode), 64 * SP_W, 0, m.st>>>(C.nnode, C.R_ptr, C.R_idx, C.R_w, C.free_mask, w.r,
                                                             wc.b);
    FL_LAUNCH_CHECK();
    rc = vcycle(m, l + 1, nullptr);
    if (rc) return rc;
    prolong_kernel<<<nb_v(L.nnode), SV_T, 0, m.st>>>(L.nnode, C.P_ptr, C.P_idx, C.P_w, L.free_mask, wc.x, w.d, w.x);
    FL_LAUNCH_CHECK();
    rc = level_cheb_step(m, l, w.d, nullptr, nullptr, 0.0, 0.0);   // r = b - A x after the correction
    if (rc) return rc;
    return chebyshev(m, l, nullptr, &dl);
}

// largest eigenvalue of D^-1 A on level l (A as the smoother sees it) by power iteration
// (device-side normalisation, one host read at the end), times a safety factor; warm:
// from the vector the previous call left (the previous Newton step's operator), else
// from a fixed pseudo-random start; either way restricted to this call's free dofs.  A
// warm vector can be exactly zero there (a lockstep solve's converged members leave the
// mask, and repeated power steps concentrate the vector on one member): the estimate
// then starts again from the pseudo-random vector with FL_MG_COLD_POWER_STEPS steps.
static int estimate_lmax(Mg &m, int l, int steps, bool warm) {
    LevelWork &w = m.w[l];
    const FlMgLevel &L = m.lv[l];
    const int64_t n = 3 * L.nnode;
    double *v = w.pv, *y = w.d2, *z = w.r, *zz = m.s + 16;
    power_start_kernel<<<nb_v(n), SV_T, 0, m.st>>>(n, L.free_mask, v, warm ? 1 : 0);
    FL_LAUNCH_CHECK();
    for (int it = 0; it < steps; ++it) {
        int rc = level_spmv(m, l, v, y, nullptr, true);
        if (rc) return rc;
        bj_norm_kernel<<<nb_v(L.nnode), SV_T, 0, m.st>>>(L.nnode, w.P, y, z, m.part);
        sum1_kernel<<<1, SV_T, 0, m.st>>>(m.part, nb_v(L.nnode), zz);
        scale_kernel<<<nb_v(n), SV_T, 0, m.st>>>(n, z, zz, v);
        FL_LAUNCH_CHECK();
    }
    double h = 0.0;
    FL_HIP_CHECK(hipMemcpyAsync(&h, zz, sizeof(double), hipMemcpyDeviceToHost, m.st));
    FL_HIP_CHECK(hipStreamSynchronize(m.st));
    if (warm && !(h > 0.0)) return estimate_lmax(m, l, std::max(steps, FL_MG_COLD_POWER_STEPS), false);
    if (!(h > 0.0) || !std::isfinite(h)) {
        set_error("fl_mg_pcg: eigenvalue estimate of level %d failed (%g)", l, h);
        return FL_E_ARG;
    }
    w.lmax = FL_MG_LMAX_SAFETY * std::sqrt(h);
    return 0;
}

static size_t carve_mg(Carve &cv, Mg &m, const FlMgLevel *lv, int nl) {
    int64_t maxpart = 0;
    for (int l = 0; l < nl; ++l) {
        const int64_t nn = lv[l].nnode;
        m.w[l].P = cv.take<double>(9 * (size_t)nn);
        double **vec[NVEC] = {&m.w[l].x, &m.w[l].r, &m.w[l].d, &m.w[l].d2, &m.w[l].b, &m.w[l].t1, &m.w[l].t2,
                              &m.w[l].pv};
        for (int k = 0; k < NVEC; ++k) *vec[k] = cv.take<double>(3 * (size_t)nn);
        // the outer iteration's partials, or the coarse CG's (p.q per SpMV block and three
        // buffers of (r.z, r.r) per update block)
        const int64_t pp = std::max<int64_t>(nb_sp(nn) + 2 * (int64_t)nb_v(3 * nn), nb_sp(nn) + 6 * (int64_t)nb_v(nn)) + 8;
        if (pp > maxpart) maxpart = pp;
    }
    m.part = cv.take<double>((size_t)maxpart);
    m.s = cv.take<double>(32);
    return cv.off + 256;
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_mg_workspace_size(const int64_t *nnodes, int32_t nlevels) {
    if (!nnodes || nlevels < 2 || nlevels > FL_MG_MAX_LEVELS) return 0;
    FlMgLevel lv[FL_MG_MAX_LEVELS] = {};
    for (int l = 0; l < nlevels; ++l) lv[l].nnode = nnodes[l];
    Mg m = {};
    Carve cv(nullptr, 0);
    // outer FCG vectors of level 0: r, r_old, p, q (carved one by one, as fl_mg_pcg does)
    for (int k = 0; k < 4; ++k) cv.take<double>(3 * (size_t)nnodes[0]);
    return carve_mg(cv, m, lv, nlevels);
}

int fl_mf_apply(const FlMgMatrixFree *op, const uint8_t *free_mask, const double *x, double *y, void *stream) {
    if (!free_mask || !x || !y) {
        set_error("fl_mf_apply: invalid argument");
        return FL_E_ARG;
    }
    MfOp mo;
    int rc = mf_prepare(op, &mo);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    rc = mf_element_apply(mo, free_mask, x, st);
    if (rc) return rc;
    mf_node_kernel<<<nb_v(mo.nnode), SV_T, 0, st>>>(mo.nnode, mo.node_elem_ptr, mo.node_elem, mo.ye, free_mask,
                                                    nullptr, nullptr, nullptr, nullptr, nullptr, y, 0.0, 0.0);
    FL_LAUNCH_CHECK();
    return 0;
}

int fl_mg_pcg(const FlMgLevel *levels, int32_t nlevels, int32_t v32, const float *const *smoother_V32,
              const FlMgMatrixFree *fine_mf, const double *b, double *x, double rtol, int32_t maxit,
              int32_t check_every, int32_t cheb_degree, double cheb_range, int32_t coarse_iterations,
              int32_t power_steps, int32_t power_warm, void *workspace, size_t workspace_bytes, int32_t *iterations,
              double *rel_residual, void *stream) {
    if (!levels || nlevels < 2 || nlevels > FL_MG_MAX_LEVELS || !b || !x || maxit <= 0 || !iterations ||
        !rel_residual || cheb_degree < 1 || cheb_degree > 16 || !(cheb_range > 1.0) || coarse_iterations < 1 ||
        power_steps < 1 || power_steps > 64) {
        set_error("fl_mg_pcg: invalid argument");
        return FL_E_ARG;
    }
    for (int l = 0; l < nlevels; ++l) {
        const FlMgLevel &L = levels[l];
        if (L.nnode <= 0 || !L.indptr || !L.node_nbr_ptr || !L.node_nbrs || !L.V || !L.free_mask) {
            set_error("fl_mg_pcg: level %d is incomplete", l);
            return FL_E_ARG;
        }
        if (l > 0 && (!L.P_ptr || !L.P_idx || !L.P_w || !L.R_ptr || !L.R_idx || !L.R_w)) {
            set_error("fl_mg_pcg: level %d has no transfer tables", l);
            return FL_E_ARG;
        }
    }
    if (check_every <= 0) check_every = 1;
    Mg m = {};
    m.lv = levels;
    m.nl = nlevels;
    m.v32 = v32 ? 1 : 0;
    for (int l = 0; l < nlevels; ++l) m.Vs32[l] = smoother_V32 && !(l == 0 && v32) ? smoother_V32[l] : nullptr;
    m.use_mf = fine_mf != nullptr;
    if (fine_mf) {
        int rc = mf_prepare(fine_mf, &m.mf);
        if (rc) return rc;
        if (fine_mf->records16) {
            rc = mf_round_records16(m.mf, fine_mf->records16, m.st);
            if (rc) return rc;
            m.mf.c16 = fine_mf->records16;
        } else if (fine_mf->records32) {
            rc = mf_round_records(m.mf, fine_mf->records32, m.st);
            if (rc) return rc;
            m.mf.c32 = fine_mf->records32;
        }
        if (m.mf.nnode != levels[0].nnode) {
            set_error("fl_mg_pcg: the matrix-free operator has %lld nodes, level 0 %lld", (long long)m.mf.nnode,
                      (long long)levels[0].nnode);
            return FL_E_ARG;
        }
    }
    m.degree = cheb_degree;
    m.range = cheb_range;
    m.coarse_its = coarse_iterations;
    m.st = as_stream(stream);
    const int64_t n0 = 3 * levels[0].nnode;
    Carve cv(workspace, workspace_bytes);
    double *rO = cv.take<double>(n0), *rOld = cv.take<double>(n0), *p = cv.take<double>(n0), *q = cv.take<double>(n0);
    const size_t need = carve_mg(cv, m, levels, nlevels);
    if (!cv.ok) {
        set_error("fl_mg_pcg: workspace too small (%zu < %zu)", workspace_bytes, need);
        return FL_E_SPACE;
    }
    hipStream_t st = m.st;
    double *s = m.s;
    zero_small_kernel<<<1, 64, 0, st>>>(s, 32);
    FL_LAUNCH_CHECK();
    // block-Jacobi inverses (a singular block is reported in the int after s[24])
    int32_t *bad = reinterpret_cast<int32_t *>(s + 24);
    for (int l = 0; l < nlevels; ++l) {
        const FlMgLevel &L = levels[l];
        const unsigned nb = (unsigned)((L.nnode + 255) / 256);
        if (l == 0 && m.v32)
            block_jacobi_kernel<float><<<nb, 256, 0, st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs,
                                                           static_cast<const float *>(L.V), L.nnode, L.free_mask,
                                                           m.w[l].P, bad);
        else
            block_jacobi_kernel<double><<<nb, 256, 0, st>>>(L.indptr, L.node_nbr_ptr, L.node_nbrs,
                                                            static_cast<const double *>(L.V), L.nnode, L.free_mask,
                                                            m.w[l].P, bad);
        FL_LAUNCH_CHECK();
    }
    for (int l = 0; l + 1 < nlevels; ++l) {
        const int rc = estimate_lmax(m, l, power_steps, power_warm != 0);
        if (rc) return rc;
    }
    int32_t hbad = 0;
    FL_HIP_CHECK(hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    // outer flexible CG
    const unsigned nbv0 = nb_v(n0), nbs0 = nb_sp(levels[0].nnode);
    fcg_start_kernel<<<nbv0, SV_T, 0, st>>>(n0, b, levels[0].free_mask, x, rO, m.part);
    sum1_kernel<<<1, SV_T, 0, st>>>(m.part, nbv0, s + 4);
    FL_LAUNCH_CHECK();
    double bb = 0.0;
    FL_HIP_CHECK(hipMemcpyAsync(&bb, s + 4, sizeof(double), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    if (hbad) {
        set_error("fl_mg_pcg: singular diagonal block or missing diagonal entry");
        return FL_E_ARG;
    }
    const double bnorm = std::sqrt(bb);
    *iterations = 0;
    *rel_residual = 0.0;
    if (bnorm == 0.0) return 0;
    double *z = m.w[0].x;
    int rc = vcycle(m, 0, rO);
    if (rc) return rc;
    fcg_dots_kernel<<<nbv0, SV_T, 0, st>>>(n0, z, rO, nullptr, m.part);
    fcg_beta_kernel<<<1, SV_T, 0, st>>>(m.part, nbv0, s, 1);
    fcg_dir_kernel<<<nbv0, SV_T, 0, st>>>(n0, s, z, p, 1);
    FL_LAUNCH_CHECK();
    double rel = 1.0;
    int it = 0;
    while (it < maxit) {
        rc = level_spmv(m, 0, p, q, m.part);
        if (rc) return rc;
        const unsigned nch = (unsigned)((nbs0 + PART_CHUNK - 1) / PART_CHUNK);
        part_sum_kernel<<<nch, SV_T, 0, st>>>(m.part, nbs0, m.part + nbs0);
        fcg_alpha_kernel<<<1, SV_T, 0, st>>>(m.part + nbs0, nch, s);
        fcg_update_kernel<<<nbv0, SV_T, 0, st>>>(n0, s, p, q, x, rO, rOld, m.part);
        sum1_kernel<<<1, SV_T, 0, st>>>(m.part, nbv0, s + 4);
        FL_LAUNCH_CHECK();
        ++it;
        if (it % check_every == 0 || it == maxit) {
            double rr = 0.0;
            FL_HIP_CHECK(hipMemcpyAsync(&rr, s + 4, sizeof(double), hipMemcpyDeviceToHost, st));
            FL_HIP_CHECK(hipStreamSynchronize(st));
            rel = std::sqrt(rr) / bnorm;
            if (!(rel == rel) || rel <= rtol) break;
        }
        if (it == maxit) break;
        rc = vcycle(m, 0, rO);
        if (rc) return rc;
        fcg_dots_kernel<<<nbv0, SV_T, 0, st>>>(n0, z, rO, rOld, m.part);
        fcg_beta_kernel<<<1, SV_T, 0, st>>>(m.part, nbv0, s, 0);
        fcg_dir_kernel<<<nbv0, SV_T, 0, st>>>(n0, s, z, p, 0);
        FL_LAUNCH_CHECK();
    }
    *iterations = it;
    *rel_residual = rel;
    return 0;
}

}  // extern "C"
