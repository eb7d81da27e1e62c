// Linear algebra of the Newton loop on the device: SpMV on the reference CSR pattern
// and a block-Jacobi preconditioned conjugate gradient for the Dirichlet-reduced
// tangent.
//
// The reference reduces K to the free dofs by scipy fancy indexing
// (BoundaryCondition.py:842-857) and calls a sparse direct solver
// (LinearSolver.py:208-300).  Here the reduction is a mask: the operator is
// A = M K M + (I - M) with M = diag(free), which keeps the CSR untouched; every
// iterate vanishes on fixed dofs.  All reductions are two-pass with a fixed order,
// so a solve is bitwise reproducible.
#include <cmath>
#include <vector>

#include "fl_common.hpp"
#include "fl_solve_kernels.hpp"

namespace fl {

template <typename VT>
static int csr_spmv(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const VT *V,
                    int64_t nnode, const uint8_t *free_mask, const double *x, double *y, void *stream,
                    const char *name) {
    if (!indptr || !node_nbr_ptr || !node_nbrs || !V || !x || !y || nnode <= 0) {
        set_error("%s: invalid argument", name);
        return FL_E_ARG;
    }
    hipStream_t st = as_stream(stream);
    spmv_kernel<VT><<<(unsigned)((nnode + SP_W - 1) / SP_W), 64 * SP_W, 0, st>>>(indptr, node_nbr_ptr, node_nbrs,
                                                                                V, nnode, free_mask, x, y, nullptr);
    FL_LAUNCH_CHECK();
    return 0;
}

template <typename VT>
static int pcg(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const VT *V,
               int64_t nnode, const uint8_t *free_mask, const double *b, double *x, double rtol, int32_t maxit,
               int32_t check_every, void *workspace, size_t workspace_bytes, int32_t *iterations,
               double *rel_residual, void *stream, const char *name) {
    hipStream_t st = as_stream(stream);
    if (!indptr || !node_nbr_ptr || !node_nbrs || !V || !b || !x || nnode <= 0 || maxit <= 0 || !iterations ||
        !rel_residual) {
        set_error("%s: invalid argument", name);
        return FL_E_ARG;
    }
    if (check_every <= 0) check_every = 1;
    const int64_t n = 3 * nnode;
    const int64_t nb_sp = (nnode + SP_W - 1) / SP_W, nb_v = (nnode + SV_T - 1) / SV_T;
    Carve cv(workspace, workspace_bytes);
    double *P = cv.take<double>(9 * nnode);
    double *r = cv.take<double>(n);
    double *z = cv.take<double>(n);
    double *p = cv.take<double>(n);
    double *q = cv.take<double>(n);
    double *part = cv.take<double>(nb_sp + 2 * nb_v + 8);
    double *s = cv.take<double>(8);
    double *p2 = cv.take<double>(n);
    if (!cv.ok) {
        set_error("%s: workspace too small (%zu < %zu)", name, workspace_bytes, cv.off);
        return FL_E_SPACE;
    }
    int32_t *bad = reinterpret_cast<int32_t *>(s + 7);
    FL_HIP_CHECK(hipMemsetAsync(s, 0, sizeof(double) * 8, st));
    block_jacobi_kernel<VT><<<(unsigned)((nnode + 255) / 256), 256, 0, st>>>(indptr, node_nbr_ptr, node_nbrs, V, nnode,
                                                                        free_mask, P, bad);
    FL_LAUNCH_CHECK();
    cg_start_kernel<<<(unsigned)nb_v, SV_T, 0, st>>>(nnode, P, b, free_mask, x, r, z, p, part);
    FL_LAUNCH_CHECK();
    reduce2_kernel<<<1, SV_T, 0, st>>>(part, nb_v, s + 0, s + 6);   // rz, bb (= r0.r0)
    FL_LAUNCH_CHECK();
    double hs[8];
    FL_HIP_CHECK(hipMemcpyAsync(hs, s, sizeof(hs), hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(hipStreamSynchronize(st));
    if (*reinterpret_cast<int32_t *>(&hs[7])) {
        set_error("%s: singular diagonal block or missing diagonal entry", name);
        return FL_E_ARG;
    }
    const double bnorm = std::sqrt(hs[6]);
    *iterations = 0;
    *rel_residual = 0.0;
    if (bnorm == 0.0) return 0;
    double rel = 1.0;
    int it = 0;
    // four launches per iteration; s[4] (beta) is 0 before the first, so p = z there
    double *pb[2] = {p, p2};
    while (it < maxit) {
        double *p_old = pb[it & 1], *p_new = pb[(it + 1) & 1];
        pcg_spmv_kernel<VT><<<(unsigned)nb_sp, 64 * SP_W, 0, st>>>(indptr, node_nbr_ptr, node_nbrs, V, nnode,
                                                                   free_mask, z, p_old, p_new, q, part, s);
        pcg_alpha_kernel<<<1, SV_T, 0, st>>>(part, nb_sp, s);
        pcg_update_kernel<<<(unsigned)nb_v, SV_T, 0, st>>>(nnode, s, P, p_new, q, x, r, z, part + nb_sp);
        pcg_beta_kernel<<<1, SV_T, 0, st>>>(part + nb_sp, nb_v, s);
        FL_LAUNCH_CHECK();
        ++it;
        if (it % check_every == 0 || it == maxit) {
            double rr = 0.0;
            FL_HIP_CHECK(hipMemcpyAsync(&rr, s + 5, sizeof(double), hipMemcpyDeviceToHost, st));
            FL_HIP_CHECK(hipStreamSynchronize(st));
            rel = std::sqrt(rr) / bnorm;
            if (!(rel == rel)) break;
            if (rel <= rtol) break;
        }
    }
    *iterations = it;
    *rel_residual = rel;
    return 0;
}


}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_pcg_workspace_size(int64_t nnode) {
    const int64_t n = 3 * nnode;
    const int64_t nb_sp = (nnode + SP_W - 1) / SP_W, nb_v = (nnode + SV_T - 1) / SV_T;
    return align_up(sizeof(double) * 9 * nnode, 256) + 5 * align_up(sizeof(double) * n, 256) +
           align_up(sizeof(double) * (nb_sp + 2 * nb_v + 8), 256) + align_up(sizeof(double) * 8, 256) + 1024;
}

int fl_csr_spmv(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const double *V,
                int64_t nnode, const uint8_t *free_mask, const double *x, double *y, void *stream) {
    return csr_spmv<double>(indptr, node_nbr_ptr, node_nbrs, V, nnode, free_mask, x, y, stream, "fl_csr_spmv");
}

int fl_csr_spmv_v32(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const float *V,
                    int64_t nnode, const uint8_t *free_mask, const double *x, double *y, void *stream) {
    return csr_spmv<float>(indptr, node_nbr_ptr, node_nbrs, V, nnode, free_mask, x, y, stream, "fl_csr_spmv_v32");
}

int fl_pcg(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const double *V,
           int64_t nnode, const uint8_t *free_mask, const double *b, double *x, double rtol, int32_t maxit,
           int32_t check_every, void *workspace, size_t workspace_bytes, int32_t *iterations, double *rel_residual,
           void *stream) {
    return pcg<double>(indptr, node_nbr_ptr, node_nbrs, V, nnode, free_mask, b, x, rtol, maxit, check_every,
                       workspace, workspace_bytes, iterations, rel_residual, stream, "fl_pcg");
}

int fl_pcg_v32(const int32_t *indptr, const int32_t *node_nbr_ptr, const int32_t *node_nbrs, const float *V,
               int64_t nnode, const uint8_t *free_mask, const double *b, double *x, double rtol, int32_t maxit,
               int32_t check_every, void *workspace, size_t workspace_bytes, int32_t *iterations,
               double *rel_residual, void *stream) {
    return pcg<float>(indptr, node_nbr_ptr, node_nbrs, V, nnode, free_mask, b, x, rtol, maxit, check_every,
                      workspace, workspace_bytes, iterations, rel_residual, stream, "fl_pcg_v32");
}

}  // extern "C"
