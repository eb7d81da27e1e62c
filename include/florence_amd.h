/*
 * florence_amd -- MI355X (gfx950) displacement-formulation assembler, C ABI.
 *
 * Every pointer named *_dev / device below is device (HBM) memory owned by the
 * caller; host pointers are marked.  Calls are asynchronous on the given HIP
 * stream (NULL = default stream) unless documented otherwise.  Return value: 0 on
 * success, a negative FL_E_* code on an argument / launch error (message in
 * fl_last_error()).  Device-side numerical conditions are reported through a
 * caller-owned int32[2] status buffer: status[0] = FL_S_* bit flags, status[1] =
 * number of quadrature points with det F <= 0.
 *
 * Indices are int32 like the reference's CSR (ComputeSparsityPattern.pyx:66-90).
 */
#ifndef FLORENCE_AMD_H
#define FLORENCE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* materials (the reference dispatches on the material class name,
 * FiniteElements/Assembly/_LowLevelAssembly_.py:47-60) */
#define FL_MAT_NEOHOOKEAN 1   /* params: mu, lamb         (_NeoHookean_.h:27-51)  */
#define FL_MAT_MOONEYRIVLIN 2 /* params: mu1, mu2, lamb   (_MooneyRivlin_.h:30-70) */

/* error codes */
#define FL_E_ARG -1     /* invalid argument / unsupported shape     */
#define FL_E_HIP -2     /* HIP runtime error                         */
#define FL_E_SPACE -3   /* workspace too small                       */
#define FL_E_LIMIT -4   /* a size exceeds what int32 indices or the kernels hold */

/* device status bits (status[0]) */
#define FL_S_NONPOSITIVE_J 1 /* some quadrature point has det F <= 0 (the reference masks it with fabs) */
#define FL_S_NONFINITE 2     /* a non-finite stress or tangent value was produced */
#define FL_S_PATTERN 4       /* an element pair was missing from the neighbour lists */

/* ---------------------------------------------------------------- library */
int fl_version(void);
const char *fl_last_error(void);   /* host string, valid until the next failing call */

/* ------------------------------------------------------- sparsity pattern
 * Replaces ComputeSparsityPattern (FiniteElements/Assembly/_Assembly_/
 * ComputeSparsityPattern.pyx:44-112, ComputeSparsityPattern.h:22-62).  Two calls:
 *
 * fl_pattern_graph: node -> element incidence (node_elem_ptr[nnode+1],
 * node_elem[nelem*npe], entries e*npe+local in ascending element order) and the
 * per-node neighbour counts turned into node_nbr_ptr[nnode+1].  Synchronises the
 * stream and returns the node-level neighbour total in *node_nnz (host) and the
 * largest neighbour count in *max_degree (host).
 * workspace: fl_pattern_workspace_size(nnode) bytes of device memory.
 *
 * fl_pattern_fill: sorted unique neighbour lists node_nbrs[node_nnz] and the CSR
 * indptr[3*nnode+1] / indices[9*node_nnz] of the reference pattern (row 3a+i lists
 * 3b+k for every neighbour b of node a in ascending order, k = 0..2).  indices may
 * be NULL.  Returns FL_E_LIMIT if 9*node_nnz does not fit int32.
 */
size_t fl_pattern_workspace_size(int64_t nnode);
int fl_pattern_graph(const int32_t *elements_dev, int64_t nelem, int32_t npe, int64_t nnode,
                     int32_t *node_elem_ptr_dev, int32_t *node_elem_dev, int32_t *node_nbr_ptr_dev,
                     int64_t *node_nnz, int32_t *max_degree, void *workspace_dev, size_t workspace_bytes,
                     void *stream);
int fl_pattern_fill(const int32_t *elements_dev, int64_t nelem, int32_t npe, int64_t nnode,
                    const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev,
                    const int32_t *node_nbr_ptr_dev, int32_t *node_nbrs_dev, int32_t *indptr_dev,
                    int32_t *indices_dev, void *stream);

/* ------------------------------------------------- implicit DF assembly
 * Replaces _GlobalAssemblyDF_<Material> (FiniteElements/Assembly/_Assembly_/
 * _LowLevelAssemblyDF_.h:8-176) in its CSR mode (recompute_sparsity_pattern=0):
 * K+residual of all elements, written (not accumulated) into V[nnz] on the
 * pattern built above and into T[3*nnode].  Values of V are summed per CSR entry
 * in ascending element order, so the result is bitwise reproducible.
 *
 * params (host): {mu, mu1, mu2, lamb}; Jm: (3, npe, ngauss) reference layout;
 * AllGauss: (ngauss).  requires_geometry_update: 1 integrates on the current
 * configuration (the reference's default), 0 on the reference configuration
 * with no traction.  workspace: fl_df_workspace_size() bytes.  status_dev: int32[2],
 * zeroed by the call.
 */
size_t fl_df_workspace_size(int32_t ngauss, int64_t nelem, int32_t npe);
int fl_df_assemble(int32_t material, const double *params, const double *points_dev,
                   const int32_t *elements_dev, const double *Eulerx_dev, const double *Jm_dev,
                   const double *AllGauss_dev, int32_t ngauss, int64_t nelem, int32_t npe, int64_t nnode,
                   int32_t requires_geometry_update, const int32_t *node_elem_ptr_dev,
                   const int32_t *node_elem_dev, const int32_t *node_nbr_ptr_dev,
                   const int32_t *node_nbrs_dev, int32_t max_degree, const int32_t *indptr_dev,
                   double *V_dev, double *T_dev, void *workspace_dev, size_t workspace_bytes,
                   int32_t *status_dev, void *stream);

/* --------------------------------------------- residual-only (explicit) pass
 * Replaces the mechanics branch of _GlobalAssemblyExplicit_DF_DPF_
 * (FiniteElements/Assembly/_Assembly_/_LowLevelAssemblyExplicit_DF_DPF_.h:459-726):
 * T[3*nnode] = internal force, matrix free.  Same workspace rules (a smaller
 * size is enough: fl_df_residual_workspace_size()).
 */
size_t fl_df_residual_workspace_size(int32_t ngauss, int64_t nelem, int32_t npe);
int fl_df_residual(int32_t material, const double *params, const double *points_dev,
                   const int32_t *elements_dev, const double *Eulerx_dev, const double *Jm_dev,
                   const double *AllGauss_dev, int32_t ngauss, int64_t nelem, int32_t npe, int64_t nnode,
                   const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev, double *T_dev,
                   void *workspace_dev, size_t workspace_bytes, int32_t *status_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLORENCE_AMD_H */
