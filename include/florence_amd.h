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
/* params: mu, lamb (_LinearElastic_.h:26-52); no geometric stiffness, as the generated
 * _GlobalAssemblyDF__LinearElastic_ (AOT_Assembler.py:84-93) */
#define FL_MAT_LINEARELASTIC 3
/* params: mu1, mu2, lamb, taken as alpha, beta, kappa of
 * _NearlyIncompressibleMooneyRivlin_.h:28-87, the constants the generated assembler
 * constructs the law from (AOT_Assembler.py:69-70) */
#define FL_MAT_NEARLYINCOMPRESSIBLEMOONEYRIVLIN 4
/* params: mu1, mu2, lamb and mu3 = params[4] (_AnisotropicMooneyRivlin_1_.h:30-95); the
 * element's fibre direction N comes from the per-element material table only (the
 * reference passes anisotropic_orientations[elem], AOT_Assembler.py:78-82), so the
 * law needs the _mt / _cs entry points with a table of stride 8 */
#define FL_MAT_ANISOTROPICMOONEYRIVLIN_1 5
/* doubles per row of a per-element material table: {mu, mu1, mu2, lamb} (stride 4) or,
 * for FL_MAT_ANISOTROPICMOONEYRIVLIN_1, {mu, mu1, mu2, lamb, mu3, N0, N1, N2} (stride 8);
 * 0 for an unknown id.  No reference counterpart. */
int32_t fl_material_table_stride(int32_t material);

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

/* Gather plan of the row-owner accumulation: node ids grouped by neighbour count
 * into FL_GATHER_BUCKETS classes (the nodes of the smallest count -- element-interior
 * nodes where the mesh has them -- then the others by at most 128, 256, 512, 1024,
 * 2048 neighbours), so each class is accumulated with an LDS footprint and kernel
 * shape sized to it.  node_order_dev:
 * int32[nnode]; bucket_offsets (host): int32[FL_GATHER_BUCKETS+1]; bucket_degree
 * (host): int32[FL_GATHER_BUCKETS], the largest neighbour count in each class.
 * Synchronous.  No reference counterpart (part of the pattern set-up above). */
#define FL_GATHER_BUCKETS 6
int fl_pattern_gather_plan(const int32_t *node_nbr_ptr_dev, int64_t nnode, int32_t *node_order_dev,
                           int32_t *bucket_offsets, int32_t *bucket_degree, void *stream);

/* Column-position table of the gather (no reference counterpart; the reference looks
 * positions up per call through its data_global_indices maps, FEMSolver.py:1659-1684):
 * ranks_dev[(e*npe + a)*npe + b] = position of element e's node b in the sorted
 * neighbour list of its node a.  int16 (neighbour lists up to 32767); bytes from
 * fl_pattern_rank_table_bytes.  status_dev bit FL_S_PATTERN on an inconsistent
 * pattern (the caller zeroes it first).  Built once per mesh, like the pattern. */
size_t fl_pattern_rank_table_bytes(int64_t nelem, int32_t npe);
int fl_pattern_rank_table(const int32_t *elements_dev, int64_t nelem, int32_t npe,
                          const int32_t *node_nbr_ptr_dev, const int32_t *node_nbrs_dev, int16_t *ranks_dev,
                          int32_t *status_dev, void *stream);

/* Element colouring of the direct scatter (no reference counterpart: the reference sums
 * duplicate COO entries on the host).  Greedy first fit in element order: element e takes
 * the smallest colour no earlier element sharing a node with it holds, so no two elements
 * of one colour share a node.  Host arrays: elements_host int32[nelem][npe];
 * colour_of_elem_host int32[nelem]; colour_elems_host int32[nelem], the elements grouped
 * by colour (ascending within a colour); colour_offsets_host int32[max_colours + 1], of
 * which [0, ncolour] are set.  max_colours <= 64; FL_E_LIMIT if more are needed. */
int fl_pattern_colour(const int32_t *elements_host, int64_t nelem, int32_t npe, int64_t nnode,
                      int32_t max_colours, int32_t *colour_of_elem_host, int32_t *colour_elems_host,
                      int32_t *colour_offsets_host, int32_t *ncolour);

/* Scatter table of the direct form (no reference counterpart), for hexes of
 * npe = p1^3 nodes with their tensor basis: for every entry the factorised stiffness
 * kernel writes, in the order it writes them, the position of the column node b in the
 * sorted neighbour list of the row node a, with the sign bit set when an element of a
 * lower colour holds both a and b.  Order: entry
 * ((((e*p1 + az)*nt + s)*64 + lane)*mt + tm)*4 + r with mt = ceil(p1^2/16),
 * nt = ceil(p1^3/16), row = 16*tm + lane/16 + 4*r = ab*p1 + bb and
 * col = 16*s + lane%16 = (ae*p1 + be)*p1 + bz: a and b are the element's tensor nodes
 * (ab*p1 + ae)*p1 + az and (bb*p1 + be)*p1 + bz (local node local_of_t of the basis);
 * slots with row >= p1^2 or col >= p1^3 hold 0.  colour_of_elem_dev: fl_pattern_colour's
 * colours on the device.  status_dev bit FL_S_PATTERN on an inconsistent pattern,
 * basis or colouring (the caller zeroes it first).  0 bytes for other element sizes. */
size_t fl_pattern_scatter_table_bytes(int64_t nelem, int32_t npe);
int fl_pattern_scatter_table(const int32_t *elements_dev, int64_t nelem, int32_t npe,
                             const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev,
                             const int32_t *node_nbr_ptr_dev, const int32_t *node_nbrs_dev,
                             const int32_t *colour_of_elem_dev, const void *tensor_basis_dev,
                             int16_t *table_dev, int32_t *status_dev, void *stream);

/* ------------------------------------------------- high-order meshes
 * Replaces GetHighOrderMesh (MeshGeneration/Mesh.py:1339-1409) -> HighOrderMeshHex /
 * HighOrderMeshTet_SEMISTABLE (HigherOrderMeshing/HigherOrderMeshingHex.py:19-133,
 * HigherOrderMeshingTet.py:19-133) with Decimals = decimals and check_duplicates: the
 * p-refined mesh of a linear one with the reference's node numbering.  Every element's
 * non-vertex nodes are placed at Neval^T x_vertices (Neval: f64 [nv][npe], the linear
 * basis at the element's p-nodes in local order, values below 1e-12 set to 0, device);
 * the new nodes are merged on their coordinates rounded to `decimals`, the whole list
 * again at decimals - 2, and every node numbered by the byte order of its rounded
 * coordinates (numpy's unique of raw-byte rows, which the reference calls).  Inputs:
 * points_lin f64 [nnode_lin][3], elements_lin int32 [nelem][nv] (ids < nnode_lin).
 * Outputs: points f64 [*nnode][3] (capacity nnode_lin + nelem (npe - nv) rows),
 * elements int32 [nelem][npe]; *nnode (host).  Synchronous.
 * workspace: fl_mesh_refine_workspace_size() bytes. */
size_t fl_mesh_refine_workspace_size(int64_t nnode_lin, int64_t nelem, int32_t nv, int32_t npe);
int fl_mesh_refine(const double *points_lin_dev, int64_t nnode_lin, const int32_t *elements_lin_dev, int64_t nelem,
                   int32_t nv, const double *Neval_dev, int32_t npe, int32_t decimals, double *points_dev,
                   int32_t *elements_dev, int64_t *nnode, void *workspace_dev, size_t workspace_bytes, void *stream);

/* ------------------------------------------------- implicit DF assembly
 * Replaces _GlobalAssemblyDF_<Material> (FiniteElements/Assembly/_Assembly_/
 * _LowLevelAssemblyDF_.h:8-176) in its CSR mode (recompute_sparsity_pattern=0):
 * K+residual of all elements, written (not accumulated) into V[nnz] on the
 * pattern built above and into T[3*nnode].  Values of V are summed per CSR entry
 * in ascending element order, so the result is bitwise reproducible.
 *
 * params (host): {mu, mu1, mu2, lamb} (mu3 in params[4] for the anisotropic law);
 * Jm: (3, npe, ngauss) reference layout;
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
                   const int32_t *node_nbrs_dev, const int32_t *node_order_dev,
                   const int32_t *bucket_offsets, const int32_t *bucket_degree, const int32_t *indptr_dev,
                   double *V_dev, double *T_dev, void *workspace_dev, size_t workspace_bytes,
                   int32_t *status_dev, void *stream);

/* Mixed-precision form of fl_df_assemble (no reference counterpart: the reference
 * is FP64 only).  precision = FL_PREC_FP64 is fl_df_assemble; FL_PREC_FP32 runs the
 * element-block contraction on FP32 MFMA with FP32 accumulation and keeps the element
 * blocks in FP32.  Kinematics, det F, the material law, the residual T and the CSR
 * sums (V) stay FP64.  workspace: fl_df_workspace_size_ex() bytes. */
#define FL_PREC_FP64 0
#define FL_PREC_FP32 1
#define FL_PREC_BF16 2   /* G and W rounded to BF16, v_mfma_f32_16x16x32_bf16, FP32 accumulation */
/* OR'ed into the precision of fl_df_assemble_mt with FL_PREC_FP32 / FL_PREC_BF16: the
 * CSR values are summed in FP64 as always but stored as FP32, so V_dev points to
 * float[nnz] (the operand of fl_csr_spmv_v32 / fl_pcg_v32 in an inexact Newton
 * method; half the bytes per Krylov iteration). */
#define FL_PREC_CSR32 16
size_t fl_df_workspace_size_ex(int32_t ngauss, int64_t nelem, int32_t npe, int32_t precision);
int fl_df_assemble_ex(int32_t material, const double *params, const double *points_dev,
                      const int32_t *elements_dev, const double *Eulerx_dev, const double *Jm_dev,
                      const double *AllGauss_dev, int32_t ngauss, int64_t nelem, int32_t npe, int64_t nnode,
                      int32_t requires_geometry_update, const int32_t *node_elem_ptr_dev,
                      const int32_t *node_elem_dev, const int32_t *node_nbr_ptr_dev,
                      const int32_t *node_nbrs_dev, const int32_t *node_order_dev,
                      const int32_t *bucket_offsets, const int32_t *bucket_degree, const int32_t *indptr_dev,
                      double *V_dev, double *T_dev, int32_t precision, void *workspace_dev,
                      size_t workspace_bytes, int32_t *status_dev, void *stream);

/* Per-element materials (no reference counterpart: the reference assembles one
 * material per call).  elem_material_dev: int32[nelem] rows into material_table_dev:
 * f64[nmat][fl_material_table_stride(material)] for the law `material`; both NULL =
 * params (not allowed for FL_MAT_ANISOTROPICMOONEYRIVLIN_1, whose rows carry N).
 * Used to assemble many meshes as one (batched form).  ranks_dev: the table of
 * fl_pattern_rank_table, or NULL (the gather then searches the neighbour lists on
 * every call; same result).  Issues no memset, so the call can be captured into a HIP
 * graph (the workspace must then stay fixed). */
int fl_df_assemble_mt(int32_t material, const double *params, const int32_t *elem_material_dev,
                      const double *material_table_dev, const double *points_dev, const int32_t *elements_dev,
                      const double *Eulerx_dev, const double *Jm_dev, const double *AllGauss_dev, int32_t ngauss,
                      int64_t nelem, int32_t npe, int64_t nnode, int32_t requires_geometry_update,
                      const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev,
                      const int32_t *node_nbr_ptr_dev, const int32_t *node_nbrs_dev, const int32_t *node_order_dev,
                      const int32_t *bucket_offsets, const int32_t *bucket_degree, const int32_t *indptr_dev,
                      double *V_dev, double *T_dev, int32_t precision, const int16_t *ranks_dev,
                      const void *tensor_basis_dev, void *workspace_dev, size_t workspace_bytes,
                      int32_t *status_dev, void *stream);

/* Tensor basis of Lagrange hexes under a tensor-product rule (no reference
 * counterpart): with it, fl_df_assemble_mt computes FP64 element blocks by sum
 * factorisation (one direction of the point sum at a time) instead of the dense
 * contraction; same K to rounding.  Layout (fl_tensor_basis_bytes(p1, q) bytes):
 *   int32 p1, q, 0, 0; int32 g_of_t[q^3] (flattened point of tensor point
 *   (qb*q + qe)*q + qz); int32 local_of_t[p1^3] (local node of tensor node
 *   (ab*p1 + ae)*p1 + az); int32 tensor_of_l[p1^3] (its inverse); zero int32
 *   padding to 8 bytes; double f1d[3][2][p1][q]
 *   (the 1D bases and their derivatives of directions zeta, eta, beta at that
 *   direction's points).  Supported (p1, q): (2,2) (3,3) (4,4) (5,5) (5,6).  NULL, an
 *   FP32/BF16 precision or fl_set_tuning(FL_TUNE_STIFF_SF, 1) select the dense path;
 *   values 3 (all-VALU contraction) and 5 (MFMA, one basis-factor type per reduction)
 *   select other forms of the factorised kernel (default: MFMA, two types per
 *   reduction). */
#define FL_TUNE_STIFF_SF 7
size_t fl_tensor_basis_bytes(int32_t p1, int32_t q);

/* fl_df_assemble_mt with the direct scatter (no reference counterpart; replaces the
 * same _GlobalAssemblyDF_ call and its duplicate-summing COO -> CSR conversion,
 * FiniteElements/Assembly/Assembly.py:207-208): with the
 * tensor basis and FL_PREC_FP64, the factorised stiffness kernel writes every CSR entry
 * itself, one launch per colour, and no element block is stored or gathered.  Each entry
 * is the sum of its elements' contributions in colour order (bitwise reproducible; equal
 * to fl_df_assemble_mt's V to rounding).  ncolour, colour_offsets (host),
 * colour_elems_dev: fl_pattern_colour's plan; scatter_dev: fl_pattern_scatter_table
 * built from the same colouring and the same tensor basis (a table of another
 * colouring gives wrong sums; the table build itself reports a colouring in which
 * elements of one colour share a node).  workspace: fl_df_workspace_size_cs() bytes on the
 * direct path; where it does not apply (other precisions, no basis, the other
 * FL_TUNE_STIFF_SF forms) the call is fl_df_assemble_mt with its workspace size. */
size_t fl_df_workspace_size_cs(int32_t ngauss, int64_t nelem, int32_t npe);
int fl_df_assemble_cs(int32_t material, const double *params, const int32_t *elem_material_dev,
                      const double *material_table_dev, const double *points_dev, const int32_t *elements_dev,
                      const double *Eulerx_dev, const double *Jm_dev, const double *AllGauss_dev, int32_t ngauss,
                      int64_t nelem, int32_t npe, int64_t nnode, int32_t requires_geometry_update,
                      const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev,
                      const int32_t *node_nbr_ptr_dev, const int32_t *node_nbrs_dev, const int32_t *node_order_dev,
                      const int32_t *bucket_offsets, const int32_t *bucket_degree, const int32_t *indptr_dev,
                      double *V_dev, double *T_dev, int32_t precision, const int16_t *ranks_dev,
                      const void *tensor_basis_dev, int32_t ncolour, const int32_t *colour_offsets,
                      const int32_t *colour_elems_dev, const int16_t *scatter_dev, void *workspace_dev,
                      size_t workspace_bytes, int32_t *status_dev, void *stream);

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
/* The same pass given the tensor basis of fl_tensor_basis_bytes (no reference
 * counterpart): tensor-product hexes then take the sum-factorised point and traction
 * kernels (same T to rounding); NULL behaves as fl_df_residual. */
int fl_df_residual_tb(int32_t material, const double *params, const double *points_dev,
                      const int32_t *elements_dev, const double *Eulerx_dev, const double *Jm_dev,
                      const double *AllGauss_dev, int32_t ngauss, int64_t nelem, int32_t npe, int64_t nnode,
                      const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev, double *T_dev,
                      const void *tensor_basis_dev, void *workspace_dev, size_t workspace_bytes,
                      int32_t *status_dev, void *stream);

/* The same pass with a per-element material table (as fl_df_assemble_mt; required for
 * FL_MAT_ANISOTROPICMOONEYRIVLIN_1, whose rows carry the fibre directions, as the
 * reference's explicit assembler takes anisotropic_orientations,
 * _LowLevelAssemblyExplicit_DF_DPF_.h:484) and the optional tensor basis. */
int fl_df_residual_mt(int32_t material, const double *params, const int32_t *elem_material_dev,
                      const double *material_table_dev, const double *points_dev, const int32_t *elements_dev,
                      const double *Eulerx_dev, const double *Jm_dev, const double *AllGauss_dev, int32_t ngauss,
                      int64_t nelem, int32_t npe, int64_t nnode, const int32_t *node_elem_ptr_dev,
                      const int32_t *node_elem_dev, double *T_dev, const void *tensor_basis_dev, void *workspace_dev,
                      size_t workspace_bytes, int32_t *status_dev, void *stream);

/* fl_df_residual_mt without a workspace argument: the pass's workspace (the transposed
 * Jm, the element tractions and, off the sum-factorised path, the point records) is a
 * fixed device array of FL_NOALLOC_SCRATCH_BYTES inside the library, one per GPU, held
 * until the process ends.  No reference counterpart: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); the call allocates nothing.  Sizes whose workspace exceeds the maximum
 * return FL_E_SPACE (fl_df_residual_workspace_size gives the need of the dense form;
 * a 24^3 p=4 hex cube with its tensor basis needs 42 MB).  Calls that overlap on one GPU
 * share the array and race; the device is the current one at the call. */
#define FL_NOALLOC_SCRATCH_BYTES (64u << 20)
int fl_df_residual_noalloc(int32_t material, const double *params, const int32_t *elem_material_dev,
                           const double *material_table_dev, const double *points_dev, const int32_t *elements_dev,
                           const double *Eulerx_dev, const double *Jm_dev, const double *AllGauss_dev,
                           int32_t ngauss, int64_t nelem, int32_t npe, int64_t nnode,
                           const int32_t *node_elem_ptr_dev, const int32_t *node_elem_dev, double *T_dev,
                           const void *tensor_basis_dev, int32_t *status_dev, void *stream);

/* ------------------------------------------------------ p-multigrid solve
 * No reference counterpart: the reference solves every Newton step with a sparse
 * direct solver (Solver/LinearSolver.py:208-300).  fl_mg_pcg solves A x = b for the
 * masked operator of fl_pcg by flexible CG preconditioned with one V-cycle over nested
 * nodal spaces on the mesh's elements (level 0 = the assembled CSR, coarser levels =
 * the p=2 lattice or the vertices of each element), Chebyshev-smoothed (degree
 * cheb_degree on [lmax/cheb_range, lmax] of the block-Jacobi-scaled operator, lmax by
 * power iteration per call, below), with coarse_iterations block-Jacobi CG steps on the
 * coarsest level.  Stopping, b, x, iterations and rel_residual as fl_pcg.
 *
 * A coarse level's operator is the Galerkin product P^T K P of the fine tangent;
 * fl_mg_galerkin forms it from the point records the last fl_df_assemble_* call on a
 * workspace left there (so it must follow the assembly whose V is level 0, on the same
 * workspace): records = fl_df_records_form(precision, tensor_basis, npe, ngauss) of
 * that assembly; JmCt: f64 [ngauss][npe_c][3], the coarse basis gradients dNc_a/dxi_m
 * at the fine quadrature points; the coarse mesh (elements_c [nelem][npe_c], npe_c =
 * 4, 8 or 27) and its pattern, gather plan and optional rank table are built by the
 * fl_pattern_* calls above.  workspace: fl_mg_galerkin_workspace_size() bytes.
 * tensor_basis_c (optional): the coarse level's own fl_tensor_basis table (a hex level:
 * its degree's 1D Lagrange bases through the lattice {-1, 0, 1} or {-1, 1} at the fine
 * rule's 1D points, g_of_t of the fine table, its local/tensor node maps); with C
 * records the operator is then formed by the factorised stiffness kernel of that space
 * (27 or 8 coarse nodes over 6^3, 5^3, 4^3 or 3^3 points) instead of point by point.
 */
#define FL_RECORDS_D 0   /* dxi/dx and the D blocks (dense point kernel) */
#define FL_RECORDS_C 1   /* the C tensors of the factorised FP64 path */
int32_t fl_df_records_form(int32_t precision, const void *tensor_basis_dev, int32_t npe, int32_t ngauss);
size_t fl_mg_galerkin_workspace_size(int64_t nelem, int32_t npe_c, int64_t nnode_c);
int fl_mg_galerkin(const void *df_workspace_dev, size_t df_workspace_bytes, int32_t ngauss, int64_t nelem,
                   int32_t npe, int32_t records, const double *JmCt_dev, const void *tensor_basis_c_dev,
                   int32_t npe_c, const int32_t *elements_c_dev,
                   int64_t nnode_c, const int32_t *node_elem_ptr_c_dev, const int32_t *node_elem_c_dev,
                   const int32_t *node_nbr_ptr_c_dev, const int32_t *node_nbrs_c_dev, const int32_t *node_order_c_dev,
                   const int32_t *bucket_offsets_c, const int32_t *bucket_degree_c, const int32_t *indptr_c_dev,
                   const int16_t *ranks_c_dev, double *V_c_dev, void *workspace_dev, size_t workspace_bytes,
                   int32_t *status_dev, void *stream);

/* One level of the hierarchy (device pointers).  V: double[nnz] (float on level 0 when
 * fl_mg_pcg's v32 is set).  P_ptr/P_idx/P_w: the prolongation from this level to the
 * next finer one, as CSR over the FINER level's nodes (node f of the finer level takes
 * sum_j P_w[j] u[P_idx[j]] of this level's nodal values); R_*: its transpose, CSR over
 * this level's nodes.  Unused on level 0. */
#define FL_MG_MAX_LEVELS 4
typedef struct {
    int64_t nnode;
    const int32_t *indptr;
    const int32_t *node_nbr_ptr;
    const int32_t *node_nbrs;
    const void *V;
    const uint8_t *free_mask;
    const int32_t *P_ptr;
    const int32_t *P_idx;
    const double *P_w;
    const int32_t *R_ptr;
    const int32_t *R_idx;
    const double *R_w;
} FlMgLevel;
/* The assembled tangent applied without its CSR (no reference counterpart): for a
 * tensor-product hex (npe = p1^3, ngauss = q^3, its tensor basis table; records form
 * FL_RECORDS_C), K x = sum_e K_e x_e with
 *   K_e^{ik}_{ab} = sum_g sum_{m,n} dN_a/dxi_m(g) C^{ik}_{mn}(g) dN_b/dxi_n(g)
 * evaluated by sum factorisation from the C tensors the last fl_df_assemble_* call on
 * df_workspace left in its point records: the per-point tangent that call contracted
 * into V, so y is V's product to rounding.  Element vectors go to element_vectors
 * (double[nelem * 3 * npe]); each node sums its elements' entries in ascending element
 * order (node_elem_ptr / node_elem of fl_pattern_graph), so a product is bitwise
 * reproducible.  Masked as the operator of fl_csr_spmv: x is read as 0 on fixed dofs
 * (free_mask 0) and y is 0 there. */
typedef struct {
    const void *df_workspace;
    size_t df_workspace_bytes;
    int32_t records;
    const void *tensor_basis;
    const int32_t *elements;
    int64_t nelem;
    int32_t npe;
    int32_t ngauss;
    int64_t nnode;
    const int32_t *node_elem_ptr;
    const int32_t *node_elem;
    double *element_vectors;
    float *records32;                /* optional, float[nelem * 54 * ngauss]: see fl_mg_pcg */
    uint32_t *records16;             /* optional, uint32[nelem * 27 * ngauss]: see fl_mg_pcg */
} FlMgMatrixFree;
int fl_mf_apply(const FlMgMatrixFree *op, const uint8_t *free_mask_dev, const double *x_dev, double *y_dev,
                void *stream);

/* smoother_V32 (optional host array of nlevels device pointers): smoother_V32[l] =
 * level l's values rounded to FP32 on the same pattern, read by level l's smoother and
 * its spectrum estimate (the outer iteration, the block-Jacobi blocks and the coarsest
 * level's CG apply the levels' own FP64 values); a NULL array or entry: the smoother
 * reads the level's own values (level 0 of v32 = 1 is FP32 already).
 * fine_matrix_free (optional): the fine level's smoother and spectrum estimate apply
 * level 0's operator in the matrix-free form above instead (smoother_V32[0] is then
 * not read); the outer iteration and the block-Jacobi blocks keep level 0's CSR.  With
 * its records32 given, the call first rounds the C tensors of the records to FP32 there
 * and the smoother reads those (half the bytes); with records16 given instead, to BF16
 * (round to nearest even), two values per word (a quarter of the bytes).  fl_mf_apply
 * always reads the FP64 records.
 * power_steps: power iterations per level for the smoother's lmax (1..64); power_warm:
 * nonzero = start them from the vectors the previous call on this workspace left (same
 * levels; the previous Newton step's operator), 0 = from a fixed pseudo-random start.  A
 * warm vector that is zero on this call's free dofs (another free mask) makes that level
 * start again from the pseudo-random vector with max(power_steps, 12) steps. */
size_t fl_mg_workspace_size(const int64_t *nnodes, int32_t nlevels);
int fl_mg_pcg(const FlMgLevel *levels, int32_t nlevels, int32_t v32, const float *const *smoother_V32,
              const FlMgMatrixFree *fine_matrix_free, const double *b_dev, double *x_dev, double rtol,
              int32_t maxit, int32_t check_every, int32_t cheb_degree, double cheb_range,
              int32_t coarse_iterations, int32_t power_steps, int32_t power_warm, void *workspace_dev,
              size_t workspace_bytes, int32_t *iterations, double *rel_residual, void *stream);

/* ------------------------------------------------------------- profiling
 * No reference counterpart (the reference only records wall time,
 * Assembly.py:39,93).  When enabled, every fl_df_assemble / fl_df_residual call
 * records HIP events on its stream at its phase boundaries.  fl_profile_read waits
 * for them and returns, per recorded call, its kind (1 assemble, 2 residual) and the
 * phase durations in ms -- assemble: qp, stiffness, traction, gather; residual: qp,
 * traction, gather (unused phases -1) -- then forgets them.  Returns the call count.
 */
int fl_profile_enable(int on);

/* Kernel-variant selection (no reference counterpart; the reference picks variants
 * at build time, setup.py:247-251).  key 1 = minimum waves per SIMD the p=4 hex
 * stiffness kernel is compiled for (1 or 2).  value 0 restores the default. */
#define FL_TUNE_STIFF_MINW 1
/* keys FL_TUNE_STIFF_VARIANT_BASE + c, c = padded node-count class (16, 32, 48, 64, 128
 * nodes -> 0..4): value v+1 selects tile-shape variant v of the dense FP64/FP32
 * stiffness kernel, written (MT x NT tiles of 16 x 16 per wave, WGM x WGN waves):
 * 32: v0 2x2/1x1, v1 1x1/2x2, v2 2x1/1x2, v3 1x2/2x1; 48: v0 3x3/1x1, v1 1x3/3x1,
 * v2 3x1/1x3; 64: v0 2x2/2x2, v1 4x4/1x1, v2 4x2/1x2, v3 2x4/2x1, v4 2x1/2x4;
 * 128: v0 default (off-diagonal pairs 4x4/2x2, then the diagonal triangles), v1 one
 * wave per SIMD, all pairs square, v2 off-diagonal pairs 8x2/1x4.  Used by the
 * autotuner; fl_stiffness_variant_count(c) gives the number of variants of class c
 * (0 for an unknown class).  No reference counterpart. */
#define FL_TUNE_STIFF_VARIANT_BASE 2
int32_t fl_stiffness_variant_count(int32_t np_class);
/* key 8: gather form -- bit 0: nodes of several elements add one element's rows at a
 * time instead of loading eight elements' rows first; bit 1: nodes inside one element
 * go through the general kernel instead of the single-element one.  Same V and T. */
#define FL_TUNE_GATHER 8
int fl_set_tuning(int32_t key, int32_t value);
int fl_profile_read(int32_t *kinds, double *phase_ms, int32_t max_calls, int32_t max_phases);

/* ------------------------------------------------------------- warm-start MLP
 * No reference counterpart (Florence starts each increment's Newton iteration from
 * the previous solution plus the Dirichlet increment, FEMSolver.py:932-946).  A
 * 3-layer MLP in -> 128 -> 128 -> out with tanh, FP32 master weights, BF16 MFMA
 * products accumulated in FP32.  Parameters P (FP32, fl_mlp_param_count() floats) in
 * the order W1[32][128] b1[128] W2[128][128] b2[128] W3[128][32] b3[32]; inputs X
 * [N][32] and outputs Y [N][32] row-major FP32 (zero-padded columns).
 *
 * fl_mlp_forward: H1, H2 [N][128], Y [N][32].
 * fl_mlp_backward: gradient Gr (FP32, param-count floats) of the loss
 *   sum_{n, c < nout} (Y - Tgt)^2 / (nout N) at the activations of a forward call,
 *   and the loss in device *loss.  N must be a multiple of 32.  Deterministic.
 * fl_mlp_adam: one Adam step (bias-corrected, step >= 1) on P with moments m, v.
 * fl_warmstart_features: F [nnode][32] = standardised (X, u1, u1 - u2, load_factor)
 *   with norm = device {mean[10], inv_std[10]}.
 * fl_warmstart_apply: u = free ? 2 u1 - u2 + out_norm[0..2] * Y + out_norm[3..5] : u_bc
 *   per dof (free_mask uint8 [3 nnode]).
 */
int64_t fl_mlp_param_count(void);
size_t fl_mlp_workspace_size(int64_t n);
int fl_mlp_forward(const float *P_dev, const float *X_dev, int64_t n, float *H1_dev, float *H2_dev, float *Y_dev,
                   void *stream);
int fl_mlp_backward(const float *P_dev, const float *X_dev, const float *H1_dev, const float *H2_dev,
                    const float *Y_dev, const float *Tgt_dev, int64_t n, int32_t nout, float *Gr_dev, float *loss_dev,
                    void *workspace_dev, size_t workspace_bytes, void *stream);
int fl_mlp_adam(float *P_dev, const float *Gr_dev, float *m_dev, float *v_dev, float lr, float beta1, float beta2,
                float eps, int32_t step, void *stream);
int fl_warmstart_features(const double *X_dev, const double *u1_dev, const double *u2_dev, double load_factor,
                          int64_t nnode, const float *norm_dev, float *F_dev, void *stream);
int fl_warmstart_apply(const double *u1_dev, const double *u2_dev, const float *Y_dev, int64_t nnode,
                       const float *out_norm_dev, const uint8_t *free_mask_dev, const double *u_bc_dev,
                       double *u_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLORENCE_AMD_H */
