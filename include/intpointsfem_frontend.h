/*
 * intpointsfem_frontend.h -- C ABI of the host front-end: the mesh stand-in
 * (TPZCompMesh) and the drop-in structural matrix
 * TPZElastoPlasticIntPointsStructMatrix built on include/intpointsfem.h.
 * Same conventions as intpointsfem.h (int status, ipf_global_last_error()).
 */
#ifndef INTPOINTSFEM_FRONTEND_H
#define INTPOINTSFEM_FRONTEND_H

#include <stdint.h>

#include "intpointsfem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ipf_cmesh ipf_cmesh;
typedef struct ipf_sm ipf_sm;

/* Structured quarter annulus r in [r_in, r_out], theta in [0, pi/2]:
 * nr radial cells growing geometrically by `ratio`, ntheta uniform angular
 * cells, straight-sided Q_p Lagrange quadrilaterals, p in 1..3.
 * Stands in for the gmsh + NeoPZ mesh of main.cpp:447-464, 542-607. */
int ipf_cmesh_quarter_annulus(int nr, int ntheta, int porder, double r_in, double r_out, double ratio,
                              ipf_cmesh **out);
/* A Gmsh 4.1 ASCII mesh (TPZGmshReader::GeometricGmshMesh, main.cpp:447-464)
 * with Lagrange elements of order porder in 1..3 on its triangles and
 * quadrilaterals (CmeshElastoplasticity, main.cpp:542-607).  Physical groups:
 * the lines "bc_wellbore" / "bc_farfield" carry the inner / outer normal
 * tractions (ipf_cmesh_set_normal_stress), the points "fixed_x" / "fixed_y"
 * fix u_x / u_y (main.cpp:394-411).  Domain elements keep the file order;
 * DoF numbering: the file's nodes, then edge nodes, then element interiors.
 * IPF_EINVAL for an unreadable file, a format other than 4.1 ASCII or an
 * element type other than point, 2-node line, 3-node triangle, 4-node
 * quadrilateral. */
int ipf_cmesh_read_gmsh(const char *path, int porder, ipf_cmesh **out);
/* What the Gmsh file held: out = nodes, elements, points, lines, triangles,
 * quadrilaterals (the counts TPZGmshReader reports, Results/.../out-1.txt). */
int ipf_cmesh_gmsh_counts(const ipf_cmesh *m, int64_t out[6]);
int ipf_cmesh_destroy(ipf_cmesh *m);
/* Material (TElastoPlasticData + TPZElasticResponse): E, nu, cohesion,
 * friction and dilation angles in radians. */
int ipf_cmesh_set_material(ipf_cmesh *m, double E, double nu, double cohesion, double phi, double psi);
/* Normal tractions sigma_n on the inner / outer boundary (arcs of the
 * quarter annulus; bc_wellbore / bc_farfield of a Gmsh mesh),
 * t = sigma_n * n_out (BC type 6 of main.cpp:397-405 in perturbation form). */
int ipf_cmesh_set_normal_stress(ipf_cmesh *m, double inner, double outer);
/* out[0..7] = nelem, nnodes, neq, porder, nr, ntheta (0 for a Gmsh mesh),
 * npts_per_element, nshape (-1 when triangles and quadrilaterals mix) */
int ipf_cmesh_info(const ipf_cmesh *m, int64_t out[8]);
/* out[0..4] = lambda, mu, phi, psi, cohesion */
int ipf_cmesh_material(const ipf_cmesh *m, double out[5]);
int ipf_cmesh_nodes(const ipf_cmesh *m, double *xy /* [2*nnodes] */);
/* connect (node) lists per element: conn_ptr [nelem+1], conn [conn_ptr[nelem]] */
int ipf_cmesh_connectivity(const ipf_cmesh *m, int64_t *conn_ptr, int64_t *conn);
/* Physical coordinates of the integration points [2*npts] (x, y), in the
 * reference's point order (element by element, eta fastest): the locations
 * of the per-point stresses and strains (post-processing, e.g. the plot over
 * y = 0 of RK_Verification). */
int ipf_cmesh_point_coords(const ipf_cmesh *m, double *xy);
/* Geometry input of ipf_transfer_geometry for this mesh: for element type
 * `type` (0 quadrilateral, 1 triangle, -1 the mesh's only type) sizes =
 * (n_ip, n_sh) and, when tab is not NULL, the shape table (n_ip*(2 n_sh + 3)
 * doubles, layout of ipf_transfer_geometry) -- the data NeoPZ's integration
 * rule and shape functions give the reference's setup
 * (TPZNumericalIntegrator.cpp:344-381).  Triangles are tabulated in the
 * reference square of the parallelogram their corners describe. */
int ipf_cmesh_shape_table(const ipf_cmesh *m, int type, int sizes[2], double *tab);
/* Element type per element: 0 quadrilateral, 1 triangle. */
int ipf_cmesh_element_types(const ipf_cmesh *m, signed char *types /* [nelem] */);
/* The 4 corners of each element's bilinear map, [8 * nelements], element
 * order (gel->Jacobian's input at TPZNumericalIntegrator.cpp:424). */
int ipf_cmesh_corners(const ipf_cmesh *m, double *corners);
/* 1 for constrained DoFs: the quarter annulus's symmetry conditions (u_y on
 * theta=0, u_x on theta=pi/2), a Gmsh mesh's fixed_x / fixed_y points */
int ipf_cmesh_dirichlet_mask(const ipf_cmesh *m, signed char *mask /* [neq] */);
/* the material's UpdateMem switch (main.cpp:487-522) */
int ipf_cmesh_set_update_mem(ipf_cmesh *m, int update_mem);

/* TPZElastoPlasticIntPointsStructMatrix(cmesh) + SetUpDataStructure
 * (TPZElastoPlasticIntPointsStructMatrix.cpp:15-25, :80-97).  With
 * device >= 0 the data is also transferred to that GPU; with device < 0 only
 * the host data is built (no GPU needed). */
int ipf_sm_create(ipf_cmesh *m, int device, ipf_sm **out);
int ipf_sm_destroy(ipf_sm *sm);
/* Device context of the structural matrix (NULL before the transfer). */
ipf_ctx *ipf_sm_context(ipf_sm *sm);
/* out[0..7] = nblocks, npts, rows, cols, neq, ncolor, storage values, 0 */
int ipf_sm_sizes(const ipf_sm *sm, int64_t out[8]);
/* IrregularBlocks arrays (TPZIrregularBlocksMatrix.h:18-30), host copies */
int ipf_sm_export_blocks(const ipf_sm *sm, int *row_sizes, int *col_sizes, int64_t *matrix_position,
                         int *row_first, int *col_first, double *storage);
/* fDoFIndexes, fColorIndexes, fElColorIndex [nblocks], fFirstColorIndex
 * [ncolor+1], weights [npts], fRhsLinear [neq]; any may be NULL */
int ipf_sm_export_indexes(const ipf_sm *sm, int *dof_indexes, int *color_indexes, int *el_color_index,
                          int64_t *first_color_index, double *weight, double *rhs_linear);
/* Assemble(rhs) for u (TPZElastoPlasticIntPointsStructMatrix.cpp:137-154),
 * host arrays [neq]; commits the state when the mesh's UpdateMem is set. */
int ipf_sm_assemble_rhs(ipf_sm *sm, const double *u_host, double *rhs_host);
/* Create() (TPZElastoPlasticIntPointsStructMatrix.cpp:34-60): the symmetric
 * upper-triangle CSR pattern of K and its per-color positions
 * (TPZNumericalIntegrator::FillLIndexes, .cpp:538-590); *nnz = CSR values.
 * Host work only: needs no GPU. */
int ipf_sm_create_tangent(ipf_sm *sm, int64_t *nnz);
/* IA [neq+1], JA [nnz], fColorLSequence [sum_e C_e (C_e+1)/2],
 * fFirstColorLIndex [ncolor+1] of that pattern; any may be NULL. */
int ipf_sm_export_tangent_pattern(const ipf_sm *sm, int64_t *ia, int64_t *ja, int64_t *color_l_sequence,
                                  int64_t *first_color_l_index);
/* Clone() (TPZElastoPlasticIntPointsStructMatrix.h:33, .cpp:30-32): a second
 * structural matrix on the same mesh that shares the first one's device
 * context (buffers, committed state, stream) by reference count, so either
 * may be destroyed first -- the reference's copy duplicates its TPZVecGPU
 * handles and frees them twice (TPZVecGPU.h:24-27).  The mesh must outlive
 * both. */
int ipf_sm_clone(const ipf_sm *sm, ipf_sm **out);
/* CreateAssemble(rhs, gui) (TPZElastoPlasticIntPointsStructMatrix.h:37,
 * .cpp:71-77) at u (host array [neq], becomes the mesh's solution): a new
 * matrix from Create() (CSR pattern, FillLIndexes, upload), rhs [neq] and the
 * matrix's values from Assemble(K, rhs).  *nnz = values of the new matrix,
 * read with ipf_sm_tangent_values; the matrix replaces any earlier one of
 * this structural matrix.  Commits the state when the mesh's UpdateMem is
 * set.  The reference's gui argument (a progress hook) has no use here. */
int ipf_sm_create_assemble(ipf_sm *sm, const double *u_host, double *rhs_host, int64_t *nnz);
/* Values [nnz] of the matrix the last ipf_sm_create_assemble /
 * ipf_sm_assemble_tangent filled (IPF_ESTATE before any). */
int ipf_sm_tangent_values(const ipf_sm *sm, double *kg_host);
/* TPZIrregularBlocksMatrix::CSRVectors (TPZIrregularBlocksMatrix.cpp:177-195):
 * the block matrix as CSR over its stored values -- row_ptr [rows+1] (row r
 * of block e starts at fMatrixPosition[e] + r * C_e), col_ind [stored values]
 * (fColFirstIndex[e] + c); the values are fStorage (ipf_sm_export_blocks).
 * int indexes, as the reference's.  Host work. */
int ipf_sm_export_csr(ipf_sm *sm, int *row_ptr, int *col_ind);
/* Assemble(K, rhs) for u (TPZElastoPlasticIntPointsStructMatrix.cpp:98-135):
 * the tangent's CSR values kg [nnz] and rhs [neq], host arrays; commits the
 * state when the mesh's UpdateMem is set. */
int ipf_sm_assemble_tangent(ipf_sm *sm, const double *u_host, double *rhs_host, double *kg_host);

#ifdef __cplusplus
}
#endif
#endif
