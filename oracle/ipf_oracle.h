/*
 * ipf_oracle.h -- CPU oracle for the integration-point residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * CPU path of IntPointsFEM (labmec/IntPointsFEM) for 2D plane-strain
 * Mohr-Coulomb elastoplasticity.  It is the checker the parity tests and the
 * bench's cpu_baseline leg compare the HIP product against.  Nothing in the
 * product (intpointsfem_amd/) may link, load or call it.
 *
 * Every function names the reference file:line it restates.  Where the
 * reference is undefined (uninitialised locals) the oracle fixes a value;
 * those places are marked DEVIATION.
 *
 * Parity pins (which test checks which function) are listed in
 * tests/test_oracle_*.py; functions without an independent pin say
 * "parity unpinned" below.
 */
#ifndef IPF_ORACLE_H
#define IPF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Voigt order of the reference's full symmetric tensors
 * (FunctionsTangentOperator.h:40-41; TPZTensor enum _XX_.._ZZ_). */
enum { ORC_XX = 0, ORC_XY = 1, ORC_XZ = 2, ORC_YY = 3, ORC_YZ = 4, ORC_ZZ = 5 };

/* Branch ids reported next to the reference's mtype flag. */
enum { ORC_BR_ELASTIC = 0, ORC_BR_MAIN = 1, ORC_BR_RIGHT = 2, ORC_BR_LEFT = 3, ORC_BR_APEX = 4 };

/* ---- per-point constitutive math (Functions*.h) ---- */
void orc_spectral_decomposition(double *sigma6, double *eigenvalues3, double *eigenvectors9);
/* Eigenvectors step alone (FunctionsSpectralDecomp.h:217-256) for given
 * eigenvalues; sigma6 is scaled by maxel in place as in the reference. */
void orc_eigenvectors(double *sigma6, const double *eigenvalues3, double *eigenvectors9, double maxel);
void orc_project_sigma(const double *eigenvalues3, double *sigma_projected3, int *m_type, double *alpha,
                       double mc_phi, double mc_psi, double mc_cohesion, double K, double G,
                       double *gradient9, int require_gradient, int *branch);
void orc_tangent_operator(const double *gradient9, const double *stress_eigenvalues3,
                          const double *strain_eigenvalues3, const double *strain_eigenvectors9,
                          double *tangent36, double G, double lambda);

/* One integration point of TPZConstitutiveLawProcessor::ComputeSigma
 * (TPZConstitutiveLawProcessor.cpp:150-211).  strain3 = [exx, gxy, eyy];
 * plastic_strain6 is read and (when update) overwritten; alpha read/updated. */
void orc_point_sigma(const double *strain3, double *plastic_strain6, double *alpha, int *mtype, int *branch,
                     double weight, double lambda, double mu, double mc_phi, double mc_psi,
                     double mc_cohesion, double *wsigma3, double *sigma6_out, double *dep9_out);

/* ---- TPZConstitutiveLawProcessor (arrays over points, AoS like the reference) ---- */
void orc_compute_sigma(int64_t npts, const double *strain3, double *wsigma3, const double *weight,
                       double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion,
                       double *plastic_strain6, double *alpha, int *mtype, int *branch, int update_mem,
                       int nthreads);
void orc_compute_sigma_dep(int64_t npts, const double *strain3, double *wsigma3, double *wdep9,
                           const double *weight, double lambda, double mu, double mc_phi, double mc_psi,
                           double mc_cohesion, double *plastic_strain6, double *alpha, int *mtype,
                           int *branch, int update_mem, int nthreads);

/* ---- TPZIrregularBlocksMatrix / TPZNumericalIntegrator ---- */
typedef struct {
    int64_t nblocks;
    const int *row_sizes;        /* nblocks   */
    const int *col_sizes;        /* nblocks   */
    const int64_t *mat_pos;      /* nblocks+1 */
    const int *row_first;        /* nblocks+1 */
    const int *col_first;        /* nblocks+1 */
    const double *storage;       /* mat_pos[nblocks] */
} orc_blocks;

void orc_gather(int64_t n, const double *x, const int *idx, double *y);
void orc_multiply_vector(const orc_blocks *b, const double *x, double *y, int opt, int nthreads);
void orc_multiply(const orc_blocks *b, const int *dof_indexes, const double *u, double *strain3,
                  int nthreads);
void orc_multiply_transpose(const orc_blocks *b, const int *color_indexes, int64_t ncolor, int64_t neq,
                            const double *wsigma3, double *res, double *slab_scratch, int nthreads);
int64_t orc_color_elements(int64_t nblocks, const int64_t *conn_ptr, const int64_t *conn,
                           int64_t nconnects, int64_t *elemcolor);
void orc_color_indexes(int64_t nblocks, const int *col_sizes, const int *col_first, const int *dof_indexes,
                       const int64_t *elemcolor, int64_t neq, int *color_indexes);
void orc_residual(const orc_blocks *b, const int *dof_indexes, const int *color_indexes, int64_t ncolor,
                  int64_t neq, const double *weight, double lambda, double mu, double mc_phi,
                  double mc_psi, double mc_cohesion, double *plastic_strain6, double *alpha, int *mtype,
                  int *branch, int update_mem, const double *u, const double *f_lin, double *rhs,
                  double *strain_scratch, double *sigma_scratch, double *slab_scratch, int nthreads);

/* Per-element tangent K_e = sum_ip B_ip^T Dep_ip B_ip, packed upper triangle
 * in color order (TPZNumericalIntegrator.cpp:139-163, :236-270). */
void orc_element_tangents(const orc_blocks *b, const double *wdep9, const int *el_color_index,
                          double *kc_packed, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
