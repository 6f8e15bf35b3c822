/*
 * intpointsfem.h -- C ABI of the MI355X (gfx950) integration-point residual
 * path for 2D plane-strain Mohr-Coulomb elastoplasticity.
 *
 * Two layers:
 *   ipf_ctx_*  / ipf_* device entry points: the data and operations the
 *     reference's GPU path keeps behind TPZCudaCalls / TPZVecGPU and the
 *     TransferDataToGPU methods.  Plain pointers and sizes only.
 *   ipf_cmesh_* / ipf_sm_* front-end entry points: the mesh stand-in and the
 *     TPZElastoPlasticIntPointsStructMatrix drop-in built on the layer above.
 *
 * Conventions
 *   - Every call returns int: IPF_OK (0) on success, a negative IPF_E* code
 *     otherwise.  ipf_last_error(ctx) / ipf_global_last_error() give the text.
 *     No C++ exception crosses this boundary.
 *   - Arrays named *_dev are device pointers (HBM), *_host are host pointers.
 *     Per-point arrays use the reference's AoS layout and point order:
 *     strain/stress [3*npts] as [xx, xy, yy] (engineering shear strain),
 *     tangent [9*npts] row-major 3x3, state [6*npts] in Voigt order
 *     [XX, XY, XZ, YY, YZ, ZZ].
 *   - Device entry points enqueue on the context's stream and return without
 *     waiting; ipf_synchronize() waits and reports asynchronous errors.  The
 *     *_host variants copy, run and wait.
 *   - A context belongs to one host thread at a time.
 */
#ifndef INTPOINTSFEM_H
#define INTPOINTSFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IPF_OK = 0,
    IPF_EINVAL = -1,      /* bad argument or inconsistent input data */
    IPF_EHIP = -2,        /* HIP runtime error */
    IPF_ESTATE = -3,      /* call out of order (e.g. residual before setup) */
    IPF_EUNSUPPORTED = -4 /* element shape without a compiled kernel */
};

/* Branch ids written next to the reference's mtype flag (0 elastic,
 * 1 main plane or edge, -1 apex; FunctionsSigmaProjection.h:341,353,371). */
enum { IPF_BR_ELASTIC = 0, IPF_BR_MAIN = 1, IPF_BR_RIGHT = 2, IPF_BR_LEFT = 3, IPF_BR_APEX = 4 };

typedef struct ipf_ctx ipf_ctx;

/* ------------------------------------------------------------------ */
/* Device context                                                      */
/* ------------------------------------------------------------------ */

/* Creates a context on HIP device `device` with its own stream.
 * Replaces the per-object TPZCudaCalls handles (TPZCudaCalls.h:15-50). */
int ipf_ctx_create(int device, ipf_ctx **out);
int ipf_ctx_destroy(ipf_ctx *ctx);
const char *ipf_last_error(const ipf_ctx *ctx);
const char *ipf_global_last_error(void);

/* Use a caller-owned stream (hipStream_t passed as void*); NULL restores the
 * context's own stream. */
int ipf_set_stream(ipf_ctx *ctx, void *hip_stream);
void *ipf_get_stream(const ipf_ctx *ctx);
int ipf_synchronize(ipf_ctx *ctx);

/* Uploads the integrator data and lays it out for the kernels.  Replaces
 * TPZNumericalIntegrator::TransferDataToGPU (TPZNumericalIntegrator.cpp:218-233),
 * TPZIrregularBlocksMatrix::TransferDataToGPU (TPZIrregularBlocksMatrix.cpp:144-174)
 * and TPZConstitutiveLawProcessor::TransferDataToGPU (TPZConstitutiveLawProcessor.cpp:739-742).
 *   blocks: IrregularBlocks fields (TPZIrregularBlocksMatrix.h:18-30) --
 *     nblocks, fRowSizes, fColSizes, fMatrixPosition (int64 here: the
 *     reference's int overflows beyond 2^31 stored values), fRowFirstIndex,
 *     fColFirstIndex, fStorage (row-major R_e x C_e blocks);
 *   dof_indexes [cols]   fDoFIndexes   (TPZNumericalIntegrator.h:88)
 *   color_indexes [cols] fColorIndexes = dof + color*neq (TPZNumericalIntegrator.cpp:503)
 *   ncolor, neq          fNColor, cmesh->NEquations()
 *   weight [npts]        fWeight = w_gauss*|detJ| (TPZNumericalIntegrator.cpp:425)
 * The state (plastic strain, alpha, mtype) is reset to zero
 * (TPZConstitutiveLawProcessor::SetUpDataByIntPoints, .cpp:86-121).
 * Fails with IPF_EINVAL when sizes are inconsistent, an index is out of range,
 * an element's color indexes disagree, or two elements of one color share a
 * DoF; with IPF_EUNSUPPORTED for a block shape without a kernel. */
int ipf_transfer_data(ipf_ctx *ctx, int64_t nblocks, const int *row_sizes, const int *col_sizes,
                      const int64_t *matrix_position, const int *row_first, const int *col_first,
                      const double *storage, const int *dof_indexes, const int *color_indexes, int64_t ncolor,
                      int64_t neq, const double *weight);

/* Setup on the device: the data of ipf_transfer_data with B and the weights
 * computed on the GPU from each element's geometry instead of uploaded.
 * Replaces the host loops of TPZNumericalIntegrator::SetUpIrregularBlocksData
 * (TPZNumericalIntegrator.cpp:272-388: B = J^-1 dphi per point, written as
 * B(3ip + c + j, 2a + c) = dN_a/dx_j) and the weights of SetUpIndexes
 * (.cpp:390-426: w |det J|), followed by TransferDataToGPU.
 *   nshape, shape_sizes [2*nshape]  (n_ip, n_sh) of each element shape;
 *   shape_tables   per shape, concatenated, n_ip*(2*n_sh + 3) doubles:
 *                  dphi[n_ip][n_sh][2] = (dN_a/dxi, dN_a/deta) at each point,
 *                  then xi[n_ip], eta[n_ip], w_q[n_ip] (the quadrature);
 *   elem_shape [nblocks]   shape index of each element;
 *   corners [8*nblocks]    the 4 corners of each element's bilinear map (x0,
 *                  y0, ..., x3, y3; counter-clockwise from (xi,eta) = (-1,-1));
 *   dof_indexes, color_indexes, ncolor, neq as in ipf_transfer_data, the
 *                  element's columns being (x, y) of each shape function.
 * Per point: J from the corners, dN_a/dx = a dN_a/dxi + b dN_a/deta,
 * dN_a/dy = c dN_a/dxi + d dN_a/deta with (a, b, c, d) the entries of J^-1
 * (reciprocal of det J, then products), weight w_q |det J|.  Blocks are
 * numbered and laid out as the reference's (row sizes 3 n_ip, column sizes
 * 2 n_sh, points in element order).  Selects IPF_B_GEOMETRY_FP64 for the
 * residual and the tangent product; the stored forms stay available.
 * An element whose det J vanishes at a point gets non-finite values there.
 * Fails like ipf_transfer_data, and with IPF_EINVAL for an unknown shape
 * index, two shapes of equal sizes or a non-finite corner coordinate. */
int ipf_transfer_geometry(ipf_ctx *ctx, int nshape, const int *shape_sizes, const double *shape_tables,
                          int64_t nblocks, const int *elem_shape, const double *corners, const int *dof_indexes,
                          const int *color_indexes, int64_t ncolor, int64_t neq);

/* Material of TPZConstitutiveLawProcessor::SetMaterial (.cpp:127-129); the
 * reference reads lambda, mu, phi, psi, c from it per call (.cpp:152-158).
 * Angles in radians. */
int ipf_set_material(ipf_ctx *ctx, double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion);

/* Several materials (the reference has one, FindMaterial(1),
 * TPZNumericalIntegrator.cpp:446): params [5*nmat] = (lambda, mu, phi, psi,
 * c) per material; ipf_set_material is the one-entry table.  The element ->
 * material map is given before the data: ipf_set_element_materials(ctx,
 * nblocks, ids [nblocks]) applies to the next ipf_transfer_data, which keeps
 * every wavefront tile within one material (each color/shape launch then
 * covers all materials).  Without it every element has material 0. */
int ipf_set_material_table(ipf_ctx *ctx, int nmat, const double *params);
int ipf_set_element_materials(ipf_ctx *ctx, int64_t nblocks, const int *element_material);
/* Order of the elements inside each (color, shape, material) run of the next
 * ipf_transfer_data / ipf_transfer_geometry: ascending keys [nblocks], then
 * element index (default: all keys 0, index order); with the patch assembly,
 * the order inside each patch (patches themselves follow the geometry).  The
 * order decides which elements share a wavefront tile -- e.g.
 * predicted-elastic elements together, so that whole tiles take the
 * certified elastic path -- and nothing else: within a color every DoF has
 * one element, so colored residuals, tangent products and states are
 * bit-identical for any keys (patch residuals: up to the order of each
 * node's sum).  No reference counterpart. */
int ipf_set_element_order_keys(ipf_ctx *ctx, int64_t nblocks, const int *keys);

/* fRhsLinear / dRhsLinear (TPZElastoPlasticIntPointsStructMatrix.cpp:63-68,156-167), [neq] host array. */
int ipf_set_rhs_linear(ipf_ctx *ctx, const double *f_lin_host);

/* The UpdateMem switch of the material (TPZConstitutiveLawProcessor.cpp:197):
 * when set, residual / compute_sigma calls commit plastic strain, alpha, mtype
 * and branch per point. */
int ipf_set_update_mem(ipf_ctx *ctx, int update_mem);

/* Sizes: out[0..7] = nblocks, npts, rows, cols, neq, ncolor, storage values,
 * device bytes held. */
int ipf_info(const ipf_ctx *ctx, int64_t out[8]);

/* ---- the hot path ---- */

/* r = f_lin - sum_e B_e^T (w * sigma(B_e u_e)), device pointers [neq].
 * TPZElastoPlasticIntPointsStructMatrix::Assemble(rhs)
 * (TPZElastoPlasticIntPointsStructMatrix.cpp:137-154) ->
 * TPZNumericalIntegrator::ResidualIntegration (TPZNumericalIntegrator.cpp:107-128).
 * One fused kernel per element color: gather, B.u, return map, B^T (w sigma)
 * and the color's atomic-free accumulation into r.  r is overwritten
 * entirely (its previous contents are never read); u and r must not
 * overlap.  Asynchronous on the context's stream. */
int ipf_residual(ipf_ctx *ctx, const double *u_dev, double *r_dev);
int ipf_residual_host(ipf_ctx *ctx, const double *u_host, double *r_host);

/* How non-committing fused residuals (ipf_residual, ipf_residual_timed with
 * UpdateMem off) evaluate the stress.  No reference counterpart.
 *   IPF_RESIDUAL_EXACT (default): the reference's return map at every point;
 *     branch flags, stresses and committed state bit-identical to the oracle.
 *   IPF_RESIDUAL_CERTIFIED_ELASTIC: a point whose trial stress is elastic
 *     with a margin -- its closed-form principal stresses below the main
 *     Mohr-Coulomb plane by 2e-3 of its largest stress entry, more than the
 *     reference's eigenvalue solver (20 Newton steps, |f| <= 1e-11 on the
 *     normalised cubic, FunctionsSpectralDecomp.h:61-111) can be off -- is
 *     elastic in the reference too, and takes its trial stress instead of
 *     the eigen-reconstruction of it (FunctionsStressStrain.h:60-68): the
 *     same stress up to that reconstruction's rounding and the solver's
 *     tolerance.  Every other point takes the exact return map; a wavefront
 *     of certified points skips the spectral decomposition altogether.
 * Committing residuals (UpdateMem on) always run the exact path, so the
 * committed state never depends on the mode; neither do ipf_residual_tangent,
 * the separate steps or the Newton solver's residuals. */
enum { IPF_RESIDUAL_EXACT = 0, IPF_RESIDUAL_CERTIFIED_ELASTIC = 1 };
int ipf_set_residual_mode(ipf_ctx *ctx, int mode);
int ipf_get_residual_mode(const ipf_ctx *ctx, int *mode);
/* How non-committing fused residuals accumulate r.  No reference
 * counterpart (the reference scatters into color slabs,
 * TPZNumericalIntegrator.cpp:81-105).
 *   IPF_ASSEMBLY_COLORED (default): one launch per element color, each
 *     adding its elements' forces to r in color order -- the reference's
 *     per-DoF summation order.
 *   IPF_ASSEMBLY_PATCH: compact patches of elements (a Morton order of the
 *     element centroids, cut per element shape and material), one
 *     workgroup per patch: each patch's nodes are gathered once, its
 *     elements' forces summed per node in LDS, nodes owned by the patch
 *     written once (r = f_lin - sum, r never read), nodes shared with other
 *     patches finished by a fix-up pass over per-patch partial sums in a
 *     fixed order.  Deterministic, atomic-free; the same residual up to the
 *     order of the sums (a few ulps of the element forces).  Needs the
 *     geometric setup (ipf_transfer_geometry) and every element's DoFs in
 *     (x, y) node pairs (IPF_EUNSUPPORTED otherwise); built on the first
 *     selection after each transfer.
 * Committing residuals always take the colored path. */
enum { IPF_ASSEMBLY_COLORED = 0, IPF_ASSEMBLY_PATCH = 1 };
int ipf_set_residual_assembly(ipf_ctx *ctx, int assembly);
int ipf_get_residual_assembly(const ipf_ctx *ctx, int *assembly);
/* out[0..7] = built (0/1), patches, patch nodes, partial slots, shared
 * nodes, DoFs of no element, patch points (padded), element shapes. */
int ipf_patch_info(const ipf_ctx *ctx, int64_t out[8]);

/* Wavefront tiles per patch of IPF_ASSEMBLY_PATCH: 2, 4 (default) or 8
 * (one workgroup of 64 * waves lanes per patch); a change drops the built
 * patch layout (rebuilt at once when the patch assembly is selected).  No
 * reference counterpart. */
int ipf_set_patch_waves(ipf_ctx *ctx, int waves);
int ipf_get_patch_waves(const ipf_ctx *ctx, int *waves);

/* Autotuner of the non-committing residual's tile shape: times reps
 * residuals at u_dev (current stress mode and B format, committed state
 * untouched) with the colored assembly (one wavefront tile per workgroup)
 * and, where the setup allows patches, with patches of 2, 4 and 8 wavefront
 * tiles, and keeps the fastest (ipf_get_residual_assembly /
 * ipf_get_patch_waves tell which).  times_ms (optional) [4]: colored, patch
 * 2, 4, 8; -1 where not timed.  The assemblies differ in the order of the
 * per-DoF sums only (a few ulps).  No reference counterpart. */
int ipf_autotune_assembly(ipf_ctx *ctx, const double *u_dev, int reps, double *times_ms);

/* The patch layout of IPF_ASSEMBLY_PATCH planned on the host (no GPU), for
 * inspection and tests: n elements with their shape (kernel shape index:
 * 0..2 Q1..Q3, 3..5 T1..T3), material, first integration point, centroid
 * and node x-DoFs (shape_ce / 2 per element, concatenated; y DoF = x + 1),
 * neq equations, waves wavefront tiles per patch (2, 4 or 8).  sizes[12] =
 * ranges, patches, tiles, slots, local node
 * entries, patch nodes, contributions, patch points, shared nodes, partial
 * slot entries of the shared nodes, orphan DoFs, partial slots.  Each output
 * array may be NULL (call once with NULL outputs for the sizes): ranges
 * [7 per range] = (shape, patches, first patch, first tile, first slot,
 * first point, first local node entry); tmeta [tiles]; slot_elem [slots]
 * (element index, -1 padding); lnode [local node entries]; node_off
 * [patches + 1]; pnode, pslot [patch nodes]; cbeg [patch nodes + 1];
 * contrib [contributions]; pt_map [patch points] (the element's first point
 * + point, -1 padding); fix_dof [shared] (ascending), fix_ptr [shared + 1],
 * fix_slot (the identity: shared node i owns the partial slots
 * [fix_ptr[i], fix_ptr[i + 1]), one per patch, in patch order); orphan
 * [orphan DoFs]. */
int ipf_patch_plan_host(int64_t n, const int *shape, const int *mat, const int64_t *first_point, const double *cx,
                        const double *cy, const int *node_xdof, int64_t neq, int waves, int64_t *sizes,
                        int64_t *ranges,
                        int *tmeta, int64_t *slot_elem, unsigned short *lnode, int64_t *node_off, int *pnode,
                        int *pslot, int *cbeg, unsigned short *contrib, int64_t *pt_map, int *fix_dof, int *fix_ptr,
                        int *fix_slot, int *orphan);

/* The certificate of IPF_RESIDUAL_CERTIFIED_ELASTIC on the host, point by
 * point (no GPU): strain3 [3n] ([xx, xy, yy], engineering shear),
 * plastic_strain4 [4n] ([XX, XY, YY, ZZ]); certified [n] = 1 where the
 * point passes, and then sigma3 [3n] (optional) = its trial stress
 * [sxx, sxy, syy], else 0. */
int ipf_certify_elastic_host(int64_t n, const double *strain3, const double *plastic_strain4, double lambda,
                             double mu, double mc_phi, double mc_psi, double mc_cohesion, int *certified,
                             double *sigma3);

/* ipf_residual with a HIP event pair around every kernel launch on the
 * context's stream; waits, then writes the duration of each launch in
 * milliseconds to launch_ms (launch 0: r = f_lin copy, exactly 0 when no
 * copy runs, then one per element group in launch order) and the group's
 * point count to launch_pts (may be NULL).  With the patch assembly
 * (non-committing): launch 0 the refresh of the patch copy of the plastic
 * strain (exactly 0 when it is current), one launch per
 * element shape, then the shared nodes' fix-up (0 points).  Returns the number of launches (>= 0) or a negative IPF_E* code.
 * Measurement aid; the plain entry point records no events. */
int ipf_residual_timed(ipf_ctx *ctx, const double *u_dev, double *r_dev, double *launch_ms, int64_t *launch_pts,
                       int max_launches);

/* ---- the reference's separate steps (same results, unfused) ---- */

/* y = B x (opt == 0; x [cols], y [rows]) or y = -B^T x (opt != 0; x [rows],
 * y [cols]) without gather/scatter, device pointers.
 * TPZIrregularBlocksMatrix::MultiplyVector (TPZIrregularBlocksMatrix.cpp:97-140). */
int ipf_multiply_vector(ipf_ctx *ctx, const double *x_dev, double *y_dev, int opt);

/* strain = B . u[dof]   TPZNumericalIntegrator::Multiply (TPZNumericalIntegrator.cpp:25-51) */
int ipf_multiply(ipf_ctx *ctx, const double *u_dev, double *strain_dev);
/* w*sigma per point     TPZConstitutiveLawProcessor::ComputeSigma (TPZConstitutiveLawProcessor.cpp:131-342, 588-659) */
int ipf_compute_sigma(ipf_ctx *ctx, const double *strain_dev, double *wsigma_dev);
/* w*sigma and w*Dep     TPZConstitutiveLawProcessor::ComputeSigmaDep (TPZConstitutiveLawProcessor.cpp:344-585, 663-735) */
int ipf_compute_sigma_dep(ipf_ctx *ctx, const double *strain_dev, double *wsigma_dev, double *wdep_dev);
/* res = -B^T wsigma, colored accumulation   TPZNumericalIntegrator::MultiplyTranspose (TPZNumericalIntegrator.cpp:54-105) */
int ipf_multiply_transpose(ipf_ctx *ctx, const double *wsigma_dev, double *res_dev);

/* ---- storage format of B ---- */

enum {
    IPF_B_DENSE_FP64 = 0,   /* every stored value of the reference's blocks (default) */
    IPF_B_COMPACT_FP64 = 1, /* (dN/dx, dN/dy) per node and point; same results */
    IPF_B_DENSE_FP32 = 2,   /* FP32 products with FP32 accumulation */
    IPF_B_COMPACT_FP32 = 3,
    IPF_B_GEOMETRY_FP64 = 4, /* nothing stored per point: the compact FP64 values and
                               the weight recomputed in the kernel from the element's
                               corners (after ipf_transfer_geometry); compact FP64 results */
    IPF_B_GEOMETRY_FP32 = 5  /* the geometric values rounded to FP32 (= compact FP32),
                               FP32 products and accumulation; as the tangent format,
                               w Dep is read from an FP32 copy too (the CG matvec in FP32) */
};

/* Selects the B representation of the fused residual (ipf_residual,
 * ipf_residual_timed, and the commit inside ipf_newton) and, separately, of
 * the tangent product (ipf_tangent_apply and the CG matvecs).  The
 * separate-step entry points and ipf_residual_tangent keep the dense FP64
 * blocks.  The compact forms hold only dN_a/dx and dN_a/dy per point -- the
 * reference's blocks carry nothing else (TPZNumericalIntegrator.cpp:371-381);
 * ipf_transfer_data checks that pattern exactly and this call fails with
 * IPF_EINVAL when it does not hold.  Compact FP64 gives the dense FP64
 * results (the skipped terms are exact zeros); the FP32 forms compute B.u and
 * B^T(w sigma) in FP32 with FP32 accumulation, everything else in FP64.  An
 * FP64 residual with an FP32 tangent makes ipf_newton an iterative refinement:
 * the FP32 products only steer the correction, the FP64 residual decides
 * convergence.  The geometric form (after ipf_transfer_geometry) reads 8
 * corner coordinates per element instead of any per-point B value or weight
 * and gives the compact FP64 results bit for bit.  No reference counterpart:
 * the reference stores dense FP64 blocks only (TPZIrregularBlocksMatrix.h:20). */
int ipf_set_b_format(ipf_ctx *ctx, int residual_format, int tangent_format);
/* Autotuner: times the fused residual (reps launches at u_dev, committed
 * state untouched) in every exact FP64 form the data allows (dense, compact,
 * geometric) and, after ipf_residual_tangent, the tangent product likewise
 * (with allow_fp32, also the FP32 forms: the CG matvecs then steer an FP64
 * Newton, see above), and selects the fastest of each.  The FP64 forms give
 * bit-identical results, so the choice changes only the time.  times_ms
 * (optional) [2 * 6]: average ms per residual by format, then per tangent
 * product by format, -1 where not timed (format codes as above).  Frees
 * stored forms neither selection uses.  No reference counterpart. */
int ipf_autotune(ipf_ctx *ctx, const double *u_dev, int reps, int allow_fp32, double *times_ms);
/* The formats in use.  ipf_transfer_data selects IPF_B_COMPACT_FP64 for both
 * when every block has the 2D elasticity pattern, IPF_B_DENSE_FP64 otherwise;
 * ipf_transfer_geometry selects IPF_B_GEOMETRY_FP64. */
int ipf_get_b_format(const ipf_ctx *ctx, int *residual_format, int *tangent_format);

/* ---- learned return-branch predictor (no reference counterpart) ---- */

/* Training pairs of the exact return map at the trial state of u (the
 * committed state is read, not changed), reference point order:
 * feat_dev [8*npts] floats = (s1/c, s2/c, s3/c, ln c, phi, psi, 1e3*alpha, 0)
 * with s1 >= s2 >= s3 the trial principal stresses
 * (FunctionsSpectralDecomp.h:261-268), branch_dev [npts] the branch id of
 * ProjectSigma (FunctionsSigmaProjection.h:337-379; IPF_BR_*), target_dev
 * [npts] 1e3 * the alpha increment of that return. */
int ipf_point_features(ipf_ctx *ctx, const double *u_dev, float *feat_dev, int *branch_dev, float *target_dev);

/* A 7 -> 64 -> 64 -> 6 MLP (ReLU; BF16 weights and activations on the
 * matrix cores, FP32 accumulation): 5 branch logits + the alpha increment.
 * Parameters are NPAR FP32 values (ipf_mlp_num_params; padded entries stay
 * zero), Glorot-uniform from `seed`.  Samples are device arrays in the
 * layout of ipf_point_features.  Loss: mean cross-entropy of the branch +
 * 0.1 * mean squared error of the increment.  ipf_mlp_gradient returns the
 * gradient of that loss over n samples (summed in a fixed order);
 * ipf_mlp_train_step applies one Adam step (beta 0.9 / 0.999, eps 1e-8);
 * ipf_mlp_predict writes argmax branches [n] and / or the 6 outputs [6n].
 * Calls wait for their own stream; loss may be NULL (then train_step does
 * not wait). */
/* Branch sorting: order_dev [npts] = the reference point indexes grouped by
 * branch_dev (keys 0..4, stable inside a key); ipf_compute_sigma_ordered is
 * ipf_compute_sigma with the points visited in that order (each wavefront
 * then sees points of one predicted branch) -- the same results. */
int ipf_branch_order(ipf_ctx *ctx, const int *branch_dev, int *order_dev);
int ipf_compute_sigma_ordered(ipf_ctx *ctx, const double *strain_dev, double *wsigma_dev, const int *order_dev);

/* ipf_branch_order that allocates nothing: no hipMalloc / hipFree or
 * allocating library call on any call, first included.  No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The counting
 * sort's per-block histograms live in a fixed 4 MiB device array of the
 * library, one per GPU, held until the process ends and shared by every
 * call on that GPU: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller.  Same order_dev as
 * ipf_branch_order.  IPF_EINVAL when the context holds more points than
 * ipf_branch_order_noalloc_max_points() (53,687,040). */
int ipf_branch_order_noalloc(ipf_ctx *ctx, const int *branch_dev, int *order_dev);
int64_t ipf_branch_order_noalloc_max_points(void);

typedef struct ipf_mlp ipf_mlp;
int ipf_mlp_create(int device, uint64_t seed, ipf_mlp **out);
int ipf_mlp_destroy(ipf_mlp *mlp);
/* caller stream (hipStream_t as void*; NULL: the object's own stream, a
 * blocking stream that orders itself with the legacy default stream) */
int ipf_mlp_set_stream(ipf_mlp *mlp, void *hip_stream);
int ipf_mlp_num_params(int *padded, int *trainable);
int ipf_mlp_set_params(ipf_mlp *mlp, const float *params_host);
int ipf_mlp_get_params(ipf_mlp *mlp, float *params_host);
int ipf_mlp_gradient(ipf_mlp *mlp, const float *feat_dev, const int *branch_dev, const float *target_dev, int64_t n,
                     float *grad_host, double *loss);
int ipf_mlp_train_step(ipf_mlp *mlp, const float *feat_dev, const int *branch_dev, const float *target_dev,
                       int64_t n, float lr, double *loss);
int ipf_mlp_predict(ipf_mlp *mlp, const float *feat_dev, int64_t n, int *branch_dev, float *out_dev);

/* ---- tangent, linear solve and Newton (the caller side of the path) ---- */

/* Symmetry / Dirichlet DoFs (1 = fixed) used by the solver entry points,
 * host array [neq].  The reference imposes them with NeoPZ penalty BCs; here
 * they are eliminated (fixed DoFs keep their value, their residual is ignored). */
int ipf_set_dirichlet_mask(ipf_ctx *ctx, const signed char *mask_host);

/* Residual as ipf_residual plus the consistent tangent w*Dep of every point,
 * kept on the device for ipf_tangent_apply / ipf_cg_solve.
 * KAssembly's Multiply -> ComputeSigmaDep -> MultiplyTranspose
 * (TPZNumericalIntegrator.cpp:185-194); the element matrices of
 * MatrixAssembleKernel (KernelsMatrixAssemble.h:24-56) are applied matrix-free. */
int ipf_residual_tangent(ipf_ctx *ctx, const double *u_dev, double *r_dev);

/* CSR positions of the packed element matrices for ipf_assemble_tangent, as
 * TPZNumericalIntegrator::FillLIndexes builds them (TPZNumericalIntegrator.cpp:
 * 538-590): el_color_index [nblocks] = fElColorIndex (elements color by
 * color), color_l_sequence [sum_e C_e (C_e + 1) / 2] = fColorLSequence (for
 * each element in that order, the position in the values array of the
 * symmetric upper-triangle CSR (TPZSYsmpMatrix) of K_e(a, c), a <= c, row by
 * row), nnz = number of CSR values.  Host arrays, copied; checked for range
 * and color order. */
int ipf_set_tangent_pattern(ipf_ctx *ctx, const int *el_color_index, const int64_t *color_l_sequence, int64_t nnz);

/* The same positions, with the pattern itself, built on the device from the
 * uploaded DoF indexes -- the symmetric upper-triangle CSR of the element
 * graph that TPZElastoPlasticIntPointsStructMatrix::Create() allocates (every
 * (i, j), i <= j, of two DoFs of one element; columns ascending in each row)
 * and the positions FillLIndexes / me() look up (TPZNumericalIntegrator.cpp:
 * 538-606) -- by a sort of the element entries' (row, column) keys: the same
 * integers as the host build.  *nnz = the pattern's entries; replaces an
 * ipf_set_tangent_pattern.  ipf_get_tangent_pattern copies the pattern to
 * host arrays ia [neq + 1] and ja [nnz] (either may be NULL). */
int ipf_build_tangent_pattern(ipf_ctx *ctx, int64_t *nnz);
int ipf_get_tangent_pattern(ipf_ctx *ctx, int64_t *ia_host, int64_t *ja_host);

/* TPZNumericalIntegrator::KAssembly (TPZNumericalIntegrator.cpp:131-214, GPU
 * variant :184-214 with MatrixAssembleKernel, KernelsMatrixAssemble.h:24-56):
 * ipf_residual_tangent, then kg_dev[nnz] = 0 and, color by color, every
 * element's K_e = sum_ip B_ip^T (w Dep_ip) B_ip added at its CSR positions.
 * Device pointers; kg_dev holds nnz doubles. */
int ipf_assemble_tangent(ipf_ctx *ctx, const double *u_dev, double *r_dev, double *kg_dev);

/* y = K v with K = sum_e B_e^T (w Dep) B_e from the last ipf_residual_tangent
 * (all DoFs, no mask), device pointers [neq].  y is overwritten entirely
 * (its contents before the call are not read); v and y must not overlap. */
int ipf_tangent_apply(ipf_ctx *ctx, const double *v_dev, double *y_dev);

/* Block-Jacobi-preconditioned CG on the free DoFs (each node's 2x2 diagonal
 * block of K inverted where its DoFs are an aligned (2k, 2k + 1) pair, the
 * diagonal elsewhere): K x = b, x0 = 0, until
 * ||r|| <= rtol * ||b|| or maxit iterations.  Device pointers [neq]; fixed DoFs
 * of x are zero.  iters / relres may be NULL. */
int ipf_cg_solve(ipf_ctx *ctx, const double *b_dev, double *x_dev, double rtol, int maxit, int *iters,
                 double *relres);

/* Newton for the current load (f_lin): u <- u + du until the free-DoF residual
 * satisfies ||r|| <= rtol * ||r0|| (r0: residual at the entry u) or ||r|| <=
 * atol (ipf_set_newton_atol, default 0), then commits the state at the
 * converged u (AcceptPseudoTimeStepSolution, main.cpp:487-522).
 * Solution / Newton loop of main.cpp:160-262 with CG in place of LDL^T; the
 * reference stops on absolute norms, ||r|| < tol and ||du|| < tol
 * (main.cpp:239; ipf_set_newton_reference_stop) -- the relative test here, the absolute floor for a call
 * that starts at an already converged u (its r0 is at rounding level, so the
 * relative test alone could not be met).  The caller's UpdateMem switch is
 * the same after the call, whatever its status.
 * stats[0..7] = Newton iterations, converged (0/1), ||r||/||r0||, ||r0||,
 * ||r||, total CG iterations, worst CG relative residual, step halvings of
 * the line search (0 without one). */
int ipf_newton(ipf_ctx *ctx, double *u_dev, double rtol, int maxit, double cg_rtol, int cg_maxit, double *stats);

/* Backtracking line search for ipf_newton (0, the default: the reference's
 * full steps): after each CG solve, u0 + t du with t = 1, 1/2, ... is taken
 * at the first t whose free-DoF residual norm is below ||r(u0)||, or at the
 * last of max_halvings halvings (0..30).  Each trial costs one residual.
 * With an inexact inner solve (a loose cg_rtol) a full step can increase the
 * residual far from the solution; the search keeps such steps from
 * diverging.  No reference counterpart (main.cpp:160-262 takes full steps). */
int ipf_set_newton_line_search(ipf_ctx *ctx, int max_halvings);

/* Progress report of ipf_newton / ipf_newton_segmented, called on the host
 * after each residual norm (iteration 0 = the entry u), like the residual
 * norms the reference's Newton loop prints every iteration (main.cpp:240-241):
 * rel = ||r|| / ||r0|| (segmented: the largest ||r_k|| / ||r0_k|| over the
 * segments), res_norm = ||r|| (segmented: over all segments), cg_iterations
 * = CG iterations so far.  fn NULL turns it off (the default). */
typedef void (*ipf_newton_monitor)(void *user, int iteration, double rel, double res_norm, int cg_iterations);
int ipf_set_newton_monitor(ipf_ctx *ctx, ipf_newton_monitor fn, void *user);

/* Absolute residual floor of ipf_newton / ipf_newton_segmented (per
 * segment): also converged when ||r|| <= atol.  atol >= 0; 0 turns it off. */
int ipf_set_newton_atol(ipf_ctx *ctx, double atol);

/* The reference's stop test for ipf_newton (tolerance > 0; 0, the default,
 * restores the relative test above): converged at the first iterate, after
 * at least one update, with ||r|| < tolerance and ||du|| < tolerance, du the
 * update that led to it -- main.cpp:239 (`norm_res < tolerance &
 * norm_delta_du < tolerance`, tolerance 1e-7 at main.cpp:139), the
 * conventional (non-accelerated) update of main.cpp:226-229.  ||r|| is taken
 * over the free DoFs, ||du|| over all DoFs (du is zero on the others); with
 * the line search du is the step taken.  rtol and ipf_set_newton_atol are
 * then unused.  Not used by ipf_newton_segmented.  EINVAL unless tolerance
 * is finite and >= 0. */
int ipf_set_newton_reference_stop(ipf_ctx *ctx, double tolerance);

/* The reference's accelerated update for ipf_newton (enable != 0; off by
 * default, as modified_thomas_accel_Q at main.cpp:76): after the Newton step
 * du of an iteration, u += alpha du, the residual there with the tangent left
 * unchanged, its correction du~ solved with that same tangent,
 * ||du|| := ||alpha du + du~|| (for ipf_set_newton_reference_stop),
 * alpha += (alpha du . du~) / (alpha du . alpha du), u += du~ -- main.cpp:
 * 194-224, with alpha = 1 at the start of each call (main.cpp:171).  Two CG
 * solves per iteration (both counted in stats[5]).  EINVAL from ipf_newton
 * when the line search is also on.  Not used by ipf_newton_segmented. */
int ipf_set_newton_thomas(ipf_ctx *ctx, int enable);

/* ipf_newton for independent blocks of equations (batched instances): segment
 * k holds the equations [seg_first[k], seg_first[k+1]) (seg_first [nseg+1],
 * from 0 to neq) and converges on its own criterion ||r_k|| <= rtol ||r0_k||
 * (norms over the DoFs the Dirichlet mask leaves free); once it has, its
 * DoFs are held -- out of the CG and of the update -- while the others
 * iterate, and the mask is restored afterwards.  One CG serves all segments
 * (the system is block diagonal): each active segment's right-hand side is
 * scaled to unit norm and the CG tolerance is cg_rtol / sqrt(active
 * segments), so that every segment's own relative CG residual is <= cg_rtol
 * as in its own ipf_newton.  The commit follows as in ipf_newton.  stats as
 * ipf_newton's, with [1] = all segments converged, [2]
 * the largest ||r_k|| / ||r0_k||, [3] / [4] the 2-norms over all segments;
 * seg_rel [nseg] (optional) the final ratio per segment, seg_iters [nseg]
 * (optional) the Newton iteration at which each converged (-1: not
 * converged).  The per-instance convergence of a batch; no reference
 * counterpart (the reference solves one mesh). */
int ipf_newton_segmented(ipf_ctx *ctx, double *u_dev, int nseg, const int64_t *seg_first, double rtol, int maxit,
                         double cg_rtol, int cg_maxit, double *stats, double *seg_rel, int *seg_iters);

/* The CG iterations of ipf_cg_solve / ipf_newton run as a captured HIP graph
 * of 10 iterations (the launches between two convergence checks), replayed
 * on the context's stream: fewer launch gaps, the same kernels in the same
 * order, hence the same results as launching them one by one (enable = 0).
 * On by default; a capture the runtime refuses falls back to single
 * launches.  No reference counterpart (the reference solves with LDL^T,
 * main.cpp:283-285). */
int ipf_set_cg_graph(ipf_ctx *ctx, int enable);

/* ipf_newton runs each Newton iteration's device work up to its first host
 * read -- the tangent residual (ipf_residual_tangent's launches), the
 * residual norm over the free DoFs, the CG's preconditioner and start and
 * its first 10 iterations -- as one HIP graph replay, captured once per
 * ipf_newton call (enable != 0, the default; used when cg_maxit >= 10), or
 * launches the same kernels one by one (enable == 0).  One host read per
 * iteration instead of three; the iterates, iteration counts and committed
 * state are the same bits either way.  The graph records no memset or copy
 * node.  ipf_newton_segmented runs its iterations eagerly.  The reference's
 * Newton step (main.cpp:160-262) is host-driven throughout. */
int ipf_set_newton_graph(ipf_ctx *ctx, int enable);

/* ---- integration-point state (TPZConstitutiveLawProcessor.h:53-63) ---- */

/* Copies the committed state to host arrays (any may be NULL):
 * plastic_strain [6*npts] Voigt, alpha [npts], mtype [npts], branch [npts]. */
int ipf_get_state(ipf_ctx *ctx, double *plastic_strain_host, double *alpha_host, int *mtype_host,
                  int *branch_host);
/* Sets the committed state (restart).  The XZ and YZ plastic components must
 * be zero (plane strain); alpha may be NULL (zero). */
int ipf_set_state(ipf_ctx *ctx, const double *plastic_strain_host, const double *alpha_host);

#ifdef __cplusplus
}
#endif
#endif
