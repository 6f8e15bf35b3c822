// ipf_device.hip -- device context behind the C ABI of include/intpointsfem.h.
//
// Owns every device buffer of one integrator instance (the reference spreads
// them over TPZVecGPU members of TPZIrregularBlocksMatrix, TPZNumericalIntegrator
// and TPZConstitutiveLawProcessor) and builds the tiled HBM layout described
// in ipf_kernels.h from the reference's IrregularBlocks arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/intpointsfem.h"
#include "ipf_geometry.h"
#include "ipf_kernels.h"
#include "ipf_learned.h"
#include "ipf_patch.h"

namespace ipf {

MatConst MatConst::make(double lambda, double mu, double mc_phi, double mc_psi, double c) {
    MatConst m{};
    const double G = mu;
    const double K = lambda + 2 * mu / 3;  // TPZConstitutiveLawProcessor.cpp:158
    m.lambda = lambda;
    m.mu = mu;
    m.G = G;
    m.K = K;
    m.lam2mu = lambda + 2 * mu;
    m.cs_a = (lambda + mu) / (3 * lambda * mu + 2 * mu * mu);
    m.cs_b = lambda / (6 * lambda * mu + 4 * mu * mu);
    m.inv_mu = 1 / mu;
    const double sinphi = sin(mc_phi);
    const double cosphi = cos(mc_phi);
    const double sinpsi = sin(mc_psi);
    m.sinphi = sinphi;
    m.cosphi = cosphi;
    m.sinpsi = sinpsi;
    m.phi_c = 2. * c * cosphi;
    m.edge_c = 2. * cosphi * c;
    m.constA = 4. * G * (1. + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    m.m0 = (2. * G * (1 + sinpsi / 3.) + 2. * K * sinpsi);
    m.m1 = (4. * G / 3. - K * 2.) * sinpsi;
    m.m2 = (2. * G * (1 - sinpsi / 3.) - 2. * K * sinpsi);
    // edge Jacobians (constant for H = 0), FunctionsSigmaProjection.h:117-136, :217-238
    auto edge = [](double ab0, double ab1, double* ji) {
        const double j00 = -ab0, j10 = -ab1, j01 = -ab1, j11 = -ab0;
        const double det = j00 * j11 - j01 * j10;
        ji[0] = j11 / det;   // [0][0]
        ji[1] = -j01 / det;  // [0][1]
        ji[2] = -j10 / det;  // [1][0]
        ji[3] = j00 / det;   // [1][1]
    };
    double ji[4];
    m.r_ab0 = 4. * G * (1 + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    m.r_ab1 = 2. * G * (1. + sinphi + sinpsi - sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    edge(m.r_ab0, m.r_ab1, ji);
    m.r_ji00 = ji[0]; m.r_ji01 = ji[1]; m.r_ji10 = ji[2]; m.r_ji11 = ji[3];
    m.l_ab0 = 4. * G * (1 + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    m.l_ab1 = 2. * G * (1. - sinphi - sinpsi - sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    edge(m.l_ab0, m.l_ab1, ji);
    m.l_ji00 = ji[0]; m.l_ji01 = ji[1]; m.l_ji10 = ji[2]; m.l_ji11 = ji[3];
    m.omsp = 1 - sin(mc_psi);
    m.opsp = 1 + sin(mc_psi);
    const double cotphi = 1. / tan(mc_phi);
    m.c_cot = c * cotphi;
    m.apex_alpha = cos(mc_phi) / sin(mc_psi);
    // principal-space tangents, FunctionsSigmaProjection.h:68-88, :163-188, :262-287
    {
        const double sin_phi = sinphi, sin_psi = sinpsi;
        double* t = m.tan_plane;
        const double den = 6.0 * G + 2.0 * (G + 3.0 * K) * sin_phi * sin_psi;
        t[0] = (sin_phi - 1.0) * (-3.0 * G + (G + 3.0 * K) * sin_psi) / den;
        t[3] = (2.0 * G - 3.0 * K) * (sin_phi + 1.0) * sin_psi / den;
        t[6] = -(sin_phi + 1.0) * (-3.0 * G + (G + 3.0 * K) * sin_psi) / den;
        t[1] = 0.0;
        t[4] = 1.0;
        t[7] = 0.0;
        t[2] = -(sin_phi - 1.0) * (3.0 * G + (G + 3.0 * K) * sin_psi) / den;
        t[5] = (2.0 * G - 3.0 * K) * (sin_phi - 1.0) * sin_psi / den;
        t[8] = (sin_phi + 1.0) * (3.0 * G + (G + 3.0 * K) * sin_psi) / den;
    }
    {
        const double sin_phi = sinphi, sin_psi = sinpsi;
        double* t = m.tan_right;
        const double a = 4.0 * G * (1.0 + (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
        const double b = 2.0 * G * (1.0 + sin_phi + sin_psi - (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
        t[0] = (3.0 * a + 3.0 * b - 4.0 * (1.0 + sin_phi) * (3.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
        t[3] = (2.0 * (1.0 + sin_phi) * (-6.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
        t[6] = (2.0 * (1.0 + sin_phi) * (-6.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
        t[1] = (-2 * (-1 + sin_phi) * (3 * K * sin_psi + G * (3 + sin_psi))) / (3. * (a + b));
        t[4] = (-3 * b * b + 3 * a * (a + 2 * G * (-1 + sin_phi)) -
                2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (-1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
        t[7] = (2 * (-1 + sin_phi) * (b * G * (-3 + sin_psi) + a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi)) /
               (3. * (a * a - b * b));
        t[2] = (-2 * (-1 + sin_phi) * (3 * K * sin_psi + G * (3 + sin_psi))) / (3. * (a + b));
        t[5] = (2 * (-1 + sin_phi) * (b * G * (-3 + sin_psi) + a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi)) /
               (3. * (a * a - b * b));
        t[8] = (-3 * b * b + 3 * a * (a + 2 * G * (-1 + sin_phi)) -
                2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (-1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
    }
    {
        const double sin_phi = sinphi, sin_psi = sinpsi;
        double* t = m.tan_left;
        const double a = 4.0 * G * (1.0 + (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
        const double b = 2.0 * G * (1.0 - sin_phi - sin_psi - (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
        t[0] = (-3 * b * b + 3 * a * (a - 2 * G * (1 + sin_phi)) -
                2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
        t[3] = (2 * (1 + sin_phi) * (a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi + b * G * (3 + sin_psi))) /
               (3. * (a * a - b * b));
        t[6] = (-2 * (1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
        t[1] = (2 * (1 + sin_phi) * (a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi + b * G * (3 + sin_psi))) /
               (3. * (a * a - b * b));
        t[4] = (-3 * b * b + 3 * a * (a - 2 * G * (1 + sin_phi)) -
                2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
        t[7] = (-2 * (1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
        t[2] = (2 * (-1 + sin_phi) * (G * (-3 + sin_psi) - 6 * K * sin_psi)) / (3. * (a + b));
        t[5] = (2 * (-1 + sin_phi) * (G * (-3 + sin_psi) - 6 * K * sin_psi)) / (3. * (a + b));
        t[8] = (3 * a + 3 * b - 4 * (-1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
    }
    return m;
}

}  // namespace ipf

using namespace ipf;

namespace {

thread_local std::string g_last_error;

struct Group {
    int color;
    int shape;
    // element slots: reference block indexes sorted by (material, index), each
    // material run padded with -1 to whole tiles
    std::vector<int64_t> elems;
    std::vector<int> tmeta;      // per tile: (material << 8) | valid elements
    int64_t nreal = 0;           // real elements
    int64_t tile0 = 0;           // first tile in the context's tile metadata
    int64_t ntiles = 0;
    int64_t pt0 = 0;
    int64_t b_off = 0;
    int64_t dof_off = 0;
    int64_t slot0 = 0;                         // first element slot of the group (geometry rows)
    int64_t f_off[FMT_COUNT] = {0, 0, 0, 0, 0, 0};  // first value of the group in each stored B format
    int64_t l_off = 0;                         // first CSR position index of the group (tangent pattern)
};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    int64_t n = 0;
};

}  // namespace

struct ipf_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;

    // sizes
    int64_t nblocks = 0, npts = 0, rows = 0, cols = 0, neq = 0, ncolor = 0, nstorage = 0, npad = 0;
    bool have_data = false;
    bool update_mem = false;
    double newton_atol = 0.0;  // absolute residual floor of ipf_newton[_segmented]
    // ipf_newton's stop test of the reference (main.cpp:239) when > 0:
    // ||r|| < tol and ||du|| < tol after an update
    double newton_ref_tol = 0.0;
    ipf_newton_monitor monitor = nullptr;  // per-iteration report of ipf_newton[_segmented]
    void* monitor_user = nullptr;
    int residual_mode = IPF_RESIDUAL_EXACT;  // ipf_set_residual_mode
    int assembly = IPF_ASSEMBLY_COLORED;     // ipf_set_residual_assembly
    // every call that may change the committed state bumps this; the patch
    // layout's copy of the plastic strain is refreshed when it is behind
    uint64_t state_version = 1;
    // patch layout: the builder's input kept from the transfer (geometric
    // data with (x, y) DoF pairs only), the device arrays built on first use
    std::vector<PatchElemIn> patch_in;
    std::vector<int> patch_in_nodes;
    bool patch_built = false;
    int patch_nw = kPatchWavesDefault;  // wavefront tiles per patch of the next build
    PatchHost patch;
    uint64_t patch_state_version = 0;
    int* dPtmeta = nullptr;
    double* dPgeo = nullptr;
    unsigned short* dPlnode = nullptr;
    int64_t* dPnodeOff = nullptr;
    int* dPnode = nullptr;
    int* dPslot = nullptr;
    int* dPcbeg = nullptr;
    unsigned short* dPcontrib = nullptr;
    int64_t* dPptmap = nullptr;
    double* dPep[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dPpartial = nullptr;
    int* dPfixDof = nullptr;
    int* dPfixPtr = nullptr;
    int* dPorphan = nullptr;
    // material table (index = element material id) and the per-element ids
    // that the next ipf_transfer_data groups by
    std::vector<MatConst> mats;
    MatConst* dMats = nullptr;
    std::vector<int> elem_mat_next;
    std::vector<int> elem_key_next;  // ipf_set_element_order_keys, for the next transfer
    int max_mat = 0;              // largest material id of the transferred data
    int* dTmeta = nullptr;
    int64_t* dNewOf = nullptr;    // reference point -> new point index
    int* dPmat = nullptr;         // material id per new point
    int* dHist = nullptr;         // counting-sort histograms

    std::vector<Group> groups;

    // device buffers
    double* dB = nullptr;
    int* dDof = nullptr;
    int* dCol = nullptr;        // column index col_first[e] + j, same layout as dDof
    double* dW = nullptr;
    double* dEp[4] = {nullptr, nullptr, nullptr, nullptr};
    double* dAlpha = nullptr;
    signed char* dMtype = nullptr;
    signed char* dBranch = nullptr;
    int64_t* dOrig = nullptr;   // orig point of each new slot (-1 for padding)
    double* dFlin = nullptr;
    // dDof with kFirstWrite on each DoF's first write in launch order, for the
    // residual launches; fold_flin: every DoF has one, so those launches start
    // r from f_lin themselves and need no copy
    int* dDofR = nullptr;
    bool fold_flin = false;
    // per element slot and node, the node's x DoF (y = x + 1), flagged like
    // dDofR, when every element's DoFs come in such pairs (GroupDev::node)
    int* dNodeR = nullptr;
    double* dU = nullptr;       // staging for host-pointer calls
    double* dR = nullptr;
    int64_t bytes = 0;

    // tangent / solver workspace (allocated on first use)
    double* dDep = nullptr;     // 9 * npad, SoA by new point index
    double* dStrain = nullptr;  // 3 * npts, reference order
    double* dSigw = nullptr;    // 3 * npts
    signed char* dMask = nullptr;  // neq
    double* dDiag = nullptr;
    // preconditioner: the pairs' coupling and z = pa r + pb r_partner
    // (launch_block_jacobi); pairs_aligned: every element numbers node a's
    // DoFs as (2k, 2k + 1), so the DoF pairs of the vectors are the nodes
    double* dOffd = nullptr;
    double* dPa = nullptr;
    double* dPb = nullptr;
    bool pairs_aligned = false;
    double* dCr = nullptr;
    double* dCz = nullptr;
    double* dCp = nullptr;
    double* dCq = nullptr;
    double* dNr = nullptr;      // Newton residual
    double* dNd = nullptr;      // Newton correction
    double* dLs = nullptr;      // u before a line-searched Newton update
    int ls_halvings = 0;        // ipf_set_newton_line_search (0: full Newton steps)
    double* dNt = nullptr;      // the Newton step while the Thomas correction is solved
    bool newton_thomas = false; // ipf_set_newton_thomas
    double* dPartial = nullptr; // dot partials
    double* dScal = nullptr;    // CG scalars
    bool have_mask = false;

    // alternative B formats (index FMT_*; FMT_D64 is dB, FMT_G64 stores nothing)
    void* dBf[FMT_COUNT] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // FP32 copy of w Dep for the FMT_G32 tangent product, current when dep32_ok
    float* dDep32 = nullptr;
    bool dep32_ok = false;
    // geometry of ipf_transfer_geometry: corners per element slot (8 doubles,
    // group slot order) and the shape tables (offset per shape, -1 if absent)
    double* dGeo = nullptr;
    double* dTab = nullptr;
    int64_t tab_off[SHAPE_COUNT] = {-1, -1, -1, -1, -1, -1};
    bool have_geometry = false;
    int rfmt = FMT_D64;  // B format of the fused residual
    int tfmt = FMT_D64;  // B format of the tangent product
    bool compact_ok = false;  // every block has the 2D elasticity pattern
    bool have_tangent = false;

    // CG iterations as a captured HIP graph (one block of kCgBlock
    // iterations between convergence checks), replayed on `stream`; rebuilt
    // when the iterate pointer or the tangent format changes, dropped with
    // the data
    bool cg_graph = true;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t cg_exec = nullptr;
    const double* cg_x = nullptr;
    int cg_tfmt = -1;
    // ipf_newton: each iteration's device work up to its first host read
    // (tangent residual, residual norm, CG start, first CG block) as one
    // graph, captured per call on cap_stream
    bool newton_graph = true;

    // tangent CSR pattern (ipf_set_tangent_pattern): per group element, the
    // CSR positions of its packed upper triangle
    int* dLidx = nullptr;
    int64_t nnz_k = 0;
    bool have_pattern = false;
    // the pattern itself when ipf_build_tangent_pattern built it on the device
    int64_t* dPatIA = nullptr;  // [neq + 1]
    int64_t* dPatJA = nullptr;  // [nnz_k]
};

namespace {

int fail(ipf_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    g_last_error = buf;
    return code;
}

#define IPF_HIP(ctx, expr)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), IPF_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

template <typename T>
int dev_alloc(ipf_ctx* ctx, T** p, int64_t n) {
    *p = nullptr;
    if (n <= 0) return IPF_OK;
    IPF_HIP(ctx, hipMalloc((void**)p, sizeof(T) * (size_t)n));
    ctx->bytes += (int64_t)sizeof(T) * n;
    return IPF_OK;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void drop_cg_graph(ipf_ctx* ctx) {
    if (ctx->cg_exec) (void)hipGraphExecDestroy(ctx->cg_exec);
    ctx->cg_exec = nullptr;
    ctx->cg_x = nullptr;
    ctx->cg_tfmt = -1;
}

void release_patch(ipf_ctx* ctx) {
    dev_free(ctx->dPtmeta);
    dev_free(ctx->dPgeo);
    dev_free(ctx->dPlnode);
    dev_free(ctx->dPnodeOff);
    dev_free(ctx->dPnode);
    dev_free(ctx->dPslot);
    dev_free(ctx->dPcbeg);
    dev_free(ctx->dPcontrib);
    dev_free(ctx->dPptmap);
    for (auto& q : ctx->dPep) dev_free(q);
    dev_free(ctx->dPpartial);
    dev_free(ctx->dPfixDof);
    dev_free(ctx->dPfixPtr);
    dev_free(ctx->dPorphan);
    ctx->patch = PatchHost();
    ctx->patch_built = false;
}

void release_data(ipf_ctx* ctx) {
    drop_cg_graph(ctx);
    release_patch(ctx);
    ctx->patch_in.clear();
    ctx->patch_in_nodes.clear();
    ctx->state_version++;
    dev_free(ctx->dB);
    dev_free(ctx->dDof);
    dev_free(ctx->dCol);
    dev_free(ctx->dW);
    for (auto& p : ctx->dEp) dev_free(p);
    dev_free(ctx->dAlpha);
    dev_free(ctx->dMtype);
    dev_free(ctx->dBranch);
    dev_free(ctx->dOrig);
    dev_free(ctx->dFlin);
    dev_free(ctx->dDofR);
    ctx->fold_flin = false;
    dev_free(ctx->dNodeR);
    dev_free(ctx->dU);
    dev_free(ctx->dR);
    dev_free(ctx->dDep);
    dev_free(ctx->dDep32);
    ctx->dep32_ok = false;
    dev_free(ctx->dStrain);
    dev_free(ctx->dSigw);
    dev_free(ctx->dMask);
    dev_free(ctx->dDiag);
    dev_free(ctx->dOffd);
    dev_free(ctx->dPa);
    dev_free(ctx->dPb);
    dev_free(ctx->dCr);
    dev_free(ctx->dCz);
    dev_free(ctx->dCp);
    dev_free(ctx->dCq);
    dev_free(ctx->dNr);
    dev_free(ctx->dNd);
    dev_free(ctx->dLs);
    dev_free(ctx->dNt);
    dev_free(ctx->dPartial);
    dev_free(ctx->dScal);
    for (int f = 1; f < FMT_COUNT; f++) dev_free(ctx->dBf[f]);
    dev_free(ctx->dGeo);
    dev_free(ctx->dTab);
    for (auto& o : ctx->tab_off) o = -1;
    ctx->have_geometry = false;
    dev_free(ctx->dLidx);
    dev_free(ctx->dPatIA);
    dev_free(ctx->dPatJA);
    dev_free(ctx->dTmeta);
    dev_free(ctx->dNewOf);
    dev_free(ctx->dPmat);
    dev_free(ctx->dHist);
    ctx->max_mat = 0;
    ctx->nnz_k = 0;
    ctx->have_pattern = false;
    ctx->rfmt = ctx->tfmt = FMT_D64;
    ctx->compact_ok = false;
    ctx->groups.clear();
    ctx->have_data = false;
    ctx->have_mask = false;
    ctx->have_tangent = false;
    ctx->bytes = 0;
}

// Solver workspace, allocated once per transferred data set.
int ensure_solver_buffers(ipf_ctx* ctx) {
    if (ctx->dDep) return IPF_OK;
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->dDep, 9 * ctx->npad)) || (rc = dev_alloc(ctx, &ctx->dStrain, 3 * ctx->npts)) ||
        (rc = dev_alloc(ctx, &ctx->dSigw, 3 * ctx->npts)) || (rc = dev_alloc(ctx, &ctx->dDiag, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dOffd, ctx->neq)) || (rc = dev_alloc(ctx, &ctx->dPa, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dPb, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dCr, ctx->neq)) || (rc = dev_alloc(ctx, &ctx->dCz, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dCp, ctx->neq)) || (rc = dev_alloc(ctx, &ctx->dCq, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dNr, ctx->neq)) || (rc = dev_alloc(ctx, &ctx->dNd, ctx->neq)) ||
        (rc = dev_alloc(ctx, &ctx->dPartial, 1024)) || (rc = dev_alloc(ctx, &ctx->dScal, 16))) {
        return rc;
    }
    IPF_HIP(ctx, hipMemset(ctx->dDep, 0, sizeof(double) * 9 * ctx->npad));
    IPF_HIP(ctx, hipMemset(ctx->dScal, 0, sizeof(double) * 16));
    if (!ctx->dMask) {
        if ((rc = dev_alloc(ctx, &ctx->dMask, ctx->neq))) return rc;
        IPF_HIP(ctx, hipMemset(ctx->dMask, 0, ctx->neq));
    }
    return IPF_OK;
}

int read_scalar(ipf_ctx* ctx, int slot, double* out) {
    IPF_HIP(ctx, hipMemcpyAsync(out, ctx->dScal + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPF_OK;
}

// scalars [0, count) in one copy
int read_scalars(ipf_ctx* ctx, int count, double* out) {
    IPF_HIP(ctx, hipMemcpyAsync(out, ctx->dScal, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPF_OK;
}

// ||x||^2 over the free DoFs into scalar 6 (uses the CG q vector as scratch)
void masked_norm_launch(ipf_ctx* ctx, const double* x, hipStream_t s) {
    launch_mask_copy(ctx->neq, ctx->dMask, x, ctx->dCq, s);
    launch_dot(ctx->neq, ctx->dCq, ctx->dCq, ctx->dPartial, dot_blocks(ctx->neq), ctx->dScal + 6, s);
}

// ||x|| over the free DoFs
int masked_norm(ipf_ctx* ctx, const double* x, double* out) {
    masked_norm_launch(ctx, x, ctx->stream);
    double v;
    int rc = read_scalar(ctx, 6, &v);
    if (rc) return rc;
    *out = std::sqrt(v);
    return IPF_OK;
}

int shape_of(int nip, int ce) {
    for (int s = 0; s < SHAPE_COUNT; s++)
        if (shape_nip(s) == nip && shape_ce(s) == ce) return s;
    return -1;
}

GroupDev group_dev(const ipf_ctx* ctx, const Group& g, bool columns = false) {
    GroupDev d;
    d.B = ctx->dB + g.b_off;
    d.dof = (columns ? ctx->dCol : ctx->dDof) + g.dof_off;
    d.nelem = (int64_t)g.elems.size();
    d.tmeta = ctx->dTmeta + g.tile0;
    d.mats = ctx->dMats;
    d.ntiles = g.ntiles;
    d.pt0 = g.pt0;
    d.tab = (ctx->dTab && ctx->tab_off[g.shape] >= 0) ? ctx->dTab + ctx->tab_off[g.shape] : nullptr;
    d.node = nullptr;
    return d;
}

PointState point_state(ipf_ctx* ctx) {
    if (ctx->update_mem) ctx->state_version++;  // the launches it feeds may commit
    PointState s;
    s.w = ctx->dW;
    s.ep0 = ctx->dEp[0];
    s.ep1 = ctx->dEp[1];
    s.ep2 = ctx->dEp[2];
    s.ep3 = ctx->dEp[3];
    s.alpha = ctx->dAlpha;
    s.mtype = ctx->dMtype;
    s.branch = ctx->dBranch;
    return s;
}

const void* group_fmt_B(const ipf_ctx* ctx, const Group& g, int fmt) {
    if (format_geometric(fmt)) return ctx->dGeo + 8 * g.slot0;
    if (fmt == FMT_D64) return ctx->dB + g.b_off;
    return static_cast<const char*>(ctx->dBf[fmt]) + format_bytes(fmt) * (size_t)g.f_off[fmt];
}

// One residual launch for a group in the selected B format (flin: see
// residual_start).
void launch_group_residual(ipf_ctx* ctx, const Group& g, const double* u, double* r, const double* flin,
                           const PointState& st) {
    GroupDev d = group_dev(ctx, g);
    if (flin) d.dof = ctx->dDofR + g.dof_off;
    if (ctx->dNodeR) d.node = ctx->dNodeR + g.dof_off / 2;  // flagged only when f_lin is folded in
    launch_residual_fmt(ctx->rfmt, g.shape, ctx->update_mem, ctx->residual_mode == IPF_RESIDUAL_CERTIFIED_ELASTIC, d,
                        group_fmt_B(ctx, g, ctx->rfmt), u, r, flin, st, ctx->stream);
}

int launch_status(ipf_ctx* ctx, const char* what);

// Builds and uploads the patch layout (ipf_patch.h) from the transfer's
// input, once per transfer.
int ensure_patch(ipf_ctx* ctx) {
    if (ctx->patch_built) return IPF_OK;
    if (ctx->patch_in.empty())
        return fail(ctx, IPF_EUNSUPPORTED,
                    "patch assembly needs the geometric setup (ipf_transfer_geometry) with every element's DoFs "
                    "in (x, y) node pairs");
    PatchHost& h = ctx->patch;
    const std::string err = build_patches(ctx->patch_in, ctx->patch_in_nodes, ctx->neq, ctx->patch_nw, h);
    if (!err.empty()) {
        ctx->patch = PatchHost();
        return fail(ctx, IPF_EUNSUPPORTED, "%s", err.c_str());
    }
    const int64_t b0 = ctx->bytes;
    int rc;
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return IPF_OK;
        IPF_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return IPF_OK;
    };
    const int64_t npn = (int64_t)h.pnode.size();
    if ((rc = dev_alloc(ctx, &ctx->dPtmeta, h.ntiles)) || (rc = dev_alloc(ctx, &ctx->dPgeo, 8 * h.nslots)) ||
        (rc = dev_alloc(ctx, &ctx->dPlnode, (int64_t)h.lnode.size())) ||
        (rc = dev_alloc(ctx, &ctx->dPnodeOff, (int64_t)h.node_off.size())) ||
        (rc = dev_alloc(ctx, &ctx->dPnode, npn)) || (rc = dev_alloc(ctx, &ctx->dPslot, npn)) ||
        (rc = dev_alloc(ctx, &ctx->dPcbeg, npn + 1)) ||
        (rc = dev_alloc(ctx, &ctx->dPcontrib, (int64_t)h.contrib.size())) ||
        (rc = dev_alloc(ctx, &ctx->dPptmap, h.npts)) || (rc = dev_alloc(ctx, &ctx->dPep[0], h.npts)) ||
        (rc = dev_alloc(ctx, &ctx->dPep[1], h.npts)) || (rc = dev_alloc(ctx, &ctx->dPep[2], h.npts)) ||
        (rc = dev_alloc(ctx, &ctx->dPep[3], h.npts)) ||
        (rc = dev_alloc(ctx, &ctx->dPpartial, 2 * std::max<int64_t>(1, h.npartial))) ||
        (rc = dev_alloc(ctx, &ctx->dPfixDof, (int64_t)h.fix_dof.size())) ||
        (rc = dev_alloc(ctx, &ctx->dPfixPtr, (int64_t)h.fix_ptr.size())) ||
        (rc = dev_alloc(ctx, &ctx->dPorphan, (int64_t)h.orphan.size())) ||
        (rc = up(ctx->dPtmeta, h.tmeta.data(), sizeof(int) * h.tmeta.size())) ||
        (rc = up(ctx->dPlnode, h.lnode.data(), sizeof(unsigned short) * h.lnode.size())) ||
        (rc = up(ctx->dPnodeOff, h.node_off.data(), sizeof(int64_t) * h.node_off.size())) ||
        (rc = up(ctx->dPnode, h.pnode.data(), sizeof(int) * npn)) ||
        (rc = up(ctx->dPslot, h.pslot.data(), sizeof(int) * npn)) ||
        (rc = up(ctx->dPcbeg, h.cbeg.data(), sizeof(int) * (npn + 1))) ||
        (rc = up(ctx->dPcontrib, h.contrib.data(), sizeof(unsigned short) * h.contrib.size())) ||
        (rc = up(ctx->dPptmap, h.pt_map.data(), sizeof(int64_t) * h.npts)) ||
        (rc = up(ctx->dPfixDof, h.fix_dof.data(), sizeof(int) * h.fix_dof.size())) ||
        (rc = up(ctx->dPfixPtr, h.fix_ptr.data(), sizeof(int) * h.fix_ptr.size())) ||
        (rc = up(ctx->dPorphan, h.orphan.data(), sizeof(int) * h.orphan.size()))) {
        release_patch(ctx);
        return rc;
    }
    // corners gathered on the device from the colored layout's slots
    int64_t* dSlotGeo = nullptr;
    if ((rc = dev_alloc(ctx, &dSlotGeo, h.nslots)) ||
        (rc = up(dSlotGeo, h.slot_geo.data(), sizeof(int64_t) * h.nslots))) {
        dev_free(dSlotGeo);
        release_patch(ctx);
        return rc;
    }
    launch_patch_geo(h.nslots, dSlotGeo, ctx->dGeo, ctx->dPgeo, ctx->stream);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    dev_free(dSlotGeo);
    if (e != hipSuccess || (rc = launch_status(ctx, "patch geometry"))) {
        release_patch(ctx);
        return rc ? rc : fail(ctx, IPF_EHIP, "patch geometry: %s", hipGetErrorString(e));
    }
    (void)b0;
    ctx->patch_built = true;
    ctx->patch_state_version = 0;
    return IPF_OK;
}

// The non-committing residual over the patch layout: refresh the patch copy
// of the plastic strain when the committed state changed, one launch per
// element shape, then the shared nodes' fix-up.
// ev (optional, ranges + 3 events): recorded before the state refresh, after
// it, after each shape's launch and after the fix-up.
int patch_residual(ipf_ctx* ctx, const double* u, double* r, hipEvent_t* ev = nullptr) {
    int rc = ensure_patch(ctx);
    if (rc) return rc;
    const PatchHost& h = ctx->patch;
    if (ev) IPF_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
    if (ctx->patch_state_version != ctx->state_version) {
        launch_patch_state(h.npts, ctx->dPptmap, ctx->dEp, ctx->dPep, ctx->stream);
        ctx->patch_state_version = ctx->state_version;
    }
    if (ev) IPF_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    int iev = 2;
    const bool fast = ctx->residual_mode == IPF_RESIDUAL_CERTIFIED_ELASTIC;
    for (const auto& R : h.ranges) {
        PatchDev p;
        const int ET = shape_et(R.shape), TP = shape_tp(R.shape);
        p.npatch = R.npatch;
        p.tmeta = ctx->dPtmeta + R.tile0;
        p.mats = ctx->dMats;
        p.geo = ctx->dPgeo + 8 * R.slot0;
        p.lnode = ctx->dPlnode + R.lnode0;
        p.node_off = ctx->dPnodeOff + R.patch0;
        p.pnode = ctx->dPnode;
        p.pslot = ctx->dPslot;
        p.cbeg = ctx->dPcbeg;
        p.contrib = ctx->dPcontrib;
        p.ep0 = ctx->dPep[0] + R.pt0;
        p.ep1 = ctx->dPep[1] + R.pt0;
        p.ep2 = ctx->dPep[2] + R.pt0;
        p.ep3 = ctx->dPep[3] + R.pt0;
        p.tab = ctx->dTab + ctx->tab_off[R.shape];
        (void)ET;
        (void)TP;
        launch_residual_patch(R.shape, h.nw, fast, p, u, r, ctx->dFlin, ctx->dPpartial, ctx->stream);
        if (ev) IPF_HIP(ctx, hipEventRecord(ev[iev++], ctx->stream));
    }
    launch_patch_fixup((int64_t)h.fix_dof.size(), ctx->dPfixDof, ctx->dPfixPtr, ctx->dPpartial,
                       (int64_t)h.orphan.size(), ctx->dPorphan, ctx->dFlin, r, ctx->stream);
    if (ev) IPF_HIP(ctx, hipEventRecord(ev[iev], ctx->stream));
    return launch_status(ctx, "patch residual");
}

// r = f_lin (TPZElastoPlasticIntPointsStructMatrix.cpp:147) before the
// residual launches: folded into them when every DoF has a first write
// (returns f_lin for launch_group_residual), else a copy (returns null).
const double* residual_start(ipf_ctx* ctx, double* r) {
    if (ctx->fold_flin) return ctx->dFlin;
    launch_copy(ctx->neq, ctx->dFlin, r, ctx->stream);
    return nullptr;
}

// y = K v on stream s: the tangent launches of every group in the selected
// B format.  When every DoF has a first write (fold_flin), the flagged entries
// start from 0 and y needs no clear; else y is cleared first.
void tangent_product(ipf_ctx* ctx, const double* v, double* y, hipStream_t s) {
    const bool zf = ctx->fold_flin;
    if (!zf) launch_fill(ctx->neq, y, 0.0, s);
    for (const auto& g : ctx->groups) {
        GroupDev d = group_dev(ctx, g);
        if (zf) d.dof = ctx->dDofR + g.dof_off;
        if (ctx->dNodeR) d.node = ctx->dNodeR + g.dof_off / 2;
        launch_tangent_fmt(ctx->tfmt, g.shape, d, group_fmt_B(ctx, g, ctx->tfmt), v, y, ctx->dDep, ctx->dDep32,
                           ctx->npad, zf, s);
    }
}

// The FP32 copy of w Dep the FMT_G32 tangent product reads, made after each
// new tangent (outside any captured graph).
int ensure_dep32(ipf_ctx* ctx) {
    if (ctx->tfmt != FMT_G32 || !ctx->have_tangent || ctx->dep32_ok) return IPF_OK;
    if (!ctx->dDep32) {
        int rc = dev_alloc(ctx, &ctx->dDep32, 9 * ctx->npad);
        if (rc) return rc;
    }
    launch_to_f32(9 * ctx->npad, ctx->dDep, ctx->dDep32, ctx->stream);
    ctx->dep32_ok = true;
    return IPF_OK;
}

constexpr int kCgBlock = 10;  // CG iterations between convergence checks

// `count` preconditioned CG iterations on stream s: q = K p (masked) and
// p.q, alpha, x / r / z update and r.z, beta, new p.  The scalars stay on
// the device, so the sequence has no host dependency.
void cg_iterations(ipf_ctx* ctx, double* x, int count, hipStream_t s) {
    const int64_t n = ctx->neq;
    const int nb = dot_blocks(n);
    for (int k = 0; k < count; k++) {
        tangent_product(ctx, ctx->dCp, ctx->dCq, s);
        launch_cg_mask_dot(n, ctx->dMask, ctx->dCp, ctx->dCq, ctx->dPartial, nb, ctx->dScal, s);
        launch_cg_update_dot(n, ctx->dScal, ctx->dCp, ctx->dCq, x, ctx->dCr, ctx->dPa, ctx->dPb, ctx->dCz, ctx->dPartial,
                             nb, s);
        launch_cg_direction(n, ctx->dScal, ctx->dCz, ctx->dCp, s);
    }
}

// The graph of kCgBlock iterations for iterate x (captured on the context's
// capture stream; nullptr when capture is off or fails, and the caller then
// launches the iterations one by one).
hipGraphExec_t cg_graph(ipf_ctx* ctx, double* x) {
    if (!ctx->cg_graph) return nullptr;
    if (ctx->cg_exec && ctx->cg_x == x && ctx->cg_tfmt == ctx->tfmt) return ctx->cg_exec;
    drop_cg_graph(ctx);
    if (!ctx->cap_stream && hipStreamCreate(&ctx->cap_stream) != hipSuccess) {
        ctx->cap_stream = nullptr;
        ctx->cg_graph = false;
        return nullptr;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        ctx->cg_graph = false;
        return nullptr;
    }
    cg_iterations(ctx, x, kCgBlock, ctx->cap_stream);
    const hipError_t e1 = hipStreamEndCapture(ctx->cap_stream, &graph);
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = (e1 == hipSuccess) ? hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) : e1;
    if (graph) (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
        (void)hipGetLastError();
        ctx->cg_graph = false;
        return nullptr;
    }
    ctx->cg_exec = exec;
    ctx->cg_x = x;
    ctx->cg_tfmt = ctx->tfmt;
    return exec;
}

int check_ready(ipf_ctx* ctx, bool need_material) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (!ctx->have_data) return fail(ctx, IPF_ESTATE, "integrator data not transferred (call ipf_transfer_data)");
    if (need_material && (int)ctx->mats.size() <= ctx->max_mat)
        return fail(ctx, IPF_ESTATE, "material %d not set (call ipf_set_material / ipf_set_material_table)",
                    ctx->mats.empty() ? 0 : ctx->max_mat);
    return IPF_OK;
}

int launch_status(ipf_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, IPF_EHIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return IPF_OK;
}

// Makes format `fmt` resident (converted on the device from the dense FP64 tiles).
int ensure_format(ipf_ctx* ctx, int fmt) {
    if (fmt < 0 || fmt >= FMT_COUNT) return fail(ctx, IPF_EINVAL, "ipf_set_b_format: unknown format %d", fmt);
    if (format_geometric(fmt))
        return ctx->have_geometry
                   ? IPF_OK
                   : fail(ctx, IPF_EINVAL, "geometric B needs the element geometry (ipf_transfer_geometry)");
    if (format_compact(fmt) && !ctx->compact_ok)
        return fail(ctx, IPF_EINVAL, "compact B needs every block in the 2D elasticity pattern; this data is not");
    if (fmt == FMT_D64 || ctx->dBf[fmt]) return IPF_OK;
    int64_t n = 0;
    for (const auto& g : ctx->groups)
        n = std::max(n, g.f_off[fmt] + g.ntiles * (int64_t)format_values(g.shape, fmt) * shape_tp(g.shape));
    char* p = nullptr;
    int rc = dev_alloc(ctx, &p, n * format_bytes(fmt));
    if (rc) return rc;
    ctx->dBf[fmt] = p;
    for (const auto& g : ctx->groups)
        launch_convert_B(fmt, g.shape, ctx->dB + g.b_off, p + (size_t)format_bytes(fmt) * g.f_off[fmt], g.ntiles,
                         ctx->stream);
    if ((rc = launch_status(ctx, "convert_B"))) return rc;
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPF_OK;
}

}  // namespace

extern "C" {

const char* ipf_last_error(const ipf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }
const char* ipf_global_last_error(void) { return g_last_error.c_str(); }
int ipf_internal_set_error(const char* msg) {
    g_last_error = msg ? msg : "";
    return IPF_OK;
}

// ------------------------------------------------- scratch of the no-alloc sort
// The per-block key histograms of ipf_branch_order_noalloc live in this
// file-scope device array instead of a hipMalloc'd buffer: 4 MiB, one copy
// on every GPU the library runs on, held until the process ends.  Host code
// never passes the array's name to a kernel (that name is a host-side
// stand-in, not a device address): each device's address is looked up once,
// with that device current, when the first context on it is created, and
// kept in g_scratch_dev[device].
constexpr int64_t kScratchInts = int64_t(1) << 20;
static __device__ __align__(256) int scratch[kScratchInts];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_dev[kMaxScratchDevices] = {};

// The device address of `scratch` on `device` (current); once per device.
int scratch_setup(int device) {
    if (device >= kMaxScratchDevices)
        return fail(nullptr, IPF_EUNSUPPORTED, "device %d beyond the %d devices the scratch table holds", device,
                    kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return IPF_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(nullptr, IPF_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                    hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<int*>(p);
    return IPF_OK;
}

int* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxScratchDevices) ? g_scratch_dev[device] : nullptr;
}
}  // namespace

int ipf_ctx_create(int device, ipf_ctx** out) {
    if (!out) return fail(nullptr, IPF_EINVAL, "ipf_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(nullptr, IPF_EHIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(nullptr, IPF_EINVAL, "device %d out of range [0,%d)", device, ndev);
    ipf_ctx* ctx = new (std::nothrow) ipf_ctx();
    if (!ctx) return fail(nullptr, IPF_EINVAL, "out of host memory");
    ctx->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&ctx->own_stream);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, IPF_EHIP, "context setup failed: %s", hipGetErrorString(e));
    }
    const int rs = scratch_setup(device);
    if (rs != IPF_OK) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return rs;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return IPF_OK;
}

int ipf_ctx_destroy(ipf_ctx* ctx) {
    if (!ctx) return IPF_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_data(ctx);
    dev_free(ctx->dMats);
    if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return IPF_OK;
}

int ipf_set_stream(ipf_ctx* ctx, void* s) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return IPF_OK;
}

void* ipf_get_stream(const ipf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int ipf_synchronize(ipf_ctx* ctx) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    IPF_HIP(ctx, hipGetLastError());
    return IPF_OK;
}

int ipf_info(const ipf_ctx* ctx, int64_t out[8]) {
    if (!ctx || !out) return fail(nullptr, IPF_EINVAL, "ipf_info: null argument");
    out[0] = ctx->nblocks;
    out[1] = ctx->npts;
    out[2] = ctx->rows;
    out[3] = ctx->cols;
    out[4] = ctx->neq;
    out[5] = ctx->ncolor;
    out[6] = ctx->nstorage;
    out[7] = ctx->bytes;
    return IPF_OK;
}

// Element geometry of ipf_transfer_geometry: corners [8 * nblocks] and one
// host shape table per kernel shape (null where no element has that shape).
struct GeoInput {
    const double* corners = nullptr;
    const double* tables[SHAPE_COUNT] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// The common part of ipf_transfer_data (storage and weight given, geo null)
// and ipf_transfer_geometry (storage and weight null, geo given: B and the
// weights are computed on the device).
static int transfer_impl(ipf_ctx* ctx, int64_t nblocks, const int* row_sizes, const int* col_sizes,
                         const int64_t* matrix_position, const int* row_first, const int* col_first,
                         const double* storage, const int* dof_indexes, const int* color_indexes, int64_t ncolor,
                         int64_t neq, const double* weight, const GeoInput* geo) {
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    release_data(ctx);

    // --- validate the IrregularBlocks prefix sums (TPZNumericalIntegrator.cpp:294-317)
    if (nblocks > 0 && (row_first[0] != 0 || col_first[0] != 0 || matrix_position[0] != 0))
        return fail(ctx, IPF_EINVAL, "first row/col/matrix position must be 0");
    std::vector<int> elem_shape(nblocks);
    for (int64_t e = 0; e < nblocks; e++) {
        const int R = row_sizes[e], C = col_sizes[e];
        if (R <= 0 || C <= 0 || R % 3 != 0 || C % 2 != 0)
            return fail(ctx, IPF_EINVAL, "block %lld: row size %d / col size %d not 3*n_ip / 2*n_shape", (long long)e,
                        R, C);
        if ((int64_t)row_first[e + 1] != (int64_t)row_first[e] + R || (int64_t)col_first[e + 1] != (int64_t)col_first[e] + C ||
            matrix_position[e + 1] != matrix_position[e] + (int64_t)R * C)
            return fail(ctx, IPF_EINVAL, "block %lld: inconsistent first-index / position prefix sums", (long long)e);
        const int s = shape_of(R / 3, C);
        if (s < 0)
            return fail(ctx, IPF_EUNSUPPORTED, "block %lld: shape (n_ip=%d, C_e=%d) has no kernel (supported: Q1/Q2/Q3, T1/T2/T3)",
                        (long long)e, R / 3, C);
        elem_shape[e] = s;
    }
    const int64_t rows = nblocks ? row_first[nblocks] : 0;
    const int64_t cols = nblocks ? col_first[nblocks] : 0;
    const int64_t npts = rows / 3;

    // --- element colors from fColorIndexes = fDoFIndexes + color*neq (TPZNumericalIntegrator.cpp:503)
    std::vector<int> elem_color(nblocks);
    for (int64_t e = 0; e < nblocks; e++) {
        int color = -1;
        for (int64_t i = col_first[e]; i < col_first[e + 1]; i++) {
            const int64_t d = dof_indexes[i];
            if (d < 0 || d >= neq)
                return fail(ctx, IPF_EINVAL, "dof index %lld of block %lld out of range [0,%lld)", (long long)d,
                            (long long)e, (long long)neq);
            const int64_t diff = (int64_t)color_indexes[i] - d;
            if (diff < 0 || diff % neq != 0 || diff / neq >= ncolor)
                return fail(ctx, IPF_EINVAL, "color index %lld of block %lld is not dof + color*neq",
                            (long long)color_indexes[i], (long long)e);
            const int c = (int)(diff / neq);
            if (color >= 0 && c != color)
                return fail(ctx, IPF_EINVAL, "block %lld has entries of colors %d and %d", (long long)e, color, c);
            color = c;
        }
        elem_color[e] = color;
    }
    // --- coloring invariant: no DoF twice within one color (race-free accumulation)
    {
        std::vector<int> owner(neq, -1);
        for (int c = 0; c < ncolor; c++) {
            for (int64_t e = 0; e < nblocks; e++) {
                if (elem_color[e] != c) continue;
                for (int64_t i = col_first[e]; i < col_first[e + 1]; i++) {
                    const int d = dof_indexes[i];
                    if (owner[d] == c)
                        return fail(ctx, IPF_EINVAL, "color %d: DoF %d is shared by two elements (block %lld)", c, d,
                                    (long long)e);
                    owner[d] = c;
                }
            }
        }
    }

    // --- element materials (ipf_set_element_materials before this call; all 0 otherwise)
    std::vector<int> emat(nblocks, 0);
    if (!ctx->elem_mat_next.empty()) {
        if ((int64_t)ctx->elem_mat_next.size() != nblocks)
            return fail(ctx, IPF_EINVAL, "element materials were set for %lld blocks, the data has %lld",
                        (long long)ctx->elem_mat_next.size(), (long long)nblocks);
        emat = ctx->elem_mat_next;
    }
    int max_mat = 0;
    for (int64_t e = 0; e < nblocks; e++) max_mat = std::max(max_mat, emat[e]);
    // --- order keys inside a material run (ipf_set_element_order_keys; all 0 otherwise)
    std::vector<int> ekey(nblocks, 0);
    if (!ctx->elem_key_next.empty()) {
        if ((int64_t)ctx->elem_key_next.size() != nblocks)
            return fail(ctx, IPF_EINVAL, "element order keys were set for %lld blocks, the data has %lld",
                        (long long)ctx->elem_key_next.size(), (long long)nblocks);
        ekey = ctx->elem_key_next;
    }

    // --- groups: color-major, then shape; inside a group elements by
    // (material, index), every material run padded to whole tiles
    std::vector<Group> groups;
    int64_t ntile_all = 0;
    for (int c = 0; c < ncolor; c++)
        for (int s = 0; s < SHAPE_COUNT; s++) {
            std::vector<int64_t> es;
            for (int64_t e = 0; e < nblocks; e++)
                if (elem_color[e] == c && elem_shape[e] == s) es.push_back(e);
            if (es.empty()) continue;
            std::stable_sort(es.begin(), es.end(), [&](int64_t a, int64_t b) {
                return emat[a] != emat[b] ? emat[a] < emat[b] : ekey[a] < ekey[b];
            });
            Group g;
            g.color = c;
            g.shape = s;
            g.nreal = (int64_t)es.size();
            const int ET = shape_et(s);
            size_t i = 0;
            while (i < es.size()) {
                const int mid = emat[es[i]];
                size_t j = i;
                while (j < es.size() && emat[es[j]] == mid) j++;
                for (size_t k = i; k < j; k += ET) {
                    const int nv = (int)std::min<size_t>(ET, j - k);
                    for (int q = 0; q < ET; q++) g.elems.push_back(q < nv ? es[k + q] : -1);
                    g.tmeta.push_back((mid << 8) | nv);
                }
                i = j;
            }
            g.tile0 = ntile_all;
            ntile_all += (int64_t)g.tmeta.size();
            groups.push_back(std::move(g));
        }
    int64_t npad = 0, nb = 0, nd = 0;
    int64_t nf[FMT_COUNT] = {0, 0, 0, 0, 0, 0};
    for (auto& g : groups) {
        const int TP = shape_tp(g.shape), CE = shape_ce(g.shape);
        g.ntiles = (int64_t)g.tmeta.size();
        g.pt0 = npad;
        g.b_off = nb;
        g.dof_off = nd;
        for (int f = 0; f < FMT_COUNT; f++) {
            g.f_off[f] = nf[f];
            nf[f] += g.ntiles * (int64_t)format_values(g.shape, f) * TP;
        }
        npad += g.ntiles * TP;
        nb += g.ntiles * (int64_t)(3 * CE) * TP;
        nd += (int64_t)g.elems.size() * CE;
    }
    // --- compact B possible? Every point of every block must have the 2D
    // elasticity pattern B(3ip + c + j, 2a + c) = dN_a/dx_j
    // (TPZNumericalIntegrator.cpp:371-381): row xx has zeros in the y columns,
    // row yy zeros in the x columns, and row xy repeats the other two.  Blocks
    // built from geometry have it by construction.
    bool compact_ok = true;
    for (int64_t e = 0; storage && e < nblocks && compact_ok; e++) {
        const int C = col_sizes[e], nip = row_sizes[e] / 3;
        const double* blk = storage + matrix_position[e];
        for (int ip = 0; ip < nip && compact_ok; ip++) {
            const double* rxx = blk + (int64_t)(3 * ip) * C;
            const double* rxy = rxx + C;
            const double* ryy = rxy + C;
            for (int a = 0; a < C / 2; a++) {
                if (rxx[2 * a + 1] != 0.0 || ryy[2 * a] != 0.0 || rxy[2 * a] != ryy[2 * a + 1] ||
                    rxy[2 * a + 1] != rxx[2 * a]) {
                    compact_ok = false;
                    break;
                }
            }
        }
    }

    // --- host staging in the tiled layout
    std::vector<double> hB;
    std::vector<int> hDof, hCol;
    std::vector<double> hW(npad, 0.0);
    std::vector<int64_t> hOrig(npad, -1);
    std::vector<double> hGeo;
    int64_t nslots = 0;
    for (auto& g : groups) {
        g.slot0 = nslots;
        nslots += (int64_t)g.elems.size();
    }
    try {
        if (storage) hB.assign(nb, 0.0);
        if (geo) hGeo.assign(8 * nslots, 0.0);
        hDof.resize(nd);
        hCol.resize(nd);
    } catch (const std::bad_alloc&) {
        return fail(ctx, IPF_EINVAL, "out of host memory staging %lld B values", (long long)nb);
    }
    for (const auto& g : groups) {
        const int NIP = shape_nip(g.shape), CE = shape_ce(g.shape), ET = shape_et(g.shape), TP = shape_tp(g.shape);
        const int NB = 3 * CE;
        const int64_t ne = (int64_t)g.elems.size();
        for (int64_t k = 0; k < ne; k++) {
            const int64_t e = g.elems[k];
            if (e < 0) {  // padding slot: zero B and corners, DoF 0 (never read), no point
                for (int j = 0; j < CE; j++) hDof[g.dof_off + k * CE + j] = hCol[g.dof_off + k * CE + j] = 0;
                continue;
            }
            std::memcpy(&hDof[g.dof_off + k * CE], &dof_indexes[col_first[e]], sizeof(int) * CE);
            for (int j = 0; j < CE; j++) hCol[g.dof_off + k * CE + j] = col_first[e] + j;
            if (geo) std::memcpy(&hGeo[8 * (g.slot0 + k)], geo->corners + 8 * e, sizeof(double) * 8);
            const int64_t t = k / ET;
            const int el = (int)(k % ET);
            for (int ip = 0; ip < NIP; ip++) {
                const int q = el * NIP + ip;
                const int64_t P = g.pt0 + t * TP + q;
                const int64_t o = row_first[e] / 3 + ip;
                hOrig[P] = o;
                if (!storage) continue;
                hW[P] = weight[o];
                double* tileB = hB.data() + g.b_off + t * (int64_t)NB * TP;
                const double* src = storage + matrix_position[e];
                for (int m = 0; m < NB; m++) {
                    // m = k*CE + j <-> row 3*ip + k, column j of the row-major block
                    tileB[((int64_t)(m >> 1) * TP + q) * 2 + (m & 1)] = src[(int64_t)(3 * ip) * CE + m];
                }
            }
        }
    }

    // --- first writes: in launch order (groups in order; inside a color each
    // DoF belongs to one element), the entry of each DoF that is written
    // first starts from f_lin; when some DoF belongs to no element the
    // residual keeps the copy r = f_lin
    std::vector<int> hDofR(hDof);
    bool fold = false;
    {
        std::vector<unsigned char> seen(neq, 0);
        int64_t touched = 0;
        for (const auto& g : groups) {
            const int CE = shape_ce(g.shape), ET = shape_et(g.shape);
            for (int64_t t = 0; t < g.ntiles; t++) {
                const int nvt = g.tmeta[t] & 255;
                for (int64_t i = g.dof_off + t * ET * CE; i < g.dof_off + (t * ET + nvt) * CE; i++) {
                    const int d = hDof[i];
                    if (seen[d]) continue;
                    seen[d] = 1;
                    touched++;
                    hDofR[i] = d | kFirstWrite;
                }
            }
        }
        fold = touched == neq;
    }
    // node pairs: every valid element slot numbers node a's DoFs (2a, 2a + 1)
    // as (d, d + 1), whose first writes fall on the same entry pair; then the
    // residual gathers one index per node
    std::vector<int> hNodeR;
    bool pairs = true, aligned = true;
    for (const auto& g : groups) {
        const int CE = shape_ce(g.shape), ET = shape_et(g.shape);
        for (int64_t t = 0; pairs && t < g.ntiles; t++) {
            const int nvt = g.tmeta[t] & 255;
            for (int64_t i = g.dof_off + t * ET * CE; pairs && i < g.dof_off + (t * ET + nvt) * CE; i += 2)
                if (hDof[i + 1] != hDof[i] + 1 || (fold && (hDofR[i] < 0) != (hDofR[i + 1] < 0))) pairs = false;
                else if (hDof[i] & 1) aligned = false;
        }
    }
    if (pairs) {
        hNodeR.resize(nd / 2);
        for (int64_t i = 0; i < nd / 2; i++) hNodeR[i] = fold ? hDofR[2 * i] : hDof[2 * i];
    }
    // input of the patch layout (built on demand): every real element slot
    // with its node x-DoFs, corner centroid, colored geometry slot and points
    std::vector<PatchElemIn> pin;
    std::vector<int> pin_nodes;
    if (geo && pairs) {
        try {
            for (const auto& g : groups) {
                const int NIP = shape_nip(g.shape), CE = shape_ce(g.shape), ET = shape_et(g.shape),
                          TP = shape_tp(g.shape);
                for (int64_t k = 0; k < (int64_t)g.elems.size(); k++) {
                    if (g.elems[k] < 0) continue;
                    PatchElemIn E;
                    E.shape = g.shape;
                    E.mat = g.tmeta[k / ET] >> 8;
                    E.geo_slot = g.slot0 + k;
                    E.P0 = g.pt0 + (k / ET) * TP + (k % ET) * NIP;
                    E.node0 = (int64_t)pin_nodes.size();
                    const double* X = &hGeo[8 * (g.slot0 + k)];
                    E.cx = 0.25 * (X[0] + X[2] + X[4] + X[6]);
                    E.cy = 0.25 * (X[1] + X[3] + X[5] + X[7]);
                    E.key = ekey[g.elems[k]];
                    for (int a = 0; a < CE / 2; a++) pin_nodes.push_back(hDof[g.dof_off + k * CE + 2 * a]);
                    pin.push_back(E);
                }
            }
        } catch (const std::bad_alloc&) {
            pin.clear();
            pin_nodes.clear();
        }
    }

    // --- device buffers
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->dB, nb)) || (rc = dev_alloc(ctx, &ctx->dDof, nd)) ||
        (rc = dev_alloc(ctx, &ctx->dCol, nd)) ||
        (rc = dev_alloc(ctx, &ctx->dW, npad)) || (rc = dev_alloc(ctx, &ctx->dEp[0], npad)) ||
        (rc = dev_alloc(ctx, &ctx->dEp[1], npad)) || (rc = dev_alloc(ctx, &ctx->dEp[2], npad)) ||
        (rc = dev_alloc(ctx, &ctx->dEp[3], npad)) || (rc = dev_alloc(ctx, &ctx->dAlpha, npad)) ||
        (rc = dev_alloc(ctx, &ctx->dMtype, npad)) || (rc = dev_alloc(ctx, &ctx->dBranch, npad)) ||
        (rc = dev_alloc(ctx, &ctx->dOrig, npad)) || (rc = dev_alloc(ctx, &ctx->dFlin, neq)) ||
        (rc = dev_alloc(ctx, &ctx->dU, neq)) || (rc = dev_alloc(ctx, &ctx->dR, neq)) ||
        (rc = dev_alloc(ctx, &ctx->dTmeta, ntile_all)) || (rc = dev_alloc(ctx, &ctx->dNewOf, npts)) ||
        (rc = dev_alloc(ctx, &ctx->dPmat, npad)) ||
        (fold && (rc = dev_alloc(ctx, &ctx->dDofR, nd))) || (pairs && (rc = dev_alloc(ctx, &ctx->dNodeR, nd / 2)))) {
        release_data(ctx);
        return rc;
    }
    int64_t ntab = 0;
    if (geo) {
        for (int sh = 0; sh < SHAPE_COUNT; sh++) {
            if (!geo->tables[sh]) continue;
            ctx->tab_off[sh] = ntab;
            ntab += (shape_table_size(shape_nip(sh), shape_ce(sh) / 2) + 31) / 32 * 32;  // 256-byte aligned rows
        }
        if ((rc = dev_alloc(ctx, &ctx->dGeo, 8 * nslots)) || (rc = dev_alloc(ctx, &ctx->dTab, ntab))) {
            release_data(ctx);
            return rc;
        }
    }
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return IPF_OK;
        IPF_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return IPF_OK;
    };
    auto zero = [&](void* dst, size_t bytes) -> int {
        if (bytes == 0) return IPF_OK;
        IPF_HIP(ctx, hipMemset(dst, 0, bytes));
        return IPF_OK;
    };
    if ((storage && (rc = up(ctx->dB, hB.data(), sizeof(double) * nb))) ||
        (rc = up(ctx->dDof, hDof.data(), sizeof(int) * nd)) || (rc = up(ctx->dCol, hCol.data(), sizeof(int) * nd)) ||
        (storage && (rc = up(ctx->dW, hW.data(), sizeof(double) * npad))) ||
        (rc = up(ctx->dOrig, hOrig.data(), sizeof(int64_t) * npad)) ||
        (rc = zero(ctx->dEp[0], sizeof(double) * npad)) || (rc = zero(ctx->dEp[1], sizeof(double) * npad)) ||
        (rc = zero(ctx->dEp[2], sizeof(double) * npad)) || (rc = zero(ctx->dEp[3], sizeof(double) * npad)) ||
        (rc = zero(ctx->dAlpha, sizeof(double) * npad)) || (rc = zero(ctx->dMtype, npad)) ||
        (rc = zero(ctx->dBranch, npad)) || (rc = zero(ctx->dFlin, sizeof(double) * neq)) ||
        (fold && (rc = up(ctx->dDofR, hDofR.data(), sizeof(int) * nd))) ||
        (pairs && (rc = up(ctx->dNodeR, hNodeR.data(), sizeof(int) * (nd / 2))))) {
        release_data(ctx);
        return rc;
    }
    {
        std::vector<int64_t> new_of(npts, -1);
        std::vector<int> pmat(npad, 0);
        for (int64_t P = 0; P < npad; P++)
            if (hOrig[P] >= 0) new_of[hOrig[P]] = P;
        for (const auto& g : groups) {
            const int TP = shape_tp(g.shape);
            for (int64_t t = 0; t < g.ntiles; t++)
                for (int q = 0; q < TP; q++) pmat[g.pt0 + t * TP + q] = g.tmeta[t] >> 8;
        }
        if ((rc = up(ctx->dNewOf, new_of.data(), sizeof(int64_t) * npts)) ||
            (rc = up(ctx->dPmat, pmat.data(), sizeof(int) * npad))) {
            release_data(ctx);
            return rc;
        }
    }
    for (const auto& g : groups)
        if ((rc = up(ctx->dTmeta + g.tile0, g.tmeta.data(), sizeof(int) * g.tmeta.size()))) {
        release_data(ctx);
        return rc;
    }
    ctx->nblocks = nblocks;
    ctx->npts = npts;
    ctx->rows = rows;
    ctx->cols = cols;
    ctx->neq = neq;
    ctx->ncolor = ncolor;
    ctx->nstorage = nblocks ? matrix_position[nblocks] : 0;
    ctx->npad = npad;
    ctx->max_mat = max_mat;
    ctx->fold_flin = fold;
    ctx->pairs_aligned = pairs && aligned;
    ctx->groups = std::move(groups);
    ctx->compact_ok = compact_ok;
    ctx->rfmt = ctx->tfmt = FMT_D64;
    if (geo) {
        // corners and tables up, then B and the weights on the device
        if ((rc = up(ctx->dGeo, hGeo.data(), sizeof(double) * 8 * nslots))) {
            release_data(ctx);
            return rc;
        }
        for (int sh = 0; sh < SHAPE_COUNT; sh++)
            if (ctx->tab_off[sh] >= 0 &&
                (rc = up(ctx->dTab + ctx->tab_off[sh], geo->tables[sh],
                         sizeof(double) * shape_table_size(shape_nip(sh), shape_ce(sh) / 2)))) {
                release_data(ctx);
                return rc;
            }
        for (const auto& g : ctx->groups)
            launch_build_B(g.shape, group_dev(ctx, g), ctx->dGeo + 8 * g.slot0, ctx->dB + g.b_off, ctx->dW,
                           ctx->stream);
        if ((rc = launch_status(ctx, "build_B"))) {
            release_data(ctx);
            return rc;
        }
        IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->have_geometry = true;
    }
    ctx->have_data = true;
    ctx->patch_in = std::move(pin);
    ctx->patch_in_nodes = std::move(pin_nodes);
    // Blocks in the 2D elasticity pattern default to compact FP64: the same
    // results from a third of the B bytes; with the geometry at hand, to the
    // geometric form (the same results from no per-point B bytes at all).
    if (geo) {
        ctx->rfmt = ctx->tfmt = FMT_G64;
    } else if (compact_ok) {
        if ((rc = ensure_format(ctx, FMT_C64))) return rc;
        ctx->rfmt = ctx->tfmt = FMT_C64;
    }
    return IPF_OK;
}

int ipf_transfer_data(ipf_ctx* ctx, int64_t nblocks, const int* row_sizes, const int* col_sizes,
                      const int64_t* matrix_position, const int* row_first, const int* col_first,
                      const double* storage, const int* dof_indexes, const int* color_indexes, int64_t ncolor,
                      int64_t neq, const double* weight) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (nblocks < 0 || ncolor < 1 || neq < 1)
        return fail(ctx, IPF_EINVAL, "bad sizes: nblocks=%lld ncolor=%lld neq=%lld", (long long)nblocks,
                    (long long)ncolor, (long long)neq);
    if (nblocks > 0 && (!row_sizes || !col_sizes || !matrix_position || !row_first || !col_first || !storage ||
                        !dof_indexes || !color_indexes || !weight))
        return fail(ctx, IPF_EINVAL, "ipf_transfer_data: null array");
    return transfer_impl(ctx, nblocks, row_sizes, col_sizes, matrix_position, row_first, col_first, storage,
                         dof_indexes, color_indexes, ncolor, neq, weight, nullptr);
}

int ipf_transfer_geometry(ipf_ctx* ctx, int nshape, const int* shape_sizes, const double* shape_tables,
                          int64_t nblocks, const int* elem_shape, const double* corners, const int* dof_indexes,
                          const int* color_indexes, int64_t ncolor, int64_t neq) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (nblocks < 0 || ncolor < 1 || neq < 1 || nshape < 1 || nshape > SHAPE_COUNT)
        return fail(ctx, IPF_EINVAL, "bad sizes: nshape=%d nblocks=%lld ncolor=%lld neq=%lld", nshape,
                    (long long)nblocks, (long long)ncolor, (long long)neq);
    if (!shape_sizes || !shape_tables || (nblocks > 0 && (!elem_shape || !corners || !dof_indexes || !color_indexes)))
        return fail(ctx, IPF_EINVAL, "ipf_transfer_geometry: null array");
    // shapes -> kernel shapes, tables by kernel shape
    GeoInput geo;
    geo.corners = corners;
    std::vector<int> kshape(nshape);
    int64_t toff = 0;
    for (int i = 0; i < nshape; i++) {
        const int nip = shape_sizes[2 * i], nsh = shape_sizes[2 * i + 1];
        const int ks = shape_of(nip, 2 * nsh);
        if (ks < 0)
            return fail(ctx, IPF_EUNSUPPORTED, "shape %d (n_ip=%d, n_sh=%d) has no kernel (supported: Q1/Q2/Q3, T1/T2/T3)", i,
                        nip, nsh);
        if (geo.tables[ks]) return fail(ctx, IPF_EINVAL, "shapes %d and another have the same sizes", i);
        geo.tables[ks] = shape_tables + toff;
        kshape[i] = ks;
        toff += shape_table_size(nip, nsh);
    }
    // the blocks' sizes and prefix sums (TPZNumericalIntegrator.cpp:294-317)
    std::vector<int> rs(nblocks), cs(nblocks), rf(nblocks + 1, 0), cf(nblocks + 1, 0);
    std::vector<int64_t> mp(nblocks + 1, 0);
    for (int64_t e = 0; e < nblocks; e++) {
        const int i = elem_shape[e];
        if (i < 0 || i >= nshape) return fail(ctx, IPF_EINVAL, "element %lld: shape %d not in [0,%d)", (long long)e, i, nshape);
        rs[e] = 3 * shape_sizes[2 * i];
        cs[e] = 2 * shape_sizes[2 * i + 1];
        if ((int64_t)rf[e] + rs[e] > INT32_MAX || (int64_t)cf[e] + cs[e] > INT32_MAX)
            return fail(ctx, IPF_EINVAL, "more than 2^31 rows or columns");
        rf[e + 1] = rf[e] + rs[e];
        cf[e + 1] = cf[e] + cs[e];
        mp[e + 1] = mp[e] + (int64_t)rs[e] * cs[e];
    }
    for (int64_t e = 0; e < nblocks; e++)
        for (int k = 0; k < 8; k++)
            if (!std::isfinite(corners[8 * e + k]))
                return fail(ctx, IPF_EINVAL, "element %lld: corner coordinate not finite", (long long)e);
    return transfer_impl(ctx, nblocks, rs.data(), cs.data(), mp.data(), rf.data(), cf.data(), nullptr, dof_indexes,
                         color_indexes, ncolor, neq, nullptr, &geo);
}

int ipf_set_material_table(ipf_ctx* ctx, int nmat, const double* params) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (nmat < 1 || nmat > (1 << 23) || !params) return fail(ctx, IPF_EINVAL, "ipf_set_material_table: nmat %d", nmat);
    std::vector<MatConst> mats(nmat);
    for (int i = 0; i < nmat; i++) {
        const double* q = params + 5 * i;
        const double lambda = q[0], mu = q[1], phi = q[2], psi = q[3], c = q[4];
        if (!(mu > 0) || !std::isfinite(lambda) || !std::isfinite(phi) || !std::isfinite(psi) || !std::isfinite(c))
            return fail(ctx, IPF_EINVAL, "bad material %d: lambda=%g mu=%g phi=%g psi=%g c=%g", i, lambda, mu, phi,
                        psi, c);
        mats[i] = MatConst::make(lambda, mu, phi, psi, c);
    }
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->mats.size() != mats.size()) {
        drop_cg_graph(ctx);  // its launches carry the table's address
        dev_free(ctx->dMats);
        int rc = dev_alloc(ctx, &ctx->dMats, (int64_t)nmat);
        if (rc) {
            ctx->mats.clear();
            return rc;
        }
    }
    IPF_HIP(ctx, hipMemcpy(ctx->dMats, mats.data(), sizeof(MatConst) * nmat, hipMemcpyHostToDevice));
    ctx->mats = std::move(mats);
    return IPF_OK;
}

int ipf_set_material(ipf_ctx* ctx, double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion) {
    const double q[5] = {lambda, mu, mc_phi, mc_psi, mc_cohesion};
    return ipf_set_material_table(ctx, 1, q);
}

int ipf_set_element_materials(ipf_ctx* ctx, int64_t nblocks, const int* element_material) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (nblocks < 0 || (nblocks > 0 && !element_material))
        return fail(ctx, IPF_EINVAL, "ipf_set_element_materials: bad arguments");
    std::vector<int> v(element_material, element_material + nblocks);
    for (int64_t e = 0; e < nblocks; e++)
        if (v[e] < 0 || v[e] >= (1 << 23))
            return fail(ctx, IPF_EINVAL, "element %lld: material id %d outside [0, 2^23)", (long long)e, v[e]);
    ctx->elem_mat_next = std::move(v);
    return IPF_OK;
}

int ipf_set_element_order_keys(ipf_ctx* ctx, int64_t nblocks, const int* keys) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (nblocks < 0 || (nblocks > 0 && !keys)) return fail(ctx, IPF_EINVAL, "ipf_set_element_order_keys: bad arguments");
    ctx->elem_key_next.assign(keys, keys + nblocks);
    return IPF_OK;
}

int ipf_set_rhs_linear(ipf_ctx* ctx, const double* f_lin_host) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!f_lin_host) return fail(ctx, IPF_EINVAL, "ipf_set_rhs_linear: NULL");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipMemcpyAsync(ctx->dFlin, f_lin_host, sizeof(double) * ctx->neq, hipMemcpyHostToDevice, ctx->stream));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPF_OK;
}

int ipf_set_residual_assembly(ipf_ctx* ctx, int assembly) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_residual_assembly: NULL context");
    if (assembly != IPF_ASSEMBLY_COLORED && assembly != IPF_ASSEMBLY_PATCH)
        return fail(ctx, IPF_EINVAL, "ipf_set_residual_assembly: unknown assembly %d", assembly);
    if (assembly == IPF_ASSEMBLY_PATCH) {
        int rc = check_ready(ctx, false);
        if (rc) return rc;
        IPF_HIP(ctx, hipSetDevice(ctx->device));
        if ((rc = ensure_patch(ctx))) return rc;
    }
    ctx->assembly = assembly;
    return IPF_OK;
}

int ipf_set_patch_waves(ipf_ctx* ctx, int waves) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_patch_waves: NULL context");
    if (!patch_waves_ok(waves)) return fail(ctx, IPF_EINVAL, "ipf_set_patch_waves: %d is not 2, 4 or 8", waves);
    if (waves == ctx->patch_nw) return IPF_OK;
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    release_patch(ctx);
    ctx->patch_nw = waves;
    // the selected patch assembly is rebuilt now, so that its errors show here
    return ctx->assembly == IPF_ASSEMBLY_PATCH ? ensure_patch(ctx) : IPF_OK;
}

int ipf_get_patch_waves(const ipf_ctx* ctx, int* waves) {
    if (!ctx || !waves) return fail(nullptr, IPF_EINVAL, "ipf_get_patch_waves: NULL argument");
    *waves = ctx->patch_nw;
    return IPF_OK;
}

int ipf_autotune_assembly(ipf_ctx* ctx, const double* u_dev, int reps, double* times_ms) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev) return fail(ctx, IPF_EINVAL, "ipf_autotune_assembly: u is NULL");
    if (reps < 1 || reps > 10000)
        return fail(ctx, IPF_EINVAL, "ipf_autotune_assembly: reps %d outside [1, 10000]", reps);
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (times_ms)
        for (int k = 0; k < 4; k++) times_ms[k] = -1.0;
    const bool commit = ctx->update_mem;
    ctx->update_mem = false;  // non-committing residuals only: the committed state must not change
    const int a0 = ctx->assembly, nw0 = ctx->patch_nw;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        ctx->update_mem = commit;
        return fail(ctx, IPF_EHIP, "ipf_autotune_assembly: event creation failed");
    }
    auto time_it = [&](double* ms) -> int {
        int r = IPF_OK;
        for (int k = 0; k < 2 && r == IPF_OK; k++) r = ipf_residual(ctx, u_dev, ctx->dR);
        if (r) return r;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) return fail(ctx, IPF_EHIP, "autotune: event record");
        for (int k = 0; k < reps && r == IPF_OK; k++) r = ipf_residual(ctx, u_dev, ctx->dR);
        if (r) return r;
        float t = 0.f;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            return fail(ctx, IPF_EHIP, "autotune: event timing");
        *ms = (double)t / reps;
        return IPF_OK;
    };
    // candidates: colored (one wavefront tile per workgroup), then patches of
    // 2, 4 and 8 wavefront tiles where the setup allows patches
    int best_a = IPF_ASSEMBLY_COLORED, best_nw = nw0;
    double best = 1e300, ms = 0.0;
    ctx->assembly = IPF_ASSEMBLY_COLORED;
    rc = time_it(&ms);
    if (rc == IPF_OK) {
        if (times_ms) times_ms[0] = ms;
        best = ms;
    }
    const int cand[3] = {2, 4, 8};
    for (int k = 0; k < 3 && rc == IPF_OK && !ctx->patch_in.empty(); k++) {
        if (cand[k] != ctx->patch_nw) {
            release_patch(ctx);
            ctx->patch_nw = cand[k];
        }
        ctx->assembly = IPF_ASSEMBLY_PATCH;
        if ((rc = ensure_patch(ctx))) break;
        if ((rc = time_it(&ms))) break;
        if (times_ms) times_ms[1 + k] = ms;
        if (ms < best) best = ms, best_a = IPF_ASSEMBLY_PATCH, best_nw = cand[k];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->update_mem = commit;
    const int fa = rc ? a0 : best_a, fnw = rc ? nw0 : best_nw;
    // keep the chosen layout built; a patch layout nobody selected goes
    if (ctx->patch_nw != fnw || fa != IPF_ASSEMBLY_PATCH) {
        (void)hipStreamSynchronize(ctx->stream);
        release_patch(ctx);
    }
    ctx->patch_nw = fnw;
    ctx->assembly = IPF_ASSEMBLY_COLORED;
    if (fa == IPF_ASSEMBLY_PATCH) {
        const int rb = ensure_patch(ctx);
        if (rb == IPF_OK) ctx->assembly = IPF_ASSEMBLY_PATCH;
        if (!rc) rc = rb;
    }
    return rc;
}

int ipf_get_residual_assembly(const ipf_ctx* ctx, int* assembly) {
    if (!ctx || !assembly) return fail(nullptr, IPF_EINVAL, "ipf_get_residual_assembly: NULL argument");
    *assembly = ctx->assembly;
    return IPF_OK;
}

int ipf_patch_info(const ipf_ctx* ctx, int64_t out[8]) {
    if (!ctx || !out) return fail(nullptr, IPF_EINVAL, "ipf_patch_info: NULL argument");
    const PatchHost& h = ctx->patch;
    int64_t np = 0;
    for (const auto& R : h.ranges) np += R.npatch;
    out[0] = ctx->patch_built ? 1 : 0;
    out[1] = np;
    out[2] = (int64_t)h.pnode.size();
    out[3] = h.npartial;
    out[4] = (int64_t)h.fix_dof.size();
    out[5] = (int64_t)h.orphan.size();
    out[6] = h.npts;
    out[7] = (int64_t)h.ranges.size();
    return IPF_OK;
}

int ipf_patch_plan_host(int64_t n, const int* shape, const int* mat, const int64_t* first_point, const double* cx,
                        const double* cy, const int* node_xdof, int64_t neq, int waves, int64_t* sizes,
                        int64_t* ranges,
                        int* tmeta, int64_t* slot_elem, unsigned short* lnode, int64_t* node_off, int* pnode,
                        int* pslot, int* cbeg, unsigned short* contrib, int64_t* pt_map, int* fix_dof, int* fix_ptr,
                        int* fix_slot, int* orphan) {
    if (n < 0 || !sizes || (n > 0 && (!shape || !mat || !first_point || !cx || !cy || !node_xdof)))
        return fail(nullptr, IPF_EINVAL, "ipf_patch_plan_host: bad arguments");
    std::vector<PatchElemIn> el(n);
    std::vector<int> nodes;
    for (int64_t e = 0; e < n; e++) {
        if (shape[e] < 0 || shape[e] >= SHAPE_COUNT) return fail(nullptr, IPF_EINVAL, "ipf_patch_plan_host: shape");
        el[e].shape = shape[e];
        el[e].mat = mat[e];
        el[e].geo_slot = e;
        el[e].P0 = first_point[e];
        el[e].cx = cx[e];
        el[e].cy = cy[e];
        el[e].node0 = (int64_t)nodes.size();
        for (int a = 0; a < shape_ce(shape[e]) / 2; a++) nodes.push_back(node_xdof[el[e].node0 + a]);
    }
    PatchHost h;
    const std::string err = build_patches(el, nodes, neq, waves, h);
    if (!err.empty()) return fail(nullptr, IPF_EUNSUPPORTED, "%s", err.c_str());
    int64_t np = 0;
    for (const auto& R : h.ranges) np += R.npatch;
    const int64_t sz[12] = {(int64_t)h.ranges.size(), np, h.ntiles, h.nslots, (int64_t)h.lnode.size(),
                            (int64_t)h.pnode.size(), (int64_t)h.contrib.size(), h.npts, (int64_t)h.fix_dof.size(),
                            (int64_t)h.fix_slot.size(), (int64_t)h.orphan.size(), h.npartial};
    std::memcpy(sizes, sz, sizeof(sz));
    auto cp = [](auto* dst, const auto& v) {
        if (dst) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
    };
    if (ranges)
        for (size_t i = 0; i < h.ranges.size(); i++) {
            const auto& R = h.ranges[i];
            const int64_t v[7] = {R.shape, R.npatch, R.patch0, R.tile0, R.slot0, R.pt0, R.lnode0};
            std::memcpy(ranges + 7 * i, v, sizeof(v));
        }
    cp(tmeta, h.tmeta);
    cp(slot_elem, h.slot_geo);
    cp(lnode, h.lnode);
    cp(node_off, h.node_off);
    cp(pnode, h.pnode);
    cp(pslot, h.pslot);
    cp(cbeg, h.cbeg);
    cp(contrib, h.contrib);
    cp(pt_map, h.pt_map);
    cp(fix_dof, h.fix_dof);
    cp(fix_ptr, h.fix_ptr);
    cp(fix_slot, h.fix_slot);
    cp(orphan, h.orphan);
    return IPF_OK;
}

int ipf_certify_elastic_host(int64_t n, const double* strain3, const double* plastic_strain4, double lambda,
                             double mu, double mc_phi, double mc_psi, double mc_cohesion, int* certified,
                             double* sigma3) {
    if (n < 0 || (n > 0 && (!strain3 || !plastic_strain4 || !certified)))
        return fail(nullptr, IPF_EINVAL, "ipf_certify_elastic_host: bad arguments");
    const MatConst m = MatConst::make(lambda, mu, mc_phi, mc_psi, mc_cohesion);
    for (int64_t i = 0; i < n; i++) {
        double ws[3] = {0.0, 0.0, 0.0};
        certified[i] = certified_elastic(strain3 + 3 * i, plastic_strain4 + 4 * i, m, 1.0, ws) ? 1 : 0;
        if (sigma3)
            for (int k = 0; k < 3; k++) sigma3[3 * i + k] = ws[k];
    }
    return IPF_OK;
}

int ipf_set_residual_mode(ipf_ctx* ctx, int mode) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_residual_mode: NULL context");
    if (mode != IPF_RESIDUAL_EXACT && mode != IPF_RESIDUAL_CERTIFIED_ELASTIC)
        return fail(ctx, IPF_EINVAL, "ipf_set_residual_mode: unknown mode %d", mode);
    ctx->residual_mode = mode;
    return IPF_OK;
}

int ipf_get_residual_mode(const ipf_ctx* ctx, int* mode) {
    if (!ctx || !mode) return fail(nullptr, IPF_EINVAL, "ipf_get_residual_mode: NULL argument");
    *mode = ctx->residual_mode;
    return IPF_OK;
}

int ipf_set_newton_monitor(ipf_ctx* ctx, ipf_newton_monitor fn, void* user) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_newton_monitor: NULL context");
    ctx->monitor = fn;
    ctx->monitor_user = user;
    return IPF_OK;
}

int ipf_set_newton_line_search(ipf_ctx* ctx, int max_halvings) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_newton_line_search: NULL context");
    if (max_halvings < 0 || max_halvings > 30)
        return fail(ctx, IPF_EINVAL, "ipf_set_newton_line_search: max_halvings %d outside [0, 30]", max_halvings);
    ctx->ls_halvings = max_halvings;
    return IPF_OK;
}

int ipf_set_newton_atol(ipf_ctx* ctx, double atol) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_newton_atol: NULL context");
    if (!(atol >= 0.0)) return fail(ctx, IPF_EINVAL, "ipf_set_newton_atol: atol must be >= 0");
    ctx->newton_atol = atol;
    return IPF_OK;
}

int ipf_set_newton_thomas(ipf_ctx* ctx, int enable) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_newton_thomas: NULL context");
    ctx->newton_thomas = enable != 0;
    return IPF_OK;
}

int ipf_set_newton_reference_stop(ipf_ctx* ctx, double tolerance) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "ipf_set_newton_reference_stop: NULL context");
    if (!(tolerance >= 0.0) || tolerance > 1.7976931348623157e308)
        return fail(ctx, IPF_EINVAL, "ipf_set_newton_reference_stop: tolerance must be finite and >= 0");
    ctx->newton_ref_tol = tolerance;
    return IPF_OK;
}

int ipf_set_update_mem(ipf_ctx* ctx, int update_mem) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    ctx->update_mem = update_mem != 0;
    return IPF_OK;
}

int ipf_residual(ipf_ctx* ctx, const double* u_dev, double* r_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev || !r_dev) return fail(ctx, IPF_EINVAL, "ipf_residual: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (ctx->assembly == IPF_ASSEMBLY_PATCH && !ctx->update_mem) return patch_residual(ctx, u_dev, r_dev);
    // r = f_lin, then each color adds -B^T(w sigma)
    const double* flin = residual_start(ctx, r_dev);
    const PointState st = point_state(ctx);
    for (const auto& g : ctx->groups)
        launch_group_residual(ctx, g, u_dev, r_dev, flin, st);
    return launch_status(ctx, "residual");
}

int ipf_residual_timed(ipf_ctx* ctx, const double* u_dev, double* r_dev, double* launch_ms, int64_t* launch_pts,
                       int max_launches) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev || !r_dev || !launch_ms) return fail(ctx, IPF_EINVAL, "ipf_residual_timed: NULL pointer");
    if (ctx->assembly == IPF_ASSEMBLY_PATCH && !ctx->update_mem) {
        IPF_HIP(ctx, hipSetDevice(ctx->device));
        if ((rc = ensure_patch(ctx))) return rc;
        const PatchHost& h = ctx->patch;
        const int nr = (int)h.ranges.size(), np = nr + 2;
        if (max_launches < np) return fail(ctx, IPF_EINVAL, "ipf_residual_timed: need room for %d launches", np);
        std::vector<hipEvent_t> ev(np + 1, nullptr);
        for (auto& e : ev) IPF_HIP(ctx, hipEventCreate(&e));
        const bool refresh = ctx->patch_state_version != ctx->state_version;  // launch 0 runs
        rc = patch_residual(ctx, u_dev, r_dev, ev.data());
        if (rc == IPF_OK) {
            IPF_HIP(ctx, hipEventSynchronize(ev[np]));
            for (int i = 0; i < np; i++) {
                float ms = 0.f;
                IPF_HIP(ctx, hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                launch_ms[i] = (i == 0 && !refresh) ? 0.0 : ms;
                if (launch_pts) {
                    int64_t pts = 0;
                    if (i >= 1 && i <= nr)
                        for (const auto& E : ctx->patch_in)
                            if (E.shape == h.ranges[i - 1].shape) pts += shape_nip(E.shape);
                    launch_pts[i] = pts;
                }
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        return rc == IPF_OK ? np : rc;
    }
    const int n = 1 + (int)ctx->groups.size();
    if (max_launches < n) return fail(ctx, IPF_EINVAL, "ipf_residual_timed: need room for %d launches", n);
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<hipEvent_t> ev(n + 1, nullptr);
    for (auto& e : ev) IPF_HIP(ctx, hipEventCreate(&e));
    const PointState st = point_state(ctx);
    IPF_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
    const bool copied = !ctx->fold_flin;
    const double* flin = residual_start(ctx, r_dev);  // launch 0: the copy, or nothing when folded
    IPF_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    for (size_t i = 0; i < ctx->groups.size(); i++) {
        const auto& g = ctx->groups[i];
        launch_group_residual(ctx, g, u_dev, r_dev, flin, st);
        IPF_HIP(ctx, hipEventRecord(ev[i + 2], ctx->stream));
    }
    rc = launch_status(ctx, "residual");
    if (rc == IPF_OK) {
        IPF_HIP(ctx, hipEventSynchronize(ev[n]));
        for (int i = 0; i < n; i++) {
            float ms = 0.f;
            IPF_HIP(ctx, hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            launch_ms[i] = (i == 0 && !copied) ? 0.0 : ms;
            if (launch_pts)
                launch_pts[i] = i == 0 ? 0
                                       : ctx->groups[i - 1].nreal * shape_nip(ctx->groups[i - 1].shape);
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc == IPF_OK ? n : rc;
}

int ipf_residual_host(ipf_ctx* ctx, const double* u_host, double* r_host) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_host || !r_host) return fail(ctx, IPF_EINVAL, "ipf_residual_host: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipMemcpyAsync(ctx->dU, u_host, sizeof(double) * ctx->neq, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = ipf_residual(ctx, ctx->dU, ctx->dR))) return rc;
    IPF_HIP(ctx, hipMemcpyAsync(r_host, ctx->dR, sizeof(double) * ctx->neq, hipMemcpyDeviceToHost, ctx->stream));
    return ipf_synchronize(ctx);
}

int ipf_multiply(ipf_ctx* ctx, const double* u_dev, double* strain_dev) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!u_dev || !strain_dev) return fail(ctx, IPF_EINVAL, "ipf_multiply: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    for (const auto& g : ctx->groups)
        launch_multiply(g.shape, group_dev(ctx, g), u_dev, strain_dev, ctx->dOrig, ctx->stream);
    return launch_status(ctx, "multiply");
}

int ipf_multiply_vector(ipf_ctx* ctx, const double* x_dev, double* y_dev, int opt) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!x_dev || !y_dev) return fail(ctx, IPF_EINVAL, "ipf_multiply_vector: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (opt == 0) {
        for (const auto& g : ctx->groups)
            launch_multiply(g.shape, group_dev(ctx, g, true), x_dev, y_dev, ctx->dOrig, ctx->stream);
    } else {
        launch_fill(ctx->cols, y_dev, 0.0, ctx->stream);
        for (const auto& g : ctx->groups)
            launch_transpose(g.shape, group_dev(ctx, g, true), x_dev, y_dev, ctx->dOrig, ctx->stream);
    }
    return launch_status(ctx, "multiply_vector");
}

int ipf_compute_sigma(ipf_ctx* ctx, const double* strain_dev, double* wsigma_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!strain_dev || !wsigma_dev) return fail(ctx, IPF_EINVAL, "ipf_compute_sigma: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    for (const auto& g : ctx->groups)
        launch_sigma(false, ctx->update_mem, g.shape, group_dev(ctx, g), ctx->npad, strain_dev, wsigma_dev, nullptr,
                     ctx->dOrig, point_state(ctx), ctx->stream);
    return launch_status(ctx, "compute_sigma");
}

int ipf_point_features(ipf_ctx* ctx, const double* u_dev, float* feat_dev, int* branch_dev, float* target_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev || !feat_dev || !branch_dev || !target_dev)
        return fail(ctx, IPF_EINVAL, "ipf_point_features: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_solver_buffers(ctx))) return rc;
    for (const auto& g : ctx->groups)
        launch_multiply(g.shape, group_dev(ctx, g), u_dev, ctx->dStrain, ctx->dOrig, ctx->stream);
    for (const auto& g : ctx->groups)
        launch_point_features(g.shape, group_dev(ctx, g), ctx->dStrain, ctx->dOrig, point_state(ctx), feat_dev,
                              branch_dev, target_dev, ctx->stream);
    return launch_status(ctx, "point_features");
}

int ipf_branch_order(ipf_ctx* ctx, const int* branch_dev, int* order_dev) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!branch_dev || !order_dev) return fail(ctx, IPF_EINVAL, "ipf_branch_order: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->dHist && (rc = dev_alloc(ctx, &ctx->dHist, (int64_t)key_sort_blocks(ctx->npts) * IPF_NBRANCH)))
        return rc;
    launch_branch_order(branch_dev, ctx->npts, ctx->dHist, order_dev, ctx->stream);
    return launch_status(ctx, "branch_order");
}

int ipf_branch_order_noalloc(ipf_ctx* ctx, const int* branch_dev, int* order_dev) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!branch_dev || !order_dev) return fail(ctx, IPF_EINVAL, "ipf_branch_order_noalloc: NULL pointer");
    const int64_t need = (int64_t)key_sort_blocks(ctx->npts) * IPF_NBRANCH;
    if (need > kScratchInts)
        return fail(ctx, IPF_EINVAL, "ipf_branch_order_noalloc: %lld points exceed its limit of %lld",
                    (long long)ctx->npts, (long long)ipf_branch_order_noalloc_max_points());
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    int* hist = scratch_of(ctx->device);
    if (!hist) return fail(ctx, IPF_ESTATE, "ipf_branch_order_noalloc: no scratch address for device %d", ctx->device);
    // every histogram entry the sort reads is written by its first kernel:
    // nothing in the scratch needs clearing
    launch_branch_order(branch_dev, ctx->npts, hist, order_dev, ctx->stream);
    return launch_status(ctx, "branch_order_noalloc");
}

int64_t ipf_branch_order_noalloc_max_points(void) { return (kScratchInts / IPF_NBRANCH) * IPF_SORT_BLOCK; }

int ipf_compute_sigma_ordered(ipf_ctx* ctx, const double* strain_dev, double* wsigma_dev, const int* order_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!strain_dev || !wsigma_dev || !order_dev) return fail(ctx, IPF_EINVAL, "ipf_compute_sigma_ordered: NULL");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    launch_sigma_order(ctx->npts, order_dev, ctx->dNewOf, ctx->dPmat, ctx->dMats, strain_dev, wsigma_dev,
                       point_state(ctx), ctx->update_mem, ctx->stream);
    return launch_status(ctx, "compute_sigma_ordered");
}

int ipf_compute_sigma_dep(ipf_ctx* ctx, const double* strain_dev, double* wsigma_dev, double* wdep_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!strain_dev || !wsigma_dev || !wdep_dev) return fail(ctx, IPF_EINVAL, "ipf_compute_sigma_dep: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    for (const auto& g : ctx->groups)
        launch_sigma(true, ctx->update_mem, g.shape, group_dev(ctx, g), ctx->npad, strain_dev, wsigma_dev, wdep_dev,
                     ctx->dOrig, point_state(ctx), ctx->stream);
    return launch_status(ctx, "compute_sigma_dep");
}

int ipf_multiply_transpose(ipf_ctx* ctx, const double* wsigma_dev, double* res_dev) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!wsigma_dev || !res_dev) return fail(ctx, IPF_EINVAL, "ipf_multiply_transpose: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    launch_fill(ctx->neq, res_dev, 0.0, ctx->stream);
    for (const auto& g : ctx->groups)
        launch_transpose(g.shape, group_dev(ctx, g), wsigma_dev, res_dev, ctx->dOrig, ctx->stream);
    return launch_status(ctx, "multiply_transpose");
}

int ipf_get_b_format(const ipf_ctx* ctx, int* residual_format, int* tangent_format) {
    if (!ctx || !residual_format || !tangent_format) return fail(nullptr, IPF_EINVAL, "ipf_get_b_format: null argument");
    *residual_format = ctx->rfmt;
    *tangent_format = ctx->tfmt;
    return IPF_OK;
}

int ipf_set_b_format(ipf_ctx* ctx, int residual_format, int tangent_format) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_format(ctx, residual_format)) || (rc = ensure_format(ctx, tangent_format))) return rc;
    ctx->rfmt = residual_format;
    ctx->tfmt = tangent_format;
    drop_cg_graph(ctx);
    if ((rc = ensure_dep32(ctx))) return rc;
    return IPF_OK;
}

int ipf_autotune(ipf_ctx* ctx, const double* u_dev, int reps, int allow_fp32, double* times_ms) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev) return fail(ctx, IPF_EINVAL, "ipf_autotune: u is NULL");
    if (reps < 1 || reps > 10000) return fail(ctx, IPF_EINVAL, "ipf_autotune: reps %d outside [1, 10000]", reps);
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (times_ms)
        for (int k = 0; k < 2 * FMT_COUNT; k++) times_ms[k] = -1.0;
    // candidates: the exact FP64 forms the data allows; FP32 forms only for
    // the tangent product and only on request
    auto allowed = [&](int f) {
        if (format_geometric(f)) return ctx->have_geometry;
        if (format_compact(f)) return ctx->compact_ok;
        return true;
    };
    hipEvent_t e0, e1;
    IPF_HIP(ctx, hipEventCreate(&e0));
    IPF_HIP(ctx, hipEventCreate(&e1));
    const bool commit = ctx->update_mem;
    ctx->update_mem = false;  // the committed state must not change
    const int r0 = ctx->rfmt, t0 = ctx->tfmt;
    auto time_it = [&](auto&& launch, double* ms) -> int {
        for (int k = 0; k < 2; k++) launch();
        IPF_HIP(ctx, hipEventRecord(e0, ctx->stream));
        for (int k = 0; k < reps; k++) launch();
        IPF_HIP(ctx, hipEventRecord(e1, ctx->stream));
        IPF_HIP(ctx, hipEventSynchronize(e1));
        float t = 0.f;
        IPF_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / reps;
        return launch_status(ctx, "autotune");
    };
    int best_r = r0, best_t = t0;
    double tr = 1e300, tt = 1e300;
    for (int f = 0; f < FMT_COUNT && rc == IPF_OK; f++) {
        const bool fp32 = format_bytes(f) == 4;
        if (!allowed(f)) continue;
        if ((rc = ensure_format(ctx, f))) break;
        double ms;
        if (!fp32) {
            ctx->rfmt = f;
            rc = time_it(
                [&] {
                    const double* flin = residual_start(ctx, ctx->dR);
                    for (const auto& g : ctx->groups)
                        launch_group_residual(ctx, g, u_dev, ctx->dR, flin, point_state(ctx));
                },
                &ms);
            if (rc) break;
            if (times_ms) times_ms[f] = ms;
            if (ms < tr) tr = ms, best_r = f;
        }
        if (ctx->have_tangent && (!fp32 || allow_fp32)) {
            ctx->tfmt = f;
            if ((rc = ensure_dep32(ctx))) break;
            rc = time_it(
                [&] {
                    tangent_product(ctx, u_dev, ctx->dR, ctx->stream);
                },
                &ms);
            if (rc) break;
            if (times_ms) times_ms[FMT_COUNT + f] = ms;
            if (ms < tt) tt = ms, best_t = f;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->update_mem = commit;
    ctx->rfmt = rc ? r0 : best_r;
    ctx->tfmt = rc ? t0 : best_t;
    drop_cg_graph(ctx);
    if (rc) return rc;
    // stored forms nobody uses any more go (dense FP64 stays: the separate steps read it)
    for (int f = 1; f < FMT_COUNT; f++)
        if (f != ctx->rfmt && f != ctx->tfmt && ctx->dBf[f]) {
            IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
            int64_t n = 0;
            for (const auto& g : ctx->groups)
                n = std::max(n, g.f_off[f] + g.ntiles * (int64_t)format_values(g.shape, f) * shape_tp(g.shape));
            ctx->bytes -= n * format_bytes(f);
            char* p = static_cast<char*>(ctx->dBf[f]);
            dev_free(p);
            ctx->dBf[f] = nullptr;
        }
    return IPF_OK;
}

int ipf_set_cg_graph(ipf_ctx* ctx, int enable) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_cg_graph(ctx);
    ctx->cg_graph = enable != 0;
    return IPF_OK;
}

int ipf_set_newton_graph(ipf_ctx* ctx, int enable) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    ctx->newton_graph = enable != 0;
    return IPF_OK;
}

int ipf_set_dirichlet_mask(ipf_ctx* ctx, const signed char* mask_host) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!mask_host) return fail(ctx, IPF_EINVAL, "ipf_set_dirichlet_mask: NULL");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->dMask && (rc = dev_alloc(ctx, &ctx->dMask, ctx->neq))) return rc;
    IPF_HIP(ctx, hipMemcpy(ctx->dMask, mask_host, ctx->neq, hipMemcpyHostToDevice));
    ctx->have_mask = true;
    return IPF_OK;
}

}  // extern "C"

namespace {

// The launches of ipf_residual_tangent on stream s: strains, stresses with
// w Dep, r = f_lin - B^T (w sigma).
void residual_tangent_launches(ipf_ctx* ctx, const double* u, double* r, hipStream_t s) {
    for (const auto& g : ctx->groups)
        launch_multiply(g.shape, group_dev(ctx, g), u, ctx->dStrain, ctx->dOrig, s);
    for (const auto& g : ctx->groups)
        launch_sigma(true, ctx->update_mem, g.shape, group_dev(ctx, g), ctx->npad, ctx->dStrain, ctx->dSigw, nullptr,
                     ctx->dOrig, point_state(ctx), s, ctx->dDep);
    launch_copy(ctx->neq, ctx->dFlin, r, s);
    for (const auto& g : ctx->groups)
        launch_transpose(g.shape, group_dev(ctx, g), ctx->dSigw, r, ctx->dOrig, s);
}

// The start of a CG solve of K x = b on stream s: the block-Jacobi
// preconditioner of the current tangent (each node's 2x2 diagonal block
// inverted where the DoFs pair up, else its diagonal), x = 0, r / z / p, r.z
// into scalar 0 and ||r||^2 into scalar bb_slot.
void cg_start(ipf_ctx* ctx, const double* b, double* x, int bb_slot, hipStream_t s) {
    const int64_t n = ctx->neq;
    const int nb = dot_blocks(n);
    double* offd = ctx->pairs_aligned ? ctx->dOffd : nullptr;
    launch_fill(n, ctx->dDiag, 0.0, s);
    if (offd) launch_fill(n, offd, 0.0, s);
    for (const auto& g : ctx->groups)
        launch_tangent_diag(g.shape, group_dev(ctx, g), ctx->dDiag, offd, ctx->dDep, ctx->npad, s);
    launch_block_jacobi(n, ctx->dMask, ctx->dDiag, offd, ctx->dPa, ctx->dPb, s);
    launch_fill(n, x, 0.0, s);
    launch_cg_init(n, ctx->dMask, b, ctx->dCr, ctx->dPa, ctx->dPb, ctx->dCz, ctx->dCp, s);
    launch_dot(n, ctx->dCr, ctx->dCz, ctx->dPartial, nb, ctx->dScal + 0, s);
    launch_dot(n, ctx->dCr, ctx->dCr, ctx->dPartial, nb, ctx->dScal + bb_slot, s);
}

// CG blocks from iteration *it (relative residual *rel there) on the
// context's stream: kCgBlock iterations per graph replay (a shorter last
// block up to maxit launched one by one), ||r|| read after each block, until
// it drops to rtol * bnorm (or turns NaN) or maxit is reached.
int cg_blocks(ipf_ctx* ctx, double* x, double rtol, int maxit, double bnorm, int* it, double* rel) {
    const int64_t n = ctx->neq;
    const int nb = dot_blocks(n);
    hipStream_t s = ctx->stream;
    int rc;
    while (*it < maxit) {
        const int blk = std::min(kCgBlock, maxit - *it);
        hipGraphExec_t exec = blk == kCgBlock ? cg_graph(ctx, x) : nullptr;
        if (exec)
            IPF_HIP(ctx, hipGraphLaunch(exec, s));
        else
            cg_iterations(ctx, x, blk, s);
        *it += blk;
        launch_dot(n, ctx->dCr, ctx->dCr, ctx->dPartial, nb, ctx->dScal + 3, s);
        double rr;
        if ((rc = launch_status(ctx, "cg")) || (rc = read_scalar(ctx, 3, &rr))) return rc;
        *rel = std::sqrt(rr) / bnorm;
        if (!(*rel > rtol)) break;  // also stops on NaN
    }
    return IPF_OK;
}

// One Newton iteration's device work at iterate u up to its first host
// read, captured on the context's capture stream: the launches of
// ipf_residual_tangent into r = dNr (and the FP32 copy of w Dep for the
// FP32 geometric tangent), ||r||^2 over the free DoFs (scalar 6), the CG
// start for K dNd = dNr with ||r||^2 in scalar 7, its first kCgBlock
// iterations and ||r_cg||^2 (scalar 3).  The same kernels in the same order
// as the eager iteration, so the iterates are the same bits.  nullptr when
// capture fails (the caller then runs the iteration eagerly).  Needs the
// solver buffers (and dDep32 for FMT_G32) allocated.
hipGraphExec_t newton_step_graph(ipf_ctx* ctx, const double* u) {
    if (!ctx->cap_stream && hipStreamCreate(&ctx->cap_stream) != hipSuccess) {
        ctx->cap_stream = nullptr;
        return nullptr;
    }
    hipStream_t s = ctx->cap_stream;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    residual_tangent_launches(ctx, u, ctx->dNr, s);
    if (ctx->tfmt == FMT_G32) launch_to_f32(9 * ctx->npad, ctx->dDep, ctx->dDep32, s);
    masked_norm_launch(ctx, ctx->dNr, s);
    cg_start(ctx, ctx->dNr, ctx->dNd, 7, s);
    cg_iterations(ctx, ctx->dNd, kCgBlock, s);
    launch_dot(ctx->neq, ctx->dCr, ctx->dCr, ctx->dPartial, dot_blocks(ctx->neq), ctx->dScal + 3, s);
    hipGraph_t graph = nullptr;
    const hipError_t e1 = hipStreamEndCapture(s, &graph);
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = (e1 == hipSuccess) ? hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) : e1;
    if (graph) (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return exec;
}

}  // namespace

extern "C" {

int ipf_residual_tangent(ipf_ctx* ctx, const double* u_dev, double* r_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev || !r_dev) return fail(ctx, IPF_EINVAL, "ipf_residual_tangent: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_solver_buffers(ctx))) return rc;
    residual_tangent_launches(ctx, u_dev, r_dev, ctx->stream);
    ctx->have_tangent = true;
    ctx->dep32_ok = false;
    if ((rc = ensure_dep32(ctx))) return rc;
    return launch_status(ctx, "residual_tangent");
}

int ipf_set_tangent_pattern(ipf_ctx* ctx, const int* el_color_index, const int64_t* color_l_sequence, int64_t nnz) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!el_color_index || !color_l_sequence) return fail(ctx, IPF_EINVAL, "ipf_set_tangent_pattern: NULL pointer");
    if (nnz < 1 || nnz > INT32_MAX)
        return fail(ctx, IPF_EINVAL, "ipf_set_tangent_pattern: nnz %lld outside [1, 2^31 - 1]", (long long)nnz);
    const int64_t n = ctx->nblocks;
    std::vector<int> grp(n, -1);
    std::vector<int64_t> pos(n, 0);
    int64_t total = 0;
    for (size_t gi = 0; gi < ctx->groups.size(); gi++) {
        Group& g = ctx->groups[gi];
        const int64_t ce = shape_ce(g.shape);
        g.l_off = total;
        total += (int64_t)g.elems.size() * (ce * (ce + 1) / 2);
        for (size_t k = 0; k < g.elems.size(); k++) {
            if (g.elems[k] < 0) continue;
            grp[g.elems[k]] = (int)gi;
            pos[g.elems[k]] = (int64_t)k;
        }
    }
    std::vector<int> L;
    std::vector<char> seen(n, 0);
    try {
        L.assign(total, 0);
    } catch (const std::bad_alloc&) {
        return fail(ctx, IPF_EINVAL, "ipf_set_tangent_pattern: host allocation failed");
    }
    int prev_color = -1;
    int64_t off = 0;
    for (int64_t i = 0; i < n; i++) {
        const int64_t e = el_color_index[i];
        if (e < 0 || e >= n || seen[e])
            return fail(ctx, IPF_EINVAL, "el_color_index[%lld] = %lld is not a permutation entry", (long long)i,
                        (long long)e);
        seen[e] = 1;
        const Group& g = ctx->groups[grp[e]];
        if (g.color < prev_color)
            return fail(ctx, IPF_EINVAL, "el_color_index is not in color order at position %lld", (long long)i);
        prev_color = g.color;
        const int64_t ce = shape_ce(g.shape);
        const int64_t nt = ce * (ce + 1) / 2;
        int* dst = L.data() + g.l_off + pos[e] * nt;
        for (int64_t t = 0; t < nt; t++) {
            const int64_t l = color_l_sequence[off + t];
            if (l < 0 || l >= nnz)
                return fail(ctx, IPF_EINVAL, "color_l_sequence[%lld] = %lld outside [0, %lld)", (long long)(off + t),
                            (long long)l, (long long)nnz);
            dst[t] = (int)l;
        }
        off += nt;
    }
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dev_free(ctx->dLidx);
    dev_free(ctx->dPatIA);
    dev_free(ctx->dPatJA);
    ctx->have_pattern = false;
    if ((rc = dev_alloc(ctx, &ctx->dLidx, total))) return rc;
    if (total) IPF_HIP(ctx, hipMemcpy(ctx->dLidx, L.data(), sizeof(int) * (size_t)total, hipMemcpyHostToDevice));
    ctx->nnz_k = nnz;
    ctx->have_pattern = true;
    return IPF_OK;
}

int ipf_build_tangent_pattern(ipf_ctx* ctx, int64_t* nnz) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!nnz) return fail(ctx, IPF_EINVAL, "ipf_build_tangent_pattern: NULL pointer");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipStream_t s = ctx->stream;
    int64_t total = 0;
    for (auto& g : ctx->groups) {
        const int64_t ce = shape_ce(g.shape);
        g.l_off = total;
        total += (int64_t)g.elems.size() * (ce * (ce + 1) / 2);
    }
    dev_free(ctx->dLidx);
    dev_free(ctx->dPatIA);
    dev_free(ctx->dPatJA);
    ctx->have_pattern = false;
    ctx->nnz_k = 0;
    if (total == 0) return fail(ctx, IPF_ESTATE, "ipf_build_tangent_pattern: no elements");
    uint64_t *keys = nullptr, *sorted = nullptr;
    unsigned int* count = nullptr;
    if ((rc = dev_alloc(ctx, &keys, total)) || (rc = dev_alloc(ctx, &sorted, total)) || (rc = dev_alloc(ctx, &count, 1))) {
        dev_free(keys);
        dev_free(sorted);
        return rc;
    }
    for (const auto& g : ctx->groups) launch_pattern_keys(g.shape, group_dev(ctx, g), ctx->neq, keys + g.l_off, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = pattern_sort_unique(keys, sorted, total, count, s);
    unsigned int nuniq = 0;
    if (e == hipSuccess) e = hipMemcpy(&nuniq, count, sizeof(nuniq), hipMemcpyDeviceToHost);
    dev_free(sorted);
    dev_free(count);
    uint64_t last = 0;
    if (e == hipSuccess && nuniq > 0) e = hipMemcpy(&last, keys + nuniq - 1, sizeof(last), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        dev_free(keys);
        return fail(ctx, IPF_EHIP, "ipf_build_tangent_pattern: %s", hipGetErrorString(e));
    }
    // the padding slots' key, if any, sorts last
    const int64_t n = (int64_t)nuniq - (nuniq > 0 && last == ~0ull ? 1 : 0);
    if (n < 1 || n > INT32_MAX) {
        dev_free(keys);
        return fail(ctx, IPF_EINVAL, "ipf_build_tangent_pattern: %lld pattern entries outside [1, 2^31 - 1]",
                    (long long)n);
    }
    if ((rc = dev_alloc(ctx, &ctx->dPatIA, ctx->neq + 1)) || (rc = dev_alloc(ctx, &ctx->dPatJA, n)) ||
        (rc = dev_alloc(ctx, &ctx->dLidx, total))) {
        dev_free(keys);
        dev_free(ctx->dPatIA);
        dev_free(ctx->dPatJA);
        return rc;
    }
    launch_pattern_rows(keys, n, ctx->neq, ctx->dPatIA, ctx->dPatJA, s);
    for (const auto& g : ctx->groups)
        launch_pattern_positions(g.shape, group_dev(ctx, g), ctx->dPatIA, ctx->dPatJA, ctx->dLidx + g.l_off, s);
    e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    dev_free(keys);
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(ctx, IPF_EHIP, "ipf_build_tangent_pattern: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    ctx->nnz_k = n;
    ctx->have_pattern = true;
    *nnz = n;
    return IPF_OK;
}

int ipf_get_tangent_pattern(ipf_ctx* ctx, int64_t* ia_host, int64_t* ja_host) {
    if (!ctx) return fail(nullptr, IPF_EINVAL, "null context");
    if (!ctx->dPatIA) return fail(ctx, IPF_ESTATE, "ipf_get_tangent_pattern: no pattern (ipf_build_tangent_pattern)");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ia_host) IPF_HIP(ctx, hipMemcpy(ia_host, ctx->dPatIA, sizeof(int64_t) * (ctx->neq + 1), hipMemcpyDeviceToHost));
    if (ja_host) IPF_HIP(ctx, hipMemcpy(ja_host, ctx->dPatJA, sizeof(int64_t) * ctx->nnz_k, hipMemcpyDeviceToHost));
    return IPF_OK;
}

int ipf_assemble_tangent(ipf_ctx* ctx, const double* u_dev, double* r_dev, double* kg_dev) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!kg_dev) return fail(ctx, IPF_EINVAL, "ipf_assemble_tangent: NULL pointer");
    if (!ctx->have_pattern) return fail(ctx, IPF_ESTATE, "no tangent pattern: call ipf_set_tangent_pattern first");
    if ((rc = ipf_residual_tangent(ctx, u_dev, r_dev))) return rc;
    launch_fill(ctx->nnz_k, kg_dev, 0.0, ctx->stream);
    for (const auto& g : ctx->groups)
        launch_assemble_K(g.shape, group_dev(ctx, g), ctx->dDep, ctx->npad, ctx->dLidx + g.l_off, kg_dev,
                          ctx->stream);
    return launch_status(ctx, "assemble_tangent");
}

int ipf_tangent_apply(ipf_ctx* ctx, const double* v_dev, double* y_dev) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!v_dev || !y_dev) return fail(ctx, IPF_EINVAL, "ipf_tangent_apply: NULL pointer");
    if (!ctx->have_tangent) return fail(ctx, IPF_ESTATE, "no tangent: call ipf_residual_tangent first");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    tangent_product(ctx, v_dev, y_dev, ctx->stream);
    return launch_status(ctx, "tangent_apply");
}

int ipf_cg_solve(ipf_ctx* ctx, const double* b_dev, double* x_dev, double rtol, int maxit, int* iters,
                 double* relres) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!b_dev || !x_dev) return fail(ctx, IPF_EINVAL, "ipf_cg_solve: NULL pointer");
    if (!ctx->have_tangent) return fail(ctx, IPF_ESTATE, "no tangent: call ipf_residual_tangent first");
    if (!(rtol > 0) || maxit < 1) return fail(ctx, IPF_EINVAL, "ipf_cg_solve: need rtol > 0 and maxit >= 1");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_solver_buffers(ctx))) return rc;
    if ((rc = ensure_dep32(ctx))) return rc;
    cg_start(ctx, b_dev, x_dev, 3, ctx->stream);
    double bb;
    if ((rc = read_scalar(ctx, 3, &bb))) return rc;
    const double bnorm = std::sqrt(bb);
    int it = 0;
    double rel = 0.0;
    // q = K p on the free DoFs, in blocks between convergence checks
    if (bnorm > 0.0 && (rc = cg_blocks(ctx, x_dev, rtol, maxit, bnorm, &it, &rel))) return rc;
    if (iters) *iters = it;
    if (relres) *relres = rel;
    return launch_status(ctx, "cg");
}

int ipf_newton(ipf_ctx* ctx, double* u_dev, double rtol, int maxit, double cg_rtol, int cg_maxit, double* stats) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev) return fail(ctx, IPF_EINVAL, "ipf_newton: NULL pointer");
    if (ctx->newton_thomas && ctx->ls_halvings > 0)
        return fail(ctx, IPF_EINVAL, "ipf_newton: the Thomas acceleration and the line search exclude each other");
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_solver_buffers(ctx))) return rc;
    if (ctx->newton_thomas && !ctx->dNt && (rc = dev_alloc(ctx, &ctx->dNt, ctx->neq))) return rc;
    // The caller's UpdateMem switch is restored on every exit: the Newton
    // residuals must not commit, only the final AcceptPseudoTimeStepSolution.
    const bool commit_flag = ctx->update_mem;
    ctx->update_mem = false;
    double r0 = 0.0, rn = 0.0, worst_cg = 0.0, du_norm = 0.0;
    double mt_alpha = 1.0;  // the Thomas acceleration factor, main.cpp:171
    int it = 0, cg_total = 0, ls_halved = 0;
    bool converged = false;
    // each iteration up to its first host read as one graph replay (when an
    // update can follow, the first CG block fits in cg_maxit and the capture
    // succeeds)
    hipGraphExec_t step = nullptr;
    if (ctx->newton_graph && maxit > 0 && cg_maxit >= kCgBlock && cg_rtol > 0) {
        if (ctx->tfmt == FMT_G32 && !ctx->dDep32) rc = dev_alloc(ctx, &ctx->dDep32, 9 * ctx->npad);
        if (!rc) step = newton_step_graph(ctx, u_dev);
    }
    for (it = 0; !rc && it <= maxit; it++) {
        double bb = 0.0, rr = 0.0;
        if (step) {
            const hipError_t e = hipGraphLaunch(step, ctx->stream);
            if (e != hipSuccess) {
                rc = fail(ctx, IPF_EHIP, "newton step graph launch failed: %s", hipGetErrorString(e));
                break;
            }
            ctx->have_tangent = true;
            ctx->dep32_ok = ctx->tfmt == FMT_G32;
            double v[8];
            if ((rc = launch_status(ctx, "newton step")) || (rc = read_scalars(ctx, 8, v))) break;
            rn = std::sqrt(v[6]);
            bb = v[7];
            rr = v[3];
        } else {
            if ((rc = ipf_residual_tangent(ctx, u_dev, ctx->dNr))) break;
            if ((rc = masked_norm(ctx, ctx->dNr, &rn))) break;
        }
        if (it == 0) r0 = rn;
        if (ctx->monitor) ctx->monitor(ctx->monitor_user, it, r0 > 0 ? rn / r0 : 0.0, rn, cg_total);
        const double ref_tol = ctx->newton_ref_tol;
        if (ref_tol > 0.0 ? (it > 0 && rn < ref_tol && du_norm < ref_tol)
                          : (rn <= rtol * r0 || rn <= ctx->newton_atol || rn == 0.0)) {
            converged = true;
            break;
        }
        if (it == maxit) break;
        int cgi = 0;
        double cgr = 0.0;
        if (step) {
            // the CG's start and first block ran in the graph: continue as
            // ipf_cg_solve does after its first block
            const double bnorm = std::sqrt(bb);
            if (bnorm > 0.0) {
                cgi = kCgBlock;
                cgr = std::sqrt(rr) / bnorm;
                if (cgr > cg_rtol && (rc = cg_blocks(ctx, ctx->dNd, cg_rtol, cg_maxit, bnorm, &cgi, &cgr))) break;
            }
            if ((rc = launch_status(ctx, "cg"))) break;
        } else if ((rc = ipf_cg_solve(ctx, ctx->dNr, ctx->dNd, cg_rtol, cg_maxit, &cgi, &cgr))) {
            break;
        }
        cg_total += cgi;
        worst_cg = std::max(worst_cg, cgr);
        if (ctx->newton_thomas) {
            // main.cpp:194-224: u += alpha du; the residual there with the
            // tangent unchanged; its correction du~ from the same tangent;
            // ||alpha du + du~||; alpha += (alpha du . du~) / |alpha du|^2;
            // u += du~.  du is kept in dNt while du~ is solved into dNd (the
            // CG graph stays keyed on dNd).
            hipStream_t s = ctx->stream;
            const int64_t n = ctx->neq;
            const int nb = dot_blocks(n);
            launch_axpy(n, mt_alpha, ctx->dNd, u_dev, s);
            launch_copy(n, ctx->dNd, ctx->dNt, s);
            if ((rc = ipf_residual(ctx, u_dev, ctx->dNr))) break;
            if ((rc = ipf_cg_solve(ctx, ctx->dNr, ctx->dNd, cg_rtol, cg_maxit, &cgi, &cgr))) break;
            cg_total += cgi;
            worst_cg = std::max(worst_cg, cgr);
            launch_copy(n, ctx->dNd, ctx->dCq, s);  // alpha du + du~ (q is free after the CG)
            launch_axpy(n, mt_alpha, ctx->dNt, ctx->dCq, s);
            launch_dot(n, ctx->dCq, ctx->dCq, ctx->dPartial, nb, ctx->dScal + 8, s);
            launch_dot(n, ctx->dNt, ctx->dNd, ctx->dPartial, nb, ctx->dScal + 9, s);
            launch_dot(n, ctx->dNt, ctx->dNt, ctx->dPartial, nb, ctx->dScal + 10, s);
            launch_axpy(n, 1.0, ctx->dNd, u_dev, s);
            double v[11];
            if ((rc = launch_status(ctx, "newton thomas")) || (rc = read_scalars(ctx, 11, v))) break;
            du_norm = std::sqrt(v[8]);
            const double num = mt_alpha * v[9], dem = mt_alpha * mt_alpha * v[10];
            if (dem > 0.0) mt_alpha += num / dem;  // (du = 0 only at r = 0, which stops first)
            continue;
        }
        if (ctx->newton_ref_tol > 0.0) {  // ||du|| of the full step (all DoFs, as main.cpp:227)
            launch_dot(ctx->neq, ctx->dNd, ctx->dNd, ctx->dPartial, dot_blocks(ctx->neq), ctx->dScal + 8,
                       ctx->stream);
            double dd;
            if ((rc = read_scalar(ctx, 8, &dd))) break;
            du_norm = std::sqrt(dd);
        }
        if (ctx->ls_halvings == 0) {
            launch_axpy(ctx->neq, 1.0, ctx->dNd, u_dev, ctx->stream);
            continue;
        }
        // backtracking: u = u0 + t du with t = 1, 1/2, ... until ||r|| drops
        // below its value at u0 (or the halvings run out: the last t stays)
        if (!ctx->dLs && (rc = dev_alloc(ctx, &ctx->dLs, ctx->neq))) break;
        launch_copy(ctx->neq, u_dev, ctx->dLs, ctx->stream);
        double t = 1.0, rt = 0.0;
        for (int h = 0;; h++) {
            launch_copy(ctx->neq, ctx->dLs, u_dev, ctx->stream);
            launch_axpy(ctx->neq, t, ctx->dNd, u_dev, ctx->stream);
            if ((rc = ipf_residual(ctx, u_dev, ctx->dNr)) || (rc = masked_norm(ctx, ctx->dNr, &rt))) break;
            if (rt < rn || h == ctx->ls_halvings) break;  // a NaN norm halves too
            t *= 0.5;
            ls_halved++;
        }
        if (rc) break;
        du_norm *= t;  // the step taken is t du (t a power of 2: exact)
    }
    if (step) (void)hipGraphExecDestroy(step);
    // AcceptPseudoTimeStepSolution: one residual with UpdateMem on
    if (!rc) {
        ctx->update_mem = true;
        rc = ipf_residual(ctx, u_dev, ctx->dNr);
    }
    ctx->update_mem = commit_flag;
    if (rc) return rc;
    if ((rc = ipf_synchronize(ctx))) return rc;
    if (stats) {
        stats[0] = it;
        stats[1] = converged ? 1.0 : 0.0;
        stats[2] = r0 > 0 ? rn / r0 : 0.0;
        stats[3] = r0;
        stats[4] = rn;
        stats[5] = cg_total;
        stats[6] = worst_cg;
        stats[7] = ls_halved;
    }
    return IPF_OK;
}

int ipf_newton_segmented(ipf_ctx* ctx, double* u_dev, int nseg, const int64_t* seg_first, double rtol, int maxit,
                         double cg_rtol, int cg_maxit, double* stats, double* seg_rel, int* seg_iters) {
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!u_dev || !seg_first) return fail(ctx, IPF_EINVAL, "ipf_newton_segmented: NULL pointer");
    if (nseg < 1) return fail(ctx, IPF_EINVAL, "ipf_newton_segmented: nseg %d < 1", nseg);
    if (seg_first[0] != 0 || seg_first[nseg] != ctx->neq)
        return fail(ctx, IPF_EINVAL, "ipf_newton_segmented: segments must cover [0, neq)");
    for (int k = 0; k < nseg; k++)
        if (seg_first[k + 1] <= seg_first[k])
            return fail(ctx, IPF_EINVAL, "ipf_newton_segmented: segment %d is empty or reversed", k);
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_solver_buffers(ctx))) return rc;
    // chunks of at most 2^16 equations inside each segment
    constexpr int64_t kChunk = 1 << 16;
    std::vector<int64_t> cb, ce;
    std::vector<int> sfc(nseg + 1, 0);
    for (int k = 0; k < nseg; k++) {
        sfc[k] = (int)cb.size();
        for (int64_t b = seg_first[k]; b < seg_first[k + 1]; b += kChunk) {
            cb.push_back(b);
            ce.push_back(std::min(b + kChunk, seg_first[k + 1]));
        }
    }
    sfc[nseg] = (int)cb.size();
    const int nchunk = (int)cb.size();
    int64_t* dcb = nullptr;
    int64_t* dce = nullptr;
    int* dsfc = nullptr;
    double* dpart = nullptr;
    double* dseg = nullptr;
    signed char* dmask0 = nullptr;
    auto cleanup = [&] {
        dev_free(dcb);
        dev_free(dce);
        dev_free(dsfc);
        dev_free(dpart);
        dev_free(dseg);
        dev_free(dmask0);
    };
    auto bail = [&](int code) {
        cleanup();
        return code;
    };
    const int64_t b0 = ctx->bytes;
    if ((rc = dev_alloc(ctx, &dcb, nchunk)) || (rc = dev_alloc(ctx, &dce, nchunk)) ||
        (rc = dev_alloc(ctx, &dsfc, nseg + 1)) || (rc = dev_alloc(ctx, &dpart, nchunk)) ||
        (rc = dev_alloc(ctx, &dseg, nseg)) || (rc = dev_alloc(ctx, &dmask0, ctx->neq)))
        return bail(rc);
    ctx->bytes = b0;  // workspace of this call only
    if (hipMemcpy(dcb, cb.data(), sizeof(int64_t) * nchunk, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dce, ce.data(), sizeof(int64_t) * nchunk, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dsfc, sfc.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpyAsync(dmask0, ctx->dMask, ctx->neq, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        return bail(fail(ctx, IPF_EHIP, "ipf_newton_segmented: workspace upload failed"));
    const bool commit_flag = ctx->update_mem;
    ctx->update_mem = false;
    std::vector<double> r0(nseg, 0.0), rn(nseg, 0.0);
    std::vector<char> done(nseg, 0);
    std::vector<int> its(nseg, -1);
    double worst_cg = 0.0;
    int it = 0, cg_total = 0;
    bool all = false;
    for (it = 0; it <= maxit; it++) {
        if ((rc = ipf_residual_tangent(ctx, u_dev, ctx->dNr))) break;
        // per-segment norms over the constrained DoFs of the caller's mask
        launch_segment_sumsq(ctx->dNr, dmask0, dcb, dce, nchunk, dpart, dsfc, nseg, dseg, ctx->stream);
        if (hipMemcpyAsync(rn.data(), dseg, sizeof(double) * nseg, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail(ctx, IPF_EHIP, "ipf_newton_segmented: norm read failed");
            break;
        }
        all = true;
        for (int k = 0; k < nseg; k++) {
            rn[k] = std::sqrt(rn[k]);
            if (it == 0) r0[k] = rn[k];
            if (!done[k] && (rn[k] <= rtol * r0[k] || rn[k] <= ctx->newton_atol || rn[k] == 0.0)) {
                done[k] = 1;
                its[k] = it;
                // hold the segment: its DoFs leave the CG and the update
                if (hipMemsetAsync(ctx->dMask + seg_first[k], 1, seg_first[k + 1] - seg_first[k], ctx->stream) !=
                    hipSuccess) {
                    rc = fail(ctx, IPF_EHIP, "ipf_newton_segmented: mask update failed");
                    break;
                }
            }
            all = all && done[k];
        }
        if (!rc && ctx->monitor) {
            double worst = 0.0, sn = 0.0;
            for (int k = 0; k < nseg; k++) {
                if (r0[k] > 0) worst = std::max(worst, rn[k] / r0[k]);
                sn += rn[k] * rn[k];
            }
            ctx->monitor(ctx->monitor_user, it, worst, std::sqrt(sn), cg_total);
        }
        if (rc || all || it == maxit) break;
        // One CG for all segments (the system is block diagonal), each active
        // segment's right-hand side scaled to unit norm and the tolerance
        // divided by sqrt(active segments): then every segment's own
        // relative residual is <= cg_rtol, as in its own Newton; with the
        // plain right-hand side the segments with large residuals would set
        // the stopping point and the others would get inaccurate updates.
        int nact = 0;
        for (int k = 0; k < nseg; k++)
            if (!done[k] && rn[k] > 0.0) {
                launch_scale(seg_first[k + 1] - seg_first[k], 1.0 / rn[k], ctx->dNr + seg_first[k], ctx->stream);
                nact++;
            }
        int cgi = 0;
        double cgr = 0.0;
        if ((rc = ipf_cg_solve(ctx, ctx->dNr, ctx->dNd, cg_rtol / std::sqrt((double)std::max(1, nact)), cg_maxit,
                               &cgi, &cgr)))
            break;
        for (int k = 0; k < nseg; k++)
            if (!done[k] && rn[k] > 0.0)
                launch_scale(seg_first[k + 1] - seg_first[k], rn[k], ctx->dNd + seg_first[k], ctx->stream);
        cg_total += cgi;
        worst_cg = std::max(worst_cg, cgr);
        launch_axpy(ctx->neq, 1.0, ctx->dNd, u_dev, ctx->stream);
    }
    // the caller's constraints back, then AcceptPseudoTimeStepSolution
    const hipError_t e = hipMemcpyAsync(ctx->dMask, dmask0, ctx->neq, hipMemcpyDeviceToDevice, ctx->stream);
    if (!rc && e != hipSuccess) rc = fail(ctx, IPF_EHIP, "ipf_newton_segmented: mask restore failed");
    if (!rc) {
        ctx->update_mem = true;
        rc = ipf_residual(ctx, u_dev, ctx->dNr);
    }
    ctx->update_mem = commit_flag;
    const int rs = ipf_synchronize(ctx);
    if (!rc) rc = rs;
    cleanup();
    if (rc) return rc;
    double worst = 0.0, s0 = 0.0, sn = 0.0;
    for (int k = 0; k < nseg; k++) {
        const double rel = r0[k] > 0 ? rn[k] / r0[k] : 0.0;
        worst = std::max(worst, rel);
        s0 += r0[k] * r0[k];
        sn += rn[k] * rn[k];
        if (seg_rel) seg_rel[k] = rel;
        if (seg_iters) seg_iters[k] = done[k] ? its[k] : -1;
    }
    if (stats) {
        stats[0] = it;
        stats[1] = all ? 1.0 : 0.0;
        stats[2] = worst;
        stats[3] = std::sqrt(s0);
        stats[4] = std::sqrt(sn);
        stats[5] = cg_total;
        stats[6] = worst_cg;
        stats[7] = 0.0;
    }
    return IPF_OK;
}

int ipf_get_state(ipf_ctx* ctx, double* plastic_strain_host, double* alpha_host, int* mtype_host, int* branch_host) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t n = ctx->npad;
    std::vector<int64_t> orig(n);
    IPF_HIP(ctx, hipMemcpy(orig.data(), ctx->dOrig, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    if (plastic_strain_host) {
        std::vector<double> ep(n);
        const int comp[4] = {XX, XY, YY, ZZ};
        for (int64_t o = 0; o < ctx->npts; o++) {
            plastic_strain_host[6 * o + XZ] = 0.0;
            plastic_strain_host[6 * o + YZ] = 0.0;
        }
        for (int c = 0; c < 4; c++) {
            IPF_HIP(ctx, hipMemcpy(ep.data(), ctx->dEp[c], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int64_t P = 0; P < n; P++)
                if (orig[P] >= 0) plastic_strain_host[6 * orig[P] + comp[c]] = ep[P];
        }
    }
    if (alpha_host) {
        std::vector<double> a(n);
        IPF_HIP(ctx, hipMemcpy(a.data(), ctx->dAlpha, sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int64_t P = 0; P < n; P++)
            if (orig[P] >= 0) alpha_host[orig[P]] = a[P];
    }
    std::vector<signed char> b(n);
    if (mtype_host) {
        IPF_HIP(ctx, hipMemcpy(b.data(), ctx->dMtype, n, hipMemcpyDeviceToHost));
        for (int64_t P = 0; P < n; P++)
            if (orig[P] >= 0) mtype_host[orig[P]] = b[P];
    }
    if (branch_host) {
        IPF_HIP(ctx, hipMemcpy(b.data(), ctx->dBranch, n, hipMemcpyDeviceToHost));
        for (int64_t P = 0; P < n; P++)
            if (orig[P] >= 0) branch_host[orig[P]] = b[P];
    }
    return IPF_OK;
}

int ipf_set_state(ipf_ctx* ctx, const double* plastic_strain_host, const double* alpha_host) {
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (!plastic_strain_host) return fail(ctx, IPF_EINVAL, "ipf_set_state: plastic strain is NULL");
    for (int64_t o = 0; o < ctx->npts; o++)
        if (plastic_strain_host[6 * o + XZ] != 0.0 || plastic_strain_host[6 * o + YZ] != 0.0)
            return fail(ctx, IPF_EINVAL, "point %lld: plastic strain XZ/YZ must be zero in plane strain",
                        (long long)o);
    IPF_HIP(ctx, hipSetDevice(ctx->device));
    IPF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t n = ctx->npad;
    std::vector<int64_t> orig(n);
    IPF_HIP(ctx, hipMemcpy(orig.data(), ctx->dOrig, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    std::vector<double> v(n);
    const int comp[4] = {XX, XY, YY, ZZ};
    for (int c = 0; c < 4; c++) {
        for (int64_t P = 0; P < n; P++) v[P] = orig[P] >= 0 ? plastic_strain_host[6 * orig[P] + comp[c]] : 0.0;
        IPF_HIP(ctx, hipMemcpy(ctx->dEp[c], v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    for (int64_t P = 0; P < n; P++) v[P] = (alpha_host && orig[P] >= 0) ? alpha_host[orig[P]] : 0.0;
    IPF_HIP(ctx, hipMemcpy(ctx->dAlpha, v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    ctx->state_version++;
    return IPF_OK;
}

}  // extern "C"
