// ipf_frontend.cpp -- C ABI of include/intpointsfem_frontend.h over the C++
// classes TPZCompMesh and TPZElastoPlasticIntPointsStructMatrix.  Exceptions
// are caught here and turned into status codes.
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/intpointsfem_frontend.h"
#include "TPZCompMesh.h"
#include "TPZElastoPlasticIntPointsStructMatrix.h"

// Last-error text shared with the device layer (ipf_device.hip).
extern "C" int ipf_internal_set_error(const char* msg);

struct ipf_cmesh {
    std::unique_ptr<TPZCompMesh> mesh;
};

struct ipf_sm {
    ipf_cmesh* mesh;
    std::unique_ptr<TPZElastoPlasticIntPointsStructMatrix> sm;
    std::unique_ptr<TPZSYsmpMatrix<STATE>> K;  // from Create(), on demand
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        ipf_internal_set_error("out of host memory");
        return IPF_EINVAL;
    } catch (const std::invalid_argument& e) {
        ipf_internal_set_error(e.what());
        return IPF_EINVAL;
    } catch (const std::exception& e) {
        ipf_internal_set_error(e.what());
        return IPF_EHIP;
    }
}

int null_arg(const char* fn) {
    ipf_internal_set_error((std::string(fn) + ": NULL argument").c_str());
    return IPF_EINVAL;
}

int fail_status(int code, const char* msg) {
    ipf_internal_set_error(msg);
    return code;
}

}  // namespace

extern "C" {

int ipf_cmesh_quarter_annulus(int nr, int ntheta, int porder, double r_in, double r_out, double ratio,
                              ipf_cmesh** out) {
    if (!out) return null_arg("ipf_cmesh_quarter_annulus");
    *out = nullptr;
    return guarded([&] {
        auto m = std::make_unique<ipf_cmesh>();
        m->mesh = std::make_unique<TPZCompMesh>(nr, ntheta, porder, r_in, r_out, ratio);
        *out = m.release();
        return IPF_OK;
    });
}

int ipf_cmesh_destroy(ipf_cmesh* m) {
    delete m;
    return IPF_OK;
}

int ipf_cmesh_set_material(ipf_cmesh* m, double E, double nu, double c, double phi, double psi) {
    if (!m) return null_arg("ipf_cmesh_set_material");
    return guarded([&] {
        m->mesh->SetMaterial(E, nu, c, phi, psi);
        return IPF_OK;
    });
}

int ipf_cmesh_set_normal_stress(ipf_cmesh* m, double inner, double outer) {
    if (!m) return null_arg("ipf_cmesh_set_normal_stress");
    m->mesh->SetNormalStress(inner, outer);
    return IPF_OK;
}

int ipf_cmesh_info(const ipf_cmesh* m, int64_t out[8]) {
    if (!m || !out) return null_arg("ipf_cmesh_info");
    const TPZCompMesh& c = *m->mesh;
    out[0] = c.NElements();
    out[1] = c.NNodes();
    out[2] = c.NEquations();
    out[3] = c.POrder();
    out[4] = c.NRadial();
    out[5] = c.NTheta();
    out[6] = c.NIntPoints();
    out[7] = c.NShapeF();
    return IPF_OK;
}

int ipf_cmesh_material(const ipf_cmesh* m, double out[5]) {
    if (!m || !out) return null_arg("ipf_cmesh_material");
    const TPZCompMesh& c = *m->mesh;
    out[0] = c.Lambda();
    out[1] = c.Mu();
    out[2] = c.Phi();
    out[3] = c.Psi();
    out[4] = c.Cohesion();
    return IPF_OK;
}

int ipf_cmesh_nodes(const ipf_cmesh* m, double* xy) {
    if (!m || !xy) return null_arg("ipf_cmesh_nodes");
    for (int64_t n = 0; n < m->mesh->NNodes(); n++) m->mesh->NodeCoords(n, xy[2 * n], xy[2 * n + 1]);
    return IPF_OK;
}

int ipf_cmesh_connectivity(const ipf_cmesh* m, int64_t* conn_ptr, int64_t* conn) {
    if (!m || !conn_ptr || !conn) return null_arg("ipf_cmesh_connectivity");
    std::vector<int64_t> nodes;
    conn_ptr[0] = 0;
    for (int64_t e = 0; e < m->mesh->NElements(); e++) {
        m->mesh->ElementConnects(e, nodes);
        std::memcpy(conn + conn_ptr[e], nodes.data(), sizeof(int64_t) * nodes.size());
        conn_ptr[e + 1] = conn_ptr[e] + (int64_t)nodes.size();
    }
    return IPF_OK;
}

int ipf_cmesh_dirichlet_mask(const ipf_cmesh* m, signed char* mask) {
    if (!m || !mask) return null_arg("ipf_cmesh_dirichlet_mask");
    std::vector<signed char> v;
    m->mesh->DirichletMask(v);
    std::memcpy(mask, v.data(), v.size());
    return IPF_OK;
}

int ipf_cmesh_point_coords(const ipf_cmesh* m, double* xy) {
    if (!m || !xy) return null_arg("ipf_cmesh_point_coords");
    const TPZCompMesh& c = *m->mesh;
    int64_t k = 0;
    for (int64_t e = 0; e < c.NElements(); e++) {
        const int nip = c.NIntPoints(c.ElementType(e));
        for (int ip = 0; ip < nip; ip++, k++) c.PointCoords(e, ip, xy[2 * k], xy[2 * k + 1]);
    }
    return IPF_OK;
}

int ipf_cmesh_shape_table(const ipf_cmesh* m, int type, int sizes[2], double* tab) {
    if (!m || !sizes) return null_arg("ipf_cmesh_shape_table");
    const TPZCompMesh& c = *m->mesh;
    if (type < 0) {
        if (c.Mixed()) return fail_status(IPF_EINVAL, "ipf_cmesh_shape_table: the mesh mixes element types; name one");
        type = c.HasType(TPZCompMesh::EQuad) ? 0 : 1;
    }
    if (type > 1) return fail_status(IPF_EINVAL, "ipf_cmesh_shape_table: type must be -1, 0 or 1");
    const auto t = (TPZCompMesh::ElType)type;
    sizes[0] = c.NIntPoints(t);
    sizes[1] = c.NShapeF(t);
    if (tab) {
        std::vector<double> v;
        c.ShapeTable(t, v);
        std::memcpy(tab, v.data(), sizeof(double) * v.size());
    }
    return IPF_OK;
}

int ipf_cmesh_element_types(const ipf_cmesh* m, signed char* types) {
    if (!m || !types) return null_arg("ipf_cmesh_element_types");
    for (int64_t e = 0; e < m->mesh->NElements(); e++) types[e] = (signed char)m->mesh->ElementType(e);
    return IPF_OK;
}

int ipf_cmesh_read_gmsh(const char* path, int porder, ipf_cmesh** out) {
    if (!path || !out) return null_arg("ipf_cmesh_read_gmsh");
    *out = nullptr;
    return guarded([&] {
        auto m = std::make_unique<ipf_cmesh>();
        m->mesh = TPZCompMesh::ReadGmsh(path, porder);
        *out = m.release();
        return IPF_OK;
    });
}

int ipf_cmesh_gmsh_counts(const ipf_cmesh* m, int64_t out[6]) {
    if (!m || !out) return null_arg("ipf_cmesh_gmsh_counts");
    const auto& k = m->mesh->Counts();
    out[0] = k.nodes;
    out[1] = k.elements;
    out[2] = k.points;
    out[3] = k.lines;
    out[4] = k.triangles;
    out[5] = k.quads;
    return IPF_OK;
}

int ipf_cmesh_corners(const ipf_cmesh* m, double* corners) {
    if (!m || !corners) return null_arg("ipf_cmesh_corners");
    const TPZCompMesh& c = *m->mesh;
    for (int64_t e = 0; e < c.NElements(); e++) {
        double X[4][2];
        c.CornerCoords(e, X);
        std::memcpy(corners + 8 * e, &X[0][0], sizeof(double) * 8);
    }
    return IPF_OK;
}

int ipf_cmesh_set_update_mem(ipf_cmesh* m, int update_mem) {
    if (!m) return null_arg("ipf_cmesh_set_update_mem");
    m->mesh->SetUpdateMem(update_mem != 0);
    return IPF_OK;
}

int ipf_sm_create(ipf_cmesh* m, int device, ipf_sm** out) {
    if (!m || !out) return null_arg("ipf_sm_create");
    *out = nullptr;
    return guarded([&] {
        auto s = std::make_unique<ipf_sm>();
        s->mesh = m;
        s->sm = std::make_unique<TPZElastoPlasticIntPointsStructMatrix>(m->mesh.get(), device);
        s->sm->SetUpDataStructure();
        if (device >= 0) s->sm->TransferDataToGPU();
        *out = s.release();
        return IPF_OK;
    });
}

int ipf_sm_destroy(ipf_sm* sm) {
    delete sm;
    return IPF_OK;
}

ipf_ctx* ipf_sm_context(ipf_sm* sm) { return sm ? sm->sm->Integrator().Context().get() : nullptr; }

int ipf_sm_sizes(const ipf_sm* sm, int64_t out[8]) {
    if (!sm || !out) return null_arg("ipf_sm_sizes");
    auto& in = sm->sm->Integrator();
    auto& B = in.BlockMatrix().Blocks();
    out[0] = B.fNumBlocks;
    out[1] = in.BlockMatrix().Rows() / 3;
    out[2] = in.BlockMatrix().Rows();
    out[3] = in.BlockMatrix().Cols();
    out[4] = sm->mesh->mesh->NEquations();
    out[5] = in.NColor();
    out[6] = B.fNumBlocks ? B.fMatrixPosition[B.fNumBlocks] : 0;
    out[7] = 0;
    return IPF_OK;
}

int ipf_sm_export_blocks(const ipf_sm* sm, int* row_sizes, int* col_sizes, int64_t* matrix_position,
                         int* row_first, int* col_first, double* storage) {
    if (!sm) return null_arg("ipf_sm_export_blocks");
    auto& B = sm->sm->Integrator().BlockMatrix().Blocks();
    const int64_t n = B.fNumBlocks;
    if (row_sizes) std::memcpy(row_sizes, B.fRowSizes.data(), sizeof(int) * n);
    if (col_sizes) std::memcpy(col_sizes, B.fColSizes.data(), sizeof(int) * n);
    if (matrix_position) std::memcpy(matrix_position, B.fMatrixPosition.data(), sizeof(int64_t) * (n + 1));
    if (row_first) std::memcpy(row_first, B.fRowFirstIndex.data(), sizeof(int) * (n + 1));
    if (col_first) std::memcpy(col_first, B.fColFirstIndex.data(), sizeof(int) * (n + 1));
    if (!storage) return IPF_OK;
    return guarded([&] {
        // the host values are filled on first request (device setups compute B on the GPU)
        const_cast<ipf_sm*>(sm)->sm->Integrator().FillHostBlocks();
        std::memcpy(storage, B.fStorage.data(), sizeof(double) * B.fStorage.size());
        return IPF_OK;
    });
}

int ipf_sm_export_indexes(const ipf_sm* sm, int* dof_indexes, int* color_indexes, int* el_color_index,
                          int64_t* first_color_index, double* weight, double* rhs_linear) {
    if (!sm) return null_arg("ipf_sm_export_indexes");
    auto& in = sm->sm->Integrator();
    if (dof_indexes) std::memcpy(dof_indexes, in.DoFIndexes().data(), sizeof(int) * in.DoFIndexes().size());
    if (color_indexes) std::memcpy(color_indexes, in.ColorIndexes().data(), sizeof(int) * in.ColorIndexes().size());
    if (el_color_index)
        std::memcpy(el_color_index, in.ElColorIndex().data(), sizeof(int) * in.BlockMatrix().Blocks().fNumBlocks);
    if (first_color_index)
        std::memcpy(first_color_index, in.FirstColorIndex().data(), sizeof(int64_t) * in.FirstColorIndex().size());
    if (weight)
        std::memcpy(weight, in.ConstitutiveLawProcessor().Weight().data(),
                    sizeof(double) * in.ConstitutiveLawProcessor().Weight().size());
    if (rhs_linear)
        std::memcpy(rhs_linear, sm->sm->RhsLinear().data(), sizeof(double) * sm->sm->RhsLinear().Rows());
    return IPF_OK;
}

int ipf_sm_assemble_rhs(ipf_sm* sm, const double* u_host, double* rhs_host) {
    if (!sm || !u_host || !rhs_host) return null_arg("ipf_sm_assemble_rhs");
    return guarded([&] {
        TPZCompMesh* m = sm->mesh->mesh.get();
        std::memcpy(m->Solution().data(), u_host, sizeof(double) * m->NEquations());
        TPZFMatrix<STATE> rhs;
        sm->sm->Assemble(rhs);
        std::memcpy(rhs_host, rhs.data(), sizeof(double) * m->NEquations());
        return IPF_OK;
    });
}

int ipf_sm_create_tangent(ipf_sm* sm, int64_t* nnz) {
    if (!sm || !nnz) return null_arg("ipf_sm_create_tangent");
    return guarded([&] {
        if (!sm->K) sm->K.reset(static_cast<TPZSYsmpMatrix<STATE>*>(sm->sm->Create()));
        *nnz = sm->K->JA().size();
        return IPF_OK;
    });
}

int ipf_sm_export_tangent_pattern(const ipf_sm* sm, int64_t* ia, int64_t* ja, int64_t* color_l_sequence,
                                  int64_t* first_color_l_index) {
    if (!sm) return null_arg("ipf_sm_export_tangent_pattern");
    if (!sm->K) {
        ipf_internal_set_error("ipf_sm_export_tangent_pattern: call ipf_sm_create_tangent first");
        return IPF_ESTATE;
    }
    auto& in = sm->sm->Integrator();
    if (ia) std::memcpy(ia, sm->K->IA().data(), sizeof(int64_t) * sm->K->IA().size());
    if (ja) std::memcpy(ja, sm->K->JA().data(), sizeof(int64_t) * sm->K->JA().size());
    if (color_l_sequence) {
        // filled on the host on first request when the device built the pattern
        auto& cl = in.ColorLSequenceFor(sm->K->IA(), sm->K->JA());
        std::memcpy(color_l_sequence, cl.data(), sizeof(int64_t) * cl.size());
    }
    if (first_color_l_index)
        std::memcpy(first_color_l_index, in.FirstColorLIndex().data(),
                    sizeof(int64_t) * in.FirstColorLIndex().size());
    return IPF_OK;
}

int ipf_sm_clone(const ipf_sm* sm, ipf_sm** out) {
    if (!sm || !out) return null_arg("ipf_sm_clone");
    *out = nullptr;
    return guarded([&] {
        auto s = std::make_unique<ipf_sm>();
        s->mesh = sm->mesh;
        s->sm.reset(static_cast<TPZElastoPlasticIntPointsStructMatrix*>(sm->sm->Clone()));
        *out = s.release();
        return IPF_OK;
    });
}

int ipf_sm_create_assemble(ipf_sm* sm, const double* u_host, double* rhs_host, int64_t* nnz) {
    if (!sm || !u_host || !rhs_host || !nnz) return null_arg("ipf_sm_create_assemble");
    return guarded([&] {
        TPZCompMesh* m = sm->mesh->mesh.get();
        std::memcpy(m->Solution().data(), u_host, sizeof(double) * m->NEquations());
        TPZFMatrix<STATE> rhs;
        std::unique_ptr<TPZMatrix<STATE>> K(sm->sm->CreateAssemble(rhs, TPZAutoPointer<TPZGuiInterface>()));
        std::memcpy(rhs_host, rhs.data(), sizeof(double) * m->NEquations());
        sm->K.reset(static_cast<TPZSYsmpMatrix<STATE>*>(K.release()));
        *nnz = sm->K->JA().size();
        return IPF_OK;
    });
}

int ipf_sm_tangent_values(const ipf_sm* sm, double* kg_host) {
    if (!sm || !kg_host) return null_arg("ipf_sm_tangent_values");
    if (!sm->K) return fail_status(IPF_ESTATE, "ipf_sm_tangent_values: no matrix (ipf_sm_create_tangent or ipf_sm_create_assemble)");
    std::memcpy(kg_host, sm->K->A().data(), sizeof(double) * sm->K->A().size());
    return IPF_OK;
}

int ipf_sm_export_csr(ipf_sm* sm, int* row_ptr, int* col_ind) {
    if (!sm) return null_arg("ipf_sm_export_csr");
    return guarded([&] {
        auto& bm = sm->sm->Integrator().BlockMatrix();
        bm.CSRVectors();
        auto& B = bm.Blocks();
        if (row_ptr) std::memcpy(row_ptr, B.fRowPtr.data(), sizeof(int) * B.fRowPtr.size());
        if (col_ind) std::memcpy(col_ind, B.fColInd.data(), sizeof(int) * B.fColInd.size());
        return IPF_OK;
    });
}

int ipf_sm_assemble_tangent(ipf_sm* sm, const double* u_host, double* rhs_host, double* kg_host) {
    if (!sm || !u_host || !rhs_host || !kg_host) return null_arg("ipf_sm_assemble_tangent");
    if (!sm->K) {
        ipf_internal_set_error("ipf_sm_assemble_tangent: call ipf_sm_create_tangent first");
        return IPF_ESTATE;
    }
    return guarded([&] {
        TPZCompMesh* m = sm->mesh->mesh.get();
        std::memcpy(m->Solution().data(), u_host, sizeof(double) * m->NEquations());
        TPZFMatrix<STATE> rhs;
        sm->sm->Assemble(*sm->K, rhs);
        std::memcpy(rhs_host, rhs.data(), sizeof(double) * m->NEquations());
        std::memcpy(kg_host, sm->K->A().data(), sizeof(double) * sm->K->A().size());
        return IPF_OK;
    });
}

}  // extern "C"
