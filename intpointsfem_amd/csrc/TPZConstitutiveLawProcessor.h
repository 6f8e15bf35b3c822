// TPZConstitutiveLawProcessor.h -- per-integration-point Mohr-Coulomb
// processor; API of TPZConstitutiveLawProcessor.h:16-82 of the reference.
// The state (plastic strain, alpha, mtype) lives on the GPU in the
// integrator's device context; the host mirror is filled on request.
#pragma once

#include <memory>

#include "../../include/intpointsfem.h"
#include "TPZCompMesh.h"
#include "tpz_types.h"

class TPZConstitutiveLawProcessor {
public:
    TPZConstitutiveLawProcessor() = default;

    void SetUpDataByIntPoints(int64_t npts);
    void SetWeightVector(TPZVec<REAL>& weight) { fWeight = weight; }
    // The reference binds a TPZMatElastoPlastic2D material (.cpp:127-129); the
    // stand-in mesh carries the same parameters and the UpdateMem switch.
    void SetMaterial(TPZCompMesh* material) { fMaterial = material; }

    // w*sigma for every point from the strain (device vectors, AoS 3 per point)
    // (TPZConstitutiveLawProcessor.cpp:588-659).  Commits the state when the
    // material's UpdateMem switch is on.
    void ComputeSigma(TPZVecGPU<REAL>& glob_delta_strain, TPZVecGPU<REAL>& glob_sigma);
    // ... plus w*Dep, 9 per point (TPZConstitutiveLawProcessor.cpp:663-735).
    void ComputeSigmaDep(TPZVecGPU<REAL>& glob_delta_strain, TPZVecGPU<REAL>& glob_sigma,
                         TPZVecGPU<REAL>& glob_dep);
    // Host-matrix overloads (TPZConstitutiveLawProcessor.cpp:131, :344): copy, run, copy back.
    void ComputeSigma(TPZFMatrix<REAL>& glob_delta_strain, TPZFMatrix<REAL>& glob_sigma);
    void ComputeSigmaDep(TPZFMatrix<REAL>& glob_delta_strain, TPZFMatrix<REAL>& glob_sigma,
                         TPZFMatrix<REAL>& glob_dep);

    // Uploads the material and weights (TPZConstitutiveLawProcessor.cpp:739-742).
    void TransferDataToGPU(std::shared_ptr<ipf_ctx> ctx);

    // Committed state, Voigt 6 per point (TPZConstitutiveLawProcessor.h:53-63).
    void GetState(TPZFMatrix<REAL>& plastic_strain, TPZFMatrix<REAL>& alpha, TPZVec<int>& mtype,
                  TPZVec<int>& branch);

    int64_t NPoints() const { return fNpts; }
    TPZVec<REAL>& Weight() { return fWeight; }
    TPZCompMesh* Material() { return fMaterial; }

private:
    void SyncUpdateMem();

    int64_t fNpts = -1;
    TPZVec<REAL> fWeight;
    TPZCompMesh* fMaterial = nullptr;
    std::shared_ptr<ipf_ctx> fCtx;
};
