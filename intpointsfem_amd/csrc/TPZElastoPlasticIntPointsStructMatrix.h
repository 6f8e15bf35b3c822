// TPZElastoPlasticIntPointsStructMatrix.h -- the drop-in structural matrix
// (API of TPZElastoPlasticIntPointsStructMatrix.h:18-73 of the reference).
// The reference derives from NeoPZ's TPZSymetricSpStructMatrix and is driven
// by TPZAnalysis through the TPZStructMatrix virtuals below; here it is driven
// directly, through the same virtuals, or through the C ABI.
#pragma once

#include <memory>

#include "TPZCompMesh.h"
#include "TPZNumericalIntegrator.h"
#include "tpz_types.h"

// The TPZStructMatrix virtuals TPZAnalysis calls on a structural matrix
// (main.cpp:181-234 through NeoPZ's TPZAnalysis::Assemble / AssembleResidual).
class TPZStructMatrix {
public:
    virtual ~TPZStructMatrix() = default;
    virtual TPZStructMatrix* Clone() = 0;
    virtual TPZMatrix<STATE>* Create() = 0;
    virtual TPZMatrix<STATE>* CreateAssemble(TPZFMatrix<STATE>& rhs, TPZAutoPointer<TPZGuiInterface> guiInterface) = 0;
    virtual void Assemble(TPZMatrix<STATE>& mat, TPZFMatrix<STATE>& rhs,
                          TPZAutoPointer<TPZGuiInterface> guiInterface) = 0;
    virtual void Assemble(TPZFMatrix<STATE>& rhs, TPZAutoPointer<TPZGuiInterface> guiInterface) = 0;
};

class TPZElastoPlasticIntPointsStructMatrix : public TPZStructMatrix {
public:
    explicit TPZElastoPlasticIntPointsStructMatrix(TPZCompMesh* cmesh, int device = 0);

    // A copy on the same mesh (.h:33, .cpp:30-32).  The copy shares the
    // original's device context -- device buffers, committed state, stream --
    // by reference count: either may be destroyed first (the reference's
    // copy duplicates the TPZVecGPU handles and frees them twice,
    // TPZVecGPU.h:24-27).  Calls through both run on the one context.
    TPZStructMatrix* Clone() override;

    // Builds the integrator data (once), the symmetric CSR pattern and its
    // per-color value positions, and uploads the data (.cpp:34-60).  Returns
    // a TPZSYsmpMatrix<STATE>; the caller owns it.
    TPZMatrix<STATE>* Create() override;
    // Create(), then rhs sized to the equations and Assemble(K, rhs) at
    // u = cmesh->Solution() (.cpp:71-77).
    TPZMatrix<STATE>* CreateAssemble(TPZFMatrix<STATE>& rhs, TPZAutoPointer<TPZGuiInterface> guiInterface) override;
    // K's values = the tangent at u = cmesh->Solution(), rhs = f_lin -
    // sum B^T (w sigma(B u)) (.cpp:98-135).  mat must be a TPZSYsmpMatrix
    // carrying the pattern Create() built (std::invalid_argument otherwise).
    // The reference's penalty-BC part fSparseMatrixLinear has no
    // counterpart: Dirichlet DoFs are eliminated in the solver.
    void Assemble(TPZMatrix<STATE>& mat, TPZFMatrix<STATE>& rhs,
                  TPZAutoPointer<TPZGuiInterface> guiInterface = TPZAutoPointer<TPZGuiInterface>()) override;
    // rhs = f_lin - sum B^T (w sigma(B u)) for u = cmesh->Solution() (.cpp:137-154).
    void Assemble(TPZFMatrix<STATE>& rhs,
                  TPZAutoPointer<TPZGuiInterface> guiInterface = TPZAutoPointer<TPZGuiInterface>()) override;
    // Builds the integrator data and uploads it without the CSR pattern
    // (enough for residual evaluations).
    void SetUpDataStructure();
    void TransferDataToGPU();

    TPZNumericalIntegrator& Integrator() { return fIntegrator; }
    TPZFMatrix<STATE>& RhsLinear() { return fRhsLinear; }
    TPZCompMesh* Mesh() { return fMesh; }

private:
    TPZElastoPlasticIntPointsStructMatrix(const TPZElastoPlasticIntPointsStructMatrix&) = default;
    void AssembleBoundaryData();
    void ComputeDomainElementIndexes(TPZVec<int>& element_indexes);

    TPZCompMesh* fMesh;
    int fDevice;
    int fDimension = 2;
    TPZNumericalIntegrator fIntegrator;
    TPZFMatrix<STATE> fRhsLinear;
    bool fTransferred = false;
};
