// TPZElastoPlasticIntPointsStructMatrix.h -- the drop-in structural matrix
// (API of TPZElastoPlasticIntPointsStructMatrix.h:18-73 of the reference).
// The reference derives from NeoPZ's TPZSymetricSpStructMatrix and is driven
// by TPZAnalysis; here it is driven directly (or through the C ABI).
#pragma once

#include <memory>

#include "TPZCompMesh.h"
#include "TPZNumericalIntegrator.h"
#include "tpz_types.h"

class TPZElastoPlasticIntPointsStructMatrix {
public:
    explicit TPZElastoPlasticIntPointsStructMatrix(TPZCompMesh* cmesh, int device = 0);

    // Builds the integrator data (once), the symmetric CSR pattern and its
    // per-color value positions, and uploads the data (.cpp:34-60).
    TPZSYsmpMatrix<STATE>* Create();
    // rhs = f_lin - sum B^T (w sigma(B u)) for u = cmesh->Solution() (.cpp:137-154).
    void Assemble(TPZFMatrix<STATE>& rhs);
    // K's values = the tangent at u = cmesh->Solution(), rhs as above
    // (.cpp:98-135; K from Create(); the reference's penalty-BC part
    // fSparseMatrixLinear has no counterpart: Dirichlet DoFs are eliminated).
    void Assemble(TPZSYsmpMatrix<STATE>& K, TPZFMatrix<STATE>& rhs);
    // Builds the integrator data and uploads it without the CSR pattern
    // (enough for residual evaluations).
    void SetUpDataStructure();
    void TransferDataToGPU();

    TPZNumericalIntegrator& Integrator() { return fIntegrator; }
    TPZFMatrix<STATE>& RhsLinear() { return fRhsLinear; }
    TPZCompMesh* Mesh() { return fMesh; }

private:
    void AssembleBoundaryData();
    void ComputeDomainElementIndexes(TPZVec<int>& element_indexes);

    TPZCompMesh* fMesh;
    int fDevice;
    int fDimension = 2;
    TPZNumericalIntegrator fIntegrator;
    TPZFMatrix<STATE> fRhsLinear;
    bool fTransferred = false;
};
