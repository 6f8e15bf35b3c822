// TPZNumericalIntegrator.h -- gather / B.u / return map / B^T(w sigma) /
// colored assembly driver; API of TPZNumericalIntegrator.h:22-111 of the
// reference.  Setup runs on the host from the mesh stand-in, the compute
// methods run on the GPU through the C ABI of include/intpointsfem.h.
#pragma once

#include <memory>

#include "TPZCompMesh.h"
#include "TPZConstitutiveLawProcessor.h"
#include "TPZIrregularBlocksMatrix.h"
#include "tpz_types.h"

class TPZNumericalIntegrator {
public:
    TPZNumericalIntegrator() = default;

    // TPZNumericalIntegrator.cpp:25-37 / :54-78 / :117-128 (device vectors)
    void Multiply(TPZVecGPU<REAL>& coef, TPZVecGPU<REAL>& delta_strain);
    void MultiplyTranspose(TPZVecGPU<REAL>& sigma, TPZVecGPU<REAL>& res);
    void ResidualIntegration(TPZFMatrix<REAL>& solution, TPZVecGPU<REAL>& rhs);
    // Host-matrix overloads (TPZNumericalIntegrator.cpp:40-51, :81-114)
    void Multiply(TPZFMatrix<REAL>& coef, TPZFMatrix<REAL>& delta_strain);
    void MultiplyTranspose(TPZFMatrix<REAL>& sigma, TPZFMatrix<REAL>& res);
    void ResidualIntegration(TPZFMatrix<REAL>& solution, TPZFMatrix<REAL>& rhs);

    // K_e = sum_ip B_ip^T Dep_ip B_ip (host, TPZNumericalIntegrator.cpp:236-270)
    void ComputeTangentMatrix(int64_t iel, TPZFMatrix<REAL>& Dep, TPZFMatrix<REAL>& K);
    // TPZNumericalIntegrator.cpp:131-214 on the GPU: Kg = the CSR values of
    // sum_e B_e^T (w Dep) B_e (zeroed first) and rhs = f_lin - B^T (w sigma)
    // (the reference adds f_lin in the caller).  Needs FillLIndexes.
    void KAssembly(TPZFMatrix<REAL>& solution, TPZVecGPU<STATE>& Kg, TPZVecGPU<STATE>& rhs);
    void KAssembly(TPZFMatrix<REAL>& solution, TPZVec<STATE>& Kg, TPZFMatrix<STATE>& rhs);

    // Setup (TPZNumericalIntegrator.cpp:272-536)
    void SetElementIndexes(TPZVec<int> element_indexes) { fElementIndex = element_indexes; }
    void SetUpIrregularBlocksData(TPZCompMesh* cmesh);
    // The two halves of SetUpIrregularBlocksData: the blocks' sizes and
    // prefix sums, and their values on the host (a no-op once filled).  The
    // device setup (TransferDataToGPU) needs only the sizes; the host values
    // are filled on first use (export, ComputeTangentMatrix).
    void SetUpIrregularBlocksSizes(TPZCompMesh* cmesh);
    void FillHostBlocks();
    int StressRateVectorSize(int dim);
    void SetUpIndexes(TPZCompMesh* cmesh);
    void SetUpColoredIndexes(TPZCompMesh* cmesh);
    void FillLIndexes(TPZVec<int64_t>& IA, TPZVec<int64_t>& JA);
    int64_t me(TPZVec<int64_t>& IA, TPZVec<int64_t>& JA, int64_t& i_dest, int64_t& j_dest);
    // FillLIndexes in two parts: the per-color offsets (fMatrixStride,
    // fFirstColorLIndex; cheap) and the positions fColorLSequence (one lookup
    // per element entry)
    void FillLOffsets();
    void FillColorLSequence(TPZVec<int64_t>& IA, TPZVec<int64_t>& JA);
    // Create()'s pattern and the entries' positions built on the device
    // (ipf_build_tangent_pattern): IA / JA filled from there, the positions
    // kept on the device; fColorLSequence is then filled only on request
    // (ColorLSequenceFor).
    void BuildPatternOnDevice(TPZVec<int64_t>& IA, TPZVec<int64_t>& JA);
    TPZVec<int64_t>& ColorLSequenceFor(TPZVec<int64_t>& IA, TPZVec<int64_t>& JA);
    bool isBuilt() { return fBlockMatrix.Rows() != 0; }

    // Creates the device context (if needed) and sets up the data there
    // (TPZNumericalIntegrator.cpp:218-233): element corners, shape table,
    // DoF and color indexes go up; B and the weights are computed on the GPU
    // (ipf_transfer_geometry) with the formulas of the host setup.
    void TransferDataToGPU(int device);
    std::shared_ptr<ipf_ctx> Context() { return fCtx; }

    TPZIrregularBlocksMatrix& BlockMatrix() { return fBlockMatrix; }
    TPZConstitutiveLawProcessor& ConstitutiveLawProcessor() { return fConstitutiveLawProcessor; }
    int64_t NColor() const { return fNColor; }
    TPZVec<int>& DoFIndexes() { return fDoFIndexes; }
    TPZVec<int>& ColorIndexes() { return fColorIndexes; }
    TPZVec<int>& ElColorIndex() { return fElColorIndex; }
    TPZVec<int64_t>& FirstColorIndex() { return fFirstColorIndex; }
    TPZVec<int64_t>& ColorLSequence() { return fColorLSequence; }
    TPZVec<int64_t>& FirstColorLIndex() { return fFirstColorLIndex; }
    int64_t NnzK() const { return fNnzK; }

private:
    TPZVec<int> fElementIndex;
    TPZIrregularBlocksMatrix fBlockMatrix;
    int64_t fNColor = -1;
    int64_t fNEquations = 0;
    TPZVec<int> fDoFIndexes;
    TPZVec<int> fColorIndexes;
    TPZConstitutiveLawProcessor fConstitutiveLawProcessor;
    TPZVec<int> fElColorIndex;
    TPZVec<int64_t> fFirstColorIndex;
    TPZVec<int64_t> fColorLSequence;
    TPZVec<int64_t> fFirstColorLIndex;
    int64_t fNnzK = 0;               // CSR values of the pattern FillLIndexes saw
    bool fPatternOnDevice = false;
    TPZCompMesh* fMesh = nullptr;
    std::shared_ptr<ipf_ctx> fCtx;
};
