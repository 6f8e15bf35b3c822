// TPZIrregularBlocksMatrix.h -- block-diagonal matrix of per-element B blocks
// with irregular sizes; API of TPZIrregularBlocksMatrix.h:13-92 of the
// reference.  Storage on the GPU is the tiled layout of ipf_kernels.h, owned
// by the integrator's device context.
#pragma once

#include <memory>

#include "../../include/intpointsfem.h"
#include "tpz_types.h"

class TPZIrregularBlocksMatrix {
public:
    // TPZIrregularBlocksMatrix.h:18-30 (fMatrixPosition widened to int64)
    struct IrregularBlocks {
        int64_t fNumBlocks = 0;
        TPZVec<REAL> fStorage;
        TPZVec<int> fRowSizes;
        TPZVec<int> fColSizes;
        TPZVec<int64_t> fMatrixPosition;
        TPZVec<int> fRowFirstIndex;
        TPZVec<int> fColFirstIndex;
        TPZVec<int64_t> fMatrixStride;
        TPZVec<int> fRowPtr;
        TPZVec<int> fColInd;
    };

    TPZIrregularBlocksMatrix() = default;
    TPZIrregularBlocksMatrix(int64_t rows, int64_t cols) : fRows(rows), fCols(cols) {}

    int64_t Rows() const { return fRows; }
    int64_t Cols() const { return fCols; }
    void Resize(int64_t rows, int64_t cols) {
        fRows = rows;
        fCols = cols;
    }

    void SetBlocks(IrregularBlocks& blocks) { fBlocksInfo = blocks; }
    IrregularBlocks& Blocks() { return fBlocksInfo; }
    const IrregularBlocks& Blocks() const { return fBlocksInfo; }

    // y = B x (opt == 0) or y = -B^T x (opt != 0) on device vectors of length
    // Cols()/Rows() (TPZIrregularBlocksMatrix.cpp:97-140).  Throws
    // std::runtime_error on failure.
    void MultiplyVector(REAL* A, REAL* res, int opt);

    // CSR form of the whole block matrix (TPZIrregularBlocksMatrix.cpp:177-195).
    void CSRVectors();

    // The device context holding this matrix after TransferDataToGPU.
    void SetContext(std::shared_ptr<ipf_ctx> ctx) { fCtx = std::move(ctx); }

private:
    int64_t fRows = 0, fCols = 0;
    IrregularBlocks fBlocksInfo;
    std::shared_ptr<ipf_ctx> fCtx;
};
