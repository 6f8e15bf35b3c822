// ipf_pattern.hip -- the tangent's upper-triangle CSR pattern and every
// element entry's position in it, built on the device from the element slots'
// DoF indexes: the pattern Create() gives the reference's TPZSYsmpMatrix (the
// DoF pairs (i, j), i <= j, that share an element) and the positions
// FillLIndexes / me() look up (TPZNumericalIntegrator.cpp:538-606), here in
// the tile layout the assembly kernel reads.
//
//   1. one 64-bit key min(i, j) * neq + max(i, j) per element entry (packed
//      row-major upper triangle of the element's DoFs; padding slots a key
//      above every real one);
//   2. radix sort and unique: the sorted distinct keys are the pattern in
//      CSR order -- ja = key % neq, ia from where the row (key / neq) changes;
//   3. each entry's position: a binary search of its column in its row.
//
// Integer work only: the same pattern and positions as the host build, bit
// for bit.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "ipf_kernels.h"

namespace ipf {

namespace {

constexpr uint64_t kNoKey = ~0ull;  // padding slots: sorts after every real key
constexpr int kBlock = 256;

// entries of the packed row-major upper triangle before row a (row a holds
// CE - a entries)
__device__ __forceinline__ int tri_row_off(int ce, int a) { return a * ce - a * (a - 1) / 2; }

// one thread per (element slot, row a): the row's keys
template <int SHAPE>
__global__ void k_pattern_keys(GroupDev g, int64_t neq, uint64_t* __restrict__ keys) {
    constexpr int CE = shape_ce(SHAPE), ET = shape_et(SHAPE), NT = CE * (CE + 1) / 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nelem * CE) return;
    const int64_t slot = t / CE;
    const int a = (int)(t - slot * CE);
    uint64_t* out = keys + slot * NT + tri_row_off(CE, a);
    if ((int)(slot % ET) >= tile_nvalid(g, slot / ET)) {
        for (int b = a; b < CE; b++) out[b - a] = kNoKey;
        return;
    }
    const int* d = g.dof + slot * CE;
    const int64_t i = d[a];
    for (int b = a; b < CE; b++) {
        const int64_t j = d[b];
        out[b - a] = (uint64_t)(i <= j ? i * neq + j : j * neq + i);
    }
}

// thread k of [0, nnz]: ia[r] = k for the rows r in (row(k - 1), row(k)]
// (row(-1) = -1, row(nnz) = neq), so empty rows get their successor's start
// and ia[neq] = nnz; ja[k] = column of key k
__global__ void k_pattern_rows(const uint64_t* __restrict__ uniq, int64_t nnz, int64_t neq, int64_t* __restrict__ ia,
                               int64_t* __restrict__ ja) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int64_t prev = k == 0 ? -1 : (int64_t)(uniq[k - 1] / (uint64_t)neq);
    const int64_t cur = k == nnz ? neq : (int64_t)(uniq[k] / (uint64_t)neq);
    for (int64_t r = prev + 1; r <= cur; r++) ia[r] = k;
    if (k < nnz) ja[k] = (int64_t)(uniq[k] % (uint64_t)neq);
}

// one thread per (element slot, row a): the CSR positions of the row's
// entries (0 for padding slots, which the assembly skips)
template <int SHAPE>
__global__ void k_pattern_positions(GroupDev g, const int64_t* __restrict__ ia, const int64_t* __restrict__ ja,
                                    int* __restrict__ lidx) {
    constexpr int CE = shape_ce(SHAPE), ET = shape_et(SHAPE), NT = CE * (CE + 1) / 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nelem * CE) return;
    const int64_t slot = t / CE;
    const int a = (int)(t - slot * CE);
    int* out = lidx + slot * NT + tri_row_off(CE, a);
    if ((int)(slot % ET) >= tile_nvalid(g, slot / ET)) {
        for (int b = a; b < CE; b++) out[b - a] = 0;
        return;
    }
    const int* d = g.dof + slot * CE;
    const int64_t i = d[a];
    for (int b = a; b < CE; b++) {
        const int64_t j = d[b];
        const int64_t row = i <= j ? i : j, col = i <= j ? j : i;
        int64_t lo = ia[row], hi = ia[row + 1];  // first ja >= col in [lo, hi)
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (ja[mid] < col)
                lo = mid + 1;
            else
                hi = mid;
        }
        out[b - a] = (int)lo;
    }
}

}  // namespace

void launch_pattern_keys(int shape, const GroupDev& g, int64_t neq, uint64_t* keys, hipStream_t s) {
    const int64_t n = g.nelem * shape_ce(shape);
    if (n == 0) return;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
    dispatch_shape(shape, [&](auto sc) {
        hipLaunchKernelGGL((k_pattern_keys<decltype(sc)::value>), grid, block, 0, s, g, neq, keys);
    });
}

void launch_pattern_positions(int shape, const GroupDev& g, const int64_t* ia, const int64_t* ja, int* lidx,
                              hipStream_t s) {
    const int64_t n = g.nelem * shape_ce(shape);
    if (n == 0) return;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
    dispatch_shape(shape, [&](auto sc) {
        hipLaunchKernelGGL((k_pattern_positions<decltype(sc)::value>), grid, block, 0, s, g, ia, ja, lidx);
    });
}

void launch_pattern_rows(const uint64_t* uniq, int64_t nnz, int64_t neq, int64_t* ia, int64_t* ja, hipStream_t s) {
    const dim3 grid((unsigned)((nnz + 1 + kBlock - 1) / kBlock)), block(kBlock);
    hipLaunchKernelGGL(k_pattern_rows, grid, block, 0, s, uniq, nnz, neq, ia, ja);
}

hipError_t pattern_sort_unique(uint64_t* keys, uint64_t* sorted, int64_t n, unsigned int* count_dev, hipStream_t s) {
    size_t b_sort = 0, b_uniq = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, b_sort, keys, sorted, (size_t)n, 0, 64, s);
    if (e == hipSuccess) e = rocprim::unique(nullptr, b_uniq, sorted, keys, count_dev, (size_t)n, rocprim::equal_to<uint64_t>(), s);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    if ((e = hipMalloc(&tmp, std::max(b_sort, b_uniq))) != hipSuccess) return e;
    e = rocprim::radix_sort_keys(tmp, b_sort, keys, sorted, (size_t)n, 0, 64, s);
    if (e == hipSuccess)
        e = rocprim::unique(tmp, b_uniq, sorted, keys, count_dev, (size_t)n, rocprim::equal_to<uint64_t>(), s);
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    return e != hipSuccess ? e : e2;
}

}  // namespace ipf
