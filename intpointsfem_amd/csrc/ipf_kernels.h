// ipf_kernels.h -- device data layout shared by the gfx950 kernels and the
// host code that builds it.
//
// HBM layout (built once by ipf_transfer_data):
//   Elements are grouped by color, then by block shape (n_ip, C_e).  Within a
//   group, elements keep the reference's ascending order and are packed into
//   tiles of ET = floor(64 / n_ip) elements; one 64-lane wavefront owns a tile
//   and lane q < TP = ET * n_ip owns integration point (q / n_ip, q % n_ip).
//   B of a tile is stored point-interleaved (structure of arrays): value m of
//   point q (m = 3 * k + ... row k in {xx, xy, yy} times column j) sits at
//   double2 index (m / 2) * TP + q, component m % 2, so every 16-byte load of
//   the wavefront reads 64 consecutive double2 (1 KiB contiguous).
//   Per-point arrays (weight, plastic strain XX/XY/YY/ZZ, alpha, mtype,
//   branch) are SoA in the same tile order ("new" point index).
#pragma once

#include <cstdint>
#include <type_traits>

#include "ipf_pointmath.h"

namespace ipf {

// One launch unit: the tiles of one color and one block shape.  Tiles never
// mix materials (elements are sorted by material inside the group and the
// last tile of every material run is padded), so a wavefront reads one
// material; tmeta[t] = (material << 8) | (valid elements of tile t), the
// valid elements being the first ones of the tile.
struct GroupDev {
    const double* B;       // tile-major, point-interleaved B values of the group
    const int* dof;        // [nelem * CE] DoF indexes by element slot (padding slots unused)
    const int* tmeta;      // [ntiles] (material << 8) | valid elements
    const MatConst* mats;  // material table
    int64_t nelem;         // element slots of the group (ntiles * ET)
    int64_t ntiles;
    int64_t pt0;           // new-point index of the group's first slot
    const double* tab;     // shape table of the group's shape (ipf_geometry.h layout), or null
    // [nelem * CE / 2] per element slot and node, the node's x DoF (its y DoF
    // is the next one), kFirstWrite-flagged like the residual's DoF copy; set
    // for the non-committing residual and the tangent product when every
    // element numbers its nodes' DoFs as consecutive (x, y) pairs, else null
    const int* node;
};

// DoF index entries of the residual launches may carry kFirstWrite: the entry
// is its DoF's first write in launch order (see launch_residual_fmt).
constexpr int kFirstWrite = int(0x80000000u);
constexpr int kDofMask = 0x7fffffff;

__device__ __forceinline__ int tile_nvalid(const GroupDev& g, int64_t tile) { return g.tmeta[tile] & 255; }
__device__ __forceinline__ const MatConst& tile_material(const GroupDev& g, int64_t tile) {
    return g.mats[g.tmeta[tile] >> 8];
}

struct PointState {
    const double* w;
    double* ep0;          // plastic strain XX
    double* ep1;          // plastic strain XY (engineering)
    double* ep2;          // plastic strain YY
    double* ep3;          // plastic strain ZZ
    double* alpha;
    signed char* mtype;
    signed char* branch;
};

// Supported block shapes (n_ip, C_e): Q1, Q2, Q3 quadrilaterals with
// (p+1)^2 Gauss points; T1, T2, T3 triangles (P_p, 3, 6, 10 shape functions)
// with the 3-, 6- and 12-point rules of degree 2p.
enum Shape { SHAPE_Q1 = 0, SHAPE_Q2 = 1, SHAPE_Q3 = 2, SHAPE_T1 = 3, SHAPE_T2 = 4, SHAPE_T3 = 5, SHAPE_COUNT = 6 };

constexpr int shape_nip(int s) {
    return s == SHAPE_Q1 ? 4 : s == SHAPE_Q2 ? 9 : s == SHAPE_Q3 ? 16 : s == SHAPE_T1 ? 3 : s == SHAPE_T2 ? 6 : 12;
}
constexpr int shape_ce(int s) {
    return s == SHAPE_Q1 ? 8 : s == SHAPE_Q2 ? 18 : s == SHAPE_Q3 ? 32 : s == SHAPE_T1 ? 6 : s == SHAPE_T2 ? 12 : 20;
}
constexpr int shape_et(int s) { return 64 / shape_nip(s); }
constexpr int shape_tp(int s) { return shape_et(s) * shape_nip(s); }

// Calls f(std::integral_constant<int, S>()) for the runtime shape s: the
// kernels are templates on the shape.
template <class F>
inline void dispatch_shape(int s, F&& f) {
    switch (s) {
        case SHAPE_Q1: f(std::integral_constant<int, SHAPE_Q1>()); break;
        case SHAPE_Q2: f(std::integral_constant<int, SHAPE_Q2>()); break;
        case SHAPE_Q3: f(std::integral_constant<int, SHAPE_Q3>()); break;
        case SHAPE_T1: f(std::integral_constant<int, SHAPE_T1>()); break;
        case SHAPE_T2: f(std::integral_constant<int, SHAPE_T2>()); break;
        case SHAPE_T3: f(std::integral_constant<int, SHAPE_T3>()); break;
        default: break;
    }
}

// Storage formats of B (same tile order, point-interleaved 16-byte vectors):
//   FMT_D64  dense FP64, every stored value of the reference's block
//   FMT_C64  compact FP64, (dN_a/dx, dN_a/dy) per node and point
//   FMT_D32  dense FP32, FMT_C32 compact FP32 (values padded to 4 per point)
//   FMT_G64  geometric FP64: nothing per point; the kernels recompute the
//            compact values (and the weight) from the element's 4 corners
//            (8 doubles per element slot, tile order) and the shape table
//            (ipf_geometry.h), bit-identical to the stored FP64 forms built
//            from the same geometry
//   FMT_G32  geometric FP32: the FMT_G64 values rounded to FP32 (the compact
//            FP32 values), FP32 products and accumulation; the tangent
//            product reads an FP32 copy of w Dep
enum BFormat { FMT_D64 = 0, FMT_C64 = 1, FMT_D32 = 2, FMT_C32 = 3, FMT_G64 = 4, FMT_G32 = 5, FMT_COUNT = 6 };
constexpr int format_bytes(int f) { return (f == FMT_D32 || f == FMT_C32 || f == FMT_G32) ? 4 : 8; }
constexpr bool format_compact(int f) { return f == FMT_C64 || f == FMT_C32 || f == FMT_G64 || f == FMT_G32; }
constexpr bool format_geometric(int f) { return f == FMT_G64 || f == FMT_G32; }
constexpr bool format_stored(int f) { return !format_geometric(f); }
// values per integration point (stored formats: padded to a whole 16-byte
// vector; geometric: the compact values the kernel computes)
constexpr int format_values(int s, int f) {
    return ((format_compact(f) ? shape_ce(s) : 3 * shape_ce(s)) + (16 / format_bytes(f)) - 1) /
           (16 / format_bytes(f)) * (16 / format_bytes(f));
}

// Kernel launchers (ipf_kernels.hip).  All enqueue on `stream`.
void launch_multiply(int shape, const GroupDev& g, const double* u, double* strain, const int64_t* orig_of_new,
                     hipStream_t stream);
void launch_transpose(int shape, const GroupDev& g, const double* wsigma, double* r, const int64_t* orig_of_new,
                      hipStream_t stream);
// Return map of one group's points (per-point arrays in the reference's AoS
// order through orig_of_new; padding slots have orig_of_new < 0).
void launch_sigma(bool dep, bool commit, int shape, const GroupDev& g, int64_t npad, const double* strain,
                  double* wsigma, double* wdep, const int64_t* orig_of_new, const PointState& st, hipStream_t stream,
                  double* dep_soa = nullptr);
void launch_copy(int64_t n, const double* src, double* dst, hipStream_t stream);
void launch_fill(int64_t n, double* dst, double value, hipStream_t stream);

// Tangent operator (matrix-free) y += sum_e B_e^T (w Dep) B_e v_e: see
// launch_tangent_fmt (every B format); w Dep is stored SoA as
// dep_soa[k * npad + P], k = 3*row + col.
// diag += diagonal of the same operator.
// offd (optional) += for each element column j, the symmetric coupling of
// columns j and j ^ 1 (a node's x and y DoFs), at column j's DoF.
void launch_tangent_diag(int shape, const GroupDev& g, double* diag, double* offd, const double* dep_soa,
                         int64_t npad, hipStream_t stream);

// Tangent stiffness into CSR values (ipf_assemble.hip): kg[lidx[e * NT + t]] +=
// entry t of element e's packed upper triangle (NT = CE (CE + 1) / 2), for
// the elements of one group (B: the dense FP64 tiles).
void launch_assemble_K(int shape, const GroupDev& g, const double* dep_soa, int64_t npad, const int* lidx,
                       double* kg, hipStream_t stream);

// The tangent's CSR pattern on the device (ipf_pattern.hip): keys of a
// group's element entries (packed upper triangles by element slot, like
// lidx), sort + unique of n keys in place (keys: the distinct keys after the
// call, *count_dev their number; sorted: n scratch keys; synchronises s),
// ia / ja from the distinct keys, and the entries' positions.
void launch_pattern_keys(int shape, const GroupDev& g, int64_t neq, uint64_t* keys, hipStream_t stream);
hipError_t pattern_sort_unique(uint64_t* keys, uint64_t* sorted, int64_t n, unsigned int* count_dev,
                               hipStream_t stream);
void launch_pattern_rows(const uint64_t* uniq, int64_t nnz, int64_t neq, int64_t* ia, int64_t* ja,
                         hipStream_t stream);
void launch_pattern_positions(int shape, const GroupDev& g, const int64_t* ia, const int64_t* ja, int* lidx,
                              hipStream_t stream);

// The fused residual and the tangent product (ipf_kernels_fmt.hip): Bv points
// at the group's first tile in format fmt (FMT_D64: the dense tiles g.B;
// FMT_C64, FMT_D32, FMT_C32: the converted copies), or for FMT_G64 at the
// corners of the group's first element slot (with g.tab set).
// With flin set, g.dof's entries flagged kFirstWrite start from flin[d]
// instead of r[d] (each DoF's first write in launch order): r = flin then
// needs no copy before the launches.
// fast: non-committing launches take the certified elastic path
// (certified_elastic, ipf_pointmath.h); ignored when commit is set.
void launch_residual_fmt(int fmt, int shape, bool commit, bool fast, const GroupDev& g, const void* Bv,
                         const double* u, double* r, const double* flin, const PointState& st, hipStream_t stream);
// The tangent launches accumulate y += B^T (w Dep) B v; with zero_first,
// g.dof's (g.node's) entries flagged kFirstWrite start from 0 instead of y[d],
// so the launches of all groups overwrite y with K v and need no clear.
void launch_tangent_fmt(int fmt, int shape, const GroupDev& g, const void* Bv, const double* v, double* y,
                        const double* dep_soa, const float* dep32_soa, int64_t npad, bool zero_first,
                        hipStream_t stream);
// dst[i] = float(src[i]), i < n (the FP32 copy of w Dep for FMT_G32)
void launch_to_f32(int64_t n, const double* src, float* dst, hipStream_t stream);
// dense FP64 tiles of one group -> format fmt
void launch_convert_B(int fmt, int shape, const double* dense, void* out, int64_t ntiles, hipStream_t stream);
// Device setup of one group from geometry (corners: 8 doubles per element
// slot, tile order; g.tab the shape table): the dense FP64 tiles of B (zeros
// included) and the weights w[g.pt0 + ...] (0 on padding slots).
void launch_build_B(int shape, const GroupDev& g, const double* corners, double* dense, double* w,
                    hipStream_t stream);

// Vector kernels of the preconditioned CG (ipf_solver.hip).  `scal` is a small
// device array of scalars (see ipf_solver.hip for the slots).
void launch_dot(int64_t n, const double* x, const double* y, double* partial, int nblocks, double* out,
                hipStream_t stream);
// The preconditioner (pa, pb): z_i = pa_i r_i + pb_i r_partner on DoF pairs
// (i, i + 1), i even -- see launch_block_jacobi.
void launch_cg_init(int64_t n, const signed char* mask, const double* b, double* r, const double* pa,
                    const double* pb, double* z, double* p, hipStream_t stream);
// pa, pb from the tangent's diagonal and, when the DoFs come in aligned
// (x, y) = (2k, 2k + 1) pairs, the pairs' coupling offd (else null): the
// inverse of each node's 2x2 diagonal block, or Jacobi where none applies.
void launch_block_jacobi(int64_t n, const signed char* mask, const double* diag, const double* offd, double* pa,
                         double* pb, hipStream_t stream);
void launch_cg_mask(int64_t n, const signed char* mask, double* q, hipStream_t stream);
void launch_cg_update(int64_t n, const double* scal, const double* p, const double* q, double* x, double* r,
                      const double* pa, const double* pb, double* z, hipStream_t stream);
void launch_cg_direction(int64_t n, const double* scal, const double* z, double* p, hipStream_t stream);
// fused forms on the CG scalars scal (rz, pq, rz_new, ., alpha, beta):
// q masked, pq = p.q and alpha = rz / pq; update with alpha, rz_new = r.z,
// beta = rz_new / rz and rz = rz_new
void launch_cg_mask_dot(int64_t n, const signed char* mask, const double* p, double* q, double* partial, int nb,
                        double* scal, hipStream_t stream);
void launch_cg_update_dot(int64_t n, double* scal, const double* p, const double* q, double* x, double* r,
                          const double* pa, const double* pb, double* z, double* partial, int nb, hipStream_t stream);
void launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t stream);
// x = a x
void launch_scale(int64_t n, double a, double* x, hipStream_t stream);
void launch_mask_copy(int64_t n, const signed char* mask, const double* src, double* dst, hipStream_t stream);
int dot_blocks(int64_t n);
// out[k] = sum of x[i]^2 over the free (mask 0) i of segment k, the segment
// being chunks [seg_first_chunk[k], seg_first_chunk[k+1]) of [chunk_begin,
// chunk_end) ranges; fixed order (deterministic).
void launch_segment_sumsq(const double* x, const signed char* mask, const int64_t* chunk_begin,
                          const int64_t* chunk_end, int nchunk, double* partial, const int* seg_first_chunk, int nseg,
                          double* out, hipStream_t stream);

}  // namespace ipf
