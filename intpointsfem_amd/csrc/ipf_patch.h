// ipf_patch.h -- the patch layout of the non-committing fused residual
// (ipf_set_residual_assembly(ctx, IPF_ASSEMBLY_PATCH)).
//
// The colored layout (ipf_kernels.h) runs one launch per element color and
// read-modify-writes r once per color: every color touches most nodes, so u
// is gathered and r read and written about four times per residual.  The
// patch layout instead cuts the elements of each (shape, material) run into
// compact patches -- consecutive elements of a Morton order of the element
// centroids, nw wavefront tiles each -- and runs one
// workgroup per patch in a single launch per shape:
//   * the patch's distinct nodes are gathered once into LDS (u read once per
//     node and patch);
//   * each wavefront computes its tile's points (geometric FP64 B, return
//     map) and reduces them to per-element column sums in LDS;
//   * each node then sums its elements' contributions from LDS in a fixed
//     order; a node whose elements all lie in the patch ("owned") gets
//     r = f_lin + (-sum) written once, without reading r; a node shared with
//     other patches writes its partial sum to its own slot, and a fix-up
//     pass sums each shared node's partials in a fixed order.
// Every DoF of r is written exactly once and no atomics are used: the result
// is deterministic; it equals the colored residual up to the order of the
// floating-point sums (a few ulps of the element forces).  The per-point
// arithmetic (geometry, strain, return map, B^T products) is the colored
// kernel's, so branch decisions are the same.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ipf_kernels.h"

namespace ipf {

// Wavefront tiles per patch: one workgroup of nw * 64 lanes, nw = 2, 4 or 8
// per context (ipf_set_patch_waves, or chosen by ipf_autotune_assembly).
// Smaller patches share more nodes with their neighbours (more partial sums
// for the fix-up); larger workgroups wait longer at their barriers for their
// slowest wave's return map and hold more LDS each.  Four is the default:
// the fastest at 800x800 Q2 and Q3 and on the 64-instance batch on MI355X.
constexpr int kPatchWavesDefault = 4;
constexpr bool patch_waves_ok(int nw) { return nw == 2 || nw == 4 || nw == 8; }
constexpr int patch_elems(int s, int nw) { return nw * shape_et(s); }
constexpr int patch_max_nodes(int s, int nw) { return patch_elems(s, nw) * (shape_ce(s) / 2); }

// One real element of the colored layout, the input of the patch builder.
struct PatchElemIn {
    int shape;
    int mat;
    int64_t geo_slot;  // its slot in the colored corner array (8 doubles per slot)
    int64_t P0;        // new-point index of its first point in the colored layout (points consecutive)
    int64_t node0;     // first of its shape_ce / 2 node x-DoFs in the builder's node array
    double cx, cy;     // centroid of its corners
    int key = 0;       // order key inside its patch (ascending, stable): elements of equal keys share tiles
};

// Device view of the patches of one element shape.
struct PatchDev {
    int64_t npatch;
    const int* tmeta;              // per tile: (material << 8) | valid elements
    const MatConst* mats;
    const double* geo;             // per slot: 4 corners (8 doubles); slot = tile * ET + element
    const unsigned short* lnode;   // per slot: shape_ce / 2 local node indexes into the patch's nodes
    const int64_t* node_off;       // [npatch + 1] first patch-node of each patch (absolute)
    const int* pnode;              // per patch-node: the node's x DoF (y DoF = x + 1)
    const int* pslot;              // per patch-node: partial slot, or -1 when the patch owns the node
    const int* cbeg;               // per patch-node: first contribution; the next patch-node's is the end
    const unsigned short* contrib; // slot-in-patch * (shape_ce / 2) + local node a, ascending
    const double* ep0;             // plastic strain XX, XY, YY, ZZ in patch point order
    const double* ep1;
    const double* ep2;
    const double* ep3;
    const double* tab;             // shape table
};

// Host result of the builder (device arrays are uploaded from it).
struct PatchHost {
    struct Range {
        int shape;
        int64_t npatch, patch0, tile0, slot0, pt0, lnode0;
    };
    std::vector<Range> ranges;
    std::vector<int> tmeta;                // per tile
    std::vector<int64_t> slot_geo;         // per slot: colored geo slot, -1 padding
    std::vector<unsigned short> lnode;     // per slot of each range: NSH entries (range offset lnode0)
    std::vector<int64_t> node_off;         // [npatch_total + 1]
    std::vector<int> pnode, pslot, cbeg;   // per patch-node (cbeg: + 1 sentinel)
    std::vector<unsigned short> contrib;
    std::vector<int64_t> pt_map;           // per patch point: colored new-point index, -1 padding
    std::vector<int> fix_dof, fix_ptr, fix_slot;  // shared nodes by x DoF: partial slots [fix_ptr[i], fix_ptr[i+1])
                                                  // (fix_slot: the identity, slots numbered in this order)
    std::vector<int> orphan;               // DoFs of no element
    int64_t npartial = 0;
    int64_t ntiles = 0, nslots = 0, npts = 0;
    int nw = kPatchWavesDefault;           // wavefront tiles per patch
};

// Builds the patches of nw wavefront tiles of the elements `el` (node x-DoFs
// in `nodes`), neq equations.  Returns "" or an error text.
std::string build_patches(const std::vector<PatchElemIn>& el, const std::vector<int>& nodes, int64_t neq, int nw,
                          PatchHost& out);

// Kernels (ipf_kernels_fmt.hip: the residual; ipf_patch.hip: the rest).
void launch_residual_patch(int shape, int nw, bool fast, const PatchDev& p, const double* u, double* r,
                           const double* flin, double* partial, hipStream_t stream);
void launch_patch_fixup(int64_t nshared, const int* fix_dof, const int* fix_ptr, const double* partial,
                        int64_t norphan, const int* orphan, const double* flin, double* r, hipStream_t stream);
// ep_patch[k][q] = ep[k][pt_map[q]] (0 for padding), k = 0..3
void launch_patch_state(int64_t npts, const int64_t* pt_map, const double* const* ep, double* const* ep_patch,
                        hipStream_t stream);
// geo_patch[8 s + c] = geo[8 slot_geo[s] + c] (0 for padding slots)
void launch_patch_geo(int64_t nslots, const int64_t* slot_geo, const double* geo, double* geo_patch,
                      hipStream_t stream);

}  // namespace ipf
