// ipf_patch.hip -- host builder of the patch layout and its small kernels
// (the fix-up of shared nodes, the state and geometry gathers); the patch
// residual kernel itself sits next to the colored one in ipf_kernels_fmt.hip.
#include "ipf_patch.h"

#include <algorithm>
#include <cstring>
#include <utility>

namespace ipf {

namespace {

// 16-bit x / y interleaved into a 32-bit Morton (Z-order) code
uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Shared node i: r = f_lin + (-(sum of its partials in patch order)); its
// partials are the consecutive slots [fix_ptr[i], fix_ptr[i + 1]).
__global__ void k_patch_fixup(int64_t nshared, const int* __restrict__ fix_dof, const int* __restrict__ fix_ptr,
                              const double* __restrict__ partial, int64_t norphan, const int* __restrict__ orphan,
                              const double* __restrict__ flin, double* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nshared) {
        double fx = 0.0, fy = 0.0;
        const int k1 = fix_ptr[i + 1];
        for (int k = fix_ptr[i]; k < k1; k++) {
            fx = fx + partial[2 * (int64_t)k];
            fy = fy + partial[2 * (int64_t)k + 1];
        }
        const int d = fix_dof[i];
        r[d] = flin[d] + (-fx);
        r[d + 1] = flin[d + 1] + (-fy);
    } else if (i < nshared + norphan) {
        const int d = orphan[i - nshared];
        r[d] = flin[d];
    }
}

struct Ep4 {
    const double* s[4];
    double* d[4];
};

__global__ void k_patch_state(int64_t n, const int64_t* __restrict__ map, Ep4 e) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t P = map[q];
#pragma unroll
    for (int k = 0; k < 4; k++) e.d[k][q] = P >= 0 ? e.s[k][P] : 0.0;
}

__global__ void k_patch_geo(int64_t nslots, const int64_t* __restrict__ slot_geo, const double* __restrict__ geo,
                            double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8 * nslots) return;
    const int64_t s = t / 8, c = t - 8 * s;
    const int64_t g = slot_geo[s];
    out[t] = g >= 0 ? geo[8 * g + c] : 0.0;
}

}  // namespace

std::string build_patches(const std::vector<PatchElemIn>& el, const std::vector<int>& nodes, int64_t neq, int nw,
                          PatchHost& out) {
    out = PatchHost();
    if (!patch_waves_ok(nw)) return "patch layout: wavefront tiles per patch must be 2, 4 or 8";
    out.nw = nw;
    const int64_t ne = (int64_t)el.size();
    // uses of each node (keyed by its x DoF) over all elements
    std::vector<int> uses(neq, 0);
    for (int64_t e = 0; e < ne; e++) {
        const int nsh = shape_ce(el[e].shape) / 2;
        for (int a = 0; a < nsh; a++) {
            const int d = nodes[el[e].node0 + a];
            if (d < 0 || d + 1 >= neq) return "patch layout: node DoF out of range";
            uses[d]++;
        }
    }
    // Morton code of each centroid in the bounding box of its (shape, material) run
    std::vector<int64_t> idx(ne);
    for (int64_t e = 0; e < ne; e++) idx[e] = e;
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        if (el[a].shape != el[b].shape) return el[a].shape < el[b].shape;
        if (el[a].mat != el[b].mat) return el[a].mat < el[b].mat;
        return el[a].geo_slot < el[b].geo_slot;
    });
    std::vector<uint32_t> code(ne, 0);
    for (int64_t i = 0; i < ne;) {
        int64_t j = i;
        double x0 = el[idx[i]].cx, x1 = x0, y0 = el[idx[i]].cy, y1 = y0;
        while (j < ne && el[idx[j]].shape == el[idx[i]].shape && el[idx[j]].mat == el[idx[i]].mat) {
            x0 = std::min(x0, el[idx[j]].cx);
            x1 = std::max(x1, el[idx[j]].cx);
            y0 = std::min(y0, el[idx[j]].cy);
            y1 = std::max(y1, el[idx[j]].cy);
            j++;
        }
        const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
        for (int64_t k = i; k < j; k++) {
            const PatchElemIn& E = el[idx[k]];
            const double qx = (E.cx - x0) * sx, qy = (E.cy - y0) * sy;
            const uint32_t ix = (uint32_t)std::min(65535.0, std::max(0.0, qx));
            const uint32_t iy = (uint32_t)std::min(65535.0, std::max(0.0, qy));
            code[idx[k]] = spread16(ix) | (spread16(iy) << 1);
        }
        i = j;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        if (el[a].shape != el[b].shape) return el[a].shape < el[b].shape;
        if (el[a].mat != el[b].mat) return el[a].mat < el[b].mat;
        return code[a] < code[b];
    });

    std::vector<std::pair<int, int>> shared;  // (x DoF, partial slot)
    out.node_off.push_back(0);
    std::vector<int> local, cnt;
    int64_t i = 0;
    while (i < ne) {
        const int shape = el[idx[i]].shape;
        const int NW = nw, ET = shape_et(shape), NIP = shape_nip(shape), TP = shape_tp(shape);
        const int NSH = shape_ce(shape) / 2, PE = NW * ET;
        PatchHost::Range R;
        R.shape = shape;
        R.npatch = 0;
        R.patch0 = (int64_t)out.node_off.size() - 1;
        R.tile0 = out.ntiles;
        R.slot0 = out.nslots;
        R.pt0 = out.npts;
        R.lnode0 = (int64_t)out.lnode.size();
        while (i < ne && el[idx[i]].shape == shape) {
            const int mat = el[idx[i]].mat;
            int64_t j = i;
            while (j < ne && el[idx[j]].shape == shape && el[idx[j]].mat == mat && j - i < PE) j++;
            const int64_t npe = j - i;
            // inside the patch: elements by order key (stable), so that equal keys fill whole tiles
            std::stable_sort(idx.begin() + i, idx.begin() + j,
                             [&](int64_t a, int64_t b) { return el[a].key < el[b].key; });
            // the patch's distinct nodes, ascending by x DoF
            local.clear();
            for (int64_t k = i; k < j; k++)
                for (int a = 0; a < NSH; a++) local.push_back(nodes[el[idx[k]].node0 + a]);
            std::sort(local.begin(), local.end());
            local.erase(std::unique(local.begin(), local.end()), local.end());
            const int nn = (int)local.size();
            if (nn > patch_max_nodes(shape, nw) || nn > 65535) return "patch layout: too many nodes in a patch";
            auto lidx = [&](int d) { return (int)(std::lower_bound(local.begin(), local.end(), d) - local.begin()); };
            // tiles, slots, points, local node indexes; contributions counted per node
            cnt.assign(nn + 1, 0);
            for (int w = 0; w < NW; w++) {
                const int64_t k0 = i + (int64_t)w * ET;
                const int nv = (int)std::max<int64_t>(0, std::min<int64_t>(ET, j - k0));
                out.tmeta.push_back((mat << 8) | nv);
                for (int q = 0; q < ET; q++) {
                    if (q < nv) {
                        const PatchElemIn& E = el[idx[k0 + q]];
                        out.slot_geo.push_back(E.geo_slot);
                        for (int a = 0; a < NSH; a++) {
                            const int l = lidx(nodes[E.node0 + a]);
                            out.lnode.push_back((unsigned short)l);
                            cnt[l + 1]++;
                        }
                    } else {
                        out.slot_geo.push_back(-1);
                        for (int a = 0; a < NSH; a++) out.lnode.push_back(0);
                    }
                }
                for (int q = 0; q < TP; q++) {
                    const int e_ = q / NIP, ip = q - e_ * NIP;
                    out.pt_map.push_back(e_ < nv ? el[idx[k0 + e_]].P0 + ip : -1);
                }
            }
            for (int l = 0; l < nn; l++) cnt[l + 1] += cnt[l];
            const int64_t c0 = (int64_t)out.contrib.size();
            out.contrib.resize(c0 + cnt[nn]);
            std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
            const int64_t slot_first = (int64_t)out.lnode.size() - (int64_t)NW * ET * NSH;
            for (int s = 0; s < npe; s++)
                for (int a = 0; a < NSH; a++) {
                    const int l = out.lnode[slot_first + (int64_t)s * NSH + a];
                    out.contrib[c0 + fillp[l]++] = (unsigned short)(s * NSH + a);
                }
            for (int l = 0; l < nn; l++) {
                const int d = local[l];
                out.pnode.push_back(d);
                out.cbeg.push_back((int)(c0 + cnt[l]));
                if (cnt[l + 1] - cnt[l] == uses[d]) {
                    out.pslot.push_back(-1);
                } else {
                    out.pslot.push_back((int)out.npartial);
                    shared.emplace_back(d, (int)out.npartial);
                    out.npartial++;
                }
            }
            out.node_off.push_back((int64_t)out.pnode.size());
            out.ntiles += NW;
            out.nslots += (int64_t)NW * ET;
            out.npts += (int64_t)NW * TP;
            R.npatch++;
            i = j;
        }
        out.ranges.push_back(R);
    }
    out.cbeg.push_back((int)out.contrib.size());
    if (out.contrib.size() >= (size_t)0x7fffffff || out.npartial >= 0x7fffffff)
        return "patch layout: too many contributions";
    // shared nodes by x DoF, each node's partials in patch order; the slots
    // are renumbered in that order, so that a shared node's partials are
    // consecutive (fix_slot is the identity) and the fix-up reads them
    // without an indirection
    std::sort(shared.begin(), shared.end());
    std::vector<int> renum(out.npartial, -1);
    out.fix_ptr.push_back(0);
    for (size_t k = 0; k < shared.size();) {
        size_t m = k;
        while (m < shared.size() && shared[m].first == shared[k].first) {
            renum[shared[m].second] = (int)out.fix_slot.size();
            out.fix_slot.push_back((int)out.fix_slot.size());
            m++;
        }
        out.fix_dof.push_back(shared[k].first);
        out.fix_ptr.push_back((int)out.fix_slot.size());
        k = m;
    }
    for (int& ps : out.pslot)
        if (ps >= 0) ps = renum[ps];
    std::vector<unsigned char> covered(neq, 0);
    for (int64_t e = 0; e < ne; e++) {
        const int nsh = shape_ce(el[e].shape) / 2;
        for (int a = 0; a < nsh; a++) {
            const int d = nodes[el[e].node0 + a];
            covered[d] = covered[d + 1] = 1;
        }
    }
    for (int64_t d = 0; d < neq; d++)
        if (!covered[d]) out.orphan.push_back((int)d);
    return "";
}

void launch_patch_fixup(int64_t nshared, const int* fix_dof, const int* fix_ptr, const double* partial,
                        int64_t norphan, const int* orphan, const double* flin, double* r, hipStream_t s) {
    const int64_t n = nshared + norphan;
    if (n == 0) return;
    hipLaunchKernelGGL(k_patch_fixup, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nshared, fix_dof, fix_ptr,
                       partial, norphan, orphan, flin, r);
}

void launch_patch_state(int64_t n, const int64_t* pt_map, const double* const* ep, double* const* ep_patch,
                        hipStream_t s) {
    if (n == 0) return;
    Ep4 e;
    for (int k = 0; k < 4; k++) {
        e.s[k] = ep[k];
        e.d[k] = ep_patch[k];
    }
    hipLaunchKernelGGL(k_patch_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, pt_map, e);
}

void launch_patch_geo(int64_t nslots, const int64_t* slot_geo, const double* geo, double* geo_patch, hipStream_t s) {
    if (nslots == 0) return;
    const int64_t n = 8 * nslots;
    hipLaunchKernelGGL(k_patch_geo, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nslots, slot_geo, geo,
                       geo_patch);
}

}  // namespace ipf
