// ipf_kernels.hip -- gfx950 kernels of the reference's separate steps on the
// dense FP64 blocks: Multiply (gather + B.u, TPZNumericalIntegrator.cpp:25-51),
// MultiplyTranspose (-B^T x with colored accumulation, :54-105), ComputeSigma
// / ComputeSigmaDep per point (ipf_pointmath.h), the diagonal (and node blocks) of the
// tangent, copies and fills.  The fused residual and the tangent product, in
// every B format, are in ipf_kernels_fmt.hip.  One 64-lane wavefront per tile
// of elements (ipf_kernels.h); B values are streamed with coalesced 16-byte
// loads and summed column by column in order (TPZIrregularBlocksMatrix.cpp:
// 97-140); elements of one color share no DoF
// (TPZNumericalIntegrator.cpp:452-536), so the read-modify-write of r needs no
// atomics and colors run as consecutive launches on one stream.
#include "ipf_kernels.h"

namespace ipf {

namespace {

constexpr int kWave = 64;

template <int SHAPE>
struct Cfg {
    static constexpr int NIP = shape_nip(SHAPE);
    static constexpr int CE = shape_ce(SHAPE);
    static constexpr int ET = shape_et(SHAPE);
    static constexpr int TP = shape_tp(SHAPE);
    static constexpr int NB = 3 * CE;            // B values per point
    static constexpr bool KEEP_B = (NB <= 64);   // keep B in registers across the return map
    // One wavefront per workgroup: the LDS reduction then waits for no other
    // wave's (divergent) return map.
    static constexpr int WPB = 1;
    // waves per SIMD the register allocation must allow (Q2: 168 VGPRs)
    static constexpr int MIN_WAVES = (CE <= 8) ? 4 : (CE <= 18) ? 3 : 1;
};

typedef double dvec2 __attribute__((ext_vector_type(2)));

// Single-use streams (B, DoF indexes, weights, committed state) are read with
// non-temporal loads so they do not evict u and r, which every color launch
// reads again, from L2 and the Infinity Cache.
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Loads the NB values of point `lane` of tile `tile` (pairs of doubles).
template <int SHAPE>
__device__ __forceinline__ void load_point_B(const double* __restrict__ Bg, int64_t tile, int lane, double* b) {
    using C = Cfg<SHAPE>;
    const dvec2* Bt = reinterpret_cast<const dvec2*>(Bg + tile * (int64_t)(C::NB * C::TP));
#pragma unroll
    for (int m2 = 0; m2 < C::NB / 2; m2++) {
        const dvec2 v = ld_stream(&Bt[m2 * C::TP + lane]);
        b[2 * m2] = v.x;
        b[2 * m2 + 1] = v.y;
    }
}

// strain_k = sum_j B[k][j] u_e[j], j ascending, starting from 0.0
template <int SHAPE>
__device__ __forceinline__ void strain_of(const double* b, const double* ue, double* e3) {
    using C = Cfg<SHAPE>;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < C::CE; j++) acc = acc + b[k * C::CE + j] * ue[j];
        e3[k] = acc;
    }
}

// Point partial of B^T x: g_j = B[0][j] x0 + B[1][j] x1 + B[2][j] x2 -> LDS row.
template <int SHAPE>
__device__ __forceinline__ void bt_partial(const double* b, const double* x, double* row) {
    using C = Cfg<SHAPE>;
#pragma unroll
    for (int j = 0; j < C::CE; j++) {
        double g = b[j] * x[0];
        g = g + b[C::CE + j] * x[1];
        g = g + b[2 * C::CE + j] * x[2];
        row[j] = g;
    }
}

// Sums the element's point partials (ip ascending), negates and adds into r.
template <int SHAPE>
__device__ __forceinline__ void reduce_scatter(const double* L, const GroupDev& g, int64_t tile, int lane,
                                               double* __restrict__ r) {
    using C = Cfg<SHAPE>;
    const int nv = tile_nvalid(g, tile);
    for (int task = lane; task < C::ET * C::CE; task += kWave) {
        const int e2 = task / C::CE;
        const int j = task - e2 * C::CE;
        const int64_t elem = tile * C::ET + e2;
        if (e2 < nv) {
            double f = L[(e2 * C::NIP) * C::CE + j];
#pragma unroll
            for (int ip = 1; ip < C::NIP; ip++) f = f + L[(e2 * C::NIP + ip) * C::CE + j];
            const int d = g.dof[elem * C::CE + j];
            r[d] = r[d] + (-f);
        }
    }
}

template <int SHAPE>
__global__ void __launch_bounds__(Cfg<SHAPE>::WPB * kWave)
k_multiply(GroupDev g, const double* __restrict__ u, double* __restrict__ strain, const int64_t* __restrict__ orig) {
    using C = Cfg<SHAPE>;
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t tile = (int64_t)blockIdx.x * C::WPB + wave;
    const int el = lane / C::NIP;
    const int64_t elem = tile * C::ET + el;
    if (!(tile < g.ntiles && lane < C::TP && el < tile_nvalid(g, tile))) return;
    double ue[C::CE];
    const int* d = g.dof + elem * C::CE;
#pragma unroll
    for (int j = 0; j < C::CE; j++) ue[j] = u[d[j]];
    double b[C::NB];
    load_point_B<SHAPE>(g.B, tile, lane, b);
    double e3[3];
    strain_of<SHAPE>(b, ue, e3);
    const int64_t o = orig[g.pt0 + tile * C::TP + lane];
    strain[3 * o + 0] = e3[0];
    strain[3 * o + 1] = e3[1];
    strain[3 * o + 2] = e3[2];
}

template <int SHAPE>
__global__ void __launch_bounds__(Cfg<SHAPE>::WPB * kWave)
k_transpose(GroupDev g, const double* __restrict__ x, double* __restrict__ r, const int64_t* __restrict__ orig) {
    using C = Cfg<SHAPE>;
    __shared__ double lds[C::WPB][C::TP * C::CE];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t tile = (int64_t)blockIdx.x * C::WPB + wave;
    const bool tile_ok = tile < g.ntiles;
    const int el = lane / C::NIP;
    const bool slot = tile_ok && lane < C::TP;
    const bool valid = slot && el < tile_nvalid(g, tile);
    double ws[3] = {0.0, 0.0, 0.0};
    if (valid) {
        const int64_t o = orig[g.pt0 + tile * C::TP + lane];
        ws[0] = x[3 * o + 0];
        ws[1] = x[3 * o + 1];
        ws[2] = x[3 * o + 2];
    }
    double* L = lds[wave];
    if (slot) {
        double b[C::NB];
        load_point_B<SHAPE>(g.B, tile, lane, b);
        bt_partial<SHAPE>(b, ws, L + lane * C::CE);
    }
    __syncthreads();
    if (tile_ok) reduce_scatter<SHAPE>(L, g, tile, lane, r);
}

// One wavefront per tile of a group; lane q < tp owns point pt0 + tile*tp + q.
template <bool DEP>
__global__ void __launch_bounds__(kWave)
k_sigma(GroupDev g, int tp, int64_t npad, const double* __restrict__ strain, double* __restrict__ wsigma,
        double* __restrict__ wdep, const int64_t* __restrict__ orig, PointState st, int commit,
        double* __restrict__ dep_soa) {
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    if (lane >= tp) return;
    const int64_t P = g.pt0 + tile * tp + lane;
    const int64_t o = orig[P];
    if (o < 0) return;
    const MatConst& m = tile_material(g, tile);
    const double e3[3] = {strain[3 * o], strain[3 * o + 1], strain[3 * o + 2]};
    double ep[4] = {st.ep0[P], st.ep1[P], st.ep2[P], st.ep3[P]};
    double alpha = st.alpha[P];
    int mt, br;
    double ws[3], d9[9];
    point_sigma<DEP>(e3, ep, alpha, mt, br, m, st.w[P], ws, d9);
    wsigma[3 * o + 0] = ws[0];
    wsigma[3 * o + 1] = ws[1];
    wsigma[3 * o + 2] = ws[2];
    if (DEP) {
        if (wdep) {
#pragma unroll
            for (int i = 0; i < 9; i++) wdep[9 * o + i] = d9[i];
        }
        if (dep_soa) {
#pragma unroll
            for (int i = 0; i < 9; i++) dep_soa[i * npad + P] = d9[i];
        }
    }
    if (commit) {
        st.ep0[P] = ep[0];
        st.ep1[P] = ep[1];
        st.ep2[P] = ep[2];
        st.ep3[P] = ep[3];
        st.alpha[P] = alpha;
        st.mtype[P] = (signed char)mt;
        st.branch[P] = (signed char)br;
    }
}

// y += sum_e B_e^T (w Dep) B_e v_e over one element group (one color): the
// matrix-free tangent of the Newton / CG solve.  Same tiling as k_residual,
// with the return map replaced by the stored 3x3 w*Dep of each point.
template <int SHAPE>
__global__ void __launch_bounds__(Cfg<SHAPE>::WPB * kWave)
k_tangent_diag(GroupDev g, double* __restrict__ diag, double* __restrict__ offd, const double* __restrict__ dep,
               int64_t npad) {
    using C = Cfg<SHAPE>;
    __shared__ double lds[C::WPB][C::TP * C::CE];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t tile = (int64_t)blockIdx.x * C::WPB + wave;
    const bool tile_ok = tile < g.ntiles;
    const int el = lane / C::NIP;
    const bool slot = tile_ok && lane < C::TP;
    const bool valid = slot && el < tile_nvalid(g, tile);
    double* L = lds[wave];
    double D[9];
    const int64_t P = g.pt0 + (tile_ok ? tile : 0) * C::TP + lane;
#pragma unroll
    for (int i = 0; i < 9; i++) D[i] = valid ? dep[i * npad + P] : 0.0;
    const dvec2* Bt = reinterpret_cast<const dvec2*>(g.B + (tile_ok ? tile : 0) * (int64_t)(C::NB * C::TP));
    auto col = [&](int j, double* bj) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int mm = k * C::CE + j;
            const dvec2 pr = Bt[(mm >> 1) * C::TP + lane];
            bj[k] = (mm & 1) ? pr.y : pr.x;
        }
    };
    // pass 0: the diagonal b_j^T D b_j; pass 1 (offd set): the symmetric
    // coupling of the node's two columns, (b_j^T D b_j' + b_j'^T D b_j) / 2
    // with j' = j ^ 1, into offd at column j's DoF
    for (int pass = 0; pass < (offd ? 2 : 1); pass++) {
        if (slot) {
            for (int j = 0; j < C::CE; j++) {
                double bj[3], bk[3];
                col(j, bj);
                col(pass ? (j ^ 1) : j, bk);
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int l = 0; l < 3; l++) {
                        s1 = s1 + bj[k] * D[3 * k + l] * bk[l];
                        s2 = s2 + bk[k] * D[3 * k + l] * bj[l];
                    }
                L[lane * C::CE + j] = pass ? -0.5 * (s1 + s2) : -s1;
            }
        }
        __syncthreads();
        if (tile_ok) reduce_scatter<SHAPE>(L, g, tile, lane, pass ? offd : diag);
        __syncthreads();
    }
}

__global__ void k_copy(int64_t n, const double* __restrict__ src, double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_fill(int64_t n, double* __restrict__ dst, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

template <int SHAPE>
dim3 tile_grid(const GroupDev& g) {
    return dim3((unsigned)((g.ntiles + Cfg<SHAPE>::WPB - 1) / Cfg<SHAPE>::WPB));
}

}  // namespace

void launch_multiply(int shape, const GroupDev& g, const double* u, double* strain, const int64_t* orig,
                     hipStream_t s) {
    if (g.ntiles == 0) return;
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL(k_multiply<S>, tile_grid<S>(g), dim3(Cfg<S>::WPB * kWave), 0, s, g, u, strain, orig);
    });
}

void launch_transpose(int shape, const GroupDev& g, const double* x, double* r, const int64_t* orig, hipStream_t s) {
    if (g.ntiles == 0) return;
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL(k_transpose<S>, tile_grid<S>(g), dim3(Cfg<S>::WPB * kWave), 0, s, g, x, r, orig);
    });
}

void launch_sigma(bool dep, bool commit, int shape, const GroupDev& g, int64_t npad, const double* strain,
                  double* wsigma, double* wdep, const int64_t* orig, const PointState& st, hipStream_t s,
                  double* dep_soa) {
    if (g.ntiles == 0) return;
    const int tp = shape_tp(shape);
    const dim3 grid((unsigned)g.ntiles);
    if (dep)
        hipLaunchKernelGGL(k_sigma<true>, grid, dim3(kWave), 0, s, g, tp, npad, strain, wsigma, wdep, orig, st,
                           commit ? 1 : 0, dep_soa);
    else
        hipLaunchKernelGGL(k_sigma<false>, grid, dim3(kWave), 0, s, g, tp, npad, strain, wsigma, wdep, orig, st,
                           commit ? 1 : 0, dep_soa);
}

void launch_tangent_diag(int shape, const GroupDev& g, double* diag, double* offd, const double* dep, int64_t npad,
                         hipStream_t s) {
    if (g.ntiles == 0) return;
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL(k_tangent_diag<S>, tile_grid<S>(g), dim3(Cfg<S>::WPB * kWave), 0, s, g, diag, offd, dep,
                           npad);
    });
}

void launch_copy(int64_t n, const double* src, double* dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, dst);
}

void launch_fill(int64_t n, double* dst, double v, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, dst, v);
}

}  // namespace ipf
