// ipf_kernels_fmt.hip -- the fused residual (the hot path) and the tangent
// product, for every storage format of B (see ipf_kernels.h, BFormat):
//
//   FMT_D64  dense FP64: every stored value of the reference's block.
//   FMT_C64  compact FP64: per point only (dN_a/dx, dN_a/dy), the values the
//            reference writes into its B block (TPZNumericalIntegrator.cpp:
//            371-381; the other entries of the 3 x 2n block are zeros or
//            copies).  The products keep the dense order term by term and
//            only skip exact zeros, so strains, stresses and forces equal the
//            dense FP64 results (up to the sign of a zero).
//   FMT_D32  dense FP32: the stored block in FP32, products accumulated in FP32.
//   FMT_C32  compact FP32.
//   FMT_G64  geometric FP64: the compact FP64 values and the weight computed
//            in the kernel from the element's corners and the shape table
//            (ipf_geometry.h) -- 8 doubles per element instead of 2 n_sh
//            doubles (+ 1 weight) per point, the same bits as the stored
//            FP64 forms built by the device setup from the same geometry.
//
// FP32 formats compute B.u and B^T(w sigma) in FP32 with FP32 accumulation;
// the return map, the assembly into r and everything else stay FP64.
#include "ipf_geometry.h"
#include "ipf_kernels.h"
#include "ipf_patch.h"

namespace ipf {

namespace {

constexpr int kWave = 64;

// Workgroup -> tile (patch) of a launch of n workgroups.  The hardware deals
// consecutive workgroups of a dispatch to the 8 XCDs in turn, each with its
// own L2; this mapping gives XCD x one contiguous range of tiles -- ranges of
// q + 1 tiles for the first n mod 8 XCDs, q = n / 8 for the others -- so
// that neighbouring tiles, whose DoF, state and geometry arrays share cache
// lines at the tile boundaries, are read through the same L2.  A bijection
// of [0, n); the results do not depend on it (no two tiles of a launch write
// the same value).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
    const int64_t q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

template <int FMT>
struct FmtTraits;
template <>
struct FmtTraits<FMT_D64> { using T = double; static constexpr bool COMPACT = false, GEOM = false; static constexpr int VEC = 2; };
template <>
struct FmtTraits<FMT_C64> { using T = double; static constexpr bool COMPACT = true, GEOM = false; static constexpr int VEC = 2; };
template <>
struct FmtTraits<FMT_D32> { using T = float; static constexpr bool COMPACT = false, GEOM = false; static constexpr int VEC = 4; };
template <>
struct FmtTraits<FMT_C32> { using T = float; static constexpr bool COMPACT = true, GEOM = false; static constexpr int VEC = 4; };
template <>
struct FmtTraits<FMT_G64> { using T = double; static constexpr bool COMPACT = true, GEOM = true; static constexpr int VEC = 2; };
template <>
struct FmtTraits<FMT_G32> { using T = float; static constexpr bool COMPACT = true, GEOM = true; static constexpr int VEC = 4; };

template <int SHAPE, int FMT>
struct FC {
    using F = FmtTraits<FMT>;
    using T = typename F::T;
    static constexpr int NIP = shape_nip(SHAPE);
    static constexpr int CE = shape_ce(SHAPE);
    static constexpr int ET = shape_et(SHAPE);
    static constexpr int TP = shape_tp(SHAPE);
    static constexpr int NV = format_values(SHAPE, FMT);  // stored values per point (padded)
    static constexpr int VEC = F::VEC;
    static constexpr int MIN_WAVES = F::COMPACT ? 4 : ((CE <= 8) ? 4 : (CE <= 18) ? 3 : 1);
};

template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

typedef double dv2 __attribute__((ext_vector_type(2)));
typedef float fv4 __attribute__((ext_vector_type(4)));

template <int SHAPE, int FMT>
__device__ __forceinline__ void load_values(const void* __restrict__ Bg, int64_t tile, int lane,
                                            typename FC<SHAPE, FMT>::T* b) {
    using C = FC<SHAPE, FMT>;
    if constexpr (C::VEC == 2) {
        const dv2* Bt = reinterpret_cast<const dv2*>(Bg) + tile * (int64_t)(C::NV / 2 * C::TP);
#pragma unroll
        for (int m = 0; m < C::NV / 2; m++) {
            const dv2 v = ld_stream(&Bt[m * C::TP + lane]);
            b[2 * m] = v.x;
            b[2 * m + 1] = v.y;
        }
    } else {
        const fv4* Bt = reinterpret_cast<const fv4*>(Bg) + tile * (int64_t)(C::NV / 4 * C::TP);
#pragma unroll
        for (int m = 0; m < C::NV / 4; m++) {
            const fv4 v = ld_stream(&Bt[m * C::TP + lane]);
            b[4 * m] = v.x;
            b[4 * m + 1] = v.y;
            b[4 * m + 2] = v.z;
            b[4 * m + 3] = v.w;
        }
    }
}

// Geometric format: the compact values (dN_a/dx, dN_a/dy) of the lane's point
// from its element's corners (Bg: 8 doubles per element slot of the group)
// and the shape table, computed in FP64 and stored as T (FP32: rounded);
// returns the weight w_q |det J|.
template <int SHAPE, class T>
__device__ __forceinline__ double geom_values(const double* __restrict__ Bg, const double* __restrict__ tab,
                                              int64_t tile, int lane, T* b) {
    using C = FC<SHAPE, FMT_G64>;
    const int el = lane / C::NIP, ip = lane - el * C::NIP;
    const dv2* Xv = reinterpret_cast<const dv2*>(Bg) + (tile * C::ET + el) * 4;
    double X[8];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const dv2 v = ld_stream(&Xv[c]);
        X[2 * c] = v.x;
        X[2 * c + 1] = v.y;
    }
    const double* xi = tab + C::NIP * C::CE;
    const QuadMap m = quad_map(X, xi[ip], xi[C::NIP + ip]);
    const dv2* dphi = reinterpret_cast<const dv2*>(tab) + ip * (C::CE / 2);
#pragma unroll
    for (int a = 0; a < C::CE / 2; a++) {
        const dv2 d = dphi[a];
        double gx, gy;
        quad_grad(m, d.x, d.y, gx, gy);
        b[2 * a] = T(gx);
        b[2 * a + 1] = T(gy);
    }
    return quad_weight(m, xi[2 * C::NIP + ip]);
}

// strain in the accumulation type T.  Dense: rows summed over columns j
// ascending.  Compact: the same sums with the zero terms left out.
template <int SHAPE, int FMT>
__device__ __forceinline__ void strain_fmt(const typename FC<SHAPE, FMT>::T* b, const double* ue, double* e3) {
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    if constexpr (FmtTraits<FMT>::COMPACT) {
        T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
        for (int a = 0; a < C::CE / 2; a++) {
            const T dx = b[2 * a], dy = b[2 * a + 1];
            const T ux = T(ue[2 * a]), uy = T(ue[2 * a + 1]);
            a0 = a0 + dx * ux;
            a1 = a1 + dy * ux;
            a1 = a1 + dx * uy;
            a2 = a2 + dy * uy;
        }
        e3[0] = double(a0);
        e3[1] = double(a1);
        e3[2] = double(a2);
    } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < C::CE; j++) acc = acc + b[k * C::CE + j] * T(ue[j]);
            e3[k] = double(acc);
        }
    }
}

// Columns [J0, J1) of one point's B^T (w sigma) partials into row[j - J0].
// Dense: rows summed in order per column; compact: the same sums without the
// zero terms.  The residual and tangent kernels stage half the columns at a
// time, which halves the LDS a wave holds (one wave per workgroup) so that
// registers, not LDS, bound the waves per CU.
template <int SHAPE, int FMT, int J0, int J1>
__device__ __forceinline__ void bt_cols(const typename FC<SHAPE, FMT>::T* b, const double* x,
                                        typename FC<SHAPE, FMT>::T* row) {
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    const T x0 = T(x[0]), x1 = T(x[1]), x2 = T(x[2]);
#pragma unroll
    for (int j = J0; j < J1; j++) {
        T g;
        if constexpr (FmtTraits<FMT>::COMPACT) {
            const int a = j / 2;
            const T dx = b[2 * a], dy = b[2 * a + 1];
            if ((j & 1) == 0) {
                g = dx * x0;
                g = g + dy * x1;
            } else {
                g = dx * x1;
                g = g + dy * x2;
            }
        } else {
            g = b[j] * x0;
            g = g + b[C::CE + j] * x1;
            g = g + b[2 * C::CE + j] * x2;
        }
        row[j - J0] = g;
    }
}

// One lane's column sum: the partials L of element e's points in column jj of
// the staged half (W columns), summed in point order and added negated to
// r[d] (TPZNumericalIntegrator::MultiplyTranspose's y = -B^T x); rv = r[d],
// read by the caller ahead of time.  Each lane owns one column sum per half
// (ET * W <= 64 for every shape), and the color's elements share no DoF, so
// no atomics.
template <int SHAPE, int FMT, int W>
__device__ __forceinline__ void reduce_lane(const typename FC<SHAPE, FMT>::T* L, int e, int jj, int d, double rv,
                                            double* __restrict__ r) {
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    T f = L[(e * C::NIP) * W + jj];
#pragma unroll
    for (int ip = 1; ip < C::NIP; ip++) f = f + L[(e * C::NIP + ip) * W + jj];
    r[d] = rv + (-double(f));
}

// Memory requests leave each wave early so that their latency overlaps the
// geometry and the return map: the tile's DoF indexes and the point's
// committed state first, the u values as soon as the indexes arrive (one
// round trip before the gather completes, not two per gather round), and the
// r values of the lane's column sums right after the gather (R_EARLY) or
// both right after the return map -- the color's elements share no DoF, so
// no other wave of the launch writes them.  R_EARLY costs 6-7 VGPRs; it is on
// at Q3 (LDS already limits the waves per SIMD to 4) and for FAST at Q2 (94
// VGPRs, still 5 waves per SIMD; a wave whose points are all certified
// reaches the scatter soon after the gather: 0.184 against 0.196 ms per
// 800x800 Q2 residual on MI355X), off for the exact Q2 kernel (no gain
// measured).
//
// FAST (never with COMMIT): a point whose trial stress passes
// certified_elastic takes its trial stress and skips the spectral
// decomposition; a wavefront whose points all pass skips the return map's
// instructions altogether.  The other points take the exact return map.
template <int SHAPE, int FMT, bool COMMIT, bool PAIR, bool FAST>
__global__ void __launch_bounds__(kWave, (FC<SHAPE, FMT>::MIN_WAVES))
k_residual_fmt(GroupDev g, const void* __restrict__ Bv, const double* __restrict__ u, double* __restrict__ r,
               const double* __restrict__ flin, PointState st) {
    static_assert(!(FAST && COMMIT), "the certified elastic path commits nothing");
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    constexpr int NSH = C::CE / 2;
    // staged halves of the columns: whole nodes with PAIR (the first half
    // holds the first ceil(NSH / 2) nodes), else CE / 2 columns each
    constexpr int H = PAIR ? 2 * ((NSH + 1) / 2) : C::CE / 2;
    constexpr int H2 = C::CE - H;
    static_assert(C::ET * NSH <= kWave, "one column sum (PAIR: node) per lane and half");
    constexpr bool R_EARLY = (SHAPE == SHAPE_Q3 || (SHAPE == SHAPE_Q2 && FAST)) && !COMMIT;
    constexpr int NE = PAIR ? C::ET * NSH : C::ET * C::CE;  // index entries per tile
    __shared__ T lds[C::TP * H];
    __shared__ double us[C::ET * C::CE];
    __shared__ int dsx[NE];
    const int lane = threadIdx.x;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int el = lane / C::NIP;
    const bool slot = lane < C::TP;
    const int nv = tile_nvalid(g, tile);
    const bool valid = slot && el < nv;
    // The tile's DoF (PAIR: node) indexes, the point's state, then the u
    // values the indexes address -- loads outside branches, so that each use
    // waits only for its own requests.  Every address is in bounds: padding
    // slots of a tile hold DoF 0 (PAIR: node DoFs 0 and 1) and points of
    // their own, lanes past the tile's points read its first point, and what
    // invalid lanes load is never used.
    constexpr int NG = (NE + kWave - 1) / kWave;
    static_assert(NE >= kWave || PAIR, "the first gather round fills the wave");
    int dl[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int t = min(lane + k * kWave, NE - 1);
        dl[k] = ld_stream(&(PAIR ? g.node : g.dof)[tile * (int64_t)NE + t]);
    }
    const int64_t P = g.pt0 + tile * C::TP + (slot ? lane : 0);
    double ep[4] = {ld_stream(&st.ep0[P]), ld_stream(&st.ep1[P]), ld_stream(&st.ep2[P]), ld_stream(&st.ep3[P])};
    // alpha only accumulates the plastic increment (H = 0: it does not enter
    // the stress); without a commit it is not needed
    double alpha = COMMIT ? ld_stream(&st.alpha[P]) : 0.0;
    const double wst = FmtTraits<FMT>::GEOM ? 0.0 : ld_stream(&st.w[P]);
    double ul[NG], vl[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        ul[k] = u[dl[k] & kDofMask];
        vl[k] = PAIR ? u[(dl[k] & kDofMask) + 1] : 0.0;
    }
    T b[C::NV];
    double wgeo = 0.0;
    if constexpr (FmtTraits<FMT>::GEOM) {
        if (valid) {
            wgeo = geom_values<SHAPE>(static_cast<const double*>(Bv), g.tab, tile, lane, b);
        } else {
#pragma unroll
            for (int k = 0; k < C::NV; k++) b[k] = T(0);
        }
    } else if (slot) {
        load_values<SHAPE, FMT>(Bv, tile, lane, b);
    } else {
#pragma unroll
        for (int k = 0; k < C::NV; k++) b[k] = T(0);
    }
    // the tile's indexes and u values, gathered once per element into LDS
    // (the lanes of an element read them from there: no per-lane copy of the
    // element's u in registers, no repeated gathers)
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int t = lane + k * kWave;
        if ((k == 0 && !PAIR) || t < NE) {
            dsx[t] = dl[k];
            if (PAIR) {
                us[2 * t] = ul[k];
                us[2 * t + 1] = vl[k];
            } else {
                us[t] = ul[k];
            }
        }
    }
    __syncthreads();
    // The lane's column sums and where their r values start (an entry flagged
    // kFirstWrite, flin set, starts from f_lin instead of r).  PAIR: node ra
    // of element re, its x and y columns, in the half that holds the node;
    // else columns rj and H + rj of element re.
    const int re = lane / (PAIR ? NSH : H), rj = lane - re * (PAIR ? NSH : H);
    const bool task = lane < C::ET * (PAIR ? NSH : H) && re < nv;
    const bool first_half = !PAIR || rj < H / 2;
    auto start = [&](int de, int off) -> double { return (de < 0 ? flin : r)[(de & kDofMask) + off]; };
    auto request_r = [&](double& r0, double& r1) {
        if (PAIR) {
            const int de = dsx[re * NSH + rj];
            r0 = start(de, 0);
            r1 = start(de, 1);
        } else {
            r0 = start(dsx[re * C::CE + rj], 0);
            r1 = start(dsx[re * C::CE + H + rj], 0);
        }
    };
    double r0 = 0.0, r1 = 0.0;
    if (R_EARLY && task) request_r(r0, r1);
    double e3[3] = {0.0, 0.0, 0.0};
    if (valid) strain_fmt<SHAPE, FMT>(b, us + el * C::CE, e3);
    double ws[3] = {0.0, 0.0, 0.0};
    if (valid) {
        int mt, br;
        const double w = FmtTraits<FMT>::GEOM ? wgeo : wst;
        const MatConst& mc = tile_material(g, tile);
        if (!(FAST && certified_elastic(e3, ep, mc, w, ws))) point_sigma<false>(e3, ep, alpha, mt, br, mc, w, ws, nullptr);
        if (COMMIT) {
            st.ep0[P] = ep[0];
            st.ep1[P] = ep[1];
            st.ep2[P] = ep[2];
            st.ep3[P] = ep[3];
            st.alpha[P] = alpha;
            st.mtype[P] = (signed char)mt;
            st.branch[P] = (signed char)br;
        }
    }
    // otherwise both r values now, in flight during the first half's products
    if (!R_EARLY && task) request_r(r0, r1);
    if (slot) bt_cols<SHAPE, FMT, 0, H>(b, ws, lds + lane * H);
    __syncthreads();
    if (PAIR) {
        if (task && first_half) {
            const int d = dsx[re * NSH + rj] & kDofMask;
            reduce_lane<SHAPE, FMT, H>(lds, re, 2 * rj, d, r0, r);
            reduce_lane<SHAPE, FMT, H>(lds, re, 2 * rj + 1, d + 1, r1, r);
        }
    } else if (task) {
        reduce_lane<SHAPE, FMT, H>(lds, re, rj, dsx[re * C::CE + rj] & kDofMask, r0, r);
    }
    __syncthreads();
    if (slot) bt_cols<SHAPE, FMT, H, C::CE>(b, ws, lds + lane * H2);
    __syncthreads();
    if (PAIR) {
        if (task && !first_half) {
            const int d = dsx[re * NSH + rj] & kDofMask;
            reduce_lane<SHAPE, FMT, H2>(lds, re, 2 * rj - H, d, r0, r);
            reduce_lane<SHAPE, FMT, H2>(lds, re, 2 * rj - H + 1, d + 1, r1, r);
        }
    } else if (task) {
        reduce_lane<SHAPE, FMT, H2>(lds, re, rj, dsx[re * C::CE + H + rj] & kDofMask, r1, r);
    }
}

// y += sum_e B_e^T (w Dep) B_e v_e with B in format FMT (FP32 formats: FP32
// products and accumulation, the matvec of the inner CG solves).  With
// zero_first, index entries flagged kFirstWrite start from 0 instead of y[d]
// (y = K v then needs no clear before the launches).
template <int SHAPE, int FMT, bool PAIR>
__global__ void __launch_bounds__(kWave, (FC<SHAPE, FMT>::MIN_WAVES))
k_tangent_fmt(GroupDev g, const void* __restrict__ Bv, const double* __restrict__ v, double* __restrict__ y,
              const double* __restrict__ dep, const float* __restrict__ dep32, int64_t npad, bool zero_first) {
    // FMT_G32: the whole element product in FP32 (w Dep from its FP32 copy)
    constexpr bool F32 = FMT == FMT_G32;
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    constexpr int NSH = C::CE / 2;
    // staged halves of the columns as in k_residual_fmt
    constexpr int H = PAIR ? 2 * ((NSH + 1) / 2) : C::CE / 2;
    constexpr int H2 = C::CE - H;
    static_assert(C::ET * NSH <= kWave, "one column sum (PAIR: node) per lane and half");
    constexpr int NE = PAIR ? C::ET * NSH : C::ET * C::CE;
    __shared__ T lds[C::TP * H];
    __shared__ double vs[C::ET * C::CE];
    __shared__ int dsx[NE];
    const int lane = threadIdx.x;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int el = lane / C::NIP;
    const bool slot = lane < C::TP;
    const int nv = tile_nvalid(g, tile);
    const bool valid = slot && el < nv;
    // DoF (PAIR: node) indexes with their flags, then the v values they
    // address, outside branches (the addresses are in bounds as in
    // k_residual_fmt)
    constexpr int NG = (NE + kWave - 1) / kWave;
    static_assert(NE >= kWave || PAIR, "the first gather round fills the wave");
    int dl[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int t = min(lane + k * kWave, NE - 1);
        dl[k] = ld_stream(&(PAIR ? g.node : g.dof)[tile * (int64_t)NE + t]);
    }
    double vl[NG], vl2[NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        vl[k] = v[dl[k] & kDofMask];
        vl2[k] = PAIR ? v[(dl[k] & kDofMask) + 1] : 0.0;
    }
    T b[C::NV];
    if constexpr (FmtTraits<FMT>::GEOM) {
        if (valid) {
            (void)geom_values<SHAPE>(static_cast<const double*>(Bv), g.tab, tile, lane, b);
        } else {
#pragma unroll
            for (int k = 0; k < C::NV; k++) b[k] = T(0);
        }
    } else if (slot) {
        load_values<SHAPE, FMT>(Bv, tile, lane, b);
    } else {
#pragma unroll
        for (int k = 0; k < C::NV; k++) b[k] = T(0);
    }
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const int t = lane + k * kWave;
        if ((k == 0 && !PAIR) || t < NE) {
            dsx[t] = dl[k];
            if (PAIR) {
                vs[2 * t] = vl[k];
                vs[2 * t + 1] = vl2[k];
            } else {
                vs[t] = vl[k];
            }
        }
    }
    __syncthreads();
    // the point's 9 tangent values and the lane's two y values (the color's
    // elements share no DoF), requested before the strain -- except next to
    // dense FP64 B of up to 18 columns, whose 3 CE values already fill the
    // registers (measured: spills, slower); there they follow the strain.
    // PAIR: node rj of element re, its x and y columns; else columns rj and
    // H + rj of element re.
    constexpr bool EARLY = !(FMT == FMT_D64 && C::CE <= 18);
    const int64_t P = g.pt0 + tile * C::TP + (slot ? lane : 0);
    const int re = lane / (PAIR ? NSH : H), rj = lane - re * (PAIR ? NSH : H);
    const bool task = lane < C::ET * (PAIR ? NSH : H) && re < nv;
    const bool first_half = !PAIR || rj < H / 2;
    double dp[9], y0 = 0.0, y1 = 0.0;
    float dpf[9];
    auto start = [&](int de, int off) -> double {
        return (zero_first && de < 0) ? 0.0 : y[(de & kDofMask) + off];
    };
    auto request = [&]() {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            if constexpr (F32)
                dpf[k] = ld_stream(&dep32[k * npad + P]);
            else
                dp[k] = ld_stream(&dep[k * npad + P]);
        }
        if (task) {
            if (PAIR) {
                const int de = dsx[re * NSH + rj];
                y0 = start(de, 0);
                y1 = start(de, 1);
            } else {
                y0 = start(dsx[re * C::CE + rj], 0);
                y1 = start(dsx[re * C::CE + H + rj], 0);
            }
        }
    };
    if constexpr (EARLY) request();
    double e3[3] = {0.0, 0.0, 0.0};
    if (valid) strain_fmt<SHAPE, FMT>(b, vs + el * C::CE, e3);
    if constexpr (!EARLY) request();
    double x[3] = {0.0, 0.0, 0.0};
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if constexpr (F32) {
                float a = dpf[3 * k] * float(e3[0]);
                a = a + dpf[3 * k + 1] * float(e3[1]);
                a = a + dpf[3 * k + 2] * float(e3[2]);
                x[k] = -double(a);
            } else {
                double a = dp[3 * k] * e3[0];
                a = a + dp[3 * k + 1] * e3[1];
                a = a + dp[3 * k + 2] * e3[2];
                x[k] = -a;  // reduce_lane adds -(sum); y accumulates +K v
            }
        }
    }
    if (slot) bt_cols<SHAPE, FMT, 0, H>(b, x, lds + lane * H);
    __syncthreads();
    if (PAIR) {
        if (task && first_half) {
            const int d = dsx[re * NSH + rj] & kDofMask;
            reduce_lane<SHAPE, FMT, H>(lds, re, 2 * rj, d, y0, y);
            reduce_lane<SHAPE, FMT, H>(lds, re, 2 * rj + 1, d + 1, y1, y);
        }
    } else if (task) {
        reduce_lane<SHAPE, FMT, H>(lds, re, rj, dsx[re * C::CE + rj] & kDofMask, y0, y);
    }
    __syncthreads();
    if (slot) bt_cols<SHAPE, FMT, H, C::CE>(b, x, lds + lane * H2);
    __syncthreads();
    if (PAIR) {
        if (task && !first_half) {
            const int d = dsx[re * NSH + rj] & kDofMask;
            reduce_lane<SHAPE, FMT, H2>(lds, re, 2 * rj - H, d, y0, y);
            reduce_lane<SHAPE, FMT, H2>(lds, re, 2 * rj - H + 1, d + 1, y1, y);
        }
    } else if (task) {
        reduce_lane<SHAPE, FMT, H2>(lds, re, rj, dsx[re * C::CE + H + rj] & kDofMask, y1, y);
    }
}

// Builds format FMT from the dense FP64 tile layout of one group (device side).
template <int SHAPE, int FMT>
__global__ void k_convert_B(const double* __restrict__ dense, void* __restrict__ out, int64_t ntiles) {
    using C = FC<SHAPE, FMT>;
    using T = typename C::T;
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    if (tile >= ntiles || lane >= C::TP) return;
    constexpr int NB = 3 * C::CE;
    const double* src = dense + tile * (int64_t)(NB * C::TP);
    T* dst = reinterpret_cast<T*>(out) + tile * (int64_t)(C::NV * C::TP);
    auto dense_at = [&](int mm) -> double { return src[((int64_t)(mm >> 1) * C::TP + lane) * 2 + (mm & 1)]; };
    for (int k = 0; k < C::NV; k++) {
        double val = 0.0;
        if (FmtTraits<FMT>::COMPACT) {
            if (k < C::CE) {
                const int a = k >> 1;
                // dx_a = B[xx][2a], dy_a = B[yy][2a+1]
                val = (k & 1) ? dense_at(2 * C::CE + 2 * a + 1) : dense_at(2 * a);
            }
        } else if (k < NB) {
            val = dense_at(k);
        }
        dst[((int64_t)(k / C::VEC) * C::TP + lane) * C::VEC + (k % C::VEC)] = T(val);
    }
}

// Device setup of one group: dense FP64 tiles of B and the weights from the
// corners (8 doubles per element slot) and the shape table g.tab.  Every value
// of a tile is written, zeros and padding slots included.
template <int SHAPE>
__global__ void k_build_B(GroupDev g, const double* __restrict__ corners, double* __restrict__ dense,
                          double* __restrict__ w) {
    using C = FC<SHAPE, FMT_D64>;
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    if (tile >= g.ntiles || lane >= C::TP) return;
    const int el = lane / C::NIP;
    const bool valid = el < tile_nvalid(g, tile);
    double cb[C::CE];
    double wt = 0.0;
    if (valid) {
        wt = geom_values<SHAPE>(corners, g.tab, tile, lane, cb);
    } else {
#pragma unroll
        for (int k = 0; k < C::CE; k++) cb[k] = 0.0;
    }
    double* dst = dense + tile * (int64_t)(3 * C::CE * C::TP);
    // value m = k * CE + j: row 3 ip + k, column j = 2 a + c of the block
#pragma unroll
    for (int m = 0; m < 3 * C::CE; m++) {
        const int k = m / C::CE, j = m % C::CE, a = j >> 1, c = j & 1;
        double v = 0.0;
        if (k == 0 && c == 0) v = cb[2 * a];          // xx: dN_a/dx
        if (k == 1) v = c ? cb[2 * a] : cb[2 * a + 1];  // xy: dN_a/dy (x column), dN_a/dx (y column)
        if (k == 2 && c == 1) v = cb[2 * a + 1];      // yy: dN_a/dy
        dst[((int64_t)(m >> 1) * C::TP + lane) * 2 + (m & 1)] = v;
    }
    w[g.pt0 + tile * C::TP + lane] = wt;
}

// Orders the LDS accesses of one wavefront (its lanes exchange values
// through LDS; LDS instructions of a wavefront execute in order, so only
// the compiler must not move them across this point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The non-committing residual over the patch layout (ipf_patch.h): one
// workgroup of NW wavefronts per patch.  Per point the
// colored kernel's arithmetic (geometric FP64 B, compact strain sums, return
// map or FAST's certified elastic path, B^T products); then per element the
// column sums over its points (point order) into LDS, and per patch node
// the sum of its elements' contributions (ascending element slot, then
// local node), written to r = f_lin + (-sum) where the patch owns the node,
// else to the node's partial slot for the fix-up pass.  A patch has at most
// one node per thread: thread t owns patch node t, whose DoF, partial slot,
// first contributions and u values it requests at the start, so that their
// latency hides behind the points' work; the slots' local node indexes go to
// LDS.  Two workgroup barriers: the nodes' u in LDS, the elements' sums in LDS.
template <int SHAPE, bool FAST, int NW>
__global__ void __launch_bounds__(kWave * NW)
k_residual_patch(PatchDev p, const double* __restrict__ u, double* __restrict__ r, const double* __restrict__ flin,
                 double* __restrict__ partial) {
    using C = FC<SHAPE, FMT_G64>;
    constexpr int ET = C::ET, NIP = C::NIP, CE = C::CE, TP = C::TP;
    constexpr int NSH = CE / 2, PE = NW * ET, MAXN = patch_max_nodes(SHAPE, NW), NT = NW * kWave;
    constexpr int H = CE / 2, H2 = CE - H;
    constexpr int MAXC = 4;  // contributions requested up front; more are read at the end
    static_assert(ET * H <= kWave && ET * H2 <= kWave, "one column sum per lane and half");
    static_assert(MAXN <= NT, "one patch node per thread");
    __shared__ double us[2 * MAXN];
    __shared__ double es[PE * CE];
    __shared__ double stg[NW * TP * H];
    __shared__ unsigned short lns[PE * NSH];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid - w * kWave;
    // XCD-contiguous patch ranges for the exact return map (800x800 Q2:
    // 0.274 -> 0.262 ms per residual); the certified-elastic variant keeps the
    // dispatch order, which measured 12-15 % faster for it at Q2 and Q3 than
    // the contiguous ranges
    const int64_t pa = FAST ? (int64_t)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
    const int64_t n0 = p.node_off[pa];
    const int nn = (int)(p.node_off[pa + 1] - n0);
    const int64_t tile = pa * NW + w;
    const int tm = p.tmeta[tile];
    const int nv = tm & 255;
    const int el = lane / NIP;
    const bool slot = lane < TP;
    const bool valid = slot && el < nv;
    // the thread's patch node: DoF, partial slot, contributions, u
    const bool owner = tid < nn;
    const int64_t k = n0 + (owner ? tid : 0);
    const int d = p.pnode[k], ps = p.pslot[k];
    const int c0 = p.cbeg[k], c1 = owner ? p.cbeg[k + 1] : c0;
    const double ux = u[d], uy = u[d + 1];
    int cv[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; j++) cv[j] = p.contrib[min(c0 + j, max(c1 - 1, c0))];
    // the point's plastic strain (padding points hold zeros), the slots' local nodes
    const int64_t P = tile * TP + (slot ? lane : 0);
    double ep[4] = {ld_stream(&p.ep0[P]), ld_stream(&p.ep1[P]), ld_stream(&p.ep2[P]), ld_stream(&p.ep3[P])};
    for (int i = tid; i < PE * NSH; i += NT) lns[i] = p.lnode[pa * (PE * NSH) + i];
    double b[C::NV];
    double wgeo = 0.0;
    if (valid) {
        wgeo = geom_values<SHAPE>(p.geo, p.tab, tile, lane, b);
    } else {
#pragma unroll
        for (int q = 0; q < C::NV; q++) b[q] = 0.0;
    }
    if (owner) {
        us[2 * tid] = ux;
        us[2 * tid + 1] = uy;
    }
    __syncthreads();
    double ws[3] = {0.0, 0.0, 0.0};
    if (valid) {
        const unsigned short* ln = lns + (w * ET + el) * NSH;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int a = 0; a < NSH; a++) {
            const double dx = b[2 * a], dy = b[2 * a + 1];
            const int l = ln[a];
            const double vx = us[2 * l], vy = us[2 * l + 1];
            a0 = a0 + dx * vx;
            a1 = a1 + dy * vx;
            a1 = a1 + dx * vy;
            a2 = a2 + dy * vy;
        }
        const double e3[3] = {a0, a1, a2};
        const MatConst& mc = p.mats[tm >> 8];
        if (!(FAST && certified_elastic(e3, ep, mc, wgeo, ws))) {
            double alpha = 0.0;
            int mt, br;
            point_sigma<false>(e3, ep, alpha, mt, br, mc, wgeo, ws, nullptr);
        }
    }
    // element column sums, half the columns at a time through the wave's own staging
    double* L = stg + w * TP * H;
    const int re = lane / H, rj = lane - re * H;
    if (slot) bt_cols<SHAPE, FMT_G64, 0, H>(b, ws, L + lane * H);
    wave_lds_sync();
    if (lane < ET * H && re < nv) {
        double f = L[(re * NIP) * H + rj];
#pragma unroll
        for (int ip = 1; ip < NIP; ip++) f = f + L[(re * NIP + ip) * H + rj];
        es[(w * ET + re) * CE + rj] = f;
    }
    wave_lds_sync();
    const int re2 = lane / H2, rj2 = lane - re2 * H2;
    if (slot) bt_cols<SHAPE, FMT_G64, H, CE>(b, ws, L + lane * H2);
    wave_lds_sync();
    if (lane < ET * H2 && re2 < nv) {
        double f = L[(re2 * NIP) * H2 + rj2];
#pragma unroll
        for (int ip = 1; ip < NIP; ip++) f = f + L[(re2 * NIP + ip) * H2 + rj2];
        es[(w * ET + re2) * CE + H + rj2] = f;
    }
    __syncthreads();
    // node sums: owned nodes straight into r (never read), shared ones to their partial slot
    if (owner) {
        double fx = 0.0, fy = 0.0;
#pragma unroll
        for (int j = 0; j < MAXC; j++)
            if (c0 + j < c1) {
                const int v = cv[j], s_ = v / NSH, a = v - s_ * NSH;
                fx = fx + es[s_ * CE + 2 * a];
                fy = fy + es[s_ * CE + 2 * a + 1];
            }
        for (int c = c0 + MAXC; c < c1; c++) {
            const int v = p.contrib[c], s_ = v / NSH, a = v - s_ * NSH;
            fx = fx + es[s_ * CE + 2 * a];
            fy = fy + es[s_ * CE + 2 * a + 1];
        }
        if (ps < 0) {
            r[d] = flin[d] + (-fx);
            r[d + 1] = flin[d + 1] + (-fy);
        } else {
            partial[2 * (int64_t)ps] = fx;
            partial[2 * (int64_t)ps + 1] = fy;
        }
    }
}

template <int SHAPE, int FMT>
void residual_fmt_shape(bool commit, bool fast, const GroupDev& g, const void* Bv, const double* u, double* r,
                        const double* flin, const PointState& st, hipStream_t s) {
    const dim3 grid((unsigned)g.ntiles), block(kWave);
    if (commit)
        hipLaunchKernelGGL((k_residual_fmt<SHAPE, FMT, true, false, false>), grid, block, 0, s, g, Bv, u, r, flin, st);
    else if (g.node && fast)
        hipLaunchKernelGGL((k_residual_fmt<SHAPE, FMT, false, true, true>), grid, block, 0, s, g, Bv, u, r, flin, st);
    else if (g.node)
        hipLaunchKernelGGL((k_residual_fmt<SHAPE, FMT, false, true, false>), grid, block, 0, s, g, Bv, u, r, flin, st);
    else if (fast)
        hipLaunchKernelGGL((k_residual_fmt<SHAPE, FMT, false, false, true>), grid, block, 0, s, g, Bv, u, r, flin, st);
    else
        hipLaunchKernelGGL((k_residual_fmt<SHAPE, FMT, false, false, false>), grid, block, 0, s, g, Bv, u, r, flin, st);
}

template <int FMT>
void residual_fmt(int shape, bool commit, bool fast, const GroupDev& g, const void* Bv, const double* u, double* r,
                  const double* flin, const PointState& st, hipStream_t s) {
    dispatch_shape(shape, [&](auto sc) {
        residual_fmt_shape<decltype(sc)::value, FMT>(commit, fast, g, Bv, u, r, flin, st, s);
    });
}

template <int FMT>
void tangent_fmt(int shape, const GroupDev& g, const void* Bv, const double* v, double* y, const double* dep,
                 const float* dep32, int64_t npad, bool zf, hipStream_t s) {
    const dim3 grid((unsigned)g.ntiles), block(kWave);
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        if (g.node)
            hipLaunchKernelGGL((k_tangent_fmt<S, FMT, true>), grid, block, 0, s, g, Bv, v, y, dep, dep32, npad, zf);
        else
            hipLaunchKernelGGL((k_tangent_fmt<S, FMT, false>), grid, block, 0, s, g, Bv, v, y, dep, dep32, npad, zf);
    });
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = float(src[i]);
}

template <int FMT>
void convert_fmt(int shape, const double* dense, void* out, int64_t ntiles, hipStream_t s) {
    const dim3 grid((unsigned)ntiles), block(kWave);
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL((k_convert_B<S, FMT>), grid, block, 0, s, dense, out, ntiles);
    });
}

}  // namespace

void launch_residual_fmt(int fmt, int shape, bool commit, bool fast, const GroupDev& g, const void* Bv,
                         const double* u, double* r, const double* flin, const PointState& st, hipStream_t s) {
    if (g.ntiles == 0) return;
    const bool f = fast && !commit;
    switch (fmt) {
        case FMT_D64: residual_fmt<FMT_D64>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        case FMT_C64: residual_fmt<FMT_C64>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        case FMT_G64: residual_fmt<FMT_G64>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        case FMT_G32: residual_fmt<FMT_G32>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        case FMT_D32: residual_fmt<FMT_D32>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        case FMT_C32: residual_fmt<FMT_C32>(shape, commit, f, g, Bv, u, r, flin, st, s); break;
        default: break;
    }
}

void launch_tangent_fmt(int fmt, int shape, const GroupDev& g, const void* Bv, const double* v, double* y,
                        const double* dep, const float* dep32, int64_t npad, bool zero_first, hipStream_t s) {
    if (g.ntiles == 0) return;
    const bool zf = zero_first;
    switch (fmt) {
        case FMT_D64: tangent_fmt<FMT_D64>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        case FMT_C64: tangent_fmt<FMT_C64>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        case FMT_G64: tangent_fmt<FMT_G64>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        case FMT_D32: tangent_fmt<FMT_D32>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        case FMT_C32: tangent_fmt<FMT_C32>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        case FMT_G32: tangent_fmt<FMT_G32>(shape, g, Bv, v, y, dep, dep32, npad, zf, s); break;
        default: break;
    }
}

void launch_to_f32(int64_t n, const double* src, float* dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, dst);
}

template <int S, int NW>
void launch_patch_nw(bool fast, const PatchDev& p, const double* u, double* r, const double* flin, double* partial,
                     hipStream_t s) {
    const dim3 grid((unsigned)p.npatch), block(kWave * NW);
    if (fast)
        hipLaunchKernelGGL((k_residual_patch<S, true, NW>), grid, block, 0, s, p, u, r, flin, partial);
    else
        hipLaunchKernelGGL((k_residual_patch<S, false, NW>), grid, block, 0, s, p, u, r, flin, partial);
}

void launch_residual_patch(int shape, int nw, bool fast, const PatchDev& p, const double* u, double* r,
                           const double* flin, double* partial, hipStream_t s) {
    if (p.npatch == 0) return;
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        if (nw == 2)
            launch_patch_nw<S, 2>(fast, p, u, r, flin, partial, s);
        else if (nw == 8)
            launch_patch_nw<S, 8>(fast, p, u, r, flin, partial, s);
        else
            launch_patch_nw<S, 4>(fast, p, u, r, flin, partial, s);
    });
}

void launch_convert_B(int fmt, int shape, const double* dense, void* out, int64_t ntiles, hipStream_t s) {
    if (ntiles == 0) return;
    switch (fmt) {
        case FMT_C64: convert_fmt<FMT_C64>(shape, dense, out, ntiles, s); break;
        case FMT_D32: convert_fmt<FMT_D32>(shape, dense, out, ntiles, s); break;
        case FMT_C32: convert_fmt<FMT_C32>(shape, dense, out, ntiles, s); break;
        default: break;
    }
}

void launch_build_B(int shape, const GroupDev& g, const double* corners, double* dense, double* w, hipStream_t s) {
    if (g.ntiles == 0) return;
    const dim3 grid((unsigned)g.ntiles), block(kWave);
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL((k_build_B<S>), grid, block, 0, s, g, corners, dense, w);
    });
}

}  // namespace ipf
