// ipf_assemble.hip -- tangent stiffness assembly into the symmetric CSR
// values (TPZNumericalIntegrator::KAssembly, TPZNumericalIntegrator.cpp:131-214,
// with ComputeTangentMatrix :236-270 and the per-color gather / add / scatter
// of :166-178).
//
// One wavefront per element.  The element's dense B rows (3 x C per point),
// its points' w*Dep (3 x 3) and the products Dep.B are staged in LDS; the
// lanes then produce the packed upper triangle K_e(a, c), a <= c, row by row
// (the order of fColorLSequence), and add each entry to Kg at its CSR
// position.  Launched once per (color, shape) group in color order, so every
// Kg entry receives its element contributions color by color as in the
// reference, and the elements of one launch never share a CSR position (they
// share no DoF).
//
// Operation order per entry (the CPU restatement in the oracle uses the same):
//   DB[r][j]  = 0 + D[r][0] B[0][j] + D[r][1] B[1][j] + D[r][2] B[2][j]
//   K_e(a, c) = sum over points, in point order, of
//               (0 + B[0][a] DB[0][c] + B[1][a] DB[1][c] + B[2][a] DB[2][c])
//   Kg[l]     = Kg[l] + K_e(a, c)
#include "ipf_kernels.h"

namespace ipf {

namespace {

constexpr int kWave = 64;

template <int SHAPE>
__global__ void __launch_bounds__(kWave)
k_assemble_K(GroupDev g, const double* __restrict__ dep, int64_t npad, const int* __restrict__ lidx,
             double* __restrict__ kg) {
    constexpr int NIP = shape_nip(SHAPE);
    constexpr int CE = shape_ce(SHAPE);
    constexpr int ET = shape_et(SHAPE);
    constexpr int TP = shape_tp(SHAPE);
    constexpr int NB = 3 * CE;
    constexpr int NT = CE * (CE + 1) / 2;
    __shared__ double Bs[NIP * NB];
    __shared__ double Ds[NIP * 9];
    __shared__ double DBs[NIP * NB];
    const int lane = threadIdx.x;
    const int64_t elem = blockIdx.x;  // grid = g.nelem element slots
    const int64_t tile = elem / ET;
    const int el = (int)(elem - tile * ET);
    if (el >= tile_nvalid(g, tile)) return;  // padding slot
    // dense tile layout: value m of tile slot q at (m / 2) * 2 TP + 2 q + (m & 1)
    const double* Bt = g.B + tile * (int64_t)(NB * TP);
    for (int t = lane; t < NIP * NB; t += kWave) {
        const int ip = t / NB;
        const int m = t - ip * NB;
        const int q = el * NIP + ip;
        Bs[t] = Bt[(m >> 1) * (2 * TP) + 2 * q + (m & 1)];
    }
    const int64_t P0 = g.pt0 + tile * TP + el * NIP;
    for (int t = lane; t < NIP * 9; t += kWave) {
        const int ip = t / 9;
        const int rk = t - ip * 9;
        Ds[t] = dep[rk * npad + P0 + ip];
    }
    __syncthreads();
    for (int t = lane; t < NIP * NB; t += kWave) {
        const int ip = t / NB;
        const int rj = t - ip * NB;
        const int r = rj / CE;
        const int j = rj - r * CE;
        double acc = 0.;
#pragma unroll
        for (int k = 0; k < 3; k++) acc = acc + Ds[ip * 9 + 3 * r + k] * Bs[ip * NB + k * CE + j];
        DBs[t] = acc;
    }
    __syncthreads();
    const int* li = lidx + elem * (int64_t)NT;
    for (int t = lane; t < NT; t += kWave) {
        // packed row-major upper triangle: row a holds CE - a entries
        int a = 0, rem = t;
        while (rem >= CE - a) {
            rem -= CE - a;
            a++;
        }
        const int c = a + rem;
        double ke = 0.;
        for (int ip = 0; ip < NIP; ip++) {
            double acc = 0.;
#pragma unroll
            for (int k = 0; k < 3; k++) acc = acc + Bs[ip * NB + k * CE + a] * DBs[ip * NB + k * CE + c];
            ke = ke + acc;
        }
        const int l = li[t];
        kg[l] = kg[l] + ke;
    }
}

}  // namespace

void launch_assemble_K(int shape, const GroupDev& g, const double* dep, int64_t npad, const int* lidx, double* kg,
                       hipStream_t s) {
    if (g.nelem == 0) return;
    const dim3 grid((unsigned)g.nelem);
    dispatch_shape(shape, [&](auto sc) {
        constexpr int S = decltype(sc)::value;
        hipLaunchKernelGGL(k_assemble_K<S>, grid, dim3(kWave), 0, s, g, dep, npad, lidx, kg);
    });
}

}  // namespace ipf
