// ipf_geometry.h -- B rows and integration weights of one integration point
// from the element's geometry, shared by the host setup, the device setup
// kernel and the geometric B format (which recomputes B inside the residual
// and tangent kernels instead of reading it).  One definition, compiled for
// host and device with contraction off, so the three produce the same bits.
//
// Geometry: the bilinear map of the element's four corners X (x0, y0, ..., x3,
// y3; counter-clockwise, corner c at reference position (-1,-1), (1,-1),
// (1,1), (-1,1)).  At reference point (xi, eta):
//   J = [[dx/dxi, dx/deta], [dy/dxi, dy/deta]],   det = |J|,
//   J^-1 = [[a, c'], [b', d]] with a = dy/deta / det, b = -dy/dxi / det,
//   c = -dx/deta / det, d = dx/dxi / det,
//   dN/dx = a dN/dxi + b dN/deta,   dN/dy = c dN/dxi + d dN/deta,
//   weight = w_q |det|.
// This is the reference's dphix = J^-1 dphi (TPZInterpolationSpace::
// ComputeRequiredData, used at TPZNumericalIntegrator.cpp:362-381) and
// w |detJ| (TPZNumericalIntegrator.cpp:420-426).
//
// Shape table of one block shape (n_ip points, n_sh shape functions), in
// doubles: dphi[n_ip][n_sh][2] = (dN_a/dxi, dN_a/deta) at each point, then
// xi[n_ip], eta[n_ip], w_q[n_ip].  dphi first keeps every row 16-byte
// aligned (2 n_sh doubles per row).
#pragma once

#include <cmath>

#include "ipf_pointmath.h"

namespace ipf {

constexpr int shape_table_size(int nip, int nsh) { return nip * (2 * nsh + 3); }

struct QuadMap {
    double a, b, c, d;  // J^-1 entries as above
    double det;
};

IPF_HD QuadMap quad_map(const double* X, double xi, double eta) {
    const double dxi[4] = {-(1 - eta) / 4, (1 - eta) / 4, (1 + eta) / 4, -(1 + eta) / 4};
    const double deta[4] = {-(1 - xi) / 4, -(1 + xi) / 4, (1 + xi) / 4, (1 - xi) / 4};
    double xx = 0, xe = 0, yx = 0, ye = 0;  // dx/dxi, dx/deta, dy/dxi, dy/deta
#pragma unroll
    for (int c = 0; c < 4; c++) {
        xx += dxi[c] * X[2 * c];
        xe += deta[c] * X[2 * c];
        yx += dxi[c] * X[2 * c + 1];
        ye += deta[c] * X[2 * c + 1];
    }
    QuadMap m;
    m.det = xx * ye - xe * yx;
    const double inv = 1.0 / m.det;
    m.a = ye * inv;
    m.b = -(yx * inv);
    m.c = -(xe * inv);
    m.d = xx * inv;
    return m;
}

IPF_HD void quad_grad(const QuadMap& m, double dxi, double deta, double& dx, double& dy) {
    dx = m.a * dxi + m.b * deta;
    dy = m.c * dxi + m.d * deta;
}

IPF_HD double quad_weight(const QuadMap& m, double wq) { return wq * std::fabs(m.det); }

}  // namespace ipf
