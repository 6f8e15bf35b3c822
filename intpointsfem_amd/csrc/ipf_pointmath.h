// ipf_pointmath.h -- per-integration-point Mohr-Coulomb return map for the
// gfx950 kernels (2D plane strain, perfect plasticity).
//
// Restates FunctionsStressStrain.h, FunctionsSpectralDecomp.h,
// FunctionsSigmaProjection.h and FunctionsTangentOperator.h of IntPointsFEM
// with the same floating-point operation order (the library is built with
// -ffp-contract=off, so no multiply-add is fused), so strains, branch flags
// and stresses are bit-identical to the CPU path.  Material trigonometry and
// every material-only subexpression are evaluated once on the host
// (MatConst::make) with the expression the reference evaluates per point.
//
// Layout conventions: full tensors are 6-vectors in Voigt order
// [XX, XY, XZ, YY, YZ, ZZ] (FunctionsTangentOperator.h:40-41); the committed
// plastic strain is stored as 4 components [XX, XY, YY, ZZ] because its XZ and
// YZ components are +0 in plane strain (0 - (+-0) == +0 in IEEE arithmetic).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define IPF_HD __host__ __device__ __forceinline__

namespace ipf {

enum { XX = 0, XY = 1, XZ = 2, YY = 3, YZ = 4, ZZ = 5 };
enum { BR_ELASTIC = 0, BR_MAIN = 1, BR_RIGHT = 2, BR_LEFT = 3, BR_APEX = 4 };

// Material constants: everything the return map needs that depends only on
// (lambda, mu, phi, psi, c).
struct MatConst {
    double lambda, mu, G, K;
    double lam2mu;          // lambda + 2*mu                      (FunctionsStressStrain.h:48-50)
    double cs_a, cs_b;      // compliance factors                  (FunctionsStressStrain.h:74-76)
    double inv_mu;          // 1/mu                                (FunctionsStressStrain.h:78-80)
    double sinphi, cosphi, sinpsi;
    double phi_c;           // 2.*c*cosphi                         (FunctionsSigmaProjection.h:12,38)
    double edge_c;          // 2.*cosphi*c                         (FunctionsSigmaProjection.h:115,220)
    double constA;          //                                     (FunctionsSigmaProjection.h:36)
    double m0, m1, m2;      // principal-stress update factors     (FunctionsSigmaProjection.h:53-55)
    double r_ab0, r_ab1, r_ji00, r_ji01, r_ji10, r_ji11;  // right edge (:117-136)
    double l_ab0, l_ab1, l_ji00, l_ji01, l_ji10, l_ji11;  // left edge  (:217-238)
    double omsp, opsp;      // 1 - sinpsi, 1 + sinpsi              (FunctionsSigmaProjection.h:359)
    double c_cot;           // c * (1/tan(phi))                    (FunctionsSigmaProjection.h:293,301)
    double apex_alpha;      // cos(phi)/sin(psi)                   (FunctionsSigmaProjection.h:300)
    double tan_plane[9], tan_right[9], tan_left[9];

    static MatConst make(double lambda, double mu, double mc_phi, double mc_psi, double c);
};

// ---------------------------------------------------------------- spectral

// FunctionsSpectralDecomp.h:7-20
IPF_HD double normalize6(double* s) {
    double maxel = s[0];
#pragma unroll
    for (int i = 1; i < 6; i++)
        if (fabs(s[i]) > fabs(maxel)) maxel = s[i];
    if (maxel != 0) {
#pragma unroll
        for (int i = 0; i < 6; i++) s[i] /= maxel;
    }
    return maxel;
}

// FunctionsSpectralDecomp.h:25-56 (row 3 uses sxy + syz as the reference does)
IPF_HD void gershgorin(const double* s, double& upper_out, double& lower_out) {
    const double row_1 = s[XY] + s[XZ];
    const double lo0 = s[XX] - fabs(row_1), up0 = s[XX] + fabs(row_1);
    const double row_2 = s[XY] + s[YZ];
    const double lo1 = s[YY] - fabs(row_2), up1 = s[YY] + fabs(row_2);
    const double row_3 = s[XY] + s[YZ];
    const double lo2 = s[ZZ] - fabs(row_3), up2 = s[ZZ] + fabs(row_3);
    double up = up0, lo = lo0;
    if (up1 > up) up = up1;
    if (lo1 < lo) lo = lo1;
    if (up2 > up) up = up2;
    if (lo2 < lo) lo = lo2;
    upper_out = up;
    lower_out = lo;
}

// Cubic of FunctionsSpectralDecomp.h:75-77, evaluated left to right.
IPF_HD double cubic(double x, const double* s, double sxy2, double sxz2, double syz2) {
    return -x * x * x - sxz2 * s[YY] - sxy2 * s[ZZ] - syz2 * s[XX] + 2 * s[XY] * s[XZ] * s[YZ] +
           s[XX] * s[YY] * s[ZZ] + x * x * (s[XX] + s[YY] + s[ZZ]) +
           x * (sxy2 + sxz2 + syz2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ]);
}

// FunctionsSpectralDecomp.h:61-111: Newton from each Gershgorin bound
// (20 iterations, |f| > 1e-11), third root from the trace, descending sort.
IPF_HD void eigenvalues_newton(double up, double lo, const double* s, double maxel, double* ev) {
    const double sxy2 = s[XY] * s[XY];
    const double sxz2 = s[XZ] * s[XZ];
    const double syz2 = s[YZ] * s[YZ];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        double x = (i == 0) ? up : lo;
        double f = cubic(x, s, sxy2, sxz2, syz2);
        double res = fabs(f);
        int it = 0;
        while (it < 20 && res > 10e-12) {
            const double df = -3 * x * x + sxy2 + sxz2 + syz2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ] +
                              2 * x * (s[XX] + s[YY] + s[ZZ]);
            x -= f / df;
            f = cubic(x, s, sxy2, sxz2, syz2);
            res = fabs(f);
            it++;
        }
        ev[i] = x;
    }
    ev[2] = s[XX] + s[YY] + s[ZZ] - ev[0] - ev[1];
    ev[0] *= maxel;
    ev[1] *= maxel;
    ev[2] *= maxel;
    // i = 0: j = 1, 2; i = 1: j = 2 (FunctionsSpectralDecomp.h:102-110)
    if (ev[0] < ev[1]) { const double a = ev[0]; ev[0] = ev[1]; ev[1] = a; }
    if (ev[0] < ev[2]) { const double a = ev[0]; ev[0] = ev[2]; ev[2] = a; }
    if (ev[1] < ev[2]) { const double a = ev[1]; ev[1] = ev[2]; ev[2] = a; }
}

// FunctionsSpectralDecomp.h:116-150
// The eigenvector helpers return their vectors as values (V3) and the callers
// store all nine components once, unconditionally: with results written
// through &vec[0] / &vec[3] / &vec[6] under branches the compiler merged the
// stores into one through a computed address and kept vec in private
// (scratch) memory.  The arithmetic is unchanged.
struct V3 {
    double x, y, z;
};

IPF_HD V3 eigvec_simple(const double* s, double lam) {
    const double d0 = (s[XX] - lam) * (s[YY] - lam) - s[XY] * s[XY];
    const double d1 = (s[XX] - lam) * (s[ZZ] - lam) - s[XZ] * s[XZ];
    const double d2 = (s[YY] - lam) * (s[ZZ] - lam) - s[YZ] * s[YZ];
    double maxdet = fabs(d0);
    if (fabs(d1) > fabs(maxdet)) maxdet = fabs(d1);
    if (fabs(d2) > fabs(maxdet)) maxdet = fabs(d2);
    double v0, v1, v2;
    if (maxdet == fabs(d0)) {
        v0 = 1 / d0 * (-(s[YY] - lam) * s[XZ] + s[XY] * s[YZ]);
        v1 = 1 / d0 * (-(s[XX] - lam) * s[YZ] + s[XY] * s[XZ]);
        v2 = 1;
    } else if (maxdet == fabs(d1)) {
        v0 = 1 / d1 * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v1 = 1;
        v2 = 1 / d1 * (-(s[XX] - lam) * s[YZ] + s[XY] * s[XZ]);
    } else {
        v0 = 1;
        v1 = 1 / d2 * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v2 = 1 / d2 * (-(s[YY] - lam) * s[XZ] + s[XY] * s[YZ]);
    }
    const double norm = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    return V3{v0 / norm, v1 / norm, v2 / norm};
}

// FunctionsSpectralDecomp.h:155-212 (second norm keeps the v2[1]*v1[1] term of :203)
IPF_HD void eigvec_double(const double* s, double lam, V3& vec1, V3& vec2) {
    const double x0 = s[XX] - lam, x1 = s[YY] - lam, x2 = s[ZZ] - lam;
    double maxx = fabs(x0);
    if (fabs(x1) > fabs(maxx)) maxx = fabs(x1);
    if (fabs(x2) > fabs(maxx)) maxx = fabs(x2);
    double a0, a1, a2, b0, b1, b2;
    if (maxx == fabs(x0)) {
        a0 = -s[XY] / x0; a1 = 1; a2 = 0;
        b0 = -s[XZ] / x0; b1 = 0; b2 = 1;
    } else if (maxx == fabs(x1)) {
        a0 = 1; a1 = -s[XY] / x1; a2 = 0;
        b0 = 0; b1 = -s[YZ] / x1; b2 = 1;
    } else {
        a0 = 1; a1 = 0; a2 = -s[XZ] / x2;
        b0 = 0; b1 = 1; b2 = -s[YZ] / x2;
    }
    const double n1 = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    const double n2 = sqrt(b0 * b0 + b1 * a1 + b2 * b2);
    vec1 = V3{a0 / n1, a1 / n1, a2 / n1};
    vec2 = V3{b0 / n2, b1 / n2, b2 / n2};
}

// FunctionsSpectralDecomp.h:217-256 followed by the normalise/scale-back of
// :261-268.  s is normalised on entry and scaled back by maxel here.
IPF_HD void eigenvectors6(double* s, const double* ev, double* vec, double maxel) {
#pragma unroll
    for (int i = 0; i < 6; i++) s[i] *= maxel;
    V3 a{1., 0., 0.}, b{0., 1., 0.}, c{0., 0., 1.};  // three equal eigenvalues: the identity
    if (!((ev[0] == ev[1]) && (ev[1] == ev[2]))) {
        if (ev[0] != ev[1] && ev[0] != ev[2]) {
            a = eigvec_simple(s, ev[0]);
        } else if (ev[0] == ev[1]) {
            eigvec_double(s, ev[0], a, b);
        } else if (ev[0] == ev[2]) {
            eigvec_double(s, ev[0], a, c);
        }
        if (ev[1] != ev[0] && ev[1] != ev[2]) {
            b = eigvec_simple(s, ev[1]);
        } else if (ev[1] == ev[2]) {
            eigvec_double(s, ev[1], b, c);
        }
        if (ev[2] != ev[0] && ev[2] != ev[1]) {
            c = eigvec_simple(s, ev[2]);
        }
    }
    vec[0] = a.x; vec[1] = a.y; vec[2] = a.z;
    vec[3] = b.x; vec[4] = b.y; vec[5] = b.z;
    vec[6] = c.x; vec[7] = c.y; vec[8] = c.z;
}

// FunctionsSpectralDecomp.h:261-268 (s is modified in place like the reference)
IPF_HD void spectral(double* s, double* ev, double* vec) {
    const double maxel = normalize6(s);
    double up, lo;
    gershgorin(s, up, lo);
    eigenvalues_newton(up, lo, s, maxel, ev);
    eigenvectors6(s, ev, vec, maxel);
}

// ---------------------------------------------------------------- plane strain
//
// spectral_ps: spectral() for a tensor whose XZ and YZ components are +0 (the
// trial stresses and elastic strains of plane strain: mu * 0.0, 0.0).  The
// results equal spectral()'s bit for bit up to the sign of exact zeros; only
// work on the zero components is left out:
//   * normalize: +0 / maxel is a zero with the sign of maxel;
//   * Gershgorin rows: |sxy + (+-0)| = |sxy|;
//   * cubic and derivative: sxz^2 = syz^2 = +0 and 2 sxy sxz syz = +-0.
//     Adding or subtracting a zero changes a partial sum only when that sum
//     is itself zero, and then only its sign, so the cubic is the same
//     nonzero value or zero in both forms (it enters only |f| and, when
//     nonzero, f / f'); the derivative's partial sum -3x^2 + sxy^2 is never
//     -0, so adding +0 to it is exact;
//   * eigenvectors: after the scale-back sxz = syz = +0, so in the branch of
//     the xy minor the vector is (zero, zero, 1) with norm exactly 1, and in
//     the other two branches its z component is a zero that the division by
//     the (finite, >= 1) norm leaves unchanged.
// Non-finite or degenerate cases (a zero minor, NaN) take the general code.

IPF_HD double cubic_ps(double x, const double* s, double sxy2) {
    return -x * x * x - sxy2 * s[ZZ] + s[XX] * s[YY] * s[ZZ] + x * x * (s[XX] + s[YY] + s[ZZ]) +
           x * (sxy2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ]);
}

IPF_HD void eigenvalues_newton_ps(double up, double lo, const double* s, double maxel, double* ev) {
    const double sxy2 = s[XY] * s[XY];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        double x = (i == 0) ? up : lo;
        double f = cubic_ps(x, s, sxy2);
        double res = fabs(f);
        int it = 0;
        while (it < 20 && res > 10e-12) {
            const double df = -3 * x * x + sxy2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ] +
                              2 * x * (s[XX] + s[YY] + s[ZZ]);
            x -= f / df;
            f = cubic_ps(x, s, sxy2);
            res = fabs(f);
            it++;
        }
        ev[i] = x;
    }
    ev[2] = s[XX] + s[YY] + s[ZZ] - ev[0] - ev[1];
    ev[0] *= maxel;
    ev[1] *= maxel;
    ev[2] *= maxel;
    if (ev[0] < ev[1]) { const double a = ev[0]; ev[0] = ev[1]; ev[1] = a; }
    if (ev[0] < ev[2]) { const double a = ev[0]; ev[0] = ev[2]; ev[2] = a; }
    if (ev[1] < ev[2]) { const double a = ev[1]; ev[1] = ev[2]; ev[2] = a; }
}

IPF_HD V3 eigvec_simple_ps(const double* s, double lam) {
    const double d0 = (s[XX] - lam) * (s[YY] - lam) - s[XY] * s[XY];
    const double d1 = (s[XX] - lam) * (s[ZZ] - lam) - s[XZ] * s[XZ];
    const double d2 = (s[YY] - lam) * (s[ZZ] - lam) - s[YZ] * s[YZ];
    double maxdet = fabs(d0);
    if (fabs(d1) > fabs(maxdet)) maxdet = fabs(d1);
    if (fabs(d2) > fabs(maxdet)) maxdet = fabs(d2);
    double v0, v1, v2;
    if (maxdet == fabs(d0)) {
        const double t0 = -(s[YY] - lam) * s[XZ] + s[XY] * s[YZ];
        const double t1 = -(s[XX] - lam) * s[YZ] + s[XY] * s[XZ];
        if ((d0 > 0.0 || d0 < 0.0) && t0 == 0.0 && t1 == 0.0) {
            // (1/d0) t: a zero with sign(d0) * sign(t); norm sqrt(0 + 0 + 1) = 1
            return V3{d0 > 0.0 ? t0 : -t0, d0 > 0.0 ? t1 : -t1, 1.0};
        }
        v0 = 1 / d0 * t0;
        v1 = 1 / d0 * t1;
        v2 = 1;
    } else if (maxdet == fabs(d1)) {
        v0 = 1 / d1 * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v1 = 1;
        v2 = 1 / d1 * (-(s[XX] - lam) * s[YZ] + s[XY] * s[XZ]);
    } else {
        v0 = 1;
        v1 = 1 / d2 * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v2 = 1 / d2 * (-(s[YY] - lam) * s[XZ] + s[XY] * s[YZ]);
    }
    const double norm = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    // v2 / norm for a zero v2 and a finite norm (>= 1 here) is v2 itself
    return V3{v0 / norm, v1 / norm, (v2 == 0.0 && norm >= 1.0 && norm <= 1.7976931348623157e308) ? v2 : v2 / norm};
}

IPF_HD void eigenvectors6_ps(double* s, const double* ev, double* vec, double maxel) {
#pragma unroll
    for (int i = 0; i < 6; i++) s[i] *= maxel;
    V3 a{1., 0., 0.}, b{0., 1., 0.}, c{0., 0., 1.};  // three equal eigenvalues: the identity
    if (!((ev[0] == ev[1]) && (ev[1] == ev[2]))) {
        if (ev[0] != ev[1] && ev[0] != ev[2]) {
            a = eigvec_simple_ps(s, ev[0]);
        } else if (ev[0] == ev[1]) {
            eigvec_double(s, ev[0], a, b);
        } else if (ev[0] == ev[2]) {
            eigvec_double(s, ev[0], a, c);
        }
        if (ev[1] != ev[0] && ev[1] != ev[2]) {
            b = eigvec_simple_ps(s, ev[1]);
        } else if (ev[1] == ev[2]) {
            eigvec_double(s, ev[1], b, c);
        }
        if (ev[2] != ev[0] && ev[2] != ev[1]) {
            c = eigvec_simple_ps(s, ev[2]);
        }
    }
    vec[0] = a.x; vec[1] = a.y; vec[2] = a.z;
    vec[3] = b.x; vec[4] = b.y; vec[5] = b.z;
    vec[6] = c.x; vec[7] = c.y; vec[8] = c.z;
}

// s[XZ] and s[YZ] must be +0 on entry.
IPF_HD void spectral_ps(double* s, double* ev, double* vec) {
    // normalize6: the zero components never win the max-abs search
    double maxel = s[XX];
    if (fabs(s[XY]) > fabs(maxel)) maxel = s[XY];
    if (fabs(s[YY]) > fabs(maxel)) maxel = s[YY];
    if (fabs(s[ZZ]) > fabs(maxel)) maxel = s[ZZ];
    if (maxel != 0) {
        s[XX] /= maxel;
        s[XY] /= maxel;
        s[YY] /= maxel;
        s[ZZ] /= maxel;
        const double z = maxel > 0 ? 0.0 : (maxel < 0 ? -0.0 : s[XZ] / maxel);
        s[XZ] = z;
        s[YZ] = z;
    }
    // Gershgorin (FunctionsSpectralDecomp.h:25-56) with |sxy + sxz| = |sxy + syz| = |sxy|
    const double a = fabs(s[XY]);
    double up = s[XX] + a, lo = s[XX] - a;
    if (s[YY] + a > up) up = s[YY] + a;
    if (s[YY] - a < lo) lo = s[YY] - a;
    if (s[ZZ] + a > up) up = s[ZZ] + a;
    if (s[ZZ] - a < lo) lo = s[ZZ] - a;
    eigenvalues_newton_ps(up, lo, s, maxel, ev);
    eigenvectors6_ps(s, ev, vec, maxel);
}

// ---------------------------------------------------------------- projection

IPF_HD bool ordered(const double* e) {
    return (e[0] > e[1] || fabs(e[0] - e[1]) < 1.e-12) && (e[1] > e[2] || fabs(e[1] - e[2]) < 1.e-12);
}

// FunctionsSigmaProjection.h:25-63 (H = 0)
IPF_HD bool main_plane(const double* tr, double* sp, double& hard, const MatConst& m) {
    double e0 = tr[0], e1 = tr[1], e2 = tr[2];
    double phi = e0 - e2 + (e0 + e2) * m.sinphi - m.phi_c;
    double gamma = 0;
    const double jac = -m.constA;
    for (int i = 0; i < 30; i++) {
        const double delta_gamma = -phi / jac;
        gamma += delta_gamma;
        phi = e0 - e2 + (e0 + e2) * m.sinphi - m.phi_c - m.constA * gamma;
        if (fabs(phi) < 1.e-12) break;
    }
    e0 -= m.m0 * gamma;
    e1 += m.m1 * gamma;
    e2 += m.m2 * gamma;
    sp[0] = e0;
    sp[1] = e1;
    sp[2] = e2;
    hard += gamma * 2. * m.cosphi;
    return ordered(sp);
}

// 2x2 Newton of the edge returns, FunctionsSigmaProjection.h:119-146 / :223-250.
// gamma starts at {0, 0} (the reference leaves it uninitialised, :103, :205).
IPF_HD void edge_newton(double sb0, double sb1, double ab0, double ab1, double ji00, double ji01,
                                            double ji10, double ji11, double edge_c, double& g0, double& g1) {
    double phi0 = sb0 - edge_c;
    double phi1 = sb1 - edge_c;
    g0 = 0.;
    g1 = 0.;
    for (int i = 0; i < 30; i++) {
        g0 -= (ji00 * phi0 + ji01 * phi1);
        g1 -= (ji10 * phi0 + ji11 * phi1);
        phi0 = sb0 - ab0 * g0 - ab1 * g1 - edge_c;
        phi1 = sb1 - ab1 * g0 - ab0 * g1 - edge_c;
        const double res = (fabs(phi0) + fabs(phi1));
        if (fabs(res) < 1.e-12) break;
    }
}

// FunctionsSigmaProjection.h:93-158
IPF_HD bool right_edge(const double* tr, double* sp, double& hard, const MatConst& m) {
    double e0 = tr[0], e1 = tr[1], e2 = tr[2];
    const double sb0 = e0 - e2 + (e0 + e2) * m.sinphi;
    const double sb1 = e0 - e1 + (e0 + e1) * m.sinphi;
    double g0, g1;
    edge_newton(sb0, sb1, m.r_ab0, m.r_ab1, m.r_ji00, m.r_ji01, m.r_ji10, m.r_ji11, m.edge_c, g0, g1);
    e0 -= m.m0 * (g0 + g1);
    e1 += m.m1 * g0 + m.m2 * g1;
    e2 += m.m2 * g0 + m.m1 * g1;
    sp[0] = e0;
    sp[1] = e1;
    sp[2] = e2;
    hard += (g0 + g1) * 2. * m.cosphi;
    return ordered(sp);
}

// FunctionsSigmaProjection.h:193-257
IPF_HD bool left_edge(const double* tr, double* sp, double& hard, const MatConst& m) {
    double e0 = tr[0], e1 = tr[1], e2 = tr[2];
    const double sb0 = e0 - e2 + (e0 + e2) * m.sinphi;
    const double sb1 = e1 - e2 + (e1 + e2) * m.sinphi;
    double g0, g1;
    edge_newton(sb0, sb1, m.l_ab0, m.l_ab1, m.l_ji00, m.l_ji01, m.l_ji10, m.l_ji11, m.edge_c, g0, g1);
    e0 += -m.m0 * g0 + m.m1 * g1;
    e1 += m.m1 * g0 - m.m0 * g1;
    e2 += m.m2 * (g0 + g1);
    sp[0] = e0;
    sp[1] = e1;
    sp[2] = e2;
    hard += (g0 + g1) * 2. * m.cosphi;
    return ordered(sp);
}

// FunctionsSigmaProjection.h:292-323
IPF_HD void apex(const double* tr, double* sp, double& hard, const MatConst& m) {
    double ptr = 0.;
    ptr += tr[0];
    ptr += tr[1];
    ptr += tr[2];
    ptr /= 3.;
    double depv = 0.;
    double res = m.c_cot - ptr;
    double p = 0.;
    for (int i = 0; i < 30; i++) {
        depv -= res / m.K;
        p = ptr - m.K * depv;
        res = m.c_cot - p;
        if (fabs(res) < 1.e-12) break;
    }
    hard += m.apex_alpha * depv;
    sp[0] = p;
    sp[1] = p;
    sp[2] = p;
}

// FunctionsSigmaProjection.h:337-379.  The gradient (principal-space
// d sigma / d sigma_trial) is produced only when WANT_GRAD.
template <bool WANT_GRAD>
IPF_HD void project(const double* ev, double* sp, int& mtype, int& branch, double& hard,
                                        const MatConst& m, double* grad) {
    const double phi = ev[0] - ev[2] + (ev[0] + ev[2]) * m.sinphi - m.phi_c;
    sp[0] = ev[0];
    sp[1] = ev[1];
    sp[2] = ev[2];
    if ((fabs(phi) < 1.e-12) || (phi < 0.0)) {
        mtype = 0;
        branch = BR_ELASTIC;
        if (WANT_GRAD) {
#pragma unroll
            for (int i = 0; i < 9; i++) grad[i] = 0.;
            grad[0] = 1.;
            grad[4] = 1.;
            grad[8] = 1.;
        }
        return;
    }
    mtype = 1;
    branch = BR_MAIN;
    if (main_plane(ev, sp, hard, m)) {
        if (WANT_GRAD) {
#pragma unroll
            for (int i = 0; i < 9; i++) grad[i] = m.tan_plane[i];
        }
        return;
    }
    bool ok;
    if ((m.omsp * ev[0] - 2. * ev[1] + m.opsp * ev[2]) > 0) {
        branch = BR_RIGHT;
        ok = right_edge(ev, sp, hard, m);
        if (ok && WANT_GRAD) {
#pragma unroll
            for (int i = 0; i < 9; i++) grad[i] = m.tan_right[i];
        }
    } else {
        branch = BR_LEFT;
        ok = left_edge(ev, sp, hard, m);
        if (ok && WANT_GRAD) {
#pragma unroll
            for (int i = 0; i < 9; i++) grad[i] = m.tan_left[i];
        }
    }
    if (!ok) {
        mtype = -1;
        branch = BR_APEX;
        apex(ev, sp, hard, m);
        if (WANT_GRAD) {
#pragma unroll
            for (int i = 0; i < 9; i++) grad[i] = 0.;
        }
    }
}

// ---------------------------------------------------------------- tangent

// FunctionsTangentOperator.h:37-123, reduced to the 3x3 [xx, xy, yy] block of
// FunctionsStressStrain.h:103-119 (TranslateDep) and scaled by w.
IPF_HD void tangent3(const double* grad, const double* sev, const double* eev, const double* v,
                                         double G, double lambda, double w, double* dep9) {
    const int kival[6] = {0, 0, 0, 1, 1, 2};
    const int kjval[6] = {0, 1, 2, 1, 2, 2};
    double T[36];
#pragma unroll
    for (int l = 0; l < 36; ++l) T[l] = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int ki = kival[k], kj = kjval[k];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double temp;
                if (ki == kj) {
                    temp = 2 * G * v[j * 3 + kj] * v[j * 3 + ki];
                    temp += lambda;
                } else {
                    temp = 2 * G * v[j * 3 + kj] * v[j * 3 + ki];
                    temp *= 2.;
                }
#pragma unroll
                for (int l = 0; l < 6; ++l) {
                    const int li = kival[l], lj = kjval[l];
                    T[l * 6 + k] += temp * grad[i * 3 + j] * v[i * 3 + li] * v[i * 3 + lj];
                }
            }
        }
    }
    double tempMat[9];
#pragma unroll
    for (int i = 0; i < 9; i++) tempMat[i] = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = i + 1; j < 3; ++j) {
            const double deps = eev[i] - eev[j];
            const double dsig = sev[i] - sev[j];
            double factor;
            if (!((fabs(deps) < 1.e-12) || (deps < 0.0)))
                factor = dsig / deps;
            else
                factor = G * (grad[i * 3 + i] - grad[i * 3 + j] - grad[j * 3 + i] + grad[j * 3 + j]);
#pragma unroll
            for (int it = 0; it < 3; ++it)
#pragma unroll
                for (int jt = 0; jt < 3; ++jt) tempMat[it * 3 + jt] += v[3 * i + it] * v[3 * j + jt];
#pragma unroll
            for (int it = 0; it < 3; ++it)
#pragma unroll
                for (int jt = 0; jt < 3; ++jt) tempMat[it * 3 + jt] += v[3 * j + it] * v[3 * i + jt];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int ki = kival[k], kj = kjval[k];
                double val;
                if (ki == kj)
                    val = v[j * 3 + ki] * v[i * 3 + kj] * factor;
                else
                    val = (v[j * 3 + ki] * v[i * 3 + kj] + v[j * 3 + kj] * v[i * 3 + ki]) * factor;
                // FromMatToVoight of val*tempMat (upper triangle, row by row)
                int kk = 0;
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = a; b < 3; b++) {
                        const double cv = val * tempMat[a * 3 + b];
                        const double aux = T[kk * 6 + k];
                        T[kk * 6 + k] = aux + cv;
                        kk++;
                    }
            }
        }
    }
    // TranslateDep: rows/cols XX, XY, YY; column XY halved; then * w
    const int ind[3] = {XX, XY, YY};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dep9[3 * i + 0] = T[XX + 6 * ind[i]];
        dep9[3 * i + 1] = T[XY + 6 * ind[i]];
        dep9[3 * i + 2] = T[YY + 6 * ind[i]];
    }
    dep9[1] /= 2;
    dep9[4] /= 2;
    dep9[7] /= 2;
#pragma unroll
    for (int i = 0; i < 9; i++) dep9[i] = dep9[i] * w;
}

// ---------------------------------------------- certified elastic points
//
// A test, from the trial stress alone, that the reference's return map would
// leave the point elastic: the closed-form principal stresses of the
// plane-strain trial stress (szz and the in-plane pair c +- R) put the main
// Mohr-Coulomb plane below zero by a margin that covers how far the
// reference's eigenvalues can sit from the exact ones.  Those come from at
// most 20 Newton steps on the characteristic cubic of the tensor normalised
// by its largest entry M (FunctionsSpectralDecomp.h:61-111), started outside
// the spectrum at the (quirked) Gershgorin bounds, at most M beyond the
// outermost eigenvalue: the iterates approach the largest / smallest root
// monotonically and shrink the error by 2/3 or better per step (the
// logarithmic derivative of a cubic with real roots), and the loop stops
// when |f| <= 1e-11 >= e^3.  So each computed eigenvalue (the third one by
// the trace included, after the sort) is within max((2/3)^20, 1e-11^(1/3)) M
// < 6e-4 M of an exact one, and the plane value within 2 * 6e-4 M.  A point
// with phi + 2e-3 M < 0 is therefore elastic in the reference (PhiPlane,
// FunctionsSigmaProjection.h:6-20: phi < 0 or |phi| < 1e-12): mtype 0, its
// stress the eigen-reconstruction of the trial stress, which equals the
// trial stress up to that reconstruction's rounding and the Newton residual
// of its eigenvalues.  Where two principal stresses nearly coincide the
// reference's eigenvector formulas (FunctionsSpectralDecomp.h:116-256)
// amplify that residual by the inverse gap and its stress departs from the
// trial stress (measured on the oracle: up to half of M for coincident
// eigenvalues, ~1e-9 M at gaps of 1e-2 M, ~1e-13 M typical), so the test
// also asks for gaps of at least 1e-2 M: such points keep the reference's
// values through the exact path.  Used only where nothing is committed.
// Returns true and ws3 = w * [sxx, sxy, syy] of the trial stress for a
// certified point; false (ws3 untouched) otherwise, NaN included.
IPF_HD bool certified_elastic(const double* e3, const double* ep4, const MatConst& m, double w, double* ws3) {
    const double exx = e3[0] - ep4[0];
    const double exy = e3[1] - ep4[1];
    const double eyy = e3[2] - ep4[2];
    const double ezz = 0 - ep4[3];
    const double sxx = m.lam2mu * exx + m.lambda * eyy + m.lambda * ezz;
    const double syy = m.lambda * exx + m.lam2mu * eyy + m.lambda * ezz;
    const double szz = m.lambda * exx + m.lambda * eyy + m.lam2mu * ezz;
    const double sxy = m.mu * exy;
    const double M = fmax(fmax(fabs(sxx), fabs(syy)), fmax(fabs(szz), fabs(sxy)));
    const double c = 0.5 * (sxx + syy), d = 0.5 * (sxx - syy);
    const double R = sqrt(d * d + sxy * sxy);
    if (!(c - R <= c + R) || !(szz == szz)) return false;  // NaN anywhere
    const double s1 = fmax(c + R, szz), s3 = fmin(c - R, szz);
    const double phi = (s1 - s3) + (s1 + s3) * m.sinphi - m.phi_c;
    if (!(phi + 2e-3 * M + 1e-12 * (M + m.phi_c) < 0.0)) return false;
    const double gap = fmin(2.0 * R, fmin(fabs(c + R - szz), fabs(c - R - szz)));
    if (!(gap >= 1e-2 * M)) return false;
    ws3[0] = sxx * w;
    ws3[1] = sxy * w;
    ws3[2] = syy * w;
    return true;
}

// ---------------------------------------------------------------- one point

// One integration point of ComputeSigma / ComputeSigmaDep
// (TPZConstitutiveLawProcessor.cpp:150-211, :364-439).
//   e3: [exx, gxy, eyy]; ep4: committed plastic strain [XX, XY, YY, ZZ],
//   updated in place; alpha: committed alpha, updated (alpha_n + d_alpha).
//   ws3: w * [sxx, sxy, syy]; sig4 (optional): [XX, XY, YY, ZZ] stress.
template <bool WANT_DEP>
IPF_HD void point_sigma(const double* e3, double* ep4, double& alpha, int& mtype, int& branch,
                                            const MatConst& m, double w, double* ws3, double* wdep9) {
    // ComposeStrain + ElasticStrain + ComputeTrialStress (FunctionsStressStrain.h:10-55)
    double el[6];
    el[XX] = e3[0] - ep4[0];
    el[XY] = e3[1] - ep4[1];
    el[XZ] = 0.0;
    el[YY] = e3[2] - ep4[2];
    el[YZ] = 0.0;
    el[ZZ] = 0 - ep4[3];
    double s[6];
    s[XX] = m.lam2mu * el[XX] + m.lambda * el[YY] + m.lambda * el[ZZ];
    s[YY] = m.lambda * el[XX] + m.lam2mu * el[YY] + m.lambda * el[ZZ];
    s[ZZ] = m.lambda * el[XX] + m.lambda * el[YY] + m.lam2mu * el[ZZ];
    s[XY] = m.mu * el[XY];
    s[XZ] = m.mu * el[XZ];
    s[YZ] = m.mu * el[YZ];
    double ev[3], vec[9];
    spectral_ps(s, ev, vec);  // s[XZ] = s[YZ] = mu * 0.0 = +0
    double sp[3];
    if (WANT_DEP) {
        double eev[3], evec[9], grad[9];
        el[XY] = el[XY] / 2;
        spectral_ps(el, eev, evec);  // el[XZ] = el[YZ] = +0
        project<true>(ev, sp, mtype, branch, alpha, m, grad);
        tangent3(grad, ev, eev, evec, m.G, m.lambda, w, wdep9);
    } else {
        project<false>(ev, sp, mtype, branch, alpha, m, nullptr);
    }
    // ReconstructStressTensor (FunctionsStressStrain.h:60-68), in-plane and ZZ
    const double sxx = (sp[0] * vec[0] * vec[0] + sp[1] * vec[3] * vec[3] + sp[2] * vec[6] * vec[6]);
    const double syy = (sp[0] * vec[1] * vec[1] + sp[1] * vec[4] * vec[4] + sp[2] * vec[7] * vec[7]);
    const double szz = (sp[0] * vec[2] * vec[2] + sp[1] * vec[5] * vec[5] + sp[2] * vec[8] * vec[8]);
    const double sxy = (sp[0] * vec[0] * vec[1] + sp[1] * vec[3] * vec[4] + sp[2] * vec[6] * vec[7]);
    // ComputeStrain + PlasticStrain (FunctionsStressStrain.h:73-81, :38-42)
    const double exx = m.cs_a * sxx - m.cs_b * (syy + szz);
    const double eyy = m.cs_a * syy - m.cs_b * (sxx + szz);
    const double ezz = m.cs_a * szz - m.cs_b * (syy + sxx);
    const double exy = m.inv_mu * sxy;
    ep4[0] = e3[0] - exx;
    ep4[1] = e3[1] - exy;
    ep4[2] = e3[2] - eyy;
    ep4[3] = 0 - ezz;
    // TranslateStress * w (FunctionsStressStrain.h:86-98; TPZConstitutiveLawProcessor.cpp:194)
    ws3[0] = sxx * w;
    ws3[1] = sxy * w;
    ws3[2] = syy * w;
}

}  // namespace ipf
