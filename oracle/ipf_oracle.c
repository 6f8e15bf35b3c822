/*
 * ipf_oracle.c -- CPU oracle (plain C restatement) of IntPointsFEM's
 * integration-point residual path.  TEST INFRASTRUCTURE ONLY: see
 * ipf_oracle.h.  Compile with -ffp-contract=off and without fast-math so the
 * floating-point operation order is exactly the one written here.
 *
 * Parity status of each function (what pins it, see tests/):
 *   orc_spectral_decomposition  pinned: Eigenvectors.nb KAT, numpy eigh,
 *                               reconstruction/trace invariants
 *   orc_project_sigma           pinned: closed forms per branch, yield
 *                               consistency, brute-force convex projection
 *   orc_tangent_operator        pinned: elastic Dep == D exactly, finite
 *                               differences of orc_point_sigma on the normal
 *                               columns of every plastic branch, frame
 *                               covariance under rotations about z (the
 *                               shear / eigenvector-rotation term)
 *   orc_point_sigma / orc_compute_sigma[_dep]  pinned through the above
 *   orc_multiply_vector / orc_multiply / orc_multiply_transpose  pinned:
 *                               dense numpy products, patch tests
 *   orc_color_elements          pinned: coloring invariant, 4 colors on
 *                               structured quads (Results/ logs)
 *   orc_element_tangents        pinned: numpy B^T D B per element
 *   Summation order of the block products follows MKL's dgemm in the
 *   reference, which is not observable here: parity unpinned at the bit
 *   level (documented; the HIP product uses the same fixed order).
 */
#include "ipf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XX ORC_XX
#define XY ORC_XY
#define XZ ORC_XZ
#define YY ORC_YY
#define YZ ORC_YZ
#define ZZ ORC_ZZ

/* ------------------------------------------------------------------ */
/* FunctionsStressStrain.h                                             */
/* ------------------------------------------------------------------ */

/* FunctionsStressStrain.h:10-24 (dim == 2 branch) */
static void compose_strain(const double *delta_strain, double *full) {
    full[XX] = delta_strain[0];
    full[XY] = delta_strain[1];
    full[XZ] = 0;
    full[YY] = delta_strain[2];
    full[YZ] = 0;
    full[ZZ] = 0;
}

/* FunctionsStressStrain.h:29-33 */
static void elastic_strain_fn(const double *plastic, const double *strain, double *elastic) {
    for (int i = 0; i < 6; i++) elastic[i] = strain[i] - plastic[i];
}

/* FunctionsStressStrain.h:38-42 */
static void plastic_strain_fn(const double *strain, const double *elastic, double *plastic) {
    for (int i = 0; i < 6; i++) plastic[i] = strain[i] - elastic[i];
}

/* FunctionsStressStrain.h:47-55 */
static void trial_stress(const double *e, double *s, double mu, double lambda) {
    s[XX] = (lambda + 2 * mu) * e[XX] + lambda * e[YY] + lambda * e[ZZ];
    s[YY] = lambda * e[XX] + (lambda + 2 * mu) * e[YY] + lambda * e[ZZ];
    s[ZZ] = lambda * e[XX] + lambda * e[YY] + (lambda + 2 * mu) * e[ZZ];
    s[XY] = mu * e[XY];
    s[XZ] = mu * e[XZ];
    s[YZ] = mu * e[YZ];
}

/* FunctionsStressStrain.h:60-68 */
static void reconstruct_stress(const double *sp, const double *v, double *s) {
    s[XX] = (sp[0] * v[0] * v[0] + sp[1] * v[3] * v[3] + sp[2] * v[6] * v[6]);
    s[YY] = (sp[0] * v[1] * v[1] + sp[1] * v[4] * v[4] + sp[2] * v[7] * v[7]);
    s[ZZ] = (sp[0] * v[2] * v[2] + sp[1] * v[5] * v[5] + sp[2] * v[8] * v[8]);
    s[XY] = (sp[0] * v[0] * v[1] + sp[1] * v[3] * v[4] + sp[2] * v[6] * v[7]);
    s[XZ] = (sp[0] * v[0] * v[2] + sp[1] * v[3] * v[5] + sp[2] * v[6] * v[8]);
    s[YZ] = (sp[0] * v[1] * v[2] + sp[1] * v[4] * v[5] + sp[2] * v[7] * v[8]);
}

/* FunctionsStressStrain.h:73-81 */
static void compute_strain(const double *s, double *e, double mu, double lambda) {
    e[XX] = (lambda + mu) / (3 * lambda * mu + 2 * mu * mu) * s[XX] - lambda / (6 * lambda * mu + 4 * mu * mu) * (s[YY] + s[ZZ]);
    e[YY] = (lambda + mu) / (3 * lambda * mu + 2 * mu * mu) * s[YY] - lambda / (6 * lambda * mu + 4 * mu * mu) * (s[XX] + s[ZZ]);
    e[ZZ] = (lambda + mu) / (3 * lambda * mu + 2 * mu * mu) * s[ZZ] - lambda / (6 * lambda * mu + 4 * mu * mu) * (s[YY] + s[XX]);
    e[XY] = 1 / mu * s[XY];
    e[XZ] = 1 / mu * s[XZ];
    e[YZ] = 1 / mu * s[YZ];
}

/* FunctionsStressStrain.h:86-98 */
static void translate_stress(const double *full, double *s3) {
    s3[0] = full[XX];
    s3[1] = full[XY];
    s3[2] = full[YY];
}

/* FunctionsStressStrain.h:103-119 */
static void translate_dep(const double *full_dep, double *dep) {
    const unsigned ind[3] = {XX, XY, YY};
    for (int i = 0; i < 3; ++i) {
        dep[3 * i + 0] = full_dep[XX + 6 * ind[i]];
        dep[3 * i + 1] = full_dep[XY + 6 * ind[i]];
        dep[3 * i + 2] = full_dep[YY + 6 * ind[i]];
    }
    dep[1] /= 2;
    dep[4] /= 2;
    dep[7] /= 2;
}

/* ------------------------------------------------------------------ */
/* FunctionsSpectralDecomp.h                                           */
/* ------------------------------------------------------------------ */

/* FunctionsSpectralDecomp.h:7-20 */
static void normalize(double *s, double *maxel) {
    *maxel = s[0];
    for (int i = 1; i < 6; i++)
        if (fabs(s[i]) > fabs(*maxel)) *maxel = s[i];
    if (*maxel != 0)
        for (int i = 0; i < 6; i++) s[i] /= *maxel;
}

/* FunctionsSpectralDecomp.h:25-56.  Row 3 uses sxy + syz as written there. */
static void interval(const double *s, double *iv) {
    double lower[3], upper[3];
    double row_1 = s[XY] + s[XZ];
    lower[0] = s[XX] - fabs(row_1);
    upper[0] = s[XX] + fabs(row_1);
    double row_2 = s[XY] + s[YZ];
    lower[1] = s[YY] - fabs(row_2);
    upper[1] = s[YY] + fabs(row_2);
    double row_3 = s[XY] + s[YZ];
    lower[2] = s[ZZ] - fabs(row_3);
    upper[2] = s[ZZ] + fabs(row_3);
    iv[0] = upper[0];
    iv[1] = lower[0];
    for (int i = 1; i < 3; i++) {
        if (upper[i] > iv[0]) iv[0] = upper[i];
        if (lower[i] < iv[1]) iv[1] = lower[i];
    }
}

/* Characteristic polynomial of the normalised tensor, FunctionsSpectralDecomp.h:75-77 / :85-87 */
static double charpoly(double x, const double *s, double sxy2, double sxz2, double syz2) {
    return -x * x * x - sxz2 * s[YY] - sxy2 * s[ZZ] - syz2 * s[XX] + 2 * s[XY] * s[XZ] * s[YZ] +
           s[XX] * s[YY] * s[ZZ] + x * x * (s[XX] + s[YY] + s[ZZ]) +
           x * (sxy2 + sxz2 + syz2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ]);
}

/* FunctionsSpectralDecomp.h:61-111 (tol written 10e-12 there, i.e. 1e-11) */
static void newton_iterations(const double *iv, const double *s, double *ev, double maxel) {
    const int numiterations = 20;
    const double tol = 10e-12;
    for (int i = 0; i < 2; i++) {
        double x = iv[i];
        int it = 0;
        double sxy2 = s[XY] * s[XY];
        double sxz2 = s[XZ] * s[XZ];
        double syz2 = s[YZ] * s[YZ];
        double f = charpoly(x, s, sxy2, sxz2, syz2);
        double res = fabs(f);
        while (it < numiterations && res > tol) {
            double df = -3 * x * x + sxy2 + sxz2 + syz2 - s[XX] * s[YY] - s[XX] * s[ZZ] - s[YY] * s[ZZ] +
                        2 * x * (s[XX] + s[YY] + s[ZZ]);
            x -= f / df;
            f = charpoly(x, s, sxy2, sxz2, syz2);
            res = fabs(f);
            it++;
        }
        ev[i] = x;
    }
    ev[2] = s[XX] + s[YY] + s[ZZ] - ev[0] - ev[1];
    ev[0] *= maxel;
    ev[1] *= maxel;
    ev[2] *= maxel;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[i] < ev[j]) {
                double a = ev[i];
                ev[i] = ev[j];
                ev[j] = a;
            }
}

/* FunctionsSpectralDecomp.h:116-150 */
static void multiplicity1(const double *s, double lam, double *vec) {
    double det[3];
    det[0] = (s[XX] - lam) * (s[YY] - lam) - s[XY] * s[XY];
    det[1] = (s[XX] - lam) * (s[ZZ] - lam) - s[XZ] * s[XZ];
    det[2] = (s[YY] - lam) * (s[ZZ] - lam) - s[YZ] * s[YZ];
    double maxdet = fabs(det[0]);
    for (int i = 1; i < 3; i++)
        if (fabs(det[i]) > fabs(maxdet)) maxdet = fabs(det[i]);
    double v[3];
    if (maxdet == fabs(det[0])) {
        v[0] = 1 / det[0] * (-(s[YY] - lam) * s[XZ] + s[XY] * s[YZ]);
        v[1] = 1 / det[0] * (-(s[XX] - lam) * s[YZ] + s[XY] * s[XZ]);
        v[2] = 1;
    } else if (maxdet == fabs(det[1])) {
        v[0] = 1 / det[1] * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v[1] = 1;
        v[2] = 1 / det[1] * (-(s[XX] - lam) * s[YZ] + s[XY] * s[XZ]);
    } else {
        v[0] = 1;
        v[1] = 1 / det[2] * (-(s[ZZ] - lam) * s[XY] + s[XZ] * s[YZ]);
        v[2] = 1 / det[2] * (-(s[YY] - lam) * s[XZ] + s[XY] * s[YZ]);
    }
    double norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    vec[0] = v[0] / norm;
    vec[1] = v[1] / norm;
    vec[2] = v[2] / norm;
}

/* FunctionsSpectralDecomp.h:155-212 (norm2 keeps the v2[1]*v1[1] term as written at :203) */
static void multiplicity2(const double *s, double lam, double *vec1, double *vec2) {
    double x[3];
    x[0] = s[XX] - lam;
    x[1] = s[YY] - lam;
    x[2] = s[ZZ] - lam;
    double maxx = fabs(x[0]);
    for (int i = 1; i < 3; i++)
        if (fabs(x[i]) > fabs(maxx)) maxx = fabs(x[i]);
    double v1[3], v2[3];
    if (maxx == fabs(x[0])) {
        v1[0] = -s[XY] / x[0]; v1[1] = 1; v1[2] = 0;
        v2[0] = -s[XZ] / x[0]; v2[1] = 0; v2[2] = 1;
    } else if (maxx == fabs(x[1])) {
        v1[0] = 1; v1[1] = -s[XY] / x[1]; v1[2] = 0;
        v2[0] = 0; v2[1] = -s[YZ] / x[1]; v2[2] = 1;
    } else {
        v1[0] = 1; v1[1] = 0; v1[2] = -s[XZ] / x[2];
        v2[0] = 0; v2[1] = 1; v2[2] = -s[YZ] / x[2];
    }
    double norm1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    double norm2 = sqrt(v2[0] * v2[0] + v2[1] * v1[1] + v2[2] * v2[2]);
    vec1[0] = v1[0] / norm1; vec1[1] = v1[1] / norm1; vec1[2] = v1[2] / norm1;
    vec2[0] = v2[0] / norm2; vec2[1] = v2[1] / norm2; vec2[2] = v2[2] / norm2;
}

/* FunctionsSpectralDecomp.h:217-256 */
static void eigenvectors_fn(double *s, const double *ev, double *vec, double maxel) {
    s[XX] *= maxel;
    s[YY] *= maxel;
    s[ZZ] *= maxel;
    s[XY] *= maxel;
    s[XZ] *= maxel;
    s[YZ] *= maxel;
    if ((ev[0] == ev[1]) && (ev[1] == ev[2])) {
        for (int i = 0; i < 9; i++) vec[i] = 0.;
        vec[0] = 1.;
        vec[4] = 1.;
        vec[8] = 1.;
    } else {
        if (ev[0] != ev[1] && ev[0] != ev[2]) {
            multiplicity1(s, ev[0], &vec[0]);
        } else if (ev[0] == ev[1]) {
            multiplicity2(s, ev[0], &vec[0], &vec[3]);
        } else if (ev[0] == ev[2]) {
            multiplicity2(s, ev[0], &vec[0], &vec[6]);
        }
        if (ev[1] != ev[0] && ev[1] != ev[2]) {
            multiplicity1(s, ev[1], &vec[3]);
        } else if (ev[1] == ev[2]) {
            multiplicity2(s, ev[1], &vec[3], &vec[6]);
        }
        if (ev[2] != ev[0] && ev[2] != ev[1]) {
            multiplicity1(s, ev[2], &vec[6]);
        }
    }
}

void orc_eigenvectors(double *sigma6, const double *eigenvalues3, double *eigenvectors9, double maxel) {
    eigenvectors_fn(sigma6, eigenvalues3, eigenvectors9, maxel);
}

/* FunctionsSpectralDecomp.h:261-268.  sigma6 is normalised and scaled back
 * in place, exactly as the reference does. */
void orc_spectral_decomposition(double *sigma6, double *eigenvalues3, double *eigenvectors9) {
    double maxel;
    double iv[2];
    normalize(sigma6, &maxel);
    interval(sigma6, iv);
    newton_iterations(iv, sigma6, eigenvalues3, maxel);
    eigenvectors_fn(sigma6, eigenvalues3, eigenvectors9, maxel);
}

/* ------------------------------------------------------------------ */
/* FunctionsSigmaProjection.h                                          */
/* ------------------------------------------------------------------ */

/* FunctionsSigmaProjection.h:6-20 */
static int phi_plane(const double *ev, double *sp, double mc_phi, double mc_cohesion) {
    const double sinphi = sin(mc_phi);
    const double cosphi = cos(mc_phi);
    double phi = ev[0] - ev[2] + (ev[0] + ev[2]) * sinphi - 2. * mc_cohesion * cosphi;
    sp[0] = ev[0];
    sp[1] = ev[1];
    sp[2] = ev[2];
    return (fabs(phi) < 1.e-12) || (phi < 0.0);
}

/* FunctionsSigmaProjection.h:25-63 */
static int return_main_plane(const double *st, double *sp, double *m_hardening, double mc_phi, double mc_psi,
                             double mc_cohesion, double K, double G) {
    double ev[3] = {st[0], st[1], st[2]};
    const double sinphi = sin(mc_phi);
    const double sinpsi = sin(mc_psi);
    const double cosphi = cos(mc_phi);
    const double sinphi2 = sinphi * sinphi;
    const double cosphi2 = 1. - sinphi2;
    const double constA = 4. * G * (1. + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    double phi = ev[0] - ev[2] + (ev[0] + ev[2]) * sinphi - 2. * mc_cohesion * cosphi;
    double gamma = 0;
    const int n_iterations = 30;
    for (int i = 0; i < n_iterations; i++) {
        double jac = -constA - 4. * cosphi2 * 0;
        double delta_gamma = -phi / jac;
        gamma += delta_gamma;
        phi = ev[0] - ev[2] + (ev[0] + ev[2]) * sinphi - 2. * mc_cohesion * cosphi - constA * gamma;
        if (fabs(phi) < 1.e-12) break;
    }
    ev[0] -= (2. * G * (1 + sinpsi / 3.) + 2. * K * sinpsi) * gamma;
    ev[1] += (4. * G / 3. - K * 2.) * sinpsi * gamma;
    ev[2] += (2. * G * (1 - sinpsi / 3.) - 2. * K * sinpsi) * gamma;
    sp[0] = ev[0];
    sp[1] = ev[1];
    sp[2] = ev[2];
    *m_hardening += gamma * 2. * cosphi;
    return (ev[0] > ev[1] || fabs(ev[0] - ev[1]) < 1.e-12) && (ev[1] > ev[2] || fabs(ev[1] - ev[2]) < 1.e-12);
}

/* FunctionsSigmaProjection.h:68-88 */
static void plane_tangent(double *t, double mc_phi, double mc_psi, double K, double G) {
    const double sin_phi = sin(mc_phi);
    const double sin_psi = sin(mc_psi);
    const double den = 6.0 * G + 2.0 * (G + 3.0 * K) * sin_phi * sin_psi;
    t[0] = (sin_phi - 1.0) * (-3.0 * G + (G + 3.0 * K) * sin_psi) / den;
    t[3] = (2.0 * G - 3.0 * K) * (sin_phi + 1.0) * sin_psi / den;
    t[6] = -(sin_phi + 1.0) * (-3.0 * G + (G + 3.0 * K) * sin_psi) / den;
    t[1] = 0.0;
    t[4] = 1.0;
    t[7] = 0.0;
    t[2] = -(sin_phi - 1.0) * (3.0 * G + (G + 3.0 * K) * sin_psi) / den;
    t[5] = (2.0 * G - 3.0 * K) * (sin_phi - 1.0) * sin_psi / den;
    t[8] = (sin_phi + 1.0) * (3.0 * G + (G + 3.0 * K) * sin_psi) / den;
}

/* FunctionsSigmaProjection.h:93-158.  DEVIATION: gamma[2] is uninitialised
 * in the reference (:103); the oracle starts from gamma = {0, 0}. */
static int return_right_edge(const double *st, double *sp, double *m_hardening, double mc_phi, double mc_psi,
                             double mc_cohesion, double K, double G) {
    double ev[3] = {st[0], st[1], st[2]};
    const double sinphi = sin(mc_phi);
    const double sinpsi = sin(mc_psi);
    const double cosphi = cos(mc_phi);
    double gamma[2] = {0., 0.}, phi[2], sigma_bar[2], ab[2];
    double jac[2][2], jac_inv[2][2];
    sigma_bar[0] = ev[0] - ev[2] + (ev[0] + ev[2]) * sinphi;
    sigma_bar[1] = ev[0] - ev[1] + (ev[0] + ev[1]) * sinphi;
    phi[0] = sigma_bar[0] - 2. * cosphi * mc_cohesion;
    phi[1] = sigma_bar[1] - 2. * cosphi * mc_cohesion;
    ab[0] = 4. * G * (1 + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    ab[1] = 2. * G * (1. + sinphi + sinpsi - sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    const int n_iterations = 30;
    for (int i = 0; i < n_iterations; i++) {
        jac[0][0] = -ab[0];
        jac[1][0] = -ab[1];
        jac[0][1] = -ab[1];
        jac[1][1] = -ab[0];
        double det_jac = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        jac_inv[0][0] = jac[1][1] / det_jac;
        jac_inv[1][0] = -jac[1][0] / det_jac;
        jac_inv[0][1] = -jac[0][1] / det_jac;
        jac_inv[1][1] = jac[0][0] / det_jac;
        gamma[0] -= (jac_inv[0][0] * phi[0] + jac_inv[0][1] * phi[1]);
        gamma[1] -= (jac_inv[1][0] * phi[0] + jac_inv[1][1] * phi[1]);
        phi[0] = sigma_bar[0] - ab[0] * gamma[0] - ab[1] * gamma[1] - 2. * cosphi * mc_cohesion;
        phi[1] = sigma_bar[1] - ab[1] * gamma[0] - ab[0] * gamma[1] - 2. * cosphi * mc_cohesion;
        double res = (fabs(phi[0]) + fabs(phi[1]));
        if (fabs(res) < 1.e-12) break;
    }
    ev[0] -= (2. * G * (1 + sinpsi / 3.) + 2. * K * sinpsi) * (gamma[0] + gamma[1]);
    ev[1] += ((4. * G / 3. - K * 2.) * sinpsi) * gamma[0] + (2. * G * (1. - sinpsi / 3.) - 2. * K * sinpsi) * gamma[1];
    ev[2] += (2. * G * (1 - sinpsi / 3.) - 2. * K * sinpsi) * gamma[0] + ((4. * G / 3. - 2. * K) * sinpsi) * gamma[1];
    sp[0] = ev[0];
    sp[1] = ev[1];
    sp[2] = ev[2];
    *m_hardening += (gamma[0] + gamma[1]) * 2. * cosphi;
    return (ev[0] > ev[1] || fabs(ev[0] - ev[1]) < 1.e-12) && (ev[1] > ev[2] || fabs(ev[1] - ev[2]) < 1.e-12);
}

/* FunctionsSigmaProjection.h:163-188 */
static void right_edge_tangent(double *t, double mc_phi, double mc_psi, double K, double G) {
    const double sin_phi = sin(mc_phi);
    const double sin_psi = sin(mc_psi);
    const double a = 4.0 * G * (1.0 + (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
    const double b = 2.0 * G * (1.0 + sin_phi + sin_psi - (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
    t[0] = (3.0 * a + 3.0 * b - 4.0 * (1.0 + sin_phi) * (3.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
    t[3] = (2.0 * (1.0 + sin_phi) * (-6.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
    t[6] = (2.0 * (1.0 + sin_phi) * (-6.0 * K * sin_psi + G * (3.0 + sin_psi))) / (3. * (a + b));
    t[1] = (-2 * (-1 + sin_phi) * (3 * K * sin_psi + G * (3 + sin_psi))) / (3. * (a + b));
    t[4] = (-3 * b * b + 3 * a * (a + 2 * G * (-1 + sin_phi)) -
            2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (-1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
    t[7] = (2 * (-1 + sin_phi) * (b * G * (-3 + sin_psi) + a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi)) /
           (3. * (a * a - b * b));
    t[2] = (-2 * (-1 + sin_phi) * (3 * K * sin_psi + G * (3 + sin_psi))) / (3. * (a + b));
    t[5] = (2 * (-1 + sin_phi) * (b * G * (-3 + sin_psi) + a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi)) /
           (3. * (a * a - b * b));
    t[8] = (-3 * b * b + 3 * a * (a + 2 * G * (-1 + sin_phi)) -
            2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (-1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
}

/* FunctionsSigmaProjection.h:193-257.  DEVIATION: gamma[2] starts at {0, 0}
 * (uninitialised in the reference, :205). */
static int return_left_edge(const double *st, double *sp, double *m_hardening, double mc_phi, double mc_psi,
                            double mc_cohesion, double K, double G) {
    double ev[3] = {st[0], st[1], st[2]};
    const double sinphi = sin(mc_phi);
    const double sinpsi = sin(mc_psi);
    const double cosphi = cos(mc_phi);
    const double sinphi2 = sinphi * sinphi;
    const double cosphi2 = 1. - sinphi2;
    double gamma[2] = {0., 0.}, phi[2], sigma_bar[2], ab[2];
    double jac[2][2], jac_inv[2][2];
    sigma_bar[0] = ev[0] - ev[2] + (ev[0] + ev[2]) * sinphi;
    sigma_bar[1] = ev[1] - ev[2] + (ev[1] + ev[2]) * sinphi;
    ab[0] = 4. * G * (1 + sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    ab[1] = 2. * G * (1. - sinphi - sinpsi - sinphi * sinpsi / 3.) + 4. * K * sinphi * sinpsi;
    phi[0] = sigma_bar[0] - 2. * cosphi * mc_cohesion;
    phi[1] = sigma_bar[1] - 2. * cosphi * mc_cohesion;
    const int n_iterations = 30;
    for (int i = 0; i < n_iterations; i++) {
        jac[0][0] = -ab[0] - 4. * cosphi2 * 0;
        jac[1][0] = -ab[1] - 4. * cosphi2 * 0;
        jac[0][1] = -ab[1] - 4. * cosphi2 * 0;
        jac[1][1] = -ab[0] - 4. * cosphi2 * 0;
        double det_jac = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        jac_inv[0][0] = jac[1][1] / det_jac;
        jac_inv[1][0] = -jac[1][0] / det_jac;
        jac_inv[0][1] = -jac[0][1] / det_jac;
        jac_inv[1][1] = jac[0][0] / det_jac;
        gamma[0] -= (jac_inv[0][0] * phi[0] + jac_inv[0][1] * phi[1]);
        gamma[1] -= (jac_inv[1][0] * phi[0] + jac_inv[1][1] * phi[1]);
        phi[0] = sigma_bar[0] - ab[0] * gamma[0] - ab[1] * gamma[1] - 2. * cosphi * mc_cohesion;
        phi[1] = sigma_bar[1] - ab[1] * gamma[0] - ab[0] * gamma[1] - 2. * cosphi * mc_cohesion;
        double res = (fabs(phi[0]) + fabs(phi[1]));
        if (fabs(res) < 1.e-12) break;
    }
    ev[0] += -(2. * G * (1 + sinpsi / 3.) + 2. * K * sinpsi) * gamma[0] + ((4. * G / 3. - 2. * K) * sinpsi) * gamma[1];
    ev[1] += ((4. * G / 3. - K * 2.) * sinpsi) * gamma[0] - (2. * G * (1. + sinpsi / 3.) + 2. * K * sinpsi) * gamma[1];
    ev[2] += (2. * G * (1 - sinpsi / 3.) - 2. * K * sinpsi) * (gamma[0] + gamma[1]);
    sp[0] = ev[0];
    sp[1] = ev[1];
    sp[2] = ev[2];
    *m_hardening += (gamma[0] + gamma[1]) * 2. * cosphi;
    return (ev[0] > ev[1] || fabs(ev[0] - ev[1]) < 1.e-12) && (ev[1] > ev[2] || fabs(ev[1] - ev[2]) < 1.e-12);
}

/* FunctionsSigmaProjection.h:262-287 */
static void left_edge_tangent(double *t, double mc_phi, double mc_psi, double K, double G) {
    const double sin_phi = sin(mc_phi);
    const double sin_psi = sin(mc_psi);
    const double a = 4.0 * G * (1.0 + (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
    const double b = 2.0 * G * (1.0 - sin_phi - sin_psi - (1.0 / 3.0) * sin_phi * sin_psi) + 4.0 * K * sin_phi * sin_psi;
    t[0] = (-3 * b * b + 3 * a * (a - 2 * G * (1 + sin_phi)) -
            2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
    t[3] = (2 * (1 + sin_phi) * (a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi + b * G * (3 + sin_psi))) /
           (3. * (a * a - b * b));
    t[6] = (-2 * (1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
    t[1] = (2 * (1 + sin_phi) * (a * (2 * G - 3 * K) * sin_psi + 3 * b * K * sin_psi + b * G * (3 + sin_psi))) /
           (3. * (a * a - b * b));
    t[4] = (-3 * b * b + 3 * a * (a - 2 * G * (1 + sin_phi)) -
            2 * (a * G + 2 * b * G + 3 * a * K - 3 * b * K) * (1 + sin_phi) * sin_psi) / (3. * (a - b) * (a + b));
    t[7] = (-2 * (1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
    t[2] = (2 * (-1 + sin_phi) * (G * (-3 + sin_psi) - 6 * K * sin_psi)) / (3. * (a + b));
    t[5] = (2 * (-1 + sin_phi) * (G * (-3 + sin_psi) - 6 * K * sin_psi)) / (3. * (a + b));
    t[8] = (3 * a + 3 * b - 4 * (-1 + sin_phi) * (G * (-3 + sin_psi) + 3 * K * sin_psi)) / (3. * (a + b));
}

/* FunctionsSigmaProjection.h:292-323 */
static void return_apex(const double *st, double *sp, double *m_hardening, double mc_phi, double mc_psi,
                        double mc_cohesion, double K) {
    const double cotphi = 1. / tan(mc_phi);
    double ptrnp1 = 0.;
    for (int i = 0; i < 3; i++) ptrnp1 += st[i];
    ptrnp1 /= 3.;
    double DEpsPV = 0.;
    double alpha = cos(mc_phi) / sin(mc_psi);
    double res = mc_cohesion * cotphi - ptrnp1;
    double pnp1 = 0.;
    const int n_iterations = 30;
    for (int i = 0; i < n_iterations; i++) {
        const double jac = K;
        DEpsPV -= res / jac;
        pnp1 = ptrnp1 - K * DEpsPV;
        res = mc_cohesion * cotphi - pnp1;
        if (fabs(res) < 1.e-12) break;
    }
    *m_hardening += alpha * DEpsPV;
    for (int i = 0; i < 3; i++) sp[i] = pnp1;
}

/* FunctionsSigmaProjection.h:337-379.  branch adds the 5-way id that the
 * reference's mtype (0 / 1 / -1) folds together. */
void orc_project_sigma(const double *ev, double *sp, int *m_type, double *alpha, double mc_phi, double mc_psi,
                       double mc_cohesion, double K, double G, double *gradient, int require_gradient,
                       int *branch) {
    int check = 0;
    *m_type = 0;
    *branch = ORC_BR_ELASTIC;
    check = phi_plane(ev, sp, mc_phi, mc_cohesion);
    if (check) {
        if (require_gradient) {
            for (int i = 0; i < 9; i++) gradient[i] = 0.;
            gradient[0] = 1.;
            gradient[4] = 1.;
            gradient[8] = 1.;
        }
        return;
    }
    *m_type = 1;
    *branch = ORC_BR_MAIN;
    check = return_main_plane(ev, sp, alpha, mc_phi, mc_psi, mc_cohesion, K, G);
    if (check && require_gradient) plane_tangent(gradient, mc_phi, mc_psi, K, G);
    if (!check) {
        if (((1 - sin(mc_psi)) * ev[0] - 2. * ev[1] + (1 + sin(mc_psi)) * ev[2]) > 0) {
            *branch = ORC_BR_RIGHT;
            check = return_right_edge(ev, sp, alpha, mc_phi, mc_psi, mc_cohesion, K, G);
            if (check && require_gradient) right_edge_tangent(gradient, mc_phi, mc_psi, K, G);
        } else {
            *branch = ORC_BR_LEFT;
            check = return_left_edge(ev, sp, alpha, mc_phi, mc_psi, mc_cohesion, K, G);
            if (check && require_gradient) left_edge_tangent(gradient, mc_phi, mc_psi, K, G);
        }
        if (!check) {
            *m_type = -1;
            *branch = ORC_BR_APEX;
            return_apex(ev, sp, alpha, mc_phi, mc_psi, mc_cohesion, K);
            if (require_gradient)
                for (int i = 0; i < 9; i++) gradient[i] = 0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* FunctionsTangentOperator.h                                          */
/* ------------------------------------------------------------------ */

/* FunctionsTangentOperator.h:37-123 */
void orc_tangent_operator(const double *gradient, const double *stress_ev, const double *strain_ev,
                          const double *strain_vec, double *T, double G, double lambda) {
    static const unsigned kival[6] = {0, 0, 0, 1, 1, 2};
    static const unsigned kjval[6] = {0, 1, 2, 1, 2, 2};
    for (int l = 0; l < 36; ++l) T[l] = 0;
    for (unsigned k = 0; k < 6; ++k) {
        const unsigned ki = kival[k];
        const unsigned kj = kjval[k];
        for (unsigned i = 0; i < 3; ++i) {
            for (unsigned j = 0; j < 3; ++j) {
                double temp;
                if (ki == kj) {
                    temp = 2 * G * strain_vec[j * 3 + kj] * strain_vec[j * 3 + ki];
                    temp += lambda;
                } else {
                    temp = 2 * G * strain_vec[j * 3 + kj] * strain_vec[j * 3 + ki];
                    temp *= 2.;
                }
                for (int l = 0; l < 6; ++l) {
                    const unsigned li = kival[l];
                    const unsigned lj = kjval[l];
                    T[l * 6 + k] += temp * gradient[i * 3 + j] * strain_vec[i * 3 + li] * strain_vec[i * 3 + lj];
                }
            }
        }
    }
    double tempMat[9], temp_mat[9], ColCorrV[6];
    for (int i = 0; i < 9; i++) tempMat[i] = 0;
    for (unsigned i = 0; i < 2; ++i) {
        for (unsigned j = i + 1; j < 3; ++j) {
            double deigeneps = strain_ev[i] - strain_ev[j];
            double deigensig = stress_ev[i] - stress_ev[j];
            double factor = 0.;
            if (!((fabs(deigeneps) < 1.e-12) || (deigeneps < 0.0))) {
                factor = deigensig / deigeneps;
            } else {
                factor = G * (gradient[i * 3 + i] - gradient[i * 3 + j] - gradient[j * 3 + i] + gradient[j * 3 + j]);
            }
            /* ProdT(v_i, v_j) and ProdT(v_j, v_i) accumulated into tempMat (:89-102) */
            for (unsigned it = 0; it < 3; ++it)
                for (unsigned jt = 0; jt < 3; ++jt)
                    temp_mat[it * 3 + jt] = strain_vec[3 * i + it] * strain_vec[3 * j + jt];
            for (unsigned it = 0; it < 9; ++it) tempMat[it] += temp_mat[it];
            for (unsigned it = 0; it < 3; ++it)
                for (unsigned jt = 0; jt < 3; ++jt)
                    temp_mat[it * 3 + jt] = strain_vec[3 * j + it] * strain_vec[3 * i + jt];
            for (unsigned it = 0; it < 9; ++it) tempMat[it] += temp_mat[it];
            for (unsigned k = 0; k < 6; ++k) {
                const unsigned ki = kival[k];
                const unsigned kj = kjval[k];
                double val;
                if (ki == kj) {
                    val = strain_vec[j * 3 + ki] * strain_vec[i * 3 + kj] * factor;
                } else {
                    val = (strain_vec[j * 3 + ki] * strain_vec[i * 3 + kj] + strain_vec[j * 3 + kj] * strain_vec[i * 3 + ki]) * factor;
                }
                for (int m = 0; m < 9; m++) temp_mat[m] = val * tempMat[m];
                /* FromMatToVoight (:28-35) */
                int kk = 0;
                for (int a = 0; a < 3; a++)
                    for (int b = a; b < 3; b++) ColCorrV[kk++] = temp_mat[a * 3 + b];
                for (int l = 0; l < 6; l++) {
                    double aux = T[l * 6 + k];
                    T[l * 6 + k] = aux + ColCorrV[l];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* TPZConstitutiveLawProcessor                                         */
/* ------------------------------------------------------------------ */

/* One point of ComputeSigma (TPZConstitutiveLawProcessor.cpp:150-211) or, when
 * dep9_out != NULL, of ComputeSigmaDep (:364-439).  DEVIATION: the local
 * alpha is uninitialised in the reference (:160, :374); here it starts from
 * the committed value, so alpha_{n+1} = alpha_n + d_alpha. */
void orc_point_sigma(const double *strain3, double *plastic_strain6, double *alpha, int *mtype, int *branch,
                     double weight, double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion,
                     double *wsigma3, double *sigma6_out, double *dep9_out) {
    const double G = mu;
    const double K = lambda + 2 * mu / 3;
    double strain[6], elastic[6], sigma[6], plastic[6], aux[3];
    double ev[3], evec[9], sp[3], gradient[9];
    compose_strain(strain3, strain);
    for (int i = 0; i < 6; i++) plastic[i] = plastic_strain6[i];
    elastic_strain_fn(plastic, strain, elastic);
    trial_stress(elastic, sigma, mu, lambda);
    orc_spectral_decomposition(sigma, ev, evec);
    if (dep9_out) {
        double strain_ev[3], strain_vec[9], dep[36];
        elastic[1] = elastic[1] / 2;
        orc_spectral_decomposition(elastic, strain_ev, strain_vec);
        orc_project_sigma(ev, sp, mtype, alpha, mc_phi, mc_psi, mc_cohesion, K, G, gradient, 1, branch);
        orc_tangent_operator(gradient, ev, strain_ev, strain_vec, dep, G, lambda);
        translate_dep(dep, gradient);
        for (int i = 0; i < 9; i++) dep9_out[i] = gradient[i] * weight;
    } else {
        orc_project_sigma(ev, sp, mtype, alpha, mc_phi, mc_psi, mc_cohesion, K, G, gradient, 0, branch);
    }
    reconstruct_stress(sp, evec, sigma);
    compute_strain(sigma, elastic, mu, lambda);
    plastic_strain_fn(strain, elastic, plastic);
    translate_stress(sigma, aux);
    for (int i = 0; i < 3; i++) wsigma3[i] = aux[i] * weight;
    for (int i = 0; i < 6; i++) plastic_strain6[i] = plastic[i];
    if (sigma6_out)
        for (int i = 0; i < 6; i++) sigma6_out[i] = sigma[i];
}

static void compute_sigma_impl(int64_t npts, const double *strain3, double *wsigma3, double *wdep9,
                               const double *weight, double lambda, double mu, double mc_phi, double mc_psi,
                               double mc_cohesion, double *plastic_strain6, double *alpha, int *mtype, int *branch,
                               int update_mem, int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t ip = 0; ip < npts; ip++) {
        double ep[6];
        for (int i = 0; i < 6; i++) ep[i] = plastic_strain6[6 * ip + i];
        double a = alpha ? alpha[ip] : 0.;
        int mt = 0, br = 0;
        orc_point_sigma(&strain3[3 * ip], ep, &a, &mt, &br, weight[ip], lambda, mu, mc_phi, mc_psi, mc_cohesion,
                        &wsigma3[3 * ip], NULL, wdep9 ? &wdep9[9 * ip] : NULL);
        if (update_mem) {
            for (int i = 0; i < 6; i++) plastic_strain6[6 * ip + i] = ep[i];
            if (alpha) alpha[ip] = a;
            if (mtype) mtype[ip] = mt;
            if (branch) branch[ip] = br;
        }
    }
}

/* TPZConstitutiveLawProcessor.cpp:131-342 */
void orc_compute_sigma(int64_t npts, const double *strain3, double *wsigma3, const double *weight, double lambda,
                       double mu, double mc_phi, double mc_psi, double mc_cohesion, double *plastic_strain6,
                       double *alpha, int *mtype, int *branch, int update_mem, int nthreads) {
    compute_sigma_impl(npts, strain3, wsigma3, NULL, weight, lambda, mu, mc_phi, mc_psi, mc_cohesion,
                       plastic_strain6, alpha, mtype, branch, update_mem, nthreads);
}

/* TPZConstitutiveLawProcessor.cpp:344-585 */
void orc_compute_sigma_dep(int64_t npts, const double *strain3, double *wsigma3, double *wdep9, const double *weight,
                           double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion,
                           double *plastic_strain6, double *alpha, int *mtype, int *branch, int update_mem,
                           int nthreads) {
    compute_sigma_impl(npts, strain3, wsigma3, wdep9, weight, lambda, mu, mc_phi, mc_psi, mc_cohesion,
                       plastic_strain6, alpha, mtype, branch, update_mem, nthreads);
}

/* ------------------------------------------------------------------ */
/* TPZIrregularBlocksMatrix / TPZNumericalIntegrator                   */
/* ------------------------------------------------------------------ */

/* cblas_dgthr semantics (TPZNumericalIntegrator.cpp:47): y[i] = x[idx[i]] */
void orc_gather(int64_t n, const double *x, const int *idx, double *y) {
    for (int64_t i = 0; i < n; i++) y[i] = x[idx[i]];
}

/* TPZIrregularBlocksMatrix::MultiplyVector (TPZIrregularBlocksMatrix.cpp:97-140,
 * MatMul.h:9-79).  Blocks are row-major R_e x C_e at mat_pos[e].
 * opt == 0: y = B x, each row summed over columns j = 0..C_e-1 in order.
 * opt != 0: y = -B^T x; rows are grouped in threes (one integration point,
 * rows 3ip..3ip+2): per point g = B[3ip]x0 + B[3ip+1]x1 + B[3ip+2]x2, then the
 * points are summed ip = 0..n_ip-1 in order, then negated (alpha = -1). */
void orc_multiply_vector(const orc_blocks *b, const double *x, double *y, int opt, int nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t e = 0; e < b->nblocks; e++) {
        const int R = b->row_sizes[e], C = b->col_sizes[e];
        const double *B = b->storage + b->mat_pos[e];
        if (opt == 0) {
            const double *xe = x + b->col_first[e];
            double *ye = y + b->row_first[e];
            for (int r = 0; r < R; r++) {
                double acc = 0.;
                for (int j = 0; j < C; j++) acc = acc + B[(int64_t)r * C + j] * xe[j];
                ye[r] = acc;
            }
        } else {
            const double *xe = x + b->row_first[e];
            double *ye = y + b->col_first[e];
            const int nip = R / 3;
            for (int j = 0; j < C; j++) {
                double f = 0.;
                for (int ip = 0; ip < nip; ip++) {
                    const double *Bp = B + (int64_t)(3 * ip) * C;
                    double g = Bp[j] * xe[3 * ip];
                    g = g + Bp[C + j] * xe[3 * ip + 1];
                    g = g + Bp[2 * C + j] * xe[3 * ip + 2];
                    f = (ip == 0) ? g : f + g;
                }
                ye[j] = -f;
            }
        }
    }
}

/* TPZNumericalIntegrator::Multiply (TPZNumericalIntegrator.cpp:40-51) */
void orc_multiply(const orc_blocks *b, const int *dof_indexes, const double *u, double *strain3, int nthreads) {
    const int64_t cols = b->col_first[b->nblocks];
    double *g = (double *)malloc(sizeof(double) * (size_t)(cols > 0 ? cols : 1));
    orc_gather(cols, u, dof_indexes, g);
    orc_multiply_vector(b, g, strain3, 0, nthreads);
    free(g);
}

/* TPZNumericalIntegrator::MultiplyTranspose (TPZNumericalIntegrator.cpp:81-105):
 * forces = -B^T sigma; scatter by assignment into ncolor slabs of neq
 * (res[color_indexes[i]] = forces[i]); pairwise-halving color reduction.
 * slab_scratch holds ncolor*neq doubles (NULL: allocated here). */
void orc_multiply_transpose(const orc_blocks *b, const int *color_indexes, int64_t ncolor, int64_t neq,
                            const double *wsigma3, double *res, double *slab_scratch, int nthreads) {
    const int64_t cols = b->col_first[b->nblocks];
    double *forces = (double *)malloc(sizeof(double) * (size_t)(cols > 0 ? cols : 1));
    double *slab = slab_scratch ? slab_scratch : (double *)malloc(sizeof(double) * (size_t)(ncolor * neq));
    orc_multiply_vector(b, wsigma3, forces, 1, nthreads);
    memset(slab, 0, sizeof(double) * (size_t)(ncolor * neq));
    for (int64_t i = 0; i < cols; i++) slab[color_indexes[i]] = forces[i]; /* cblas_dsctr, :93 */
    int64_t nc = ncolor;
    int64_t colorassemb = nc / 2;
    while (colorassemb > 0) { /* :95-103 */
        int64_t firsteq = (nc - colorassemb) * neq;
        for (int64_t i = 0; i < colorassemb * neq; i++) slab[i] = slab[i] + slab[firsteq + i];
        nc -= colorassemb;
        colorassemb = nc / 2;
    }
    for (int64_t i = 0; i < neq; i++) res[i] = slab[i];
    free(forces);
    if (!slab_scratch) free(slab);
}

/* Greedy multi-pass element coloring, TPZNumericalIntegrator.cpp:452-493.
 * conn_ptr/conn: CSR connect list per element (BuildConnectList order).
 * Returns the number of colors. */
int64_t orc_color_elements(int64_t nblocks, const int64_t *conn_ptr, const int64_t *conn, int64_t nconnects,
                           int64_t *elemcolor) {
    char *marked = (char *)calloc((size_t)(nconnects > 0 ? nconnects : 1), 1);
    for (int64_t e = 0; e < nblocks; e++) elemcolor[e] = -1;
    int64_t contcolor = 0;
    int needstocontinue = 1;
    while (needstocontinue) {
        needstocontinue = 0;
        for (int64_t e = 0; e < nblocks; e++) {
            if (elemcolor[e] != -1) continue;
            int64_t icon;
            for (icon = conn_ptr[e]; icon < conn_ptr[e + 1]; icon++)
                if (marked[conn[icon]]) break;
            if (icon != conn_ptr[e + 1]) {
                needstocontinue = 1;
                continue;
            }
            elemcolor[e] = contcolor;
            for (icon = conn_ptr[e]; icon < conn_ptr[e + 1]; icon++) marked[conn[icon]] = 1;
        }
        contcolor++;
        memset(marked, 0, (size_t)(nconnects > 0 ? nconnects : 1));
    }
    free(marked);
    return contcolor;
}

/* fColorIndexes = fDoFIndexes + color * neq, TPZNumericalIntegrator.cpp:496-505 */
void orc_color_indexes(int64_t nblocks, const int *col_sizes, const int *col_first, const int *dof_indexes,
                       const int64_t *elemcolor, int64_t neq, int *color_indexes) {
    for (int64_t e = 0; e < nblocks; e++)
        for (int c = 0; c < col_sizes[e]; c++)
            color_indexes[col_first[e] + c] = (int)(dof_indexes[col_first[e] + c] + elemcolor[e] * neq);
}

/* TPZElastoPlasticIntPointsStructMatrix::Assemble(rhs) (.cpp:137-154) ->
 * ResidualIntegration (TPZNumericalIntegrator.cpp:107-114):
 * rhs = MultiplyTranspose(ComputeSigma(Multiply(u))) + f_lin. */
void orc_residual(const orc_blocks *b, const int *dof_indexes, const int *color_indexes, int64_t ncolor, int64_t neq,
                  const double *weight, double lambda, double mu, double mc_phi, double mc_psi, double mc_cohesion,
                  double *plastic_strain6, double *alpha, int *mtype, int *branch, int update_mem, const double *u,
                  const double *f_lin, double *rhs, double *strain_scratch, double *sigma_scratch,
                  double *slab_scratch, int nthreads) {
    const int64_t rows = b->row_first[b->nblocks];
    const int64_t npts = rows / 3;
    double *strain = strain_scratch ? strain_scratch : (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
    double *sigma = sigma_scratch ? sigma_scratch : (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
    orc_multiply(b, dof_indexes, u, strain, nthreads);
    orc_compute_sigma(npts, strain, sigma, weight, lambda, mu, mc_phi, mc_psi, mc_cohesion, plastic_strain6, alpha,
                      mtype, branch, update_mem, nthreads);
    orc_multiply_transpose(b, color_indexes, ncolor, neq, sigma, rhs, slab_scratch, nthreads);
    if (f_lin)
        for (int64_t i = 0; i < neq; i++) rhs[i] = rhs[i] + f_lin[i];
    if (!strain_scratch) free(strain);
    if (!sigma_scratch) free(sigma);
}

/* ComputeTangentMatrix (TPZNumericalIntegrator.cpp:236-270) + packing into Kc
 * at fMatrixStride (KAssembly :139-163).  Element i of the color order is
 * el_color_index[i]; Kc receives its upper triangle row by row. */
void orc_element_tangents(const orc_blocks *b, const double *wdep9, const int *el_color_index, double *kc_packed,
                          int nthreads) {
    /* offsets of the packed triangles in color order (FillLIndexes :546-565) */
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(b->nblocks + 1));
    off[0] = 0;
    for (int64_t i = 0; i < b->nblocks; i++) {
        int64_t C = b->col_sizes[el_color_index[i]];
        off[i + 1] = off[i] + (C * C + C) / 2;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < b->nblocks; i++) {
        const int64_t e = el_color_index[i];
        const int C = b->col_sizes[e];
        const int nip = b->row_sizes[e] / 3;
        const int64_t first_ip = b->row_first[e] / 3;
        const double *B = b->storage + b->mat_pos[e];
        double *Ke = (double *)calloc((size_t)C * C, sizeof(double));
        double *DB = (double *)malloc(sizeof(double) * 3 * (size_t)C);
        for (int ip = 0; ip < nip; ip++) {
            const double *Bp = B + (int64_t)ip * 3 * C;
            const double *D = wdep9 + 9 * (first_ip + ip);
            /* DeBip = dep * Bip (3 x C) */
            for (int r = 0; r < 3; r++)
                for (int j = 0; j < C; j++) {
                    double acc = 0.;
                    for (int k = 0; k < 3; k++) acc = acc + D[3 * r + k] * Bp[k * C + j];
                    DB[r * C + j] = acc;
                }
            /* K += Bip^T DeBip */
            for (int a = 0; a < C; a++)
                for (int c = 0; c < C; c++) {
                    double acc = 0.;
                    for (int k = 0; k < 3; k++) acc = acc + Bp[k * C + a] * DB[k * C + c];
                    Ke[a * C + c] = Ke[a * C + c] + acc;
                }
        }
        int64_t c = off[i];
        for (int a = 0; a < C; a++)
            for (int d = a; d < C; d++) kc_packed[c++] = Ke[a * C + d];
        free(Ke);
        free(DB);
    }
    free(off);
}
