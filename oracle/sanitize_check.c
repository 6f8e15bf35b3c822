/*
 * sanitize_check.c -- exercises every entry point of the CPU oracle on small
 * inputs, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_oracle_sanitizers.py compiles and runs it).  TEST
 * INFRASTRUCTURE ONLY, like the oracle itself.  Prints "ok" and exits 0 when
 * the stresses, residual and block products came out finite; a sanitizer
 * report (out-of-bounds access, leak, undefined behaviour) ends the run with
 * an error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "ipf_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double urand(void) {  /* xorshift64*, uniform in [-1, 1) */
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 0x2545f4914f6cdd1dull) >> 11) / 4503599627370496.0 - 1.0;
}

static int bad = 0;

static void check_finite(const char *what, const double *x, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (!isfinite(x[i])) {
            printf("%s[%lld] not finite\n", what, (long long)i);
            bad = 1;
            return;
        }
}

int main(void) {
    const double lambda = 1388.9, mu = 2083.3, phi = 0.5236, psi = 0.5236, c = 10.0;
    /* 1. spectral decomposition: random, repeated, zero and plane-strain tensors */
    for (int k = 0; k < 2000; k++) {
        double s[6], ev[3], vec[9];
        for (int i = 0; i < 6; i++) s[i] = urand() * pow(10.0, (k % 7) - 3);
        if (k % 5 == 1) s[ORC_XZ] = s[ORC_YZ] = 0.0;
        if (k % 5 == 2) { s[ORC_XX] = s[ORC_YY] = s[ORC_ZZ]; s[ORC_XY] = s[ORC_XZ] = s[ORC_YZ] = 0.0; }
        if (k % 5 == 3) for (int i = 0; i < 6; i++) s[i] = 0.0;
        if (k % 5 == 4) { s[ORC_YY] = s[ORC_XX]; s[ORC_XY] = s[ORC_XZ] = s[ORC_YZ] = 0.0; }
        orc_spectral_decomposition(s, ev, vec);
        check_finite("eigenvalues", ev, 3);
    }
    /* 2. one point with its tangent, across the branches (strain scales) */
    for (int k = 0; k < 2000; k++) {
        const double sc = 1e-4 * pow(10.0, k % 4);
        double e3[3] = {urand() * sc, urand() * sc, urand() * sc};
        double ep[6] = {0, 0, 0, 0, 0, 0}, alpha = 0.0, ws[3], sig[6], dep[9];
        int mt, br;
        orc_point_sigma(e3, ep, &alpha, &mt, &br, 0.5, lambda, mu, phi, psi, c, ws, sig, dep);
        check_finite("wsigma", ws, 3);  /* the tangent may divide by equal eigenvalues: not checked */
    }
    /* 3. a 2 x 2 patch of 4-node elements (9 nodes, 18 DoFs, 4 points each):
     * coloring, color indexes, residual with commit, element tangents,
     * MultiplyVector both ways; B values random */
    enum { NE = 4, NIP = 4, NC = 8, NEQ = 18 };
    int row_sizes[NE], col_sizes[NE], row_first[NE + 1], col_first[NE + 1], dof[NE * NC];
    int64_t mat_pos[NE + 1], conn_ptr[NE + 1], conn[NE * 4], color[NE];
    double storage[NE * 3 * NIP * NC];
    row_first[0] = col_first[0] = 0;
    mat_pos[0] = conn_ptr[0] = 0;
    for (int e = 0; e < NE; e++) {
        const int i = e % 2, j = e / 2;
        const int nodes[4] = {3 * j + i, 3 * j + i + 1, 3 * (j + 1) + i + 1, 3 * (j + 1) + i};
        row_sizes[e] = 3 * NIP;
        col_sizes[e] = NC;
        row_first[e + 1] = row_first[e] + 3 * NIP;
        col_first[e + 1] = col_first[e] + NC;
        mat_pos[e + 1] = mat_pos[e] + 3 * NIP * NC;
        for (int a = 0; a < 4; a++) {
            dof[e * NC + 2 * a] = 2 * nodes[a];
            dof[e * NC + 2 * a + 1] = 2 * nodes[a] + 1;
            conn[conn_ptr[e] + a] = nodes[a];
        }
        conn_ptr[e + 1] = conn_ptr[e] + 4;
    }
    for (int k = 0; k < NE * 3 * NIP * NC; k++) storage[k] = urand() * 10.0;
    const int64_t ncolor = orc_color_elements(NE, conn_ptr, conn, 9, color);
    int cidx[NE * NC];
    orc_color_indexes(NE, col_sizes, col_first, dof, color, NEQ, cidx);
    orc_blocks b = {NE, row_sizes, col_sizes, mat_pos, row_first, col_first, storage};
    const int npts = NE * NIP;
    double weight[NE * NIP], u[NEQ], flin[NEQ], rhs[NEQ], ep[6 * NE * NIP] = {0}, al[NE * NIP] = {0};
    double strain[3 * NE * NIP], sigma[3 * NE * NIP];
    double *slab = (double *)calloc((size_t)ncolor * NEQ, sizeof(double));  /* one residual slab per color */
    int mt[NE * NIP], brs[NE * NIP];
    for (int k = 0; k < npts; k++) weight[k] = 0.25;
    for (int k = 0; k < NEQ; k++) { u[k] = urand() * 1e-4; flin[k] = urand(); }
    orc_residual(&b, dof, cidx, ncolor, NEQ, weight, lambda, mu, phi, psi, c, ep, al, mt, brs, 1, u, flin, rhs,
                 strain, sigma, slab, 2);
    check_finite("residual", rhs, NEQ);
    double wdep[9 * NE * NIP], ws[3 * NE * NIP];
    orc_compute_sigma_dep(npts, strain, ws, wdep, weight, lambda, mu, phi, psi, c, ep, al, mt, brs, 0, 2);
    int elc[NE + 1];
    for (int e = 0; e <= NE; e++) elc[e] = e < NE ? e : 0;
    double kc[NE * NC * (NC + 1) / 2];
    orc_element_tangents(&b, wdep, elc, kc, 2);
    double x[NE * NC], y[3 * NE * NIP];
    for (int k = 0; k < NE * NC; k++) x[k] = urand();
    orc_multiply_vector(&b, x, y, 0, 2);
    check_finite("B x", y, 3 * NE * NIP);
    orc_multiply_vector(&b, y, x, 1, 2);
    check_finite("B^T y", x, NE * NC);
    free(slab);
    if (bad) return 1;
    printf("ok: %lld colors, %d points\n", (long long)ncolor, npts);
    return 0;
}
