// TPZCompMesh.cpp -- structured quarter-annulus mesh with Q_p Lagrange
// elements (see TPZCompMesh.h).
#include "TPZCompMesh.h"

#include <cmath>
#include <stdexcept>

namespace {

// Gauss-Legendre nodes (ascending) and weights by Newton iteration on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; i++) {
        long double z = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1, p1 = z;
            dp = n * (z * p1 - p0) / (z * z - 1);
            long double dz = p1 / dp;
            z -= dz;
            if (std::fabs((double)dz) < 1e-19) break;
        }
        long double p0 = 1, p1 = z;
        for (int k = 2; k <= n; k++) {
            long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[n - 1 - i] = (double)z;  // z decreases with i
        w[n - 1 - i] = (double)(2 / ((1 - z * z) * dp * dp));
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

}  // namespace

TPZCompMesh::TPZCompMesh(int nr, int ntheta, int porder, double r_in, double r_out, double ratio)
    : fNr(nr), fNt(ntheta), fP(porder), fRin(r_in), fRout(r_out), fRatio(ratio) {
    if (nr < 1 || ntheta < 1) throw std::invalid_argument("TPZCompMesh: need at least one cell per direction");
    if (porder < 1 || porder > 3) throw std::invalid_argument("TPZCompMesh: order must be 1, 2 or 3");
    if (!(r_in > 0) || !(r_out > r_in) || !(ratio > 0))
        throw std::invalid_argument("TPZCompMesh: need 0 < r_in < r_out and ratio > 0");
    fNodesR = (int64_t)fP * nr + 1;
    fNodesT = (int64_t)fP * ntheta + 1;
    // corner radii: h_k = h0 * ratio^k, sum = r_out - r_in
    fR.resize(nr + 1);
    fR[0] = r_in;
    const double L = r_out - r_in;
    double h0 = (ratio == 1.0) ? L / nr : L * (ratio - 1.0) / (std::pow(ratio, nr) - 1.0);
    double h = h0;
    for (int k = 0; k < nr; k++) {
        fR[k + 1] = fR[k] + h;
        h *= ratio;
    }
    fR[nr] = r_out;
    fTheta.resize(ntheta + 1);
    for (int j = 0; j <= ntheta; j++) fTheta[j] = (M_PI / 2) * j / ntheta;
    // local node order: vertices, edge interiors (edges 0..3), interior
    const int p = fP;
    auto add = [&](int a, int b) {
        fGa.push_back(a);
        fGb.push_back(b);
    };
    add(0, 0);
    add(p, 0);
    add(p, p);
    add(0, p);
    for (int a = 1; a < p; a++) add(a, 0);
    for (int b = 1; b < p; b++) add(p, b);
    for (int a = p - 1; a >= 1; a--) add(a, p);
    for (int b = p - 1; b >= 1; b--) add(0, b);
    for (int b = 1; b < p; b++)
        for (int a = 1; a < p; a++) add(a, b);
    gauss_legendre(p + 1, fGx, fGw);
    fSolution.Redim(NEquations(), 1);
}

void TPZCompMesh::ElementConnects(int64_t e, std::vector<int64_t>& nodes) const {
    const int64_t i = e % fNr, j = e / fNr;
    const int n = (int)fGa.size();
    nodes.resize(n);
    for (int a = 0; a < n; a++) nodes[a] = NodeIndex(fP * i + fGa[a], fP * j + fGb[a]);
}

void TPZCompMesh::CornerCoords(int64_t e, double X[4][2]) const {
    const int64_t i = e % fNr, j = e / fNr;
    auto pt = [&](int64_t ii, int64_t jj, double* out) {
        const double r = fR[ii];
        if (jj == 0) {
            out[0] = r;
            out[1] = 0.0;
        } else if (jj == fNt) {
            out[0] = 0.0;
            out[1] = r;
        } else {
            out[0] = r * std::cos(fTheta[jj]);
            out[1] = r * std::sin(fTheta[jj]);
        }
    };
    pt(i, j, X[0]);
    pt(i + 1, j, X[1]);
    pt(i + 1, j + 1, X[2]);
    pt(i, j + 1, X[3]);
}

void TPZCompMesh::NodeCoords(int64_t node, double& x, double& y) const {
    const int64_t I = node % fNodesR, J = node / fNodesR;
    const int64_t i = std::min<int64_t>(I / fP, fNr - 1), j = std::min<int64_t>(J / fP, fNt - 1);
    const double xi = -1.0 + 2.0 * (double)(I - fP * i) / fP;
    const double eta = -1.0 + 2.0 * (double)(J - fP * j) / fP;
    double X[4][2];
    CornerCoords(j * fNr + i, X);
    const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4, (1 + xi) * (1 + eta) / 4,
                         (1 - xi) * (1 + eta) / 4};
    x = N[0] * X[0][0] + N[1] * X[1][0] + N[2] * X[2][0] + N[3] * X[3][0];
    y = N[0] * X[0][1] + N[1] * X[1][1] + N[2] * X[2][1] + N[3] * X[3][1];
}

void TPZCompMesh::PointCoords(int64_t e, int ip, double& x, double& y) const {
    const int n1 = fP + 1;
    const double xi = fGx[ip / n1], eta = fGx[ip % n1];
    double X[4][2];
    CornerCoords(e, X);
    const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4, (1 + xi) * (1 + eta) / 4,
                         (1 - xi) * (1 + eta) / 4};
    x = N[0] * X[0][0] + N[1] * X[1][0] + N[2] * X[2][0] + N[3] * X[3][0];
    y = N[0] * X[0][1] + N[1] * X[1][1] + N[2] * X[2][1] + N[3] * X[3][1];
}

double TPZCompMesh::Lagrange(int k, double x) const {
    double v = 1.0;
    const double xk = -1.0 + 2.0 * k / fP;
    for (int m = 0; m <= fP; m++) {
        if (m == k) continue;
        const double xm = -1.0 + 2.0 * m / fP;
        v *= (x - xm) / (xk - xm);
    }
    return v;
}

double TPZCompMesh::LagrangeDeriv(int k, double x) const {
    const double xk = -1.0 + 2.0 * k / fP;
    double s = 0.0;
    for (int l = 0; l <= fP; l++) {
        if (l == k) continue;
        const double xl = -1.0 + 2.0 * l / fP;
        double t = 1.0 / (xk - xl);
        for (int m = 0; m <= fP; m++) {
            if (m == k || m == l) continue;
            const double xm = -1.0 + 2.0 * m / fP;
            t *= (x - xm) / (xk - xm);
        }
        s += t;
    }
    return s;
}

void TPZCompMesh::ShapeTable(std::vector<double>& tab) const {
    const int n1 = fP + 1, nip = NIntPoints(), nsh = (int)NShapeF();
    std::vector<double> L(n1 * n1), dL(n1 * n1);  // [q][k]: basis k at Gauss point q
    for (int q = 0; q < n1; q++)
        for (int k = 0; k < n1; k++) {
            L[q * n1 + k] = Lagrange(k, fGx[q]);
            dL[q * n1 + k] = LagrangeDeriv(k, fGx[q]);
        }
    tab.assign((size_t)nip * (2 * nsh + 3), 0.0);
    double* dphi = tab.data();
    double* xi = dphi + (size_t)nip * 2 * nsh;
    double* eta = xi + nip;
    double* wq = eta + nip;
    for (int qi = 0; qi < n1; qi++)
        for (int qj = 0; qj < n1; qj++) {
            const int ip = qi * n1 + qj;  // eta fastest
            xi[ip] = fGx[qi];
            eta[ip] = fGx[qj];
            wq[ip] = fGw[qi] * fGw[qj];
            for (int a = 0; a < nsh; a++) {
                dphi[((size_t)ip * nsh + a) * 2 + 0] = dL[qi * n1 + fGa[a]] * L[qj * n1 + fGb[a]];
                dphi[((size_t)ip * nsh + a) * 2 + 1] = L[qi * n1 + fGa[a]] * dL[qj * n1 + fGb[a]];
            }
        }
}

void TPZCompMesh::SetMaterial(double E, double nu, double c, double phi, double psi) {
    if (!(E > 0) || !(nu > -1.0) || !(nu < 0.5)) throw std::invalid_argument("TPZCompMesh: bad E / nu");
    fMu = E / (2. * (1. + nu));
    fLambda = nu * E / ((1. + nu) * (1. - 2. * nu));
    fC = c;
    fPhi = phi;
    fPsi = psi;
    fHasMaterial = true;
}

void TPZCompMesh::AssembleBoundaryLoads(TPZFMatrix<STATE>& f) const {
    f.Redim(NEquations(), 1);
    const int p = fP;
    for (int side = 0; side < 2; side++) {
        const double sn = side == 0 ? fSnInner : fSnOuter;
        if (sn == 0.0) continue;
        const int64_t i = side == 0 ? 0 : fNr - 1;
        for (int64_t j = 0; j < fNt; j++) {
            double X[4][2];
            CornerCoords(j * fNr + i, X);
            // inner: edge v3 -> v0 (xi = -1); outer: edge v1 -> v2 (xi = +1)
            const double* A = side == 0 ? X[3] : X[1];
            const double* B = side == 0 ? X[0] : X[2];
            const double dx = B[0] - A[0], dy = B[1] - A[1];
            const double len = std::sqrt(dx * dx + dy * dy);
            const double n[2] = {dy / len, -dx / len};
            const int64_t I = side == 0 ? 0 : fNodesR - 1;
            for (int b = 0; b <= p; b++) {
                // along the edge eta runs from -1 (v0 / v1) to +1 (v3 / v2)
                double s = 0.0;
                for (int q = 0; q <= p; q++) s += fGw[q] * Lagrange(b, fGx[q]);
                s *= len / 2.0;
                const int64_t node = NodeIndex(I, p * j + b);
                f(2 * node + 0, 0) += s * sn * n[0];
                f(2 * node + 1, 0) += s * sn * n[1];
            }
        }
    }
}

void TPZCompMesh::DirichletMask(std::vector<signed char>& mask) const {
    mask.assign(NEquations(), 0);
    for (int64_t I = 0; I < fNodesR; I++) {
        mask[2 * NodeIndex(I, 0) + 1] = 1;             // u_y = 0 on theta = 0
        mask[2 * NodeIndex(I, fNodesT - 1) + 0] = 1;   // u_x = 0 on theta = pi/2
    }
}
