// TPZCompMesh.cpp -- the mesh stand-in (see TPZCompMesh.h): the structured
// quarter annulus, the Gmsh 4.1 reader, Lagrange shape tables and loads.
#include "TPZCompMesh.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <array>

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


void quad_table(const TPZCompMesh& m, std::vector<double>& tab) {
    const int fP = m.POrder();
    const auto& fGx = m.GaussPoints();
    const auto& fGw = m.GaussWeights();
    const auto& fGa = m.LocalGridA();
    const auto& fGb = m.LocalGridB();
    const int n1 = fP + 1, nip = n1 * n1, nsh = n1 * n1;
    std::vector<double> L(n1 * n1), dL(n1 * n1);  // [q][k]: basis k at Gauss point q
    for (int q = 0; q < n1; q++)
        for (int k = 0; k < n1; k++) {
            L[q * n1 + k] = m.Lagrange(k, fGx[q]);
            dL[q * n1 + k] = m.LagrangeDeriv(k, fGx[q]);
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


// Symmetric triangle rules of degree 2p on the reference triangle
// (0,0), (1,0), (0,1) (area 1/2): p = 1, 3 points (degree 2); p = 2, 6
// points (degree 4); p = 3, 12 points (degree 6) -- D. A. Dunavant, "High
// degree efficient symmetrical Gaussian quadrature rules for the triangle",
// IJNME 21 (1985), rules 2, 4 and 6.  Weights sum to 1/2.
void triangle_rule(int p, std::vector<double>& s, std::vector<double>& t, std::vector<double>& w) {
    s.clear();
    t.clear();
    w.clear();
    auto orbit3 = [&](double a, double wt) {  // (a, a, 1 - 2a) and its rotations
        const double b = 1.0 - 2.0 * a;
        const double pts[3][2] = {{a, a}, {b, a}, {a, b}};
        for (auto& q : pts) {
            s.push_back(q[0]);
            t.push_back(q[1]);
            w.push_back(wt / 2);
        }
    };
    auto orbit6 = [&](double a, double b, double wt) {  // permutations of (a, b, 1 - a - b)
        const double c = 1.0 - a - b;
        const double pts[6][2] = {{a, b}, {b, a}, {b, c}, {c, b}, {c, a}, {a, c}};
        for (auto& q : pts) {
            s.push_back(q[0]);
            t.push_back(q[1]);
            w.push_back(wt / 2);
        }
    };
    if (p == 1) {
        orbit3(1.0 / 6.0, 1.0 / 3.0);
    } else if (p == 2) {
        orbit3(0.445948490915965, 0.223381589678011);
        orbit3(0.091576213509771, 0.109951743655322);
    } else {
        orbit3(0.249286745170910, 0.116786275726379);
        orbit3(0.063089014491502, 0.050844906370207);
        orbit6(0.310352451033784, 0.053145049844817, 0.082851075618374);
    }
}

// Local nodes of a P_p triangle as barycentric integer coordinates (i0, i1,
// i2), i0 + i1 + i2 = p, lambda = (1 - s - t, s, t): vertices v0, v1, v2,
// edge interiors v0->v1, v1->v2, v2->v0, interior.
std::vector<std::array<int, 3>> triangle_nodes(int p) {
    std::vector<std::array<int, 3>> n = {{p, 0, 0}, {0, p, 0}, {0, 0, p}};
    for (int k = 1; k < p; k++) n.push_back({p - k, k, 0});
    for (int k = 1; k < p; k++) n.push_back({0, p - k, k});
    for (int k = 1; k < p; k++) n.push_back({k, 0, p - k});
    for (int j = 1; j < p; j++)
        for (int i = 1; i + j < p; i++) n.push_back({p - i - j, i, j});
    return n;
}

// prod_{m < i} (p l - m) / (m + 1) and its derivative in l
void lagrange_factor(int p, int i, double l, double& f, double& df) {
    f = 1.0;
    df = 0.0;
    for (int m = 0; m < i; m++) {
        const double g = (p * l - m) / (m + 1), dg = (double)p / (m + 1);
        df = df * g + f * dg;
        f = f * g;
    }
}

// Triangle table in the coordinates of the parallelogram's reference square:
// xi = 2s - 1, eta = 2t - 1, dN/dxi = dN/ds / 2, w_q = 4 w.
void triangle_table(int p, std::vector<double>& tab) {
    std::vector<double> s, t, w;
    triangle_rule(p, s, t, w);
    const auto nodes = triangle_nodes(p);
    const int nip = (int)s.size(), nsh = (int)nodes.size();
    tab.assign((size_t)nip * (2 * nsh + 3), 0.0);
    double* dphi = tab.data();
    double* xi = dphi + (size_t)nip * 2 * nsh;
    double* eta = xi + nip;
    double* wq = eta + nip;
    for (int ip = 0; ip < nip; ip++) {
        const double l[3] = {1.0 - s[ip] - t[ip], s[ip], t[ip]};
        xi[ip] = 2.0 * s[ip] - 1.0;
        eta[ip] = 2.0 * t[ip] - 1.0;
        wq[ip] = 4.0 * w[ip];
        for (int a = 0; a < nsh; a++) {
            double f[3], df[3];
            for (int k = 0; k < 3; k++) lagrange_factor(p, nodes[a][k], l[k], f[k], df[k]);
            const double d0 = df[0] * f[1] * f[2], d1 = f[0] * df[1] * f[2], d2 = f[0] * f[1] * df[2];
            dphi[((size_t)ip * nsh + a) * 2 + 0] = (d1 - d0) / 2;
            dphi[((size_t)ip * nsh + a) * 2 + 1] = (d2 - d0) / 2;
        }
    }
}

}  // namespace

// ============================================================== construction

TPZCompMesh::TPZCompMesh(int porder) : fP(porder) {
    if (porder < 1 || porder > 3) throw std::invalid_argument("TPZCompMesh: order must be 1, 2 or 3");
    SetUpReference();
}

void TPZCompMesh::SetUpReference() {
    const int p = fP;
    fGa.clear();
    fGb.clear();
    // local node order: vertices, edge interiors (edges 0..3), interior
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
}

TPZCompMesh::TPZCompMesh(int nr, int ntheta, int porder, double r_in, double r_out, double ratio) : fP(porder) {
    if (nr < 1 || ntheta < 1) throw std::invalid_argument("TPZCompMesh: need at least one cell per direction");
    if (porder < 1 || porder > 3) throw std::invalid_argument("TPZCompMesh: order must be 1, 2 or 3");
    if (!(r_in > 0) || !(r_out > r_in) || !(ratio > 0))
        throw std::invalid_argument("TPZCompMesh: need 0 < r_in < r_out and ratio > 0");
    fNr = nr;
    fNt = ntheta;
    fRin = r_in;
    fRout = r_out;
    fRatio = ratio;
    SetUpReference();
    const int p = fP;
    const int64_t nodesR = (int64_t)p * nr + 1, nodesT = (int64_t)p * ntheta + 1;
    auto node_index = [&](int64_t I, int64_t J) { return J * nodesR + I; };
    // corner radii: h_k = h0 * ratio^k, sum = r_out - r_in
    std::vector<double> R(nr + 1), Th(ntheta + 1);
    R[0] = r_in;
    const double L = r_out - r_in;
    double h0 = (ratio == 1.0) ? L / nr : L * (ratio - 1.0) / (std::pow(ratio, nr) - 1.0);
    double h = h0;
    for (int k = 0; k < nr; k++) {
        R[k + 1] = R[k] + h;
        h *= ratio;
    }
    R[nr] = r_out;
    for (int j = 0; j <= ntheta; j++) Th[j] = (M_PI / 2) * j / ntheta;
    auto cell_corners = [&](int64_t i, int64_t j, double X[4][2]) {
        auto pt = [&](int64_t ii, int64_t jj, double* out) {
            const double r = R[ii];
            if (jj == 0) {
                out[0] = r;
                out[1] = 0.0;
            } else if (jj == ntheta) {
                out[0] = 0.0;
                out[1] = r;
            } else {
                out[0] = r * std::cos(Th[jj]);
                out[1] = r * std::sin(Th[jj]);
            }
        };
        pt(i, j, X[0]);
        pt(i + 1, j, X[1]);
        pt(i + 1, j + 1, X[2]);
        pt(i, j + 1, X[3]);
    };
    // nodes: Q1 map of the owning cell at the node's reference position
    fXY.resize(2 * nodesR * nodesT);
    for (int64_t node = 0; node < nodesR * nodesT; node++) {
        const int64_t I = node % nodesR, J = node / nodesR;
        const int64_t i = std::min<int64_t>(I / p, nr - 1), j = std::min<int64_t>(J / p, ntheta - 1);
        const double xi = -1.0 + 2.0 * (double)(I - p * i) / p;
        const double eta = -1.0 + 2.0 * (double)(J - p * j) / p;
        double X[4][2];
        cell_corners(i, j, X);
        const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4, (1 + xi) * (1 + eta) / 4,
                             (1 - xi) * (1 + eta) / 4};
        fXY[2 * node] = N[0] * X[0][0] + N[1] * X[1][0] + N[2] * X[2][0] + N[3] * X[3][0];
        fXY[2 * node + 1] = N[0] * X[0][1] + N[1] * X[1][1] + N[2] * X[2][1] + N[3] * X[3][1];
    }
    // elements, radial index fastest
    const int64_t nel = (int64_t)nr * ntheta;
    const int nconn = (int)fGa.size();
    fType.assign(nel, (unsigned char)EQuad);
    fHas[EQuad] = true;
    fConnPtr.resize(nel + 1);
    fConn.resize(nel * nconn);
    fCorner.resize(8 * nel);
    for (int64_t e = 0; e < nel; e++) {
        const int64_t i = e % nr, j = e / nr;
        fConnPtr[e + 1] = fConnPtr[e] + nconn;
        for (int a = 0; a < nconn; a++) fConn[e * nconn + a] = node_index(p * i + fGa[a], p * j + fGb[a]);
        double X[4][2];
        cell_corners(i, j, X);
        std::copy(&X[0][0], &X[0][0] + 8, &fCorner[8 * e]);
    }
    // tractions on the inner (xi = -1: edge v3 -> v0) and outer (xi = +1:
    // edge v1 -> v2) arcs; nodes along eta from theta_j to theta_j+1
    for (int side = 0; side < 2; side++) {
        const int64_t i = side == 0 ? 0 : nr - 1;
        const int64_t I = side == 0 ? 0 : nodesR - 1;
        for (int64_t j = 0; j < ntheta; j++) {
            double X[4][2];
            cell_corners(i, j, X);
            std::vector<int64_t> nodes(p + 1);
            for (int b = 0; b <= p; b++) nodes[b] = node_index(I, p * j + b);
            AddLoadEdge(side, side == 0 ? X[3] : X[1], side == 0 ? X[0] : X[2], nodes);
        }
    }
    // symmetry: u_y = 0 on theta = 0, u_x = 0 on theta = pi/2
    for (int64_t I = 0; I < nodesR; I++) {
        fFixedDofs.push_back(2 * node_index(I, 0) + 1);
        fFixedDofs.push_back(2 * node_index(I, nodesT - 1) + 0);
    }
    fSolution.Redim(NEquations(), 1);
}

void TPZCompMesh::AddLoadEdge(int side, const double* A, const double* B, const std::vector<int64_t>& nodes) {
    const double dx = B[0] - A[0], dy = B[1] - A[1];
    const double len = std::sqrt(dx * dx + dy * dy);
    fEdgeSide.push_back(side);
    fEdgeNormal.push_back(dy / len);
    fEdgeNormal.push_back(-dx / len);
    fEdgeLen.push_back(len);
    fEdgeNodes.insert(fEdgeNodes.end(), nodes.begin(), nodes.end());
}

// =================================================================== Gmsh

namespace {

struct GmshFile {
    std::map<std::pair<int, int>, std::string> phys_name;  // (dim, tag) -> name
    std::map<std::pair<int, int>, int> entity_phys;         // (dim, entity tag) -> first physical tag
    std::vector<int64_t> node_tag;
    std::vector<double> xy;
    struct Elem {
        int type, dim, entity;
        std::vector<int64_t> nodes;  // node tags
    };
    std::vector<Elem> elems;
};

[[noreturn]] void gmsh_fail(const std::string& file, const std::string& what) {
    throw std::invalid_argument("Gmsh reader: " + file + ": " + what);
}

GmshFile parse_gmsh(const std::string& file) {
    std::ifstream in(file);
    if (!in) gmsh_fail(file, "cannot open");
    GmshFile g;
    std::string tok;
    bool have_format = false, have_nodes = false, have_elems = false;
    while (in >> tok) {
        if (tok == "$MeshFormat") {
            double version;
            int binary, dsize;
            if (!(in >> version >> binary >> dsize)) gmsh_fail(file, "bad $MeshFormat");
            if (version < 4.1 - 1e-9 || version >= 5.0) gmsh_fail(file, "format version must be 4.1");
            if (binary != 0) gmsh_fail(file, "binary files are not supported");
            have_format = true;
        } else if (tok == "$PhysicalNames") {
            int n;
            in >> n;
            for (int k = 0; k < n; k++) {
                int dim, tag;
                std::string name;
                in >> dim >> tag;
                std::getline(in, name);
                const auto a = name.find('"'), b = name.rfind('"');
                if (a == std::string::npos || b <= a) gmsh_fail(file, "bad physical name");
                g.phys_name[{dim, tag}] = name.substr(a + 1, b - a - 1);
            }
        } else if (tok == "$Entities") {
            int64_t cnt[4];
            in >> cnt[0] >> cnt[1] >> cnt[2] >> cnt[3];
            for (int dim = 0; dim < 4; dim++)
                for (int64_t k = 0; k < cnt[dim]; k++) {
                    int tag;
                    double v;
                    in >> tag;
                    for (int c = 0; c < (dim == 0 ? 3 : 6); c++) in >> v;
                    int nphys;
                    in >> nphys;
                    for (int q = 0; q < nphys; q++) {
                        int ph;
                        in >> ph;
                        if (q == 0) g.entity_phys[{dim, tag}] = ph;
                    }
                    if (dim > 0) {
                        int nb, b;
                        in >> nb;
                        for (int q = 0; q < nb; q++) in >> b;
                    }
                }
            if (!in) gmsh_fail(file, "bad $Entities");
        } else if (tok == "$Nodes") {
            int64_t nblocks, nnodes, mn, mx;
            in >> nblocks >> nnodes >> mn >> mx;
            g.node_tag.reserve(nnodes);
            g.xy.reserve(2 * nnodes);
            for (int64_t b = 0; b < nblocks; b++) {
                int edim, etag, param;
                int64_t n;
                in >> edim >> etag >> param >> n;
                if (param) gmsh_fail(file, "parametric nodes are not supported");
                const size_t first = g.node_tag.size();
                for (int64_t k = 0; k < n; k++) {
                    int64_t t;
                    in >> t;
                    g.node_tag.push_back(t);
                }
                for (int64_t k = 0; k < n; k++) {
                    double x, y, z;
                    in >> x >> y >> z;
                    g.xy.push_back(x);
                    g.xy.push_back(y);
                }
                (void)first;
            }
            if (!in || (int64_t)g.node_tag.size() != nnodes) gmsh_fail(file, "bad $Nodes");
            have_nodes = true;
        } else if (tok == "$Elements") {
            int64_t nblocks, nel, mn, mx;
            in >> nblocks >> nel >> mn >> mx;
            g.elems.reserve(nel);
            for (int64_t b = 0; b < nblocks; b++) {
                int edim, etag, etype;
                int64_t n;
                in >> edim >> etag >> etype >> n;
                int nn;
                switch (etype) {
                    case 15: nn = 1; break;
                    case 1: nn = 2; break;
                    case 2: nn = 3; break;
                    case 3: nn = 4; break;
                    default: gmsh_fail(file, "element type " + std::to_string(etype) + " is not supported");
                }
                for (int64_t k = 0; k < n; k++) {
                    int64_t t;
                    in >> t;
                    GmshFile::Elem e{etype, edim, etag, std::vector<int64_t>(nn)};
                    for (int q = 0; q < nn; q++) in >> e.nodes[q];
                    g.elems.push_back(std::move(e));
                }
            }
            if (!in || (int64_t)g.elems.size() != nel) gmsh_fail(file, "bad $Elements");
            have_elems = true;
        }
    }
    if (!have_format || !have_nodes || !have_elems) gmsh_fail(file, "missing $MeshFormat, $Nodes or $Elements");
    return g;
}

}  // namespace

std::unique_ptr<TPZCompMesh> TPZCompMesh::ReadGmsh(const std::string& file, int porder) {
    const GmshFile g = parse_gmsh(file);
    std::unique_ptr<TPZCompMesh> m(new TPZCompMesh(porder));
    const int p = porder;
    const int64_t nv = (int64_t)g.node_tag.size();
    std::unordered_map<int64_t, int64_t> idx;
    idx.reserve(2 * nv);
    for (int64_t k = 0; k < nv; k++) idx[g.node_tag[k]] = k;
    auto vertex = [&](int64_t tag) {
        auto it = idx.find(tag);
        if (it == idx.end()) gmsh_fail(file, "element refers to unknown node " + std::to_string(tag));
        return it->second;
    };
    auto phys = [&](const GmshFile::Elem& e) -> std::string {
        auto it = g.entity_phys.find({e.dim, e.entity});
        if (it == g.entity_phys.end()) return "";
        auto nt = g.phys_name.find({e.dim, it->second});
        return nt == g.phys_name.end() ? "" : nt->second;
    };
    auto& cnt = m->fCounts;
    cnt.nodes = nv;
    cnt.elements = (int64_t)g.elems.size();
    // domain elements (2D, file order) and their edges in order of first use
    std::vector<const GmshFile::Elem*> dom;
    for (const auto& e : g.elems) {
        cnt.points += e.type == 15;
        cnt.lines += e.type == 1;
        cnt.triangles += e.type == 2;
        cnt.quads += e.type == 3;
        if (e.type == 2 || e.type == 3) dom.push_back(&e);
    }
    if (dom.empty()) gmsh_fail(file, "no triangles or quadrilaterals");
    std::map<std::pair<int64_t, int64_t>, int64_t> edge_id;  // (min, max) vertex -> edge
    std::vector<int64_t> edge_owner;
    for (size_t k = 0; k < dom.size(); k++) {
        const auto& e = *dom[k];
        const int nvx = (int)e.nodes.size();
        for (int c = 0; c < nvx; c++) {
            const int64_t a = vertex(e.nodes[c]), b = vertex(e.nodes[(c + 1) % nvx]);
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            if (edge_id.emplace(key, (int64_t)edge_owner.size()).second) edge_owner.push_back((int64_t)k);
        }
    }
    const int64_t ne = (int64_t)edge_owner.size();
    // node numbering: vertices, edge nodes (p - 1 per edge, from its lower
    // vertex to its higher), element interiors
    m->fXY = g.xy;
    m->fXY.resize(2 * (nv + ne * (p - 1)));
    for (const auto& kv : edge_id) {
        const int64_t a = kv.first.first, b = kv.first.second, id = kv.second;
        for (int k = 1; k < p; k++) {
            const double f = (double)k / p;
            const int64_t n = nv + id * (p - 1) + (k - 1);
            m->fXY[2 * n] = g.xy[2 * a] + f * (g.xy[2 * b] - g.xy[2 * a]);
            m->fXY[2 * n + 1] = g.xy[2 * a + 1] + f * (g.xy[2 * b + 1] - g.xy[2 * a + 1]);
        }
    }
    const int64_t nel = (int64_t)dom.size();
    m->fType.resize(nel);
    m->fConnPtr.assign(1, 0);
    m->fCorner.resize(8 * nel);
    std::vector<double> centroid(2 * nel);
    for (int64_t k = 0; k < nel; k++) {
        const auto& e = *dom[k];
        const bool quad = e.type == 3;
        const int nvx = quad ? 4 : 3;
        m->fType[k] = (unsigned char)(quad ? EQuad : ETriangle);
        m->fHas[quad ? EQuad : ETriangle] = true;
        int64_t vtx[4];
        for (int c = 0; c < nvx; c++) {
            vtx[c] = vertex(e.nodes[c]);
            m->fConn.push_back(vtx[c]);
        }
        for (int c = 0; c < nvx; c++) {  // edge c: vertex c -> vertex c + 1
            const int64_t a = vtx[c], b = vtx[(c + 1) % nvx];
            const int64_t id = edge_id.at({std::min(a, b), std::max(a, b)});
            for (int k2 = 1; k2 < p; k2++) {
                const int kk = a < b ? k2 : p - k2;
                m->fConn.push_back(nv + id * (p - 1) + (kk - 1));
            }
        }
        double* X = &m->fCorner[8 * k];
        for (int c = 0; c < nvx; c++) {
            X[2 * c] = g.xy[2 * vtx[c]];
            X[2 * c + 1] = g.xy[2 * vtx[c] + 1];
        }
        if (!quad) {  // (v0, v1, v1 + v2 - v0, v2)
            const double v2x = X[4], v2y = X[5];
            X[4] = X[2] + v2x - X[0];
            X[5] = X[3] + v2y - X[1];
            X[6] = v2x;
            X[7] = v2y;
        }
        // interior nodes
        auto add_node = [&](double x, double y) {
            m->fConn.push_back((int64_t)m->fXY.size() / 2);
            m->fXY.push_back(x);
            m->fXY.push_back(y);
        };
        if (quad) {
            for (int b = 1; b < p; b++)
                for (int a = 1; a < p; a++) {
                    const double xi = -1.0 + 2.0 * a / p, eta = -1.0 + 2.0 * b / p;
                    const double N[4] = {(1 - xi) * (1 - eta) / 4, (1 + xi) * (1 - eta) / 4,
                                         (1 + xi) * (1 + eta) / 4, (1 - xi) * (1 + eta) / 4};
                    add_node(N[0] * X[0] + N[1] * X[2] + N[2] * X[4] + N[3] * X[6],
                             N[0] * X[1] + N[1] * X[3] + N[2] * X[5] + N[3] * X[7]);
                }
        } else {
            for (int j = 1; j < p; j++)
                for (int i = 1; i + j < p; i++) {
                    const double s = (double)i / p, t = (double)j / p;
                    add_node(X[0] + s * (X[2] - X[0]) + t * (X[6] - X[0]), X[1] + s * (X[3] - X[1]) + t * (X[7] - X[1]));
                }
        }
        m->fConnPtr.push_back((int64_t)m->fConn.size());
        double cx = 0, cy = 0;
        for (int c = 0; c < nvx; c++) {
            cx += g.xy[2 * vtx[c]];
            cy += g.xy[2 * vtx[c] + 1];
        }
        centroid[2 * k] = cx / nvx;
        centroid[2 * k + 1] = cy / nvx;
    }
    // boundary lines -> traction edges (outward normal: away from the owning element)
    for (const auto& e : g.elems) {
        if (e.type == 15) {
            const std::string name = phys(e);
            if (name == "fixed_x") m->fFixedDofs.push_back(2 * vertex(e.nodes[0]) + 0);
            if (name == "fixed_y") m->fFixedDofs.push_back(2 * vertex(e.nodes[0]) + 1);
            continue;
        }
        if (e.type != 1) continue;
        const std::string name = phys(e);
        const int side = name == "bc_wellbore" ? 0 : name == "bc_farfield" ? 1 : -1;
        if (side < 0) continue;
        const int64_t a = vertex(e.nodes[0]), b = vertex(e.nodes[1]);
        auto it = edge_id.find({std::min(a, b), std::max(a, b)});
        if (it == edge_id.end()) gmsh_fail(file, "boundary line is not an edge of a 2D element");
        std::vector<int64_t> nodes{a};
        for (int k = 1; k < p; k++) nodes.push_back(nv + it->second * (p - 1) + ((a < b ? k : p - k) - 1));
        nodes.push_back(b);
        double A[2] = {g.xy[2 * a], g.xy[2 * a + 1]}, B[2] = {g.xy[2 * b], g.xy[2 * b + 1]};
        const int64_t own = edge_owner[it->second];
        const double nx = B[1] - A[1], ny = -(B[0] - A[0]);
        const double tx = centroid[2 * own] - 0.5 * (A[0] + B[0]), ty = centroid[2 * own + 1] - 0.5 * (A[1] + B[1]);
        if (nx * tx + ny * ty > 0) {  // (dy, -dx) points into the element: walk B -> A instead
            std::reverse(nodes.begin(), nodes.end());
            std::swap(A[0], B[0]);
            std::swap(A[1], B[1]);
        }
        m->AddLoadEdge(side, A, B, nodes);
    }
    std::sort(m->fFixedDofs.begin(), m->fFixedDofs.end());
    m->fFixedDofs.erase(std::unique(m->fFixedDofs.begin(), m->fFixedDofs.end()), m->fFixedDofs.end());
    m->fSolution.Redim(m->NEquations(), 1);
    return m;
}

// ============================================================ reference data

void TPZCompMesh::ShapeTable(ElType t, std::vector<double>& tab) const {
    if (t == EQuad)
        quad_table(*this, tab);
    else
        triangle_table(fP, tab);
}

void TPZCompMesh::ShapeTable(std::vector<double>& tab) const {
    if (Mixed()) throw std::logic_error("TPZCompMesh::ShapeTable: the mesh mixes triangles and quadrilaterals");
    ShapeTable(fHas[EQuad] ? EQuad : ETriangle, tab);
}

void TPZCompMesh::PointCoords(int64_t e, int ip, double& x, double& y) const {
    double xi, eta;
    if (ElementType(e) == EQuad) {
        const int n1 = fP + 1;
        xi = fGx[ip / n1];
        eta = fGx[ip % n1];
    } else {
        std::vector<double> s, t, w;
        triangle_rule(fP, s, t, w);
        xi = 2.0 * s[ip] - 1.0;
        eta = 2.0 * t[ip] - 1.0;
    }
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
    for (size_t k = 0; k < fEdgeSide.size(); k++) {
        const double sn = fEdgeSide[k] == 0 ? fSnInner : fSnOuter;
        if (sn == 0.0) continue;
        const double* n = &fEdgeNormal[2 * k];
        const double len = fEdgeLen[k];
        for (int b = 0; b <= p; b++) {
            double s = 0.0;
            for (int q = 0; q <= p; q++) s += fGw[q] * Lagrange(b, fGx[q]);
            s *= len / 2.0;
            const int64_t node = fEdgeNodes[k * (p + 1) + b];
            f(2 * node + 0, 0) += s * sn * n[0];
            f(2 * node + 1, 0) += s * sn * n[1];
        }
    }
}

void TPZCompMesh::DirichletMask(std::vector<signed char>& mask) const {
    mask.assign(NEquations(), 0);
    for (int64_t d : fFixedDofs) mask[d] = 1;
}
