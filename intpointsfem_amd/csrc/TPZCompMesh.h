// TPZCompMesh.h -- the computational-mesh stand-in the integrator is built
// from.  The reference takes shape functions, integration rules, Jacobians,
// DoF positions and boundary loads from NeoPZ (TPZNumericalIntegrator.cpp:
// 307-386, 420-443; TPZElastoPlasticIntPointsStructMatrix.cpp:156-167), which
// is not available; this class supplies the same information for
// straight-sided Lagrange elements of order p = 1..3:
//   - geometry: quadrilaterals, the bilinear map of their four corners;
//     triangles, the affine map of their three corners, held as the
//     parallelogram (v0, v1, v1 + v2 - v0, v2) whose bilinear map it is;
//   - field: tensor-product Lagrange polynomials on equispaced nodes
//     (quadrilaterals), P_p Lagrange polynomials on equispaced barycentric
//     nodes (triangles);
//   - integration: (p+1) x (p+1) Gauss-Legendre points, eta fastest
//     (quadrilaterals; NeoPZ's order-2p rule), the symmetric 3-, 6- and
//     12-point rules of degree 2p (triangles; Dunavant 1985);
//   - connects: one per node, ordered vertices, edge interiors (edge k from
//     vertex k to vertex k+1, its nodes in that direction), element interior
//     (NeoPZ's connect order), two DoFs (x, y) per node, DoF = 2*node + dir;
//   - loads: normal traction sigma_n * n (n the domain's outward normal) on
//     the inner and outer boundary, integrated with the element's 1D Gauss
//     rule; displacement constraints by DoF.
// Two producers fill it: the structured quarter annulus (symmetry
// constraints u_y = 0 on theta = 0, u_x = 0 on theta = pi/2) and a Gmsh 4.1
// reader for the reference's meshes (main.cpp:447-464: physical groups
// "Omega" the domain, "bc_wellbore" / "bc_farfield" the inner / outer
// tractions, "fixed_x" / "fixed_y" points with u_x / u_y = 0, main.cpp:
// 394-411).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "tpz_types.h"

class TPZCompMesh {
public:
    enum ElType { EQuad = 0, ETriangle = 1 };

    // Quarter annulus r in [r_in, r_out], theta in [0, pi/2]; nr radial cells
    // with geometric growth `ratio` (cell k has size h0 * ratio^k), ntheta
    // uniform angular cells, quadrilaterals of order porder in 1..3.  Element
    // e = j * nr + i (radial index fastest); node (I, J) of the (p nr + 1) x
    // (p ntheta + 1) grid is J * (p nr + 1) + I.
    TPZCompMesh(int nr, int ntheta, int porder, double r_in, double r_out, double ratio);

    // A Gmsh 4.1 ASCII mesh (element types: 15 point, 1 line, 2 triangle,
    // 3 quadrilateral) with elements of order porder.  Domain elements are
    // the 2D elements in file order; vertices are the file's nodes in file
    // order, then edge nodes in order of first use, then element interiors.
    // Throws std::invalid_argument on a malformed or unsupported file.
    static std::unique_ptr<TPZCompMesh> ReadGmsh(const std::string& file, int porder);

    // What a Gmsh file held (the counts TPZGmshReader prints).
    struct GmshCounts {
        int64_t nodes = 0, elements = 0, points = 0, lines = 0, triangles = 0, quads = 0;
    };
    const GmshCounts& Counts() const { return fCounts; }

    int Dimension() const { return 2; }
    int POrder() const { return fP; }
    bool IsStructured() const { return fNr > 0; }
    int NRadial() const { return fNr; }  // 0 for a Gmsh mesh
    int NTheta() const { return fNt; }
    int64_t NElements() const { return (int64_t)fType.size(); }
    int64_t NNodes() const { return (int64_t)fXY.size() / 2; }
    int64_t NConnects() const { return NNodes(); }
    int64_t NEquations() const { return 2 * NNodes(); }

    ElType ElementType(int64_t e) const { return (ElType)fType[e]; }
    bool HasType(ElType t) const { return fHas[t]; }
    int NShapeF(ElType t) const { return t == EQuad ? (fP + 1) * (fP + 1) : (fP + 1) * (fP + 2) / 2; }
    int NIntPoints(ElType t) const { return t == EQuad ? (fP + 1) * (fP + 1) : (fP == 1 ? 3 : fP == 2 ? 6 : 12); }
    // The counts of a mesh of one element type; -1 when types are mixed.
    int64_t NShapeF() const { return Mixed() ? -1 : NShapeF(fHas[EQuad] ? EQuad : ETriangle); }
    int NIntPoints() const { return Mixed() ? -1 : NIntPoints(fHas[EQuad] ? EQuad : ETriangle); }
    bool Mixed() const { return fHas[EQuad] && fHas[ETriangle]; }

    // Connect (node) list of element e in NeoPZ connect order.
    void ElementConnects(int64_t e, std::vector<int64_t>& nodes) const {
        nodes.assign(fConn.begin() + fConnPtr[e], fConn.begin() + fConnPtr[e + 1]);
    }
    // Corners X[c][0..1] of the element's bilinear map, in the element's
    // vertex order; a triangle's third corner is v1 + v2 - v0 and its fourth v2.
    void CornerCoords(int64_t e, double X[4][2]) const {
        for (int k = 0; k < 8; k++) (&X[0][0])[k] = fCorner[8 * e + k];
    }
    void NodeCoords(int64_t node, double& x, double& y) const {
        x = fXY[2 * node];
        y = fXY[2 * node + 1];
    }
    // Coordinates of integration point ip of element e (its type's point order).
    void PointCoords(int64_t e, int ip, double& x, double& y) const;

    // Shape table of an element type (layout of ipf_geometry.h, size
    // NIntPoints(t) * (2 NShapeF(t) + 3)): (dN_a/dxi, dN_a/deta) at each
    // point in the coordinates (xi, eta) of the bilinear map's reference
    // square, the points' (xi, eta) and weights w_q with
    // integral = sum w_q f |det J| (triangles: xi = 2s - 1, eta = 2t - 1 for
    // the triangle's (s, t), dN/dxi = dN/ds / 2, w_q = 4 w_triangle).
    void ShapeTable(ElType t, std::vector<double>& tab) const;
    // The table of a mesh of one element type (throws when types are mixed).
    void ShapeTable(std::vector<double>& tab) const;

    // Quadrilateral reference data: local node a has grid position
    // (ga[a], gb[a]) in 0..p along (xi, eta).
    const std::vector<int>& LocalGridA() const { return fGa; }
    const std::vector<int>& LocalGridB() const { return fGb; }
    // 1D Gauss-Legendre rule with p+1 points, ascending.
    const std::vector<double>& GaussPoints() const { return fGx; }
    const std::vector<double>& GaussWeights() const { return fGw; }
    // 1D Lagrange basis on equispaced nodes xi_k = -1 + 2k/p.
    double Lagrange(int k, double x) const;
    double LagrangeDeriv(int k, double x) const;

    // Material (TElastoPlasticData): Young's modulus, Poisson ratio, cohesion,
    // friction and dilation angles (radians).  lambda and mu follow the
    // standard isotropic formulas (ElasticElementaryContribution.nb:1561-1575).
    void SetMaterial(double E, double nu, double c, double phi, double psi);
    double Lambda() const { return fLambda; }
    double Mu() const { return fMu; }
    double Phi() const { return fPhi; }
    double Psi() const { return fPsi; }
    double Cohesion() const { return fC; }
    bool HasMaterial() const { return fHasMaterial; }

    // Normal traction sigma_n on the inner and outer boundary: t = sigma_n * n
    // with n the outward unit normal of the domain.
    void SetNormalStress(double inner, double outer) {
        fSnInner = inner;
        fSnOuter = outer;
    }
    // External load vector of the tractions above ([neq]).
    void AssembleBoundaryLoads(TPZFMatrix<STATE>& f) const;
    // mask[dof] = 1 for a constrained DoF.
    void DirichletMask(std::vector<signed char>& mask) const;

    TPZFMatrix<STATE>& Solution() { return fSolution; }
    bool GetUpdateMem() const { return fUpdateMem; }
    void SetUpdateMem(bool v) { fUpdateMem = v; }

    double RIn() const { return fRin; }
    double ROut() const { return fRout; }
    double Ratio() const { return fRatio; }

private:
    explicit TPZCompMesh(int porder);
    void SetUpReference();  // quadrilateral grid, 1D Gauss rule
    void AddLoadEdge(int side, const double* A, const double* B, const std::vector<int64_t>& nodes);

    int fP = 1;
    int fNr = 0, fNt = 0;
    double fRin = 0, fRout = 0, fRatio = 0;
    std::vector<int> fGa, fGb;
    std::vector<double> fGx, fGw;

    std::vector<double> fXY;           // node coordinates
    std::vector<unsigned char> fType;  // element types
    bool fHas[2] = {false, false};
    std::vector<int64_t> fConnPtr{0}, fConn;
    std::vector<double> fCorner;       // 8 per element
    // load edges: side (0 inner, 1 outer), unit normal, length, p+1 nodes from A to B
    std::vector<int> fEdgeSide;
    std::vector<double> fEdgeNormal, fEdgeLen;
    std::vector<int64_t> fEdgeNodes;
    std::vector<int64_t> fFixedDofs;
    GmshCounts fCounts;

    double fLambda = 0, fMu = 0, fPhi = 0, fPsi = 0, fC = 0;
    bool fHasMaterial = false;
    double fSnInner = 0, fSnOuter = 0;
    TPZFMatrix<STATE> fSolution;
    bool fUpdateMem = false;
};
