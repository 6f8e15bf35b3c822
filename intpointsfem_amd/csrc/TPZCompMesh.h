// TPZCompMesh.h -- the computational-mesh stand-in the integrator is built
// from.  The reference takes shape functions, integration rules, Jacobians,
// DoF positions and boundary loads from NeoPZ (TPZNumericalIntegrator.cpp:
// 307-386, 420-443; TPZElastoPlasticIntPointsStructMatrix.cpp:156-167), which
// is not available; this class supplies the same information for a
// structured quarter annulus meshed with straight-sided Q_p Lagrange
// quadrilaterals:
//   - geometry: bilinear (Q1) map of each cell's four corners;
//   - field: tensor-product Lagrange polynomials on equispaced nodes;
//   - integration: (p+1) x (p+1) Gauss-Legendre points (NeoPZ's order-2p rule),
//     eta varying fastest, 1D points ascending;
//   - connects: one per node, ordered vertices, edge interiors, cell interior
//     (NeoPZ's connect order), two DoFs (x, y) per node, DoF = 2*node + dir;
//   - loads: normal traction on the inner (r = r_in) and outer (r = r_out)
//     arcs, integrated with the element's 1D Gauss rule.
#pragma once

#include <cstdint>
#include <vector>

#include "tpz_types.h"

class TPZCompMesh {
public:
    // Quarter annulus r in [r_in, r_out], theta in [0, pi/2]; nr radial cells
    // with geometric growth `ratio` (cell k has size h0 * ratio^k), ntheta
    // uniform angular cells, order porder in 1..3.
    TPZCompMesh(int nr, int ntheta, int porder, double r_in, double r_out, double ratio);

    int Dimension() const { return 2; }
    int POrder() const { return fP; }
    int NRadial() const { return fNr; }
    int NTheta() const { return fNt; }
    int64_t NElements() const { return (int64_t)fNr * fNt; }
    int64_t NNodes() const { return fNodesR * fNodesT; }
    int64_t NConnects() const { return NNodes(); }
    int64_t NEquations() const { return 2 * NNodes(); }
    int64_t NShapeF() const { return (int64_t)(fP + 1) * (fP + 1); }
    int NIntPoints() const { return (fP + 1) * (fP + 1); }

    // Element e = j * nr + i (radial index fastest).  Connect (node) list in
    // NeoPZ connect order.
    void ElementConnects(int64_t e, std::vector<int64_t>& nodes) const;
    // Corner coordinates X[c][0..1], corners counter-clockwise starting at
    // (r_i, theta_j): (i,j), (i+1,j), (i+1,j+1), (i,j+1).
    void CornerCoords(int64_t e, double X[4][2]) const;
    // Node coordinates (Q1 map of the owning cell at the node's reference position).
    void NodeCoords(int64_t node, double& x, double& y) const;
    // Coordinates of integration point ip (eta fastest) of element e.
    void PointCoords(int64_t e, int ip, double& x, double& y) const;

    // Reference-element data for the element order: local node a has grid
    // position (ga[a], gb[a]) in 0..p along (xi, eta).
    const std::vector<int>& LocalGridA() const { return fGa; }
    const std::vector<int>& LocalGridB() const { return fGb; }
    // 1D Gauss-Legendre rule with p+1 points, ascending.
    const std::vector<double>& GaussPoints() const { return fGx; }
    const std::vector<double>& GaussWeights() const { return fGw; }

    // 1D Lagrange basis on equispaced nodes xi_k = -1 + 2k/p.
    double Lagrange(int k, double x) const;
    double LagrangeDeriv(int k, double x) const;

    // Shape table of the element order (layout of ipf_geometry.h, size
    // NIntPoints() * (2 NShapeF() + 3)): dN_a/dxi = L'_ga(xi) L_gb(eta),
    // dN_a/deta = L_ga(xi) L'_gb(eta) at each point, the points' (xi, eta)
    // and the tensor Gauss weights w_i w_j.
    void ShapeTable(std::vector<double>& tab) const;

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

    // Normal traction sigma_n on the inner and outer arcs: t = sigma_n * n with
    // n the outward unit normal of the domain.
    void SetNormalStress(double inner, double outer) {
        fSnInner = inner;
        fSnOuter = outer;
    }
    // External load vector of the tractions above ([neq]).
    void AssembleBoundaryLoads(TPZFMatrix<STATE>& f) const;
    // Symmetry constraints: u_y = 0 on theta = 0, u_x = 0 on theta = pi/2.
    // mask[dof] = 1 for a constrained DoF.
    void DirichletMask(std::vector<signed char>& mask) const;

    TPZFMatrix<STATE>& Solution() { return fSolution; }
    bool GetUpdateMem() const { return fUpdateMem; }
    void SetUpdateMem(bool v) { fUpdateMem = v; }

    double RIn() const { return fRin; }
    double ROut() const { return fRout; }
    double Ratio() const { return fRatio; }

private:
    int64_t NodeIndex(int64_t I, int64_t J) const { return J * fNodesR + I; }

    int fNr, fNt, fP;
    double fRin, fRout, fRatio;
    int64_t fNodesR, fNodesT;
    std::vector<double> fR;      // nr+1 radii of the cell corners
    std::vector<double> fTheta;  // ntheta+1 angles
    std::vector<int> fGa, fGb;
    std::vector<double> fGx, fGw;
    double fLambda = 0, fMu = 0, fPhi = 0, fPsi = 0, fC = 0;
    bool fHasMaterial = false;
    double fSnInner = 0, fSnOuter = 0;
    TPZFMatrix<STATE> fSolution;
    bool fUpdateMem = false;
};
