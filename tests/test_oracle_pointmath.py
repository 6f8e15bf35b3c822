"""Pins of the CPU oracle's per-point math (oracle/ipf_oracle.c) against
reference-held known answers, closed forms and brute force -- not against a
retyped copy of the oracle's own formulas.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

LAM, MU = 1388.888888888889, 2083.3333333333335  # E = 5000, nu = 0.2
K = LAM + 2 * MU / 3
G = MU
PHI = PSI = math.radians(30.0)
C = 10.0


def voigt(m):
    """3x3 symmetric -> [XX, XY, XZ, YY, YZ, ZZ]"""
    return np.array([m[0][0], m[0][1], m[0][2], m[1][1], m[1][2], m[2][2]], dtype=float)


def full(v):
    return np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])


# ---------------------------------------------------------------- spectral

def test_eigen_known_answer_notebook():
    """mathematica_notebooks/Eigenvectors/Eigenvectors.nb: A = {{5,3,0},{3,3,0},{0,0,1}}
    eigenvalues {7.16227766016838, 1, 0.8377223398316205} (:375) and the
    eigenvectors printed by the notebook's implementation of the same formulas
    (:389-397, :430-440, :472-484).  The notebook's eigenvalues are exact; the
    reference's Newton stops at |f| <= 1e-11 on the cubic of the tensor
    normalised by its largest entry (FunctionsSpectralDecomp.h:62-91), which
    leaves ~1e-10 absolute error after scaling back by 5: the bar is 1e-9."""
    ev, vec, _ = O.spectral_decomposition(voigt([[5, 3, 0], [3, 3, 0], [0, 0, 1]]))
    np.testing.assert_allclose(ev, [7.16227766016838, 1.0, 0.8377223398316205], rtol=0, atol=1e-9)
    # the third root comes from the trace (FunctionsSpectralDecomp.h:93), so the sum is exact to rounding
    assert abs(ev.sum() - 9.0) < 1e-14
    np.testing.assert_allclose(vec[0], [0.8112421851755608, 0.584710284663765, 0.0], atol=1e-9)
    np.testing.assert_allclose(np.abs(vec[1]), [0.0, 0.0, 1.0], atol=0)
    np.testing.assert_allclose(vec[2], [-0.584710284663765, 0.8112421851755609, 0.0], atol=1e-9)
    # with the exact eigenvalues, the eigenvector formulas reproduce the notebook to machine precision
    lam = [7.16227766016838, 1.0, 0.8377223398316205]
    v = O.eigenvectors(voigt([[5, 3, 0], [3, 3, 0], [0, 0, 1]]), lam)
    np.testing.assert_allclose(v[0], [0.8112421851755608, 0.584710284663765, 0.0], atol=3e-15)
    np.testing.assert_allclose(v[2], [-0.584710284663765, 0.8112421851755609, 0.0], atol=3e-15)


def test_multiplicity2_formula():
    """Eigenvectors.nb:521: B = {{3,-2,4},{-2,6,2},{4,2,3}} has the double
    eigenvalue 7 (and -2).  With the exact eigenvalues the double-root branch
    must return two vectors of the eigenspace of 7 and a unit eigenvector of -2.
    The second vector's norm uses v2[1]*v1[1] (FunctionsSpectralDecomp.h:203):
    it is an eigenvector but not of unit length."""
    Bm = np.array([[3, -2, 4], [-2, 6, 2], [4, 2, 3]], dtype=float)
    vec = O.eigenvectors(voigt(Bm), [7.0, 7.0, -2.0], maxel=1.0)
    for i, lam in ((0, 7.0), (1, 7.0), (2, -2.0)):
        np.testing.assert_allclose(Bm @ vec[i], lam * vec[i], atol=1e-13)
        assert abs(np.linalg.norm(vec[i]) - 1) < 1e-15  # x[0] branch: v2[1] = 0, the typo is inert
    # x[1] branch (|syy - lambda| largest): the typo changes the second norm
    q = np.array([1.0, 3.0, 2.0]) / math.sqrt(14.0)  # sxy != syz so v2[1]*v1[1] != v2[1]**2
    M = 2.0 * np.eye(3) - 3.0 * np.outer(q, q)
    vec = O.eigenvectors(voigt(M), [2.0, 2.0, -1.0], maxel=1.0)
    for i, lam in ((0, 2.0), (1, 2.0), (2, -1.0)):
        np.testing.assert_allclose(M @ vec[i], lam * vec[i], atol=1e-13)
    assert abs(np.linalg.norm(vec[0]) - 1) < 1e-15
    assert abs(np.linalg.norm(vec[1]) - 1) > 1e-3  # FunctionsSpectralDecomp.h:203, reproduced


def test_triple_eigenvalue_identity():
    ev, vec, _ = O.spectral_decomposition(voigt(np.eye(3) * -4.0))
    np.testing.assert_array_equal(ev, [-4.0, -4.0, -4.0])
    np.testing.assert_array_equal(vec, np.eye(3))


@pytest.mark.parametrize("seed", range(4))
def test_plane_strain_tensors_vs_lapack(seed):
    """Random plane-strain stress tensors (XZ = YZ = 0, the only kind the path
    produces): eigenvalues match LAPACK (numpy eigh), descending, and
    sum lambda v v^T reconstructs the tensor."""
    rng = np.random.default_rng(seed)
    for _ in range(200):
        s = rng.standard_normal(4) * 10 ** rng.uniform(-3, 3)
        m = np.array([[s[0], s[1], 0], [s[1], s[2], 0], [0, 0, s[3]]])
        ev, vec, after = O.spectral_decomposition(voigt(m))
        ref = np.sort(np.linalg.eigvalsh(m))[::-1]
        scale = np.abs(ref).max()
        gaps = np.abs(np.diff(np.sort(ref))).min()
        if gaps < 1e-3 * scale:
            continue  # near-multiple roots: Newton is only linearly convergent there
        np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-9 * scale)
        rec = sum(ev[i] * np.outer(vec[i], vec[i]) for i in range(3))
        np.testing.assert_allclose(rec, m, rtol=0, atol=1e-8 * scale)
        np.testing.assert_allclose(after, voigt(m), rtol=1e-15, atol=0)  # normalise + scale back


# ---------------------------------------------------------------- projection

def mc_planes(s, phi=PHI, c=C):
    """All six Mohr-Coulomb planes (sa - sb) + (sa + sb) sin(phi) - 2 c cos(phi)."""
    out = []
    for a in range(3):
        for b in range(3):
            if a != b:
                out.append(s[a] - s[b] + (s[a] + s[b]) * math.sin(phi) - 2 * c * math.cos(phi))
    return np.array(out)


def principal_compliance():
    D = np.full((3, 3), LAM) + 2 * MU * np.eye(3)
    return np.linalg.inv(D)


def closest_point(tr):
    """Brute force: the point of the Mohr-Coulomb set closest to the trial
    stress in the complementary-energy norm (associated flow, psi = phi)."""
    from scipy.optimize import minimize
    Cm = principal_compliance()
    cons = [{"type": "ineq", "fun": (lambda s, k=k: -mc_planes(s)[k])} for k in range(6)]
    best = None
    p0 = C / math.tan(PHI)
    for x0 in (np.array(tr) * 0.5, np.full(3, p0 - 1.0), np.array(tr) - 1.0):
        res = minimize(lambda s: (s - tr) @ Cm @ (s - tr), x0, constraints=cons, method="SLSQP",
                       options={"ftol": 1e-15, "maxiter": 500})
        if best is None or res.fun < best.fun:
            best = res
    return best.x


def sorted_trial(rng, scale):
    s = np.sort(rng.standard_normal(3) * scale)[::-1]
    return s


def test_elastic_region_unchanged():
    tr = np.array([-5.0, -10.0, -20.0])
    assert mc_planes(tr).max() < 0
    sp, mt, alpha, br, _ = O.project_sigma(tr, PHI, PSI, C, K, G)
    assert mt == 0 and br == 0
    np.testing.assert_array_equal(sp, tr)


def test_main_plane_closed_form():
    """H = 0: the main-plane return is linear in gamma, gamma = phi_tr / constA,
    and the projected stress lies on the plane."""
    tr = np.array([30.0, -10.0, -40.0])
    sphi, spsi = math.sin(PHI), math.sin(PSI)
    phi_tr = tr[0] - tr[2] + (tr[0] + tr[2]) * sphi - 2 * C * math.cos(PHI)
    constA = 4 * G * (1 + sphi * spsi / 3) + 4 * K * sphi * spsi
    gamma = phi_tr / constA
    sp, mt, alpha, br, _ = O.project_sigma(tr, PHI, PSI, C, K, G)
    assert mt == 1 and br == 1
    assert abs(alpha - gamma * 2 * math.cos(PHI)) < 1e-14 * abs(alpha)
    plane = sp[0] - sp[2] + (sp[0] + sp[2]) * sphi - 2 * C * math.cos(PHI)
    assert abs(plane) < 1e-11 * np.abs(tr).max()


def test_apex_closed_form():
    """Apex return: all principal stresses equal c*cot(phi)."""
    tr = np.array([60.0, 55.0, 50.0])
    sp, mt, alpha, br, _ = O.project_sigma(tr, PHI, PSI, C, K, G)
    assert mt == -1 and br == 4
    np.testing.assert_allclose(sp, C / math.tan(PHI), rtol=1e-14)


@pytest.mark.parametrize("seed", range(3))
def test_return_map_is_closest_point(seed):
    """For psi = phi (associated flow) the return map is the closest-point
    projection in the energy norm; compare against constrained minimisation
    for trial states in every branch region."""
    rng = np.random.default_rng(100 + seed)
    seen = set()
    for _ in range(40):
        tr = sorted_trial(rng, 40.0)
        sp, mt, alpha, br, _ = O.project_sigma(tr, PHI, PSI, C, K, G)
        seen.add(br)
        ref = closest_point(tr)
        np.testing.assert_allclose(np.sort(sp)[::-1], np.sort(ref)[::-1], rtol=0, atol=2e-5 * max(1.0, np.abs(tr).max()))
        assert mc_planes(sp).max() < 1e-9 * max(1.0, np.abs(tr).max())
    assert len(seen) >= 3, seen


# ---------------------------------------------------------------- tangent

def dmat3(w=1.0):
    return w * np.array([[LAM + 2 * MU, 0, LAM], [0, MU, 0], [LAM, 0, LAM + 2 * MU]])


def test_elastic_tangent_is_D():
    """Elastic point: Dep = D in [xx, xy, yy] with engineering shear strain."""
    out = O.point_sigma([1e-4, 2e-4, -3e-4], np.zeros(6), LAM, MU, PHI, PSI, C, weight=0.7, dep=True)
    assert out["mtype"] == 0
    D = dmat3(0.7)
    # the spectral form rebuilds D from eigenvectors carrying the Newton tolerance: 1e-11 relative to max(D)
    np.testing.assert_allclose(out["wdep"], D, rtol=0, atol=1e-11 * np.abs(D).max())


def _sigma(e3, ep6):
    return O.point_sigma(e3, ep6, LAM, MU, PHI, PSI, C)["wsigma"]


@pytest.mark.parametrize("e3", [[4e-3, 0.0, -5e-3], [6e-3, 0.0, -8e-3], [2e-3, 0.0, -9e-3]])
def test_plastic_tangent_coaxial_columns_match_finite_differences(e3):
    """Main-plane points with principal axes along x, y, z (zero shear): the
    xx and yy columns of Dep carry no eigenvector-rotation term, so they must
    equal central finite differences of the stress update.  The xy column is
    not the exact derivative in the reference: its rotation term uses the trial
    stress eigenvalues and a tempMat accumulated over the eigen-pairs
    (FunctionsTangentOperator.h:79-121); that term is pinned only by the
    elastic case below and the GPU-vs-oracle bit-exact tests."""
    ep6 = np.zeros(6)
    out = O.point_sigma(e3, ep6, LAM, MU, PHI, PSI, C, dep=True)
    assert out["branch"] == 1, out["branch"]
    h = 1e-8
    for j in (0, 2):
        ep = np.array(e3, float)
        em = np.array(e3, float)
        ep[j] += h
        em[j] -= h
        fd = (_sigma(ep, ep6) - _sigma(em, ep6)) / (2 * h)
        np.testing.assert_allclose(out["wdep"][:, j], fd, rtol=0, atol=2e-5 * np.abs(fd).max())


def test_plastic_tangent_shear_uses_trial_eigenvalues():
    """Documents the reference's shear term: for a main-plane point with axes
    along x, y, z the xy-xy entry is the elastic shear modulus G (the rotation
    factor (s_i - s_j)/(e_i - e_j) is evaluated with trial stresses, which
    equals 2G), not the exact consistent value from the projected stresses."""
    e3 = [4e-3, 0.0, -5e-3]
    out = O.point_sigma(e3, np.zeros(6), LAM, MU, PHI, PSI, C, dep=True)
    assert out["branch"] == 1
    assert abs(out["wdep"][1, 1] - G) < 1e-9 * G
    h = 1e-8
    fd = (_sigma([e3[0], h, e3[2]], np.zeros(6)) - _sigma([e3[0], -h, e3[2]], np.zeros(6)))[1] / (2 * h)
    assert fd < 0.99 * G  # the exact derivative is softer


# ------------------------------------------- tangent on every branch (FD + covariance)

def _coaxial_states(psi, rng, per_branch=3):
    """Seeded coaxial strains (zero shear, so the principal axes are x, y, z)
    whose branch survives a 1e-3 relative change of either normal strain:
    each state lies inside its branch region, away from the switches."""
    found = {}
    while any(len(found.get(b, [])) < per_branch for b in (1, 2, 3, 4)):
        exx, eyy = rng.uniform(-0.03, 0.03, 2)
        b = O.point_sigma([exx, 0.0, eyy], np.zeros(6), LAM, MU, PHI, psi, C)["branch"]
        if b == 0 or len(found.get(b, [])) >= per_branch:
            continue
        if all(O.point_sigma([exx * (1 + dx), 0.0, eyy * (1 + dy)], np.zeros(6), LAM, MU, PHI, psi, C)["branch"] == b
               for dx, dy in ((1e-3, 0), (-1e-3, 0), (0, 1e-3), (0, -1e-3))):
            found.setdefault(b, []).append([exx, 0.0, eyy])
    return found


@pytest.mark.parametrize("psi_deg", [30.0, 10.0])
def test_tangent_normal_columns_match_fd_on_every_plastic_branch(psi_deg):
    """Main plane, right edge, left edge and apex (FunctionsSigmaProjection.h:
    68-88, 163-188, 262-287, 328-332) through TangentOperator
    (FunctionsTangentOperator.h:37-123): with principal axes along x, y, z the
    xx and yy columns of Dep carry no eigenvector-rotation term, so they are
    the exact derivative of the stress update, which is piecewise linear
    inside a branch (H = 0).  Central differences with h = 1e-6 on strains of
    order 1e-2 agree to rounding (measured <= 3e-11 of max|Dep|); a dropped
    term, swapped sin(phi) / sin(psi) or a wrong edge formula moves an entry
    by a sizeable fraction of it.  psi = 10 deg separates the phi and psi
    terms of the closed forms.  Apex: both columns are zero."""
    psi = math.radians(psi_deg)
    states = _coaxial_states(psi, np.random.default_rng(int(psi_deg)))
    h = 1e-6
    for b, lst in states.items():
        for e3 in lst:
            out = O.point_sigma(e3, np.zeros(6), LAM, MU, PHI, psi, C, dep=True)
            assert out["branch"] == b
            scale = np.abs(out["wdep"]).max()
            for j in (0, 2):
                ep = np.array(e3)
                em = np.array(e3)
                ep[j] += h
                em[j] -= h
                fd = (O.point_sigma(ep, np.zeros(6), LAM, MU, PHI, psi, C)["wsigma"]
                      - O.point_sigma(em, np.zeros(6), LAM, MU, PHI, psi, C)["wsigma"]) / (2 * h)
                np.testing.assert_allclose(out["wdep"][:, j], fd, rtol=0, atol=1e-8 * scale,
                                           err_msg=f"branch {b} strain {e3} column {j}")
            if b == 4:
                np.testing.assert_array_equal(out["wdep"][:, [0, 2]], 0.0)


def _t_sigma(th):
    """[sxx, sxy, syy] of R s R^T for a rotation by th about z."""
    c, s = math.cos(th), math.sin(th)
    return np.array([[c * c, -2 * c * s, s * s], [c * s, c * c - s * s, -c * s], [s * s, 2 * c * s, c * c]])


def _t_eps(th):
    """[exx, gxy, eyy] (engineering shear) of R e R^T."""
    c, s = math.cos(th), math.sin(th)
    return np.array([[c * c, -c * s, s * s], [2 * c * s, c * c - s * s, -2 * c * s], [s * s, c * s, c * c]])


@pytest.mark.parametrize("psi_deg", [30.0, 10.0])
def test_tangent_is_frame_covariant_on_every_branch(psi_deg):
    """The stress update is an isotropic function of the strain, so its
    tangent must transform as Dep(R e R^T) = T_s Dep(e) T_e^-1 for a rotation
    R about z.  Starting from the coaxial states above (normal columns pinned
    by finite differences, the xy-xy entry = G from the trial-eigenvalue
    factor, test above), this pins the eigenvector-rotation term of
    FunctionsTangentOperator.h:79-121 -- eigenvector products, the factor 2
    of the off-diagonal Voigt columns, signs -- for every orientation, on
    every branch and the elastic one.  In plane strain the in-plane block
    does not see the accumulated tempMat quirk (the extra pairs all involve
    the z eigenvector, whose products with in-plane columns vanish).
    Measured agreement <= 3e-9 of max|Dep| (the eigenvalues carry the
    Newton tolerance, amplified by 1 / gap in the rotation factor where two
    strain eigenvalues are close); a tangent with the xy-xy entry halved
    misses by ~0.1."""
    psi = math.radians(psi_deg)
    states = _coaxial_states(psi, np.random.default_rng(10 + int(psi_deg)), per_branch=2)
    states[0] = [[-0.012, 0.0, -0.0046], [0.001, 0.0, -0.002]]
    for b, lst in states.items():
        for e3 in lst:
            out = O.point_sigma(e3, np.zeros(6), LAM, MU, PHI, psi, C, dep=True)
            assert out["branch"] == b
            D0 = out["wdep"]
            scale = np.abs(D0).max()
            wrong = D0.copy()
            wrong[1, 1] *= 0.5
            for th in np.linspace(0.15, 3.0, 9):
                E = np.array([[e3[0], e3[1] / 2], [e3[1] / 2, e3[2]]])
                R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
                Er = R @ E @ R.T
                o2 = O.point_sigma([Er[0, 0], 2 * Er[0, 1], Er[1, 1]], np.zeros(6), LAM, MU, PHI, psi, C, dep=True)
                assert o2["branch"] == b
                Ti = np.linalg.inv(_t_eps(th))
                np.testing.assert_allclose(o2["wdep"], _t_sigma(th) @ D0 @ Ti, rtol=0, atol=1e-8 * scale,
                                           err_msg=f"branch {b} strain {e3} theta {th}")
                np.testing.assert_allclose(o2["wsigma"], _t_sigma(th) @ out["wsigma"], rtol=0,
                                           atol=1e-9 * np.abs(out["wsigma"]).max())
                if abs(math.sin(2 * th)) > 0.3:  # the check above would catch a halved shear entry
                    assert np.abs(o2["wdep"] - _t_sigma(th) @ wrong @ Ti).max() > 1e-2 * scale
