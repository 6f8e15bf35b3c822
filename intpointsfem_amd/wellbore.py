"""Wellbore problem definitions (the configurations the path is exercised on).

Quarter annulus r in [0.1, 4] m, theta in [0, pi/2], plane strain, E = 5000
MPa, nu = 0.2, Mohr-Coulomb c = 10 MPa, phi = psi = 30 deg, perfect
plasticity.  As in the reference driver (main.cpp:395-405) the far-field
stress sigma0 = -50 MPa is the reference state: the mesh carries zero initial
stress and the load is the increment of the wellbore pressure.  The
reference's normal-stress boundary value is -(pw - sigma0) and it is the
radial stress sigma_rr imposed on the wall (its IntPoints CPU profile,
mathematica_notebooks/RK_Verification/ElastoPlasticIntPoints-CPU.csv:2, has
sigma_rr(0.1) = -9.99 for the value -10 of main.cpp:400).  pw ramps -50 ->
-19.5 MPa over 10 load steps, so sigma_rr(r_in) = -3.05 k at step k; the outer
arc carries no increment.  Symmetry: u_y = 0 on theta = 0, u_x = 0 on
theta = pi/2.
"""
import math

import numpy as np

E = 5000.0
NU = 0.2
COHESION = 10.0
PHI = math.radians(30.0)
PSI = math.radians(30.0)
R_IN = 0.1
R_OUT = 4.0
SIGMA0 = -50.0
PW_FINAL = -19.5
NSTEPS = 10

# structured meshes: (nr, ntheta, p, radial growth ratio)
MESHES = {
    "40x40-q2": (40, 40, 2, 1.1),
    "800x800-q2": (800, 800, 2, 1.1 ** (40.0 / 800.0)),
    "800x800-q3": (800, 800, 3, 1.1 ** (40.0 / 800.0)),
}


def lame_parameters(E_=E, nu=NU):
    mu = E_ / (2.0 * (1.0 + nu))
    lam = nu * E_ / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return lam, mu


def normal_stress(step, nsteps=NSTEPS, pw_final=PW_FINAL, sigma0=SIGMA0):
    """Radial stress imposed on the inner arc at load step `step` (1-based):
    the reference's boundary value -(pw - sigma0) (main.cpp:399-400)."""
    pw = sigma0 + (pw_final - sigma0) * step / nsteps
    return -(pw - sigma0)


def make_mesh(name_or_tuple, material=True, step=NSTEPS):
    """CompMesh for a named configuration with material and step-`step` load."""
    from . import CompMesh
    nr, nt, p, q = MESHES[name_or_tuple] if isinstance(name_or_tuple, str) else name_or_tuple
    m = CompMesh(nr, nt, p, R_IN, R_OUT, q)
    if material:
        m.set_material(E, NU, COHESION, PHI, PSI)
        m.set_normal_stress(normal_stress(step), 0.0)
    return m


def lame_displacement(xy, sigma_n, scale=1.0, E_=E, nu=NU, r_in=R_IN, r_out=R_OUT):
    """Linear-elastic thick-cylinder displacement with sigma_rr(r_in) = sigma_n
    and sigma_rr(r_out) = 0 (plane strain), at nodes `xy` [n, 2]; returns the
    DoF vector [2n] (x, y interleaved) times `scale`.
    sigma_rr(r) = 2(l+m)A - 2mB/r^2, u_r = A r + B / r."""
    lam, mu = lame_parameters(E_, nu)
    # sigma_rr(r_in) = sigma_n (outward normal -e_r => t = sigma_rr * n), sigma_rr(r_out) = 0
    M = np.array([[2 * (lam + mu), -2 * mu / r_in ** 2], [2 * (lam + mu), -2 * mu / r_out ** 2]])
    A, B = np.linalg.solve(M, [sigma_n, 0.0])
    r = np.hypot(xy[:, 0], xy[:, 1])
    ur = A * r + B / r
    u = np.zeros(2 * len(xy))
    u[0::2] = ur * xy[:, 0] / r * scale
    u[1::2] = ur * xy[:, 1] / r * scale
    return u
