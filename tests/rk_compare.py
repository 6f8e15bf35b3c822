"""Helpers shared by the RK-profile tests: the reference's elastoplastic
verification case (tests/golden/rk_wellbore_profile.json) on this build's
wellbore mesh, and the comparison of a converged solution with the RK profile
along the first row of integration points (theta ~ 0, where sigma_xx ~
sigma_rr and sigma_yy ~ sigma_tt)."""
import json
import math
import os

import numpy as np

import intpointsfem_amd as ipf

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    with open(os.path.join(HERE, "golden", "rk_wellbore_profile.json")) as f:
        return json.load(f)


def rk_mesh(nr):
    c = fixture()["case"]
    m = ipf.CompMesh(nr, nr, 2, c["r_in"], c["r_out"], 1.1 ** (40.0 / nr))
    m.set_material(c["E"], c["nu"], c["cohesion"], math.radians(c["phi_deg"]), math.radians(c["psi_deg"]))
    m.set_normal_stress(c["sigma_n_inner"], c["sigma_n_outer"])
    return m


def compare(m, sigma3, plastic_strain6):
    """sigma3 [npts, 3] (xx, xy, yy) and committed plastic strain [npts, 6] of a
    converged solution; returns (r, max |srr - rk| and |stt - rk| for r >= 0.2,
    plastic radius of the first point row, RK plastic radius, radial cell size
    there, max |sxx - reference IntPoints CPU| for r >= 0.2)."""
    fx = fixture()
    rk = fx["rk"]
    xy = m.point_coords()
    r = np.hypot(xy[:, 0], xy[:, 1])
    th = np.arctan2(xy[:, 1], xy[:, 0])
    row = th < th.min() + 1e-9
    o = np.argsort(r[row])
    rr = r[row][o]
    s = np.asarray(sigma3).reshape(-1, 3)[row][o]
    far = rr >= 0.2
    e_rr = np.abs(s[far, 0] - np.interp(rr[far], rk["radius"], rk["sigma_rr"])).max()
    e_tt = np.abs(s[far, 2] - np.interp(rr[far], rk["radius"], rk["sigma_tt"])).max()
    ep = np.asarray(plastic_strain6).reshape(-1, 6)[row][o]
    rp = rr[(np.abs(ep[:, 0]) + np.abs(ep[:, 3])) > 1e-9].max()
    rp_rk = fx["case"]["rk_plastic_radius"]
    # radial cell containing the RK plastic radius
    nodes_r = np.unique(np.round(np.hypot(*m.nodes().T), 12))
    edges = nodes_r[::m.porder] if m.porder > 1 else nodes_r
    k = np.searchsorted(edges, rp_rk)
    h = edges[min(k, len(edges) - 1)] - edges[max(k - 1, 0)]
    ip = fx["intpoints_cpu"]
    e_ip = np.abs(s[far, 0] - np.interp(rr[far], ip["x"], ip["sigma_xx"])).max()
    return rr, e_rr, e_tt, rp, rp_rk, h, e_ip
