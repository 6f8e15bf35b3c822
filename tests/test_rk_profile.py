"""The reference's elastoplastic verification case (main.cpp:374-445:
E = 2000, nu = 0.2, c = 5, phi = psi = 20 deg, wall stress increment -10)
solved by the CPU oracle's Newton on a 20 x 20 Q2 wellbore in 10 load steps,
against the reference-held Runge-Kutta profile and the reference IntPointsFEM
CPU result (tests/golden/rk_wellbore_profile.json, written by
tests/golden/make_rk_profile.py).

Bars (the RK and IntPoints files are float32 text and the meshes differ, so
they pin the physics, not bits): for r >= 0.2, sigma_rr within 0.04 MPa and
sigma_tt within 0.02 MPa of RK (measured 0.029 / 0.012 of a 10 MPa load) and
sigma_xx within 0.06 MPa of the reference's own output (measured 0.049); the
plastic radius within one radial cell of RK's.  Near the wall (r < 0.2) the
Gauss-point stresses of both FE codes sit up to 0.6 MPa from RK (the
reference's own file: -9.99 at r = 0.1 against RK's -9.75)."""
import numpy as np

import intpointsfem_amd as ipf
import rk_compare
from oracle import oracle as O
from oracle.newton import OracleNewton


def test_oracle_newton_matches_rk_profile():
    m = rk_compare.rk_mesh(20)
    sm = ipf.StructMatrix(m, device=-1)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    orc = OracleNewton(B, ix, s["neq"], s["ncolor"], m.material(), m.dirichlet_mask(), nthreads=8)
    for k in range(1, 11):
        st = orc.load_step(ix["rhs_linear"] * k / 10, rtol=1e-10, maxit=60)
        assert st["converged"], (k, st)
    mat = m.material()
    strain = O.multiply(B, ix["dof_indexes"], orc.u)
    ws, *_ = O.compute_sigma(strain, np.ones(s["npts"]), mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                             mat["cohesion"], plastic_strain6=orc.ep.copy(), alpha=orc.alpha.copy())
    r, e_rr, e_tt, rp, rp_rk, h, e_ip = rk_compare.compare(m, ws, orc.ep)
    print(f"r >= 0.2: |srr - RK| {e_rr:.4f}, |stt - RK| {e_tt:.4f}, |sxx - reference IntPoints| {e_ip:.4f}; "
          f"plastic radius {rp:.4f} vs RK {rp_rk:.4f} (cell {h:.4f})")
    assert e_rr <= 0.04 and e_tt <= 0.02 and e_ip <= 0.06
    assert abs(rp - rp_rk) <= h
