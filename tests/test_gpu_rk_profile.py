"""The reference's elastoplastic verification case solved on the GPU (40 x 40
Q2 wellbore, 10 load steps of Newton + block-Jacobi PCG), against the reference-held
Runge-Kutta profile and the reference IntPointsFEM CPU result along y = 0
(tests/golden/rk_wellbore_profile.json).  Bars: for r >= 0.2, sigma_rr and
sigma_tt within 0.012 MPa of RK (0.12 % of the 10 MPa load; the CPU oracle on
this mesh gives 0.008 / 0.0074) and sigma_xx within 0.03 MPa of the
reference's own output (oracle: 0.024); the plastic radius within one radial
cell of RK's (0.1640)."""
import numpy as np
import pytest

import intpointsfem_amd as ipf
import rk_compare

pytestmark = pytest.mark.gpu


def test_gpu_newton_matches_rk_profile(gpu):
    import torch
    m = rk_compare.rk_mesh(40)
    sm = ipf.StructMatrix(m, device=0)
    ctx = sm.context()
    ctx.set_dirichlet_mask(m.dirichlet_mask())
    ix = sm.export_indexes()
    s = sm.sizes()
    u = torch.zeros(s["neq"], dtype=torch.float64, device=gpu)
    for k in range(1, 11):
        ctx.set_rhs_linear(ix["rhs_linear"] * k / 10)
        st = ctx.newton(u, rtol=1e-10, maxit=60, cg_rtol=1e-12, cg_maxit=40000)
        assert st["converged"] == 1.0, (k, st)
    npts = s["npts"]
    strain = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ws = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.multiply(u, strain)
    ctx.compute_sigma(strain, ws)
    ctx.synchronize()
    sigma = ws.cpu().numpy().reshape(-1, 3) / ix["weight"][:, None]
    ep, _, _, _ = ctx.get_state()
    r, e_rr, e_tt, rp, rp_rk, h, e_ip = rk_compare.compare(m, sigma, ep)
    print(f"r >= 0.2: |srr - RK| {e_rr:.4f}, |stt - RK| {e_tt:.4f}, |sxx - reference IntPoints| {e_ip:.4f}; "
          f"plastic radius {rp:.4f} vs RK {rp_rk:.4f} (cell {h:.4f})")
    assert e_rr <= 0.012 and e_tt <= 0.012 and e_ip <= 0.03
    assert abs(rp - rp_rk) <= h
