"""GPU tangent assembly into CSR values (KAssembly) against the CPU oracle.

Bars:
  * CSR values: bit-identical to the oracle's element matrices
    (orc_element_tangents, the same operation order) accumulated color by
    color at fColorLSequence, the reference's gather / add / scatter
    (TPZNumericalIntegrator.cpp:166-178);
  * rhs of Assemble(K, rhs): the residual, 1e-12 relative (color order);
  * the assembled K applied to a vector equals the matrix-free K v to 1e-12
    relative (different summation order).
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_kg(B, ix, pat, wdep, nnz, ncolor):
    kc = O.element_tangents(B, wdep, ix["el_color_index"])
    kg = np.zeros(nnz)
    fl = pat["first_color_l_index"]
    cl = pat["color_l_sequence"]
    for c in range(ncolor):
        idx = cl[fl[c]:fl[c + 1]]
        assert len(np.unique(idx)) == len(idx)  # one color touches each CSR value once
        kg[idx] = kg[idx] + kc[fl[c]:fl[c + 1]]
    return kg


@pytest.mark.parametrize("mesh", [(6, 5, 1, 1.2), (12, 10, 2, 1.1), (5, 4, 3, 1.3)],
                         ids=lambda t: f"{t[0]}x{t[1]}-q{t[2]}")
def test_assemble_tangent_matches_oracle(gpu, mesh):
    import scipy.sparse as sp
    import torch
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    pat = sm.tangent_pattern()
    nnz = len(pat["ja"])
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    rhs, kg = sm.assemble_tangent(u)
    # oracle: w*Dep per point, element matrices, color-ordered accumulation
    mat = m.material()
    strain = O.multiply(B, ix["dof_indexes"], u)
    ws, wd, ep, al, mt, br = O.compute_sigma(strain, ix["weight"], mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                                             mat["cohesion"], dep=True, update_mem=True)
    assert (mt != 0).any()
    kg_o = _oracle_kg(B, ix, pat, wd, nnz, s["ncolor"])
    mism = int((kg != kg_o).sum())
    print(f"{mesh}: nnz {nnz}, {mism} CSR values differ from the oracle, max |diff| "
          f"{np.abs(kg - kg_o).max():.3e}")
    assert mism == 0
    r_o, *_ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"],
                         mat["lambda"], mat["mu"], mat["phi"], mat["psi"], mat["cohesion"], u,
                         f_lin=ix["rhs_linear"])
    assert np.abs(rhs - r_o).max() <= 1e-12 * max(np.abs(r_o - ix["rhs_linear"]).max(), np.abs(r_o).max())
    # the CSR K agrees with the matrix-free operator of the same tangent
    ia, ja = pat["ia"], pat["ja"]
    U = sp.csr_matrix((kg, ja, ia), shape=(s["neq"], s["neq"]))
    K = U + U.T - sp.diags(U.diagonal())
    ctx = sm.context()
    du = torch.as_tensor(u, dtype=torch.float64, device=gpu)
    r = torch.zeros(s["neq"], dtype=torch.float64, device=gpu)
    ctx.residual_tangent(du, r)
    v = np.random.default_rng(3).standard_normal(s["neq"])
    y = torch.zeros(s["neq"], dtype=torch.float64, device=gpu)
    ctx.tangent_apply(torch.as_tensor(v, dtype=torch.float64, device=gpu), y)
    ctx.synchronize()
    ref = K @ v
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_assemble_tangent_errors(gpu):
    import torch
    m = wb.make_mesh((4, 3, 2, 1.2))
    sm = ipf.StructMatrix(m, device=0)
    ctx = sm.context()
    neq = sm.sizes()["neq"]
    z = torch.zeros(neq, dtype=torch.float64, device=gpu)
    kg = torch.zeros(16, dtype=torch.float64, device=gpu)
    with pytest.raises(ipf.IPFError, match="no tangent pattern"):
        ctx.assemble_tangent(z, z, kg)
    ix = sm.export_indexes()
    pat = sm.tangent_pattern()
    bad = pat["color_l_sequence"].copy()
    bad[5] = len(pat["ja"])
    with pytest.raises(ipf.IPFError, match="outside"):
        ctx.set_tangent_pattern(ix["el_color_index"], bad, len(pat["ja"]))
    with pytest.raises(ipf.IPFError, match="color order"):
        ctx.set_tangent_pattern(ix["el_color_index"][::-1].copy(), pat["color_l_sequence"], len(pat["ja"]))


def test_create_assemble_equals_create_then_assemble(gpu):
    """CreateAssemble(rhs, gui) (TPZElastoPlasticIntPointsStructMatrix.cpp:
    71-77) = Create() then Assemble(K, rhs): the same CSR values and rhs bit
    for bit as the two calls on a second structural matrix of an identical
    mesh, at a plastic state."""
    mesh = (12, 10, 2, 1.1)
    m1, m2 = wb.make_mesh(mesh), wb.make_mesh(mesh)
    u = wb.lame_displacement(m1.nodes(), wb.normal_stress(10), scale=6.0)
    sm1 = ipf.StructMatrix(m1, device=0)
    r1, k1 = sm1.assemble_tangent(u)
    sm2 = ipf.StructMatrix(m2, device=0)
    r2, k2 = sm2.create_assemble(u)
    assert len(k1) == len(k2)
    np.testing.assert_array_equal(k2, k1)
    np.testing.assert_array_equal(r2, r1)
    # a second call builds a new matrix with the same values
    r3, k3 = sm2.create_assemble(u)
    np.testing.assert_array_equal(k3, k1)
    np.testing.assert_array_equal(r3, r1)


def test_clone_shares_the_device_context(gpu):
    """Clone() shares the device context by reference count: residuals
    through the clone equal the original's, a commit through one is seen by
    the other, and the clone keeps working after the original is destroyed."""
    m = wb.make_mesh((10, 8, 2, 1.1))
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    sm = ipf.StructMatrix(m, device=0)
    c = sm.clone()
    r0 = sm.assemble_rhs(u)
    np.testing.assert_array_equal(c.assemble_rhs(u), r0)
    m.set_update_mem(1)
    c.assemble_rhs(u)  # commits the plastic strain at u through the clone
    m.set_update_mem(0)
    ep_c = c.context().get_state()[0]
    assert np.abs(ep_c).max() > 0
    np.testing.assert_array_equal(sm.context().get_state()[0], ep_c)
    r_committed = sm.assemble_rhs(u)
    assert np.abs(r_committed - r0).max() > 0  # the original sees the committed state
    del sm
    import gc
    gc.collect()
    np.testing.assert_array_equal(c.assemble_rhs(u), r_committed)
