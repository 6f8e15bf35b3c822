"""How the residual starts from f_lin (r = f_lin, then every color adds
-B^T(w sigma); TPZElastoPlasticIntPointsStructMatrix.cpp:147).

When every DoF belongs to an element, the color launches read f_lin
themselves at each DoF's first write in launch order (no copy of f_lin into r
before them); when some DoF belongs to no element, the library copies f_lin
into r first.  Bars: both forms give bit-identical residuals on the shared
DoFs, the unowned DoFs hold f_lin exactly, the result does not depend on what
r held before the call (NaN-filled here), and the residual matches the oracle
within 1e-12 of max(|f_int|, |f_lin|) (color accumulation order).

The tangent product y = K v starts the same way: the first writes start
from 0 when every DoF belongs to an element, else y is cleared first.  Bars:
both forms give bit-identical K v on the shared DoFs, exact zeros on the
unowned ones, whatever y held before (NaN-filled).
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", [(6, 5, 2, 1.2), (5, 4, 3, 1.3)], ids=["q2", "q3"])
@pytest.mark.parametrize("fmt", ["geometry-fp64", "dense-fp64"])
def test_unowned_dofs_keep_f_lin(gpu, mesh, fmt):
    import torch
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    neq, extra = s["neq"], 3
    d = m.material()
    mat = (d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"])
    rng = np.random.default_rng(7)
    flin = np.concatenate([ix["rhs_linear"], rng.standard_normal(extra)])
    u = np.concatenate([wb.lame_displacement(m.nodes(), wb.normal_stress(10)), rng.standard_normal(extra)])
    v = np.concatenate([u[:neq] * 3.0 - 1e-4, rng.standard_normal(extra)])
    out, kv = {}, {}
    (nip, nsh), tab = m.shape_table()
    corners = m.corners()
    ci = np.asarray(ix["color_indexes"], dtype=np.int64)
    for n in (neq, neq + extra):
        # the reference's color indexes are dof + color * neq (TPZNumericalIntegrator.cpp:452-536)
        cin = (ci % neq + (ci // neq) * n).astype(np.int32)
        ctx = ipf.DeviceContext(device=0)
        if fmt == "geometry-fp64":
            ctx.transfer_geometry([(nip, nsh)], tab, np.zeros(corners.shape[0], np.int32), corners,
                                  ix["dof_indexes"], cin, s["ncolor"], n)
        else:
            ctx.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                              b["storage"], ix["dof_indexes"], cin, s["ncolor"], n, ix["weight"])
        ctx.set_material(*mat)
        ctx.set_b_format(fmt)
        ctx.set_rhs_linear(flin[:n])
        r = torch.full((n,), float("nan"), dtype=torch.float64, device=gpu)
        ctx.residual(torch.as_tensor(u[:n], device=gpu), r)
        ctx.synchronize()
        out[n] = r.cpu().numpy()
        # ipf_residual_timed: the same residual; its launch 0 (the copy) reads
        # exactly 0 ms when nothing runs there, a positive time when it does
        r.fill_(float("nan"))
        ms, pts = ctx.residual_timed(torch.as_tensor(u[:n], device=gpu), r)
        assert np.array_equal(r.cpu().numpy(), out[n])
        assert pts[0] == 0 and pts[1:].sum() == s["npts"] and (ms[1:] > 0).all()
        assert (ms[0] == 0.0) if n == neq else (ms[0] > 0.0)
        ctx.set_b_format(fmt, tangent=fmt)
        ctx.residual_tangent(torch.as_tensor(u[:n], device=gpu), torch.empty_like(r))
        y = torch.full((n,), float("nan"), dtype=torch.float64, device=gpu)
        ctx.tangent_apply(torch.as_tensor(v[:n], device=gpu), y)
        ctx.synchronize()
        kv[n] = y.cpu().numpy()
    folded, copied = out[neq], out[neq + extra]
    assert np.array_equal(folded, copied[:neq])
    assert np.array_equal(copied[neq:], flin[neq:])
    assert np.array_equal(kv[neq], kv[neq + extra][:neq])
    assert (kv[neq + extra][neq:] == 0).all() and np.abs(kv[neq]).max() > 0
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    r_o, *_ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], neq, ix["weight"], *mat, u[:neq],
                         f_lin=flin[:neq])
    tol = 1e-12 * max(np.abs(r_o - flin[:neq]).max(), np.abs(flin[:neq]).max())
    assert np.abs(folded - r_o).max() <= tol
