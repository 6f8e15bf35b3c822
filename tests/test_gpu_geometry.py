"""Setup on the device (ipf_transfer_geometry) and the geometric B format.

Bars, all bit-exact:
  * a context set up on the GPU from the element corners and the shape table
    and one given the host-built blocks and weights (ipf_transfer_data) give
    the same strains B u (every B value enters), the same w sigma (every
    weight enters), and the same residual, branch flags and plastic strains
    in dense and compact FP64;
  * the geometric format (B and w recomputed inside the kernels) gives the
    compact FP64 residual, flags and plastic strains, with and without the
    commit, and the same tangent product K v;
  * the same for a batch of instances of orders 1, 2 and 3 (mixed shapes and
    materials in one context).
The host blocks themselves are pinned on the CPU (tests/test_frontend.py:
the reference's emitted B of one element, patch tests, and the geometry
formula restated in numpy).
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import batch
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu

MESHES = [(6, 5, 1, 1.2), (5, 4, 2, 1.1), (4, 4, 3, 1.3)]


def _host_ctx(sm, mesh):
    s = sm.sizes()
    b = sm.export_blocks()
    ix = sm.export_indexes()
    ctx = ipf.DeviceContext(device=0)
    ctx.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                      b["storage"], ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"])
    mat = mesh.material()
    ctx.set_material(mat["lambda"], mat["mu"], mat["phi"], mat["psi"], mat["cohesion"])
    ctx.set_rhs_linear(ix["rhs_linear"])
    return ctx


def _mesh(spec):
    nr, nt, p, q = spec
    m = ipf.CompMesh(nr, nt, p, wb.R_IN, wb.R_OUT, q)
    m.set_material(wb.E, wb.NU, wb.COHESION, wb.PHI, wb.PSI)
    m.set_normal_stress(wb.normal_stress(wb.NSTEPS), 0.0)
    return m


def _residual(ctx, u, neq, fmt, commit, gpu):
    import torch
    ctx.set_b_format(fmt)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.set_update_mem(commit)
    ctx.residual(u, r)
    ctx.set_update_mem(False)
    ctx.synchronize()
    return r.cpu().numpy()


def _same_state(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("spec", MESHES, ids=["q1", "q2", "q3"])
def test_device_setup_equals_host_blocks(gpu, spec):
    import torch
    m = _mesh(spec)
    sm = ipf.StructMatrix(m, device=0)
    s = sm.sizes()
    neq, npts = s["neq"], s["npts"]
    cg = sm.context()
    ch = _host_ctx(sm, m)
    assert cg.b_format() == ("geometry-fp64", "geometry-fp64")
    u = torch.as_tensor(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), device=gpu)
    out = []
    for ctx in (cg, ch):
        strain = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
        ws = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
        ctx.multiply(u, strain)
        ctx.compute_sigma(strain, ws)
        ctx.synchronize()
        out.append((strain.cpu().numpy(), ws.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    for fmt in ("dense-fp64", "compact-fp64"):
        for ctx in (cg, ch):
            ctx.set_state(np.zeros(6 * npts))
        rg = _residual(cg, u, neq, fmt, True, gpu)
        rh = _residual(ch, u, neq, fmt, True, gpu)
        assert np.array_equal(rg, rh), fmt
        sg, shs = cg.get_state(), ch.get_state()
        assert _same_state(sg, shs), fmt
    assert (sg[2] != 0).any(), "the load should make some points plastic"


@pytest.mark.parametrize("spec", MESHES, ids=["q1", "q2", "q3"])
def test_geometric_format_equals_compact(gpu, spec):
    import torch
    m = _mesh(spec)
    sm = ipf.StructMatrix(m, device=0)
    s = sm.sizes()
    neq, npts = s["neq"], s["npts"]
    ctx = sm.context()
    u = torch.as_tensor(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), device=gpu)
    # not committing, from the zero state and from a plastic state
    for state in ("zero", "plastic"):
        if state == "plastic":
            _residual(ctx, u * 1.3, neq, "compact-fp64", True, gpu)
        st0 = ctx.get_state()
        rc = _residual(ctx, u, neq, "compact-fp64", False, gpu)
        rg = _residual(ctx, u, neq, "geometry-fp64", False, gpu)
        assert np.array_equal(rc, rg), state
        assert _same_state(st0, ctx.get_state())
    # committing
    ctx.set_state(np.zeros(6 * npts))
    rc = _residual(ctx, u, neq, "compact-fp64", True, gpu)
    stc = ctx.get_state()
    ctx.set_state(np.zeros(6 * npts))
    rg = _residual(ctx, u, neq, "geometry-fp64", True, gpu)
    assert np.array_equal(rc, rg)
    assert _same_state(stc, ctx.get_state())
    assert (stc[2] != 0).any()
    # tangent product at the committed state's trial tangent
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(u, r)
    v = torch.as_tensor(np.random.default_rng(3).normal(size=neq), device=gpu)
    ys = []
    for fmt in ("compact-fp64", "geometry-fp64"):
        ctx.set_b_format("compact-fp64", fmt)
        y = torch.zeros(neq, dtype=torch.float64, device=gpu)
        ctx.tangent_apply(v, y)
        ctx.synchronize()
        ys.append(y.cpu().numpy())
    assert np.array_equal(ys[0], ys[1])
    assert np.abs(ys[0]).max() > 0


def test_batch_device_setup(gpu):
    """Instances of orders 1-3 in one context: device setup == host blocks,
    geometric == compact."""
    import torch
    table = [dict(N=7, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
             dict(N=5, p=3, phi_deg=35.0, cohesion=2.0, E=15000.0, nu=0.28, pw_final=-28.0),
             dict(N=6, p=2, phi_deg=20.0, cohesion=1.5, E=8000.0, nu=0.16, pw_final=-16.0)]
    meshes = [batch.instance_mesh(i) for i in table]
    Bd = batch.Batch(meshes, device=0, setup="device")
    Bh = batch.Batch(meshes, device=0, setup="host")
    us = [wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, i["pw_final"], wb.SIGMA0),
                               E_=i["E"], nu=i["nu"]) * 3.0 for i, m in zip(table, meshes)]
    u = torch.as_tensor(np.concatenate(us), device=gpu)
    rd = _residual(Bd.ctx, u, Bd.neq, "compact-fp64", True, gpu)
    rh = _residual(Bh.ctx, u, Bh.neq, "compact-fp64", True, gpu)
    assert np.array_equal(rd, rh)
    std = Bd.ctx.get_state()
    assert _same_state(std, Bh.ctx.get_state())
    assert (std[2] != 0).any()
    Bd.ctx.set_state(np.zeros(6 * Bd.npts))
    rg = _residual(Bd.ctx, u, Bd.neq, "geometry-fp64", True, gpu)
    assert np.array_equal(rg, rd)
    assert _same_state(std, Bd.ctx.get_state())


def test_geometric_format_needs_geometry(gpu):
    """A context given host blocks has no geometry: the format is refused."""
    m = _mesh(MESHES[1])
    sm = ipf.StructMatrix(m, device=0)
    ch = _host_ctx(sm, m)
    with pytest.raises(ipf.IPFError, match="geometry"):
        ch.set_b_format("geometry-fp64")
    assert ch.b_format() == ("compact-fp64", "compact-fp64")


def test_transfer_geometry_errors_and_empty_mesh(gpu):
    """ipf_transfer_geometry rejects an unknown shape index, a non-finite
    corner, a shape without a kernel and two shapes of equal sizes, each with
    IPF_EINVAL / IPF_EUNSUPPORTED and a message; a mesh without elements
    transfers and its residual is f_lin."""
    import torch
    m = _mesh(MESHES[1])
    (nip, nsh), tab = m.shape_table()
    sm = ipf.StructMatrix(m, device=-1)
    s = sm.sizes()
    ix = sm.export_indexes()
    corners = m.corners()
    nb = corners.shape[0]
    args = dict(shape_sizes=[(nip, nsh)], shape_tables=tab, elem_shape=np.zeros(nb, np.int32), corners=corners,
                dof_indexes=ix["dof_indexes"], color_indexes=ix["color_indexes"], ncolor=s["ncolor"], neq=s["neq"])
    ctx = ipf.DeviceContext(device=0)
    ctx.transfer_geometry(**args)  # the good input
    bad = dict(args, elem_shape=np.full(nb, 1, np.int32))
    with pytest.raises(ipf.IPFError, match="shape 1 not in"):
        ctx.transfer_geometry(**bad)
    c2 = corners.copy()
    c2[3, 5] = np.nan
    with pytest.raises(ipf.IPFError, match="not finite"):
        ctx.transfer_geometry(**dict(args, corners=c2))
    with pytest.raises(ipf.IPFError, match="no kernel"):
        ctx.transfer_geometry(**dict(args, shape_sizes=[(5, 5)], shape_tables=np.zeros(5 * 13)))
    with pytest.raises(ipf.IPFError, match="same sizes"):
        ctx.transfer_geometry(**dict(args, shape_sizes=[(nip, nsh), (nip, nsh)], shape_tables=np.concatenate([tab, tab])))
    # no elements: the residual is f_lin
    e = ipf.DeviceContext(device=0)
    e.transfer_geometry([(nip, nsh)], tab, np.zeros(0, np.int32), np.zeros((0, 8)), np.zeros(0, np.int32),
                        np.zeros(0, np.int32), 1, 4)
    e.set_material(1.0, 1.0, 0.5, 0.5, 1.0)
    f = np.array([1.0, -2.0, 3.0, 0.5])
    e.set_rhs_linear(f)
    r = torch.zeros(4, dtype=torch.float64, device=gpu)
    e.residual(torch.zeros(4, dtype=torch.float64, device=gpu), r)
    e.synchronize()
    assert np.array_equal(r.cpu().numpy(), f)
