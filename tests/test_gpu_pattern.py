"""The tangent's CSR pattern built on the device (ipf_build_tangent_pattern)
against the host build of Create() / FillLIndexes
(TPZNumericalIntegrator.cpp:538-606): integer arrays, compared for equality.

  * ia, ja and the reference's per-color positions (fColorLSequence,
    fFirstColorLIndex) of StructMatrix(device=0) -- which now builds on the
    device -- equal those of StructMatrix(device=-1), the host build, on
    quarter-annulus meshes of orders 1-3, the mixed triangle / quadrilateral
    gmsh mesh and a batch of uneven instances;
  * the assembled K (ipf_assemble_tangent) is the same bits with the
    device-built positions as with the host positions uploaded through
    ipf_set_tangent_pattern;
  * 800 x 800 Q2 (84.5 M pattern entries): ia and ja equal the host build's;
    the times of both builds are printed.
"""
import os
import time

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import batch as bt
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _meshes(name):
    """The meshes of a case, built inside the test (not at collection: the
    GPU tests initialise torch before the library loads)."""
    if name in ("q1", "q2", "q3"):
        return [wb.make_mesh({"q1": (7, 5, 1, 1.2), "q2": (9, 8, 2, 1.1), "q3": (6, 5, 3, 1.3)}[name])]
    if name == "gmsh-mix":
        mix = ipf.CompMesh.from_gmsh(os.path.join(HERE, "golden", "gmsh", "wellbore_mix.msh"), 2)
        mix.set_material(5000, 0.2, 10, 0.5, 0.5)
        return [mix]
    table = [dict(N=9, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
             dict(N=6, p=3, phi_deg=35.0, cohesion=2.0, E=15000.0, nu=0.28, pw_final=-28.0),
             dict(N=7, p=2, phi_deg=30.0, cohesion=5.0, E=8000.0, nu=0.25, pw_final=-24.0)]
    return [bt.instance_mesh(i) for i in table]


def _pattern(mesh, device):
    sm = ipf.StructMatrix(mesh, device=device)
    return sm, sm.tangent_pattern()


@pytest.mark.parametrize("name", ["q1", "q2", "q3", "gmsh-mix", "batch"])
def test_device_pattern_equals_host_build(gpu, name):
    for m in _meshes(name):
        _, host = _pattern(m, -1)
        sm, dev = _pattern(m, 0)
        for k in ("ia", "ja", "color_l_sequence", "first_color_l_index"):
            assert np.array_equal(host[k], dev[k]), (name, k)
        assert host["ia"][-1] == len(host["ja"]) and (np.diff(host["ia"]) >= 0).all()


@pytest.mark.parametrize("mesh", [(9, 8, 2, 1.1), (6, 5, 3, 1.3)], ids=["q2", "q3"])
def test_assembled_K_same_bits_with_device_positions(gpu, mesh):
    import torch
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    pat = sm.tangent_pattern()  # device build; color_l_sequence filled on the host on request
    ix = sm.export_indexes()
    nnz = len(pat["ja"])
    ctx = sm.context()
    u = torch.as_tensor(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), device=gpu)
    out = {}
    for how in ("device", "host"):
        if how == "device":
            assert ctx.build_tangent_pattern() == nnz
            ia, ja = ctx.get_tangent_pattern(sm.sizes()["neq"], nnz)
            assert np.array_equal(ia, pat["ia"]) and np.array_equal(ja, pat["ja"])
        else:
            ctx.set_tangent_pattern(ix["el_color_index"], pat["color_l_sequence"], nnz)
        r = torch.zeros_like(u)
        kg = torch.zeros(nnz, dtype=torch.float64, device=gpu)
        ctx.assemble_tangent(u, r, kg)
        torch.cuda.synchronize()
        out[how] = kg.cpu().numpy()
    assert np.abs(out["device"]).max() > 0
    assert np.array_equal(out["device"], out["host"])


def test_production_size_pattern(gpu):
    m = wb.make_mesh("800x800-q2")
    t0 = time.time()
    smh = ipf.StructMatrix(m, device=-1)
    nnz_h = smh.create_tangent()
    t1 = time.time()
    smd = ipf.StructMatrix(m, device=0)
    smd.context()  # device setup outside the timing
    t2 = time.time()
    nnz_d = smd.create_tangent()
    t3 = time.time()
    print(f"800x800 Q2: {nnz_h} pattern entries; host Create {t1 - t0:.2f} s (with the host setup), "
          f"device Create {t3 - t2:.2f} s (pattern, positions, ia / ja to the host)", flush=True)
    assert nnz_d == nnz_h
    ctx = smd.context()
    neq = smd.sizes()["neq"]
    ia_d, ja_d = ctx.get_tangent_pattern(neq, nnz_d)
    full = smh.tangent_pattern()
    assert np.array_equal(full["ia"], ia_d) and np.array_equal(full["ja"], ja_d)
