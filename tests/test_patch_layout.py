"""The patch layout of IPF_ASSEMBLY_PATCH (ipf_patch_plan_host, the host
planner the device build uses), checked on the CPU against the meshes'
connectivity:

  * every element sits in exactly one slot, every integration point in
    exactly one patch point; tiles hold one material;
  * each slot's local node indexes point at its own nodes in its patch's
    node list, which is sorted and duplicate-free and fits the kernel's LDS
    bound (patch_max_nodes);
  * each patch node's contribution list holds exactly the (slot, node)
    pairs of the patch that reference it, ascending, decodable within the
    patch;
  * a patch node is owned (written by the patch) exactly when every element
    using it lies in the patch; the fix-up lists of the shared nodes hold
    exactly their partial slots; orphan DoFs are those of no element;
  * the assembly the layout encodes -- owned nodes from their patch's sums,
    shared nodes from the fix-up over partials -- reproduces a plain
    scatter-add (np.add.at) of random element forces to 1e-13.
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import batch as bt
from intpointsfem_amd import wellbore as wb

SHAPES = {(4, 8): 0, (9, 18): 1, (16, 32): 2, (3, 6): 3, (6, 12): 4, (12, 20): 5}
NIP = [4, 9, 16, 3, 6, 12]
CE = [8, 18, 32, 6, 12, 20]


def et(s):
    return 64 // NIP[s]


def elements(meshes):
    """shape, material, first point, centroid, node x-DoFs of stacked meshes (instance = material)"""
    shape, mat, fp, cx, cy, nodes, neq0, pt0 = [], [], [], [], [], [], 0, 0
    for i, m in enumerate(meshes):
        sm = ipf.StructMatrix(m, device=-1)
        b = sm.export_blocks()
        ix = sm.export_indexes()
        s = sm.sizes()
        corners = m.corners().reshape(-1, 4, 2)
        for e in range(s["nblocks"]):
            sh = SHAPES[(b["row_sizes"][e] // 3, b["col_sizes"][e])]
            shape.append(sh)
            mat.append(i)
            fp.append(pt0 + b["row_first"][e] // 3)
            c = corners[e].mean(axis=0)
            cx.append(c[0])
            cy.append(c[1])
            d = ix["dof_indexes"][b["col_first"][e]:b["col_first"][e + 1]]
            assert (d[1::2] == d[0::2] + 1).all()
            nodes.extend(int(x) + neq0 for x in d[0::2])
        neq0 += s["neq"]
        pt0 += s["npts"]
    return (np.array(shape), np.array(mat), np.array(fp, dtype=np.int64), np.array(cx), np.array(cy),
            np.array(nodes, dtype=np.int64), neq0, pt0)


def check(meshes, waves=4):
    shape, mat, fp, cx, cy, nodes, neq, npts = elements(meshes)
    n = len(shape)
    noff = np.concatenate([[0], np.cumsum([CE[s] // 2 for s in shape])])
    L = ipf.patch_plan(shape, mat, fp, cx, cy, nodes, neq, waves=waves)
    uses = np.zeros(neq, dtype=np.int64)
    np.add.at(uses, nodes, 1)
    seen_elem = np.zeros(n, dtype=np.int64)
    seen_pt = np.zeros(npts, dtype=np.int64)
    rng = np.random.default_rng(0)
    force = [rng.standard_normal(CE[s]) for s in shape]  # element forces, (x, y) per node
    r_ref = np.zeros(neq)
    for e in range(n):
        np.add.at(r_ref, nodes[noff[e]:noff[e + 1]], force[e][0::2])
        np.add.at(r_ref, nodes[noff[e]:noff[e + 1]] + 1, force[e][1::2])
    r = np.full(neq, np.nan)
    partial = np.full((max(L["npartial"], 1), 2), np.nan)
    writer = np.full(max(L["npartial"], 1), -1, dtype=np.int64)  # the patch that writes each partial slot
    for (sh, npch, patch0, tile0, slot0, pt0, ln0) in L["ranges"]:
        NW, ET, NSH, TP = waves, et(sh), CE[sh] // 2, et(sh) * NIP[sh]
        for p in range(npch):
            P = patch0 + p
            n0, n1 = L["node_off"][P], L["node_off"][P + 1]
            pn = L["pnode"][n0:n1]
            assert len(pn) <= NW * ET * NSH
            assert (np.diff(pn) > 0).all()
            mats = set()
            es = {}
            for w in range(NW):
                t = tile0 + p * NW + w
                nv = L["tmeta"][t] & 255
                mats.add(L["tmeta"][t] >> 8)
                for q in range(ET):
                    s_ = (tile0 + p * NW + w) * ET + q - tile0 * ET + slot0
                    e = L["slot_elem"][s_]
                    if q >= nv:
                        assert e == -1
                        continue
                    assert e >= 0 and shape[e] == sh and mat[e] == (L["tmeta"][t] >> 8)
                    seen_elem[e] += 1
                    ln = L["lnode"][ln0 + (s_ - slot0) * NSH: ln0 + (s_ - slot0 + 1) * NSH].astype(np.int64)
                    assert (ln < len(pn)).all()
                    assert np.array_equal(pn[ln], nodes[noff[e]:noff[e + 1]])
                    es[w * ET + q] = force[e]
                for lane in range(TP):
                    q = pt0 + (p * NW + w) * TP + lane
                    m_ = L["pt_map"][q]
                    if lane // NIP[sh] < nv:
                        e = L["slot_elem"][(p * NW + w) * ET + lane // NIP[sh] + slot0]
                        assert m_ == fp[e] + lane % NIP[sh]
                        seen_pt[m_] += 1
                    else:
                        assert m_ == -1
            assert len(mats) == 1
            # contributions: exactly the (slot, node) references, ascending
            refs = {}
            for sl, f in es.items():
                s_ = (tile0 + p * NW) * ET - tile0 * ET + slot0 + sl
                for a, l in enumerate(L["lnode"][ln0 + (s_ - slot0) * NSH: ln0 + (s_ - slot0 + 1) * NSH]):
                    refs.setdefault(int(l), []).append(sl * NSH + a)
            for k in range(n0, n1):
                c = L["contrib"][L["cbeg"][k]:L["cbeg"][k + 1]].astype(np.int64)
                assert list(c) == refs.get(k - n0, [])
                fx = sum(es[v // NSH][2 * (v % NSH)] for v in c)
                fy = sum(es[v // NSH][2 * (v % NSH) + 1] for v in c)
                d = L["pnode"][k]
                owned = len(c) == uses[d]
                assert (L["pslot"][k] < 0) == owned
                if owned:
                    assert np.isnan(r[d])
                    r[d], r[d + 1] = fx, fy
                else:
                    partial[L["pslot"][k]] = (fx, fy)
                    writer[L["pslot"][k]] = P
    assert (seen_elem == 1).all() and (seen_pt == 1).all()
    slots_seen = np.zeros(max(L["npartial"], 1), dtype=np.int64)
    # a shared node's partials are consecutive slots, in patch order, nodes by DoF
    assert np.array_equal(L["fix_slot"], np.arange(len(L["fix_slot"])))
    assert (np.diff(L["fix_dof"]) > 0).all()
    for i, d in enumerate(L["fix_dof"]):
        sl = L["fix_slot"][L["fix_ptr"][i]:L["fix_ptr"][i + 1]]
        assert len(sl) >= 2 and (np.diff(writer[sl]) > 0).all()
        slots_seen[sl] += 1
        assert np.isnan(r[d])
        r[d] = partial[sl, 0].sum()
        r[d + 1] = partial[sl, 1].sum()
    assert (slots_seen[:L["npartial"]] == 1).all()
    covered = np.zeros(neq, dtype=bool)
    covered[nodes] = covered[nodes + 1] = True
    assert np.array_equal(np.sort(L["orphan"]), np.nonzero(~covered)[0])
    r[L["orphan"]] = 0.0
    assert not np.isnan(r).any()
    np.testing.assert_allclose(r, r_ref, rtol=0, atol=1e-13 * np.abs(r_ref).max())
    shared_frac = len(L["fix_dof"]) / max(1, np.count_nonzero(covered) // 2)
    print(f"{n} elements in {len(L['node_off']) - 1} patches, {len(L['fix_dof'])} shared nodes ({shared_frac:.2f})")
    return L


@pytest.mark.parametrize("waves", [2, 4, 8])
@pytest.mark.parametrize("mesh", [(30, 20, 1, 1.1), (24, 22, 2, 1.1), (13, 11, 3, 1.2)],
                         ids=lambda t: f"{t[0]}x{t[1]}-q{t[2]}")
def test_patch_layout_quarter_annulus(mesh, waves):
    check([wb.make_mesh(mesh)], waves)


def test_patch_layout_rejects_other_wave_counts():
    m = wb.make_mesh((6, 5, 2, 1.1))
    with pytest.raises(ipf.IPFError, match="2, 4 or 8"):
        check([m], waves=3)


def test_patch_layout_mixed_mesh_and_instances():
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    mix = ipf.CompMesh.from_gmsh(os.path.join(here, "golden", "gmsh", "wellbore_mix.msh"), 2)
    mix.set_material(5000, 0.2, 10, 0.5, 0.5)
    table = [dict(N=9, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
             dict(N=6, p=3, phi_deg=35.0, cohesion=2.0, E=15000.0, nu=0.28, pw_final=-28.0)]
    for waves in (2, 4, 8):
        check([mix] + [bt.instance_mesh(i) for i in table], waves)
