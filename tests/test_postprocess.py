"""Post-processing output on the CPU: the VTK files hold the mesh, the nodal
displacement and the weighted element averages of the point fields (checked
against averages computed here element by element); the integration-point
file holds every point's values; plot_over_line selects and orders the points
near a segment; polar() inverts a stress built from known polar components."""
import math
import os

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import postprocess as pp

HERE = os.path.dirname(os.path.abspath(__file__))


def _parse_vtk(path):
    with open(path) as f:
        lines = [l.strip() for l in f]
    out = {}
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if not tok:
            i += 1
            continue
        if tok[0] == "POINTS":
            n = int(tok[1])
            out["points"] = np.array([[float(v) for v in l.split()] for l in lines[i + 1:i + 1 + n]])
            i += 1 + n
        elif tok[0] == "CELLS":
            n = int(tok[1])
            out["cells"] = [[int(v) for v in l.split()] for l in lines[i + 1:i + 1 + n]]
            i += 1 + n
        elif tok[0] == "CELL_TYPES":
            n = int(tok[1])
            out["types"] = np.array([int(l) for l in lines[i + 1:i + 1 + n]])
            i += 1 + n
        elif tok[0] in ("POINT_DATA", "CELL_DATA"):
            out["n_" + tok[0]] = int(tok[1])
            i += 1
        elif tok[0] == "VECTORS":
            n = out["n_POINT_DATA"]
            out[tok[1]] = np.array([[float(v) for v in l.split()] for l in lines[i + 1:i + 1 + n]])
            i += 1 + n
        elif tok[0] == "SCALARS":
            n = out.get("n_CELL_DATA", out.get("n_POINT_DATA"))
            out[tok[1]] = np.array([float(l) for l in lines[i + 2:i + 2 + n]])
            i += 2 + n
        elif tok[0] == "TENSORS":
            n = out.get("n_CELL_DATA", out.get("n_POINT_DATA"))
            out[tok[1]] = np.array([[float(v) for v in l.split()] for l in lines[i + 1:i + 1 + n]])
            i += 1 + n
        else:
            i += 1
    return out


def _meshes():
    m1 = ipf.CompMesh(3, 2, 2, 0.1, 4.0, 1.2)
    m2 = ipf.CompMesh.from_gmsh(os.path.join(HERE, "golden", "gmsh", "wellbore_mix.msh"), 1)
    for m in (m1, m2):
        m.set_material(2000.0, 0.2, 5.0, 0.35, 0.35)
    return [m1, m2]


@pytest.mark.parametrize("k", [0, 1], ids=["quarter-annulus-q2", "mixed-p1"])
def test_vtk_files(tmp_path, k):
    m = _meshes()[k]
    sm = ipf.StructMatrix(m, device=-1)
    ix = sm.export_indexes()
    npts = sm.sizes()["npts"]
    rng = np.random.default_rng(k)
    fields = dict(xy=m.point_coords(), strain=rng.normal(size=(npts, 3)), stress=rng.normal(size=(npts, 4)),
                  plastic_strain=rng.normal(size=(npts, 4)), failure_type=rng.integers(0, 2, npts),
                  weight=ix["weight"])
    u = rng.normal(size=m.info()["neq"])
    pp.write_vtk(str(tmp_path / "m.vtk"), m, u, fields)
    v = _parse_vtk(str(tmp_path / "m.vtk"))
    xy = m.nodes()
    assert np.array_equal(v["points"][:, :2], xy)
    ptr, conn = m.connectivity()
    t = m.element_types()
    assert len(v["cells"]) == len(t) and np.array_equal(v["types"], np.where(t == 1, 5, 9))
    for e in range(len(t)):
        nv = 3 if t[e] == 1 else 4
        assert v["cells"][e] == [nv] + list(conn[ptr[e]:ptr[e] + nv])
    np.testing.assert_array_equal(v["Displacement"][:, :2], u.reshape(-1, 2))
    # element averages, element by element
    n = pp.points_per_element(m)
    off = np.concatenate([[0], np.cumsum(n)])
    w = ix["weight"]
    for e in range(len(t)):
        sl = slice(off[e], off[e + 1])
        st = (fields["stress"][sl] * w[sl, None]).sum(axis=0) / w[sl].sum()
        np.testing.assert_allclose(v["Stress"][e][[0, 1, 4, 8]], st, rtol=1e-13, atol=1e-15)
        ft = (fields["failure_type"][sl] * w[sl]).sum() / w[sl].sum()
        assert abs(v["FailureType"][e] - ft) < 1e-14
        es = (fields["strain"][sl] * w[sl, None]).sum(axis=0) / w[sl].sum()
        np.testing.assert_allclose(v["Strain"][e][[0, 1, 4]], [es[0], es[1] / 2, es[2]], rtol=1e-13, atol=1e-15)
    pp.write_points_vtk(str(tmp_path / "p.vtk"), fields)
    q = _parse_vtk(str(tmp_path / "p.vtk"))
    np.testing.assert_array_equal(q["points"][:, :2], fields["xy"])
    np.testing.assert_array_equal(q["FailureType"], fields["failure_type"])
    np.testing.assert_array_equal(q["StrainPlastic"][:, [0, 4, 8]], fields["plastic_strain"][:, [0, 2, 3]])


def test_plot_over_line_and_polar():
    rng = np.random.default_rng(0)
    xy = rng.uniform(-1, 1, (2000, 2))
    idx, s = pp.plot_over_line(xy, (0.0, 0.0), (1.0, 1.0), 0.05)
    d = np.abs(xy[:, 0] - xy[:, 1]) / math.sqrt(2)
    t = (xy[:, 0] + xy[:, 1]) / math.sqrt(2)
    want = np.nonzero((d <= 0.05) & (t >= -0.05) & (t <= math.sqrt(2) + 0.05))[0]
    assert set(idx) == set(want) and (np.diff(s) >= 0).all()
    np.testing.assert_allclose(s, t[idx], rtol=0, atol=1e-15)
    # polar components round trip
    th = rng.uniform(0, 2 * math.pi, 50)
    r = rng.uniform(0.1, 4, 50)
    pxy = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    srr, stt, srt = rng.normal(size=(3, 50))
    c, sn = np.cos(th), np.sin(th)
    sxx = srr * c * c + stt * sn * sn - 2 * srt * sn * c
    syy = srr * sn * sn + stt * c * c + 2 * srt * sn * c
    sxy = (srr - stt) * sn * c + srt * (c * c - sn * sn)
    a, b, g = pp.polar(np.stack([sxx, sxy, syy, 0 * sxx], axis=1), pxy)
    np.testing.assert_allclose(a, srr, atol=1e-12)
    np.testing.assert_allclose(b, stt, atol=1e-12)
    np.testing.assert_allclose(g, srt, atol=1e-12)
