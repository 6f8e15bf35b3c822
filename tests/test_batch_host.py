"""Batched instances, host side (no GPU): the drawn table equals the committed
fixture, and the concatenated integrator data is a valid single data set
(offsets, colors, materials)."""
import json
import os

import numpy as np

from intpointsfem_amd import batch

HERE = os.path.dirname(os.path.abspath(__file__))

SMALL = [dict(N=5, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
         dict(N=4, p=3, phi_deg=35.0, cohesion=12.0, E=15000.0, nu=0.28, pw_final=-28.0),
         dict(N=6, p=2, phi_deg=20.0, cohesion=1.5, E=8000.0, nu=0.16, pw_final=-16.0)]


def test_draw_matches_fixture():
    with open(os.path.join(HERE, "golden", "batch64_instances.json")) as f:
        fx = json.load(f)
    assert fx["seed"] == 0
    assert batch.draw_instances(64, seed=0) == fx["instances"]
    N = np.array([i["N"] for i in fx["instances"]])
    p = np.array([i["p"] for i in fx["instances"]])
    assert N.min() >= 50 and N.max() <= 400 and set(p) <= {1, 2, 3}
    assert all(20 <= i["phi_deg"] <= 40 and 1 <= i["cohesion"] <= 20 and -30 <= i["pw_final"] <= -15
               for i in fx["instances"])


def test_concatenation_is_a_valid_data_set():
    meshes = [batch.instance_mesh(i) for i in SMALL]
    h = batch.concat_instances(meshes)
    NEQ, nc = h["neq"], h["ncolor"]
    nb = len(h["row_sizes"])
    assert h["matrix_position"][-1] == len(h["storage"])
    assert np.array_equal(np.diff(h["matrix_position"]), h["row_sizes"].astype(np.int64) * h["col_sizes"])
    assert np.array_equal(np.diff(h["row_first"]), h["row_sizes"])
    assert np.array_equal(np.diff(h["col_first"]), h["col_sizes"])
    dof = h["dof_indexes"].astype(np.int64)
    col = (h["color_indexes"].astype(np.int64) - dof) // NEQ
    assert ((h["color_indexes"] - dof) % NEQ == 0).all() and col.min() >= 0 and col.max() < nc
    # every instance's DoFs stay in its own range; no DoF twice inside a color
    for i in range(len(meshes)):
        b0, b1 = h["block_offsets"][i], h["block_offsets"][i + 1]
        d = dof[h["col_first"][b0]:h["col_first"][b1]]
        assert d.min() >= h["eq_offsets"][i] and d.max() < h["eq_offsets"][i + 1]
        assert (h["element_material"][b0:b1] == i).all()
    for c in range(nc):
        dc = dof[col == c]
        assert len(np.unique(dc)) == len(dc)
    assert h["material_table"].shape == (len(meshes), 5)
    assert nb == sum(i["N"] ** 2 for i in SMALL)


def test_geometry_concatenation_matches_blocks():
    """The device-setup input of a batch (corners, shape tables) describes the
    same data set as the host-built one: equal sizes, DoF / color indexes,
    f_lin and materials; one table per element order."""
    meshes = [batch.instance_mesh(i) for i in SMALL]
    h = batch.concat_instances(meshes)
    g = batch.concat_geometry(meshes)
    for k in ("row_sizes", "col_sizes", "dof_indexes", "color_indexes", "rhs_linear", "element_material",
              "material_table", "eq_offsets", "pt_offsets", "block_offsets"):
        assert np.array_equal(h[k], g[k]), k
    assert (h["ncolor"], h["neq"]) == (g["ncolor"], g["neq"])
    assert sorted(map(tuple, g["shape_sizes"].tolist())) == [(4, 4), (9, 9), (16, 16)]
    assert g["shape_tables"].size == sum(n * (2 * m + 3) for n, m in g["shape_sizes"])
    assert g["corners"].shape == (len(h["row_sizes"]), 8)
    assert np.array_equal(3 * g["shape_sizes"][g["elem_shape"], 0], h["row_sizes"])
