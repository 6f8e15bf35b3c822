"""The cache-line floor of a colored launch (intpointsfem_amd.traffic_floor)
against a per-element loop over the same indexes on small meshes: the color
of each element from its own color indexes, the set of 128-byte lines its
DoFs fall in, the points and node counts summed element by element."""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import traffic_floor as tf
from intpointsfem_amd import wellbore as wb


@pytest.mark.parametrize("mesh", [(6, 5, 1, 1.1), (7, 4, 2, 1.2), (5, 5, 3, 1.1)])
def test_floor_matches_element_loop(mesh):
    sm = ipf.StructMatrix(wb.make_mesh(mesh), device=-1)
    ix, b, neq = sm.export_indexes(), sm.export_blocks(), sm.sizes()["neq"]
    got = tf.color_floors(sm)
    want = {}
    for e in range(sm.sizes()["nblocks"]):
        c0, c1 = b["col_first"][e], b["col_first"][e + 1]
        cols = ix["color_indexes"][c0:c1]
        c = int(cols[0]) // neq
        assert all(int(x) // neq == c for x in cols)  # one color per element
        w = want.setdefault(c, {"elements": 0, "points": 0, "lines": set(), "nodes": 0})
        w["elements"] += 1
        w["points"] += b["row_sizes"][e] // 3
        w["lines"].update(int(d) * 8 // 128 for d in ix["dof_indexes"][c0:c1])
        w["nodes"] += (c1 - c0) // 2
    assert sorted(want) == [f["color"] for f in got]
    for f in got:
        w = want[f["color"]]
        assert (f["elements"], f["points"]) == (w["elements"], w["points"])
        assert f["u_read"] == f["r_read"] == f["r_write"] == 128 * len(w["lines"])
        assert f["node_indexes"] == 4 * w["nodes"] and f["corners"] == 64 * w["elements"]
        assert f["plastic_strain"] == 32 * w["points"]
        assert f["total"] == 3 * f["u_read"] + f["plastic_strain"] + f["corners"] + f["node_indexes"]
    # the colors together touch every line of u at least once
    all_lines = np.unique(np.asarray(ix["dof_indexes"], dtype=np.int64) * 8 // 128).size
    assert sum(f["u_read"] for f in got) >= 128 * all_lines
