"""Cache-line floor of the colored residual's HBM traffic.

A color launch of the fused residual (one wavefront tile of elements of one
color per workgroup) reads, at the least, every 128-byte line of u and of r
that its elements' DoFs fall in, writes every such line of r, and streams the
points' plastic strain (4 doubles per point), the elements' corners (8
doubles) and their node indexes (one int per element node).  The elements of
one color share no node, but their DoFs share cache lines; this module counts
those lines from the integrator's own DoF and color indexes, so the floor
follows the mesh's numbering and the reference's greedy coloring exactly.

  python -m intpointsfem_amd.traffic_floor 800x800-q2
"""
import sys

import numpy as np

LINE = 128  # bytes per L2 line / HBM request


def color_floors(sm):
    """Per color: elements, points and the floor's parts in bytes."""
    ix = sm.export_indexes()
    b = sm.export_blocks()
    sz = sm.sizes()
    neq = sz["neq"]
    cf = np.asarray(b["col_first"], dtype=np.int64)
    dof = np.asarray(ix["dof_indexes"], dtype=np.int64)
    # the color of an element is the block its color indexes (dof + color * neq) fall in
    color = np.asarray(ix["color_indexes"], dtype=np.int64)[cf[:-1]] // neq
    npts = np.asarray(b["row_sizes"], dtype=np.int64) // 3
    ncols = np.diff(cf)
    out = []
    for c in range(int(color.max()) + 1):
        els = np.nonzero(color == c)[0]
        # positions of the color's DoF indexes: each element's [cf[e], cf[e + 1])
        pos = np.repeat(cf[els] - np.cumsum(np.concatenate([[0], ncols[els][:-1]])), ncols[els]) + \
            np.arange(ncols[els].sum())
        lines = np.unique(dof[pos] * 8 // LINE).size * LINE
        p = int(npts[els].sum())
        out.append({"color": c, "elements": len(els), "points": p, "u_read": lines, "r_read": lines,
                    "r_write": lines, "plastic_strain": 32 * p, "corners": 64 * len(els),
                    "node_indexes": 4 * int(ncols[els].sum() // 2)})
    for f in out:
        f["total"] = sum(v for k, v in f.items() if k not in ("color", "elements", "points"))
    return out


def main():
    from . import StructMatrix
    from . import wellbore as wb
    mesh = sys.argv[1] if len(sys.argv) > 1 else "800x800-q2"
    sm = StructMatrix(wb.make_mesh(mesh), device=-1)
    for f in color_floors(sm):
        parts = ", ".join(f"{k} {f[k] / 1e6:.1f}" for k in ("u_read", "r_read", "r_write", "plastic_strain",
                                                           "corners", "node_indexes"))
        print(f"{mesh} color {f['color']}: {f['elements']} elements, {f['points']} points; floor "
              f"{f['total'] / 1e6:.1f} MB ({parts} MB)")


if __name__ == "__main__":
    main()
