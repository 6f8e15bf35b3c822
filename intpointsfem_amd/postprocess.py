"""Post-processing of a solution: the fields the reference's PostProcess
writes (main.cpp:289-324: FailureType, Displacement, Stress, Strain,
StrainPlastic) and the values along a line that its RK verification plots
(mathematica_notebooks/RK_Verification).

The reference projects the integration-point fields onto an L2 space per
element (TPZPostProcAnalysis::TransferSolution) before writing VTK.  Here the
integration-point values are written exactly (one VTK vertex per point,
``write_points_vtk``) and, on the mesh, as element averages weighted by the
integration weights -- the L2 projection onto piecewise constants -- next to
the nodal displacement (``write_vtk``).  Files are legacy ASCII VTK.

Tensor components: stress and plastic strain as (xx, xy, yy, zz); strain as
(xx, xy, yy) with xy the engineering shear (B u); VTK tensors take the
tensorial shear (xy / 2 for strains).
"""
import numpy as np

TRIANGLE_POINTS = {1: 3, 2: 6, 3: 12}


def points_per_element(mesh):
    """Integration points of each element (reference point order: element by element)."""
    p = mesh.porder
    t = mesh.element_types()
    return np.where(t == 1, TRIANGLE_POINTS[p], (p + 1) ** 2).astype(np.int64)


def point_fields(sm, u, device="cuda:0"):
    """Fields at the integration points for displacement u (host array or
    device tensor) and the committed state: dict with xy [npts, 2], strain
    [npts, 3], stress [npts, 4], plastic_strain [npts, 4], failure_type
    [npts] (0 elastic, 1 plastic: the committed mtype), branch [npts] and
    weight [npts].  Stress is w sigma / w from ComputeSigma (state not
    changed); sigma_zz = lambda (e_xx + e_yy) + (lambda + 2 mu) e_zz of the
    elastic strain e = strain - plastic strain (plane strain)."""
    import torch
    ctx = sm.context()
    s = sm.sizes()
    npts = s["npts"]
    uu = u if isinstance(u, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(u, dtype=np.float64), device=device)
    strain = torch.zeros(3 * npts, dtype=torch.float64, device=uu.device)
    ws = torch.zeros(3 * npts, dtype=torch.float64, device=uu.device)
    ctx.set_update_mem(False)
    ctx.multiply(uu, strain)
    ctx.compute_sigma(strain, ws)
    ctx.synchronize()
    ix = sm.export_indexes()
    w = ix["weight"]
    eps = strain.cpu().numpy().reshape(-1, 3)
    sig3 = ws.cpu().numpy().reshape(-1, 3) / w[:, None]
    ep6, _, mtype, branch = ctx.get_state()
    ep6 = ep6.reshape(-1, 6)  # XX, XY, XZ, YY, YZ, ZZ
    ep = np.stack([ep6[:, 0], ep6[:, 1], ep6[:, 3], ep6[:, 5]], axis=1)
    d = sm.mesh.material()
    lam, mu = d["lambda"], d["mu"]
    exx, eyy, ezz = eps[:, 0] - ep[:, 0], eps[:, 2] - ep[:, 2], 0.0 - ep[:, 3]
    szz = lam * (exx + eyy) + (lam + 2 * mu) * ezz
    stress = np.stack([sig3[:, 0], sig3[:, 1], sig3[:, 2], szz], axis=1)
    return dict(xy=sm.mesh.point_coords(), strain=eps, stress=stress, plastic_strain=ep,
                failure_type=np.asarray(mtype), branch=np.asarray(branch), weight=w)


def element_averages(mesh, values, weight):
    """Weighted element averages of per-point values [npts, ...] (the L2
    projection onto piecewise constants)."""
    n = points_per_element(mesh)
    start = np.concatenate([[0], np.cumsum(n)[:-1]])
    v = np.asarray(values, dtype=np.float64)
    w = np.asarray(weight, dtype=np.float64)
    wv = v * (w[:, None] if v.ndim > 1 else w)
    return np.add.reduceat(wv, start, axis=0) / (np.add.reduceat(w, start)[:, None] if v.ndim > 1
                                                 else np.add.reduceat(w, start))


def _tensor(c, shear_half):
    """[n, 4] (xx, xy, yy, zz) -> [n, 3, 3] rows of a symmetric tensor."""
    xy = c[:, 1] / 2 if shear_half else c[:, 1]
    z = np.zeros(len(c))
    return np.stack([c[:, 0], xy, z, xy, c[:, 2], z, z, z, c[:, 3]], axis=1)


def write_vtk(path, mesh, u, fields):
    """The mesh (corner vertices of every element; all mesh nodes as points)
    with POINT_DATA Displacement and CELL_DATA FailureType, Stress, Strain,
    StrainPlastic as element averages of ``fields`` (from point_fields)."""
    xy = mesh.nodes()
    u = np.asarray(u, dtype=np.float64).reshape(-1, 2)
    ptr, conn = mesh.connectivity()
    types = mesh.element_types()
    nel = len(types)
    nv = np.where(types == 1, 3, 4)
    w = fields["weight"]
    avg = {k: element_averages(mesh, fields[k], w) for k in ("stress", "strain", "plastic_strain")}
    ft = element_averages(mesh, fields["failure_type"].astype(np.float64), w)
    strain4 = np.concatenate([avg["strain"], np.zeros((nel, 1))], axis=1)
    lines = ["# vtk DataFile Version 3.0", "intpointsfem_amd solution", "ASCII", "DATASET UNSTRUCTURED_GRID",
             f"POINTS {len(xy)} double"]
    lines += [f"{x:.17g} {y:.17g} 0" for x, y in xy]
    lines.append(f"CELLS {nel} {int((nv + 1).sum())}")
    for e in range(nel):
        lines.append(" ".join(str(int(v)) for v in [nv[e]] + list(conn[ptr[e]:ptr[e] + nv[e]])))
    lines.append(f"CELL_TYPES {nel}")
    lines += ["5" if t == 1 else "9" for t in types]
    lines += [f"POINT_DATA {len(xy)}", "VECTORS Displacement double"]
    lines += [f"{a:.17g} {b:.17g} 0" for a, b in u]
    lines += [f"CELL_DATA {nel}", "SCALARS FailureType double 1", "LOOKUP_TABLE default"]
    lines += [f"{v:.17g}" for v in ft]
    for name, comp, half in (("Stress", avg["stress"], False), ("Strain", strain4, True),
                             ("StrainPlastic", avg["plastic_strain"], True)):
        lines.append(f"TENSORS {name} double")
        lines += [" ".join(f"{v:.17g}" for v in row) for row in _tensor(comp, half)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_points_vtk(path, fields):
    """The integration points as VTK vertices with their exact values."""
    xy = fields["xy"]
    n = len(xy)
    lines = ["# vtk DataFile Version 3.0", "intpointsfem_amd integration points", "ASCII",
             "DATASET UNSTRUCTURED_GRID", f"POINTS {n} double"]
    lines += [f"{x:.17g} {y:.17g} 0" for x, y in xy]
    lines += [f"CELLS {n} {2 * n}"] + [f"1 {i}" for i in range(n)]
    lines += [f"CELL_TYPES {n}"] + ["1"] * n
    lines += [f"POINT_DATA {n}", "SCALARS FailureType int 1", "LOOKUP_TABLE default"]
    lines += [str(int(v)) for v in fields["failure_type"]]
    strain4 = np.concatenate([fields["strain"], np.zeros((n, 1))], axis=1)
    for name, comp, half in (("Stress", fields["stress"], False), ("Strain", strain4, True),
                             ("StrainPlastic", fields["plastic_strain"], True)):
        lines.append(f"TENSORS {name} double")
        lines += [" ".join(f"{v:.17g}" for v in row) for row in _tensor(comp, half)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def plot_over_line(xy, a, b, tol):
    """Indexes of the points within `tol` of the segment a -> b, ordered along
    it, and their parameter s (distance from a along the segment)."""
    xy = np.asarray(xy, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = b - a
    L = np.hypot(*d)
    t = (xy - a) @ d / L
    dist = np.abs((xy[:, 0] - a[0]) * d[1] - (xy[:, 1] - a[1]) * d[0]) / L
    sel = np.nonzero((dist <= tol) & (t >= -tol) & (t <= L + tol))[0]
    o = np.argsort(t[sel], kind="stable")
    return sel[o], t[sel][o]


def polar(stress, xy):
    """sigma_rr, sigma_tt, sigma_rt of (xx, xy, yy, ...) stresses at points xy."""
    r = np.hypot(xy[:, 0], xy[:, 1])
    c, s = xy[:, 0] / r, xy[:, 1] / r
    sxx, sxy, syy = stress[:, 0], stress[:, 1], stress[:, 2]
    srr = sxx * c * c + syy * s * s + 2 * sxy * s * c
    stt = sxx * s * s + syy * c * c - 2 * sxy * s * c
    srt = (syy - sxx) * s * c + sxy * (c * c - s * s)
    return srr, stt, srt


__all__ = ["point_fields", "element_averages", "write_vtk", "write_points_vtk", "plot_over_line", "polar",
           "points_per_element"]
