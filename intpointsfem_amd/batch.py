"""Many independent wellbore instances of uneven size, order and material in
one device context.

The instances' integrator data are concatenated: blocks, points and
equations are stacked with offsets, every element keeps its instance's color
(instances share no DoF, so the union of the colorings is a coloring), and
each element carries its instance's material id.  The device context then
runs one launch per (color, element shape) for the whole batch -- at most
4 colors x 3 shapes = 12 residual launches however many instances there are
-- with every wavefront tile inside one instance's material
(ipf_set_element_materials).  Residuals, branch flags and plastic strains of
an instance are those of the instance alone: each residual entry still
receives its contributions in color order, one per launch.

The reference solves one mesh with one material (FindMaterial(1),
TPZNumericalIntegrator.cpp:446); the batch is an extension built on the same
entry points.
"""
import math

import numpy as np

from . import DeviceContext, StructMatrix
from . import wellbore as wb

SIGMA0 = wb.SIGMA0


def draw_instances(n=64, seed=0):
    """The batched-instances table: N_r = N_theta ~ U{50..400}, p ~ U{1,2,3},
    phi ~ U(20, 40) deg with psi = phi, c ~ LogU(1, 20) MPa,
    E ~ U(2000, 20000) MPa, nu ~ U(0.15, 0.3), final wellbore pressure
    ~ U(-30, -15) MPa; numpy PCG64(seed), drawn column by column."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = rng.integers(50, 400, size=n, endpoint=True)
    p = rng.integers(1, 3, size=n, endpoint=True)
    phi = rng.uniform(20.0, 40.0, size=n)
    c = np.exp(rng.uniform(math.log(1.0), math.log(20.0), size=n))
    E = rng.uniform(2000.0, 20000.0, size=n)
    nu = rng.uniform(0.15, 0.3, size=n)
    pw = rng.uniform(-30.0, -15.0, size=n)
    return [dict(N=int(N[i]), p=int(p[i]), phi_deg=float(phi[i]), cohesion=float(c[i]), E=float(E[i]),
                 nu=float(nu[i]), pw_final=float(pw[i])) for i in range(n)]


def instance_mesh(inst, step=wb.NSTEPS, nsteps=wb.NSTEPS):
    """CompMesh of one instance: the wellbore geometry with N x N cells of
    order p, the radial grading profile of the 40-cell mesh (ratio 1.1^(40/N))
    and the instance's material; inner-wall load of load step `step`."""
    from . import CompMesh
    N, p = inst["N"], inst["p"]
    m = CompMesh(N, N, p, wb.R_IN, wb.R_OUT, 1.1 ** (40.0 / N))
    phi = math.radians(inst["phi_deg"])
    m.set_material(inst["E"], inst["nu"], inst["cohesion"], phi, phi)
    m.set_normal_stress(wb.normal_stress(step, nsteps, inst["pw_final"], SIGMA0), 0.0)
    return m


def concat_instances(meshes):
    """Host side of a batch: the instances' IrregularBlocks, DoF / color
    indexes, weights and f_lin stacked with offsets (the arrays
    ipf_transfer_data takes), the per-block material id and the material
    table.  Needs no GPU."""
    parts = []
    for m in meshes:
        sm = StructMatrix(m, device=-1)
        parts.append((m, sm.sizes(), sm.export_blocks(), sm.export_indexes()))
    n = len(parts)
    nb = np.array([p[1]["nblocks"] for p in parts], dtype=np.int64)
    neq = np.array([p[1]["neq"] for p in parts], dtype=np.int64)
    npt = np.array([p[1]["npts"] for p in parts], dtype=np.int64)
    rows = np.array([p[1]["rows"] for p in parts], dtype=np.int64)
    cols = np.array([p[1]["cols"] for p in parts], dtype=np.int64)
    nsto = np.array([p[1]["nstorage"] for p in parts], dtype=np.int64)

    def off(a):
        return np.concatenate([[0], np.cumsum(a)]).astype(np.int64)

    eq0, r0, c0, s0, pt0 = off(neq), off(rows), off(cols), off(nsto), off(npt)
    ncolor = max(p[1]["ncolor"] for p in parts)
    NEQ = int(eq0[-1])
    if r0[-1] >= 2 ** 31 or c0[-1] >= 2 ** 31 or NEQ * ncolor >= 2 ** 31:
        raise ValueError("batch too large for the int32 row / column / color indexes of the block format")
    dof = np.concatenate([p[3]["dof_indexes"].astype(np.int64) + eq0[i] for i, p in enumerate(parts)])
    # color of every column entry from the instance's fColorIndexes = dof + color * neq_i
    colr = np.concatenate([(p[3]["color_indexes"].astype(np.int64) - p[3]["dof_indexes"]) // neq[i]
                           for i, p in enumerate(parts)])
    params = []
    for m, *_ in parts:
        d = m.material()
        params.append([d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"]])
    return dict(
        row_sizes=np.concatenate([p[2]["row_sizes"] for p in parts]),
        col_sizes=np.concatenate([p[2]["col_sizes"] for p in parts]),
        matrix_position=np.concatenate([p[2]["matrix_position"][:-1] + s0[i] for i, p in enumerate(parts)]
                                       + [[s0[-1]]]),
        row_first=np.concatenate([p[2]["row_first"][:-1] + r0[i] for i, p in enumerate(parts)] + [[r0[-1]]]),
        col_first=np.concatenate([p[2]["col_first"][:-1] + c0[i] for i, p in enumerate(parts)] + [[c0[-1]]]),
        storage=np.concatenate([p[2]["storage"] for p in parts]),
        dof_indexes=dof.astype(np.int32),
        color_indexes=(dof + colr * NEQ).astype(np.int32),
        ncolor=ncolor, neq=NEQ,
        weight=np.concatenate([p[3]["weight"] for p in parts]),
        rhs_linear=np.concatenate([p[3]["rhs_linear"] for p in parts]),
        element_material=np.repeat(np.arange(n, dtype=np.int32), nb),
        material_table=np.array(params),
        eq_offsets=eq0, pt_offsets=pt0, block_offsets=off(nb))


def concat_geometry(meshes):
    """Host side of a batch for the device setup (ipf_transfer_geometry): the
    instances' element corners, one shape table per element order, the shape
    index per element, DoF / color indexes and f_lin stacked with offsets, the
    per-block material id and the material table.  No B values on the host."""
    parts = []
    for m in meshes:
        sm = StructMatrix(m, device=-1)
        parts.append((m, sm.sizes(), sm.export_indexes(), m.shape_table(), m.corners()))
    n = len(parts)
    nb = np.array([p[1]["nblocks"] for p in parts], dtype=np.int64)
    neq = np.array([p[1]["neq"] for p in parts], dtype=np.int64)
    npt = np.array([p[1]["npts"] for p in parts], dtype=np.int64)
    rows = np.array([p[1]["rows"] for p in parts], dtype=np.int64)
    cols = np.array([p[1]["cols"] for p in parts], dtype=np.int64)

    def off(a):
        return np.concatenate([[0], np.cumsum(a)]).astype(np.int64)

    eq0, r0, c0, pt0 = off(neq), off(rows), off(cols), off(npt)
    ncolor = max(p[1]["ncolor"] for p in parts)
    NEQ = int(eq0[-1])
    if r0[-1] >= 2 ** 31 or c0[-1] >= 2 ** 31 or NEQ * ncolor >= 2 ** 31:
        raise ValueError("batch too large for the int32 row / column / color indexes of the block format")
    sizes, tables, shape_of = [], [], []
    for p in parts:
        key = p[3][0]
        if key not in sizes:
            sizes.append(key)
            tables.append(p[3][1])
        shape_of.append(sizes.index(key))
    dof = np.concatenate([p[2]["dof_indexes"].astype(np.int64) + eq0[i] for i, p in enumerate(parts)])
    colr = np.concatenate([(p[2]["color_indexes"].astype(np.int64) - p[2]["dof_indexes"]) // neq[i]
                           for i, p in enumerate(parts)])
    params = []
    for m, *_ in parts:
        d = m.material()
        params.append([d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"]])
    return dict(
        shape_sizes=np.array(sizes, dtype=np.int32), shape_tables=np.concatenate(tables),
        elem_shape=np.repeat(np.array(shape_of, dtype=np.int32), nb),
        corners=np.concatenate([p[4] for p in parts]),
        dof_indexes=dof.astype(np.int32), color_indexes=(dof + colr * NEQ).astype(np.int32),
        ncolor=ncolor, neq=NEQ,
        rhs_linear=np.concatenate([p[2]["rhs_linear"] for p in parts]),
        element_material=np.repeat(np.arange(n, dtype=np.int32), nb),
        material_table=np.array(params),
        eq_offsets=eq0, pt_offsets=pt0, block_offsets=off(nb),
        row_sizes=np.repeat(np.array([3 * s[0] for s in sizes], dtype=np.int32)[shape_of], nb),
        col_sizes=np.repeat(np.array([2 * s[1] for s in sizes], dtype=np.int32)[shape_of], nb))


class Batch:
    """Independent instances in one DeviceContext (see the module docstring).

    meshes: CompMesh objects (material and boundary load set).  Attributes:
    ctx, n, neq (total), npts, eq_offsets [n+1], pt_offsets [n+1] (each
    instance's points in its reference order, stacked), rhs_linear."""

    def __init__(self, meshes, device=0, setup="device"):
        """setup "device": B and the weights computed on the GPU from the
        element corners (ipf_transfer_geometry); "host": the host-built
        blocks uploaded (ipf_transfer_data).  Same data either way."""
        if setup not in ("device", "host"):
            raise ValueError(f"setup must be 'device' or 'host', not {setup!r}")
        h = concat_geometry(meshes) if setup == "device" else concat_instances(meshes)
        self.ctx = DeviceContext(device=device)
        self._h = h
        self._setup = setup
        self.mask = None
        self._transfer()
        self.n = len(meshes)
        self.element_offsets = h["block_offsets"]
        self.neq = h["neq"]
        self.npts = int(h["pt_offsets"][-1])
        self.ncolor = h["ncolor"]
        self.eq_offsets = h["eq_offsets"]
        self.pt_offsets = h["pt_offsets"]
        self.rhs_linear = h["rhs_linear"]
        self.row_sizes = h["row_sizes"]
        self.col_sizes = h["col_sizes"]
        self.meshes = list(meshes)

    def _transfer(self, keys=None):
        h = self._h
        self.ctx.set_element_materials(h["element_material"])
        if keys is not None:
            self.ctx.set_element_order_keys(keys)
        if self._setup == "device":
            self.ctx.transfer_geometry(h["shape_sizes"], h["shape_tables"], h["elem_shape"], h["corners"],
                                       h["dof_indexes"], h["color_indexes"], h["ncolor"], h["neq"])
        else:
            self.ctx.transfer_data(h["row_sizes"], h["col_sizes"], h["matrix_position"], h["row_first"],
                                   h["col_first"], h["storage"], h["dof_indexes"], h["color_indexes"], h["ncolor"],
                                   h["neq"], h["weight"])
        self.ctx.set_material_table(h["material_table"])
        self.ctx.set_rhs_linear(h["rhs_linear"])
        if self.mask is not None:
            self.ctx.set_dirichlet_mask(self.mask)

    def set_dirichlet_mask(self, mask):
        """The batch's constrained DoFs (kept across relayout)."""
        self.mask = np.ascontiguousarray(mask, dtype=np.int8)
        self.ctx.set_dirichlet_mask(self.mask)

    def dirichlet_mask(self):
        """The instances' symmetry constraints, stacked."""
        return np.concatenate([m.dirichlet_mask() for m in self.meshes])

    def relayout(self, keys):
        """Re-transfers the batch with the elements of every (color, shape,
        material) run ordered by `keys` [elements] (ipf_set_element_order_keys:
        e.g. predicted-plastic elements last, so that whole tiles take the
        certified elastic path).  The committed state and the Dirichlet mask
        carry over; f_lin returns to the instances' own loads (rhs_linear) and
        the B formats to the library's default.  Results are bit-identical for
        any keys (within a color every DoF has one element)."""
        st = self.ctx.get_state()
        self._transfer(keys)
        self.ctx.set_state(st[0], st[1])

    def newton(self, u, rtol=1e-10, maxit=80, cg_rtol=1e-12, cg_maxit=40000):
        """Newton over the batch with per-instance convergence
        (ipf_newton_segmented over the instances' equation ranges)."""
        return self.ctx.newton_segmented(u, self.eq_offsets, rtol=rtol, maxit=maxit, cg_rtol=cg_rtol,
                                         cg_maxit=cg_maxit)
