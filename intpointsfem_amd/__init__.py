"""MI355X (gfx950) integration-point residual path for 2D plane-strain
Mohr-Coulomb elastoplasticity, behind IntPointsFEM's operator API.

The compute lives in ``libintpointsfem.so`` (host C++ + hand-written HIP
kernels, C ABI in ``include/intpointsfem.h`` and
``include/intpointsfem_frontend.h``).  This module is a thin ctypes binding:

* :class:`CompMesh` -- the structured quarter-annulus mesh stand-in
  (TPZCompMesh) with material and boundary tractions;
* :class:`StructMatrix` -- TPZElastoPlasticIntPointsStructMatrix: builds the
  integrator data, uploads it and evaluates residuals on the GPU;
* :class:`DeviceContext` -- the raw device context (``ipf_ctx``): residual,
  Multiply / ComputeSigma / MultiplyTranspose steps and state access on
  device pointers (e.g. ``torch.Tensor.data_ptr()``).

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "library_path", "CompMesh", "StructMatrix", "DeviceContext", "MLP", "IPFError", "build"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libintpointsfem.so")

IPF_OK = 0
BRANCH_NAMES = ("elastic", "main", "right_edge", "left_edge", "apex")
B_FORMATS = ("dense-fp64", "compact-fp64", "dense-fp32", "compact-fp32", "geometry-fp64", "geometry-fp32")
RESIDUAL_MODES = ("exact", "certified-elastic")
ASSEMBLIES = ("colored", "patch")


class IPFError(RuntimeError):
    pass


def library_path():
    return _LIB_PATH


def build(verbose=False):
    from . import buildlib
    return buildlib.build(verbose=verbose)


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_bp = ctypes.POINTER(ctypes.c_byte)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_int = ctypes.c_int
# ipf_newton_monitor: (user, iteration, rel, res_norm, cg_iterations)
NEWTON_MONITOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int)
_MONITORS = {}

_lib = None


def lib():
    """Loads libintpointsfem.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise IPFError(f"{_LIB_PATH} is missing: build it with `python -m intpointsfem_amd.buildlib` "
                       "(there is no CPU fallback)")
    # torch brings its own HIP runtime: loaded after this library's (the
    # system one), the two do not share the GPU -- the library's contexts
    # then find no device once torch has initialised.  Loading torch first
    # makes the library bind to torch's runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(_LIB_PATH)
    sig = {
        # device layer
        "ipf_ctx_create": (_int, [_int, ctypes.POINTER(_vp)]),
        "ipf_ctx_destroy": (_int, [_vp]),
        "ipf_last_error": (ctypes.c_char_p, [_vp]),
        "ipf_global_last_error": (ctypes.c_char_p, []),
        "ipf_set_stream": (_int, [_vp, _vp]),
        "ipf_get_stream": (_vp, [_vp]),
        "ipf_synchronize": (_int, [_vp]),
        "ipf_transfer_data": (_int, [_vp, _i64, _ip, _ip, _lp, _ip, _ip, _dp, _ip, _ip, _i64, _i64, _dp]),
        "ipf_transfer_geometry": (_int, [_vp, _int, _ip, _dp, _i64, _ip, _dp, _ip, _ip, _i64, _i64]),
        "ipf_set_material": (_int, [_vp, _d, _d, _d, _d, _d]),
        "ipf_set_material_table": (_int, [_vp, _int, _dp]),
        "ipf_set_element_materials": (_int, [_vp, _i64, _ip]),
        "ipf_set_element_order_keys": (_int, [_vp, _i64, _ip]),
        "ipf_set_rhs_linear": (_int, [_vp, _dp]),
        "ipf_set_update_mem": (_int, [_vp, _int]),
        "ipf_info": (_int, [_vp, _lp]),
        "ipf_residual": (_int, [_vp, _vp, _vp]),
        "ipf_residual_host": (_int, [_vp, _dp, _dp]),
        "ipf_residual_timed": (_int, [_vp, _vp, _vp, _dp, _lp, _int]),
        "ipf_multiply_vector": (_int, [_vp, _vp, _vp, _int]),
        "ipf_multiply": (_int, [_vp, _vp, _vp]),
        "ipf_compute_sigma": (_int, [_vp, _vp, _vp]),
        "ipf_compute_sigma_dep": (_int, [_vp, _vp, _vp, _vp]),
        "ipf_multiply_transpose": (_int, [_vp, _vp, _vp]),
        "ipf_set_b_format": (_int, [_vp, _int, _int]),
        "ipf_get_b_format": (_int, [_vp, _ip, _ip]),
        "ipf_set_dirichlet_mask": (_int, [_vp, _bp]),
        "ipf_residual_tangent": (_int, [_vp, _vp, _vp]),
        "ipf_tangent_apply": (_int, [_vp, _vp, _vp]),
        "ipf_set_tangent_pattern": (_int, [_vp, _ip, _lp, _i64]),
        "ipf_build_tangent_pattern": (_int, [_vp, _lp]),
        "ipf_get_tangent_pattern": (_int, [_vp, _lp, _lp]),
        "ipf_assemble_tangent": (_int, [_vp, _vp, _vp, _vp]),
        "ipf_cg_solve": (_int, [_vp, _vp, _vp, _d, _int, _ip, _dp]),
        "ipf_newton": (_int, [_vp, _vp, _d, _int, _d, _int, _dp]),
        "ipf_set_cg_graph": (_int, [_vp, _int]),
        "ipf_set_newton_graph": (_int, [_vp, _int]),
        "ipf_set_newton_atol": (_int, [_vp, _d]),
        "ipf_set_newton_reference_stop": (_int, [_vp, _d]),
        "ipf_set_newton_thomas": (_int, [_vp, _int]),
        "ipf_set_newton_line_search": (_int, [_vp, _int]),
        "ipf_set_newton_monitor": (_int, [_vp, NEWTON_MONITOR, _vp]),
        "ipf_set_residual_mode": (_int, [_vp, _int]),
        "ipf_get_residual_mode": (_int, [_vp, _ip]),
        "ipf_certify_elastic_host": (_int, [_i64, _dp, _dp, _d, _d, _d, _d, _d, _ip, _dp]),
        "ipf_set_residual_assembly": (_int, [_vp, _int]),
        "ipf_get_residual_assembly": (_int, [_vp, _ip]),
        "ipf_patch_info": (_int, [_vp, _lp]),
        "ipf_patch_plan_host": (_int, [_i64, _ip, _ip, _lp, _dp, _dp, _ip, _i64, _int, _lp, _lp, _ip, _lp, _vp,
                                       _lp, _ip, _ip, _ip, _vp, _lp, _ip, _ip, _ip, _ip]),
        "ipf_set_patch_waves": (_int, [_vp, _int]),
        "ipf_get_patch_waves": (_int, [_vp, _ip]),
        "ipf_autotune_assembly": (_int, [_vp, _vp, _int, _dp]),
        "ipf_newton_segmented": (_int, [_vp, _vp, _int, _lp, _d, _int, _d, _int, _dp, _dp, _ip]),
        "ipf_autotune": (_int, [_vp, _vp, _int, _int, _dp]),
        "ipf_get_state": (_int, [_vp, _dp, _dp, _ip, _ip]),
        "ipf_point_features": (_int, [_vp, _vp, _vp, _vp, _vp]),
        "ipf_branch_order": (_int, [_vp, _vp, _vp]),
        "ipf_compute_sigma_ordered": (_int, [_vp, _vp, _vp, _vp]),
        "ipf_branch_order_noalloc": (_int, [_vp, _vp, _vp]),
        "ipf_branch_order_noalloc_max_points": (_i64, []),
        "ipf_mlp_create": (_int, [_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
        "ipf_mlp_destroy": (_int, [_vp]),
        "ipf_mlp_set_stream": (_int, [_vp, _vp]),
        "ipf_mlp_num_params": (_int, [_ip, _ip]),
        "ipf_mlp_set_params": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
        "ipf_mlp_get_params": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
        "ipf_mlp_gradient": (_int, [_vp, _vp, _vp, _vp, _i64, ctypes.POINTER(ctypes.c_float), _dp]),
        "ipf_mlp_train_step": (_int, [_vp, _vp, _vp, _vp, _i64, ctypes.c_float, _dp]),
        "ipf_mlp_predict": (_int, [_vp, _vp, _i64, _vp, _vp]),
        "ipf_set_state": (_int, [_vp, _dp, _dp]),
        # front-end
        "ipf_cmesh_quarter_annulus": (_int, [_int, _int, _int, _d, _d, _d, ctypes.POINTER(_vp)]),
        "ipf_cmesh_destroy": (_int, [_vp]),
        "ipf_cmesh_set_material": (_int, [_vp, _d, _d, _d, _d, _d]),
        "ipf_cmesh_set_normal_stress": (_int, [_vp, _d, _d]),
        "ipf_cmesh_info": (_int, [_vp, _lp]),
        "ipf_cmesh_material": (_int, [_vp, _dp]),
        "ipf_cmesh_nodes": (_int, [_vp, _dp]),
        "ipf_cmesh_connectivity": (_int, [_vp, _lp, _lp]),
        "ipf_cmesh_dirichlet_mask": (_int, [_vp, _bp]),
        "ipf_cmesh_set_update_mem": (_int, [_vp, _int]),
        "ipf_cmesh_point_coords": (_int, [_vp, _dp]),
        "ipf_cmesh_shape_table": (_int, [_vp, _int, _ip, _dp]),
        "ipf_cmesh_element_types": (_int, [_vp, _bp]),
        "ipf_cmesh_read_gmsh": (_int, [ctypes.c_char_p, _int, ctypes.POINTER(_vp)]),
        "ipf_cmesh_gmsh_counts": (_int, [_vp, _lp]),
        "ipf_cmesh_corners": (_int, [_vp, _dp]),
        "ipf_sm_create": (_int, [_vp, _int, ctypes.POINTER(_vp)]),
        "ipf_sm_destroy": (_int, [_vp]),
        "ipf_sm_context": (_vp, [_vp]),
        "ipf_sm_sizes": (_int, [_vp, _lp]),
        "ipf_sm_export_blocks": (_int, [_vp, _ip, _ip, _lp, _ip, _ip, _dp]),
        "ipf_sm_export_indexes": (_int, [_vp, _ip, _ip, _ip, _lp, _dp, _dp]),
        "ipf_sm_assemble_rhs": (_int, [_vp, _dp, _dp]),
        "ipf_sm_create_tangent": (_int, [_vp, _lp]),
        "ipf_sm_export_tangent_pattern": (_int, [_vp, _lp, _lp, _lp, _lp]),
        "ipf_sm_assemble_tangent": (_int, [_vp, _dp, _dp, _dp]),
        "ipf_sm_clone": (_int, [_vp, ctypes.POINTER(_vp)]),
        "ipf_sm_create_assemble": (_int, [_vp, _dp, _dp, _lp]),
        "ipf_sm_tangent_values": (_int, [_vp, _dp]),
        "ipf_sm_export_csr": (_int, [_vp, _ip, _ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def patch_plan(shape, mat, first_point, cx, cy, node_xdof, neq, waves=4):
    """The patch layout planned on the host (ipf_patch_plan_host), patches of
    `waves` wavefront tiles: a dict of numpy arrays."""
    n = len(shape)
    sh = np.ascontiguousarray(shape, dtype=np.int32)
    mt = np.ascontiguousarray(mat, dtype=np.int32)
    fp = np.ascontiguousarray(first_point, dtype=np.int64)
    x = np.ascontiguousarray(cx, dtype=np.float64)
    y = np.ascontiguousarray(cy, dtype=np.float64)
    nd = np.ascontiguousarray(node_xdof, dtype=np.int32)
    sizes = np.zeros(12, dtype=np.int64)
    args = [n, _p(sh, _ip), _p(mt, _ip), _p(fp, _lp), _p(x, _dp), _p(y, _dp), _p(nd, _ip), int(neq), int(waves),
            _p(sizes, _lp)]
    _check(lib().ipf_patch_plan_host(*args, *([None] * 14)), what="ipf_patch_plan_host")
    nr, npch, nt, ns, nl, npn, nc, npt, nsh, nfs, nor, npart = (int(v) for v in sizes)
    out = dict(ranges=np.zeros(7 * nr, np.int64), tmeta=np.zeros(nt, np.int32), slot_elem=np.zeros(ns, np.int64),
               lnode=np.zeros(nl, np.uint16), node_off=np.zeros(npch + 1, np.int64), pnode=np.zeros(npn, np.int32),
               pslot=np.zeros(npn, np.int32), cbeg=np.zeros(npn + 1, np.int32), contrib=np.zeros(nc, np.uint16),
               pt_map=np.zeros(npt, np.int64), fix_dof=np.zeros(nsh, np.int32), fix_ptr=np.zeros(nsh + 1, np.int32),
               fix_slot=np.zeros(nfs, np.int32), orphan=np.zeros(nor, np.int32))
    ptrs = [_p(out["ranges"], _lp), _p(out["tmeta"], _ip), _p(out["slot_elem"], _lp), out["lnode"].ctypes.data,
            _p(out["node_off"], _lp), _p(out["pnode"], _ip), _p(out["pslot"], _ip), _p(out["cbeg"], _ip),
            out["contrib"].ctypes.data, _p(out["pt_map"], _lp), _p(out["fix_dof"], _ip), _p(out["fix_ptr"], _ip),
            _p(out["fix_slot"], _ip), _p(out["orphan"], _ip)]
    _check(lib().ipf_patch_plan_host(*args, *ptrs), what="ipf_patch_plan_host")
    out["ranges"] = out["ranges"].reshape(-1, 7)
    out["npartial"] = npart
    return out


def certify_elastic(strain3, plastic_strain4, lam, mu, phi, psi, cohesion):
    """Host evaluation of the certified-elastic test (ipf_certify_elastic_host):
    returns (certified [n] bool, trial stress [n, 3])."""
    e = np.ascontiguousarray(strain3, dtype=np.float64).reshape(-1, 3)
    ep = np.ascontiguousarray(plastic_strain4, dtype=np.float64).reshape(-1, 4)
    n = e.shape[0]
    cert = np.zeros(n, dtype=np.int32)
    sig = np.zeros((n, 3))
    _check(lib().ipf_certify_elastic_host(n, _p(e, _dp), _p(ep, _dp), lam, mu, phi, psi, cohesion, _p(cert, _ip),
                                          _p(sig, _dp)), what="ipf_certify_elastic_host")
    return cert.astype(bool), sig


def exported_symbols():
    """Names the C ABI headers declare (checked against the library by the tests)."""
    import re
    names = []
    inc = os.path.join(os.path.dirname(_HERE), "include")
    for h in ("intpointsfem.h", "intpointsfem_frontend.h"):
        with open(os.path.join(inc, h)) as f:
            text = f.read()
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(ipf_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def _err(ctx=None):
    L = lib()
    msg = L.ipf_last_error(ctx) if ctx else L.ipf_global_last_error()
    return (msg or b"").decode(errors="replace")


def _check(rc, ctx=None, what=""):
    if rc != IPF_OK:
        raise IPFError(f"{what} failed ({rc}): {_err(ctx)}")


def _p(a, t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


def _ptr(x):
    """Device pointer of a torch tensor or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


class CompMesh:
    """The TPZCompMesh stand-in: a structured quarter annulus with Q_p
    Lagrange elements (the constructor) or a Gmsh 4.1 mesh of triangles and
    quadrilaterals (``CompMesh.from_gmsh``)."""

    def __init__(self, nr, ntheta, porder, r_in=0.1, r_out=4.0, ratio=1.0):
        L = lib()
        h = _vp()
        _check(L.ipf_cmesh_quarter_annulus(nr, ntheta, porder, r_in, r_out, ratio, ctypes.byref(h)),
               what="ipf_cmesh_quarter_annulus")
        self._h = h
        self.nr, self.ntheta, self.porder = nr, ntheta, porder
        self.r_in, self.r_out, self.ratio = r_in, r_out, ratio

    @classmethod
    def from_gmsh(cls, path, porder):
        """Elements of order porder on the triangles and quadrilaterals of a
        Gmsh 4.1 ASCII file (ipf_cmesh_read_gmsh)."""
        h = _vp()
        _check(lib().ipf_cmesh_read_gmsh(os.fsencode(path), int(porder), ctypes.byref(h)), what="ipf_cmesh_read_gmsh")
        m = cls.__new__(cls)
        m._h = h
        m.nr = m.ntheta = None
        m.porder = int(porder)
        m.r_in = m.r_out = m.ratio = None
        m.path = os.fspath(path)
        return m

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ipf_cmesh_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_material(self, E, nu, cohesion, phi, psi):
        _check(lib().ipf_cmesh_set_material(self._h, E, nu, cohesion, phi, psi), what="ipf_cmesh_set_material")

    def set_normal_stress(self, inner, outer=0.0):
        _check(lib().ipf_cmesh_set_normal_stress(self._h, inner, outer), what="ipf_cmesh_set_normal_stress")

    def set_update_mem(self, flag):
        _check(lib().ipf_cmesh_set_update_mem(self._h, 1 if flag else 0), what="ipf_cmesh_set_update_mem")

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        _check(lib().ipf_cmesh_info(self._h, _p(out, _lp)), what="ipf_cmesh_info")
        keys = ("nelem", "nnodes", "neq", "porder", "nr", "ntheta", "npts_per_element", "nshape")
        return dict(zip(keys, (int(v) for v in out)))

    def material(self):
        out = np.zeros(5)
        _check(lib().ipf_cmesh_material(self._h, _p(out, _dp)), what="ipf_cmesh_material")
        return dict(zip(("lambda", "mu", "phi", "psi", "cohesion"), (float(v) for v in out)))

    def nodes(self):
        n = self.info()["nnodes"]
        xy = np.zeros(2 * n)
        _check(lib().ipf_cmesh_nodes(self._h, _p(xy, _dp)), what="ipf_cmesh_nodes")
        return xy.reshape(n, 2)

    def point_coords(self):
        """Integration-point coordinates [npts, 2] in the reference point order."""
        inf = self.info()
        if inf["npts_per_element"] >= 0:
            npts = inf["nelem"] * inf["npts_per_element"]
        else:  # triangles and quadrilaterals
            p = inf["porder"]
            t = self.element_types()
            npts = int((t == 1).sum()) * {1: 3, 2: 6, 3: 12}[p] + int((t == 0).sum()) * (p + 1) ** 2
        xy = np.zeros(2 * npts)
        _check(lib().ipf_cmesh_point_coords(self._h, _p(xy, _dp)), what="ipf_cmesh_point_coords")
        return xy.reshape(-1, 2)

    def connectivity(self):
        inf = self.info()
        ptr = np.zeros(inf["nelem"] + 1, dtype=np.int64)
        conn = np.zeros(inf["nelem"] * (inf["porder"] + 1) ** 2, dtype=np.int64)  # the largest element
        _check(lib().ipf_cmesh_connectivity(self._h, _p(ptr, _lp), _p(conn, _lp)), what="ipf_cmesh_connectivity")
        return ptr, conn[:ptr[-1]]

    def shape_table(self, etype=-1):
        """((n_ip, n_sh), table) -- the shape table of ipf_transfer_geometry
        for element type etype (0 quadrilateral, 1 triangle, -1 the mesh's
        only type)."""
        sz = np.zeros(2, dtype=np.int32)
        _check(lib().ipf_cmesh_shape_table(self._h, int(etype), _p(sz, _ip), None), what="ipf_cmesh_shape_table")
        nip, nsh = int(sz[0]), int(sz[1])
        tab = np.zeros(nip * (2 * nsh + 3))
        _check(lib().ipf_cmesh_shape_table(self._h, int(etype), _p(sz, _ip), _p(tab, _dp)),
               what="ipf_cmesh_shape_table")
        return (nip, nsh), tab

    def element_types(self):
        """[nelem] int8: 0 quadrilateral, 1 triangle."""
        t = np.zeros(self.info()["nelem"], dtype=np.int8)
        _check(lib().ipf_cmesh_element_types(self._h, _p(t, _bp)), what="ipf_cmesh_element_types")
        return t

    def gmsh_counts(self):
        """What the Gmsh file held: nodes, elements, points, lines, triangles, quads."""
        out = np.zeros(6, dtype=np.int64)
        _check(lib().ipf_cmesh_gmsh_counts(self._h, _p(out, _lp)), what="ipf_cmesh_gmsh_counts")
        return dict(zip(("nodes", "elements", "points", "lines", "triangles", "quads"), (int(v) for v in out)))

    def corners(self):
        """[nelem, 8] corners of each element's bilinear map (x0, y0, ..., x3, y3)."""
        c = np.zeros(8 * self.info()["nelem"])
        _check(lib().ipf_cmesh_corners(self._h, _p(c, _dp)), what="ipf_cmesh_corners")
        return c.reshape(-1, 8)

    def dirichlet_mask(self):
        m = np.zeros(self.info()["neq"], dtype=np.int8)
        _check(lib().ipf_cmesh_dirichlet_mask(self._h, _p(m, _bp)), what="ipf_cmesh_dirichlet_mask")
        return m.astype(bool)


class DeviceContext:
    """Non-owning view of an ``ipf_ctx`` (owned by a StructMatrix) or an owning one."""

    def __init__(self, handle=None, device=0, owner=None):
        L = lib()
        self._own = handle is None
        if handle is None:
            h = _vp()
            _check(L.ipf_ctx_create(device, ctypes.byref(h)), what="ipf_ctx_create")
            handle = h
        self._h = handle if isinstance(handle, _vp) else _vp(handle)
        self._owner = owner

    def __del__(self):
        if getattr(self, "_own", False) and self._h and _lib is not None:
            _lib.ipf_ctx_destroy(self._h)
            _MONITORS.pop(int(self._h.value or 0), None)
            self._h = None

    @property
    def handle(self):
        return self._h

    def _c(self, rc, what):
        _check(rc, self._h, what)

    def transfer_data(self, row_sizes, col_sizes, matrix_position, row_first, col_first, storage, dof_indexes,
                      color_indexes, ncolor, neq, weight):
        a = [np.ascontiguousarray(x, dtype=t) for x, t in
             ((row_sizes, np.int32), (col_sizes, np.int32), (matrix_position, np.int64), (row_first, np.int32),
              (col_first, np.int32), (storage, np.float64), (dof_indexes, np.int32), (color_indexes, np.int32),
              (weight, np.float64))]
        self._c(lib().ipf_transfer_data(self._h, a[0].size, _p(a[0], _ip), _p(a[1], _ip), _p(a[2], _lp),
                                        _p(a[3], _ip), _p(a[4], _ip), _p(a[5], _dp), _p(a[6], _ip),
                                        _p(a[7], _ip), int(ncolor), int(neq), _p(a[8], _dp)), "ipf_transfer_data")

    def transfer_geometry(self, shape_sizes, shape_tables, elem_shape, corners, dof_indexes, color_indexes, ncolor,
                          neq):
        """Device setup from geometry (ipf_transfer_geometry): shape_sizes
        [(n_ip, n_sh), ...], shape_tables the tables concatenated, elem_shape
        [nblocks] shape index per element, corners [nblocks, 8]."""
        sz = np.ascontiguousarray(shape_sizes, dtype=np.int32).reshape(-1, 2)
        tab = np.ascontiguousarray(shape_tables, dtype=np.float64)
        es = np.ascontiguousarray(elem_shape, dtype=np.int32)
        cr = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(dof_indexes, dtype=np.int32)
        c = np.ascontiguousarray(color_indexes, dtype=np.int32)
        if tab.size != sum(int(n) * (2 * int(m) + 3) for n, m in sz) or cr.size != 8 * es.size:
            raise IPFError("transfer_geometry: table or corner array of the wrong size")
        self._c(lib().ipf_transfer_geometry(self._h, sz.shape[0], _p(sz, _ip), _p(tab, _dp), es.size, _p(es, _ip),
                                            _p(cr, _dp), _p(d, _ip), _p(c, _ip), int(ncolor), int(neq)),
                "ipf_transfer_geometry")

    def set_material(self, lam, mu, phi, psi, cohesion):
        self._c(lib().ipf_set_material(self._h, lam, mu, phi, psi, cohesion), "ipf_set_material")

    def set_material_table(self, params):
        """params [nmat, 5]: (lambda, mu, phi, psi, cohesion) per material (radians)."""
        q = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 5)
        self._c(lib().ipf_set_material_table(self._h, q.shape[0], _p(q, _dp)), "ipf_set_material_table")

    def set_element_materials(self, ids):
        """Material id per block; applies to the next transfer_data."""
        a = np.ascontiguousarray(ids, dtype=np.int32)
        self._c(lib().ipf_set_element_materials(self._h, a.size, _p(a, _ip)), "ipf_set_element_materials")

    def set_element_order_keys(self, keys):
        """Order key per block inside each (color, shape, material) run; applies to the next transfer."""
        a = np.ascontiguousarray(keys, dtype=np.int32)
        self._c(lib().ipf_set_element_order_keys(self._h, a.size, _p(a, _ip)), "ipf_set_element_order_keys")

    def set_rhs_linear(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._c(lib().ipf_set_rhs_linear(self._h, _p(f, _dp)), "ipf_set_rhs_linear")

    def set_update_mem(self, flag):
        self._c(lib().ipf_set_update_mem(self._h, 1 if flag else 0), "ipf_set_update_mem")

    def set_stream(self, stream_ptr):
        self._c(lib().ipf_set_stream(self._h, stream_ptr), "ipf_set_stream")

    def stream(self):
        return lib().ipf_get_stream(self._h)

    def synchronize(self):
        self._c(lib().ipf_synchronize(self._h), "ipf_synchronize")

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        self._c(lib().ipf_info(self._h, _p(out, _lp)), "ipf_info")
        keys = ("nblocks", "npts", "rows", "cols", "neq", "ncolor", "nstorage", "device_bytes")
        return dict(zip(keys, (int(v) for v in out)))

    # ---- compute on device pointers (torch tensors or raw ints)
    def residual(self, u, r):
        self._c(lib().ipf_residual(self._h, _ptr(u), _ptr(r)), "ipf_residual")

    def residual_timed(self, u, r, max_launches=64):
        """Residual with per-launch HIP events; returns (ms[n], points[n])."""
        ms = np.zeros(max_launches)
        pts = np.zeros(max_launches, dtype=np.int64)
        n = lib().ipf_residual_timed(self._h, _ptr(u), _ptr(r), _p(ms, _dp), _p(pts, _lp), max_launches)
        if n < 0:
            _check(n, self._h, "ipf_residual_timed")
        return ms[:n], pts[:n]

    def residual_host(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        self._c(lib().ipf_residual_host(self._h, _p(u, _dp), _p(r, _dp)), "ipf_residual_host")
        return r

    def multiply_vector(self, x, y, opt):
        self._c(lib().ipf_multiply_vector(self._h, _ptr(x), _ptr(y), int(opt)), "ipf_multiply_vector")

    def multiply(self, u, strain):
        self._c(lib().ipf_multiply(self._h, _ptr(u), _ptr(strain)), "ipf_multiply")

    def compute_sigma(self, strain, wsigma):
        self._c(lib().ipf_compute_sigma(self._h, _ptr(strain), _ptr(wsigma)), "ipf_compute_sigma")

    def compute_sigma_dep(self, strain, wsigma, wdep):
        self._c(lib().ipf_compute_sigma_dep(self._h, _ptr(strain), _ptr(wsigma), _ptr(wdep)),
                "ipf_compute_sigma_dep")

    def multiply_transpose(self, wsigma, res):
        self._c(lib().ipf_multiply_transpose(self._h, _ptr(wsigma), _ptr(res)), "ipf_multiply_transpose")

    def set_b_format(self, fmt, tangent=None):
        """B format of the fused residual and (``tangent``, default: the same)
        of the tangent product: 0 dense FP64, 1 compact FP64, 2 dense FP32,
        3 compact FP32, 4 geometric FP64 (after transfer_geometry), or the
        names in B_FORMATS."""
        def code(f):
            return B_FORMATS.index(f) if isinstance(f, str) else int(f)
        fr = code(fmt)
        ft = fr if tangent is None else code(tangent)
        self._c(lib().ipf_set_b_format(self._h, fr, ft), "ipf_set_b_format")

    def b_format(self):
        """(residual format, tangent format) names in use."""
        fr = ctypes.c_int()
        ft = ctypes.c_int()
        self._c(lib().ipf_get_b_format(self._h, ctypes.byref(fr), ctypes.byref(ft)), "ipf_get_b_format")
        return B_FORMATS[fr.value], B_FORMATS[ft.value]

    # ---- tangent / solver
    def set_dirichlet_mask(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.int8)
        self._c(lib().ipf_set_dirichlet_mask(self._h, _p(m, _bp)), "ipf_set_dirichlet_mask")

    def residual_tangent(self, u, r):
        self._c(lib().ipf_residual_tangent(self._h, _ptr(u), _ptr(r)), "ipf_residual_tangent")

    def set_tangent_pattern(self, el_color_index, color_l_sequence, nnz):
        """fElColorIndex, fColorLSequence and the CSR value count (FillLIndexes)."""
        e = np.ascontiguousarray(el_color_index, dtype=np.int32)
        l = np.ascontiguousarray(color_l_sequence, dtype=np.int64)
        self._c(lib().ipf_set_tangent_pattern(self._h, _p(e, _ip), _p(l, _lp), int(nnz)), "ipf_set_tangent_pattern")

    def build_tangent_pattern(self):
        """The tangent's CSR pattern and the element entries' positions built
        on the device; returns the number of pattern entries."""
        n = np.zeros(1, dtype=np.int64)
        self._c(lib().ipf_build_tangent_pattern(self._h, _p(n, _lp)), "ipf_build_tangent_pattern")
        return int(n[0])

    def get_tangent_pattern(self, neq, nnz):
        """(ia [neq + 1], ja [nnz]) of the pattern build_tangent_pattern built."""
        ia = np.zeros(neq + 1, dtype=np.int64)
        ja = np.zeros(nnz, dtype=np.int64)
        self._c(lib().ipf_get_tangent_pattern(self._h, _p(ia, _lp), _p(ja, _lp)), "ipf_get_tangent_pattern")
        return ia, ja

    def assemble_tangent(self, u, r, kg):
        """KAssembly: residual into r and the CSR values of K into kg (device tensors)."""
        self._c(lib().ipf_assemble_tangent(self._h, _ptr(u), _ptr(r), _ptr(kg)), "ipf_assemble_tangent")

    def tangent_apply(self, v, y):
        """y = K v from the last residual_tangent (ipf_tangent_apply); y is
        overwritten whatever it held."""
        self._c(lib().ipf_tangent_apply(self._h, _ptr(v), _ptr(y)), "ipf_tangent_apply")

    def autotune(self, u, reps=20, allow_fp32=False):
        """Selects the fastest B format for the residual and (after
        residual_tangent) the tangent product (ipf_autotune); returns
        {"residual": {format: ms}, "tangent": {format: ms}, "selected": (r, t)}."""
        t = np.zeros(2 * len(B_FORMATS))
        self._c(lib().ipf_autotune(self._h, _ptr(u), int(reps), 1 if allow_fp32 else 0, _p(t, _dp)), "ipf_autotune")
        n = len(B_FORMATS)
        return {"residual": {B_FORMATS[k]: t[k] for k in range(n) if t[k] >= 0},
                "tangent": {B_FORMATS[k]: t[n + k] for k in range(n) if t[n + k] >= 0},
                "selected": self.b_format()}

    def newton_segmented(self, u, seg_first, rtol=1e-10, maxit=50, cg_rtol=1e-10, cg_maxit=20000):
        """ipf_newton_segmented: every segment [seg_first[k], seg_first[k+1])
        of the equations converges on its own criterion.  Returns the stats
        dict of newton() plus seg_rel and seg_iters arrays."""
        sf = np.ascontiguousarray(seg_first, dtype=np.int64)
        n = sf.size - 1
        st = np.zeros(8)
        rel = np.zeros(n)
        its = np.zeros(n, dtype=np.int32)
        self._c(lib().ipf_newton_segmented(self._h, _ptr(u), n, _p(sf, _lp), rtol, maxit, cg_rtol, cg_maxit,
                                           _p(st, _dp), _p(rel, _dp), _p(its, _ip)), "ipf_newton_segmented")
        keys = ("iterations", "converged", "rel_residual", "r0", "rn", "cg_iterations", "worst_cg_rel", "reserved")
        out = dict(zip(keys, (float(v) for v in st)))
        out["seg_rel"] = rel
        out["seg_iters"] = its
        return out

    def set_cg_graph(self, enable):
        """CG iterations from a captured HIP graph (default) or launched one by one."""
        self._c(lib().ipf_set_cg_graph(self._h, 1 if enable else 0), "ipf_set_cg_graph")

    def set_newton_graph(self, enable):
        """Each Newton iteration up to its first host read (tangent residual,
        residual norm, CG start and first CG block) as one captured HIP graph
        replay (default) or launched eagerly; the same bits either way."""
        self._c(lib().ipf_set_newton_graph(self._h, 1 if enable else 0), "ipf_set_newton_graph")

    def cg_solve(self, b, x, rtol=1e-10, maxit=10000):
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        self._c(lib().ipf_cg_solve(self._h, _ptr(b), _ptr(x), rtol, maxit, ctypes.byref(it), ctypes.byref(rel)),
                "ipf_cg_solve")
        return it.value, rel.value

    def set_residual_mode(self, mode):
        """'exact' (0) or 'certified-elastic' (1): see ipf_set_residual_mode."""
        code = RESIDUAL_MODES.index(mode) if isinstance(mode, str) else int(mode)
        self._c(lib().ipf_set_residual_mode(self._h, code), "ipf_set_residual_mode")

    def residual_mode(self):
        m = ctypes.c_int()
        self._c(lib().ipf_get_residual_mode(self._h, ctypes.byref(m)), "ipf_get_residual_mode")
        return RESIDUAL_MODES[m.value]

    def set_residual_assembly(self, assembly):
        """'colored' (0) or 'patch' (1): see ipf_set_residual_assembly."""
        code = ASSEMBLIES.index(assembly) if isinstance(assembly, str) else int(assembly)
        self._c(lib().ipf_set_residual_assembly(self._h, code), "ipf_set_residual_assembly")

    def residual_assembly(self):
        m = ctypes.c_int()
        self._c(lib().ipf_get_residual_assembly(self._h, ctypes.byref(m)), "ipf_get_residual_assembly")
        return ASSEMBLIES[m.value]

    def patch_info(self):
        out = np.zeros(8, dtype=np.int64)
        self._c(lib().ipf_patch_info(self._h, _p(out, _lp)), "ipf_patch_info")
        keys = ("built", "patches", "patch_nodes", "partials", "shared_nodes", "orphan_dofs", "points", "shapes")
        return dict(zip(keys, (int(v) for v in out)))

    def set_newton_line_search(self, max_halvings):
        """Backtracking line search of newton(): up to max_halvings step
        halvings per iteration (0: full Newton steps)."""
        self._c(lib().ipf_set_newton_line_search(self._h, int(max_halvings)), "ipf_set_newton_line_search")

    def set_patch_waves(self, waves):
        """Wavefront tiles per patch of the patch assembly: 2, 4 or 8."""
        self._c(lib().ipf_set_patch_waves(self._h, int(waves)), "ipf_set_patch_waves")

    def patch_waves(self):
        m = ctypes.c_int()
        self._c(lib().ipf_get_patch_waves(self._h, ctypes.byref(m)), "ipf_get_patch_waves")
        return m.value

    def autotune_assembly(self, u, reps=10):
        """Times the non-committing residual with the colored assembly and
        with patches of 2, 4 and 8 wavefront tiles (where available) and
        keeps the fastest.  Returns {candidate: ms} and the selection."""
        t = np.zeros(4)
        self._c(lib().ipf_autotune_assembly(self._h, _ptr(u), int(reps), _p(t, _dp)), "ipf_autotune_assembly")
        names = ("colored", "patch-2", "patch-4", "patch-8")
        times = {n: float(v) for n, v in zip(names, t) if v >= 0}
        sel = self.residual_assembly()
        return {"times_ms": times, "selected": sel if sel == "colored" else f"patch-{self.patch_waves()}"}

    def set_newton_atol(self, atol):
        self._c(lib().ipf_set_newton_atol(self._h, float(atol)), "ipf_set_newton_atol")

    def set_newton_thomas(self, enable):
        """The reference's modified Thomas acceleration of the Newton update
        (main.cpp:194-224); off by default."""
        self._c(lib().ipf_set_newton_thomas(self._h, 1 if enable else 0), "ipf_set_newton_thomas")

    def set_newton_reference_stop(self, tolerance):
        """Newton stops as the reference's loop does (main.cpp:239): after an
        update, ||r|| < tolerance and ||du|| < tolerance; 0 restores the
        relative test."""
        self._c(lib().ipf_set_newton_reference_stop(self._h, float(tolerance)), "ipf_set_newton_reference_stop")

    def set_newton_monitor(self, fn):
        """fn(iteration, rel, res_norm, cg_iterations) after every residual
        norm of newton / newton_segmented; None turns it off."""
        # the callback object lives as long as the context may call it: in a
        # table keyed by the context handle (several wrappers may share one)
        key = int(self._h.value or 0)
        cb = NEWTON_MONITOR() if fn is None else NEWTON_MONITOR(lambda _u, it, rel, rn, cg: fn(it, rel, rn, cg))
        self._c(lib().ipf_set_newton_monitor(self._h, cb, None), "ipf_set_newton_monitor")
        if fn is None:
            _MONITORS.pop(key, None)
        else:
            _MONITORS[key] = cb

    def newton(self, u, rtol=1e-10, maxit=50, cg_rtol=1e-10, cg_maxit=20000):
        """Newton for the current load; commits the state. Returns a stats dict."""
        st = np.zeros(8)
        self._c(lib().ipf_newton(self._h, _ptr(u), rtol, maxit, cg_rtol, cg_maxit, _p(st, _dp)), "ipf_newton")
        keys = ("iterations", "converged", "rel_residual", "r0", "r", "cg_iterations", "cg_worst_relres",
                "line_search_halvings")
        return dict(zip(keys, (float(v) for v in st[:8])))

    def point_features(self, u, feat, branch, target):
        """Training pairs of the exact return map at the trial state of u
        (device tensors: feat float32 [npts, 8], branch int32 [npts], target float32 [npts])."""
        self._c(lib().ipf_point_features(self._h, _ptr(u), _ptr(feat), _ptr(branch), _ptr(target)),
                "ipf_point_features")

    def branch_order(self, branch, order):
        self._c(lib().ipf_branch_order(self._h, _ptr(branch), _ptr(order)), "ipf_branch_order")

    def branch_order_noalloc(self, branch, order):
        self._c(lib().ipf_branch_order_noalloc(self._h, _ptr(branch), _ptr(order)), "ipf_branch_order_noalloc")

    def compute_sigma_ordered(self, strain, wsigma, order):
        self._c(lib().ipf_compute_sigma_ordered(self._h, _ptr(strain), _ptr(wsigma), _ptr(order)),
                "ipf_compute_sigma_ordered")

    def get_state(self):
        n = self.info()["npts"]
        ep = np.zeros(6 * n)
        al = np.zeros(n)
        mt = np.zeros(n, dtype=np.int32)
        br = np.zeros(n, dtype=np.int32)
        self._c(lib().ipf_get_state(self._h, _p(ep, _dp), _p(al, _dp), _p(mt, _ip), _p(br, _ip)), "ipf_get_state")
        return ep, al, mt, br

    def set_state(self, plastic_strain6, alpha=None):
        ep = np.ascontiguousarray(plastic_strain6, dtype=np.float64).reshape(-1)
        al = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        self._c(lib().ipf_set_state(self._h, _p(ep, _dp), _p(al, _dp)), "ipf_set_state")


class MLP:
    """The learned return-branch predictor (``ipf_mlp``): 7 -> 64 -> 64 -> 6,
    ReLU, BF16 matrix-core operands, FP32 accumulation, Adam."""

    def __init__(self, device=0, seed=0):
        h = _vp()
        _check(lib().ipf_mlp_create(device, seed, ctypes.byref(h)), what="ipf_mlp_create")
        self._h = h
        a = ctypes.c_int()
        b = ctypes.c_int()
        lib().ipf_mlp_num_params(ctypes.byref(a), ctypes.byref(b))
        self.npar, self.ntrainable = a.value, b.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ipf_mlp_destroy(h)
            self._h = None

    def set_stream(self, stream_ptr):
        _check(lib().ipf_mlp_set_stream(self._h, stream_ptr), what="ipf_mlp_set_stream")

    def params(self):
        p = np.zeros(self.npar, dtype=np.float32)
        _check(lib().ipf_mlp_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               what="ipf_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().ipf_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               what="ipf_mlp_set_params")

    def gradient(self, feat, branch, target, n=None):
        n = len(branch) if n is None else n
        g = np.zeros(self.npar, dtype=np.float32)
        loss = np.zeros(1)
        _check(lib().ipf_mlp_gradient(self._h, _ptr(feat), _ptr(branch), _ptr(target), int(n),
                                      g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _p(loss, _dp)),
               what="ipf_mlp_gradient")
        return float(loss[0]), g

    def train_step(self, feat, branch, target, lr, n=None, want_loss=True):
        n = len(branch) if n is None else n
        loss = np.zeros(1)
        _check(lib().ipf_mlp_train_step(self._h, _ptr(feat), _ptr(branch), _ptr(target), int(n), float(lr),
                                        _p(loss, _dp) if want_loss else None), what="ipf_mlp_train_step")
        return float(loss[0]) if want_loss else None

    def predict(self, feat, branch=None, out=None, n=None):
        n = (feat.shape[0] if hasattr(feat, "shape") else 0) if n is None else n
        _check(lib().ipf_mlp_predict(self._h, _ptr(feat), int(n), _ptr(branch), _ptr(out)), what="ipf_mlp_predict")


class StructMatrix:
    """TPZElastoPlasticIntPointsStructMatrix over a CompMesh.

    ``device >= 0`` uploads to that GPU; ``device < 0`` builds the host data
    only (exports work without a GPU)."""

    def __init__(self, mesh, device=0, _handle=None):
        if _handle is None:
            h = _vp()
            _check(lib().ipf_sm_create(mesh.handle, device, ctypes.byref(h)), what="ipf_sm_create")
        else:
            h = _handle
        self._h = h
        self.mesh = mesh
        self.device = device

    def clone(self):
        """Clone(): a structural matrix sharing this one's device context."""
        h = _vp()
        _check(lib().ipf_sm_clone(self._h, ctypes.byref(h)), what="ipf_sm_clone")
        return StructMatrix(self.mesh, self.device, _handle=h)

    def create_assemble(self, u):
        """CreateAssemble(rhs, gui) at u: returns (rhs, kg) of a new matrix."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        n = np.zeros(1, dtype=np.int64)
        _check(lib().ipf_sm_create_assemble(self._h, _p(u, _dp), _p(r, _dp), _p(n, _lp)),
               what="ipf_sm_create_assemble")
        kg = np.zeros(int(n[0]))
        _check(lib().ipf_sm_tangent_values(self._h, _p(kg, _dp)), what="ipf_sm_tangent_values")
        return r, kg

    def export_csr(self):
        """CSRVectors(): (row_ptr [rows+1], col_ind [stored values]) of the block matrix."""
        s = self.sizes()
        rp = np.zeros(s["rows"] + 1, np.int32)
        ci = np.zeros(s["nstorage"], np.int32)
        _check(lib().ipf_sm_export_csr(self._h, _p(rp, _ip), _p(ci, _ip)), what="ipf_sm_export_csr")
        return rp, ci

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ipf_sm_destroy(h)
            self._h = None

    def sizes(self):
        out = np.zeros(8, dtype=np.int64)
        _check(lib().ipf_sm_sizes(self._h, _p(out, _lp)), what="ipf_sm_sizes")
        keys = ("nblocks", "npts", "rows", "cols", "neq", "ncolor", "nstorage")
        return dict(zip(keys, (int(v) for v in out[:7])))

    def context(self):
        h = lib().ipf_sm_context(self._h)
        if not h:
            raise IPFError("structural matrix has no device context (created with device < 0)")
        return DeviceContext(handle=h, owner=self)

    def export_blocks(self):
        s = self.sizes()
        nb = s["nblocks"]
        rs = np.zeros(nb, np.int32)
        cs = np.zeros(nb, np.int32)
        mp = np.zeros(nb + 1, np.int64)
        rf = np.zeros(nb + 1, np.int32)
        cf = np.zeros(nb + 1, np.int32)
        st = np.zeros(s["nstorage"])
        _check(lib().ipf_sm_export_blocks(self._h, _p(rs, _ip), _p(cs, _ip), _p(mp, _lp), _p(rf, _ip), _p(cf, _ip),
                                          _p(st, _dp)), what="ipf_sm_export_blocks")
        return dict(row_sizes=rs, col_sizes=cs, matrix_position=mp, row_first=rf, col_first=cf, storage=st)

    def export_indexes(self):
        s = self.sizes()
        dof = np.zeros(s["cols"], np.int32)
        col = np.zeros(s["cols"], np.int32)
        elc = np.zeros(s["nblocks"], np.int32)
        fci = np.zeros(s["ncolor"] + 1, np.int64)
        w = np.zeros(s["npts"])
        fl = np.zeros(s["neq"])
        _check(lib().ipf_sm_export_indexes(self._h, _p(dof, _ip), _p(col, _ip), _p(elc, _ip), _p(fci, _lp),
                                           _p(w, _dp), _p(fl, _dp)), what="ipf_sm_export_indexes")
        return dict(dof_indexes=dof, color_indexes=col, el_color_index=elc, first_color_index=fci, weight=w,
                    rhs_linear=fl)

    def assemble_rhs(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        _check(lib().ipf_sm_assemble_rhs(self._h, _p(u, _dp), _p(r, _dp)), what="ipf_sm_assemble_rhs")
        return r

    def create_tangent(self):
        """Create(): builds the symmetric CSR pattern and FillLIndexes; returns nnz."""
        n = np.zeros(1, dtype=np.int64)
        _check(lib().ipf_sm_create_tangent(self._h, _p(n, _lp)), what="ipf_sm_create_tangent")
        return int(n[0])

    def tangent_pattern(self):
        """dict(ia, ja, color_l_sequence, first_color_l_index) of the pattern Create() built."""
        nnz = self.create_tangent()
        s = self.sizes()
        C = self.export_blocks()["col_sizes"].astype(np.int64)
        ia = np.zeros(s["neq"] + 1, dtype=np.int64)
        ja = np.zeros(nnz, dtype=np.int64)
        cl = np.zeros(int(((C * C + C) // 2).sum()), dtype=np.int64)
        fl = np.zeros(s["ncolor"] + 1, dtype=np.int64)
        _check(lib().ipf_sm_export_tangent_pattern(self._h, _p(ia, _lp), _p(ja, _lp), _p(cl, _lp), _p(fl, _lp)),
               what="ipf_sm_export_tangent_pattern")
        return {"ia": ia, "ja": ja, "color_l_sequence": cl, "first_color_l_index": fl}

    def assemble_tangent(self, u):
        """Assemble(K, rhs) at u (host arrays): returns (rhs, kg) with kg the
        upper-triangle CSR values in the pattern of tangent_pattern()."""
        nnz = self.create_tangent()
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros_like(u)
        kg = np.zeros(nnz)
        _check(lib().ipf_sm_assemble_tangent(self._h, _p(u, _dp), _p(r, _dp), _p(kg, _dp)),
               what="ipf_sm_assemble_tangent")
        return r, kg
