"""ctypes wrapper of the CPU oracle (oracle/libipf_oracle.so).

TEST INFRASTRUCTURE ONLY: the parity tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker.  The product
package ``intpointsfem_amd`` never imports it.

The C sources restate the CPU path of IntPointsFEM; see ipf_oracle.c for the
reference file:line of each function.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libipf_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_d = ctypes.c_double
_int = ctypes.c_int


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Blocks(ctypes.Structure):
    _fields_ = [
        ("nblocks", ctypes.c_int64),
        ("row_sizes", _ip),
        ("col_sizes", _ip),
        ("mat_pos", _lp),
        ("row_first", _ip),
        ("col_first", _ip),
        ("storage", _dp),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_spectral_decomposition.argtypes = [_dp, _dp, _dp]
        L.orc_eigenvectors.argtypes = [_dp, _dp, _dp, _d]
        L.orc_project_sigma.argtypes = [_dp, _dp, _ip, _dp, _d, _d, _d, _d, _d, _dp, _int, _ip]
        L.orc_tangent_operator.argtypes = [_dp, _dp, _dp, _dp, _dp, _d, _d]
        L.orc_point_sigma.argtypes = [_dp, _dp, _dp, _ip, _ip, _d, _d, _d, _d, _d, _d, _dp, _dp, _dp]
        L.orc_compute_sigma.argtypes = [_i64, _dp, _dp, _dp, _d, _d, _d, _d, _d, _dp, _dp, _ip, _ip, _int, _int]
        L.orc_compute_sigma_dep.argtypes = [_i64, _dp, _dp, _dp, _dp, _d, _d, _d, _d, _d, _dp, _dp, _ip, _ip,
                                            _int, _int]
        L.orc_multiply_vector.argtypes = [ctypes.POINTER(_Blocks), _dp, _dp, _int, _int]
        L.orc_multiply.argtypes = [ctypes.POINTER(_Blocks), _ip, _dp, _dp, _int]
        L.orc_multiply_transpose.argtypes = [ctypes.POINTER(_Blocks), _ip, _i64, _i64, _dp, _dp, _dp, _int]
        L.orc_color_elements.argtypes = [_i64, _lp, _lp, _i64, _lp]
        L.orc_color_elements.restype = _i64
        L.orc_color_indexes.argtypes = [_i64, _ip, _ip, _ip, _lp, _i64, _ip]
        L.orc_residual.argtypes = [ctypes.POINTER(_Blocks), _ip, _ip, _i64, _i64, _dp, _d, _d, _d, _d, _d,
                                   _dp, _dp, _ip, _ip, _int, _dp, _dp, _dp, _dp, _dp, _dp, _int]
        L.orc_element_tangents.argtypes = [ctypes.POINTER(_Blocks), _dp, _ip, _dp, _int]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------- point math

def spectral_decomposition(sigma6):
    """FunctionsSpectralDecomp.h:261-268. Returns (eigenvalues, eigenvectors(3x3 rows), sigma6 after)."""
    s = _f64(sigma6).copy()
    ev = np.zeros(3)
    vec = np.zeros(9)
    lib().orc_spectral_decomposition(_p(s, _dp), _p(ev, _dp), _p(vec, _dp))
    return ev, vec.reshape(3, 3), s


def eigenvectors(sigma6, eigenvalues, maxel=1.0):
    """FunctionsSpectralDecomp.h:217-256 for given eigenvalues. Returns 3x3 (rows = vectors)."""
    s = _f64(sigma6).copy()
    vec = np.zeros(9)
    lib().orc_eigenvectors(_p(s, _dp), _p(_f64(eigenvalues), _dp), _p(vec, _dp), maxel)
    return vec.reshape(3, 3)


def project_sigma(eigenvalues, phi, psi, c, K, G, alpha=0.0, gradient=False):
    """FunctionsSigmaProjection.h:337-379. Returns (sigma_projected, mtype, alpha, branch, gradient)."""
    ev = _f64(eigenvalues)
    sp = np.zeros(3)
    mt = ctypes.c_int(0)
    br = ctypes.c_int(0)
    a = ctypes.c_double(alpha)
    g = np.zeros(9)
    lib().orc_project_sigma(_p(ev, _dp), _p(sp, _dp), ctypes.byref(mt), ctypes.byref(a), phi, psi, c, K, G,
                            _p(g, _dp), 1 if gradient else 0, ctypes.byref(br))
    return sp, mt.value, a.value, br.value, g.reshape(3, 3)


def point_sigma(strain3, plastic_strain6, lam, mu, phi, psi, c, weight=1.0, alpha=0.0, dep=False):
    """One point of ComputeSigma / ComputeSigmaDep. Returns dict."""
    e = _f64(strain3)
    ep = _f64(plastic_strain6).copy()
    a = ctypes.c_double(alpha)
    mt = ctypes.c_int(0)
    br = ctypes.c_int(0)
    ws = np.zeros(3)
    s6 = np.zeros(6)
    d9 = np.zeros(9) if dep else None
    lib().orc_point_sigma(_p(e, _dp), _p(ep, _dp), ctypes.byref(a), ctypes.byref(mt), ctypes.byref(br), weight,
                          lam, mu, phi, psi, c, _p(ws, _dp), _p(s6, _dp), _p(d9, _dp))
    out = dict(wsigma=ws, sigma6=s6, plastic_strain=ep, alpha=a.value, mtype=mt.value, branch=br.value)
    if dep:
        out["wdep"] = d9.reshape(3, 3)
    return out


def compute_sigma(strain3, weight, lam, mu, phi, psi, c, plastic_strain6=None, alpha=None, update_mem=False,
                  nthreads=1, dep=False):
    """TPZConstitutiveLawProcessor::ComputeSigma[Dep] over all points (AoS arrays)."""
    strain3 = _f64(strain3).reshape(-1)
    npts = strain3.size // 3
    w = _f64(weight)
    ep = np.zeros(6 * npts) if plastic_strain6 is None else _f64(plastic_strain6).reshape(-1)
    al = np.zeros(npts) if alpha is None else _f64(alpha)
    mt = np.zeros(npts, dtype=np.int32)
    br = np.zeros(npts, dtype=np.int32)
    ws = np.zeros(3 * npts)
    if dep:
        wd = np.zeros(9 * npts)
        lib().orc_compute_sigma_dep(npts, _p(strain3, _dp), _p(ws, _dp), _p(wd, _dp), _p(w, _dp), lam, mu, phi,
                                    psi, c, _p(ep, _dp), _p(al, _dp), _p(mt, _ip), _p(br, _ip),
                                    1 if update_mem else 0, nthreads)
        return ws, wd, ep, al, mt, br
    lib().orc_compute_sigma(npts, _p(strain3, _dp), _p(ws, _dp), _p(w, _dp), lam, mu, phi, psi, c, _p(ep, _dp),
                            _p(al, _dp), _p(mt, _ip), _p(br, _ip), 1 if update_mem else 0, nthreads)
    return ws, ep, al, mt, br


def tangent_operator(gradient, stress_ev, strain_ev, strain_vec, G, lam):
    g = _f64(gradient).reshape(-1)
    T = np.zeros(36)
    lib().orc_tangent_operator(_p(g, _dp), _p(_f64(stress_ev), _dp), _p(_f64(strain_ev), _dp),
                               _p(_f64(strain_vec).reshape(-1), _dp), _p(T, _dp), G, lam)
    return T.reshape(6, 6)


# ---------------------------------------------------------------- blocks

class Blocks:
    """Irregular block matrix in the reference layout (TPZIrregularBlocksMatrix.h:18-30)."""

    def __init__(self, row_sizes, col_sizes, mat_pos, row_first, col_first, storage):
        self.row_sizes = _i32(row_sizes)
        self.col_sizes = _i32(col_sizes)
        self.mat_pos = _i64a(mat_pos)
        self.row_first = _i32(row_first)
        self.col_first = _i32(col_first)
        self.storage = _f64(storage)
        self.nblocks = self.row_sizes.size
        self._s = _Blocks(self.nblocks, _p(self.row_sizes, _ip), _p(self.col_sizes, _ip), _p(self.mat_pos, _lp),
                          _p(self.row_first, _ip), _p(self.col_first, _ip), _p(self.storage, _dp))

    @property
    def rows(self):
        return int(self.row_first[-1])

    @property
    def cols(self):
        return int(self.col_first[-1])

    def ref(self):
        return ctypes.byref(self._s)


def multiply_vector(blocks, x, opt, nthreads=1):
    x = _f64(x)
    y = np.zeros(blocks.rows if opt == 0 else blocks.cols)
    lib().orc_multiply_vector(blocks.ref(), _p(x, _dp), _p(y, _dp), opt, nthreads)
    return y


def multiply(blocks, dof_indexes, u, nthreads=1):
    u = _f64(u)
    strain = np.zeros(blocks.rows)
    lib().orc_multiply(blocks.ref(), _p(_i32(dof_indexes), _ip), _p(u, _dp), _p(strain, _dp), nthreads)
    return strain


def multiply_transpose(blocks, color_indexes, ncolor, neq, wsigma, nthreads=1):
    res = np.zeros(neq)
    lib().orc_multiply_transpose(blocks.ref(), _p(_i32(color_indexes), _ip), ncolor, neq, _p(_f64(wsigma), _dp),
                                 _p(res, _dp), None, nthreads)
    return res


def color_elements(conn_ptr, conn, nconnects):
    conn_ptr = _i64a(conn_ptr)
    conn = _i64a(conn)
    nb = conn_ptr.size - 1
    colors = np.zeros(nb, dtype=np.int64)
    nc = lib().orc_color_elements(nb, _p(conn_ptr, _lp), _p(conn, _lp), nconnects, _p(colors, _lp))
    return int(nc), colors


def color_indexes(col_sizes, col_first, dof_indexes, elemcolor, neq):
    col_sizes = _i32(col_sizes)
    ci = np.zeros(int(col_first[-1]), dtype=np.int32)
    lib().orc_color_indexes(col_sizes.size, _p(col_sizes, _ip), _p(_i32(col_first), _ip), _p(_i32(dof_indexes), _ip),
                            _p(_i64a(elemcolor), _lp), neq, _p(ci, _ip))
    return ci


def residual(blocks, dof_indexes, color_indexes_, ncolor, neq, weight, lam, mu, phi, psi, c, u, f_lin=None,
             plastic_strain6=None, alpha=None, update_mem=False, nthreads=1):
    """Assemble(rhs): MultiplyTranspose(ComputeSigma(Multiply(u))) + f_lin.
    Returns (rhs, plastic_strain6, alpha, mtype, branch)."""
    npts = blocks.rows // 3
    ep = np.zeros(6 * npts) if plastic_strain6 is None else _f64(plastic_strain6).reshape(-1).copy()
    al = np.zeros(npts) if alpha is None else _f64(alpha).copy()
    mt = np.zeros(npts, dtype=np.int32)
    br = np.zeros(npts, dtype=np.int32)
    rhs = np.zeros(neq)
    lib().orc_residual(blocks.ref(), _p(_i32(dof_indexes), _ip), _p(_i32(color_indexes_), _ip), ncolor, neq,
                       _p(_f64(weight), _dp), lam, mu, phi, psi, c, _p(ep, _dp), _p(al, _dp), _p(mt, _ip),
                       _p(br, _ip), 1 if update_mem else 0, _p(_f64(u), _dp),
                       _p(None if f_lin is None else _f64(f_lin), _dp), _p(rhs, _dp), None, None, None, nthreads)
    return rhs, ep, al, mt, br


def element_tangents(blocks, wdep9, el_color_index, nthreads=1):
    C = blocks.col_sizes.astype(np.int64)
    total = int(((C * C + C) // 2).sum())
    kc = np.zeros(total)
    lib().orc_element_tangents(blocks.ref(), _p(_f64(wdep9), _dp), _p(_i32(el_color_index), _ip), _p(kc, _dp),
                               nthreads)
    return kc
