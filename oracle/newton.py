"""Newton driver over the CPU oracle (TEST INFRASTRUCTURE ONLY).

Restates the reference's caller side: Solution (main.cpp:160-262) with the
tangent K = sum_e B_e^T (w Dep) B_e (KAssembly, TPZNumericalIntegrator.cpp:
131-182) and a direct sparse solve in place of NeoPZ's LDL^T
(main.cpp:283-285); symmetry DoFs are eliminated.  After convergence the state
is committed by one residual evaluation with UpdateMem
(AcceptPseudoTimeStepSolution, main.cpp:487-522).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O


def assemble_tangent(blocks, dof_indexes, wdep9, neq):
    R = blocks.row_sizes.astype(np.int64)
    C = blocks.col_sizes.astype(np.int64)
    if not (np.all(R == R[0]) and np.all(C == C[0])):
        raise ValueError("assemble_tangent: uniform block shapes only")
    nb, r, c = blocks.nblocks, int(R[0]), int(C[0])
    nip = r // 3
    Bt = blocks.storage.reshape(nb, nip, 3, c)
    D = np.asarray(wdep9).reshape(nb, nip, 3, 3)
    Ke = np.einsum("epkc,epkl,epld->ecd", Bt, D, Bt)
    dof = np.asarray(dof_indexes).reshape(nb, c)
    rows = np.repeat(dof, c, axis=1).reshape(-1)
    cols = np.tile(dof, (1, c)).reshape(-1)
    return sp.csr_matrix((Ke.reshape(-1), (rows, cols)), shape=(neq, neq))


class OracleNewton:
    def __init__(self, blocks, ix, neq, ncolor, material, mask, nthreads=1):
        self.B = blocks
        self.ix = ix
        self.neq = neq
        self.ncolor = ncolor
        self.mat = material
        self.free = ~np.asarray(mask, dtype=bool)
        npts = blocks.rows // 3
        self.u = np.zeros(neq)
        self.ep = np.zeros(6 * npts)
        self.alpha = np.zeros(npts)
        self.nthreads = nthreads

    def residual(self, u, f_lin, update_mem=False):
        m = self.mat
        return O.residual(self.B, self.ix["dof_indexes"], self.ix["color_indexes"], self.ncolor, self.neq,
                          self.ix["weight"], m["lambda"], m["mu"], m["phi"], m["psi"], m["cohesion"], u, f_lin=f_lin,
                          plastic_strain6=self.ep, alpha=self.alpha, update_mem=update_mem, nthreads=self.nthreads)

    def tangent(self, u):
        m = self.mat
        strain = O.multiply(self.B, self.ix["dof_indexes"], u, self.nthreads)
        ws, wd, *_ = O.compute_sigma(strain, self.ix["weight"], m["lambda"], m["mu"], m["phi"], m["psi"],
                                     m["cohesion"], plastic_strain6=self.ep.copy(), alpha=self.alpha.copy(), dep=True,
                                     nthreads=self.nthreads)
        return assemble_tangent(self.B, self.ix["dof_indexes"], wd, self.neq)

    def load_step(self, f_lin, rtol=1e-12, maxit=60):
        """Newton at load f_lin from the current u; commits. Returns stats."""
        free = self.free
        u = self.u.copy()
        r0 = None
        rn = 0.0
        it = 0
        for it in range(maxit + 1):
            r = self.residual(u, f_lin)[0]
            rn = np.linalg.norm(r[free])
            if it == 0:
                r0 = rn
            if rn <= rtol * r0 or rn == 0.0 or it == maxit:
                break
            K = self.tangent(u)[free][:, free]
            du = np.zeros(self.neq)
            du[free] = spla.spsolve(K.tocsc(), r[free])
            u = u + du
        r, ep, al, mt, br = self.residual(u, f_lin, update_mem=True)
        self.u = u
        self.ep = ep
        self.alpha = al
        return dict(iterations=it, rel_residual=(rn / r0 if r0 else 0.0), r0=r0, mtype=mt, branch=br,
                    converged=bool(rn <= rtol * r0 or rn == 0.0))
