"""Pins of the oracle's block products, colored scatter and coloring
(TPZIrregularBlocksMatrix / TPZNumericalIntegrator restatements) against
dense numpy algebra and the invariants the reference relies on."""
import numpy as np
import pytest

from oracle import oracle as O


def random_blocks(rng, nb, shapes=((12, 8), (27, 18), (48, 32))):
    R = []
    Cc = []
    for _ in range(nb):
        r, c = shapes[rng.integers(len(shapes))]
        R.append(r)
        Cc.append(c)
    R = np.array(R, np.int32)
    Cc = np.array(Cc, np.int32)
    mp = np.concatenate([[0], np.cumsum(R.astype(np.int64) * Cc)])
    rf = np.concatenate([[0], np.cumsum(R)]).astype(np.int32)
    cf = np.concatenate([[0], np.cumsum(Cc)]).astype(np.int32)
    st = rng.standard_normal(int(mp[-1]))
    return O.Blocks(R, Cc, mp, rf, cf, st)


def dense(B):
    M = np.zeros((B.rows, B.cols))
    for e in range(B.nblocks):
        r, c = B.row_sizes[e], B.col_sizes[e]
        blk = B.storage[B.mat_pos[e]:B.mat_pos[e + 1]].reshape(r, c)  # row-major block
        M[B.row_first[e]:B.row_first[e] + r, B.col_first[e]:B.col_first[e] + c] = blk
    return M


def test_multiply_vector_matches_dense():
    rng = np.random.default_rng(0)
    B = random_blocks(rng, 17)
    M = dense(B)
    x = rng.standard_normal(B.cols)
    np.testing.assert_allclose(O.multiply_vector(B, x, 0), M @ x, rtol=1e-13, atol=1e-13)
    s = rng.standard_normal(B.rows)
    # opt = 1 is -B^T x (alpha = -1, TPZIrregularBlocksMatrix.cpp:110,136)
    np.testing.assert_allclose(O.multiply_vector(B, s, 1), -(M.T @ s), rtol=1e-13, atol=1e-13)


def test_gather_multiply_and_colored_transpose():
    """Multiply = B . u[dof]; MultiplyTranspose = scatter-add of -B^T s into
    neq, through ncolor slabs and the pairwise-halving reduction."""
    rng = np.random.default_rng(1)
    neq = 40
    # 5 colors, each with DoF-disjoint blocks of 8 columns
    blocks = []
    dof = []
    color = []
    for c in range(5):
        perm = rng.permutation(neq)
        for b in range(4):
            dof.extend(perm[8 * b:8 * b + 8])
            color.append(c)
    nb = len(color)
    R = np.full(nb, 12, np.int32)
    Cc = np.full(nb, 8, np.int32)
    mp = np.arange(nb + 1, dtype=np.int64) * 96
    rf = (np.arange(nb + 1) * 12).astype(np.int32)
    cf = (np.arange(nb + 1) * 8).astype(np.int32)
    B = O.Blocks(R, Cc, mp, rf, cf, rng.standard_normal(96 * nb))
    dof = np.array(dof, np.int32)
    ci = O.color_indexes(Cc, cf, dof, np.repeat(np.array(color, np.int64), 1), neq)
    np.testing.assert_array_equal(ci, dof + np.repeat(color, 8) * neq)
    M = dense(B)
    u = rng.standard_normal(neq)
    np.testing.assert_allclose(O.multiply(B, dof, u), M @ u[dof], rtol=1e-13, atol=1e-13)
    s = rng.standard_normal(B.rows)
    ref = np.zeros(neq)
    np.add.at(ref, dof, -(M.T @ s))
    np.testing.assert_allclose(O.multiply_transpose(B, ci, 5, neq, s), ref, rtol=1e-12, atol=1e-12)


def grid_connectivity(nx, ny, p=2):
    """Structured quad mesh, nodes (p*nx+1) x (p*ny+1), radial-fastest elements."""
    npx = p * nx + 1
    ptr = [0]
    conn = []
    for j in range(ny):
        for i in range(nx):
            for b in range(p + 1):
                for a in range(p + 1):
                    conn.append((p * j + b) * npx + p * i + a)
            ptr.append(len(conn))
    return np.array(ptr), np.array(conn), npx * (p * ny + 1)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_coloring_structured_quads(p):
    """Greedy coloring (TPZNumericalIntegrator.cpp:452-536): 4 colors on
    structured quadrilaterals (the count every Results/*/out-*.txt log prints,
    e.g. Results/CPU_NOT_MODIFIED/order2-mesh4/out-1.txt:64), a checkerboard
    pattern, and no connect shared inside a color."""
    ptr, conn, nconn = grid_connectivity(9, 7, p)
    nc, col = O.color_elements(ptr, conn, nconn)
    assert nc == 4
    ij = [(e % 9, e // 9) for e in range(63)]
    for e, (i, j) in enumerate(ij):
        assert col[e] == (i % 2) + 2 * (j % 2)
    for c in range(nc):
        seen = set()
        for e in np.nonzero(col == c)[0]:
            nodes = set(conn[ptr[e]:ptr[e + 1]])
            assert not (nodes & seen)
            seen |= nodes


def test_coloring_irregular_graph():
    rng = np.random.default_rng(5)
    ptr = [0]
    conn = []
    for e in range(200):
        conn.extend(rng.choice(120, size=6, replace=False))
        ptr.append(len(conn))
    ptr = np.array(ptr)
    conn = np.array(conn)
    nc, col = O.color_elements(ptr, conn, 120)
    assert col.min() == 0 and col.max() == nc - 1
    for c in range(nc):
        seen = set()
        for e in np.nonzero(col == c)[0]:
            nodes = set(conn[ptr[e]:ptr[e + 1]])
            assert not (nodes & seen)
            seen |= nodes
    # multi-pass greedy (TPZNumericalIntegrator.cpp:463-493): pass c visits the
    # uncolored elements in index order, so an element left out of pass c
    # conflicts with an element of color c that precedes it
    for e in range(200):
        ne = set(conn[ptr[e]:ptr[e + 1]])
        for c in range(col[e]):
            assert any(ne & set(conn[ptr[f]:ptr[f + 1]]) for f in np.nonzero(col == c)[0] if f < e)


def test_element_tangents_match_numpy():
    rng = np.random.default_rng(7)
    B = random_blocks(rng, 6, shapes=((27, 18),))
    npts = B.rows // 3
    D = rng.standard_normal((npts, 3, 3))
    order = rng.permutation(B.nblocks).astype(np.int32)
    kc = O.element_tangents(B, D.reshape(-1), order)
    off = 0
    for i in order:
        r, c = B.row_sizes[i], B.col_sizes[i]
        blk = B.storage[B.mat_pos[i]:B.mat_pos[i + 1]].reshape(r // 3, 3, c)
        Ke = sum(blk[ip].T @ D[B.row_first[i] // 3 + ip] @ blk[ip] for ip in range(r // 3))
        iu = np.triu_indices(c)
        np.testing.assert_allclose(kc[off:off + len(iu[0])], Ke[iu], rtol=1e-12, atol=1e-12)
        off += len(iu[0])
