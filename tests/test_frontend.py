"""Host front-end of the product (mesh stand-in, B blocks, weights, DoF and
color indexes, boundary loads) on the CPU -- no GPU needed (device < 0)."""
import math

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

# B matrix of one Q1 element emitted by IntPointsFEM
# (mathematica_notebooks/Elementary_Contribution/ElasticElementaryContribution.nb:2734-2886),
# element nodes 1, 9, 10, 2 of nodes.txt: (0.1, 0), (0.446667, 0), (0, 0.446667), (0, 0.1).
BEL_DX = [[-1.31307491164053, 2.53277784303890, 0.35183736217565, -1.57154029357401],
          [2.98045301654539, 1.57154029357401, 1.31307491164053, -5.86506822175993],
          [-2.15542783615350, 2.72136430957332, 0.16325089564123, -0.72918736906105],
          [-0.16325089564123, 2.27535456830035, 0.60926063691420, -2.72136430957332]]
BEL_DY = [[-5.86506822175993, 1.31307491164053, 1.57154029357401, 2.98045301654539],
          [-1.57154029357401, 0.35183736217565, 2.53277784303890, -1.31307491164053],
          [-2.72136430957332, 0.60926063691420, 2.27535456830035, -0.16325089564123],
          [-0.72918736906105, 0.16325089564123, 2.72136430957332, -2.15542783615350]]


def host_sm(mesh):
    sm = ipf.StructMatrix(mesh, device=-1)
    return sm, sm.export_blocks(), sm.export_indexes()


def oracle_blocks(b):
    return O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                    b["storage"])


def test_bel_known_answer():
    """One Q1 quarter-annulus cell reproduces the B matrix IntPointsFEM emitted
    for that element (12 x 8, rows 3*ip + {xx, xy, yy}, columns 2*a + {x, y});
    the notebook's node coordinates are given to 6 digits, hence 1e-5."""
    m = ipf.CompMesh(1, 1, 1, 0.1, 0.446667, 1.0)
    m.set_material(2000, 0.2, 1e10, 0.3, 0.3)
    sm, b, ix = host_sm(m)
    Bel = np.zeros((12, 8))
    for ip in range(4):
        for a in range(4):
            Bel[3 * ip + 0, 2 * a] = BEL_DX[ip][a]
            Bel[3 * ip + 1, 2 * a] = BEL_DY[ip][a]
            Bel[3 * ip + 1, 2 * a + 1] = BEL_DX[ip][a]
            Bel[3 * ip + 2, 2 * a + 1] = BEL_DY[ip][a]
    got = b["storage"].reshape(12, 8)
    np.testing.assert_allclose(got, Bel, rtol=0, atol=1e-5 * np.abs(Bel).max())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_patch_and_rigid_modes(p):
    """Linear displacement fields give the exact constant strain at every point;
    translations and infinitesimal rotations give zero strain."""
    m = ipf.CompMesh(5, 4, p, 0.1, 4.0, 1.3)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm, b, ix = host_sm(m)
    B = oracle_blocks(b)
    xy = m.nodes()
    a = np.array([1e-3, 2e-3, -3e-3, 0.5e-3])
    u = np.zeros(2 * len(xy))
    u[0::2] = a[0] * xy[:, 0] + a[1] * xy[:, 1]
    u[1::2] = a[2] * xy[:, 0] + a[3] * xy[:, 1]
    eps = O.multiply(B, ix["dof_indexes"], u).reshape(-1, 3)
    np.testing.assert_allclose(eps, np.tile([a[0], a[1] + a[2], a[3]], (len(eps), 1)), rtol=0, atol=1e-14)
    th = 1e-3
    u_rb = np.zeros_like(u)
    u_rb[0::2] = 0.3 - th * xy[:, 1]
    u_rb[1::2] = -0.2 + th * xy[:, 0]
    np.testing.assert_allclose(O.multiply(B, ix["dof_indexes"], u_rb), 0.0, atol=1e-14)
    # partition of unity: sum_a dN_a/dx = 0 at every point
    for e in range(b["row_sizes"].size):
        blk = b["storage"][b["matrix_position"][e]:b["matrix_position"][e + 1]].reshape(b["row_sizes"][e], -1)
        np.testing.assert_allclose(blk[0::3, 0::2].sum(axis=1), 0.0, atol=1e-12)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_weights_sum_to_polygon_area(p):
    nr, nt = 7, 9
    m = ipf.CompMesh(nr, nt, p, 0.1, 4.0, 1.2)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm, b, ix = host_sm(m)
    # straight-sided cells: the mesh is the annular polygon with nt chords per arc
    area = 0.5 * nt * math.sin(math.pi / 2 / nt) * (4.0 ** 2 - 0.1 ** 2)
    assert abs(ix["weight"].sum() - area) < 1e-12 * area
    assert (ix["weight"] > 0).all()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sizes_dofs_and_colors(p):
    nr, nt = 6, 5
    m = ipf.CompMesh(nr, nt, p, 0.1, 4.0, 1.1)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm, b, ix = host_sm(m)
    s = sm.sizes()
    V = (nr + 1) * (nt + 1)
    E = nr * (nt + 1) + nt * (nr + 1)
    F = nr * nt
    # the closed form the reference's DoF counts obey (Results/*/order*-mesh1: 2 (V + (p-1)E + (p-1)^2 F))
    assert s["neq"] == 2 * (V + (p - 1) * E + (p - 1) ** 2 * F)
    assert s["nblocks"] == nr * nt
    assert s["npts"] == nr * nt * (p + 1) ** 2
    assert s["ncolor"] == 4
    assert (b["row_sizes"] == 3 * (p + 1) ** 2).all() and (b["col_sizes"] == 2 * (p + 1) ** 2).all()
    # DoF indexes: 2*node + dir in connect order
    ptr, conn = m.connectivity()
    dof = ix["dof_indexes"].reshape(nr * nt, -1)
    np.testing.assert_array_equal(dof[:, 0::2], conn.reshape(nr * nt, -1) * 2)
    np.testing.assert_array_equal(dof[:, 1::2], conn.reshape(nr * nt, -1) * 2 + 1)
    # colors: the product's coloring equals the oracle's restatement, bit for bit
    nc, col = O.color_elements(ptr, conn, m.info()["nnodes"])
    ci = O.color_indexes(b["col_sizes"], b["col_first"], ix["dof_indexes"], col, s["neq"])
    np.testing.assert_array_equal(ci, ix["color_indexes"])
    # elements grouped by color, ascending inside a color (fElColorIndex / fFirstColorIndex)
    np.testing.assert_array_equal(ix["el_color_index"], np.argsort(col, kind="stable"))
    np.testing.assert_array_equal(ix["first_color_index"], np.concatenate([[0], np.cumsum(np.bincount(col))]))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tangent_pattern_and_l_indexes(p):
    """Create() / FillLIndexes: the upper-triangle CSR pattern is exactly the
    set of DoF pairs sharing an element (rows sorted, diagonal present), and
    fColorLSequence maps each element's packed K_e(a, c), a <= c, in
    fElColorIndex order, to the CSR position of (min, max) of its DoFs --
    checked against a dictionary built independently."""
    m = ipf.CompMesh(5, 4, p, 0.1, 4.0, 1.2)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm, b, ix = host_sm(m)
    s = sm.sizes()
    pat = sm.tangent_pattern()
    ia, ja = pat["ia"], pat["ja"]
    neq = s["neq"]
    C = b["col_sizes"]
    dof = ix["dof_indexes"]
    pairs = set()
    for e in range(s["nblocks"]):
        d = dof[b["col_first"][e]:b["col_first"][e + 1]]
        for x in d:
            for y in d:
                if x <= y:
                    pairs.add((int(x), int(y)))
    assert len(ja) == len(pairs) and ia[0] == 0 and ia[-1] == len(ja)
    pos = {}
    for row in range(neq):
        cols = ja[ia[row]:ia[row + 1]]
        assert (np.diff(cols) > 0).all() and cols[0] == row
        for k, c in enumerate(cols):
            pos[(row, int(c))] = ia[row] + k
    assert set(pos) == pairs
    expect = []
    for e in ix["el_color_index"]:
        d = dof[b["col_first"][e]:b["col_first"][e + 1]]
        for a in range(C[e]):
            for c in range(a, C[e]):
                expect.append(pos[(min(d[a], d[c]), max(d[a], d[c]))])
    np.testing.assert_array_equal(pat["color_l_sequence"], np.array(expect))
    tri = (C * C + C) // 2
    col = np.zeros(s["nblocks"], dtype=np.int64)
    fci = ix["first_color_index"]
    for c in range(s["ncolor"]):
        col[ix["el_color_index"][fci[c]:fci[c + 1]]] = c
    np.testing.assert_array_equal(pat["first_color_l_index"],
                                  np.concatenate([[0], np.cumsum(np.bincount(col, weights=tri).astype(np.int64))]))


def test_config1_sizes():
    """40 x 40 Q2 with geometric ratio 1.1: 1,600 elements, 14,400 points,
    13,122 equations; first radial cell h0 = 3.9*0.1/(1.1^40 - 1)."""
    m = wb.make_mesh("40x40-q2")
    sm, b, ix = host_sm(m)
    s = sm.sizes()
    assert (s["nblocks"], s["npts"], s["neq"], s["ncolor"]) == (1600, 14400, 13122, 4)
    xy = m.nodes()
    r = np.hypot(xy[:, 0], xy[:, 1])
    h0 = 3.9 * 0.1 / (1.1 ** 40 - 1)
    assert abs(np.unique(np.round(r[xy[:, 1] == 0], 12))[2] - (0.1 + h0)) < 1e-12  # node index 2 = first corner


def test_boundary_load_resultant():
    """Normal traction sigma_n on the inner polygonal arc: the load vector's
    resultant is sigma_n times the integral of the outward normal, which for
    the chord polyline from (r,0) to (0,r) is (-r, -r)."""
    for p in (1, 2, 3):
        m = ipf.CompMesh(4, 6, p, 0.1, 4.0, 1.2)
        m.set_material(5000, 0.2, 10, 0.5, 0.5)
        m.set_normal_stress(30.5, -7.0)
        sm, b, ix = host_sm(m)
        f = ix["rhs_linear"]
        # inner: outward normal points to the hole, outer: away from it
        fx = 30.5 * (-0.1) + (-7.0) * 4.0
        np.testing.assert_allclose([f[0::2].sum(), f[1::2].sum()], [fx, fx], rtol=1e-13)


def test_dirichlet_mask():
    m = ipf.CompMesh(3, 2, 2, 0.1, 4.0, 1.0)
    mask = m.dirichlet_mask()
    xy = m.nodes()
    on_x = np.isclose(xy[:, 1], 0.0)
    on_y = np.isclose(xy[:, 0], 0.0, atol=1e-15)
    np.testing.assert_array_equal(mask[1::2], on_x)
    np.testing.assert_array_equal(mask[0::2], on_y)
    assert not mask[0::2][on_x & ~on_y].any()


def test_abi_exports_every_declared_symbol():
    L = ipf.lib()
    names = ipf.exported_symbols()
    assert len(names) >= 39
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_errors_are_reported():
    with pytest.raises(ipf.IPFError, match="order"):
        ipf.CompMesh(4, 4, 5)
    m = ipf.CompMesh(2, 2, 1)
    with pytest.raises(ipf.IPFError, match="E / nu"):
        m.set_material(-1.0, 0.2, 1, 0.5, 0.5)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_blocks_from_corners_and_shape_table(p):
    """The host blocks are the geometry formula applied to the mesh's corners
    and shape table (the input of the device setup): per point J from the
    bilinear map, inv = 1/det, (a, b, c, d) = (ye, -yx, -xe, xx) * inv,
    dN/dx = a dN/dxi + b dN/deta, dN/dy = c dN/dxi + d dN/deta -- restated
    here in numpy in the same operation order, so the bits must agree; the
    weights are w_q |det|."""
    m = ipf.CompMesh(4, 3, p, 0.1, 4.0, 1.2)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm, b, ix = host_sm(m)
    (nip, nsh), tab = m.shape_table()
    X = m.corners()
    dphi = tab[:nip * nsh * 2].reshape(nip, nsh, 2)
    xi, eta, wq = tab[nip * nsh * 2:].reshape(3, nip)
    C = 2 * nsh
    for e in range(X.shape[0]):
        blk = b["storage"][b["matrix_position"][e]:b["matrix_position"][e + 1]].reshape(3 * nip, C)
        for ip in range(nip):
            x, y = X[e, 0::2], X[e, 1::2]
            dxi = np.array([-(1 - eta[ip]) / 4, (1 - eta[ip]) / 4, (1 + eta[ip]) / 4, -(1 + eta[ip]) / 4])
            det_ = np.array([-(1 - xi[ip]) / 4, -(1 + xi[ip]) / 4, (1 + xi[ip]) / 4, (1 - xi[ip]) / 4])
            xx = xe = yx = ye = 0.0
            for c in range(4):
                xx += dxi[c] * x[c]
                xe += det_[c] * x[c]
                yx += dxi[c] * y[c]
                ye += det_[c] * y[c]
            det = xx * ye - xe * yx
            inv = 1.0 / det
            A, Bc, Cc, D = ye * inv, -(yx * inv), -(xe * inv), xx * inv
            dx = A * dphi[ip, :, 0] + Bc * dphi[ip, :, 1]
            dy = Cc * dphi[ip, :, 0] + D * dphi[ip, :, 1]
            assert np.array_equal(blk[3 * ip, 0::2], dx) and np.array_equal(blk[3 * ip + 2, 1::2], dy)
            assert np.array_equal(blk[3 * ip + 1, 0::2], dy) and np.array_equal(blk[3 * ip + 1, 1::2], dx)
            assert not blk[3 * ip, 1::2].any() and not blk[3 * ip + 2, 0::2].any()
            assert ix["weight"][b["row_first"][e] // 3 + ip] == wq[ip] * abs(det)


def test_clone_keeps_host_data_after_original_is_gone():
    """Clone() (TPZElastoPlasticIntPointsStructMatrix.cpp:30-32) copies the
    host data; the original may be destroyed first (no shared handle is
    freed twice, unlike TPZVecGPU's implicit copy, TPZVecGPU.h:24-27)."""
    m = ipf.CompMesh(5, 4, 2, 0.1, 4.0, 1.1)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    m.set_normal_stress(3.0, 0.0)
    sm = ipf.StructMatrix(m, device=-1)
    ref = sm.export_indexes()
    refb = sm.export_blocks()
    c = sm.clone()
    del sm
    got = c.export_indexes()
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(c.export_blocks()["storage"], refb["storage"])
    assert c.create_tangent() > 0


@pytest.mark.parametrize("p", [1, 3])
def test_csr_vectors_describe_the_block_matrix(p):
    """CSRVectors (TPZIrregularBlocksMatrix.cpp:177-195): (row_ptr, col_ind)
    over fStorage is the block-diagonal matrix of the blocks -- checked
    against scipy's block_diag of the exported blocks, entry by entry, and
    as an operator on a random vector."""
    import scipy.sparse as sp
    m = ipf.CompMesh(3, 4, p, 0.1, 4.0, 1.2)
    m.set_material(5000, 0.2, 10, 0.5, 0.5)
    sm = ipf.StructMatrix(m, device=-1)
    s = sm.sizes()
    b = sm.export_blocks()
    rp, ci = sm.export_csr()
    assert rp[0] == 0 and rp[-1] == s["nstorage"]
    A = sp.csr_matrix((b["storage"], ci, rp), shape=(s["rows"], s["cols"]))
    blocks = [b["storage"][b["matrix_position"][e]:b["matrix_position"][e + 1]].reshape(b["row_sizes"][e],
                                                                                         b["col_sizes"][e])
              for e in range(s["nblocks"])]
    D = sp.block_diag(blocks, format="csr")
    assert abs(A - D).max() == 0
    x = np.random.default_rng(p).standard_normal(s["cols"])
    np.testing.assert_allclose(A @ x, D @ x, rtol=0, atol=0)
