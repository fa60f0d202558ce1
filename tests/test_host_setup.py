"""CPU tests of the product library's host side (no GPU calls): the C-ABI
loads and exports everything include/cp_abi.h declares; its mesh generators,
load/BC rules and reference CSR pattern agree with the oracle bit for bit."""
import os

import numpy as np
import pytest

import oracle_lib as O
from cudaplasticity_amd import abi as A


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = A.declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.cp_abi_version() == 1


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU the context cannot be created and no CPU path exists."""
    from conftest import HAS_GPU
    if HAS_GPU:
        pytest.skip("a device is visible")
    with pytest.raises(A.CPError) as e:
        A.Context(0)
    assert e.value.code == A.CP_ERR_NO_DEVICE


MESHES = [
    ("rect", (-1, 1, -1, 1, 4, 4)),
    ("rect", (0, 3, -2, 1, 7, 5)),
    ("hole", (-3, 3, -2, 2, 0.5, 48, 32)),
    ("hole", (-1, 2, -1, 1, 0.2, 10, 6)),
    ("ring", (0.5, 1.0, 40, 5)),
    ("ring", (0.1, 0.2, 41, 3)),
    ("qplate", (1.0, 0.1, 100, 49)),
    ("qplate", (2.0, 0.3, 17, 4)),
    ("qring", (0.5, 1.0, 64, 20, 1.05)),
    ("qring", (0.5, 1.0, 16, 8, 1.0)),
]


def _pair(kind, args):
    o = {"rect": O.gen_rect, "hole": O.gen_rect_with_hole, "ring": O.gen_ring,
         "qplate": O.gen_quarter_plate, "qring": O.gen_quarter_ring}[kind](*args)
    p = {"rect": A.gen_rect, "hole": A.gen_rect_with_hole, "ring": A.gen_ring,
         "qplate": A.gen_quarter_plate, "qring": A.gen_quarter_ring}[kind](*args)
    return o, p


@pytest.mark.parametrize("kind,args", MESHES)
def test_generators_match_oracle(kind, args):
    o, p = _pair(kind, args)
    assert np.array_equal(o.conn, p.conn)
    assert len(o.borders) == len(p.borders)
    for bo, bp in zip(o.borders, p.borders):
        assert np.array_equal(bo, bp)
    assert np.array_equal(o.x, p.x) and np.array_equal(o.y, p.y)


@pytest.mark.parametrize("kind,args", MESHES)
def test_csr_pattern_matches_oracle(kind, args):
    o, p = _pair(kind, args)
    ro, co = O.csr_pattern(o)
    rp, cp = A.csr_pattern(p)
    assert np.array_equal(ro, rp) and np.array_equal(co, cp)


def test_loads_match_oracle():
    o, p = _pair("hole", (-3, 3, -2, 2, 0.5, 24, 16))
    lo = O.Problem(o).force(1, 170.0, 3.0).force(3, -170.0, 0.0).force(4, 5.0, 1.0, norm_orient=True)
    lo = lo.gravity().fix_vert_axis(0).fix_hor_axis(0).displacement(2, 1e-3, 0.0, False, True)
    lp = A.Loads(p).force(1, 170.0, 3.0).force(3, -170.0, 0.0).force(4, 5.0, 1.0, norm_orient=True)
    lp = lp.gravity().fix_vert_axis(0).fix_hor_axis(0).displacement(2, 1e-3, 0.0, False, True)
    assert np.array_equal(lo.rp, lp.rp)
    assert np.array_equal(lo.mask, lp.mask)
    assert np.array_equal(lo.u, lp.u)


def test_degenerate_and_out_of_range_connectivity_rejected():
    m = A.gen_rect(0, 1, 0, 1, 2, 2)
    bad = m.conn.copy()
    bad[1, 2] = bad[1, 0]
    with pytest.raises(A.CPError):
        A.csr_pattern(A.Mesh(m.x, m.y, bad, m.borders))
    bad = m.conn.copy()
    bad[0, 0] = m.nn
    with pytest.raises(A.CPError):
        A.csr_pattern(A.Mesh(m.x, m.y, bad, m.borders))


def test_problem_bytes_budget_monotone():
    import ctypes as C
    b1, b2 = C.c_size_t(), C.c_size_t()
    A.check(A.lib().cp_problem_bytes(5050, 4900, 180000, A.MIXED, C.byref(b1)))
    A.check(A.lib().cp_problem_bytes(201000, 200000, 7200000, A.MIXED, C.byref(b2)))
    assert 0 < b1.value < b2.value
    b3 = C.c_size_t()
    A.check(A.lib().cp_problem_bytes(201000, 200000, 7200000, A.FP64, C.byref(b3)))
    assert b3.value > b2.value


def test_nodal_load_entry_points_match_constant_ones():
    """cp_rhs_body_force_nodal with a constant field equals cp_rhs_body_force;
    cp_displacement_border_values with constant values equals
    cp_displacement_border (both bit-exact, every fix-mid combination)."""
    m = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 24, 16)
    lib = A.lib()
    import ctypes as C
    rp1 = np.zeros(2 * m.nn)
    rp2 = np.zeros(2 * m.nn)
    A.check(lib.cp_rhs_body_force(m.ne, A.dp(m.x), A.dp(m.y), A.ip(m.conn), C.c_double(0.5), C.c_double(-9.81),
                                  C.c_double(2.0), C.c_double(0.3), A.dp(rp1)))
    rn = np.tile([0.5, -9.81], m.nn)
    A.check(lib.cp_rhs_body_force_nodal(m.ne, A.dp(m.x), A.dp(m.y), A.ip(m.conn), A.dp(rn), C.c_double(2.0),
                                        C.c_double(0.3), A.dp(rp2)))
    assert np.array_equal(rp1, rp2)
    b = m.borders[2]
    for mx in (0, 1):
        for my in (0, 1):
            k1, u1 = np.ones(2 * m.nn, np.uint8), np.zeros(2 * m.nn)
            k2, u2 = np.ones(2 * m.nn, np.uint8), np.zeros(2 * m.nn)
            A.check(lib.cp_displacement_border(A.ip(b), b.size, C.c_double(1e-3), C.c_double(-2e-3), mx, my,
                                               A.u8p(k1), A.dp(u1)))
            uv = np.tile([1e-3, -2e-3], b.size)
            A.check(lib.cp_displacement_border_values(A.ip(b), b.size, A.dp(uv), mx, my, A.u8p(k2), A.dp(u2)))
            assert np.array_equal(k1, k2) and np.array_equal(u1, u2)


MESHES2 = [
    ("rect", (-2, 2, -1, 1, 12, 6)),
    ("rect", (0, 3, -2, 1, 5, 3)),
    ("hole", (-3, 3, -2, 2, 0.5, 12, 8)),
    ("hole", (-3, 3, -2, 2, 0.5, 18, 12)),
    ("ring", (1, 2, 40, 6)),
    ("ring", (0.1, 0.2, 9, 2)),
]


@pytest.mark.parametrize("kind,args", MESHES2)
def test_order2_generators_match_oracle(kind, args):
    o = {"rect": O.gen_rect, "hole": O.gen_rect_with_hole, "ring": O.gen_ring}[kind](*args, order=2)
    p = {"rect": A.gen_rect, "hole": A.gen_rect_with_hole, "ring": A.gen_ring}[kind](*args, order=2)
    assert p.npe == 8 and np.array_equal(o.conn, p.conn)
    assert np.array_equal(o.x, p.x) and np.array_equal(o.y, p.y)
    assert np.array_equal(o.sec, p.sec)
    assert len(o.borders) == len(p.borders)
    for bo, bp in zip(o.borders, p.borders):
        assert np.array_equal(bo, bp)
    # reference CSR pattern over elem8
    ro, co = O.csr_pattern(o)
    rp, cp = A.csr_pattern(p)
    assert np.array_equal(ro, rp) and np.array_equal(co, cp)


def test_order2_generator_matches_reference_mesh_file():
    """hole1.vtk: genRectWithHole(-3, 3, -2, 2, 0.5, 12, 8, 2) saved by Mesh::saveAsVtk."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mesh_hole1.npz"))
    p = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2)
    assert np.array_equal(p.conn[:, [0, 4, 1, 5, 2, 6, 3, 7]], g["polys"])


def test_order2_loads_match_oracle():
    o = O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2)
    p = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2)
    po = O.Problem(o, h=0.7).force(1, 170.0, 3.0).force(3, -170.0, 0.0).force(4, 5.0, 1.0, norm_orient=True)
    pa = A.Loads(p, h=0.7).force(1, 170.0, 3.0).force(3, -170.0, 0.0).force(4, 5.0, 1.0, norm_orient=True)
    po.fix_vert_axis(0.0).fix_hor_axis(0.0)
    pa.fix_vert_axis(0.0).fix_hor_axis(0.0)
    assert np.array_equal(po.rp, pa.rp)
    assert np.array_equal(po.mask, pa.mask)


def test_remap_rejects_odd_strips():
    m = A.gen_rect(0, 1, 0, 1, 3, 2)
    x8, y8 = np.zeros(m.nn), np.zeros(m.nn)
    c8 = np.zeros(2 * m.conn.size, dtype=np.int32)
    sec = np.zeros(m.nn, dtype=np.uint8)
    nn8 = A.C.c_int()
    code = A.lib().cp_mesh_remap_order(m.nn, m.ne, A.dp(m.x), A.dp(m.y), A.ip(m.conn), 2, 0, None, A.dp(x8),
                                       A.dp(y8), A.ip(c8), A.u8p(sec), A.C.byref(nn8))
    assert code == A.CP_ERR_INVALID


def write_inp(path, m, elem_type, extra_nodes=True, blank_between=True):
    """An Abaqus-style deck of mesh m: *NODE (1-based, then one MPC reference
    node per border), the elements, one *MPC block per border."""
    lines = ["*HEADING", "test deck", "*NODE"]
    for i in range(m.nn):
        lines.append("%d, %.17g, %.17g, 0" % (i + 1, m.x[i], m.y[i]))
    if extra_nodes:
        for b in range(len(m.borders)):
            lines.append("%d, 0, 0, 0" % (m.nn + b + 1))
    if blank_between:
        lines.append("")
    lines.append("*ELEMENT, TYPE=%s, ELSET=PLATE" % elem_type)
    for e, c in enumerate(m.conn):
        lines.append("%d, " % (e + 1) + ", ".join(str(v + 1) for v in c))
    lines.append("**")
    for b, bd in enumerate(m.borders):
        lines.append("*MPC")
        for v in bd:
            lines.append("TIE, %d, %d" % (v + 1, m.nn + b + 1))
    lines.append("*END")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("order", [1, 2])
def test_inp_loader_round_trip(tmp_path, order):
    """Mesh::loadFromFile (Mesh.cu:725-937): a deck written from a generated
    mesh loads back to the same nodes, elements, mid-side flags and borders."""
    m = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=order)
    p = os.path.join(str(tmp_path), "plate.inp")
    write_inp(p, m, "S8R" if order == 2 else "S4")
    g = A.load_inp(p)
    assert g.npe == m.npe and np.array_equal(g.conn, m.conn)
    assert np.array_equal(g.x, m.x) and np.array_equal(g.y, m.y)
    assert len(g.borders) == len(m.borders)
    for a, b in zip(g.borders, m.borders):
        assert np.array_equal(a, b)
    if order == 2:
        assert np.array_equal(g.sec, m.sec)
    # the load vector built on the loaded mesh equals the generated one's
    la = A.Loads(g).force(1, 170.0, 0.0).force(3, -170.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    lb = A.Loads(m).force(1, 170.0, 0.0).force(3, -170.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    assert np.array_equal(la.rp, lb.rp) and np.array_equal(la.mask, lb.mask)


def test_inp_loader_errors(tmp_path):
    with pytest.raises(A.CPError) as e:
        A.load_inp(os.path.join(str(tmp_path), "mesh.msh"))
    assert "Unknown mesh file format" in str(e.value)
    with pytest.raises(A.CPError) as e:
        A.load_inp(os.path.join(str(tmp_path), "missing.inp"))
    assert "There is no such file" in str(e.value)
    # an element through a border's reference node (an id past the mesh nodes)
    m = A.gen_rect(0, 1, 0, 1, 2, 2)
    bad = A.Mesh(m.x, m.y, m.conn.copy(), m.borders)
    bad.conn[0, 2] = m.nn + 1
    p = os.path.join(str(tmp_path), "bad.inp")
    write_inp(p, bad, "S4")
    with pytest.raises(A.CPError) as e:
        A.load_inp(p)
    assert "out of range" in str(e.value)


def test_body_force_8node_matches_oracle_and_closed_form():
    m = A.gen_ring(1, 2, 12, 3, order=2)
    o = O.gen_ring(1, 2, 12, 3, order=2)
    la = A.Loads(m, h=0.5).gravity(9.81, 7.8)
    lo = O.Problem(o, h=0.5).gravity(9.81, 7.8)
    assert np.array_equal(la.rp, lo.rp)
    area = sum(O.elem_space(o, e) for e in range(o.ne))
    assert abs(la.rp[1::2].sum() + 7.8 * 0.5 * 9.81 * area) <= 1e-12 * 7.8 * 0.5 * 9.81 * area
    assert la.rp[0::2].max() == 0.0 and la.rp[0::2].min() == 0.0
    # consistent loads of one element: corners -1/12, mid-sides 1/3 of its weight
    one = A.Mesh(m.x, m.y, m.conn[:1].ravel(), [], npe=8, sec=m.sec)
    l1 = A.Loads(one).gravity(1.0, 1.0)
    w = O.elem_space(o, 0)
    np.testing.assert_allclose(l1.rp[2 * m.conn[0, :4] + 1], w / 12, rtol=1e-14)
    np.testing.assert_allclose(l1.rp[2 * m.conn[0, 4:] + 1], -w / 3, rtol=1e-14)


@pytest.mark.parametrize("mesh_fn,d", [
    (lambda M: M.gen_rect(0, 1, 0, 1, 12, 8), (1.0, 1.0)),          # ties on every anti-diagonal
    (lambda M: M.gen_rect(0, 2, 0, 1, 9, 7), (0.0, 1.0)),           # ties along every row
    (lambda M: M.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8), (1.0, 1.0)),
    (lambda M: M.gen_ring(0.5, 1.0, 24, 6, order=2), (1.0, -0.5)),
])
def test_renum_by_direction_matches_oracle(mesh_fn, d):
    """Mesh::renumByDirection (Mesh.cu:537-630) against the oracle's step by
    step restatement of the reference's merge sort (renumSort,
    Mesh.cu:505-535, right run first on ties): equal numbering, also where
    many nodes have the same projection; the projections come out sorted
    and the geometry (element areas) is unchanged."""
    pm, om = mesh_fn(A), mesh_fn(O)
    (rp, perm), ro = A.renum_by_direction(pm, *d, with_perm=True), O.renum_by_direction(om, *d)
    assert np.array_equal(rp.x, ro.x) and np.array_equal(rp.y, ro.y) and np.array_equal(rp.conn, ro.conn)
    # the returned permutation: node `old` is node perm[old] afterwards
    assert np.array_equal(np.sort(perm), np.arange(pm.nn))
    assert np.array_equal(rp.x[perm], pm.x) and np.array_equal(rp.y[perm], pm.y)
    assert np.array_equal(rp.conn, perm[pm.conn])
    assert all(np.array_equal(a, b) for a, b in zip(rp.borders, ro.borders))
    key = rp.x * d[0] + rp.y * d[1]
    assert np.all(np.diff(key) >= 0)
    n_ties = int((np.diff(np.sort(pm.x * d[0] + pm.y * d[1])) == 0).sum())
    print("nodes", pm.nn, "tied neighbours", n_ties)
    orig = [O.elem_space(O.Mesh(pm.x, pm.y, pm.conn, [], npe=pm.npe), e) for e in range(pm.ne)]
    ren = [O.elem_space(O.Mesh(rp.x, rp.y, rp.conn, [], npe=pm.npe), e) for e in range(pm.ne)]
    assert np.array_equal(orig, ren)
    if pm.sec is not None:
        assert rp.sec.sum() == pm.sec.sum() and np.all(rp.sec[rp.conn[:, 4:]] == 1)


def test_renumerate_ring_matches_oracle_and_refuses_bad_counts():
    m = A.gen_rect(0, 1, 0, 1, 7, 5)                                # 8 x 6 nodes
    om = O.gen_rect(0, 1, 0, 1, 7, 5)
    r = A.renumerate_ring(m, 8)
    blen = np.array([b.size for b in om.borders], dtype=np.int32)
    x, y, conn = om.x.copy(), om.y.copy(), om.conn.copy()
    bord = np.concatenate(om.borders).astype(np.int32)
    O.lib.or_renumerate_ring(om.nn, om.ne, O.dp(x), O.dp(y), O.ip(conn), blen.size, O.ip(bord), O.ip(blen), 8)
    assert np.array_equal(r.x, x) and np.array_equal(r.conn, conn)
    with pytest.raises(A.CPError):
        A.renumerate_ring(m, 7)


@pytest.mark.parametrize("kind,args,mesh_args,periodic", [
    ("rect", (7, 5), ("gen_rect", (0, 1, 0, 1, 7, 5)), False),
    ("hole", (-3., 3., -2., 2., 0.5, 12, 8), ("gen_rect_with_hole", (-3, 3, -2, 2, 0.5, 12, 8)), True),
    ("hole", (-3., 3., -2., 2., 0.5, 96, 64), ("gen_rect_with_hole", (-3, 3, -2, 2, 0.5, 96, 64)), True),
    ("ring", (24, 5), ("gen_ring", (0.5, 1.0, 24, 5)), True)])
def test_corner_lattice_of_8_node_meshes(kind, args, mesh_args, periodic):
    """cp_mesh_*2_lattice: the corner nodes of the 8-node generators form an
    m1 x m2 lattice (direction 1 periodic around a hole or a ring): every
    element edge joins two corners one lattice step apart, and the lattice
    holds exactly the corners."""
    fn, margs = mesh_args
    mesh = getattr(A, fn)(*margs, order=2)
    m1, m2, corner_of = A.corner_lattice2(kind, *args)
    conn = mesh.conn.reshape(-1, 8)
    assert np.array_equal(np.sort(corner_of), np.unique(conn[:, :4]))
    lg = -np.ones(mesh.nn, dtype=np.int64)
    lg[corner_of] = np.arange(m1 * m2)
    c = lg[conn[:, :4]]
    i1, i2 = c // m2, c % m2
    d1 = np.roll(i1, -1, axis=1) - i1
    if periodic:
        d1 = (d1 + m1 // 2) % m1 - m1 // 2
    d2 = np.roll(i2, -1, axis=1) - i2
    assert np.all(np.abs(d1) + np.abs(d2) == 1)
    # each mid-side node belongs to one edge: every element through it has
    # the same two corners at its ends
    edges = {}
    for e in conn:
        for k in range(4):
            key = tuple(sorted((int(e[k]), int(e[(k + 1) % 4]))))
            assert edges.setdefault(int(e[4 + k]), key) == key
    assert len(edges) == mesh.nn - m1 * m2
