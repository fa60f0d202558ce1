"""The drop-in C++ API (include/cudaplasticity/, libcudaplasticity.so),
driven by tests/cpp/api_test.cpp the way the reference's experiments drive
its classes (Main.cpp).  CPU cases check meshes, material laws and the load
conditions the solver builds against the oracle; the GPU case runs the
reference's plasticityTest through PlasticitySolver::solve() with the
reference's defaults and compares with the reference's own output file."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from cudaplasticity_amd import abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "api_test")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(mode, outdir):
    assert os.path.exists(BIN), "tests/cpp/_build/api_test is not built: run __graft_entry__.build()"
    print("api_test", mode, flush=True)
    r = subprocess.run([BIN, mode, str(outdir)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.stderr


def rd(d, name):
    kind = {"f64": np.float64, "i32": np.int32, "u8": np.uint8}[name.rsplit(".", 1)[1]]
    return np.fromfile(os.path.join(d, name), dtype=kind)


def run_in(mode, outdir, indir):
    assert os.path.exists(BIN), "tests/cpp/_build/api_test is not built: run __graft_entry__.build()"
    r = subprocess.run([BIN, mode, str(outdir), str(indir)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.stderr


@pytest.fixture(scope="module")
def cpu_out(tmp_path_factory):
    d = tmp_path_factory.mktemp("api_cpu")
    src = tmp_path_factory.mktemp("api_in")
    ring = np.load(os.path.join(GOLD, "ring.npz"))
    np.ascontiguousarray(ring["points"], dtype=np.float64).tofile(os.path.join(src, "ring_points.f64"))
    np.ascontiguousarray(ring["polys"], dtype=np.int32).tofile(os.path.join(src, "ring_polys.i32"))
    from test_host_setup import write_inp
    write_inp(os.path.join(src, "plate8.inp"), A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2), "S8R")
    run_in("cpu", d, src)
    return d


def read_vtk(path):
    """Section header lines (every non-numeric line, in order) and the
    numeric blocks of a legacy VTK polydata file."""
    layout, fields, points, polys = [], {}, None, None
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln[:1].isalpha() or ln.startswith("#"):
            layout.append(ln)
            parts = ln.split()
            if parts[0] == "POINTS":
                n = int(parts[1])
                points = np.array([list(map(float, lines[i + 1 + k].split()))[:2] for k in range(n)])
                i += n + 1
                continue
            if parts[0] == "POLYGONS":
                n = int(parts[1])
                polys = np.array([list(map(int, lines[i + 1 + k].split()))[1:] for k in range(n)])
                i += n + 1
                continue
            if len(parts) == 4 and parts[3] == "float":
                fields[parts[0]] = np.array(list(map(float, lines[i + 1].split())))
                i += 2
                continue
        i += 1
    return layout, fields, points, polys


def layouts():
    import json
    with open(os.path.join(GOLD, "vtk_layout.json")) as f:
        return json.load(f)


def sigma_errors_restated(mesh, sx, sy, txy, max_r):
    """PlasticitySolver::sigmaError and polarSigmaError of the reference for
    the test functions of api_test.cpp: the largest error at the element
    centres (mean of the 4 corners) of each component, times the mesh area
    over that component's area-weighted integral (signed for the Cartesian
    measure, of absolute values for the polar one, which keeps elements with
    centre radius <= max_r and takes the functions of (r, phi))."""
    conn = mesh.conn
    X, Y = mesh.x[conn], mesh.y[conn]
    cx, cy = X.mean(axis=1), Y.mean(axis=1)
    # bilinear quad area by the shoelace formula (exact for straight edges)
    area = 0.5 * np.abs((X[:, 0] * Y[:, 1] - X[:, 1] * Y[:, 0]) + (X[:, 1] * Y[:, 2] - X[:, 2] * Y[:, 1]) +
                        (X[:, 2] * Y[:, 3] - X[:, 3] * Y[:, 2]) + (X[:, 3] * Y[:, 0] - X[:, 0] * Y[:, 3]))
    fx = lambda a, b: 100. + 40. * a - 7. * b
    fy = lambda a, b: 30. + 12. * b * b
    fxy = lambda a, b: -5. + 9. * a * b
    vol = area.sum()
    cart = max(np.abs(f(cx, cy) - s).max() * vol / (s * area).sum() for f, s in ((fx, sx), (fy, sy), (fxy, txy)))
    r = np.hypot(cx, cy)
    phi = np.arctan(cy / cx) + np.where(cx < 0, np.pi, 0.0)
    keep = r <= max_r
    co, si = cx / r, cy / r
    srr = co * co * sx + si * si * sy + 2 * co * si * txy
    spp = si * si * sx + co * co * sy - 2 * co * si * txy
    srp = co * si * (sy - sx) + txy * (co * co - si * si)
    pv = area[keep].sum()
    polar = max(np.abs(f(r, phi) - s)[keep].max() * pv / (np.abs(s) * area)[keep].sum()
                for f, s in ((fx, srr), (fy, spp), (fxy, srp)))
    return np.array([cart, polar])


def test_mesh_analysis_and_vtk_writer(cpu_out):
    """Mesh::printAnalysis + Mesh::saveAsVtk (Mesh.cu:939-989, 1032-1080) on
    the points and quads of the reference's data/ring.vtk: the same section
    layout, and the quality columns the reference printed (6 digits)."""
    fx = np.load(os.path.join(GOLD, "ring.npz"))
    lay, F, pts, polys = read_vtk(os.path.join(cpu_out, "ring_mesh.vtk"))
    assert lay == layouts()["ring"]
    assert np.array_equal(polys, fx["polys"])
    assert np.abs(pts - fx["points"]).max() < 1e-6
    np.testing.assert_allclose(F["Element_space"], fx["f_Element_space"], rtol=1e-4)
    np.testing.assert_allclose(F["Element_aspect_ratio"], fx["f_Element_aspect_ratio"], rtol=1e-4)
    np.testing.assert_allclose(F["Element_skew_angle"], fx["f_Element_skew_angle"], atol=1e-4)


def check_mesh(d, tag, om):
    xy = rd(d, f"{tag}_xy.f64").reshape(-1, 2)
    assert np.array_equal(xy[:, 0], om.x) and np.array_equal(xy[:, 1], om.y)
    assert np.array_equal(rd(d, f"{tag}_elem4.i32").reshape(-1, 4), om.conn)
    meta = rd(d, f"{tag}_meta.i32")
    assert meta[0] == om.nn and meta[1] == om.ne and meta[2] == len(om.borders)
    assert meta[3] == O.lib.or_max_index_diff(om.ne, O.ip(om.conn))
    q = rd(d, f"{tag}_quality.f64").reshape(-1, 3)
    area = np.array([O.elem_space4(om, e) for e in range(om.ne)])
    assert np.array_equal(q[:, 0], area)
    P = np.stack([om.x[om.conn], om.y[om.conn]], axis=-1)          # ne x 4 x 2
    sides = np.linalg.norm(np.roll(P, -1, axis=1) - P, axis=-1)
    np.testing.assert_allclose(q[:, 1], sides.max(1) / sides.min(1), rtol=1e-14)
    v1 = 0.5 * (P[:, 0] + P[:, 1] - P[:, 2] - P[:, 3])
    v2 = 0.5 * (P[:, 1] + P[:, 2] - P[:, 3] - P[:, 0])
    d12 = (v1 * v2).sum(1)
    np.testing.assert_allclose(q[:, 2], 1 - d12 ** 2 / ((v1 * v1).sum(1) * (v2 * v2).sum(1)), rtol=1e-12,
                               atol=1e-15)


def check_conditions(d, tag, prob):
    assert np.array_equal(rd(d, f"{tag}_rp.f64"), prob.rp), tag
    assert np.array_equal(rd(d, f"{tag}_mask.u8"), prob.mask), tag
    assert np.array_equal(rd(d, f"{tag}_u0.f64"), prob.u), tag


def test_renumberings_and_crossz_through_cpp_api(cpu_out):
    """Mesh::renumByDirection() with the default direction {1, 1} (as
    Main.cpp:1029 calls it), Mesh::renumerateRing and vec2::crossZ through
    the C++ classes, against the oracle's restatements."""
    om = O.renum_by_direction(O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8), 1.0, 1.0)
    assert np.array_equal(rd(cpu_out, "renum_xy.f64"), np.c_[om.x, om.y].ravel())
    assert np.array_equal(rd(cpu_out, "renum_elem4.i32"), om.conn.ravel())
    r = O.gen_rect(0, 1, 0, 1, 7, 5)
    blen = np.array([b.size for b in r.borders], dtype=np.int32)
    x, y, conn, bord = r.x.copy(), r.y.copy(), r.conn.copy(), np.concatenate(r.borders).astype(np.int32)
    O.lib.or_renumerate_ring(r.nn, r.ne, O.dp(x), O.dp(y), O.ip(conn), blen.size, O.ip(bord), O.ip(blen), 8)
    assert np.array_equal(rd(cpu_out, "ringrenum_xy.f64"), np.c_[x, y].ravel())
    assert np.array_equal(rd(cpu_out, "ringrenum_elem4.i32"), conn.ravel())
    assert rd(cpu_out, "crossz.f64").tolist() == [3.0 * 4.0 - (-2.0) * 0.5, 0.5 * -2.0 - 4.0 * 3.0, 0.0]


def test_hole_plate_mesh_and_conditions(cpu_out):
    """plasticityTest's mesh (Main.cpp:256) and loads (:266-272)."""
    om = O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 48, 32)
    check_mesh(cpu_out, "hole", om)
    check_mesh(cpu_out, "holecopy", om)
    fx = np.load(os.path.join(GOLD, "holePlatePlastPow.npz"))
    assert np.array_equal(rd(cpu_out, "hole_elem4.i32").reshape(-1, 4), fx["polys"])
    check_conditions(cpu_out, "hole",
                     O.Problem(om).force(1, 170.0, 0.0).force(3, -170.0, 0.0).fix_vert_axis(0).fix_hor_axis(0))


def test_ring_function_force(cpu_out):
    """testRing (Main.cpp:75-106): setForce with a lambda, thickness 0.1."""
    om = O.gen_ring(0.1, 0.2, 40, 5)
    check_mesh(cpu_out, "ring", om)
    b = om.borders[0]
    fv = np.stack([1e6 * np.abs(om.x[b]), np.zeros(b.size)], axis=1)
    check_conditions(cpu_out, "ring", O.Problem(om, h=0.1).force_values(0, fv).fix_vert_axis(0).fix_hor_axis(0))


def test_gravity_and_displacements(cpu_out):
    """diagramTest level (Main.cpp:929-994): setGravity + setDisplacement."""
    om = O.gen_rect(-1, 1, -1, 1, 2, 2)
    check_mesh(cpu_out, "diag", om)
    check_conditions(cpu_out, "diag", O.Problem(om).gravity().displacement(0, 0.0015, 0.0, False, True)
                     .displacement(2, 0.0, 0.0, False, True))


def test_normal_pressure_quarter_ring(cpu_out):
    om = O.gen_quarter_ring(0.5, 1.0, 48, 12, 1.05)
    check_mesh(cpu_out, "qring", om)
    check_conditions(cpu_out, "qring",
                     O.Problem(om).force(0, -180.0, 0.0, norm_orient=True).fix_vert_axis(0).fix_hor_axis(0))


def test_points_fixed_border_rotation_function_displacement(cpu_out):
    """setRotation, fixBorderVert, setForcePoint, setDisplPoint, fixPoint and
    a displacement lambda (LoadConditions.cuh:69-179; their effect on the
    load vector and mask as PlasticitySolver.cu:895-1003 applies them)."""
    om = O.gen_rect(0, 1, 0, 1, 4, 4)
    check_mesh(cpu_out, "pts", om)
    nn = om.nn
    rp = np.zeros(2 * nn)
    u = np.zeros(2 * nn)
    mask = np.ones(2 * nn, dtype=np.uint8)
    rho, h, w2, axis = 3.0, 0.5, 4.0, np.array([0.5, 0.0])
    for e in range(om.ne):
        c = om.conn[e]
        wgt = 0.25 * rho * h * O.elem_space4(om, e)
        for n in c:
            R = w2 * (np.array([om.x[n], om.y[n]]) - axis)
            rp[2 * n] += wgt * R[0]
            rp[2 * n + 1] += wgt * R[1]
    b2 = om.borders[2]
    end2 = b2.size - (1 if b2[0] == b2[-1] else 0)
    for n in b2[:end2]:
        u[2 * n], u[2 * n + 1] = 0.01 * om.x[n], -0.02 * om.x[n] * om.x[n]
        mask[2 * n] = mask[2 * n + 1] = 0
    b3 = om.borders[3]
    end3 = b3.size - (1 if b3[0] == b3[-1] else 0)
    mask[2 * b3[:end3]] = 0
    node = lambda x, y: int(np.flatnonzero((om.x == x) & (om.y == y))[0])
    n11 = node(1.0, 1.0)
    rp[2 * n11] += 3.0
    rp[2 * n11 + 1] += 4.0
    for (x, y), val in (((0.0, 0.0), (0.1, 0.2)), ((1.0, 0.0), (0.0, 0.0))):
        n = node(x, y)
        u[2 * n], u[2 * n + 1] = val
        mask[2 * n] = mask[2 * n + 1] = 0
    np.testing.assert_allclose(rd(cpu_out, "pts_rp.f64"), rp, rtol=1e-14, atol=1e-15)
    assert np.array_equal(rd(cpu_out, "pts_mask.u8"), mask)
    assert np.array_equal(rd(cpu_out, "pts_u0.f64"), u)


def test_material_laws_on_host(cpu_out):
    got = rd(cpu_out, "laws.f64").reshape(-1, 3)
    mats = [O.material(2.1e5, 0.3, 240.0, O.LAW_LINEAR, K_T=2.1e5 / 101),
            O.material(193e3, 0.3, 170.0, O.LAW_POWER, A=1930.0, m=0.015),
            O.material(2.1e5, 0.3, 240.0, O.LAW_SWIFT, n=0.2)]
    for i, eps in enumerate([1e-4, 1e-3, 2e-3, 5e-3, 2e-2, 0.1]):
        for j, m in enumerate(mats):
            assert got[i, j] == O.material_f(m, eps), (i, j)


def test_unsupported_paths_refuse(cpu_out):
    """solveCPU, the 4-node mesh analysis of an 8-node mesh and multigrid on
    a mesh with no lattice throw instead of silently degrading."""
    assert rd(cpu_out, "refused.i32")[0] == 3


def lattice_cells_ok(n1, n2, periodic, nodes, elem4):
    """Every quad occupies one cell of the lattice: its corners sit at
    (a, b), (a + 1, b), (a, b + 1), (a + 1, b + 1), direction 1 taken modulo
    n1 when it is periodic.  Returns the number of quads that do not."""
    nn = nodes.size
    assert n1 * n2 == nn and np.array_equal(np.sort(nodes), np.arange(nn))
    lg = np.empty(nn, dtype=np.int64)
    lg[nodes] = np.arange(nn)
    i1, i2 = lg[elem4] // n2, lg[elem4] % n2
    bad = 0
    for a, b in zip(i1, i2):
        if np.ptp(b) != 1 or sorted(np.bincount(b - b.min()).tolist()) != [2, 2]:
            bad += 1
            continue
        lo = a.min()
        d = (a - lo) % n1 if periodic else a - lo
        if periodic and d.max() > 1:   # the cell across the seam
            d = (a - a.max()) % n1
        if sorted(d.tolist()) != [0, 0, 1, 1]:
            bad += 1
    return bad


def test_mesh_lattices_through_cpp_api(cpu_out):
    """Mesh records the lattice of the structured meshes it generates
    (multigrid's input): rectangles and quarter meshes node by node, plates
    with a hole and rings as 2 (N1 + N2) or N_phi rays closing around the
    hole; renumByDirection and renumerateRing carry it along.  Checked by
    the defining property: every quad is one lattice cell."""
    want = {"lat_hole": (160, 33, 1), "lat_ring": (40, 6, 1), "lat_rect": (10, 5, 0),
            "lat_qplate": (17, 8, 0), "lat_qring": (13, 7, 0)}
    for tag, dims in want.items():
        n1, n2, per = rd(cpu_out, f"{tag}_lattice_meta.i32").tolist()
        assert (n1, n2, per) == dims, tag
        nodes = rd(cpu_out, f"{tag}_lattice.i32")
        elem4 = rd(cpu_out, f"{tag}_elem4.i32").reshape(-1, 4)
        assert lattice_cells_ok(n1, n2, per, nodes, elem4) == 0, tag
        if not per:
            assert np.array_equal(nodes, np.arange(nodes.size)), tag
    for tag in ("renum", "ringrenum"):
        n1, n2, per = rd(cpu_out, f"{tag}_lattice_meta.i32").tolist()
        nodes = rd(cpu_out, f"{tag}_lattice.i32")
        elem4 = rd(cpu_out, f"{tag}_elem4.i32").reshape(-1, 4)
        assert n1 > 0 and lattice_cells_ok(n1, n2, per, nodes, elem4) == 0, tag
        assert not np.array_equal(nodes, np.arange(nodes.size)), tag
    # a hole plate's seam lies inside the lattice: the property is not vacuous
    n1, n2, per = want["lat_hole"]
    nodes = rd(cpu_out, "lat_hole_lattice.i32")
    elem4 = rd(cpu_out, "lat_hole_elem4.i32").reshape(-1, 4)
    assert lattice_cells_ok(n1, n2, 0, nodes, elem4) == n2 - 1
    # an 8-node plate (order 2): the lattice of its corner nodes, periodic
    assert rd(cpu_out, "lat_hole2_lattice_meta.i32").tolist() == [20, 4, 1]
    corners = rd(cpu_out, "lat_hole2_lattice.i32")
    m = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 6, 4, order=2)
    assert np.array_equal(np.sort(corners), np.unique(m.conn.reshape(-1, 8)[:, :4]))


def test_load_from_file_through_cpp_api(cpu_out):
    """Mesh::loadFromFile (Mesh.cu:725-937) on an .inp deck of the order-2
    hole plate; a wrong extension and a missing file throw const char*."""
    m = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2)
    assert np.array_equal(rd(cpu_out, "inp_xy.f64"), np.stack([m.x, m.y], 1).ravel())
    assert np.array_equal(rd(cpu_out, "inp_sec.u8"), m.sec)
    assert np.array_equal(rd(cpu_out, "inp_elem8.i32"), m.conn.ravel())
    assert list(rd(cpu_out, "inp_blen.i32")) == [b.size for b in m.borders]
    assert np.array_equal(rd(cpu_out, "inp_border1.i32"), m.borders[1])
    assert list(rd(cpu_out, "inp_counts.i32")) == [0, 0, m.ne, m.ne, 2]


@pytest.mark.parametrize("name", ["rect1", "hole1", "ring1"])
def test_order2_meshes_through_cpp_api(cpu_out, name):
    """Mesh::genRectangle / genRectWithHole / genRing with order 2 and
    Mesh::saveAsVtk write the reference's data/rect1.vtk, hole1.vtk and
    ring1.vtk meshes: same polygons, points to the printed digits."""
    g = np.load(os.path.join(GOLD, "mesh_%s.npz" % name))
    lay, _, pts, polys = read_vtk(os.path.join(cpu_out, name + ".vtk"))
    assert np.array_equal(polys, g["polys"])
    assert np.abs(pts - g["points"]).max() <= 5e-6 * max(1.0, np.abs(g["points"]).max())
    assert "POLYGONS %d %d" % (g["polys"].shape[0], 9 * g["polys"].shape[0]) in lay


def test_order2_conditions_through_cpp_api(cpu_out):
    om = O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2)
    op = O.Problem(om, h=0.7).force(1, 170.0, 3.0).force(3, -170.0, 0.0).force(4, 5.0, 1.0, norm_orient=True)
    op.fix_vert_axis(0.0).fix_hor_axis(0.0)
    assert np.array_equal(rd(cpu_out, "hole1_sec.u8"), om.sec)
    assert np.array_equal(rd(cpu_out, "hole1_rp.f64"), op.rp)
    assert np.array_equal(rd(cpu_out, "hole1_mask.u8"), op.mask)
    assert np.array_equal(rd(cpu_out, "hole1copy_elem8.i32"), om.conn.ravel())
    avr, maxdiff, s0 = rd(cpu_out, "hole1_q8.f64")
    areas = np.array([O.elem_space(om, e) for e in range(om.ne)])
    assert abs(avr - areas.mean()) <= 1e-12 * avr and abs(s0 - areas[0]) <= 1e-12 * s0
    assert maxdiff == max(int(np.ptp(c)) for c in om.conn)


@pytest.mark.gpu
def test_order2_energy_through_cpp_api(tmp_path):
    """energyConvTest on the 660-element 8-node plate through
    PlasticitySolver::solve() with the reference's defaults and
    deformEnergy(): data/aitken__plast2.txt prints 2.45491e+06."""
    run("gpu", tmp_path)
    en, ne, status = rd(tmp_path, "q8_energy.f64")
    print("8-node energy %.1f on %d elements" % (en, ne), flush=True)
    assert ne == 660 and status == 0
    assert abs(en - 2.45491e6) <= 5.0
    lay, VF, _, polys = read_vtk(os.path.join(tmp_path, "q8_plate.vtk"))
    assert polys.shape == (660, 8) and VF["Sigma_intensity"].size == 660
    # SolverOptions::multigrid on the 8-node mesh (its corner lattice, which
    # genRectWithHole records): the same table entry, far fewer CG iterations
    enm, stm, cgm, cg0 = rd(tmp_path, "q8mg_energy.f64")
    print("8-node energy with multigrid %.1f, CG %d (plain CG %d)" % (enm, cgm, cg0), flush=True)
    assert stm == 0 and abs(enm - 2.45491e6) <= 5.0
    assert cgm * 10 < cg0


@pytest.mark.gpu
def test_plasticity_test_through_cpp_api(tmp_path):
    """The reference's plasticityTest (Main.cpp:252-296) through
    PlasticitySolver::solve() with the reference's defaults reproduces
    data/holePlatePlastPow.vtk; config-1 elastic solve within 1e-8 of the
    oracle; load steps with the additive options equal the C-ABI path."""
    run("gpu", tmp_path)
    fx = np.load(os.path.join(GOLD, "holePlatePlastPow.npz"))
    om = O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 48, 32)
    u = rd(tmp_path, "plast_u.f64")
    F = rd(tmp_path, "plast_fields.f64").reshape(len(A.FIELDS), -1)
    st = rd(tmp_path, "plast_stats.f64")
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(u, fx["f_Displacement"][:, :2].reshape(-1)) < 2e-7
    for k, n in [(O.FIELDS.index("sxx"), "Sigma_xx"), (O.FIELDS.index("sigI"), "Sigma_intensity"),
                 (O.FIELDS.index("psi"), "Psi"), (O.FIELDS.index("epsI_p"), "Eps_p_intensity")]:
        assert rel(F[k], fx["f_" + n]) < 2e-7, n
    E = 193e9 * 1e-6
    mat = O.material(E, 0.29, 170.0, O.LAW_POWER, A=E * 0.01, m=0.015)
    flags = F[O.FIELDS.index("epsI")] >= mat.sigmaT / mat.E
    assert flags.sum() == 3528
    assert np.array_equal(flags, fx["f_Eps_intensity"] >= mat.sigmaT / mat.E)
    assert st[2] == 0 and abs(st[0] - 41) <= 1 and st[1] <= 1e-5
    assert abs(st[3] - O.deform_energy(om, mat, F)) <= 1e-12 * abs(st[3])
    assert st[4] == F[O.FIELDS.index("sxx"), 0] and st[5] == F[O.FIELDS.index("exx"), 0]
    assert os.path.getsize(os.path.join(tmp_path, "plast_displs.txt")) > 0

    # saveAsVtk (PlasticitySolver.cu:1784-1945): same layout as the reference's files
    lay, VF, pts, polys = read_vtk(os.path.join(tmp_path, "holePlatePlastPow.vtk"))
    assert lay == layouts()["holePlatePlastPow"]
    assert np.array_equal(polys, fx["polys"]) and np.abs(pts - fx["points"]).max() < 2e-5
    for name in ("Eps_xx", "Gamma_xy", "Eps_p_intensity", "Psi", "E_c", "nu_c", "Sigma_xx", "Tau_xy",
                 "Sigma_intensity"):
        assert rel(VF[name], fx["f_" + name]) < 2e-7, name
    assert rel(VF["Displacement"].reshape(-1, 3)[:, :2].ravel(), fx["f_Displacement"][:, :2].ravel()) < 2e-7
    # SolverOptions::multigrid: the plate's periodic lattice, then the plate
    # renumbered by direction, FP64 and mixed: the fixture to the same bar.
    key = om.x + om.y
    old_of_new = np.lexsort((-np.arange(om.nn), key))
    ro = O.renum_by_direction(om, 1.0, 1.0)
    assert np.array_equal(om.x[old_of_new], ro.x) and np.array_equal(om.y[old_of_new], ro.y)
    u_fx = fx["f_Displacement"][:, :2]
    for k in range(3):
        um = rd(tmp_path, f"plastmg{k}_u.f64").reshape(-1, 2)
        Fm = rd(tmp_path, f"plastmg{k}_fields.f64").reshape(len(A.FIELDS), -1)
        sm = rd(tmp_path, f"plastmg{k}_stats.f64")
        ref_u = u_fx if k == 0 else u_fx[old_of_new]
        errs = [rel(um.ravel(), ref_u.ravel())] + [rel(Fm[O.FIELDS.index(f)], fx["f_" + n]) for f, n in
                                                   [("sxx", "Sigma_xx"), ("sigI", "Sigma_intensity"),
                                                    ("psi", "Psi"), ("epsI_p", "Eps_p_intensity")]]
        fl = Fm[O.FIELDS.index("epsI")] >= mat.sigmaT / mat.E
        print("multigrid %d: outer %d rel_err %.3g cg %d, %.2f s; errors %s, plastic %d (%d differ)"
              % (k, sm[0], sm[1], sm[3], sm[4], " ".join("%.2e" % e for e in errs), fl.sum(),
                 (fl != flags).sum()), flush=True)
        assert sm[2] == 0 and abs(sm[0] - 41) <= 1 and sm[1] <= 1e-5
        assert max(errs) < 1e-6
        assert fl.sum() == 3528 and np.array_equal(fl, flags)
        assert sm[3] < 41 * 20

    play, PF, _, _ = read_vtk(os.path.join(tmp_path, "holePlatePlastPow_polar.vtk"))
    rename = {"Eps_xx": "Eps_rr", "Eps_yy": "Eps_phiphi", "Gamma_xy": "Gamma_rphi", "Eps_p_xx": "Eps_p_rr",
              "Eps_p_yy": "Eps_p_phiphi", "Gamma_p_xy": "Gamma_p_rphi", "Sigma_xx": "Sigma_rr",
              "Sigma_yy": "Sigma_phiphi", "Tau_xy": "Tau_rphi"}
    want = [" ".join([rename.get(w.split()[0], w.split()[0])] + w.split()[1:]) for w in lay]
    assert play == want
    P = np.stack([om.x[om.conn], om.y[om.conn]], axis=-1).sum(1) / 4
    c, s_ = P[:, 0] / np.hypot(P[:, 0], P[:, 1]), P[:, 1] / np.hypot(P[:, 0], P[:, 1])
    sx, sy, txy = VF["Sigma_xx"], VF["Sigma_yy"], VF["Tau_xy"]
    np.testing.assert_allclose(PF["Sigma_rr"], c * c * sx + s_ * s_ * sy + 2 * c * s_ * txy, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(PF["Tau_rphi"], c * s_ * (sy - sx) + txy * (c * c - s_ * s_), rtol=1e-12, atol=1e-9)
    ex, ey, g = VF["Eps_xx"], VF["Eps_yy"], VF["Gamma_xy"]
    np.testing.assert_allclose(PF["Gamma_rphi"], c * s_ * (ey - ex) + 0.5 * g * (c * c - s_ * s_), rtol=1e-12,
                               atol=1e-15)

    # solverTest (Main.cpp:14-38) through solveElast + saveAsVtk against data/solverTest.vtk
    st_fx = np.load(os.path.join(GOLD, "solverTest.npz"))
    lay, SF, pts, polys = read_vtk(os.path.join(tmp_path, "solverTest.vtk"))
    assert lay == layouts()["solverTest"]
    assert np.array_equal(polys, st_fx["polys"]) and np.array_equal(pts, st_fx["points"])
    for name in ("Eps_xx", "Sigma_xx", "Sigma_intensity"):
        assert rel(SF[name], st_fx["f_" + name]) < 1e-9, name
    # nu = 2.9e-7 here: the lateral strain is 1e-6 of the axial one, so it is
    # measured against the axial strain's size
    assert np.linalg.norm(SF["Eps_yy"] - st_fx["f_Eps_yy"]) < 1e-9 * np.linalg.norm(st_fx["f_Eps_xx"])
    assert np.abs(SF["Sigma_xx"] - 1.0).max() < 1e-9

    # config-1 elastic
    qm = O.gen_quarter_plate(1.0, 0.1, 100, 49)
    qp = O.Problem(qm).force(1, 100.0, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    uo, so, _ = O.solve(qp, O.material(2.1e5, 0.3), O.options(cg_tol=1e-10, elastic=True))
    assert rel(rd(tmp_path, "elast_u.f64"), uo) < 1e-8
    assert rel(rd(tmp_path, "elast_sxx.f64"), O.fields(so)["sxx"]) < 1e-8
    assert rd(tmp_path, "elast_polar_err.f64")[0] < 0.2   # Kirsch (infinite plate) near the hole
    # fixBorderVert / fixBorderHor: the exact uniform state of uniaxial tension
    pu = rd(tmp_path, "patch_u.f64")
    right, left, bottom = pu[:3].astype(int)
    assert len({right, left, bottom}) == 3 and min(right, left, bottom) >= 0
    X, Y, UX, UY = pu[3:].reshape(-1, 4).T
    assert np.abs(UX - 50.0 * X / 2.1e5).max() <= 1e-9 * 50.0 * 2.0 / 2.1e5
    assert np.abs(UY + 0.3 * 50.0 * Y / 2.1e5).max() <= 1e-9 * 50.0 * 2.0 / 2.1e5
    ue = rd(tmp_path, "patch_u_strain.f64")                 # setPlainCondition(strain)
    assert np.abs(ue[0::2] - 50.0 * (1 - 0.09) * X / 2.1e5).max() <= 1e-9 * 50.0 * 2.0 / 2.1e5
    assert np.abs(ue[1::2] + 50.0 * 0.3 * 1.3 * Y / 2.1e5).max() <= 1e-9 * 50.0 * 2.0 / 2.1e5
    # fixBorder + setVolumeForce == setGravity bit for bit; solveCUDA_FD within 1e-6
    g, v, fd = rd(tmp_path, "vol_gravity.f64"), rd(tmp_path, "vol_volume.f64"), rd(tmp_path, "vol_fd.f64")
    assert np.array_equal(g, v)
    assert np.abs(g).max() > 0
    assert rel(fd, g) < 1e-6
    # sigmaError / polarSigmaError against their definitions
    # (PlasticitySolver.cuh:464-494, 497-541) restated over the same fields
    sx, sy, txy = rd(tmp_path, "elast_sxx.f64"), rd(tmp_path, "elast_syy.f64"), rd(tmp_path, "elast_tau.f64")
    got = rd(tmp_path, "elast_sigma_errors.f64")
    want = sigma_errors_restated(A.gen_quarter_plate(1.0, 0.1, 100, 49), sx, sy, txy, 0.6)
    assert np.allclose(got, want, rtol=1e-12, atol=0), (got, want)

    # additive API == C-ABI path
    pm = A.gen_quarter_plate(1.0, 0.1, 40, 20)
    pl = A.Loads(pm).force(1, 300.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    ctx = A.Context(0)
    P = A.Problem(ctx, pm, pl, A.material(2.1e5, 0.3, 240.0, A.LAW_LINEAR, K_T=2.1e5 / 101))
    P.solve_steps(4, 1.0, A.options(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=1000, precision=A.MIXED,
                                    precond=True, predictor=1, cg_forcing=0.1, inner_tol=1e-4))
    assert np.array_equal(rd(tmp_path, "steps_u.f64"), P.field("u"))
    assert np.array_equal(rd(tmp_path, "steps_sigi.f64"), P.field("sigI"))
    P.close()
    ctx.close()
