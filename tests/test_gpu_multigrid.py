"""Multigrid-preconditioned CG (precond = 2) on the GPU through the C-ABI:
the same solutions as the oracle and as block-Jacobi CG within the solve
tolerances, far fewer iterations, deterministic, and refusals where the
lattice does not fit."""
import numpy as np
import pytest

import oracle_lib as O
from cudaplasticity_amd import abi as A

pytestmark = pytest.mark.gpu

E = 2.1e5


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def plate(ctx, Nt, Nr, S=300.0, sY=240.0):
    mesh = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    mat = A.material(E, 0.3, sY, A.LAW_LINEAR, K_T=E / 101)
    P = A.Problem(ctx, mesh, loads, mat)
    P.set_grid(Nt + 1, Nr + 1)
    return mesh, loads, mat, P


def test_elastic_plate_matches_oracle(ctx):
    """config-1 mesh, one elastic solve to 1e-10: u and sigma within 1e-8 of
    the oracle; a few dozen iterations instead of thousands."""
    mesh, loads, mat, P = plate(ctx, 100, 49, S=100.0, sY=1e300)
    print("levels", P.mg_levels(), flush=True)
    assert len(P.mg_levels()) >= 3
    st = P.solve(A.options(cg_tol=1e-10, precond=2), elastic=True)
    st_bj = A.Problem(ctx, mesh, loads, mat).solve(A.options(cg_tol=1e-10, precond=1), elastic=True)
    om = O.gen_quarter_plate(1.0, 0.1, 100, 49)
    ol = O.Problem(om).force(1, 100.0, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    uo, so, _ = O.solve(ol, O.material(E, 0.3), O.options(cg_tol=1e-10, elastic=True))
    print("mg iterations", st.cg_iters, "block-Jacobi", st_bj.cg_iters, "smoothers", P.mg_smoothers(), flush=True)
    assert rel(P.field("u"), uo) < 1e-8
    assert rel(P.field("sxx"), O.fields(so)["sxx"]) < 1e-8
    assert st.cg_iters * 10 < st_bj.cg_iters
    # the cells near the hole are stretched along the lattice direction 1
    # (about 1:16): the segment smoother keeps the V-cycle's rate there
    assert P.mg_smoothers()[0] == 1
    assert st.cg_iters <= 30


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_plastic_steps_match_block_jacobi(ctx, prec):
    """Load steps with Aitken: the multigrid run reaches the same converged
    levels as the block-Jacobi run (same fixed point, tolerances 1e-9/1e-6)."""
    mesh, loads, mat, P = plate(ctx, 60, 30)
    opt = dict(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=500, precision=prec, predictor=1, cg_forcing=0.1,
               inner_tol=1e-4)
    per = P.solve_steps(6, 1.0, A.options(precond=2, **opt))
    Q = A.Problem(ctx, mesh, loads, mat)
    per_bj = Q.solve_steps(6, 1.0, A.options(precond=1, **opt))
    print("mg", [(s.outer_iters, s.cg_iters) for s in per], "bj", [(s.outer_iters, s.cg_iters) for s in per_bj],
          flush=True)
    assert all(s.status == A.CP_OK for s in per)
    assert rel(P.field("u"), Q.field("u")) < 2e-6
    assert rel(P.field("sigI"), Q.field("sigI")) < 2e-6
    assert sum(s.cg_iters for s in per) * 5 < sum(s.cg_iters for s in per_bj)


def test_graded_ring_and_rectangle(ctx):
    """Quarter ring (radially graded) under pressure and a rectangle with
    displacement-driven loading, multigrid against the oracle."""
    for kind in ("ring", "rect"):
        if kind == "ring":
            m, om = A.gen_quarter_ring(0.5, 1.0, 64, 16, 1.05), O.gen_quarter_ring(0.5, 1.0, 64, 16, 1.05)
            n1, n2 = 65, 17
            l = A.Loads(m).force(0, -150.0, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
            ol = O.Problem(om).force(0, -150.0, 0.0, norm_orient=True).fix_vert_axis(0).fix_hor_axis(0)
        else:
            m, om = A.gen_rect(0, 2, 0, 1, 40, 20), O.gen_rect(0, 2, 0, 1, 40, 20)
            n1, n2 = 41, 21
            l = A.Loads(m).displacement(0, 0.0, 0.0).displacement(2, 2e-3, 1e-4)
            ol = O.Problem(om).displacement(0, 0.0, 0.0).displacement(2, 2e-3, 1e-4)
        mat = A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)
        P = A.Problem(ctx, m, l, mat)
        P.set_grid(n1, n2)
        st = P.solve(A.options(cg_tol=1e-12, fixed_outer=6, precond=2))
        uo, so, _ = O.solve(ol, O.material(E, 0.3, 240.0, O.LAW_LINEAR, K_T=E / 101),
                            O.options(cg_tol=1e-12, fixed_outer=6))
        print(kind, "cg", st.cg_iters, "rel u", rel(P.field("u"), uo), flush=True)
        assert rel(P.field("u"), uo) < 1e-8
        assert rel(P.field("sigI"), O.fields(so)["sigI"]) < 1e-8


def test_deterministic(ctx):
    out = []
    for _ in range(2):
        mesh, loads, mat, P = plate(ctx, 48, 24)
        P.solve_steps(3, 1.0, A.options(cg_tol=1e-9, tol=1e-6, aitken=True, precision=A.MIXED, precond=2,
                                        predictor=1, cg_forcing=0.1))
        out.append((P.field("u"), P.field("sigI")))
        P.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_refusals(ctx):
    mesh, loads, mat, P = plate(ctx, 20, 10)
    with pytest.raises(A.CPError) as e:
        P.set_grid(20, 11)                      # wrong node count
    assert e.value.code == A.CP_ERR_INVALID
    hm = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 24, 16)     # renumbered: not the identity lattice
    H = A.Problem(ctx, hm, A.Loads(hm).force(1, 100.0, 0.0).fix_vert_axis(0).fix_hor_axis(0), mat)
    with pytest.raises(A.CPError) as e:
        H.set_grid(16, hm.nn // 16)
    assert e.value.code == A.CP_ERR_INVALID
    Q = A.Problem(ctx, mesh, loads, mat)        # no lattice
    with pytest.raises(A.CPError) as e:
        Q.solve(A.options(precond=2))
    assert e.value.code == A.CP_ERR_STATE
    srv = A.Server(ctx, 1 << 30)
    with pytest.raises(A.CPError) as e:
        srv.submit(mesh, loads, mat, 1, 1.0, A.options(precond=2))
    assert e.value.code == A.CP_ERR_INVALID
    srv.close()


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_graph_replay_equals_eager(ctx, prec):
    """The multigrid CG iteration replayed as a HIP graph gives bitwise the
    eager launches' results.  Four load steps replay it hundreds of times,
    with assembly, refinement residuals, return maps and host
    synchronisations between the replays."""
    out = []
    for graphs in (False, True):
        mesh, loads, mat, P = plate(ctx, 60, 30)
        P.set_graphs(graphs)
        per = P.solve_steps(4, 1.0, A.options(cg_tol=1e-9, tol=1e-6, aitken=True, precision=prec, precond=2,
                                              predictor=1, cg_forcing=0.1, inner_tol=1e-4))
        out.append((P.field("u"), P.field("sigI"), [(s.outer_iters, s.cg_iters) for s in per]))
        P.close()
    print("eager", out[0][2], "graph", out[1][2], flush=True)
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_permuted_lattice_segment_smoother(ctx):
    """The same plate with its nodes renumbered at random: the lattice is
    handed over through node_of, the segment smoother and the Galerkin
    levels gather through it, and the solution is the renumbered one."""
    Nt, Nr = 96, 48
    m = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    perm = np.random.default_rng(7).permutation(m.nn).astype(np.int32)   # old node -> new node
    x, y = np.empty(m.nn), np.empty(m.nn)
    x[perm], y[perm] = m.x, m.y
    pm = A.Mesh(x, y, perm[m.conn], [perm[b] for b in m.borders])
    mat = A.material(E, 0.3)
    out = []
    for mesh, node_of in ((m, None), (pm, perm)):
        loads = A.Loads(mesh).force(1, 100.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(Nt + 1, Nr + 1, node_of)
        st = P.solve(A.options(cg_tol=1e-10, precond=2), elastic=True)
        print("cg", st.cg_iters, "smoothers", P.mg_smoothers(), flush=True)
        assert P.mg_smoothers()[:2] == [1, 1]
        out.append((P.field("u").reshape(-1, 2), st.cg_iters))
        P.close()
    u0, u1 = out[0][0], out[1][0][perm]
    assert rel(u1, u0) < 1e-8
    assert abs(out[0][1] - out[1][1]) <= 2


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_coarse_rebuild_interval(ctx, prec):
    """Reusing the coarse levels between rebuilds changes the iteration
    path, not the converged load levels."""
    res = []
    for every in (1, 8):
        mesh, loads, mat, P = plate(ctx, 60, 30)
        P.set_mg_rebuild(every)
        per = P.solve_steps(6, 1.0, A.options(cg_tol=1e-9, tol=1e-7, aitken=True, precision=prec, precond=2,
                                              predictor=1, cg_forcing=0.1, inner_tol=1e-4))
        assert all(s.status == A.CP_OK for s in per)
        res.append((P.field("u"), P.field("sigI"), sum(s.cg_iters for s in per)))
        P.close()
    print("cg iterations, rebuild 1 / 8:", res[0][2], res[1][2], flush=True)
    assert rel(res[1][0], res[0][0]) < 1e-6 and rel(res[1][1], res[0][1]) < 1e-6
    mesh, loads, mat, P = plate(ctx, 20, 10)
    with pytest.raises(A.CPError) as e:
        P.set_mg_rebuild(0)
    assert e.value.code == A.CP_ERR_INVALID


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_tail_level_read_from_global_memory(ctx, prec):
    """A coarse tail whose finest level keeps its rows and pivot inverses in
    global memory.  The 79x49 lattice has level 1 at 40x25 = 1000 nodes, so
    the one-workgroup tail holds 40x25, 21x13, 11x7 and the 6x4 inverse
    (1374 nodes).  In FP64 their six vectors take 96 B a node (132 KB of the
    144 KB the vectors may use) and the masks 2 B a node; of the LDS that
    remains the pivot inverses (32 B a node) fit for the coarse levels but not
    for the 1000-node one, so its steps read rows and pivots from global
    memory (in mixed precision its pivots fit and only its rows stay out).
    Either way the solve matches the oracle."""
    mesh, loads, mat, P = plate(ctx, 78, 48, S=100.0, sY=1e300)
    levels = P.mg_levels()
    print("levels", levels, flush=True)
    assert levels[1] == (40, 25) and levels[2] == (21, 13) and levels[-1] == (6, 4)
    st = P.solve(A.options(cg_tol=1e-10, precond=2, precision=prec), elastic=True)
    assert st.status == A.CP_OK
    om = O.gen_quarter_plate(1.0, 0.1, 78, 48)
    ol = O.Problem(om).force(1, 100.0, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    uo, so, _ = O.solve(ol, O.material(E, 0.3), O.options(cg_tol=1e-10, elastic=True))
    print("cg iterations", st.cg_iters, flush=True)
    assert rel(P.field("u"), uo) < 1e-8
    assert rel(P.field("sxx"), O.fields(so)["sxx"]) < 1e-8
    P.close()
