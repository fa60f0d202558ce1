"""Multigrid CG (precond = 2) on the reference's own meshes: plates with a
hole (genRectWithHole, Mesh.cu:112-245) and rings (genRing, Mesh.cu:249-329),
whose rays close around the hole, so that their lattice is periodic along
direction 1 (cp_problem_set_grid_periodic, the lattice from
cp_mesh_ring_lattice).  Checked against the reference's plasticityTest result
(data/holePlatePlastPow.vtk), the oracle's secant loop with tight CG, the
oracle's secant loop with a direct solve on a speedTest2 plate
(Main.cpp:458-600, 192 x 128, 81,920 quads), and against the same problem
renumbered and served."""
import os

import numpy as np
import pytest

import oracle_lib as O
from cudaplasticity_amd import abi as A
from test_gpu_parity import hole_fixture_case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def with_lattice(P, mesh, border_n):
    n1, n2, node_of = A.ring_lattice(mesh, border_n)
    P.set_grid(n1, n2, node_of, periodic=True)
    return n1, n2, node_of


def test_hole_plate_fixture_with_multigrid(ctx):
    """plasticityTest (Main.cpp:252-296) with the reference's tolerances
    (CG to 1e-7, secant exit 1e-5) and the multigrid preconditioner on the
    plate's 160 x 33 periodic lattice: the fixture's iteration count, exit
    measure, plastic zone (bit-exact flags) and fields (1e-6, the bar of the
    plain-CG path), in FP64 and mixed precision, with a few hundred CG
    iterations in all."""
    fx = np.load(os.path.join(GOLD, "holePlatePlastPow.npz"))
    pm, pl, pmat, om, ol, omat = hole_fixture_case()
    P = A.Problem(ctx, pm, pl, pmat)
    with_lattice(P, pm, 160)
    levels = P.mg_levels()
    print("levels", levels, "smoothers", P.mg_smoothers(), flush=True)
    assert levels == [(160, 33), (80, 17), (40, 9), (20, 5), (10, 3)]
    for prec in (A.FP64, A.MIXED):
        P.reset()
        st = P.solve(A.options(cg_tol=1e-7, tol=1e-5, precond=2, precision=prec))
        errs = {n: rel(P.field(k), fx["f_" + n]) for k, n in
                [("sxx", "Sigma_xx"), ("sigI", "Sigma_intensity"), ("Ec", "E_c"), ("epsI_p", "Eps_p_intensity")]}
        errs["u"] = rel(P.field("u"), fx["f_Displacement"][:, :2].reshape(-1))
        flags = P.field("epsI") >= pmat.sigmaT / pmat.E
        ref_flags = fx["f_Eps_intensity"] >= pmat.sigmaT / pmat.E
        print("precision", prec, "outer", st.outer_iters, "cg", st.cg_iters, "relErr %.4g" % st.rel_err,
              " ".join("%s %.2e" % kv for kv in errs.items()), "plastic", flags.sum(),
              "differ", (flags != ref_flags).sum(), flush=True)
        assert st.status == 0 and abs(st.outer_iters - 41) <= 1
        assert abs(st.rel_err - 9.68e-6) < 0.05e-6
        assert max(errs.values()) < 1e-6
        assert flags.sum() == 3528 and np.array_equal(flags, ref_flags)
        assert st.cg_iters < 41 * 20
    P.close()


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_hole_plate_fixed_iterations_vs_oracle(ctx, prec):
    """12 plain secant iterations, CG to 1e-13 on both sides (the mixed
    path refines in FP64 to its rounding floor): displacements and every
    stress field within 1e-8 of the oracle, plastic flags equal."""
    pm, pl, pmat, om, ol, omat = hole_fixture_case()
    P = A.Problem(ctx, pm, pl, pmat)
    with_lattice(P, pm, 160)
    st = P.solve(A.options(cg_tol=1e-13, fixed_outer=12, precond=2, precision=prec))
    u_o, s_o, _ = O.solve(ol, omat, O.options(cg_tol=1e-13, fixed_outer=12))
    F_o = O.fields(s_o)
    print("cg", st.cg_iters, "rel u %.2e" % rel(P.field("u"), u_o), flush=True)
    assert rel(P.field("u"), u_o) < 1e-8
    for f in ("sxx", "syy", "tau", "sigI", "Ec", "psi", "epsI_p"):
        assert rel(P.field(f), F_o[f]) < 1e-8, f
    epsT = pmat.sigmaT / pmat.E
    assert np.array_equal(P.field("epsI") >= epsT, F_o["epsI"] >= epsT)
    P.close()


def speedtest2_case(scale):
    """speedTest2's first-order plastic plate (Main.cpp:458-476, 503-590):
    genRectWithHole(-3, 3, -2, 2, 0.5, 6 s, 4 s), 304L steel in SI units
    with linear hardening K_T = E / 10, traction 170 MPa on x = +-3."""
    E, sT, nu = 193e9, 170e6, 0.29
    N1, N2 = 6 * scale, 4 * scale
    pm = A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, N1, N2)
    om = O.gen_rect_with_hole(-3, 3, -2, 2, 0.5, N1, N2)
    pl = A.Loads(pm).force(1, 170e6, 0.0).force(3, -170e6, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    ol = O.Problem(om).force(1, 170e6, 0.0).force(3, -170e6, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    kw = dict(E=E, nu=nu, sigmaT=sT, law=A.LAW_LINEAR, K_T=E * 0.1)
    return pm, pl, A.material(**kw), om, ol, O.material(**kw), 2 * (N1 + N2)


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_speedtest2_plate_vs_direct_oracle(ctx, prec):
    """The speedTest2 plate at scale 32 (192 x 128, 82,560 nodes): three
    plain secant iterations with CG to 1e-12 against the oracle's loop with
    a direct solve of each secant system (u and stresses within 1e-8)."""
    pm, pl, pmat, om, ol, omat, bn = speedtest2_case(32)
    P = A.Problem(ctx, pm, pl, pmat)
    with_lattice(P, pm, bn)
    print("case: speedTest2 plate 192x128, precision", prec, "levels", P.mg_levels(), flush=True)
    st = P.solve(A.options(cg_tol=1e-12, fixed_outer=3, precond=2, precision=prec))
    u_o, s_o, rels = O.solve_direct(ol, omat, 3)
    F_o = O.fields(s_o)
    print("cg", st.cg_iters, "rel u %.2e" % rel(P.field("u"), u_o), "oracle exit measures", rels, flush=True)
    assert st.status == 0 and st.outer_iters == 3
    assert rel(P.field("u"), u_o) < 1e-8
    for f in ("sxx", "syy", "tau", "sigI", "Ec"):
        assert rel(P.field(f), F_o[f]) < 1e-8, f
    assert (P.field("epsI") >= pmat.sigmaT / pmat.E).sum() > 0
    P.close()


def test_speedtest2_plate_solve_multigrid_vs_plain_cg(ctx):
    """The same plate solved as speedTest2 solves it (floatBoost: mixed
    precision, the reference's tolerances) with multigrid and with the
    reference's unpreconditioned CG: both converge to the same plastic
    state (fields within the 1e-5 secant exit of each other)."""
    pm, pl, pmat, om, ol, omat, bn = speedtest2_case(32)
    P = A.Problem(ctx, pm, pl, pmat)
    with_lattice(P, pm, bn)
    st = P.solve(A.options(precond=2, precision=A.MIXED))
    Q = A.Problem(ctx, pm, pl, pmat)
    sq = Q.solve(A.options(precision=A.MIXED))
    print("multigrid: outer %d cg %d %.2f s; plain CG: outer %d cg %d %.2f s"
          % (st.outer_iters, st.cg_iters, st.seconds, sq.outer_iters, sq.cg_iters, sq.seconds), flush=True)
    assert st.status == 0 and sq.status == 0
    assert abs(st.outer_iters - sq.outer_iters) <= 1
    assert rel(P.field("u"), Q.field("u")) < 5e-5
    assert rel(P.field("sigI"), Q.field("sigI")) < 5e-5
    assert st.cg_iters * 20 < sq.cg_iters
    P.close()
    Q.close()


def ring_case(Nphi, Nr, p=150.0):
    """A full ring (genRing) under inner pressure, linear hardening."""
    E = 2.1e5
    pm, om = A.gen_ring(0.5, 1.0, Nphi, Nr), O.gen_ring(0.5, 1.0, Nphi, Nr)
    pl = A.Loads(pm).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
    ol = O.Problem(om).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0).fix_hor_axis(0)
    kw = dict(E=E, nu=0.3, sigmaT=240.0, law=A.LAW_LINEAR, K_T=E / 101)
    return pm, pl, A.material(**kw), om, ol, O.material(**kw)


def test_ring_vs_oracle(ctx):
    """genRing 96 x 24: the rays are long and thin, so the first level is
    coarsened radially only (96 x 25 -> 96 x 13), then in both directions,
    the periodic one 96 -> 48 -> 24 -> 12; six secant iterations with CG to
    1e-12 against the oracle within 1e-8."""
    pm, pl, pmat, om, ol, omat = ring_case(96, 24)
    P = A.Problem(ctx, pm, pl, pmat)
    with_lattice(P, pm, 96)
    print("levels", P.mg_levels(), flush=True)
    assert P.mg_levels() == [(96, 25), (96, 13), (48, 7), (24, 4), (12, 3)]
    st = P.solve(A.options(cg_tol=1e-12, fixed_outer=6, precond=2))
    u_o, s_o, _ = O.solve(ol, omat, O.options(cg_tol=1e-12, fixed_outer=6))
    print("cg", st.cg_iters, "rel u %.2e" % rel(P.field("u"), u_o), flush=True)
    assert rel(P.field("u"), u_o) < 1e-8
    assert rel(P.field("sigI"), O.fields(s_o)["sigI"]) < 1e-8
    assert (P.field("epsI") >= pmat.sigmaT / pmat.E).sum() > 0
    P.close()


def test_ring_load_steps_mixed_vs_fp64(ctx):
    """A 256 x 32 ring in load steps with Aitken and the predictor: mixed
    precision reaches the FP64 run's converged levels."""
    pm, pl, pmat, om, ol, omat = ring_case(256, 32, p=170.0)
    out = []
    for prec in (A.FP64, A.MIXED):
        P = A.Problem(ctx, pm, pl, pmat)
        with_lattice(P, pm, 256)
        per = P.solve_steps(4, 1.0, A.options(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=500, precision=prec,
                                              precond=2, predictor=1, cg_forcing=0.1, inner_tol=1e-4))
        print("precision", prec, [(s.outer_iters, s.cg_iters) for s in per], flush=True)
        assert all(s.status == A.CP_OK for s in per)
        out.append((P.field("u"), P.field("sigI")))
        P.close()
    assert rel(out[1][0], out[0][0]) < 2e-6 and rel(out[1][1], out[0][1]) < 2e-6


def test_renumbered_plate_lattice(ctx):
    """The plate renumbered by direction (contactTest's numbering,
    Main.cpp:1029): the lattice follows the permutation and the solve equals
    the original numbering's up to rounding."""
    pm, pl, pmat, om, ol, omat = hole_fixture_case()
    hm, new_of_old = A.renum_by_direction(pm, 1.0, 1.0, with_perm=True)
    assert np.array_equal(hm.x[new_of_old], pm.x) and np.array_equal(hm.y[new_of_old], pm.y)
    hl = A.Loads(hm).force(1, 170.0, 0.0).force(3, -170.0, 0.0).fix_vert_axis(0).fix_hor_axis(0)
    n1, n2, node_of = A.ring_lattice(pm, 160)
    opt = A.options(cg_tol=1e-13, fixed_outer=12, precond=2)
    P = A.Problem(ctx, pm, pl, pmat)
    P.set_grid(n1, n2, node_of, periodic=True)
    P.solve(opt)
    H = A.Problem(ctx, hm, hl, pmat)
    H.set_grid(n1, n2, new_of_old[node_of], periodic=True)
    H.solve(opt)
    u0, uh = P.field("u").reshape(-1, 2), H.field("u").reshape(-1, 2)
    print("rel u %.2e" % rel(uh[new_of_old].ravel(), u0.ravel()), flush=True)
    assert rel(uh[new_of_old].ravel(), u0.ravel()) < 1e-9
    assert rel(H.field("sigI"), P.field("sigI")) < 1e-9
    P.close()
    H.close()


def test_served_periodic_equals_alone(ctx):
    """A plate with a hole and a ring served with their periodic lattices
    (cp_server_submit_grid_periodic) give bitwise the results of the same
    problems solved alone."""
    opt = A.options(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=500, precision=A.MIXED, precond=2,
                    predictor=1, cg_forcing=0.3, inner_tol=1e-5)
    pm, pl, pmat, _, _, _ = hole_fixture_case()
    rm, rl, rmat, _, _, _ = ring_case(96, 24)
    cases = [(pm, pl, pmat, 160), (rm, rl, rmat, 96)]
    alone = []
    for m, l, mat, bn in cases:
        P = A.Problem(ctx, m, l, mat)
        with_lattice(P, m, bn)
        P.solve_steps(3, 1.0, opt)
        alone.append((P.field("u"), P.field("sigI")))
        P.close()
    srv = A.Server(ctx, hbm_budget=1 << 32, max_slots=2, batch_mg=True)
    ids = []
    for m, l, mat, bn in cases:
        n1, n2, node_of = A.ring_lattice(m, bn)
        ids.append(srv.submit(m, l, mat, 3, 1.0, opt, grid=(n1, n2), node_of=node_of, periodic=True))
    srv.run_until_done()
    for i, rid in enumerate(ids):
        code, info, u, sig = srv.result(rid)
        print("request", rid, "code", code, "outer", info.outer_iters, "cg", info.cg_iters, flush=True)
        assert code == A.CP_OK
        assert np.array_equal(u, alone[i][0]) and np.array_equal(sig, alone[i][1]), i
    srv.close()


def test_periodic_refusals(ctx):
    """A ring of 45 rays cannot halve its periodic direction and is too big
    to be a coarsest level: refused; the plate's lattice without its seam
    (cp_problem_set_grid) leaves the seam's couplings off the lattice:
    refused."""
    pm, pl, pmat, _, _, _ = ring_case(45, 8)
    P = A.Problem(ctx, pm, pl, pmat)
    n1, n2, node_of = A.ring_lattice(pm, 45)
    with pytest.raises(A.CPError) as e:
        P.set_grid(n1, n2, node_of, periodic=True)
    assert e.value.code == A.CP_ERR_INVALID and "too small or too thin" in str(e.value)
    hm, hl, hmat, _, _, _ = hole_fixture_case()
    H = A.Problem(ctx, hm, hl, hmat)
    n1, n2, node_of = A.ring_lattice(hm, 160)
    with pytest.raises(A.CPError) as e:
        H.set_grid(n1, n2, node_of)
    assert e.value.code == A.CP_ERR_INVALID and "not lattice neighbours" in str(e.value)
    P.close()
    H.close()
