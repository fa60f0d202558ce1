"""Serving loop (continuous batching) on the GPU through the C-ABI: every
result is the same alone and in the mix, admission respects the HBM budget
and slot count, oversize requests are refused."""
import numpy as np
import pytest

from cudaplasticity_amd import abi as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def plate(Nt, Nr, r=0.1, S=300.0, sY=240.0, law=A.LAW_LINEAR):
    mesh = A.gen_quarter_plate(1.0, r, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = A.material(E, 0.3, sY, law, K_T=E / 101, n=0.2)
    return mesh, loads, mat


def ring(Nt, Nr, p=180.0, sY=240.0):
    mesh = A.gen_quarter_ring(0.5, 1.0, Nt, Nr, 1.05)
    loads = A.Loads(mesh).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = A.material(E, 0.3, sY, A.LAW_LINEAR, K_T=E / 101)
    return mesh, loads, mat


CASES = [
    (plate(40, 20), 4, A.MIXED),
    (plate(24, 10, r=0.2, sY=300.0, law=A.LAW_SWIFT), 3, A.MIXED),
    (ring(48, 12), 3, A.MIXED),
    (plate(60, 30, r=0.15), 4, A.FP64),
    (plate(30, 12), 2, A.FP64),
]


def opts(prec):
    return A.options(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=500, precision=prec, precond=True,
                     predictor=1, cg_forcing=0.1, inner_tol=1e-4)


@pytest.mark.parametrize("batch", [True, False])
def test_served_results_equal_alone(ctx, batch):
    alone = []
    for (mesh, loads, mat), nsteps, prec in CASES:
        P = A.Problem(ctx, mesh, loads, mat)
        P.solve_steps(nsteps, 1.0, opts(prec))
        alone.append((P.field("u"), P.field("sigI")))
        P.close()
    srv = A.Server(ctx, hbm_budget=1 << 30, max_slots=3, batch_cg=batch)
    ids = [srv.submit(m, l, mat, n, 1.0, opts(p)) for (m, l, mat), n, p in CASES]
    srv.run_until_done()
    for i, rid in enumerate(ids):
        code, info, u, sig = srv.result(rid)
        print("request", rid, "status", info.status, "outer", info.outer_iters, "cg", info.cg_iters,
              "latency", info.t_done - info.t_submit, flush=True)
        assert code == A.CP_OK
        assert np.array_equal(u, alone[i][0]), (i, np.abs(u - alone[i][0]).max())
        assert np.array_equal(sig, alone[i][1])
        assert info.t_submit <= info.t_admit <= info.t_done
    srv.close()


MG_CASES = [
    (plate(40, 20), (41, 21), 4, A.MIXED),
    (ring(48, 12), (49, 13), 3, A.MIXED),
    (plate(60, 30, r=0.15), (61, 31), 4, A.FP64),
    (plate(24, 10, r=0.2, sY=300.0, law=A.LAW_SWIFT), (25, 11), 3, A.MIXED),
]


def mg_opts(prec):
    return A.options(cg_tol=1e-9, tol=1e-6, aitken=True, max_outer=500, precision=prec, precond=2,
                     predictor=1, cg_forcing=0.3, inner_tol=1e-5)


@pytest.mark.parametrize("batch_mg", [True, False])
def test_served_multigrid_equal_alone(ctx, batch_mg):
    """Multigrid requests, served next to block-Jacobi requests in the
    batched kernels, give bitwise the results of the same problems solved
    alone: with lockstep batches (batch_mg: the CG iterations of the
    requests in their CG phase advance together, each step one launch over
    all of them) and with each request replaying its own V-cycle graphs."""
    alone = []
    for (mesh, loads, mat), grid, nsteps, prec in MG_CASES:
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(*grid)
        P.solve_steps(nsteps, 1.0, mg_opts(prec))
        alone.append((P.field("u"), P.field("sigI")))
        P.close()
    srv = A.Server(ctx, hbm_budget=1 << 30, max_slots=3, batch_cg=True, batch_mg=batch_mg)
    ids = [srv.submit(m, l, mat, n, 1.0, mg_opts(p), grid=g) for (m, l, mat), g, n, p in MG_CASES]
    (bm, bl, bmat), bn, bp = CASES[0]
    bj = srv.submit(bm, bl, bmat, bn, 1.0, opts(bp))          # a block-Jacobi request in the same rounds
    srv.run_until_done()
    rounds, members, iters = srv.batch_stats()
    print("lockstep batches", rounds, "members", members, "iterations", iters, flush=True)
    if batch_mg:
        assert rounds > 0 and members >= 2 * rounds
    else:
        assert rounds == 0
    for i, rid in enumerate(ids):
        code, info, u, sig = srv.result(rid)
        print("request", rid, "status", info.status, "outer", info.outer_iters, "cg", info.cg_iters,
              "latency", info.t_done - info.t_submit, flush=True)
        assert code == A.CP_OK
        assert np.array_equal(u, alone[i][0]), (i, np.abs(u - alone[i][0]).max())
        assert np.array_equal(sig, alone[i][1])
    assert srv.result(bj)[0] == A.CP_OK
    with pytest.raises(A.CPError) as e:                        # lattice of the wrong size
        (m, l, mat), g, n, p = MG_CASES[0]
        srv.submit(m, l, mat, n, 1.0, mg_opts(p), grid=(g[0], g[1] + 1))
    assert e.value.code == A.CP_ERR_INVALID
    srv.close()


def test_served_stream_overlapping_preparation(ctx):
    """A stream of 24 multigrid requests (plates and graded rings of
    2k-40k quads, both hardening laws, FP64 and mixed) submitted at once to
    6 slots: a worker thread prepares the next requests (pattern, upload,
    lattice, graph capture) while the others run, each request advances as
    soon as its own readback lands, and the CG iterations of the requests
    that reach their CG phase together run in lockstep batches.  Every
    result equals the problem solved alone, bit for bit, and failures carry
    their message."""
    from cudaplasticity_amd import serving
    stream = serving.make_stream(24, seed=99, qmin=2e3, qmax=4e4)
    alone = []
    for k, (mesh, loads, mat, desc) in enumerate(stream):
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(*serving.lattice(desc))
        P.solve_steps(3, 1.0, mg_opts(A.FP64 if k % 3 == 0 else A.MIXED))
        alone.append((P.field("u"), P.field("sigI")))
        P.close()
    srv = A.Server(ctx, hbm_budget=1 << 34, max_slots=6, batch_mg=True)
    ids = [srv.submit(m, l, mat, 3, 1.0, mg_opts(A.FP64 if k % 3 == 0 else A.MIXED), grid=serving.lattice(d))
           for k, (m, l, mat, d) in enumerate(stream)]
    srv.run_until_done()
    for i, rid in enumerate(ids):
        code, info, u, sig = srv.result(rid)
        assert code == A.CP_OK, (i, srv.last_message)
        assert np.array_equal(u, alone[i][0]) and np.array_equal(sig, alone[i][1]), i
        assert info.t_submit <= info.t_admit <= info.t_done
    rounds, members, iters = srv.batch_stats()
    print("served", len(ids), "requests; last done at %.3f s" % max(srv.result(r, False)[1].t_done for r in ids),
          "; lockstep batches", rounds, "members", members, "iterations", iters, flush=True)
    assert rounds > 0 and members >= 2 * rounds
    srv.close()


def test_failed_requests_carry_their_message(ctx):
    """A request whose level cannot converge (one secant iteration allowed,
    a plastic problem) fails with CP_ERR_NOT_CONVERGED and the failure's
    message, on the multigrid path and on the batched block-Jacobi path,
    while a healthy request served beside them finishes with the bits it
    has alone."""
    (m, l, mat), grid, n, p = MG_CASES[0]
    bad = A.options(cg_tol=1e-9, tol=1e-12, max_outer=1, precision=A.MIXED, precond=2, inner_tol=1e-5)
    bad_bj = A.options(cg_tol=1e-9, tol=1e-12, max_outer=1, precision=A.MIXED, precond=True, inner_tol=1e-5)
    (gm, gl, gmat), ggrid, gn, gp = MG_CASES[1]
    P = A.Problem(ctx, gm, gl, gmat)
    P.set_grid(*ggrid)
    P.solve_steps(gn, 1.0, mg_opts(gp))
    good_alone = P.field("u")
    P.close()
    srv = A.Server(ctx, hbm_budget=1 << 32, max_slots=4)
    i_mg = srv.submit(m, l, mat, n, 1.0, bad, grid=grid)
    i_bj = srv.submit(m, l, mat, n, 1.0, bad_bj)
    i_ok = srv.submit(gm, gl, gmat, gn, 1.0, mg_opts(gp), grid=ggrid)
    srv.run_until_done()
    for rid in (i_mg, i_bj):
        code, info, _, _ = srv.result(rid)
        print("request", rid, "code", code, srv.last_message, flush=True)
        assert code == A.CP_ERR_NOT_CONVERGED and info.status == A.CP_ERR_NOT_CONVERGED
        assert "did not converge" in srv.last_message
    code, info, u, _ = srv.result(i_ok)
    assert code == A.CP_OK and np.array_equal(u, good_alone)
    srv.close()


@pytest.mark.parametrize("precond", [1])
def test_stalled_refinement_same_alone_and_served(ctx, precond):
    """A mixed-precision solve whose refinement stalls (block-Jacobi CG,
    inner_tol 0.9: each FP32 solve reduces the residual by 10 % only) takes
    the same passes alone and served: cp_solve and the serving loop share the
    refinement decision (one multigrid iteration reduces the residual far
    more than that, so the multigrid path does not stall this way).  The
    level's last linear solve did not converge, so both report
    CP_ERR_NOT_CONVERGED after the same secant and CG iterations."""
    (m, l, mat), grid, n, p = MG_CASES[0]
    stall = A.options(cg_tol=1e-9, tol=1e-6, fixed_outer=3, precision=A.MIXED, precond=precond, inner_tol=0.9)
    P = A.Problem(ctx, m, l, mat)
    if precond == 2:
        P.set_grid(*grid)
    P.set_load(1.0 / n)
    st = A.Stats()
    code = A.lib().cp_solve(P.ptr, A.C.byref(stall), A.C.byref(st))
    msg = A.lib().cp_last_error().decode()
    print("alone: code", code, "outer", st.outer_iters, "cg", st.cg_iters, msg, flush=True)
    assert code == A.CP_ERR_NOT_CONVERGED and "stalled" in msg
    P.close()
    srv = A.Server(ctx, hbm_budget=1 << 32, max_slots=2)
    rid = srv.submit(m, l, mat, n, 1.0, stall, grid=grid if precond == 2 else None)
    srv.run_until_done()
    code, info, _, _ = srv.result(rid)
    print("served: code", code, "outer", info.outer_iters, "cg", info.cg_iters, srv.last_message, flush=True)
    assert code == A.CP_ERR_NOT_CONVERGED and "stalled" in srv.last_message
    assert (info.outer_iters, info.cg_iters) == (st.outer_iters, st.cg_iters)
    srv.close()


def test_refinement_floor_same_alone_and_served(ctx):
    """cg_tol 1e-15 lies below what an FP64 residual of this system can
    show: the refinement ends at the rounding floor of its own FP64
    evaluation (16 eps64 ||(|K| |u| + |b|)||) and the level converges, with
    the same bits alone and served (multigrid CG)."""
    (m, l, mat), grid, n, p = MG_CASES[0]
    opt = A.options(cg_tol=1e-15, tol=1e-6, aitken=True, max_outer=100, precision=A.MIXED, precond=2,
                    inner_tol=1e-5)
    P = A.Problem(ctx, m, l, mat)
    P.set_grid(*grid)
    per = P.solve_steps(2, 1.0, opt)
    print("alone:", [(s.outer_iters, s.cg_iters, s.refinements) for s in per], flush=True)
    assert all(s.status == A.CP_OK for s in per)
    u, sig = P.field("u"), P.field("sigI")
    P.close()
    srv = A.Server(ctx, hbm_budget=1 << 32, max_slots=2)
    rid = srv.submit(m, l, mat, 2, 1.0, opt, grid=grid)
    srv.run_until_done()
    code, info, su, ssig = srv.result(rid)
    assert code == A.CP_OK, srv.last_message
    assert np.array_equal(su, u) and np.array_equal(ssig, sig)
    srv.close()


def test_budget_and_slots(ctx):
    (m, l, mat), _, _ = CASES[0]
    import ctypes as C
    need = C.c_size_t()
    A.check(A.lib().cp_problem_bytes(m.nn, m.ne, 4 * 9 * m.nn, A.MIXED, C.byref(need)))
    srv = A.Server(ctx, hbm_budget=int(need.value * 2.5), max_slots=8)
    ids = [srv.submit(m, l, mat, 2, 1.0, opts(A.MIXED)) for _ in range(5)]
    big = A.gen_quarter_plate(1.0, 0.1, 400, 200)
    bl = A.Loads(big).force(1, 300.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    big_id = srv.submit(big, bl, mat, 1, 1.0, opts(A.MIXED))
    srv.run_until_done()
    admits = []
    for rid in ids:
        code, info, u, sig = srv.result(rid)
        assert code == A.CP_OK
        admits.append((info.t_admit, info.t_done))
    # at most two fit at once: the third is admitted only after one finished
    admits.sort()
    assert admits[2][0] >= min(admits[0][1], admits[1][1])
    code, info, _, _ = srv.result(big_id)
    assert code == A.CP_ERR_BUDGET and info.status == A.CP_ERR_BUDGET
    srv.close()


def test_served_stream_full_range_two_rates(ctx):
    """The config-5 size range at its ends: 8 requests drawn log-uniform over
    2k-500k quads (plates and graded rings, both laws, mixed precision,
    multigrid CG, 5 increments each as in the bench's stream) plus a
    500k-quad plate and a 500k-quad ring, served with Poisson arrivals at two
    rates (all at once, and spread at about the back-to-back service rate).
    Each served result equals the problem solved alone, bit for bit, at both
    rates."""
    import time
    from cudaplasticity_amd import serving
    rng = np.random.Generator(np.random.PCG64(20261204))
    stream = [serving.make_problem(rng, 2e3, 5e5) for _ in range(8)]
    # one plate and one ring of about 500k quads: the first of each kind
    # drawn from a second generator (each draw is a plate with probability 1/2)
    big = np.random.Generator(np.random.PCG64(5))
    kinds = {}
    for _ in range(64):
        mesh, loads, mat, desc = serving.make_problem(big, 4.9e5, 5e5)
        kinds.setdefault(desc["kind"], (mesh, loads, mat, desc))
        if len(kinds) == 2:
            break
    assert len(kinds) == 2
    stream += [kinds["plate"], kinds["ring"]]
    quads = [d["quads"] for _, _, _, d in stream]
    print("stream quads", sorted(quads), flush=True)
    assert min(quads) < 1e4 and max(quads) > 4.8e5
    nsteps = 5
    opt = mg_opts(A.MIXED)
    alone = []
    t0 = time.time()
    for mesh, loads, mat, desc in stream:
        ts = time.time()
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(*serving.lattice(desc))
        P.solve_steps(nsteps, 1.0, opt)
        alone.append((P.field("u"), P.field("sigI")))
        P.close()
        print("  alone: %s %d quads %.2f s" % (desc["kind"], desc["quads"], time.time() - ts), flush=True)
    span = time.time() - t0
    print("alone back to back: %.2f s" % span, flush=True)
    for rate in (0.0, len(stream) / span):
        recs = serving.serve(ctx, stream, nsteps, opt, rate, hbm_budget=40 << 30, max_slots=8)
        lat = sorted(r["latency"] for r in recs)
        print("rate %.2f/s: p50 %.3f s, p99 %.3f s" % (rate, lat[len(lat) // 2], lat[-1]), flush=True)
        for i, r in enumerate(recs):
            assert r["code"] == A.CP_OK, (i, r["message"])
            assert np.array_equal(r["u"], alone[i][0]) and np.array_equal(r["sig"], alone[i][1]), (rate, i)
