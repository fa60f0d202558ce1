"""GPU parity of the 8-node (order-2) element path through the C-ABI.

Inputs: the reference's energyConvTest hole plate, genRectWithHole(-3, 3,
-2, 2, 0.5, 3k, 2k, order 2), 304L material (Main.cpp:8-10, 105-178).
Bars: FP64 assembly within 1e-13 of the oracle's CSR values; FP64 solutions
within relative L2 1e-9 (elastic) and 1e-6 (plastic, secant exit 1e-8) of
the oracle; the mixed path (FP32 CG with block-Jacobi + FP64 refinement)
within 1e-6 of the FP64 path; and deformEnergy of the GPU solution on the
reference's printed 6 digits (data/aitken__elast2.txt, aitken__plast2.txt).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from cudaplasticity_amd import abi as A

pytestmark = pytest.mark.gpu
E304, ST304, NU304 = 193e9, 170e6, 0.29


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def case(k, law, S):
    pm = A.gen_rect_with_hole(-3., 3., -2., 2., 0.5, 3 * k, 2 * k, order=2)
    om = O.gen_rect_with_hole(-3., 3., -2., 2., 0.5, 3 * k, 2 * k, order=2)
    pl = A.Loads(pm).force(1, S, 0.).force(3, -S, 0.).fix_vert_axis(0.).fix_hor_axis(0.)
    ol = O.Problem(om).force(1, S, 0.).force(3, -S, 0.).fix_vert_axis(0.).fix_hor_axis(0.)
    assert np.array_equal(pl.rp, ol.rp) and np.array_equal(pl.mask, ol.mask)
    if law == "elastic":
        kw = dict(E=E304, nu=NU304)
    else:
        kw = dict(E=E304, nu=NU304, sigmaT=ST304, law=A.LAW_POWER, A=0.01 * E304, m=0.015)
    print("8-node hole plate k=%d: %d nodes, %d elements, %s" % (k, pm.nn, pm.ne, law), flush=True)
    return pm, pl, A.material(**kw), om, ol, O.material(**kw)


def energy(om, omat, P):
    F = P.fields()
    st = np.stack([F[name] for name in O.FIELDS])
    return O.deform_energy(om, omat, st)


def within_printed(value, printed):
    unit = 10.0 ** (math.floor(math.log10(abs(printed))) - 5)
    return abs(value - printed) <= 0.5 * unit


def test_q8_assembly_matches_oracle(ctx):
    pm, pl, pmat, om, ol, omat = case(4, "power", 170e6)
    P = A.Problem(ctx, pm, pl, pmat)
    rows, cols = P.csr()
    ro, co = O.csr_pattern(om)
    assert np.array_equal(rows, ro) and np.array_equal(cols, co)
    P.assemble(A.FP64)
    got = P.values()
    want = O.assemble(om, ro, co, np.full(om.ne, E304), np.full(om.ne, NU304))
    err = np.abs(got - want).max() / np.abs(want).max()
    print("assembly max rel err %.2e" % err, flush=True)
    assert err < 1e-13
    P.close()


def test_q8_elastic_matches_oracle_and_energy_table(ctx):
    pm, pl, pmat, om, ol, omat = case(6, "elastic", 1e6)
    P = A.Problem(ctx, pm, pl, pmat)
    P.solve(A.options(cg_tol=1e-12), elastic=True)
    u_o, st_o, _ = O.solve(ol, omat, O.options(cg_tol=1e-12, elastic=True))
    e_u = relnorm(P.field("u"), u_o)
    e_s = relnorm(P.field("sigI"), st_o[O.FIELDS.index("sigI")])
    en = energy(om, omat, P)
    print("elastic u %.2e sigma_i %.2e energy %.7f (table 68.6683)" % (e_u, e_s, en), flush=True)
    assert e_u < 1e-9 and e_s < 1e-9
    assert within_printed(en, 68.6683)
    # equilibrium of the internal forces on the free DOFs
    f = P.internal_force()
    free = pl.mask.astype(bool)
    assert np.linalg.norm((f - pl.rp)[free]) <= 1e-9 * np.linalg.norm(pl.rp[free])
    P.close()


@pytest.mark.parametrize("precision,precond", [(A.FP64, 0), (A.FP64, 1), (A.MIXED, 1)])
def test_q8_plastic_matches_oracle(ctx, precision, precond):
    pm, pl, pmat, om, ol, omat = case(4, "power", 170e6)
    opt = dict(cg_tol=1e-10, tol=1e-8, aitken=True)
    P = A.Problem(ctx, pm, pl, pmat)
    st = P.solve(A.options(precision=precision, precond=precond, inner_tol=1e-5, **opt))
    u_o, st_o, so = O.solve(ol, omat, O.options(max_outer=1000, **opt))
    assert st.status == A.CP_OK and so.status == 0
    e_u = relnorm(P.field("u"), u_o)
    e_s = relnorm(P.field("sigI"), st_o[O.FIELDS.index("sigI")])
    print("precision %d precond %d: outer %d (oracle %d), cg %d; u %.2e sigma_i %.2e" %
          (precision, precond, st.outer_iters, so.outer_iters, st.cg_iters, e_u, e_s), flush=True)
    assert e_u < 1e-6 and e_s < 1e-6
    P.close()


def test_q8_energy_table_at_reference_tolerances(ctx):
    """The reference's loop: FP64, unpreconditioned CG to 1e-7, plain secant
    iterations to relErr 1e-5; aitken__plast2.txt prints 2.45491e+06 for the
    660-element mesh."""
    pm, pl, pmat, om, ol, omat = case(6, "power", 170e6)
    P = A.Problem(ctx, pm, pl, pmat)
    st = P.solve(A.options())
    en = energy(om, omat, P)
    print("outer %d cg %d energy %.1f (table 2.45491e+06)" % (st.outer_iters, st.cg_iters, en), flush=True)
    assert st.status == A.CP_OK
    assert within_printed(en, 2.45491e6)
    P.close()


def q8_grid(P, k):
    """The corner lattice of the k plate (periodic around the hole)."""
    m1, m2, corners = A.corner_lattice2("hole", -3., 3., -2., 2., 0.5, 3 * k, 2 * k)
    P.set_grid(m1, m2, corners, periodic=True)
    return m1, m2


@pytest.mark.parametrize("precision", [A.FP64, A.MIXED])
def test_q8_multigrid_matches_oracle(ctx, precision):
    """Multigrid on an 8-node mesh: level 0 is the mesh itself (2x2
    block-Jacobi Chebyshev smoothing on its rows), level 1 its corner
    lattice (the Galerkin product of the transfer that gives a corner its
    own value and a mid-side node the mean of its edge's corners), then the
    lattice hierarchy.  The plastic plate of test_q8_plastic_matches_oracle
    to the same bars, in a fraction of block-Jacobi's CG iterations."""
    pm, pl, pmat, om, ol, omat = case(4, "power", 170e6)
    opt = dict(cg_tol=1e-10, tol=1e-8, aitken=True)
    P = A.Problem(ctx, pm, pl, pmat)
    m1, m2 = q8_grid(P, 4)
    levels = P.mg_levels()
    print("levels", levels, "smoothers", P.mg_smoothers(), flush=True)
    assert levels[0] == (0, 0) and levels[1] == (m1, m2) and len(levels) >= 3
    st = P.solve(A.options(precision=precision, precond=2, inner_tol=1e-5, **opt))
    u_o, st_o, so = O.solve(ol, omat, O.options(max_outer=1000, **opt))
    Q = A.Problem(ctx, pm, pl, pmat)
    sq = Q.solve(A.options(precision=precision, precond=1, inner_tol=1e-5, **opt))
    e_u = relnorm(P.field("u"), u_o)
    e_s = relnorm(P.field("sigI"), st_o[O.FIELDS.index("sigI")])
    print("multigrid: outer %d cg %d; block-Jacobi: outer %d cg %d; u %.2e sigma_i %.2e" %
          (st.outer_iters, st.cg_iters, sq.outer_iters, sq.cg_iters, e_u, e_s), flush=True)
    assert st.status == A.CP_OK and so.status == 0
    assert e_u < 1e-6 and e_s < 1e-6
    assert st.cg_iters * 3 < sq.cg_iters
    P.close()
    Q.close()


def test_q8_multigrid_fixed_iterations_vs_oracle(ctx):
    """Six plain secant iterations with CG to 1e-12 on both sides: the
    multigrid path within 1e-8 of the oracle (a preconditioner changes the
    CG path, not the solution it converges to)."""
    pm, pl, pmat, om, ol, omat = case(4, "power", 170e6)
    P = A.Problem(ctx, pm, pl, pmat)
    q8_grid(P, 4)
    st = P.solve(A.options(cg_tol=1e-12, fixed_outer=6, precond=2))
    u_o, st_o, so = O.solve(ol, omat, O.options(cg_tol=1e-12, fixed_outer=6))
    print("cg %d, u %.2e" % (st.cg_iters, relnorm(P.field("u"), u_o)), flush=True)
    assert relnorm(P.field("u"), u_o) < 1e-8
    for f in ("sxx", "sigI", "Ec", "psi"):
        assert relnorm(P.field(f), st_o[O.FIELDS.index(f)]) < 1e-8, f
    P.close()


def test_q8_multigrid_energy_table(ctx):
    """The 660-element plate at the reference's tolerances (CG to 1e-7,
    secant exit 1e-5) with multigrid: aitken__plast2.txt's 2.45491e+06."""
    pm, pl, pmat, om, ol, omat = case(6, "power", 170e6)
    P = A.Problem(ctx, pm, pl, pmat)
    q8_grid(P, 6)
    st = P.solve(A.options(precond=2))
    en = energy(om, omat, P)
    print("outer %d cg %d energy %.1f (table 2.45491e+06)" % (st.outer_iters, st.cg_iters, en), flush=True)
    assert st.status == A.CP_OK
    assert within_printed(en, 2.45491e6)
    P.close()


def test_q8_multigrid_refusals(ctx):
    """An 8-node mesh's lattice is its corner lattice, given with node_of;
    a lattice that misses that shape is refused."""
    pm, pl, pmat, om, ol, omat = case(2, "elastic", 1e6)
    P = A.Problem(ctx, pm, pl, pmat)
    m1, m2, corners = A.corner_lattice2("hole", -3., 3., -2., 2., 0.5, 6, 4)
    with pytest.raises(A.CPError) as e:
        P.set_grid(m1, m2, None, periodic=True)                     # node_of required
    assert e.value.code == A.CP_ERR_INVALID
    with pytest.raises(A.CPError) as e:
        P.set_grid(m1, m2, corners, periodic=False)                 # the seam's edges are not lattice steps
    assert e.value.code == A.CP_ERR_INVALID and "lattice step" in str(e.value)
    mid = int(pm.conn.reshape(-1, 8)[0, 4])
    bad = corners.copy()
    bad[0] = mid
    with pytest.raises(A.CPError) as e:
        P.set_grid(m1, m2, bad, periodic=True)                      # a mid-side node as a lattice point
    assert e.value.code == A.CP_ERR_INVALID
    with pytest.raises(A.CPError):
        P.solve(A.options(precond=2))                               # no lattice set
    P.close()


def test_q8_gravity_equilibrium(ctx):
    """An 8-node column under its own weight (consistent body-force loads,
    corners -1/12 and mid-sides 1/3 of an element's weight), top edge fixed:
    the reactions on the fixed DOFs carry the whole weight."""
    pm = A.gen_rect(0.0, 1.0, 0.0, 4.0, 3, 8, order=2)
    top = pm.borders[1]
    assert np.allclose(pm.y[top], 4.0)
    pl = A.Loads(pm, h=0.2).gravity(9.81, 7.8).displacement(1, 0.0, 0.0)
    P = A.Problem(ctx, pm, pl, A.material(2.1e5, 0.3, h=0.2))
    P.solve(A.options(cg_tol=1e-12), elastic=True)
    f = P.internal_force()
    weight = 7.8 * 9.81 * 0.2 * 4.0
    fixed_y = 2 * top + 1
    reaction = (f[fixed_y] - pl.rp[fixed_y]).sum()   # support force: internal force less the lumped load
    print("reaction %.12e weight %.12e" % (reaction, weight), flush=True)
    assert abs(reaction - weight) <= 1e-9 * weight
    assert abs(pl.rp[1::2].sum() + weight) <= 1e-12 * weight
    free = pl.mask.astype(bool)
    assert np.linalg.norm((f - pl.rp)[free]) <= 1e-9 * np.linalg.norm(pl.rp[free])
    P.close()
