"""Failure paths of the device solver, each forced from the host so that
every kernel still exits normally and the status carries the failure.

The reference checks no CUDA status, has no cap on its GPU loops and leaves
them silently on a NaN, since NaN > eps is false (PlasticitySolver.cu:489-503;
CG: ConjGradCuda.cu:203-222).  Here each case ends with its own status code
and message:
* an invalid material (E <= 0, nu outside (-1, 1/2), h <= 0) is refused at
  creation with CP_ERR_INVALID;
* a NaN in the load gives CP_ERR_NAN (the CG scalar step flags it);
* a body with no Dirichlet constraint under an unbalanced load (K singular,
  the system inconsistent) stops at cg_maxit with CP_ERR_NOT_CONVERGED;
* a level whose linear solves never reach their tolerance is not reported
  converged, even where the secant exit test alone would pass (an elastic
  level satisfies it for any u);
* the unloaded body converges at once with u = 0 (the reference's 0/0 exit
  measure is taken as converged)."""
import numpy as np
import pytest

from cudaplasticity_amd import abi as A

pytestmark = pytest.mark.gpu
E = 2.1e5


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def plate(S=300.0, Nt=30, Nr=12, fix=True):
    mesh = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0)
    if fix:
        loads = loads.fix_vert_axis(0.0).fix_hor_axis(0.0)
    return mesh, loads


@pytest.mark.parametrize("kw", [dict(E=0.0), dict(E=-2.1e5), dict(nu=0.5), dict(nu=-1.0), dict(h=0.0),
                                dict(sigmaT=0.0)])
def test_invalid_material_refused(ctx, kw):
    mesh, loads = plate()
    args = dict(E=E, nu=0.3, sigmaT=240.0, law=A.LAW_LINEAR, K_T=E / 101)
    args.update(kw)
    print("case: material", kw, flush=True)
    with pytest.raises(A.CPError) as e:
        A.Problem(ctx, mesh, loads, A.material(**args))
    assert e.value.code == A.CP_ERR_INVALID
    assert "invalid material" in str(e.value)


@pytest.mark.parametrize("prec,precond", [(A.FP64, 0), (A.FP64, 2), (A.MIXED, 1), (A.MIXED, 2)])
def test_nan_load_reports_nan(ctx, prec, precond):
    mesh, loads = plate()
    loads.rp[2 * mesh.borders[1][3]] = np.nan
    print("case: NaN load, precision", prec, "precond", precond, flush=True)
    P = A.Problem(ctx, mesh, loads, A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101))
    if precond == 2:
        P.set_grid(31, 13)
    with pytest.raises(A.CPError) as e:
        P.solve(A.options(cg_tol=1e-9, precision=prec, precond=precond, max_outer=20))
    assert e.value.code == A.CP_ERR_NAN
    assert "NaN" in str(e.value)
    P.close()


@pytest.mark.parametrize("prec,precond", [(A.FP64, 0), (A.FP64, 1), (A.MIXED, 1)])
def test_unconstrained_body_stops_at_cap(ctx, prec, precond):
    mesh, loads = plate(fix=False)              # traction on x = 1 only: no support, no balance
    assert loads.mask.all()
    print("case: unconstrained body, precision", prec, "precond", precond, flush=True)
    P = A.Problem(ctx, mesh, loads, A.material(E, 0.3))
    st = A.Stats()
    code = A.lib().cp_solve_elastic(P.ptr, A.C.byref(A.options(cg_tol=1e-9, cg_maxit=300, precision=prec,
                                                                 precond=precond)), A.C.byref(st))
    msg = A.lib().cp_last_error().decode()
    print("status", code, "cg", st.cg_iters, msg, flush=True)
    assert code == A.CP_ERR_NOT_CONVERGED and st.status == A.CP_ERR_NOT_CONVERGED
    assert "did not converge" in msg or "stalled" in msg or "without convergence" in msg
    P.close()


@pytest.mark.parametrize("prec", [A.FP64, A.MIXED])
def test_unconverged_linear_solves_do_not_end_a_level(ctx, prec):
    """Below yield the secant exit measure is 0 for any u, so before the
    linear solve's own status was taken into account a level with 5 CG
    iterations per solve ended as converged."""
    mesh, loads = plate(S=50.0)
    print("case: 50 MPa (elastic), cg_maxit=5, precision", prec, flush=True)
    P = A.Problem(ctx, mesh, loads, A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101))
    st = A.Stats()
    code = A.lib().cp_solve(P.ptr, A.C.byref(A.options(cg_tol=1e-9, cg_maxit=5, max_outer=4, precision=prec,
                                                         inner_tol=1e-3)), A.C.byref(st))
    msg = A.lib().cp_last_error().decode()
    print("status", code, "outer", st.outer_iters, "rel_err", st.rel_err, msg, flush=True)
    assert code == A.CP_ERR_NOT_CONVERGED and st.status == A.CP_ERR_NOT_CONVERGED
    assert st.outer_iters == 4
    # the same level with enough iterations converges at its first secant iteration
    st = P.solve(A.options(cg_tol=1e-9, max_outer=4, precision=prec))
    assert st.status == A.CP_OK and st.outer_iters == 1
    P.close()


@pytest.mark.parametrize("prec,precond", [(A.FP64, 0), (A.MIXED, 2)])
def test_zero_load_converges(ctx, prec, precond):
    mesh, loads = plate(S=0.0)
    assert not loads.rp.any()
    print("case: zero load, precision", prec, "precond", precond, flush=True)
    P = A.Problem(ctx, mesh, loads, A.material(E, 0.3, 240.0, A.LAW_SWIFT, n=0.2))
    if precond == 2:
        P.set_grid(31, 13)
    st = P.solve(A.options(cg_tol=1e-9, precision=prec, precond=precond, tol=1e-6, aitken=True))
    assert st.status == A.CP_OK and st.rel_err == 0.0 and st.outer_iters == 1
    assert not P.field("u").any() and not P.field("sigI").any()
    assert (P.field("psi") == 1.0).all()
    P.close()
