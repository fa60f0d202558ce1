"""cp_internal_force_noalloc (the entry point that allocates nothing) through
the C-ABI against cp_internal_force, the existing entry point it varies: the
same kernel on the same state, so bit for bit, on two calls in a row (the
second on a smaller problem, after the staging array held the first one's
force), on a fresh problem before any solve (cp_internal_force would
allocate the mixed-precision arena there) and on 8-node quads; a problem
above the staging limit is refused with CP_ERR_INVALID."""
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


def plate(Nt, Nr, S=300.0):
    mesh = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    return mesh, loads, A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)


def test_noalloc_force_equals_allocating_entry_point(ctx):
    for Nt, Nr in ((120, 60), (30, 12)):          # second call: a smaller problem
        print("case: plate %dx%d" % (Nt, Nr), flush=True)
        mesh, loads, mat = plate(Nt, Nr)
        P = A.Problem(ctx, mesh, loads, mat)
        P.upload_state(u=np.linspace(0.0, 1e-3, 2 * mesh.nn))
        f0 = P.internal_force_noalloc()           # fresh problem: no solver arena yet
        assert np.array_equal(f0, P.internal_force())
        P.solve(A.options(cg_tol=1e-10, fixed_outer=4, precision=A.MIXED, inner_tol=1e-5))
        f1 = P.internal_force_noalloc()
        f2 = P.internal_force()
        assert np.array_equal(f1, f2) and np.abs(f1).max() > 0
        P.close()


def test_noalloc_force_8node(ctx):
    m8 = A.gen_rect_with_hole(-3., 3., -2., 2., 0.5, 12, 8, order=2)
    l8 = A.Loads(m8).force(1, 170.0, 0.0).force(3, -170.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    P = A.Problem(ctx, m8, l8, A.material(193e3, 0.29, 170.0, A.LAW_POWER, A=0.01 * 193e3, m=0.015))
    P.solve(A.options(cg_tol=1e-10, fixed_outer=3, precond=1))
    assert np.array_equal(P.internal_force_noalloc(), P.internal_force())
    P.close()


def test_noalloc_refuses_problems_above_the_staging_limit(ctx):
    mesh, loads, mat = plate(800, 700)           # 801 x 701 = 561,501 nodes > 524,288
    assert mesh.nn > 524288
    P = A.Problem(ctx, mesh, loads, mat)
    with pytest.raises(A.CPError) as e:
        P.internal_force_noalloc()
    assert e.value.code == A.CP_ERR_INVALID and "524288 nodes" in str(e.value)
    P.close()
