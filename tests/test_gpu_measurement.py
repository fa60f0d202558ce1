"""The measurement entry points on the GPU: the bandwidth yardstick
(cp_bench_copy) reports physically possible figures and refuses bad sizes,
and the timing entry points of the roofline kernel (cp_bench_mg) and of the
CG's operator product (cp_bench_mg_spmv) return byte counts of the level-0
storage."""
import ctypes as C

import pytest

from cudaplasticity_amd import abi as A

pytestmark = pytest.mark.gpu
PEAK_GBS = 8000.0


@pytest.fixture(scope="module")
def ctx():
    c = A.Context(0)
    yield c
    c.close()


def test_copy_yardstick_is_below_peak(ctx):
    """1 GiB buffers, four times the Infinity Cache: a device copy and a
    read-only pass both run below the 8 TB/s HBM peak and above 2 TB/s."""
    n = 1 << 30
    ms, msr = A.bench_copy(ctx, n, 5)
    copy, read = 2 * n / ms / 1e6, n / msr / 1e6
    print("copy %.0f GB/s, read %.0f GB/s" % (copy, read), flush=True)
    assert 2000.0 < copy < PEAK_GBS and 2000.0 < read < PEAK_GBS
    with pytest.raises(A.CPError) as e:
        A.bench_copy(ctx, 1000, 1)                   # not a multiple of 16
    assert e.value.code == A.CP_ERR_INVALID


def test_smoothing_step_bytes_match_the_storage_formula(ctx):
    """cp_bench_mg's byte count of the finest smoothing step lies between
    the stored level-0 operator and that operator plus eight vectors, and
    the V-cycle takes longer than one smoothing step.  On a lattice numbered
    node by node the FP32 step reads the symmetric lattice layout: five
    16-byte blocks per node (64-node slices), fewer bytes than the sliced-ELL
    rows (a 16-byte block and a 4-byte column for each of up to 9 blocks)."""
    mesh = A.gen_quarter_plate(1.0, 0.1, 64, 32)
    loads = A.Loads(mesh).force(1, 100.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    P = A.Problem(ctx, mesh, loads, A.material(2.1e5, 0.3))
    P.set_grid(65, 33)
    P.solve(A.options(cg_tol=1e-8, precision=A.MIXED, precond=2), elastic=True)
    v, sm, b = C.c_double(), C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_mg(P.ptr, A.MIXED, 10, C.byref(v), C.byref(sm), C.byref(b)))
    lay = P.layout()
    print("V-cycle %.3f ms, smoothing step %.4f ms, %d bytes" % (v.value, sm.value, b.value), flush=True)
    assert v.value > sm.value > 0.0
    sym = 5 * 16.0 * 64 * ((mesh.nn + 63) // 64)
    assert sym <= b.value <= sym + 8 * 8.0 * mesh.nn + 64 * 1024
    assert b.value < 20.0 * lay["nblk_stored"]
    P.close()


def test_cg_operator_product_bytes_and_results(ctx):
    """cp_bench_mg_spmv times the multigrid CG's w = mask (A u) kernel on the
    symmetric lattice layout: its byte count is the five 16-byte blocks per
    node (64-node slices) plus three float2 vectors, the two-byte mask and
    the three partial sums of each 256-node block; zero launches are
    refused."""
    mesh = A.gen_quarter_plate(1.0, 0.1, 64, 32)
    loads = A.Loads(mesh).force(1, 100.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    P = A.Problem(ctx, mesh, loads, A.material(2.1e5, 0.3))
    P.set_grid(65, 33)
    P.solve(A.options(cg_tol=1e-8, precision=A.MIXED, precond=2), elastic=True)
    ms, b = C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_mg_spmv(P.ptr, A.MIXED, 10, C.byref(ms), C.byref(b)))
    nb = (mesh.nn + 255) // 256
    expect = 5 * 16.0 * 64 * ((mesh.nn + 63) // 64) + 3 * 8.0 * mesh.nn + 2.0 * mesh.nn + 24.0 * nb
    print("operator product %.4f ms, %d bytes" % (ms.value, b.value), flush=True)
    assert ms.value > 0.0
    assert b.value == pytest.approx(expect, rel=1e-12)
    with pytest.raises(A.CPError) as e:
        A.check(A.lib().cp_bench_mg_spmv(P.ptr, A.MIXED, 0, C.byref(ms), C.byref(b)))
    assert e.value.code == A.CP_ERR_INVALID
    P.close()
