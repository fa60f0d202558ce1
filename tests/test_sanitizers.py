"""Host code under sanitizers (CPU only).

* The library's host setup code (cp_mesh.cpp: generators, renumberings, CSR
  pattern, load and boundary-condition rules, the .inp reader) driven by
  tests/cpp/host_check.cpp, and the C++ drop-in classes
  (cudaplasticity_amd/cpp/) driven by the CPU cases of tests/cpp/api_test.cpp,
  under AddressSanitizer + UndefinedBehaviorSanitizer.
* The oracle (oracle/cp_oracle.c) driven by tests/cpp/oracle_check.c under
  ASan + UBSan, and under ThreadSanitizer with its OpenMP loops on four
  threads: LLVM's OpenMP runtime with its Archer tool, which tells TSan about
  the runtime's own synchronisation (with libgomp TSan reports every
  reduction as a race).
Every run must end with status 0 and no sanitizer report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "cpp", "_build", "san")
LLVM = "/opt/rocm/llvm"
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:", "WARNING: ThreadSanitizer",
           "SUMMARY:")


@pytest.fixture(scope="module")
def built():
    if shutil.which("make") is None:
        pytest.skip("make not available")
    need_tsan = os.path.exists(os.path.join(LLVM, "bin", "clang")) and os.path.exists(
        os.path.join(LLVM, "lib", "libarcher.so"))
    targets = ["host_check", "api_test", "oracle_check_asan"] + (["oracle_check_tsan"] if need_tsan else [])
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")] +
                          [os.path.join(SAN, t) for t in targets])
    return set(targets)


def run(args, env_extra, timeout=600):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    for rep in REPORTS:
        assert rep not in out, out[-4000:]
    return out


def test_host_setup_code_asan_ubsan(built):
    out = run([os.path.join(SAN, "host_check")], {"ASAN_OPTIONS": "detect_leaks=1"})
    assert "0 failures" in out


def test_cpp_classes_asan_ubsan(built, tmp_path):
    """The CPU cases of the C++ API driver (meshes, renumberings, loads and
    conditions, .inp files, mesh analysis and VTK output).  The HIP runtime the
    classes link is loaded but not instrumented, so leak checks are off here."""
    src, out = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    out.mkdir()
    ring = np.load(os.path.join(ROOT, "tests", "golden", "ring.npz"))
    np.ascontiguousarray(ring["points"], dtype=np.float64).tofile(str(src / "ring_points.f64"))
    np.ascontiguousarray(ring["polys"], dtype=np.int32).tofile(str(src / "ring_polys.i32"))
    from cudaplasticity_amd import abi as A
    from test_host_setup import write_inp
    write_inp(str(src / "plate8.inp"), A.gen_rect_with_hole(-3, 3, -2, 2, 0.5, 12, 8, order=2), "S8R")
    text = run([os.path.join(SAN, "api_test"), "cpu", str(out), str(src)], {"ASAN_OPTIONS": "detect_leaks=0"})
    assert "cpu cases written" in text


def test_oracle_asan_ubsan(built):
    out = run([os.path.join(SAN, "oracle_check_asan")], {"ASAN_OPTIONS": "detect_leaks=1", "OMP_NUM_THREADS": "4"})
    assert "all solves converged" in out


def test_oracle_openmp_tsan(built):
    if "oracle_check_tsan" not in built:
        pytest.skip("LLVM clang with the Archer OpenMP tool not available")
    out = run([os.path.join(SAN, "oracle_check_tsan")],
              {"TSAN_OPTIONS": "ignore_noninstrumented_modules=1 halt_on_error=1",
               "OMP_TOOL_LIBRARIES": os.path.join(LLVM, "lib", "libarcher.so"), "OMP_NUM_THREADS": "4"})
    assert "all solves converged" in out
