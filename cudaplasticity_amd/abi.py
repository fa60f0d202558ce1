"""ctypes binding of the C-ABI in include/cp_abi.h (libcp_hip.so).

The library is the product: every call here goes straight to the gfx950
kernels or to the library's own host-side setup code.  Loading fails loudly
when the library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libcp_hip.so")
HEADER = os.path.join(ROOT, "include", "cp_abi.h")

CP_OK, CP_ERR_INVALID, CP_ERR_NOMEM, CP_ERR_DEVICE, CP_ERR_NOT_CONVERGED, CP_ERR_NAN, CP_ERR_NO_DEVICE, \
    CP_ERR_BUDGET, CP_ERR_STATE = range(9)
LAW_LINEAR, LAW_POWER, LAW_SWIFT = 0, 1, 2
PLANE_STRESS, PLANE_STRAIN = 0, 1
FP64, MIXED = 0, 1
FIELDS = ["exx", "eyy", "gxy", "epsI", "sxx", "syy", "tau", "sigI", "sigTab", "psi", "Ec", "nuc",
          "exx_p", "eyy_p", "gxy_p", "epsI_p"]
F_U, F_RHS = 100, 101

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)


class Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("h", C.c_double), ("sigmaT", C.c_double),
                ("law", C.c_int), ("plane", C.c_int), ("K_T", C.c_double), ("A", C.c_double),
                ("m", C.c_double), ("n", C.c_double)]


class Options(C.Structure):
    _fields_ = [("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("tol", C.c_double), ("max_outer", C.c_int),
                ("fixed_outer", C.c_int), ("aitken", C.c_int), ("precision", C.c_int), ("warm", C.c_int),
                ("inner_tol", C.c_double), ("precond", C.c_int), ("predictor", C.c_int),
                ("cg_forcing", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("status", C.c_int), ("cg_iters", C.c_longlong),
                ("refinements", C.c_longlong), ("rel_err", C.c_double), ("seconds", C.c_double),
                ("cg_seconds", C.c_double), ("asm_seconds", C.c_double), ("rm_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_slots", C.c_int), ("batch_cg", C.c_int),
                ("batch_mg", C.c_int), ("batch_mg_max_nodes", C.c_int), ("reserved", C.c_int * 3)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("steps_done", C.c_int), ("cg_iters", C.c_longlong),
                ("outer_iters", C.c_int), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


class CPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    lib.cp_last_error.restype = C.c_char_p
    lib.cp_abi_version.restype = C.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def check(code):
    if code != CP_OK:
        raise CPError(code, lib().cp_last_error().decode())
    return code


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def u8p(a):
    return a.ctypes.data_as(_U8)


def material(E, nu, sigmaT=1e300, law=LAW_LINEAR, K_T=None, A=0.0, m=0.0, n=0.0, h=1.0, plane=PLANE_STRESS):
    return Material(E, nu, h, sigmaT, law, plane, E if K_T is None else K_T, A, m, n)


def options(cg_tol=1e-7, cg_maxit=0, tol=1e-5, max_outer=1000, fixed_outer=0, aitken=False, precision=FP64,
            warm=False, inner_tol=1e-3, precond=False, predictor=0, cg_forcing=0.0):
    """Defaults are the reference's (unpreconditioned CG to 1e-7, exit 1e-5)."""
    o = Options()
    o.cg_tol, o.cg_maxit, o.tol, o.max_outer, o.fixed_outer = cg_tol, cg_maxit, tol, max_outer, fixed_outer
    o.aitken, o.precision, o.warm, o.inner_tol = int(aitken), precision, int(warm), inner_tol
    o.precond, o.predictor, o.cg_forcing = int(precond), int(predictor), cg_forcing
    return o


# ------------------------------------------------------------------ host-side setup (library code)

class Mesh:
    """Node coordinates, connectivity (ne x npe, reference local order) and
    borders.  npe = 4 (bilinear quads, elem4) or 8 (8-node quads, elem8:
    corners, then the mid-sides of edges 1-2, 2-3, 3-4, 4-1); sec marks the
    mid-side nodes of an 8-node mesh (Mesh::secOrdNodes)."""

    def __init__(self, x, y, conn, borders, npe=4, sec=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.npe = npe
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, npe)
        self.borders = [np.ascontiguousarray(b, dtype=np.int32) for b in borders]
        self.sec = None if sec is None else np.ascontiguousarray(sec, dtype=np.uint8)

    @property
    def nn(self):
        return self.x.size

    @property
    def ne(self):
        return self.conn.shape[0]


def _gen(sizes_name, sizes_args, gen_name, gen_args):
    L = lib()
    nn, ne, nb = C.c_int(), C.c_int(), C.c_int()
    blen = np.zeros(8, dtype=np.int32)
    check(getattr(L, sizes_name)(*sizes_args, C.byref(nn), C.byref(ne), C.byref(nb), ip(blen)))
    x, y = np.zeros(nn.value), np.zeros(nn.value)
    conn = np.zeros(4 * ne.value, dtype=np.int32)
    bord = np.zeros(int(blen[:nb.value].sum()), dtype=np.int32)
    check(getattr(L, gen_name)(*gen_args, dp(x), dp(y), ip(conn), ip(bord), ip(blen)))
    borders, off = [], 0
    for b in range(nb.value):
        borders.append(bord[off:off + blen[b]].copy())
        off += blen[b]
    return Mesh(x, y, conn, borders)


def _gen2(sizes_name, sizes_args, gen_name, gen_args):
    """Order-2 generators: 8 nodes per element and the mid-side flags."""
    L = lib()
    nn, ne, nb = C.c_int(), C.c_int(), C.c_int()
    blen = np.zeros(8, dtype=np.int32)
    check(getattr(L, sizes_name)(*sizes_args, C.byref(nn), C.byref(ne), C.byref(nb), ip(blen)))
    x, y = np.zeros(nn.value), np.zeros(nn.value)
    conn = np.zeros(8 * ne.value, dtype=np.int32)
    sec = np.zeros(nn.value, dtype=np.uint8)
    bord = np.zeros(int(blen[:nb.value].sum()), dtype=np.int32)
    check(getattr(L, gen_name)(*gen_args, dp(x), dp(y), ip(conn), u8p(sec), ip(bord), ip(blen)))
    borders, off = [], 0
    for b in range(nb.value):
        borders.append(bord[off:off + blen[b]].copy())
        off += blen[b]
    return Mesh(x, y, conn, borders, npe=8, sec=sec)


_f = C.c_double


def gen_rect(x1, x2, y1, y2, N1, N2, order=1):
    """Mesh::genRectangle (Mesh.cu:17-109); order 2: 8-node quads (remapOrder)."""
    a = (_f(x1), _f(x2), _f(y1), _f(y2), N1, N2)
    if order == 2:
        return _gen2("cp_mesh_rect2_sizes", (N1, N2), "cp_mesh_rect2", a)
    return _gen("cp_mesh_rect_sizes", (N1, N2), "cp_mesh_rect", a)


def gen_rect_with_hole(x1, x2, y1, y2, r, N1, N2, order=1):
    """Mesh::genRectWithHole (Mesh.cu:112-245); order 2: 8-node quads."""
    a = (_f(x1), _f(x2), _f(y1), _f(y2), _f(r), N1, N2)
    if order == 2:
        return _gen2("cp_mesh_hole2_sizes", a, "cp_mesh_hole2", a)
    return _gen("cp_mesh_hole_sizes", a, "cp_mesh_hole", a)


def gen_ring(a, b, Nphi, Nr, order=1):
    """Mesh::genRing (Mesh.cu:249-329); order 2: 8-node quads."""
    if order == 2:
        return _gen2("cp_mesh_ring2_sizes", (Nphi, Nr), "cp_mesh_ring2", (_f(a), _f(b), Nphi, Nr))
    return _gen("cp_mesh_ring_sizes", (Nphi, Nr), "cp_mesh_ring", (_f(a), _f(b), Nphi, Nr))


def load_inp(path):
    """Mesh::loadFromFile (Mesh.cu:725-937): an .inp file of S4 (4-node) or
    S8R (8-node) quads with *MPC borders."""
    L = lib()
    p = os.fsencode(path)
    nn, n3, n4, n8, nb = (C.c_int() for _ in range(5))
    check(L.cp_mesh_inp_sizes(p, C.byref(nn), C.byref(n3), C.byref(n4), C.byref(n8), C.byref(nb), None, 0))
    blen = np.zeros(max(nb.value, 1), dtype=np.int32)
    check(L.cp_mesh_inp_sizes(p, C.byref(nn), C.byref(n3), C.byref(n4), C.byref(n8), C.byref(nb), ip(blen),
                              nb.value))
    if n3.value or (n4.value and n8.value):
        raise ValueError("%s: the solver takes meshes of 4-node or of 8-node quads only" % path)
    x, y = np.zeros(nn.value), np.zeros(nn.value)
    c3 = np.zeros(1, dtype=np.int32)
    c4 = np.zeros(max(4 * n4.value, 1), dtype=np.int32)
    c8 = np.zeros(max(8 * n8.value, 1), dtype=np.int32)
    sec = np.zeros(max(nn.value, 1), dtype=np.uint8)
    bord = np.zeros(max(int(blen[:nb.value].sum()), 1), dtype=np.int32)
    check(L.cp_mesh_load_inp(p, dp(x), dp(y), ip(c3), ip(c4), ip(c8), u8p(sec), ip(bord)))
    borders, off = [], 0
    for b in range(nb.value):
        borders.append(bord[off:off + blen[b]].copy())
        off += blen[b]
    if n8.value:
        return Mesh(x, y, c8[:8 * n8.value], borders, npe=8, sec=sec[:nn.value])
    return Mesh(x, y, c4[:4 * n4.value], borders)


def gen_quarter_plate(L, r, Nt, Nr):
    return _gen("cp_mesh_quarter_plate_sizes", (Nt, Nr), "cp_mesh_quarter_plate", (_f(L), _f(r), Nt, Nr))


def gen_quarter_ring(a, b, Nt, Nr, ratio):
    return _gen("cp_mesh_quarter_ring_sizes", (Nt, Nr), "cp_mesh_quarter_ring", (_f(a), _f(b), Nt, Nr, _f(ratio)))


def renum_by_direction(mesh, dx=1.0, dy=1.0, with_perm=False):
    """Mesh::renumByDirection (Mesh.cu:537-630): a copy of `mesh` with its
    nodes renumbered by increasing x dx + y dy (with_perm: and the new
    number of every old node)."""
    x, y, conn = mesh.x.copy(), mesh.y.copy(), mesh.conn.copy()
    bord = np.concatenate(mesh.borders).astype(np.int32) if mesh.borders else np.zeros(0, dtype=np.int32)
    perm = np.zeros(mesh.nn, dtype=np.int32)
    check(lib().cp_mesh_renum_by_direction(mesh.nn, dp(x), dp(y), mesh.ne, mesh.npe, ip(conn), bord.size, ip(bord),
                                           _f(dx), _f(dy), ip(perm)))
    out, off = [], 0
    for b in mesh.borders:
        out.append(bord[off: off + b.size].copy())
        off += b.size
    sec = None
    if mesh.sec is not None:
        sec = np.zeros(mesh.nn, dtype=np.uint8)
        sec[conn[:, 4:].ravel()] = 1
    m = Mesh(x, y, conn, out, npe=mesh.npe, sec=sec)
    return (m, perm) if with_perm else m


def ring_lattice(mesh, border_n):
    """cp_mesh_ring_lattice: (n1, n2, node_of) of a mesh numbered by
    renumerateRing (periodic along direction 1)."""
    n1, n2 = C.c_int(), C.c_int()
    node_of = np.zeros(mesh.nn, dtype=np.int32)
    check(lib().cp_mesh_ring_lattice(mesh.nn, int(border_n), C.byref(n1), C.byref(n2), ip(node_of)))
    return n1.value, n2.value, node_of


def corner_lattice2(kind, *args):
    """cp_mesh_{rect2,hole2,ring2}_lattice: (m1, m2, corner_of) of an 8-node
    generator's corner nodes; args as the generator takes them (rect2:
    N1, N2; hole2: x1, x2, y1, y2, r, N1, N2; ring2: Nphi, Nr)."""
    fn = getattr(lib(), f"cp_mesh_{kind}2_lattice")
    conv = [_f(a) if isinstance(a, float) else int(a) for a in args]
    m1, m2 = C.c_int(), C.c_int()
    check(fn(*conv, C.byref(m1), C.byref(m2), None))
    corner_of = np.zeros(m1.value * m2.value, dtype=np.int32)
    check(fn(*conv, C.byref(m1), C.byref(m2), ip(corner_of)))
    return m1.value, m2.value, corner_of


def renumerate_ring(mesh, border_n):
    """Mesh::renumerateRing (Mesh.cu:332-384) on a 4-node mesh: a renumbered copy."""
    x, y, conn = mesh.x.copy(), mesh.y.copy(), mesh.conn.copy()
    blen = np.array([b.size for b in mesh.borders], dtype=np.int32)
    bord = np.concatenate(mesh.borders).astype(np.int32) if mesh.borders else np.zeros(0, dtype=np.int32)
    check(lib().cp_mesh_renumerate_ring(mesh.nn, mesh.ne, dp(x), dp(y), ip(conn), blen.size, ip(bord), ip(blen),
                                        int(border_n), None))
    out, off = [], 0
    for n in blen:
        out.append(bord[off: off + n].copy())
        off += n
    return Mesh(x, y, conn, out)


def csr_pattern(mesh):
    nnz = C.c_longlong()
    check(lib().cp_csr_nnz_n(mesh.nn, mesh.ne, mesh.npe, ip(mesh.conn), C.byref(nnz)))
    rows = np.zeros(2 * mesh.nn + 1, dtype=np.int32)
    cols = np.zeros(nnz.value, dtype=np.int32)
    check(lib().cp_csr_pattern_n(mesh.nn, mesh.ne, mesh.npe, ip(mesh.conn), ip(rows), ip(cols)))
    return rows, cols


class Loads:
    """Load vector (at load factor 1), Dirichlet mask and prescribed values,
    built by the library's initConditions rules (PlasticitySolver.cu:1007-1216)."""

    def __init__(self, mesh, h=1.0):
        self.mesh, self.h = mesh, h
        self.rp = np.zeros(2 * mesh.nn)
        self.mask = np.ones(2 * mesh.nn, dtype=np.uint8)
        self.u = np.zeros(2 * mesh.nn)

    def force(self, border_id, fx, fy, norm_orient=False):
        b = self.mesh.borders[border_id]
        fv = np.tile(np.array([fx, fy], dtype=np.float64), b.size)
        return self.force_values(border_id, fv, norm_orient)

    def force_values(self, border_id, fv, norm_orient=False):
        b = self.mesh.borders[border_id]
        fv = np.ascontiguousarray(fv, dtype=np.float64).reshape(-1)
        sec = None if self.mesh.sec is None else u8p(self.mesh.sec)
        check(lib().cp_rhs_border_force_q(dp(self.mesh.x), dp(self.mesh.y), ip(b), b.size, sec, dp(fv),
                                          int(norm_orient), _f(self.h), dp(self.rp)))
        return self

    def gravity(self, g=9.81, rho=1.0):
        if self.mesh.npe == 8:
            rn = np.tile(np.array([0.0, -g]), self.mesh.nn)
            check(lib().cp_rhs_body_force_nodal8(self.mesh.ne, dp(self.mesh.x), dp(self.mesh.y), ip(self.mesh.conn),
                                                 dp(rn), _f(rho), _f(self.h), dp(self.rp)))
            return self
        check(lib().cp_rhs_body_force(self.mesh.ne, dp(self.mesh.x), dp(self.mesh.y), ip(self.mesh.conn),
                                      _f(0.0), _f(-g), _f(rho), _f(self.h), dp(self.rp)))
        return self

    def displacement(self, border_id, ux, uy, fix_mid_x=False, fix_mid_y=False):
        b = self.mesh.borders[border_id]
        check(lib().cp_displacement_border(ip(b), b.size, _f(ux), _f(uy), int(fix_mid_x), int(fix_mid_y),
                                           u8p(self.mask), dp(self.u)))
        return self

    def fix_vert_axis(self, x):
        check(lib().cp_fix_axis(self.mesh.nn, dp(self.mesh.x), dp(self.mesh.y), 1, _f(x), u8p(self.mask)))
        return self

    def fix_hor_axis(self, y):
        check(lib().cp_fix_axis(self.mesh.nn, dp(self.mesh.x), dp(self.mesh.y), 0, _f(y), u8p(self.mask)))
        return self


# ------------------------------------------------------------------ device objects

class Context:
    def __init__(self, device=0):
        self.ptr = C.c_void_p()
        check(lib().cp_ctx_create(device, C.byref(self.ptr)))

    def info(self):
        name = C.create_string_buffer(256)
        mem, cus = C.c_size_t(), C.c_int()
        check(lib().cp_device_info(self.ptr, name, 256, C.byref(mem), C.byref(cus)))
        return {"name": name.value.decode(), "total_mem": mem.value, "cu_count": cus.value}

    def close(self):
        if self.ptr:
            lib().cp_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()


class Problem:
    """One device-resident problem (PlasticitySolver construction)."""

    def __init__(self, ctx, mesh, loads, mat):
        self.mesh, self.mat = mesh, mat
        self.ptr = C.c_void_p()
        check(lib().cp_problem_create_n(ctx.ptr, mesh.npe, mesh.nn, mesh.ne, dp(mesh.x), dp(mesh.y),
                                        ip(mesh.conn), dp(loads.rp), dp(loads.u), u8p(loads.mask), C.byref(mat),
                                        C.byref(self.ptr)))

    def close(self):
        if self.ptr:
            lib().cp_problem_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csr(self):
        nnz = C.c_longlong()
        check(lib().cp_problem_csr_nnz(self.ptr, C.byref(nnz)))
        rows = np.zeros(2 * self.mesh.nn + 1, dtype=np.int32)
        cols = np.zeros(nnz.value, dtype=np.int32)
        check(lib().cp_problem_get_csr(self.ptr, ip(rows), ip(cols)))
        return rows, cols

    def values(self):
        nnz = C.c_longlong()
        check(lib().cp_problem_csr_nnz(self.ptr, C.byref(nnz)))
        v = np.zeros(nnz.value)
        check(lib().cp_problem_get_values(self.ptr, dp(v)))
        return v

    def layout(self):
        a, b, w = C.c_longlong(), C.c_longlong(), C.c_int()
        check(lib().cp_problem_layout(self.ptr, C.byref(a), C.byref(b), C.byref(w)))
        return {"nblk_stored": a.value, "nblk": b.value, "max_width": w.value}

    def set_grid(self, n1, n2, node_of=None, periodic=False):
        """Logical lattice of a structured mesh (for precond=2, multigrid);
        periodic: direction 1 closes on itself (rings, plates with a hole)."""
        no = None if node_of is None else np.ascontiguousarray(node_of, dtype=np.int32)
        fn = lib().cp_problem_set_grid_periodic if periodic else lib().cp_problem_set_grid
        check(fn(self.ptr, int(n1), int(n2), ip(no) if no is not None else None))

    def set_graphs(self, enable):
        check(lib().cp_problem_set_graphs(self.ptr, int(enable)))

    def mg_levels(self):
        n = C.c_int()
        dims = (C.c_int * 128)()
        check(lib().cp_problem_mg_levels(self.ptr, C.byref(n), dims))
        return [(dims[2 * i], dims[2 * i + 1]) for i in range(n.value)]

    def set_mg_rebuild(self, every):
        check(lib().cp_problem_set_mg_rebuild(self.ptr, int(every)))

    def mg_level0_sym(self):
        """True where the finest level's FP32 operator uses the symmetric
        lattice layout (five blocks per node)."""
        v = C.c_int()
        check(lib().cp_problem_mg_level0_layout(self.ptr, C.byref(v)))
        return bool(v.value)

    def mg_smoothers(self):
        """Per level: 0 = 2x2 block Jacobi, 1 = segments along direction 1."""
        n = len(self.mg_levels())
        kinds = (C.c_int * max(n, 1))()
        check(lib().cp_problem_mg_smoothers(self.ptr, kinds))
        return [kinds[i] for i in range(n)]

    def set_load(self, factor):
        check(lib().cp_set_load(self.ptr, _f(factor)))

    def reset(self):
        check(lib().cp_reset_state(self.ptr))

    def advance_load(self, factor, predictor=True):
        check(lib().cp_advance_load(self.ptr, _f(factor), int(predictor)))

    def assemble(self, precision=FP64):
        check(lib().cp_assemble(self.ptr, precision))

    def cg(self, tol=1e-7, maxit=0, precision=FP64, inner_tol=1e-3):
        it, rr = C.c_longlong(), C.c_double()
        check(lib().cp_cg(self.ptr, _f(tol), maxit, precision, _f(inner_tol), C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def return_map(self):
        r = C.c_double()
        check(lib().cp_return_map(self.ptr, C.byref(r)))
        return r.value

    def plastic_strains(self):
        check(lib().cp_plastic_strains(self.ptr))

    def solve(self, opt=None, elastic=False, allow=(CP_OK,)):
        st = Stats()
        fn = lib().cp_solve_elastic if elastic else lib().cp_solve
        code = fn(self.ptr, C.byref(opt) if opt is not None else None, C.byref(st))
        if code not in allow:
            check(code)
        return st

    def solve_steps(self, nsteps, final_factor=1.0, opt=None):
        per = (Stats * nsteps)()
        check(lib().cp_solve_steps(self.ptr, nsteps, _f(final_factor), C.byref(opt) if opt is not None else None,
                                   per))
        return [s for s in per]

    def field(self, name):
        if name == "u":
            out = np.zeros(2 * self.mesh.nn)
            check(lib().cp_download(self.ptr, F_U, dp(out)))
            return out
        if name == "rhs":
            out = np.zeros(2 * self.mesh.nn)
            check(lib().cp_download(self.ptr, F_RHS, dp(out)))
            return out
        out = np.zeros(self.mesh.ne)
        check(lib().cp_download(self.ptr, FIELDS.index(name), dp(out)))
        return out

    def fields(self):
        return {n: self.field(n) for n in FIELDS}

    def upload_state(self, u=None, psi=None):
        check(lib().cp_upload_state(self.ptr, dp(np.ascontiguousarray(u, dtype=np.float64)) if u is not None else None,
                                    dp(np.ascontiguousarray(psi, dtype=np.float64)) if psi is not None else None))

    def internal_force(self):
        f = np.zeros(2 * self.mesh.nn)
        check(lib().cp_internal_force(self.ptr, dp(f)))
        return f

    def internal_force_noalloc(self):
        """cp_internal_force_noalloc: the same force, staged in the library's
        static device array (no device allocation)."""
        f = np.zeros(2 * self.mesh.nn)
        check(lib().cp_internal_force_noalloc(self.ptr, dp(f)))
        return f

    def last_cg_ms_per_iter(self):
        ms = C.c_double()
        check(lib().cp_last_cg_kernel_ms(self.ptr, C.byref(ms)))
        return ms.value

    def bench_spmv(self, precision=FP64, reps=50):
        ms, bs, ba = C.c_double(), C.c_double(), C.c_double()
        check(lib().cp_bench_spmv(self.ptr, precision, reps, C.byref(ms), C.byref(bs), C.byref(ba)))
        return ms.value, bs.value, ba.value


def bench_copy(ctx, nbytes, reps=20):
    """Average ms of a plain device copy of nbytes (read + write once) and of
    a read-only pass over nbytes."""
    ms, msr = C.c_double(), C.c_double()
    check(lib().cp_bench_copy(ctx.ptr, C.c_size_t(nbytes), reps, C.byref(ms), C.byref(msr)))
    return ms.value, msr.value


def declared_functions(header=HEADER):
    """Names of every function declared in include/cp_abi.h."""
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cp_\w+)\s*\(", txt, re.M)))


class Server:
    """Continuous-batching serving loop (cp_server_*)."""

    def __init__(self, ctx, hbm_budget, max_slots=16, batch_cg=True, batch_mg=False, batch_mg_max_nodes=0):
        cfg = ServerConfig()
        cfg.hbm_budget, cfg.max_slots, cfg.batch_cg = int(hbm_budget), int(max_slots), int(batch_cg)
        cfg.batch_mg, cfg.batch_mg_max_nodes = int(batch_mg), int(batch_mg_max_nodes)
        self.ptr = C.c_void_p()
        check(lib().cp_server_create(ctx.ptr, C.byref(cfg), C.byref(self.ptr)))
        self.sizes = {}
        self.last_message = ""   # message of the last failed result()

    def submit(self, mesh, loads, mat, nsteps, final_factor, opt, grid=None, node_of=None, periodic=False):
        """grid = (n1, n2): the mesh's lattice, needed for multigrid CG (precond 2);
        periodic: direction 1 closes on itself (rings, plates with a hole)."""
        rid = C.c_longlong()
        if grid is None:
            check(lib().cp_server_submit(self.ptr, mesh.nn, mesh.ne, dp(mesh.x), dp(mesh.y), ip(mesh.conn),
                                         dp(loads.rp), dp(loads.u), u8p(loads.mask), C.byref(mat), int(nsteps),
                                         _f(final_factor), C.byref(opt), C.byref(rid)))
        else:
            no = None if node_of is None else np.ascontiguousarray(node_of, dtype=np.int32)
            fn = lib().cp_server_submit_grid_periodic if periodic else lib().cp_server_submit_grid
            check(fn(self.ptr, mesh.nn, mesh.ne, dp(mesh.x), dp(mesh.y), ip(mesh.conn), dp(loads.rp), dp(loads.u),
                     u8p(loads.mask), C.byref(mat), int(nsteps), _f(final_factor), C.byref(opt), int(grid[0]),
                     int(grid[1]), ip(no) if no is not None else None, C.byref(rid)))
        self.sizes[rid.value] = (mesh.nn, mesh.ne)
        return rid.value

    def batch_stats(self):
        """(batches, members summed over batches, lockstep iterations) of the
        lockstep multigrid batches run so far."""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        check(lib().cp_server_batch_stats(self.ptr, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def step(self):
        n = C.c_int()
        check(lib().cp_server_step(self.ptr, C.byref(n)))
        return n.value

    def clock(self):
        t = C.c_double()
        check(lib().cp_server_clock(self.ptr, C.byref(t)))
        return t.value

    def result(self, rid, want_fields=True):
        nn, ne = self.sizes[rid]
        info = RequestInfo()
        u = np.zeros(2 * nn) if want_fields else None
        sig = np.zeros(ne) if want_fields else None
        code = lib().cp_server_result(self.ptr, C.c_longlong(rid), C.byref(info),
                                      dp(u) if want_fields else None, dp(sig) if want_fields else None)
        self.last_message = lib().cp_last_error().decode() if code != CP_OK else ""
        return code, info, u, sig

    def run_until_done(self, max_rounds=10 ** 7):
        for _ in range(max_rounds):
            if self.step() == 0:
                return
        raise RuntimeError("server did not drain")

    def close(self):
        if self.ptr:
            lib().cp_server_destroy(self.ptr)
            self.ptr = C.c_void_p()
