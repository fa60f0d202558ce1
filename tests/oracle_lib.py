"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() (as the checker)
and bench.py's cpu_baseline leg use it.  The oracle restates the reference
(SATOH-SD/CudaPlasticity) in plain C; see oracle/cp_oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)

LAW_LINEAR, LAW_POWER, LAW_SWIFT = 0, 1, 2
PLANE_STRESS, PLANE_STRAIN = 0, 1
FIELDS = ["exx", "eyy", "gxy", "epsI", "sxx", "syy", "tau", "sigI", "sigTab", "psi", "Ec", "nuc",
          "exx_p", "eyy_p", "gxy_p", "epsI_p"]


class Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("h", C.c_double), ("sigmaT", C.c_double),
                ("law", C.c_int), ("plane", C.c_int), ("K_T", C.c_double), ("A", C.c_double),
                ("m", C.c_double), ("n", C.c_double)]


class Options(C.Structure):
    _fields_ = [("cg_tol", C.c_double), ("cg_maxit", C.c_int), ("tol", C.c_double),
                ("max_outer", C.c_int), ("fixed_outer", C.c_int), ("aitken", C.c_int),
                ("cg_quirk", C.c_int), ("elastic", C.c_int), ("warm", C.c_int), ("cg_forcing", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("cg_iters", C.c_longlong), ("rel_err", C.c_double),
                ("status", C.c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = C.CDLL(LIB_PATH)
    lib.or_material_f.restype = C.c_double
    lib.or_material_f.argtypes = [C.POINTER(Material), C.c_double]
    lib.or_hole_radN.restype = C.c_int
    lib.or_csr_nnz.restype = C.c_longlong
    lib.or_max_index_diff.restype = C.c_int
    lib.or_elem_space4.restype = C.c_double
    lib.or_elem_space8.restype = C.c_double
    lib.or_csr_nnz_n.restype = C.c_longlong
    lib.or_deform_energy_n.restype = C.c_double
    lib.or_solve_n.restype = C.c_int
    lib.or_deform_energy.restype = C.c_double
    lib.or_solve.restype = C.c_int
    lib.or_cg.restype = C.c_int
    return lib


lib = _load()


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def u8p(a):
    return a.ctypes.data_as(_U8)


class Mesh:
    """Plain arrays: x, y (nn), conn (ne, npe) int32 with npe = 4 (bilinear
    quads) or 8 (8-node quads), borders (list of int32 arrays), sec (nn
    uint8, 1 on mid-side nodes; None for order 1)."""

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


def _gen(sizes_fn, sizes_args, gen_fn, gen_args):
    nn, ne, nb = C.c_int(), C.c_int(), C.c_int()
    blen = np.zeros(8, dtype=np.int32)
    sizes_fn(*sizes_args, C.byref(nn), C.byref(ne), C.byref(nb), ip(blen))
    x = np.zeros(nn.value)
    y = np.zeros(nn.value)
    conn = np.zeros(4 * ne.value, dtype=np.int32)
    bord = np.zeros(int(blen[: nb.value].sum()), dtype=np.int32)
    gen_fn(*gen_args, dp(x), dp(y), ip(conn), ip(bord), ip(blen))
    borders, off = [], 0
    for b in range(nb.value):
        borders.append(bord[off: off + blen[b]].copy())
        off += blen[b]
    return Mesh(x, y, conn, borders)


def remap_order(m, width):
    """remapOrder (Mesh.cu:443-503): order-1 mesh of doubled resolution ->
    8-node mesh; `width` = elements per strip of the order-1 numbering."""
    ne8 = m.ne // 4
    nn8 = m.nn - ne8
    x8, y8 = np.zeros(nn8), np.zeros(nn8)
    conn8 = np.zeros(8 * ne8, dtype=np.int32)
    sec = np.zeros(nn8, dtype=np.uint8)
    bord = np.concatenate(m.borders).astype(np.int32) if m.borders else np.zeros(0, dtype=np.int32)
    got = lib.or_remap_order(m.nn, m.ne, dp(m.x), dp(m.y), ip(m.conn), width, bord.size, ip(bord), dp(x8),
                             dp(y8), ip(conn8), u8p(sec))
    assert got == nn8
    borders, off = [], 0
    for b in m.borders:
        borders.append(bord[off: off + b.size].copy())
        off += b.size
    return Mesh(x8, y8, conn8, borders, npe=8, sec=sec)


def gen_rect(x1, x2, y1, y2, N1, N2, order=1):
    """genRectangle (Mesh.cu:17-109); order 2 doubles N1, N2, then remapOrder(2 N2)."""
    f = C.c_double
    k = 2 if order == 2 else 1
    m = _gen(lib.or_rect_sizes, (k * N1, k * N2), lib.or_gen_rect, (f(x1), f(x2), f(y1), f(y2), k * N1, k * N2))
    return remap_order(m, 2 * N2) if order == 2 else m


def gen_rect_with_hole(x1, x2, y1, y2, r, N1, N2, order=1):
    """genRectWithHole + renumerateRing (Mesh.cu:112-245, 332-384); order 2
    doubles N1, N2 and the radial count, then remapOrder(2 radN)."""
    f = C.c_double
    if order != 2:
        return _gen(lib.or_hole_sizes, (f(x1), f(x2), f(y1), f(y2), f(r), N1, N2),
                    lib.or_gen_rect_with_hole, (f(x1), f(x2), f(y1), f(y2), f(r), N1, N2))
    radN = 2 * lib.or_hole_radN(f(x1), f(x2), f(y1), f(y2), f(r), N1, N2)
    borderN = 4 * (N1 + N2)
    nn, ne = borderN * (radN + 1), borderN * radN
    blen = np.array([2 * N1 + 1, 2 * N2 + 1, 2 * N1 + 1, 2 * N2 + 1, borderN + 1, 0, 0, 0], dtype=np.int32)
    x, y = np.zeros(nn), np.zeros(nn)
    conn = np.zeros(4 * ne, dtype=np.int32)
    bord = np.zeros(int(blen[:5].sum()), dtype=np.int32)
    lib.or_gen_rect_with_hole_r(f(x1), f(x2), f(y1), f(y2), f(r), 2 * N1, 2 * N2, radN, dp(x), dp(y), ip(conn),
                                ip(bord), ip(blen))
    borders, off = [], 0
    for b in range(5):
        borders.append(bord[off: off + blen[b]].copy())
        off += blen[b]
    return remap_order(Mesh(x, y, conn, borders), radN)


def gen_ring(a, b, Nphi, Nr, order=1):
    """genRing + renumerateRing (Mesh.cu:249-329); order 2 doubles Nphi, Nr,
    then remapOrder(2 Nr)."""
    f = C.c_double
    k = 2 if order == 2 else 1
    m = _gen(lib.or_ring_sizes, (k * Nphi, k * Nr), lib.or_gen_ring, (f(a), f(b), k * Nphi, k * Nr))
    return remap_order(m, 2 * Nr) if order == 2 else m


def gen_quarter_plate(L, r, Nt, Nr):
    f = C.c_double
    return _gen(lib.or_quarter_plate_sizes, (Nt, Nr), lib.or_gen_quarter_plate, (f(L), f(r), Nt, Nr))


def gen_quarter_ring(a, b, Nt, Nr, ratio):
    f = C.c_double
    return _gen(lib.or_quarter_ring_sizes, (Nt, Nr), lib.or_gen_quarter_ring,
                (f(a), f(b), Nt, Nr, f(ratio)))


def material(E, nu, sigmaT=1e300, law=LAW_LINEAR, K_T=None, A=0.0, m=0.0, n=0.0, h=1.0, plane=PLANE_STRESS):
    return Material(E, nu, h, sigmaT, law, plane, E if K_T is None else K_T, A, m, n)


def material_f(mat, eps):
    return lib.or_material_f(C.byref(mat), C.c_double(eps))


def renum_by_direction(mesh, dx=1.0, dy=1.0):
    """or_renum_by_direction: a renumbered copy of the mesh."""
    x, y, conn = mesh.x.copy(), mesh.y.copy(), mesh.conn.copy()
    bord = np.concatenate(mesh.borders).astype(np.int32) if mesh.borders else np.zeros(0, dtype=np.int32)
    lib.or_renum_by_direction(mesh.nn, dp(x), dp(y), mesh.ne, mesh.npe, ip(conn), bord.size, ip(bord),
                              C.c_double(dx), C.c_double(dy))
    out, off = [], 0
    for b in mesh.borders:
        out.append(bord[off: off + b.size].copy())
        off += b.size
    return Mesh(x, y, conn, out, npe=mesh.npe, sec=mesh.sec)


def csr_pattern(mesh):
    nnz = lib.or_csr_nnz_n(mesh.nn, mesh.ne, mesh.npe, ip(mesh.conn))
    rows = np.zeros(2 * mesh.nn + 1, dtype=np.int32)
    cols = np.zeros(nnz, dtype=np.int32)
    lib.or_csr_pattern_n(mesh.nn, mesh.ne, mesh.npe, ip(mesh.conn), ip(rows), ip(cols))
    return rows, cols


def assemble(mesh, rows, cols, Ec, nuc, plane=PLANE_STRESS, h=1.0):
    vals = np.zeros(cols.size)
    lib.or_assemble_n(mesh.nn, mesh.ne, mesh.npe, dp(mesh.x), dp(mesh.y), ip(mesh.conn), ip(rows), ip(cols),
                      dp(np.ascontiguousarray(Ec, dtype=np.float64)),
                      dp(np.ascontiguousarray(nuc, dtype=np.float64)), plane, C.c_double(h), dp(vals))
    return vals


def calc_b8(mesh, e):
    B = np.zeros((9, 2, 8))
    dJ = np.zeros(9)
    lib.or_calc_b8(dp(mesh.x), dp(mesh.y), ip(mesh.conn), e, dp(B), dp(dJ))
    return B, dJ


def form_ke8(B, dJ, Cm, h=1.0):
    Ke = np.zeros(136)
    lib.or_form_ke8(dp(np.ascontiguousarray(B)), dp(np.ascontiguousarray(dJ)),
                    dp(np.ascontiguousarray(Cm, dtype=np.float64).reshape(9)), C.c_double(h), dp(Ke))
    full = np.zeros((16, 16))
    for a in range(16):
        for b in range(a + 1):
            full[a, b] = full[b, a] = Ke[a * (a + 1) // 2 + b]
    return full


def elem_space(mesh, e):
    f = lib.or_elem_space8 if mesh.npe == 8 else lib.or_elem_space4
    return f(dp(mesh.x), dp(mesh.y), ip(mesh.conn), e)


def calc_b4(mesh, e):
    B = np.zeros((4, 2, 4))
    dJ = np.zeros(4)
    lib.or_calc_b4(dp(mesh.x), dp(mesh.y), ip(mesh.conn), e, dp(B), dp(dJ))
    return B, dJ


def form_c(plane, E, nu):
    Cm = np.zeros(9)
    lib.or_form_c(plane, C.c_double(E), C.c_double(nu), dp(Cm))
    return Cm.reshape(3, 3)


def form_ke4(B, dJ, Cm, h=1.0):
    Ke = np.zeros(36)
    lib.or_form_ke4(dp(np.ascontiguousarray(B)), dp(np.ascontiguousarray(dJ)),
                    dp(np.ascontiguousarray(Cm, dtype=np.float64).reshape(9)), C.c_double(h), dp(Ke))
    full = np.zeros((8, 8))
    for a in range(8):
        for b in range(a + 1):
            full[a, b] = full[b, a] = Ke[a * (a + 1) // 2 + b]
    return full


class Problem:
    """Load vector, Dirichlet mask and prescribed displacements for a mesh,
    built with the reference's initConditions rules (PlasticitySolver.cu:812-1005)."""

    def __init__(self, mesh, h=1.0):
        self.mesh = mesh
        self.h = h
        self.rp = np.zeros(2 * mesh.nn)
        self.mask = np.ones(2 * mesh.nn, dtype=np.uint8)
        self.u = np.zeros(2 * mesh.nn)

    def _sec(self):
        return None if self.mesh.sec is None else u8p(self.mesh.sec)

    def force(self, border_id, fx, fy, norm_orient=False):
        b = self.mesh.borders[border_id]
        fv = np.tile(np.array([fx, fy], dtype=np.float64), b.size)
        lib.or_rhs_border_force_q(dp(self.mesh.x), dp(self.mesh.y), ip(b), b.size, self._sec(), dp(fv),
                                  int(norm_orient), C.c_double(self.h), dp(self.rp))
        return self

    def force_values(self, border_id, fv, norm_orient=False):
        b = self.mesh.borders[border_id]
        fv = np.ascontiguousarray(fv, dtype=np.float64).reshape(-1)
        lib.or_rhs_border_force_q(dp(self.mesh.x), dp(self.mesh.y), ip(b), b.size, self._sec(), dp(fv),
                                  int(norm_orient), C.c_double(self.h), dp(self.rp))
        return self

    def gravity(self, g=9.81, rho=1.0):
        if self.mesh.npe == 8:
            lib.or_rhs_body_force8(self.mesh.ne, dp(self.mesh.x), dp(self.mesh.y), ip(self.mesh.conn),
                                   C.c_double(0.0 * g), C.c_double(-g), C.c_double(rho), C.c_double(self.h),
                                   dp(self.rp))
            return self
        lib.or_rhs_body_force(self.mesh.ne, dp(self.mesh.x), dp(self.mesh.y), ip(self.mesh.conn),
                              C.c_double(0.0 * g), C.c_double(-g), C.c_double(rho), C.c_double(self.h),
                              dp(self.rp))
        return self

    def displacement(self, border_id, ux, uy, fix_mid_x=False, fix_mid_y=False):
        b = self.mesh.borders[border_id]
        lib.or_displacement_border(ip(b), b.size, C.c_double(ux), C.c_double(uy), int(fix_mid_x),
                                   int(fix_mid_y), u8p(self.mask), dp(self.u))
        return self

    def fix_vert_axis(self, x):
        lib.or_fix_axis(self.mesh.nn, dp(self.mesh.x), dp(self.mesh.y), 1, C.c_double(x), u8p(self.mask))
        return self

    def fix_hor_axis(self, y):
        lib.or_fix_axis(self.mesh.nn, dp(self.mesh.x), dp(self.mesh.y), 0, C.c_double(y), u8p(self.mask))
        return self


def options(cg_tol=1e-7, cg_maxit=1000000, tol=1e-5, max_outer=50, fixed_outer=0, aitken=False,
            cg_quirk=False, elastic=False, warm=False, cg_forcing=0.0):
    return Options(cg_tol, cg_maxit, tol, max_outer, fixed_outer, int(aitken), int(cg_quirk),
                   int(elastic), int(warm), cg_forcing)


def solve(prob, mat, opt, u=None, state=None, rp=None):
    """One load level of the secant loop.  Returns (u, state dict, stats)."""
    m = prob.mesh
    u = prob.u.copy() if u is None else np.ascontiguousarray(u, dtype=np.float64).copy()
    st = np.zeros((len(FIELDS), m.ne)) if state is None else np.ascontiguousarray(state).copy()
    rp = prob.rp if rp is None else np.ascontiguousarray(rp, dtype=np.float64)
    stats = Stats()
    lib.or_solve_n(m.nn, m.ne, m.npe, dp(m.x), dp(m.y), ip(m.conn), dp(rp), u8p(prob.mask), dp(u),
                   C.byref(mat), C.byref(opt), dp(st), C.byref(stats))
    return u, st, stats


def fields(st):
    return {name: st[i] for i, name in enumerate(FIELDS)}


def cg(rows, cols, vals, b, mask, x0, tol, maxit=1000000, quirk=False):
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    it = C.c_int()
    status = lib.or_cg(b.size, ip(rows), ip(cols), dp(vals), dp(b), u8p(mask), dp(x), C.c_double(tol),
                       maxit, int(quirk), C.byref(it))
    return x, it.value, status


def deform_energy(mesh, mat, st):
    return lib.or_deform_energy_n(mesh.ne, mesh.npe, dp(mesh.x), dp(mesh.y), ip(mesh.conn), C.byref(mat),
                                dp(np.ascontiguousarray(st)))


def elem_space4(mesh, e):
    return lib.or_elem_space4(dp(mesh.x), dp(mesh.y), ip(mesh.conn), e)


def solve_steps(prob, mat, nsteps, opt, final_factor=1.0, predictor=False):
    """Incremental loading: level n applies n/nsteps of the load, warm-started
    from level n-1's displacements and secant moduli; with `predictor`, u and
    log(psi) are extrapolated linearly in the load factor from the last two
    levels (the level before the first is the unloaded state); predictor 1
    extrapolates u only, 2 also log(psi)."""
    u = prob.u * 0.0
    st = None
    per = []
    lam_prev, u_prev, psi_prev = 0.0, np.zeros(2 * prob.mesh.nn), np.ones(prob.mesh.ne)
    lam_cur = 0.0
    for n in range(1, nsteps + 1):
        lam = final_factor * n / nsteps
        if st is not None and predictor:
            ratio = (lam - lam_cur) / (lam_cur - lam_prev)
            psi = st[FIELDS.index("psi")].copy()
            u_new = u + ratio * (u - u_prev)
            psi_new = psi
            if predictor >= 2:
                psi_new = np.maximum(np.exp(np.log(psi) + ratio * (np.log(psi) - np.log(psi_prev))), 1.0)
            u_prev, psi_prev = u.copy(), psi
            u = u_new
            st[FIELDS.index("psi")] = psi_new
            st[FIELDS.index("Ec")] = mat.E / psi_new
            st[FIELDS.index("nuc")] = 0.5 * (1 - (1 - 2 * mat.nu) / psi_new)
            lam_prev = lam_cur
        u = np.where(prob.mask == 0, prob.u * lam, u)
        o = Options.from_buffer_copy(opt)
        o.warm = 1 if st is not None else 0
        u, st, s = solve(prob, mat, o, u=u, state=st, rp=prob.rp * lam)
        lam_cur = lam
        per.append((s.outer_iters, s.cg_iters, s.rel_err, s.status))
    return u, st, per


lib.or_return_map_n.restype = C.c_double


def return_map(mesh, mat, u, state):
    """One secant update from u (or_return_map_n): strains, stresses and the
    plain psi / E_c / nu_c update in `state` (modified in place); returns the
    exit measure."""
    return lib.or_return_map_n(mesh.nn, mesh.ne, mesh.npe, dp(mesh.x), dp(mesh.y), ip(mesh.conn),
                               dp(np.ascontiguousarray(u, dtype=np.float64)), C.byref(mat), dp(state))


def solve_direct(prob, mat, n_secant, rp=None):
    """`n_secant` plain secant iterations of the oracle's loop (or_solve_n
    with fixed_outer = n_secant) with the masked CG replaced by a direct
    sparse solve of the same system: the oracle's CSR assembly
    (or_assemble_n), scipy's SuperLU on the free DOFs (K_ff u_f = b_f -
    K_fc u_c), the oracle's secant update (or_return_map_n).  CG run to a
    tight tolerance converges to this solution; the direct solve reaches it
    in seconds on meshes where unpreconditioned CG needs 10^5 iterations.
    Returns (u, state, [exit measure per iteration])."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    m = prob.mesh
    rows, cols = csr_pattern(m)
    N = 2 * m.nn
    rp = prob.rp if rp is None else np.ascontiguousarray(rp, dtype=np.float64)
    free = prob.mask.astype(bool)
    u = prob.u.copy()
    st = np.zeros((len(FIELDS), m.ne))
    st[FIELDS.index("psi")] = 1.0
    st[FIELDS.index("Ec")] = mat.E
    st[FIELDS.index("nuc")] = mat.nu
    rels = []
    for _ in range(n_secant):
        Ec = np.ascontiguousarray(st[FIELDS.index("Ec")])
        nuc = np.ascontiguousarray(st[FIELDS.index("nuc")])
        vals = np.zeros(cols.size)
        lib.or_assemble_n(m.nn, m.ne, m.npe, dp(m.x), dp(m.y), ip(m.conn), ip(rows), ip(cols), dp(Ec), dp(nuc),
                          mat.plane, C.c_double(mat.h), dp(vals))
        K = sp.csr_matrix((vals, cols, rows[: N + 1]), shape=(N, N))
        Kff = K[free][:, free].tocsc()
        b = rp[free] - K[free][:, ~free] @ u[~free]
        u[free] = spl.spsolve(Kff, b, permc_spec="MMD_AT_PLUS_A")
        rels.append(return_map(m, mat, u, st))
    return u, st, rels
