/*
 * cp_abi.h -- C-ABI of the MI355X (gfx950) plasticity hot path.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (CP_OK == 0); cp_last_error() gives the message of the calling thread's
 * last failure.  Nothing here throws across the boundary.
 *
 * The reference (SATOH-SD/CudaPlasticity) has no FFI: its boundary is the
 * C++ classes Mesh / Material / LoadConditions / PlasticitySolver.  The C++
 * drop-in API in include/cudaplasticity/ implements those classes on top of
 * the entry points below; each entry point names the reference interface
 * it replaces (paths relative to the reference's src/).
 */
#ifndef CP_ABI_H
#define CP_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CP_ABI_VERSION 1

/* status codes */
#define CP_OK                 0
#define CP_ERR_INVALID        1   /* bad argument / unsupported mesh */
#define CP_ERR_NOMEM          2   /* device or host allocation failed */
#define CP_ERR_DEVICE         3   /* HIP runtime error */
#define CP_ERR_NOT_CONVERGED  4   /* iteration cap reached */
#define CP_ERR_NAN            5   /* NaN/Inf residual or stress */
#define CP_ERR_NO_DEVICE      6   /* no HIP device / library built without one */
#define CP_ERR_BUDGET         7   /* request larger than the serving HBM budget */
#define CP_ERR_STATE          8   /* call out of order */

/* material laws (Material.cu:3-44; CP_LAW_SWIFT is new) */
#define CP_LAW_LINEAR  0   /* sigmaT + (eps - epsT) K_T above epsT = sigmaT/E */
#define CP_LAW_POWER   1   /* reference power law with A, m */
#define CP_LAW_SWIFT   2   /* sigma = sigmaT (1 + eps_p/eps0)^n, eps0 = sigmaT/E */

#define CP_PLANE_STRESS 0  /* PlasticitySolver.cuh:15 plainCondition::stress */
#define CP_PLANE_STRAIN 1

/* precision of the linear solve */
#define CP_FP64   0        /* FP64 values, vectors and arithmetic */
#define CP_MIXED  1        /* FP32 CG (values, SpMV, vector updates) + FP64 iterative refinement */

/* per-element fields for cp_download (PlasticitySolver.cuh:49-52) */
#define CP_F_EXX    0
#define CP_F_EYY    1
#define CP_F_GXY    2
#define CP_F_EPSI   3
#define CP_F_SXX    4
#define CP_F_SYY    5
#define CP_F_TAU    6
#define CP_F_SIGI   7
#define CP_F_SIGTAB 8
#define CP_F_PSI    9
#define CP_F_EC     10
#define CP_F_NUC    11
#define CP_F_EXXP   12
#define CP_F_EYYP   13
#define CP_F_GXYP   14
#define CP_F_EPSIP  15
#define CP_NFIELDS  16
#define CP_F_U      100    /* nodal displacements, 2*nn doubles (u0,v0,u1,v1,...) */
#define CP_F_RHS    101    /* current load vector, 2*nn doubles */

typedef struct cp_ctx cp_ctx;
typedef struct cp_problem cp_problem;
typedef struct cp_server cp_server;

/* Material.cuh:9-41 (E, nu, h, sigmaT, K_T, A, m) plus the plane condition
 * (PlasticitySolver::setPlainCondition) and the Swift exponent n. */
typedef struct {
    double E, nu, h, sigmaT;
    int law, plane;
    double K_T, A, m, n;
} cp_material;

/* Tolerances default to the reference's hard-coded values
 * (PlasticitySolver.cu:454,495,503): cg_tol 1e-7, tol 1e-5. */
typedef struct {
    double cg_tol;       /* ||r||/||b|| stop of the linear solve */
    int    cg_maxit;     /* per linear solve */
    double tol;          /* plastic exit: ||sigma_i - f(eps_i)|| / ||f(eps_i)|| */
    int    max_outer;    /* cap on secant iterations per load level */
    int    fixed_outer;  /* > 0: run exactly this many secant iterations */
    int    aitken;       /* 1: safeguarded Irons-Tuck acceleration of psi */
    int    precision;    /* CP_FP64 or CP_MIXED */
    int    warm;         /* 1: keep u and the secant moduli from the previous call */
    double inner_tol;    /* CP_MIXED: smallest relative reduction one FP32 inner solve is asked
                            for (each asks for what the FP64 residual test still needs) */
    int    precond;      /* 0: unpreconditioned CG (the reference's); 1: 2x2 block-Jacobi PCG;
                            2: CG preconditioned by a geometric multigrid V-cycle (needs
                            cp_problem_set_grid; served through cp_server_submit_grid) */
    int    predictor;    /* load stepping initial guess from the last two levels: 0 none (previous
                            level's state), 1 extrapolate u, 2 extrapolate u and log(psi) */
    double cg_forcing;   /* > 0: inexact secant iterations -- secant iteration k solves to
                            max(cg_tol, cg_forcing * rel_err_{k-1}) (rel_err_{-1} = 1); a level
                            only exits after an iteration solved at cg_tol itself */
} cp_options;

typedef struct {
    int       outer_iters;
    int       status;        /* CP_OK, CP_ERR_NOT_CONVERGED, CP_ERR_NAN */
    long long cg_iters;      /* inner CG iterations (all refinements) */
    long long refinements;   /* CP_MIXED: FP64 residual evaluations */
    double    rel_err;       /* last plastic exit measure */
    double    seconds;       /* wall time of the call */
    double    cg_seconds;    /* device time in CG kernels (HIP events) */
    double    asm_seconds;   /* device time in assembly kernels */
    double    rm_seconds;    /* device time in return-map kernels */
} cp_stats;

/* ------------------------------------------------------------------ */
/* context / device                                                     */
/* ------------------------------------------------------------------ */
/* Mesh::Mesh() device probe (Mesh.cuh:86-104). */
int         cp_ctx_create(int device, cp_ctx** out);
int         cp_ctx_destroy(cp_ctx* ctx);
const char* cp_last_error(void);
int         cp_device_info(cp_ctx* ctx, char* name, int name_len, size_t* total_mem, int* cu_count);
int         cp_abi_version(void);
/* Material::f (Material.cu:3-23; Swift law new): the stress intensity the
 * device kernels use for strain intensity eps, evaluated on the host. */
int         cp_material_f(const cp_material* mat, double eps, double* sigma);

/* ------------------------------------------------------------------ */
/* host-side mesh generators (order 1) -- Mesh.cu                      */
/* Sizes first (nn, ne, nb, blen[nb]); then the generator fills x, y    */
/* (nn each), conn (4*ne, reference local order), bord (sum blen).     */
/* ------------------------------------------------------------------ */
/* Mesh::genRectangle (Mesh.cu:17-109) */
int cp_mesh_rect_sizes(int N1, int N2, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_rect(double x1, double x2, double y1, double y2, int N1, int N2,
                 double* x, double* y, int* conn, int* bord, int* blen);
/* Mesh::genRectWithHole + renumerateRing (Mesh.cu:112-245, 332-384) */
int cp_mesh_hole_sizes(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                       int* nn, int* ne, int* nb, int* blen);
int cp_mesh_hole(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                 double* x, double* y, int* conn, int* bord, int* blen);
/* Mesh::genRing + renumerateRing (Mesh.cu:249-329) */
int cp_mesh_ring_sizes(int Nphi, int Nr, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_ring(double a, double b, int Nphi, int Nr,
                 double* x, double* y, int* conn, int* bord, int* blen);
/* New: quarter plate [0,L]^2 minus a quarter disc (borders: 0 y=0, 1 x=L, 2 y=L, 3 x=0, 4 hole) */
int cp_mesh_quarter_plate_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_quarter_plate(double L, double r, int Nt, int Nr,
                          double* x, double* y, int* conn, int* bord, int* blen);
/* New: graded quarter ring (borders: 0 inner arc, 1 theta=pi/2, 2 outer arc, 3 theta=0) */
int cp_mesh_quarter_ring_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_quarter_ring(double a, double b, int Nt, int Nr, double ratio,
                         double* x, double* y, int* conn, int* bord, int* blen);
/* Mesh::renumerateRing (Mesh.cu:332-384) on a 4-node mesh, in place: the
 * nn / borderN rows of borderN columns are numbered column by column,
 * alternately from the first and the last column (the generators of plates
 * with a hole and rings already apply it).  new_of_old (nn, may be NULL):
 * the new number of every old node. */
int cp_mesh_renumerate_ring(int nn, int ne, double* x, double* y, int* conn, int nb, int* bord, const int* blen,
                            int borderN, int* new_of_old);
/* The lattice of a mesh numbered by renumerateRing (genRectWithHole with
 * borderN = 2 (N1 + N2), genRing with borderN = N_phi, order 1): n1 =
 * borderN rays along direction 1, periodic, n2 = nn / borderN nodes per ray,
 * node_of (nn entries, may be NULL) as cp_problem_set_grid_periodic takes
 * it.  New: the reference has no multigrid. */
int cp_mesh_ring_lattice(int nn, int borderN, int* n1, int* n2, int* node_of);
/* The corner-node lattice of the 8-node generators (order 2: the order-1
 * mesh of doubled resolution regrouped by remapOrder, Mesh.cu:443-503): its
 * lattice points with both coordinates even, m1 x m2, corner_of[i1 * m2 + i2]
 * the 8-node mesh's node there (may be NULL); the input of multigrid on an
 * 8-node problem (cp_problem_set_grid / _periodic: direction 1 periodic for
 * plates with a hole and rings).  New: the reference has no multigrid. */
int cp_mesh_rect2_lattice(int N1, int N2, int* m1, int* m2, int* corner_of);
int cp_mesh_hole2_lattice(double x1, double x2, double y1, double y2, double r, int N1, int N2, int* m1, int* m2,
                          int* corner_of);
int cp_mesh_ring2_lattice(int Nphi, int Nr, int* m1, int* m2, int* corner_of);
/* Mesh::renumByDirection (Mesh.cu:537-630) in place: nodes renumbered by
 * increasing x dx + y dy, ties in decreasing original number (the order the
 * reference's merge sort, Mesh.cu:505-535, leaves them in); conn (npe = 4 or
 * 8 nodes per element) and the nbord border entries follow; new_of_old
 * (nn, may be NULL): the new number of every old node. */
int cp_mesh_renum_by_direction(int nn, double* x, double* y, int ne, int npe, int* conn, int nbord, int* bord,
                               double dx, double dy, int* new_of_old);
/* Mesh::findMaxIndexDiff (Mesh.cu:417-441) */
int cp_mesh_max_index_diff(int ne, const int* conn, int* out);
/* Mesh::remapOrder (Mesh.cu:443-503): the order-1 mesh of doubled resolution
 * (ne4 quads in strips of `width`, both counts even) regrouped into ne4 / 4
 * 8-node quads; the patch centres are dropped (nn8 = nn - ne4 / 4 nodes),
 * bord (nbord entries) is renumbered in place, sec[n] = 1 on mid-side nodes. */
int cp_mesh_remap_order(int nn, int ne4, const double* x, const double* y, const int* conn4, int width,
                        int nbord, int* bord, double* x8, double* y8, int* conn8, unsigned char* sec,
                        int* nn8);
/* Mesh::loadFromFile (Mesh.cu:725-937): an Abaqus-style .inp file with
 * *NODE (1-based "id, x, y"), *ELEMENT TYPE=S3 / S4 / S8R and *MPC blocks
 * (one border per block, its node after the first four characters of each
 * line).  nn = node lines less one reference node per border.  Sizes first
 * (blen: max_borders entries), then the arrays: conn3 (3 n3), conn4 (4 n4),
 * conn8 (8 n8), sec (nn: mid-side nodes of the S8R elements), bord (sum of
 * blen), all 0-based.  A path not ending in ".inp" or a missing file is
 * CP_ERR_INVALID. */
int cp_mesh_inp_sizes(const char* path, int* nn, int* n3, int* n4, int* n8, int* nb, int* blen,
                      int max_borders);
int cp_mesh_load_inp(const char* path, double* x, double* y, int* conn3, int* conn4, int* conn8,
                     unsigned char* sec, int* bord);
/* genRectangle / genRectWithHole / genRing with order = 2 (Mesh.cu:17-23,
 * 112-121, 249-253 + remapOrder): 8 nodes per element in conn (corners, then
 * the mid-sides of edges 1-2, 2-3, 3-4, 4-1), sec marks the mid-side nodes;
 * border lists run through the mid-side nodes. */
int cp_mesh_rect2_sizes(int N1, int N2, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_rect2(double x1, double x2, double y1, double y2, int N1, int N2,
                  double* x, double* y, int* conn, unsigned char* sec, int* bord, int* blen);
int cp_mesh_hole2_sizes(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                        int* nn, int* ne, int* nb, int* blen);
int cp_mesh_hole2(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                  double* x, double* y, int* conn, unsigned char* sec, int* bord, int* blen);
int cp_mesh_ring2_sizes(int Nphi, int Nr, int* nn, int* ne, int* nb, int* blen);
int cp_mesh_ring2(double a, double b, int Nphi, int Nr,
                  double* x, double* y, int* conn, unsigned char* sec, int* bord, int* blen);

/* ------------------------------------------------------------------ */
/* host-side loads / boundary conditions -- PlasticitySolver.cu:1007-1216 */
/* ------------------------------------------------------------------ */
/* Border traction: fval holds the force value at every border node (2 each),
 * as LoadConditions::setForce's function would return it (:1070-1103). */
int cp_rhs_border_force(const double* x, const double* y, const int* border, int len,
                        const double* fval, int norm_orient, double h, double* rp);
/* The same on order-2 borders: where sec[border[k]] is set, border[k] is a
 * mid-side node and the segment [border[k-1], border[k], border[k+1]] takes
 * the quadratic loads f L/6, 2 f L/3, f L/6 (:1025-1069).  sec may be NULL. */
int cp_rhs_border_force_q(const double* x, const double* y, const int* border, int len,
                          const unsigned char* sec, const double* fval, int norm_orient, double h,
                          double* rp);
/* Constant body force on 4-node elements (:1106-1117). */
int cp_rhs_body_force(int ne, const double* x, const double* y, const int* conn,
                      double gx, double gy, double rho, double h, double* rp);
/* Body force given at every node (rnode: 2 per node), 4-node elements
 * (:1106-1117; R is evaluated at the element's own nodes -- the reference
 * indexes node[2 * elem] there, which only constant R survives). */
int cp_rhs_body_force_nodal(int ne, const double* x, const double* y, const int* conn, const double* rnode,
                            double rho, double h, double* rp);
/* Body force given at every node on 8-node quads (:1118-1136): corners
 * -rho h A / 12, mid-sides rho h A / 3 of R at the node, A from elemSpace8. */
int cp_rhs_body_force_nodal8(int ne, const double* x, const double* y, const int* conn, const double* rnode,
                             double rho, double h, double* rp);
/* LoadConditions::fixVertAxis / fixHorAxis (:1165-1181). */
int cp_fix_axis(int nn, const double* x, const double* y, int vertical, double coord, unsigned char* mask);
/* LoadConditions::setDisplacement on a border (:1138-1164). */
int cp_displacement_border(const int* border, int len, double ux, double uy,
                           int fix_mid_x, int fix_mid_y, unsigned char* mask, double* u);
/* The same with a displacement per border entry (uval: 2 per entry), as
 * LoadConditions::setDisplacement's function would return it. */
int cp_displacement_border_values(const int* border, int len, const double* uval,
                                  int fix_mid_x, int fix_mid_y, unsigned char* mask, double* u);

/* ------------------------------------------------------------------ */
/* CSR pattern in the reference layout -- CudaSparseSLAE.cuh:303-458    */
/* ------------------------------------------------------------------ */
int cp_csr_nnz(int nn, int ne, const int* conn, long long* nnz);
int cp_csr_pattern(int nn, int ne, const int* conn, int* rows, int* cols);
/* npe = 4 or 8 nodes per element (the elem8 scan of :339-391). */
int cp_csr_nnz_n(int nn, int ne, int npe, const int* conn, long long* nnz);
int cp_csr_pattern_n(int nn, int ne, int npe, const int* conn, int* rows, int* cols);

/* ------------------------------------------------------------------ */
/* device problems -- replaces the PlasticitySolver constructor         */
/* (PlasticitySolver.cuh:236-338), CudaSparseSLAE / CudaSLAE /          */
/* ConjGradCuda construction and the host-side initConditions_ upload.  */
/* Inputs are copied; the library owns the device memory.               */
/* rp: load vector at load factor 1; u0: prescribed displacements on    */
/* fixed DOFs (scaled by the load factor too); mask: 1 free, 0 fixed.   */
/* ------------------------------------------------------------------ */
int cp_problem_create(cp_ctx* ctx, int nn, int ne, const double* x, const double* y, const int* conn,
                      const double* rp, const double* u0, const unsigned char* mask,
                      const cp_material* mat, cp_problem** out);
/* The same for npe = 4 (bilinear quads, elem4) or npe = 8 (8-node quads,
 * elem8: formKe8 / insertKe8 / strain8, PlasticitySolver.cu:66-224,
 * 1452-1490, 1551-1588).  8-node problems take block-Jacobi or plain CG
 * (precond 0 / 1); multigrid and the serving loop take 4-node meshes. */
int cp_problem_create_n(cp_ctx* ctx, int npe, int nn, int ne, const double* x, const double* y,
                        const int* conn, const double* rp, const double* u0, const unsigned char* mask,
                        const cp_material* mat, cp_problem** out);
int cp_problem_destroy(cp_problem* prob);
/* Device bytes held by a problem of this size (serving admission). */
int cp_problem_bytes(int nn, int ne, long long nnz, int precision, size_t* bytes);

/* Integer CSR structure as the reference builds it (bit-exact target). */
int cp_problem_csr_nnz(cp_problem* prob, long long* nnz);
int cp_problem_get_csr(cp_problem* prob, int* rows, int* cols);
/* FP64 values of the last FP64 assembly, in reference CSR order. */
int cp_problem_get_values(cp_problem* prob, double* vals);

/* New: the logical lattice of a structured quad mesh, for the multigrid
 * preconditioner (precond = 2): node_of[i1 * n2 + i2] is the node at lattice
 * point (i1, i2) (NULL: node = i1 * n2 + i2, the numbering of the rectangle,
 * quarter-plate and quarter-ring generators, with n1 = N1 + 1 or N_theta + 1
 * and n2 = N2 + 1 or N_r + 1).  Every pair of nodes sharing an element must
 * be lattice neighbours; the call builds the level hierarchy. */
int cp_problem_set_grid(cp_problem* prob, int n1, int n2, const int* node_of);
/* The same for a lattice periodic in direction 1: lattice point (n1 - 1, i2)
 * neighbours (0, i2), as the rays of genRing and genRectWithHole close
 * around the ring or the hole (Mesh.cu:112-245, 249-329).  New. */
int cp_problem_set_grid_periodic(cp_problem* prob, int n1, int n2, const int* node_of);
/* Replay of the multigrid CG iteration as a HIP graph (default 1); 0 launches
 * its kernels one by one (same kernels, same results). */
int cp_problem_set_graphs(cp_problem* prob, int enable);
/* Number of multigrid levels (0 without a lattice) and their lattice sizes
 * (dims: 2 per level, may be NULL). */
int cp_problem_mg_levels(cp_problem* prob, int* levels, int* dims);
/* Layout of the finest level's FP32 operator in the multigrid iterations:
 * 1 = symmetric lattice layout (five 2x2 blocks per node, no column indices;
 * lattices numbered node by node, not periodic), 0 = the sliced-ELL rows.
 * New. */
int cp_problem_mg_level0_layout(cp_problem* prob, int* sym);
/* Smoother of each multigrid level except the coarsest (kinds: one per
 * level, 0 = 2x2 block Jacobi, 1 = block-tridiagonal segments along lattice
 * direction 1; the coarsest level, solved exactly, reports 0). */
int cp_problem_mg_smoothers(cp_problem* prob, int* kinds);
/* Multigrid set-up policy: the coarse levels are rebuilt at the first secant
 * iteration of every load level and then every `every`-th assembly (1 =
 * every assembly, default 8); in between only the finest level's smoother
 * follows the new secant moduli. */
int cp_problem_set_mg_rebuild(cp_problem* prob, int every);

/* Scale the load (rp, prescribed u) -- new: load stepping. */
int cp_set_load(cp_problem* prob, double factor);
/* Move to the next load level: with predictor >= 1 and a converged level
 * behind, u (and with predictor = 2 also log(psi)) is extrapolated linearly
 * in the load factor before the level's secant iterations start (initial
 * guess only; the converged level does not depend on it). */
int cp_advance_load(cp_problem* prob, double factor, int predictor);
/* Reset u and the secant moduli to the initial state (initPlastParams :385-410). */
int cp_reset_state(cp_problem* prob);

/* One assembly with the current secant moduli: fillGlobalStiffness
 * (PlasticitySolver.cu:1492-1508) without the band matrix. */
int cp_assemble(cp_problem* prob, int precision);
/* One masked CG solve from the current u: ConjGradCuda::solve (ConjGradCuda.cu:185-231). */
int cp_cg(cp_problem* prob, double tol, int maxit, int precision, double inner_tol,
          long long* iters, double* relres);
/* One secant update (updateParamDouble :1613-1625 + exitCondition :1722-1735)
 * followed by the plain update of psi, E_c, nu_c; returns the exit measure. */
int cp_return_map(cp_problem* prob, double* rel_err);
/* calcPlastDeformCuda (:1655-1662). */
int cp_plastic_strains(cp_problem* prob);
/* f = K(E_c, nu_c) u computed matrix-free in FP64 (element internal force). */
int cp_internal_force(cp_problem* prob, double* f_host);
/* The same f with no device allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  cp_internal_force may
 * allocate the problem's mixed-precision arena on first use to stage f;
 * this call stages it in the library's static 8 MiB device array instead,
 * so problems of more than 524,288 nodes are refused with CP_ERR_INVALID.
 * The array is shared by every call on one GPU (its copy is held until the
 * process ends): calls that overlap on one GPU, from several threads, race. */
int cp_internal_force_noalloc(cp_problem* prob, double* f_host);

/* PlasticitySolver::solve / solveElast for one load level (solveCUDA :469-512). */
int cp_solve(cp_problem* prob, const cp_options* opt, cp_stats* stats);
int cp_solve_elastic(cp_problem* prob, const cp_options* opt, cp_stats* stats);
/* nsteps load levels factor*n/nsteps, n = 1..nsteps, each warm-started. */
int cp_solve_steps(cp_problem* prob, int nsteps, double final_factor, const cp_options* opt,
                   cp_stats* per_step);

/* dataToRAM (:1738-1781): one field (CP_F_*) to host memory. */
int cp_download(cp_problem* prob, int field, double* host);
/* Warm-start state from host: u (2*nn) and psi (ne); either may be NULL. */
int cp_upload_state(cp_problem* prob, const double* u, const double* psi);

/* Kernel timings of the last CG solve (HIP events on the problem's stream):
 * average duration of one CG iteration kernel, and of the SpMV-only kernel
 * when cp_bench_spmv was used. */
int cp_last_cg_kernel_ms(cp_problem* prob, double* ms_per_iter);
/* Times `reps` standalone SpMV launches y = A x (precision) and returns the
 * average kernel ms, the stored-format bytes per launch and the algorithmic
 * bytes per launch. */
int cp_bench_spmv(cp_problem* prob, int precision, int reps, double* ms, double* bytes_stored,
                  double* bytes_algo);
/* Times `reps` launches of the fused CG iteration kernel alone (precision;
 * precond 1: block-Jacobi variant) on the matrix of the current secant
 * moduli (assembled here), HIP events on the problem's stream; returns the
 * average kernel ms and the bytes one launch moves in this library's
 * storage format. */
int cp_bench_cg(cp_problem* prob, int precision, int precond, int reps, double* ms, double* bytes_format);
/* Times `reps` assemblies (element stiffness + deterministic assembly into
 * the FP64 or FP32 matrix, no preconditioner set-up) at the current secant
 * moduli; returns the average ms of one assembly. */
int cp_bench_assemble(cp_problem* prob, int precision, int reps, double* ms);
/* Multigrid timing (precond 2, lattice set): assembles and sets up the
 * hierarchy, then times `reps` launches of the finest level's Chebyshev
 * smoothing step and `reps` V-cycles (HIP events on the problem's stream);
 * returns their average ms and the bytes one smoothing step moves in this
 * library's storage format. */
int cp_bench_mg(cp_problem* prob, int precision, int reps, double* ms_vcycle, double* ms_smooth,
                double* bytes_smooth);
/* The same set-up, then `reps` launches of the multigrid CG's operator
 * product alone (w = mask (A u) with the iteration's three dot products, on
 * level 0's storage: the symmetric lattice layout in FP32 where it is
 * built); returns their average ms and the bytes one launch must move in
 * that layout.  Replaces the timing of cg1sparse (ConjGradCuda.cu:39-60),
 * the reference's SpMV. */
int cp_bench_mg_spmv(cp_problem* prob, int precision, int reps, double* ms, double* bytes);
/* Achievable HBM bandwidth: average ms of `reps` launches of a plain device
 * copy of `bytes` (a multiple of 16; read once and written once per launch)
 * between two buffers this call allocates and frees, and (ms_read, may be
 * NULL) of a read-only pass over the same `bytes`.  New: a yardstick for the
 * bandwidth-bound kernels, no counterpart in the reference. */
int cp_bench_copy(cp_ctx* ctx, size_t bytes, int reps, double* ms, double* ms_read);
/* Matrix storage info: nodes, stored blocks (with padding), true blocks. */
int cp_problem_layout(cp_problem* prob, long long* nblk_stored, long long* nblk, int* max_width);

/* ------------------------------------------------------------------ */
/* serving: continuous batching of independent problems (new)          */
/* ------------------------------------------------------------------ */
/* Call the cp_server_* functions of one server from one thread.  The server
 * owns a worker thread of its own that prepares admitted requests (pattern,
 * upload, multigrid lattice and graphs); cp_server_destroy stops it.
 * Multigrid requests advance asynchronously, each on a stream of its own;
 * block-Jacobi requests advance in batched rounds on the server's stream.
 * A request's results do not depend on what else is served. */
typedef struct {
    size_t hbm_budget;    /* bytes of device memory the server may hold for problems */
    int    max_slots;     /* problems resident at once */
    int    batch_cg;      /* 1: batched multi-problem CG kernels; 0: one launch per problem */
    int    batch_mg;      /* 1: the multigrid requests in their CG phase advance in lockstep, each
                             step of the iteration one launch over all of them; 0: each on its own */
    int    batch_mg_max_nodes;   /* > 0: only problems of at most this many nodes join the lockstep
                                    batches (larger ones replay their own graphs); 0: no limit */
    int    reserved[3];
} cp_server_config;

typedef struct {
    long long id;
    int       status;          /* CP_OK / CP_ERR_* */
    int       steps_done;
    long long cg_iters;
    int       outer_iters;
    double    t_submit, t_admit, t_done;   /* seconds on the server clock */
} cp_request_info;

int cp_server_create(cp_ctx* ctx, const cp_server_config* cfg, cp_server** out);
/* Waits for the device work of every request, then frees everything. */
int cp_server_destroy(cp_server* srv);
/* Submit one problem: mesh + loads + material + nsteps load levels to
 * final_factor with options (precond 0 or 1); inputs are copied. */
int cp_server_submit(cp_server* srv, int nn, int ne, const double* x, const double* y, const int* conn,
                     const double* rp, const double* u0, const unsigned char* mask,
                     const cp_material* mat, int nsteps, double final_factor, const cp_options* opt,
                     long long* id);
/* The same with the mesh's lattice for multigrid CG (opt->precond = 2; n1 x
 * n2 nodes, node_of as in cp_problem_set_grid, may be NULL); precond 0 and 1
 * ignore the lattice.  New: the reference has no serving loop. */
int cp_server_submit_grid(cp_server* srv, int nn, int ne, const double* x, const double* y, const int* conn,
                          const double* rp, const double* u0, const unsigned char* mask, const cp_material* mat,
                          int nsteps, double final_factor, const cp_options* opt, int n1, int n2, const int* node_of,
                          long long* id);
/* The same with a lattice periodic in direction 1, as
 * cp_problem_set_grid_periodic takes it (rings, plates with a hole).  New. */
int cp_server_submit_grid_periodic(cp_server* srv, int nn, int ne, const double* x, const double* y, const int* conn,
                                   const double* rp, const double* u0, const unsigned char* mask,
                                   const cp_material* mat, int nsteps, double final_factor, const cp_options* opt,
                                   int n1, int n2, const int* node_of, long long* id);
/* Advance the server without waiting for multigrid work: start the
 * requests the worker has prepared, admit queued requests that fit (FIFO,
 * HBM budget, slots), advance every multigrid request whose last readback
 * has landed, and run one batched round of the block-Jacobi requests
 * (which waits for that round).  *in_flight = queued + preparing + active
 * requests.  Memory of finished problems returns to the budget in batches
 * (when admission needs it, every max_slots problems, or when idle). */
int cp_server_step(cp_server* srv, int* in_flight);
/* Completed request: info and (optional) results; returns CP_ERR_STATE if
 * not done, or the request's failure code, with cp_last_error() giving the
 * message of the failure that ended it. */
int cp_server_result(cp_server* srv, long long id, cp_request_info* info, double* u, double* sig_i);
int cp_server_clock(cp_server* srv, double* now);
/* Lockstep multigrid batches run so far (batch_mg): batches, their members
 * summed over batches, and the CG iterations they ran in lockstep.  New. */
int cp_server_batch_stats(cp_server* srv, long long* rounds, long long* members, long long* iters);

#ifdef __cplusplus
}
#endif
#endif
