/*
 * cp_oracle.h -- CPU FP64 restatement of the reference's 2-D elasto-plastic
 * FEM hot path (SATOH-SD/CudaPlasticity), used ONLY as test infrastructure:
 * the parity checker in tests/, the checker inside __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  Nothing in the product library
 * links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference's src/ directory).  Deliberate deviations from the
 * reference (all documented in the repository's design notes):
 *   - masked CG returns x_k (the reference GPU path leaves x_{k-1} in the
 *     caller's buffer when k is odd, ConjGradCuda.cu:203-227); the quirk is
 *     available as an option (OR_CG_QUIRK) so the reference's own fixtures
 *     can be reproduced;
 *   - assembly is sequential in element order (the reference uses FP64
 *     atomics / omp atomics, PlasticitySolver.cu:711-746,1427-1449);
 *   - no data races in the CSR pattern build (CudaSparseSLAE.cuh:316-324);
 *   - CG and the plastic loop have iteration caps and a status code;
 *   - a zero intensity stress (sigma_tab == 0) keeps psi = 1 instead of
 *     producing 0/0 (PlasticitySolver.cu:1607);
 *   - an unloaded body (sigma_i = sigma_tab = 0 everywhere) has exit
 *     measure 0 instead of 0/0; the reference's loop leaves there as well,
 *     since NaN > 1e-5 is false (PlasticitySolver.cu:503);
 *   - renumerateRing with an odd border count assigns the middle column
 *     correctly (the reference's Mesh.cu:347-349 writes the map reversed).
 *
 * Parity pins for each function are listed in tests/test_oracle_*.py.
 */
#ifndef CP_ORACLE_H
#define CP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- material law (Material.cu:3-44) ---------------- */
enum { OR_LAW_LINEAR = 0, OR_LAW_POWER = 1, OR_LAW_SWIFT = 2 };
enum { OR_PLANE_STRESS = 0, OR_PLANE_STRAIN = 1 };

typedef struct {
    double E, nu, h, sigmaT;
    int law;          /* OR_LAW_* */
    int plane;        /* OR_PLANE_* */
    double K_T;       /* linear: post-yield slope of sigma(eps) (Material.cu:14-17) */
    double A, m;      /* power:  Material.cu:19-23 */
    double n;         /* swift:  sigma = sigmaT (1 + eps_p/eps0)^n, eps0 = sigmaT/E */
} or_material;

double or_material_f(const or_material* mat, double eps);

/* ---------------- mesh generators (Mesh.cu) ---------------- */
/* Sizes of the reference's generators, order 1. */
void or_rect_sizes(int N1, int N2, int* nn, int* ne, int* nb, int* blen);
void or_gen_rect(double x1, double x2, double y1, double y2, int N1, int N2,
                 double* x, double* y, int* conn, int* bord, int* blen);

int  or_hole_radN(double x1, double x2, double y1, double y2, double r, int N1, int N2);
void or_hole_sizes(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                   int* nn, int* ne, int* nb, int* blen);
void or_gen_rect_with_hole(double x1, double x2, double y1, double y2, double r, int N1, int N2,
                           double* x, double* y, int* conn, int* bord, int* blen);
/* the same with the radial segment count given (sizes: borderN (radN + 1) nodes) */
void or_gen_rect_with_hole_r(double x1, double x2, double y1, double y2, double r, int N1, int N2, int radN,
                             double* x, double* y, int* conn, int* bord, int* blen);

void or_ring_sizes(int Nphi, int Nr, int* nn, int* ne, int* nb, int* blen);
void or_gen_ring(double a, double b, int Nphi, int Nr,
                 double* x, double* y, int* conn, int* bord, int* blen);

/* renumerateRing (Mesh.cu:332-384), applied in place. */
void or_renumerate_ring(int nn, int ne, double* x, double* y, int* conn,
                        int nb, int* bord, const int* blen, int borderN);

/* Quarter plate [0,L]^2 minus a disc of radius r at the origin (new; the
 * reference's ray construction Mesh.cu:169-190 restricted to 0..pi/2).
 * borders: 0 y=0 edge, 1 x=L edge, 2 y=L edge, 3 x=0 edge, 4 hole arc. */
void or_quarter_plate_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen);
void or_gen_quarter_plate(double L, double r, int Nt, int Nr,
                          double* x, double* y, int* conn, int* bord, int* blen);

/* Quarter ring a<=rho<=b, radial element sizes growing geometrically by
 * `ratio` from the inner radius (new).  borders: 0 inner arc, 1 theta=pi/2
 * edge, 2 outer arc, 3 theta=0 edge. */
void or_quarter_ring_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen);
void or_gen_quarter_ring(double a, double b, int Nt, int Nr, double ratio,
                         double* x, double* y, int* conn, int* bord, int* blen);

/* renumByDirection + renumSort (Mesh.cu:505-630), in place; npe = 4 or 8. */
void or_renum_by_direction(int nn, double* x, double* y, int ne, int npe, int* conn, int nbord, int* bord,
                           double dx, double dy);
int    or_max_index_diff(int ne, const int* conn);             /* Mesh.cu:417-441 */
double or_elem_space4(const double* x, const double* y, const int* conn, int e); /* Mesh.cuh:258-267 */
double or_elem_space8(const double* x, const double* y, const int* conn, int e); /* Mesh.cuh:269-278 */
/* remapOrder (Mesh.cu:443-503): order-1 mesh of doubled resolution -> 8-node
 * mesh; returns the new node count (nn - ne4/4). */
int or_remap_order(int nn, int ne4, const double* x, const double* y, const int* conn4, int width,
                   int nbord, int* bord, double* x8, double* y8, int* conn8, unsigned char* sec);

/* ---------------- CSR pattern (CudaSparseSLAE.cuh:303-458) ---------------- */
long long or_csr_nnz(int nn, int ne, const int* conn);
void      or_csr_pattern(int nn, int ne, const int* conn, int* rows, int* cols);
/* npe nodes per element (4 or 8). */
long long or_csr_nnz_n(int nn, int ne, int npe, const int* conn);
void      or_csr_pattern_n(int nn, int ne, int npe, const int* conn, int* rows, int* cols);

/* ---------------- element math (PlasticitySolver.cu) ---------------- */
void or_calc_b4(const double* x, const double* y, const int* conn, int e,
                double B[4][2][4], double detJ[4]);                     /* :25-64 */
void or_form_c(int plane, double E, double nu, double C[9]);          /* :226-242 */
void or_form_ke4(const double B[4][2][4], const double detJ[4], const double C[9],
                 double h, double Ke[36]);                              /* :156-186 */
void or_calc_b8(const double* x, const double* y, const int* conn, int e,
                double B[9][2][8], double detJ[9]);                     /* :66-134 */
void or_form_ke8(const double B[9][2][8], const double detJ[9], const double C[9],
                 double h, double Ke[136]);                             /* :188-224 */
/* Assemble K(E_c, nu_c) into CSR values (fixed element order). */
void or_assemble(int nn, int ne, const double* x, const double* y, const int* conn,
                 const int* rows, const int* cols, const double* Ec, const double* nuc,
                 int plane, double h, double* vals);
void or_assemble_n(int nn, int ne, int npe, const double* x, const double* y, const int* conn,
                   const int* rows, const int* cols, const double* Ec, const double* nuc,
                   int plane, double h, double* vals);

/* ---------------- loads / boundary conditions (PlasticitySolver.cu:812-1005) ---------------- */
/* Border traction: force values given per border node (2 per node). */
void or_rhs_border_force(const double* x, const double* y, const int* border, int len,
                         const double* fval, int normOrient, double h, double* rp);
/* The same with order-2 (3-node) border segments where sec[border[i]] marks
 * a mid-side node (:1025-1069); sec may be NULL. */
void or_rhs_border_force_q(const double* x, const double* y, const int* border, int len,
                           const unsigned char* sec, const double* fval, int normOrient, double h,
                           double* rp);
void or_rhs_body_force(int ne, const double* x, const double* y, const int* conn,
                       double gx, double gy, double rho, double h, double* rp);
void or_rhs_body_force8(int ne, const double* x, const double* y, const int* conn,
                        double gx, double gy, double rho, double h, double* rp);
void or_fix_axis(int nn, const double* x, const double* y, int vertical, double coord,
                 unsigned char* mask);
void or_displacement_border(const int* border, int len, double ux, double uy,
                            int fixMidX, int fixMidY, unsigned char* mask, double* u);

/* ---------------- masked CG (ConjGradSolver.h:52-101, ConjGradCuda.cu:185-231) ---------------- */
enum { OR_CG_OK = 0, OR_CG_MAXIT = 1, OR_CG_NAN = 2 };
int or_cg(int N, const int* rows, const int* cols, const double* vals, const double* b,
          const unsigned char* mask, double* x, double tol, int maxit, int quirk, int* iters);

/* ---------------- full solve (solveCPU :325-366 / solveCUDA :469-512) ---------------- */
enum {
    OR_F_EXX = 0, OR_F_EYY, OR_F_GXY, OR_F_EPSI, OR_F_SXX, OR_F_SYY, OR_F_TAU,
    OR_F_SIGI, OR_F_SIGTAB, OR_F_PSI, OR_F_EC, OR_F_NUC,
    OR_F_EXXP, OR_F_EYYP, OR_F_GXYP, OR_F_EPSIP, OR_NFIELDS
};

typedef struct {
    double cg_tol;      /* ||r||/||b|| stop (reference: 1e-7) */
    int    cg_maxit;
    double tol;         /* plastic exit tolerance (reference: 1e-5) */
    int    max_outer;   /* reference CPU cap: 50 */
    int    fixed_outer; /* >0: run exactly this many outer iterations */
    int    aitken;      /* 1: Aitken (Irons-Tuck) relaxation of psi */
    int    cg_quirk;    /* 1: reference GPU x_{k-1} behaviour */
    int    elastic;     /* 1: solveElast (one assembly, one CG, stresses) */
    int    warm;        /* 1: start from given u / psi state; 0: reset (initPlastParams) */
    double cg_forcing;  /* > 0: secant iteration k solves to max(cg_tol, cg_forcing rel_err_{k-1}),
                           rel_err_{-1} = 1; exit only after an iteration solved at cg_tol */
} or_options;

typedef struct {
    int outer_iters;
    long long cg_iters;
    double rel_err;
    int status;         /* 0 converged, 1 cap reached, 2 NaN */
} or_stats;

/* state: OR_NFIELDS*ne doubles, field f of element e at state[f*ne+e].
 * u: 2*nn doubles, in: prescribed/initial values, out: solution. */
int or_solve(int nn, int ne, const double* x, const double* y, const int* conn,
             const double* rp, const unsigned char* mask, double* u,
             const or_material* mat, const or_options* opt, double* state, or_stats* st);

/* The same for npe = 4 or 8 nodes per element (strain4 / strain8). */
int or_solve_n(int nn, int ne, int npe, const double* x, const double* y, const int* conn,
               const double* rp, const unsigned char* mask, double* u,
               const or_material* mat, const or_options* opt, double* state, or_stats* st);

/* One secant iteration's update from a given u, as inside or_solve_n:
 * strain4 / strain8 + plastStress (:1517-1610) of every element from the
 * state's E_c, nu_c, then the plain update of psi, E_c, nu_c
 * (updateParamDouble :1613-1625).  Returns the exit measure (exitCondition
 * :1722-1735).  Lets a test pair the oracle's assembly and secant update
 * with a direct sparse solve where its CG would take too long. */
double or_return_map_n(int nn, int ne, int npe, const double* x, const double* y, const int* conn,
                       const double* u, const or_material* mat, double* state);

/* Secant update for one element given its strains (updateParameters :1320-1335). */
void or_secant_update(const or_material* mat, double Ec, double nuc,
                      double exx, double eyy, double gxy, double out[8]);

/* Strain energy (PlasticitySolver.cuh:415-445). */
double or_deform_energy(int ne, const double* x, const double* y, const int* conn,
                        const or_material* mat, const double* state);
double or_deform_energy_n(int ne, int npe, const double* x, const double* y, const int* conn,
                          const or_material* mat, const double* state);

/* y = A x over CSR (diagnostic). */
void or_spmv(int N, const int* rows, const int* cols, const double* vals, const double* x, double* y);

#ifdef __cplusplus
}
#endif
#endif
