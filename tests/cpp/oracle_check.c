/* oracle_check.c -- drives the CPU oracle (oracle/cp_oracle.c) through its
 * generators, pattern, assembly, loads, masked CG and secant loops (4- and
 * 8-node elements, plain and Aitken, load steps through warm starts) on
 * small meshes, for builds under AddressSanitizer + UndefinedBehaviorSanitizer
 * and under ThreadSanitizer with the oracle's OpenMP loops on several
 * threads (tests/test_sanitizers.py).  Prints a checksum; exits non-zero if
 * a solve that must converge does not. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cp_oracle.h"

static double checksum = 0.;

static void add(const double* v, size_t n)
{
    for (size_t i = 0; i < n; ++i) checksum += v[i] * (double)(1 + i % 5);
}

typedef struct {
    int nn, ne, nb, npe;
    int blen[8];
    double *x, *y;
    int *conn, *bord;
    unsigned char* sec;
} mesh;

static void free_mesh(mesh* m)
{
    free(m->x); free(m->y); free(m->conn); free(m->bord); free(m->sec);
}

static mesh quarter_plate(int Nt, int Nr)
{
    mesh m;
    memset(&m, 0, sizeof m);
    m.npe = 4;
    or_quarter_plate_sizes(Nt, Nr, &m.nn, &m.ne, &m.nb, m.blen);
    int tot = 0;
    for (int b = 0; b < m.nb; ++b) tot += m.blen[b];
    m.x = malloc(sizeof(double) * m.nn);
    m.y = malloc(sizeof(double) * m.nn);
    m.conn = malloc(sizeof(int) * 4 * m.ne);
    m.bord = malloc(sizeof(int) * tot);
    or_gen_quarter_plate(1.0, 0.1, Nt, Nr, m.x, m.y, m.conn, m.bord, m.blen);
    return m;
}

/* traction S on border 1, symmetry on x = 0 and y = 0 */
static void loads(const mesh* m, int border, double S, double* rp, unsigned char* mask, double* u)
{
    int off = 0;
    for (int b = 0; b < border; ++b) off += m->blen[b];
    int len = m->blen[border];
    double* f = malloc(sizeof(double) * 2 * len);
    for (int i = 0; i < len; ++i) { f[2 * i] = S; f[2 * i + 1] = 0.; }
    memset(rp, 0, sizeof(double) * 2 * m->nn);
    memset(u, 0, sizeof(double) * 2 * m->nn);
    memset(mask, 1, 2 * m->nn);
    if (m->npe == 8) or_rhs_border_force_q(m->x, m->y, m->bord + off, len, m->sec, f, 0, 1.0, rp);
    else or_rhs_border_force(m->x, m->y, m->bord + off, len, f, 0, 1.0, rp);
    or_fix_axis(m->nn, m->x, m->y, 1, 0.0, mask);
    or_fix_axis(m->nn, m->x, m->y, 0, 0.0, mask);
    free(f);
}

static int solve_case(const mesh* m, const or_material* mat, or_options opt, int steps, double S)
{
    double* rp = malloc(sizeof(double) * 2 * m->nn);
    double* rps = malloc(sizeof(double) * 2 * m->nn);
    double* u = malloc(sizeof(double) * 2 * m->nn);
    unsigned char* mask = malloc(2 * m->nn);
    double* st = calloc((size_t)OR_NFIELDS * m->ne, sizeof(double));
    loads(m, 1, S, rp, mask, u);
    int bad = 0;
    for (int n = 1; n <= steps; ++n) {
        for (int i = 0; i < 2 * m->nn; ++i) rps[i] = rp[i] * n / steps;
        opt.warm = n > 1;
        or_stats s;
        int rc = or_solve_n(m->nn, m->ne, m->npe, m->x, m->y, m->conn, rps, mask, u, mat, &opt, st, &s);
        if (rc != 0 || s.status != 0) bad = 1;
        checksum += s.rel_err + s.outer_iters;
    }
    add(u, 2 * m->nn);
    add(st, (size_t)OR_NFIELDS * m->ne);
    checksum += or_deform_energy_n(m->ne, m->npe, m->x, m->y, m->conn, mat, st);
    double r = or_return_map_n(m->nn, m->ne, m->npe, m->x, m->y, m->conn, u, mat, st);
    checksum += r;
    free(rp); free(rps); free(u); free(mask); free(st);
    return bad;
}

int main(void)
{
    int bad = 0;
    const double E = 2.1e5;
    or_material lin = {E, 0.3, 1.0, 240.0, OR_LAW_LINEAR, OR_PLANE_STRESS, E / 101, 0., 0., 0.};
    or_material swift = {E, 0.3, 1.0, 240.0, OR_LAW_SWIFT, OR_PLANE_STRESS, 0., 0., 0., 0.2};
    or_material pow304 = {193e3, 0.29, 1.0, 170.0, OR_LAW_POWER, OR_PLANE_STRAIN, 0., 1930., 0.015, 0.};
    or_options base = {1e-12, 1000000, 1e-8, 500, 0, 0, 0, 0, 0, 0.0};

    mesh qp = quarter_plate(24, 12);
    or_options o = base;
    bad |= solve_case(&qp, &lin, o, 3, 300.0);
    o.aitken = 1;
    o.cg_forcing = 0.1;
    bad |= solve_case(&qp, &swift, o, 3, 300.0);
    o = base;
    o.fixed_outer = 4;
    o.cg_quirk = 1;
    bad |= solve_case(&qp, &pow304, o, 1, 170.0);

    /* 8-node quads: genRectWithHole(..., 6, 4, order 2) = the order-1 plate of
     * doubled resolution and radial count, regrouped by remapOrder(radN) */
    {
        const int N1 = 6, N2 = 4;
        const int radN = 2 * or_hole_radN(-3., 3., -2., 2., 0.5, N1, N2);
        const int borderN = 4 * (N1 + N2);
        const int nn = borderN * (radN + 1), ne = borderN * radN;
        int blen[8] = {2 * N1 + 1, 2 * N2 + 1, 2 * N1 + 1, 2 * N2 + 1, borderN + 1, 0, 0, 0};
        int tot = 0;
        for (int b = 0; b < 5; ++b) tot += blen[b];
        double* x = malloc(sizeof(double) * nn);
        double* y = malloc(sizeof(double) * nn);
        int* conn = malloc(sizeof(int) * 4 * ne);
        int* bord = malloc(sizeof(int) * tot);
        or_gen_rect_with_hole_r(-3., 3., -2., 2., 0.5, 2 * N1, 2 * N2, radN, x, y, conn, bord, blen);
        mesh m8;
        memset(&m8, 0, sizeof m8);
        m8.npe = 8;
        m8.nb = 5;
        memcpy(m8.blen, blen, sizeof blen);
        m8.ne = ne / 4;
        m8.x = malloc(sizeof(double) * nn);
        m8.y = malloc(sizeof(double) * nn);
        m8.conn = malloc(sizeof(int) * 8 * m8.ne);
        m8.sec = calloc(nn, 1);
        m8.bord = bord;
        m8.nn = or_remap_order(nn, ne, x, y, conn, radN, tot, bord, m8.x, m8.y, m8.conn, m8.sec);
        long long nnz = or_csr_nnz_n(m8.nn, m8.ne, 8, m8.conn);
        int* rows = malloc(sizeof(int) * (2 * m8.nn + 1));
        int* cols = malloc(sizeof(int) * nnz);
        or_csr_pattern_n(m8.nn, m8.ne, 8, m8.conn, rows, cols);
        double* Ec = malloc(sizeof(double) * m8.ne);
        double* nuc = malloc(sizeof(double) * m8.ne);
        for (int e = 0; e < m8.ne; ++e) { Ec[e] = E; nuc[e] = 0.3; }
        double* vals = malloc(sizeof(double) * nnz);
        or_assemble_n(m8.nn, m8.ne, 8, m8.x, m8.y, m8.conn, rows, cols, Ec, nuc, OR_PLANE_STRESS, 1.0, vals);
        add(vals, (size_t)nnz);
        bad |= solve_case(&m8, &lin, base, 2, 150.0);
        free(rows); free(cols); free(Ec); free(nuc); free(vals);
        free(x); free(y); free(conn);
        free_mesh(&m8);
    }
    /* renumbering and the ring generators */
    {
        int nn, ne, nb, blen[8];
        or_ring_sizes(24, 6, &nn, &ne, &nb, blen);
        int tot = 0;
        for (int b = 0; b < nb; ++b) tot += blen[b];
        double* x = malloc(sizeof(double) * nn);
        double* y = malloc(sizeof(double) * nn);
        int* conn = malloc(sizeof(int) * 4 * ne);
        int* bord = malloc(sizeof(int) * tot);
        or_gen_ring(0.5, 1.0, 24, 6, x, y, conn, bord, blen);
        or_renum_by_direction(nn, x, y, ne, 4, conn, tot, bord, 1.0, 1.0);
        add(x, nn);
        checksum += or_max_index_diff(ne, conn);
        free(x); free(y); free(conn); free(bord);
    }
    free_mesh(&qp);
    printf("oracle_check: checksum %.17g, %s\n", checksum, bad ? "a solve failed" : "all solves converged");
    return bad;
}
