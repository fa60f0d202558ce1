// host_check.cpp -- drives every host-side entry point of the library's
// setup code (cp_mesh.cpp: generators, renumberings, CSR pattern, load and
// boundary-condition rules, the .inp reader) on small meshes, valid and
// invalid arguments alike, for a build under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py).  Prints a checksum
// of the outputs; exits non-zero if a call that must succeed fails.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cp_abi.h"

namespace {

double sum = 0.;
int failures = 0;

void expect(int rc, int want, const char* what)
{
    if (rc != want) {
        std::fprintf(stderr, "%s: status %d, expected %d (%s)\n", what, rc, want, cp_last_error());
        ++failures;
    }
}

template <typename T>
void add(const std::vector<T>& v)
{
    for (size_t i = 0; i < v.size(); ++i) sum += double(v[i]) * double(1 + (i % 7));
}

struct M {
    int nn = 0, ne = 0, nb = 0, npe = 4;
    std::vector<double> x, y;
    std::vector<int> conn, bord, blen = std::vector<int>(8);
    std::vector<unsigned char> sec;
};

template <typename Sizes, typename Gen>
M gen(Sizes sizes, Gen g, const char* what)
{
    M m;
    expect(sizes(&m.nn, &m.ne, &m.nb, m.blen.data()), CP_OK, what);
    int total = 0;
    for (int b = 0; b < m.nb; ++b) total += m.blen[b];
    m.x.resize(m.nn);
    m.y.resize(m.nn);
    m.conn.resize(4 * (size_t)m.ne);
    m.bord.resize(total);
    expect(g(m), CP_OK, what);
    add(m.x);
    add(m.y);
    add(m.conn);
    add(m.bord);
    return m;
}

void check_mesh(M& m, const char* what)
{
    long long nnz = 0;
    expect(cp_csr_nnz_n(m.nn, m.ne, m.npe, m.conn.data(), &nnz), CP_OK, what);
    std::vector<int> rows(2 * (size_t)m.nn + 1), cols(nnz);
    expect(cp_csr_pattern_n(m.nn, m.ne, m.npe, m.conn.data(), rows.data(), cols.data()), CP_OK, what);
    add(rows);
    add(cols);
    int diff = 0;
    if (m.npe == 4) {
        expect(cp_mesh_max_index_diff(m.ne, m.conn.data(), &diff), CP_OK, what);
        sum += diff;
    }
    std::vector<double> rp(2 * (size_t)m.nn, 0.), u(2 * (size_t)m.nn, 0.);
    std::vector<unsigned char> mask(2 * (size_t)m.nn, 1);
    int off = 0;
    for (int b = 0; b < m.nb; ++b) {
        std::vector<double> f(2 * (size_t)m.blen[b]);
        for (size_t i = 0; i < f.size(); ++i) f[i] = 1.0 + 0.25 * double(i % 5);
        if (m.npe == 8)
            expect(cp_rhs_border_force_q(m.x.data(), m.y.data(), m.bord.data() + off, m.blen[b], m.sec.data(), f.data(),
                                         b & 1, 0.5, rp.data()),
                   CP_OK, what);
        else
            expect(cp_rhs_border_force(m.x.data(), m.y.data(), m.bord.data() + off, m.blen[b], f.data(), b & 1, 0.5,
                                       rp.data()),
                   CP_OK, what);
        off += m.blen[b];
    }
    std::vector<double> rnode(2 * (size_t)m.nn, 0.3);
    if (m.npe == 4) {
        expect(cp_rhs_body_force(m.ne, m.x.data(), m.y.data(), m.conn.data(), 0., -9.81, 2., 1., rp.data()), CP_OK, what);
        expect(cp_rhs_body_force_nodal(m.ne, m.x.data(), m.y.data(), m.conn.data(), rnode.data(), 2., 1., rp.data()),
               CP_OK, what);
    } else {
        expect(cp_rhs_body_force_nodal8(m.ne, m.x.data(), m.y.data(), m.conn.data(), rnode.data(), 2., 1., rp.data()),
               CP_OK, what);
    }
    expect(cp_fix_axis(m.nn, m.x.data(), m.y.data(), 1, 0.0, mask.data()), CP_OK, what);
    expect(cp_fix_axis(m.nn, m.x.data(), m.y.data(), 0, 0.0, mask.data()), CP_OK, what);
    if (m.nb > 0) {
        expect(cp_displacement_border(m.bord.data(), m.blen[0], 1e-3, 0., 1, 0, mask.data(), u.data()), CP_OK, what);
        std::vector<double> uv(2 * (size_t)m.blen[0], 2e-3);
        expect(cp_displacement_border_values(m.bord.data(), m.blen[0], uv.data(), 0, 1, mask.data(), u.data()), CP_OK,
               what);
    }
    add(rp);
    add(u);
    add(mask);
    // renumberings on a copy
    M r = m;
    expect(cp_mesh_renum_by_direction(r.nn, r.x.data(), r.y.data(), r.ne, r.npe, r.conn.data(), (int)r.bord.size(),
                                      r.bord.data(), 1., 1., nullptr),
           CP_OK, what);
    add(r.conn);
    add(r.x);
}

}  // namespace

int main()
{
    std::vector<M> meshes;
    meshes.push_back(gen([](int* a, int* b, int* c, int* d) { return cp_mesh_rect_sizes(7, 5, a, b, c, d); },
                         [](M& m) {
                             return cp_mesh_rect(0., 1., 0., 2., 7, 5, m.x.data(), m.y.data(), m.conn.data(),
                                                 m.bord.data(), m.blen.data());
                         },
                         "rect"));
    meshes.push_back(gen(
        [](int* a, int* b, int* c, int* d) { return cp_mesh_hole_sizes(-3., 3., -2., 2., 0.5, 12, 8, a, b, c, d); },
        [](M& m) {
            return cp_mesh_hole(-3., 3., -2., 2., 0.5, 12, 8, m.x.data(), m.y.data(), m.conn.data(), m.bord.data(),
                                m.blen.data());
        },
        "hole"));
    meshes.push_back(gen([](int* a, int* b, int* c, int* d) { return cp_mesh_ring_sizes(24, 5, a, b, c, d); },
                         [](M& m) {
                             return cp_mesh_ring(0.5, 1., 24, 5, m.x.data(), m.y.data(), m.conn.data(), m.bord.data(),
                                                 m.blen.data());
                         },
                         "ring"));
    meshes.push_back(gen([](int* a, int* b, int* c, int* d) { return cp_mesh_quarter_plate_sizes(16, 8, a, b, c, d); },
                         [](M& m) {
                             return cp_mesh_quarter_plate(1., 0.1, 16, 8, m.x.data(), m.y.data(), m.conn.data(),
                                                          m.bord.data(), m.blen.data());
                         },
                         "quarter plate"));
    meshes.push_back(gen([](int* a, int* b, int* c, int* d) { return cp_mesh_quarter_ring_sizes(12, 9, a, b, c, d); },
                         [](M& m) {
                             return cp_mesh_quarter_ring(0.5, 1., 12, 9, 1.05, m.x.data(), m.y.data(), m.conn.data(),
                                                         m.bord.data(), m.blen.data());
                         },
                         "quarter ring"));
    // order 2: the 8-node generators
    {
        M m;
        m.npe = 8;
        expect(cp_mesh_hole2_sizes(-3., 3., -2., 2., 0.5, 6, 4, &m.nn, &m.ne, &m.nb, m.blen.data()), CP_OK, "hole2");
        int total = 0;
        for (int b = 0; b < m.nb; ++b) total += m.blen[b];
        m.x.resize(m.nn);
        m.y.resize(m.nn);
        m.conn.resize(8 * (size_t)m.ne);
        m.sec.resize(m.nn);
        m.bord.resize(total);
        expect(cp_mesh_hole2(-3., 3., -2., 2., 0.5, 6, 4, m.x.data(), m.y.data(), m.conn.data(), m.sec.data(),
                             m.bord.data(), m.blen.data()),
               CP_OK, "hole2");
        add(m.conn);
        add(m.sec);
        meshes.push_back(m);
    }
    for (M& m : meshes) check_mesh(m, "mesh");
    // renumerateRing on the rectangle (8 x 6 nodes)
    {
        M r = meshes[0];
        expect(cp_mesh_renumerate_ring(r.nn, r.ne, r.x.data(), r.y.data(), r.conn.data(), r.nb, r.bord.data(),
                                       r.blen.data(), 8, nullptr),
               CP_OK, "renumerate ring");
        add(r.conn);
        expect(cp_mesh_renumerate_ring(r.nn, r.ne, r.x.data(), r.y.data(), r.conn.data(), r.nb, r.bord.data(),
                                       r.blen.data(), 7, nullptr),
               CP_ERR_INVALID, "renumerate ring, bad count");
    }
    // .inp round trip and the reader's errors
    {
        const M& m = meshes[1];
        const std::string path = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") + "/host_check.inp";
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return 2;
        std::fprintf(f, "*NODE\n");
        for (int n = 0; n < m.nn; ++n) std::fprintf(f, "%d, %.17g, %.17g\n", n + 1, m.x[n], m.y[n]);
        std::fprintf(f, "*ELEMENT, TYPE=S4, ELSET=A\n");
        for (int e = 0; e < m.ne; ++e)
            std::fprintf(f, "%d, %d, %d, %d, %d\n", e + 1, m.conn[4 * e] + 1, m.conn[4 * e + 1] + 1,
                         m.conn[4 * e + 2] + 1, m.conn[4 * e + 3] + 1);
        std::fclose(f);
        int nn = 0, n3 = 0, n4 = 0, n8 = 0, nb = 0;
        std::vector<int> blen(4);
        expect(cp_mesh_inp_sizes(path.c_str(), &nn, &n3, &n4, &n8, &nb, blen.data(), 4), CP_OK, "inp sizes");
        std::vector<double> x(nn), y(nn);
        std::vector<int> c4(4 * (size_t)n4), bord(1);
        std::vector<unsigned char> sec(nn);
        expect(cp_mesh_load_inp(path.c_str(), x.data(), y.data(), nullptr, c4.data(), nullptr, sec.data(), bord.data()),
               CP_OK, "inp load");
        add(x);
        add(c4);
        std::remove(path.c_str());
        expect(cp_mesh_inp_sizes("missing.inp", &nn, &n3, &n4, &n8, &nb, blen.data(), 4), CP_ERR_INVALID, "inp missing");
        expect(cp_mesh_inp_sizes("mesh.vtk", &nn, &n3, &n4, &n8, &nb, blen.data(), 4), CP_ERR_INVALID, "inp format");
    }
    // refusals
    {
        M r = meshes[3];
        expect(cp_mesh_renum_by_direction(r.nn, r.x.data(), r.y.data(), r.ne, 4, r.conn.data(), 0, nullptr, NAN, 1., nullptr),
               CP_ERR_INVALID, "renum NaN direction");
        r.conn[5] = r.nn + 3;
        long long nnz = 0;
        expect(cp_csr_nnz_n(r.nn, r.ne, 4, r.conn.data(), &nnz), CP_ERR_INVALID, "connectivity out of range");
        int nn = 0, ne = 0, nb = 0;
        std::vector<int> bl(8);
        expect(cp_mesh_rect_sizes(0, 3, &nn, &ne, &nb, bl.data()), CP_ERR_INVALID, "rect size 0");
    }
    std::printf("host_check: %zu meshes, checksum %.17g, %d failures\n", meshes.size(), sum, failures);
    return failures == 0 ? 0 : 1;
}
