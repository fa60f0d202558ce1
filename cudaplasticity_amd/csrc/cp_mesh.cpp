// cp_mesh.cpp -- host-side setup of the plasticity hot path: mesh generators,
// load / boundary-condition rules and the reference CSR sparsity pattern.
// sin/cos pairs go through glibc sincos so that every build (gcc or clang)
// produces the same coordinates.  These run once per problem on the host;
// the device path starts at
// cp_problem_create.  Reference citations are relative to its src/.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cp_abi.h"
#include "cp_internal.h"

namespace {

constexpr double kPi = 3.141592653589793;

struct BorderWriter {
    int* bord;
    const int* blen;
    int* at(int b) const {
        int off = 0;
        for (int i = 0; i < b; ++i) off += blen[i];
        return bord + off;
    }
};

// Mesh::renumerateRing (Mesh.cu:332-384): columns of the ring are numbered
// alternately from the front and the back so the band stays narrow.  An odd
// column count numbers the middle column last (the reference writes that
// map reversed, Mesh.cu:347-349).
// The new number of old node i * columns + j (column j, row i).
std::vector<int> ring_permutation(int nn, int columns)
{
    const int rowsPerCol = nn / columns;
    std::vector<int> perm(nn);
    int next = 0;
    int lo = 0, hi = columns - 1;
    for (; lo < hi; ++lo, --hi) {
        for (int i = 0; i < rowsPerCol; ++i) perm[i * columns + lo] = next++;
        for (int i = 0; i < rowsPerCol; ++i) perm[i * columns + hi] = next++;
    }
    if (lo == hi)
        for (int i = 0; i < rowsPerCol; ++i) perm[i * columns + lo] = next++;
    return perm;
}

void renumber_ring(int nn, int ne, double* x, double* y, int* conn, int nb, int* bord, const int* blen,
                   int columns)
{
    const std::vector<int> perm = ring_permutation(nn, columns);
    for (int k = 0; k < 4 * ne; ++k) conn[k] = perm[conn[k]];
    int total = 0;
    for (int b = 0; b < nb; ++b) total += blen[b];
    for (int k = 0; k < total; ++k) bord[k] = perm[bord[k]];
    std::vector<double> nx(nn), ny(nn);
    for (int i = 0; i < nn; ++i) {
        nx[perm[i]] = x[i];
        ny[perm[i]] = y[i];
    }
    std::memcpy(x, nx.data(), sizeof(double) * nn);
    std::memcpy(y, ny.data(), sizeof(double) * nn);
}

int radial_segments(double x1, double x2, double y1, double y2, double r, int N1, int N2)
{
    // Mesh.cu:115-116
    int radN = (N1 + N2) / 4;
    return (int)std::round(double(radN) * (x2 - x1 + y2 - y1 - 4. * r) * 2. / (x2 - x1 + y2 - y1));
}

// Mesh::genRectWithHole (Mesh.cu:112-245), order 1, radial segment count
// given: outer rectangle nodes counter-clockwise from (x1, y1), hole nodes
// on a circle, straight rays between outer node j and hole node j, then
// renumerateRing.  Order 2 doubles N1, N2 and radN (:117-121).
int hole_mesh(double x1, double x2, double y1, double y2, double r, int N1, int N2, int radN, double* x,
              double* y, int* conn, int* bord, int* blen)
{
    const int borderN = 2 * (N1 + N2);
    const int nn = borderN * (radN + 1), ne = borderN * radN, nb = 5;
    blen[0] = blen[2] = N1 + 1;
    blen[1] = blen[3] = N2 + 1;
    blen[4] = borderN + 1;
    BorderWriter bw{bord, blen};
    int *bBottom = bw.at(0), *bRight = bw.at(1), *bTop = bw.at(2), *bLeft = bw.at(3), *bHole = bw.at(4);

    const double hx = (x2 - x1) / N1, hy = (y2 - y1) / N2;
    int id = 0;
    for (int i = 0; i < N1; ++i, ++id) { x[id] = x1 + i * hx; y[id] = y1; bBottom[i] = id; }
    bBottom[N1] = id;
    for (int j = 0; j < N2; ++j, ++id) { x[id] = x2; y[id] = y1 + j * hy; bRight[j] = id; }
    bRight[N2] = id;
    for (int i = 0; i < N1; ++i, ++id) { x[id] = x2 - i * hx; y[id] = y2; bTop[i] = id; }
    bTop[N1] = id;
    for (int j = 0; j < N2; ++j, ++id) { x[id] = x1; y[id] = y2 - j * hy; bLeft[j] = id; }
    bLeft[N2] = 0;

    const double dphi = 2. * kPi / borderN;
    const double cx = 0.5 * (x1 + x2), cy = 0.5 * (y1 + y2);
    const double phi0 = dphi * (N1 + N2 * 1.5);
    const int holeBase = radN * borderN;
    for (int k = 0; k < borderN; ++k) {
        const double a = phi0 + double(k) * dphi;
        double sa, ca;
        ::sincos(a, &sa, &ca);
        x[holeBase + k] = cx + ca * r;
        y[holeBase + k] = cy + sa * r;
        bHole[borderN - 1 - k] = holeBase + k;
    }
    bHole[borderN] = nn - 1;
    for (int j = 0; j < borderN; ++j) {
        const double sx = (x[holeBase + j] - x[j]) / radN, sy = (y[holeBase + j] - y[j]) / radN;
        for (int i = 1; i < radN; ++i) {
            x[i * borderN + j] = x[j] + sx * double(i);
            y[i * borderN + j] = y[j] + sy * double(i);
        }
    }
    for (int j = 0; j < borderN; ++j) {
        const int jn = (j + 1 == borderN) ? 0 : j + 1;
        for (int i = 0; i < radN; ++i) {
            int* c = conn + 4 * (j * radN + i);
            c[0] = (i + 1) * borderN + jn;
            c[1] = (i + 1) * borderN + j;
            c[2] = i * borderN + j;
            c[3] = i * borderN + jn;
        }
    }
    renumber_ring(nn, ne, x, y, conn, nb, bord, blen, borderN);
    return CP_OK;
}

}  // namespace

namespace {
struct InpCounts {
    int nodeLines = 0, n3 = 0, n4 = 0, n8 = 0;
    std::vector<int> blen;
};

void blank_commas(std::string& s)
{
    for (char& c : s)
        if (c == ',') c = ' ';
}

bool body_line(const std::string& l) { return !l.empty() && l[0] != '*'; }

int elem_kind(const std::string& line)
{
    if (line.compare(0, 8, "*ELEMENT") != 0) return 0;
    const size_t t = line.find("TYPE=");
    if (t == std::string::npos) return 0;
    if (line.compare(t + 5, 3, "S8R") == 0) return 8;
    if (line.compare(t + 5, 2, "S4") == 0) return 4;
    if (line.compare(t + 5, 2, "S3") == 0) return 3;
    return 0;
}

// Walks the file once; on every block line calls fn(kind, index-in-block,
// block-number, line): kind 1 node, 3/4/8 element, 9 border.
template <typename Fn>
int walk_inp(const char* path, Fn fn)
{
    const std::string name(path ? path : "");
    if (name.size() < 4 || name.compare(name.size() - 4, 4, ".inp") != 0)
        return cp::fail(CP_ERR_INVALID, "Unknown mesh file format");
    std::ifstream f(name);
    if (!f) return cp::fail(CP_ERR_INVALID, "There is no such file");
    std::string line;
    bool have = false;   // `line` already holds the next line to dispatch
    int nb = 0;
    while (have || std::getline(f, line)) {
        have = false;
        int kind = 0;
        if (line == "*NODE") kind = 1;
        else if (line == "*MPC") kind = 9;
        else kind = elem_kind(line);
        if (!kind) continue;
        int k = 0;
        while (std::getline(f, line)) {
            if (!body_line(line)) {
                have = true;
                break;
            }
            if (int s = fn(kind, k++, nb, line)) return s;
        }
        if (kind == 9) ++nb;
    }
    return CP_OK;
}
}  // namespace

extern "C" {

int cp_mesh_rect_sizes(int N1, int N2, int* nn, int* ne, int* nb, int* blen)
{
    if (N1 < 1 || N2 < 1) return cp::fail(CP_ERR_INVALID, "cp_mesh_rect_sizes: N1, N2 must be >= 1");
    *nn = (N1 + 1) * (N2 + 1);
    *ne = N1 * N2;
    *nb = 4;
    if (blen) {
        blen[0] = blen[2] = N2 + 1;
        blen[1] = blen[3] = N1 + 1;
    }
    return CP_OK;
}

// Mesh::genRectangle (Mesh.cu:42-101), order 1.
int cp_mesh_rect(double x1, double x2, double y1, double y2, int N1, int N2, double* x, double* y, int* conn,
                 int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_rect_sizes(N1, N2, &nn, &ne, &nb, blen)) return s;
    const double hx = (x2 - x1) / N1, hy = (y2 - y1) / N2;
    const int stride = N2 + 1;
    for (int i = 0; i <= N1; ++i)
        for (int j = 0; j <= N2; ++j) {
            x[i * stride + j] = x1 + hx * i;
            y[i * stride + j] = y1 + hy * j;
        }
    for (int i = 0; i < N1; ++i)
        for (int j = 0; j < N2; ++j) {
            int* c = conn + 4 * (i * N2 + j);
            c[0] = (i + 1) * stride + j + 1;
            c[1] = i * stride + j + 1;
            c[2] = i * stride + j;
            c[3] = (i + 1) * stride + j;
        }
    BorderWriter bw{bord, blen};
    int *right = bw.at(0), *top = bw.at(1), *left = bw.at(2), *bottom = bw.at(3);
    for (int j = 0; j <= N2; ++j) {
        right[j] = N1 * stride + j;      // (x2, y1) -> (x2, y2)
        left[j] = N2 - j;                // (x1, y2) -> (x1, y1)
    }
    for (int i = 0; i <= N1; ++i) {
        top[i] = (N1 - i) * stride + N2; // (x2, y2) -> (x1, y2)
        bottom[i] = i * stride;          // (x1, y1) -> (x2, y1)
    }
    return CP_OK;
}

int cp_mesh_hole_sizes(double x1, double x2, double y1, double y2, double r, int N1, int N2, int* nn, int* ne,
                       int* nb, int* blen)
{
    if (N1 < 1 || N2 < 1 || !(r > 0)) return cp::fail(CP_ERR_INVALID, "cp_mesh_hole_sizes: bad arguments");
    const int borderN = 2 * (N1 + N2);
    const int radN = radial_segments(x1, x2, y1, y2, r, N1, N2);
    if (radN < 1) return cp::fail(CP_ERR_INVALID, "cp_mesh_hole_sizes: hole too large for the plate");
    *nn = borderN * (radN + 1);
    *ne = borderN * radN;
    *nb = 5;
    if (blen) {
        blen[0] = blen[2] = N1 + 1;
        blen[1] = blen[3] = N2 + 1;
        blen[4] = borderN + 1;
    }
    return CP_OK;
}

// Mesh::genRectWithHole (Mesh.cu:112-245), order 1.
int cp_mesh_hole(double x1, double x2, double y1, double y2, double r, int N1, int N2, double* x, double* y,
                 int* conn, int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_hole_sizes(x1, x2, y1, y2, r, N1, N2, &nn, &ne, &nb, blen)) return s;
    return hole_mesh(x1, x2, y1, y2, r, N1, N2, radial_segments(x1, x2, y1, y2, r, N1, N2), x, y, conn, bord,
                     blen);
}

int cp_mesh_ring_sizes(int Nphi, int Nr, int* nn, int* ne, int* nb, int* blen)
{
    if (Nphi < 3 || Nr < 1) return cp::fail(CP_ERR_INVALID, "cp_mesh_ring_sizes: bad arguments");
    *nn = (Nr + 1) * Nphi;
    *ne = Nr * Nphi;
    *nb = 2;
    if (blen) blen[0] = blen[1] = Nphi + 1;
    return CP_OK;
}

// Mesh::genRing (Mesh.cu:249-329), order 1.
int cp_mesh_ring(double a, double b, int Nphi, int Nr, double* x, double* y, int* conn, int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_ring_sizes(Nphi, Nr, &nn, &ne, &nb, blen)) return s;
    BorderWriter bw{bord, blen};
    int *inner = bw.at(0), *outer = bw.at(1);
    const double hr = (b - a) / Nr, hphi = 2. * kPi / Nphi;
    for (int j = 0; j < Nphi; ++j) {
        for (int i = 0; i <= Nr; ++i) {
            const double rr = b - hr * i, phi = hphi * j;
            double sp, cp;
            ::sincos(phi, &sp, &cp);
            x[i * Nphi + j] = rr * cp;
            y[i * Nphi + j] = rr * sp;
        }
        outer[Nphi - j] = j;
        inner[j] = Nr * Nphi + j;
    }
    outer[0] = 0;
    inner[Nphi] = Nr * Nphi;
    for (int j = 0; j < Nphi; ++j) {
        const int jn = (j + 1 == Nphi) ? 0 : j + 1;
        for (int i = 0; i < Nr; ++i) {
            int* c = conn + 4 * (j * Nr + i);
            c[0] = (i + 1) * Nphi + jn;
            c[1] = (i + 1) * Nphi + j;
            c[2] = i * Nphi + j;
            c[3] = i * Nphi + jn;
        }
    }
    renumber_ring(nn, ne, x, y, conn, nb, bord, blen, Nphi);
    return CP_OK;
}

int cp_mesh_quarter_plate_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen)
{
    if (Nt < 2 || Nr < 1) return cp::fail(CP_ERR_INVALID, "cp_mesh_quarter_plate_sizes: bad arguments");
    *nn = (Nt + 1) * (Nr + 1);
    *ne = Nt * Nr;
    *nb = 5;
    if (blen) {
        blen[0] = Nr + 1;
        blen[1] = Nt / 2 + 1;
        blen[2] = Nt - Nt / 2 + 1;
        blen[3] = Nr + 1;
        blen[4] = Nt + 1;
    }
    return CP_OK;
}

// Quarter plate with a quarter hole (new).  Ray j joins outer-edge point j
// (x = L for the first Nt/2 segments, then y = L) to hole point j at angle
// (pi/2) j/Nt; radial index i runs from the outer edge (0) to the hole (Nr).
int cp_mesh_quarter_plate(double L, double r, int Nt, int Nr, double* x, double* y, int* conn, int* bord,
                          int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_quarter_plate_sizes(Nt, Nr, &nn, &ne, &nb, blen)) return s;
    if (!(r > 0) || !(r < L)) return cp::fail(CP_ERR_INVALID, "cp_mesh_quarter_plate: need 0 < r < L");
    const int half = Nt / 2, col = Nr + 1;
    for (int j = 0; j <= Nt; ++j) {
        double ox, oy, hx, hy;
        if (j <= half) {
            ox = L;
            oy = L * double(j) / double(half);
        } else {
            ox = L - L * double(j - half) / double(Nt - half);
            oy = L;
        }
        if (j == 0) { hx = r; hy = 0.; }
        else if (j == Nt) { hx = 0.; hy = r; }
        else {
            const double t = 0.5 * kPi * double(j) / double(Nt);
            double st, ct;
            ::sincos(t, &st, &ct);
            hx = r * ct;
            hy = r * st;
        }
        const double sx = (hx - ox) / Nr, sy = (hy - oy) / Nr;
        x[j * col] = ox;
        y[j * col] = oy;
        for (int i = 1; i < Nr; ++i) {
            x[j * col + i] = ox + sx * double(i);
            y[j * col + i] = oy + sy * double(i);
        }
        x[j * col + Nr] = hx;
        y[j * col + Nr] = hy;
    }
    for (int j = 0; j < Nt; ++j)
        for (int i = 0; i < Nr; ++i) {
            int* c = conn + 4 * (j * Nr + i);
            c[0] = (j + 1) * col + i + 1;
            c[1] = j * col + i + 1;
            c[2] = j * col + i;
            c[3] = (j + 1) * col + i;
        }
    BorderWriter bw{bord, blen};
    int *b0 = bw.at(0), *b1 = bw.at(1), *b2 = bw.at(2), *b3 = bw.at(3), *b4 = bw.at(4);
    for (int k = 0; k <= Nr; ++k) { b0[k] = Nr - k; b3[k] = Nt * col + k; }
    for (int k = 0; k <= half; ++k) b1[k] = k * col;
    for (int k = 0; k <= Nt - half; ++k) b2[k] = (half + k) * col;
    for (int k = 0; k <= Nt; ++k) b4[k] = (Nt - k) * col + Nr;
    return CP_OK;
}

int cp_mesh_quarter_ring_sizes(int Nt, int Nr, int* nn, int* ne, int* nb, int* blen)
{
    if (Nt < 1 || Nr < 1) return cp::fail(CP_ERR_INVALID, "cp_mesh_quarter_ring_sizes: bad arguments");
    *nn = (Nt + 1) * (Nr + 1);
    *ne = Nt * Nr;
    *nb = 4;
    if (blen) {
        blen[0] = blen[2] = Nt + 1;
        blen[1] = blen[3] = Nr + 1;
    }
    return CP_OK;
}

// Graded quarter ring (new): radial sizes h0 * ratio^k from the inner radius.
int cp_mesh_quarter_ring(double a, double b, int Nt, int Nr, double ratio, double* x, double* y, int* conn,
                         int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_quarter_ring_sizes(Nt, Nr, &nn, &ne, &nb, blen)) return s;
    if (!(a > 0) || !(b > a) || !(ratio > 0)) return cp::fail(CP_ERR_INVALID, "cp_mesh_quarter_ring: bad radii/ratio");
    std::vector<double> rad(Nr + 1);
    const double h0 = (ratio == 1.0) ? (b - a) / Nr : (b - a) * (ratio - 1.0) / (std::pow(ratio, Nr) - 1.0);
    rad[0] = a;
    double h = h0;
    for (int i = 1; i < Nr; ++i) {
        rad[i] = rad[i - 1] + h;
        h *= ratio;
    }
    rad[Nr] = b;
    const int col = Nr + 1;
    for (int j = 0; j <= Nt; ++j) {
        double c, s;
        if (j == 0) { c = 1.; s = 0.; }
        else if (j == Nt) { c = 0.; s = 1.; }
        else {
            const double t = 0.5 * kPi * double(j) / double(Nt);
            ::sincos(t, &s, &c);
        }
        for (int i = 0; i <= Nr; ++i) {
            x[j * col + i] = rad[i] * c;
            y[j * col + i] = rad[i] * s;
        }
    }
    for (int j = 0; j < Nt; ++j)
        for (int i = 0; i < Nr; ++i) {
            int* q = conn + 4 * (j * Nr + i);
            q[0] = (j + 1) * col + i + 1;
            q[1] = (j + 1) * col + i;
            q[2] = j * col + i;
            q[3] = j * col + i + 1;
        }
    BorderWriter bw{bord, blen};
    int *inner = bw.at(0), *left = bw.at(1), *outer = bw.at(2), *bottom = bw.at(3);
    for (int k = 0; k <= Nt; ++k) {
        inner[k] = (Nt - k) * col;
        outer[k] = k * col + Nr;
    }
    for (int k = 0; k <= Nr; ++k) {
        bottom[k] = k;
        left[k] = Nt * col + Nr - k;
    }
    return CP_OK;
}

int cp_mesh_max_index_diff(int ne, const int* conn, int* out)
{
    int mx = 0;
    for (int e = 0; e < ne; ++e)
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) mx = std::max(mx, std::abs(conn[4 * e + i] - conn[4 * e + j]));
    *out = mx;
    return CP_OK;
}

// Border traction (PlasticitySolver.cu:1022-1103): each segment [n1, n2]
// adds f(n) * h * |n2 - n1| / 2 to both end nodes; with norm_orient the
// value is (normal, tangential) with normal = (t.y, -t.x).  Where sec marks
// border[k] as a mid-side node (order-2 meshes) the segment is [n1, n2, n3]
// with the quadratic split f1 L/6, f2 2L/3, f3 L/6 (:1025-1069), the mid
// node's frame being the mean of the two half segments' frames.
int cp_rhs_border_force_q(const double* x, const double* y, const int* border, int len, const unsigned char* sec,
                          const double* fval, int norm_orient, double h, double* rp)
{
    auto frame = [&](int a, int b, double& tx, double& ty, double& nx, double& ny) {
        tx = x[b] - x[a];
        ty = y[b] - y[a];
        const double it = 1. / std::sqrt(tx * tx + ty * ty);
        tx *= it;
        ty *= it;
        nx = ty;
        ny = -tx;
        const double in = 1. / std::sqrt(nx * nx + ny * ny);
        nx *= in;
        ny *= in;
    };
    for (int k = 1; k < len; ++k) {
        if (sec && sec[border[k]] && k + 1 < len) {
            ++k;
            const int a = border[k - 2], m = border[k - 1], b = border[k];
            const double d1x = 0.5 * (x[a] - x[m]), d1y = 0.5 * (y[a] - y[m]);
            const double d2x = 0.5 * (x[m] - x[b]), d2y = 0.5 * (y[m] - y[b]);
            const double dl1 = std::sqrt(d1x * d1x + d1y * d1y), dl2 = std::sqrt(d2x * d2x + d2y * d2y);
            double fa[2] = {fval[2 * (k - 2)], fval[2 * (k - 2) + 1]};
            double fm[2] = {fval[2 * (k - 1)], fval[2 * (k - 1) + 1]};
            double fb[2] = {fval[2 * k], fval[2 * k + 1]};
            if (norm_orient) {
                double t1x, t1y, n1x, n1y, t2x, t2y, n2x, n2y;
                frame(a, m, t1x, t1y, n1x, n1y);
                frame(m, b, t2x, t2y, n2x, n2y);
                const double ax = n1x * fa[0] + t1x * fa[1], ay = n1y * fa[0] + t1y * fa[1];
                const double mx = 0.5 * ((n1x + n2x) * fm[0] + (t1x + t2x) * fm[1]);
                const double my = 0.5 * ((n1y + n2y) * fm[0] + (t1y + t2y) * fm[1]);
                const double bx = n2x * fb[0] + t2x * fb[1], by = n2y * fb[0] + t2y * fb[1];
                fa[0] = ax; fa[1] = ay; fm[0] = mx; fm[1] = my; fb[0] = bx; fb[1] = by;
            }
            const double wa = h * dl1 * 2., wm = h * (dl1 + dl2), wb = h * dl2 * 2.;
            rp[2 * a] += fa[0] / 3. * wa;
            rp[2 * a + 1] += fa[1] / 3. * wa;
            rp[2 * m] += fm[0] * 4. / 3. * wm;
            rp[2 * m + 1] += fm[1] * 4. / 3. * wm;
            rp[2 * b] += fb[0] / 3. * wb;
            rp[2 * b + 1] += fb[1] / 3. * wb;
            continue;
        }
        const int a = border[k - 1], b = border[k];
        const double hdx = 0.5 * (x[a] - x[b]), hdy = 0.5 * (y[a] - y[b]);
        const double dl = std::sqrt(hdx * hdx + hdy * hdy);
        double fa[2] = {fval[2 * (k - 1)], fval[2 * (k - 1) + 1]};
        double fb[2] = {fval[2 * k], fval[2 * k + 1]};
        if (norm_orient) {
            double tx, ty, nx, ny;
            frame(a, b, tx, ty, nx, ny);
            const double ax = nx * fa[0] + tx * fa[1], ay = ny * fa[0] + ty * fa[1];
            const double bx = nx * fb[0] + tx * fb[1], by = ny * fb[0] + ty * fb[1];
            fa[0] = ax; fa[1] = ay; fb[0] = bx; fb[1] = by;
        }
        const double w = h * dl;
        rp[2 * a] += fa[0] * w;
        rp[2 * a + 1] += fa[1] * w;
        rp[2 * b] += fb[0] * w;
        rp[2 * b + 1] += fb[1] * w;
    }
    return CP_OK;
}

int cp_rhs_border_force(const double* x, const double* y, const int* border, int len, const double* fval,
                        int norm_orient, double h, double* rp)
{
    return cp_rhs_border_force_q(x, y, border, len, nullptr, fval, norm_orient, h, rp);
}

int cp_rhs_body_force(int ne, const double* x, const double* y, const int* conn, double gx, double gy,
                      double rho, double h, double* rp)
{
    for (int e = 0; e < ne; ++e) {
        const int* c = conn + 4 * e;
        // elemSpace4 (Mesh.cuh:258-267)
        double s1 = 0., s2 = 0.;
        for (int i = 1; i < 4; ++i) {
            s1 += x[c[i - 1]] * y[c[i]];
            s2 += y[c[i - 1]] * x[c[i]];
        }
        s1 += x[c[3]] * y[c[0]];
        s2 += y[c[3]] * x[c[0]];
        const double w = 0.25 * rho * h * (std::fabs(s1 - s2) * 0.5);
        for (int i = 0; i < 4; ++i) {
            rp[2 * c[i]] += w * gx;
            rp[2 * c[i] + 1] += w * gy;
        }
    }
    return CP_OK;
}

int cp_rhs_body_force_nodal(int ne, const double* x, const double* y, const int* conn, const double* rnode,
                            double rho, double h, double* rp)
{
    for (int e = 0; e < ne; ++e) {
        const int* c = conn + 4 * e;
        double s1 = 0., s2 = 0.;
        for (int i = 1; i < 4; ++i) {
            s1 += x[c[i - 1]] * y[c[i]];
            s2 += y[c[i - 1]] * x[c[i]];
        }
        s1 += x[c[3]] * y[c[0]];
        s2 += y[c[3]] * x[c[0]];
        const double w = 0.25 * rho * h * (std::fabs(s1 - s2) * 0.5);
        for (int i = 0; i < 4; ++i) {
            rp[2 * c[i]] += w * rnode[2 * c[i]];
            rp[2 * c[i] + 1] += w * rnode[2 * c[i] + 1];
        }
    }
    return CP_OK;
}

// Body force on 8-node quads (PlasticitySolver.cu:1118-1136): corner nodes
// take -rho h A / 12, mid-side nodes rho h A / 3 of R at the node, A the
// octagon area of elemSpace8 (Mesh.cuh:269-278).  R is evaluated at the
// element's own nodes (the reference reads node[2 * index] there, which
// only a constant R survives).
int cp_rhs_body_force_nodal8(int ne, const double* x, const double* y, const int* conn, const double* rnode,
                             double rho, double h, double* rp)
{
    for (int e = 0; e < ne; ++e) {
        const int* c = conn + 8 * (size_t)e;
        double s1 = 0., s2 = 0.;
        for (int i = 0; i < 3; ++i) {
            s1 += x[c[i]] * y[c[i + 4]] + x[c[i + 4]] * y[c[i + 1]];
            s2 += y[c[i]] * x[c[i + 4]] + y[c[i + 4]] * x[c[i + 1]];
        }
        s1 += x[c[3]] * y[c[7]] + x[c[7]] * y[c[0]];
        s2 += y[c[3]] * x[c[7]] + y[c[7]] * x[c[0]];
        const double space = std::fabs(s1 - s2) * 0.5;
        const double w1 = -rho * h * space / 12., w2 = rho * h * space / 3.;
        for (int i = 0; i < 8; ++i) {
            const double w = i < 4 ? w1 : w2;
            rp[2 * c[i]] += w * rnode[2 * c[i]];
            rp[2 * c[i] + 1] += w * rnode[2 * c[i] + 1];
        }
    }
    return CP_OK;
}

int cp_fix_axis(int nn, const double* x, const double* y, int vertical, double coord, unsigned char* mask)
{
    const double* v = vertical ? x : y;
    const int comp = vertical ? 0 : 1;
    for (int i = 0; i < nn; ++i)
        if (std::fabs(v[i] - coord) < 1e-14) mask[2 * i + comp] = 0;
    return CP_OK;
}

int cp_displacement_border(const int* border, int len, double ux, double uy, int fix_mid_x, int fix_mid_y,
                           unsigned char* mask, double* u)
{
    if (len < 1) return cp::fail(CP_ERR_INVALID, "cp_displacement_border: empty border");
    const int end = len - (border[0] == border[len - 1] ? 1 : 0);
    const double val[2] = {ux, uy};
    const int mid[2] = {fix_mid_x, fix_mid_y};
    for (int comp = 0; comp < 2; ++comp) {
        if (mid[comp]) {
            const int d = 2 * border[len / 2] + comp;
            mask[d] = 0;
            u[d] = val[comp];
        } else {
            for (int i = 0; i < end; ++i) {
                const int d = 2 * border[i] + comp;
                mask[d] = 0;
                u[d] = val[comp];
            }
        }
    }
    return CP_OK;
}

int cp_displacement_border_values(const int* border, int len, const double* uval, int fix_mid_x, int fix_mid_y,
                                  unsigned char* mask, double* u)
{
    if (len < 1) return cp::fail(CP_ERR_INVALID, "cp_displacement_border_values: empty border");
    const int end = len - (border[0] == border[len - 1] ? 1 : 0);
    const int mid[2] = {fix_mid_x, fix_mid_y};
    for (int comp = 0; comp < 2; ++comp) {
        if (mid[comp]) {
            const int d = 2 * border[len / 2] + comp;
            mask[d] = 0;
            u[d] = uval[2 * (len / 2) + comp];
        } else {
            for (int i = 0; i < end; ++i) {
                const int d = 2 * border[i] + comp;
                mask[d] = 0;
                u[d] = uval[2 * i + comp];
            }
        }
    }
    return CP_OK;
}

int cp_csr_nnz_n(int nn, int ne, int npe, const int* conn, long long* nnz)
{
    cp::Adjacency adj;
    if (int s = cp::build_adjacency(nn, ne, npe, conn, adj)) return s;
    *nnz = 4LL * (long long)adj.ptr[nn];
    return CP_OK;
}

int cp_csr_pattern_n(int nn, int ne, int npe, const int* conn, int* rows, int* cols)
{
    cp::Adjacency adj;
    if (int s = cp::build_adjacency(nn, ne, npe, conn, adj)) return s;
    cp::export_csr(adj, nn, rows, cols);
    return CP_OK;
}

int cp_csr_nnz(int nn, int ne, const int* conn, long long* nnz) { return cp_csr_nnz_n(nn, ne, 4, conn, nnz); }

int cp_csr_pattern(int nn, int ne, const int* conn, int* rows, int* cols)
{
    return cp_csr_pattern_n(nn, ne, 4, conn, rows, cols);
}

// Mesh::remapOrder (Mesh.cu:443-503).  In the strip numbering k = i*width + j
// of the order-1 quads, the 2x2 patch with i, j even becomes 8-node quad
// (i/2)(width/2) + j/2: corners = local nodes 0, 1, 2, 3 of quads (i+1, j+1),
// (i, j+1), (i, j), (i+1, j); mid-sides = local nodes 1, 2, 3, 0 of the same
// quads.  The patch centre (local node 0 of quad (i, j)) is dropped and the
// remaining nodes are renumbered in their order; bord (nbord entries) is
// renumbered in place and sec marks the mid-side nodes.
int cp_mesh_remap_order(int nn, int ne4, const double* x, const double* y, const int* conn4, int width, int nbord,
                        int* bord, double* x8, double* y8, int* conn8, unsigned char* sec, int* nn8)
{
    if (nn < 1 || width < 2 || (width & 1) || ne4 % width || ((ne4 / width) & 1) || !x || !y || !conn4 || !x8 ||
        !y8 || !conn8 || !sec || !nn8 || (nbord > 0 && !bord))
        return cp::fail(CP_ERR_INVALID, "cp_mesh_remap_order: the order-1 mesh must be 2m x 2n strips of quads");
    const int strips = ne4 / width;
    std::vector<unsigned char> keep(nn, 1);
    for (int i = 0; i < strips; i += 2)
        for (int j = 0; j < width; j += 2) {
            const int* qa = conn4 + 4 * ((size_t)(i + 1) * width + j + 1);
            const int* qb = conn4 + 4 * ((size_t)i * width + j + 1);
            const int* qc = conn4 + 4 * ((size_t)i * width + j);
            const int* qd = conn4 + 4 * ((size_t)(i + 1) * width + j);
            int* c = conn8 + 8 * ((size_t)(i / 2) * (width / 2) + j / 2);
            const int v[8] = {qa[0], qb[1], qc[2], qd[3], qa[1], qb[2], qc[3], qd[0]};
            for (int t = 0; t < 8; ++t) {
                if (v[t] < 0 || v[t] >= nn) return cp::fail(CP_ERR_INVALID, "cp_mesh_remap_order: bad node index");
                c[t] = v[t];
            }
            keep[qc[0]] = 0;
        }
    std::vector<int> id(nn, -1);
    int m = 0;
    for (int n = 0; n < nn; ++n)
        if (keep[n]) {
            id[n] = m;
            x8[m] = x[n];
            y8[m] = y[n];
            ++m;
        }
    const size_t ne8 = (size_t)ne4 / 4;
    for (size_t k = 0; k < 8 * ne8; ++k) conn8[k] = id[conn8[k]];
    for (int k = 0; k < nbord; ++k) {
        if (bord[k] < 0 || bord[k] >= nn || id[bord[k]] < 0)
            return cp::fail(CP_ERR_INVALID, "cp_mesh_remap_order: border through a removed node");
        bord[k] = id[bord[k]];
    }
    std::memset(sec, 0, (size_t)m);
    for (size_t e = 0; e < ne8; ++e)
        for (int t = 4; t < 8; ++t) sec[conn8[8 * e + t]] = 1;
    *nn8 = m;
    return CP_OK;
}

// Mesh::loadFromFile (Mesh.cu:725-937): an Abaqus-style .inp file.  Lines
// are matched as the reference matches them: a block starts at a line equal
// to "*NODE" or "*MPC", or beginning with "*ELEMENT" and holding "TYPE=S3",
// "TYPE=S4" or "TYPE=S8R", and runs to the next empty line or line starting
// with '*'.  Commas count as blanks.  Node lines are "id x y ..." with 1-based
// ids; element lines "id n1 n2 ..." (1-based); every *MPC block is one border,
// its lines holding the border node after the first four characters.  One
// node per border (the MPC reference node) is not a mesh node: the node count
// is the number of node lines less the number of borders, and node lines with
// a larger id are skipped.
int cp_mesh_inp_sizes(const char* path, int* nn, int* n3, int* n4, int* n8, int* nb, int* blen, int max_borders)
{
    if (!nn || !n3 || !n4 || !n8 || !nb) return cp::fail(CP_ERR_INVALID, "cp_mesh_inp_sizes: null argument");
    InpCounts c;
    int s = walk_inp(path, [&](int kind, int, int b, const std::string&) {
        if (kind == 1) ++c.nodeLines;
        else if (kind == 3) ++c.n3;
        else if (kind == 4) ++c.n4;
        else if (kind == 8) ++c.n8;
        else {
            if ((int)c.blen.size() <= b) c.blen.resize(b + 1, 0);
            ++c.blen[b];
        }
        return CP_OK;
    });
    if (s) return s;
    *nn = c.nodeLines - (int)c.blen.size();
    *n3 = c.n3;
    *n4 = c.n4;
    *n8 = c.n8;
    *nb = (int)c.blen.size();
    if (*nn < 0) return cp::fail(CP_ERR_INVALID, "cp_mesh_inp_sizes: fewer nodes than borders");
    if (blen) {
        if (*nb > max_borders) return cp::fail(CP_ERR_INVALID, "cp_mesh_inp_sizes: more borders than blen holds");
        for (int b = 0; b < *nb; ++b) blen[b] = c.blen[b];
    }
    return CP_OK;
}

int cp_mesh_load_inp(const char* path, double* x, double* y, int* conn3, int* conn4, int* conn8, unsigned char* sec,
                     int* bord)
{
    int nn, n3, n4, n8, nb;
    if (int s = cp_mesh_inp_sizes(path, &nn, &n3, &n4, &n8, &nb, nullptr, 0)) return s;
    std::vector<int> boff;   // start of each border in bord
    {
        InpCounts c;
        walk_inp(path, [&](int kind, int, int b, const std::string&) {
            if (kind == 9) {
                if ((int)c.blen.size() <= b) c.blen.resize(b + 1, 0);
                ++c.blen[b];
            }
            return CP_OK;
        });
        int off = 0;
        for (int b = 0; b < nb; ++b) {
            boff.push_back(off);
            off += c.blen[b];
        }
    }
    if (sec) std::memset(sec, 0, (size_t)std::max(nn, 0));
    int e3 = 0, e4 = 0, e8 = 0;
    auto node_index = [&](long v, int& out) {
        if (v < 1 || v > nn) return cp::fail(CP_ERR_INVALID, "cp_mesh_load_inp: node index out of range");
        out = (int)(v - 1);
        return (int)CP_OK;
    };
    return walk_inp(path, [&](int kind, int k, int b, const std::string& raw) {
        std::string line = raw;
        blank_commas(line);
        if (kind == 9) {
            for (size_t i = 0; i < 4 && i < line.size(); ++i) line[i] = ' ';
            std::istringstream ss(line);
            long v = 0;
            if (!(ss >> v)) return cp::fail(CP_ERR_INVALID, "cp_mesh_load_inp: bad *MPC line");
            return node_index(v, bord[boff[b] + k]);
        }
        std::istringstream ss(line);
        long id = 0;
        if (!(ss >> id)) return cp::fail(CP_ERR_INVALID, "cp_mesh_load_inp: bad line");
        if (kind == 1) {
            double px = 0., py = 0.;
            if (!(ss >> px >> py)) return cp::fail(CP_ERR_INVALID, "cp_mesh_load_inp: bad *NODE line");
            if (id >= 1 && id <= nn) {   // larger ids: the borders' reference nodes
                x[id - 1] = px;
                y[id - 1] = py;
            }
            return (int)CP_OK;
        }
        int* c = kind == 3 ? conn3 + 3 * (size_t)e3++ : kind == 4 ? conn4 + 4 * (size_t)e4++ : conn8 + 8 * (size_t)e8++;
        for (int i = 0; i < kind; ++i) {
            long v = 0;
            if (!(ss >> v)) return cp::fail(CP_ERR_INVALID, "cp_mesh_load_inp: short element line");
            if (int s = node_index(v, c[i])) return s;
        }
        if (kind == 8 && sec)
            for (int i = 4; i < 8; ++i) sec[c[i]] = 1;
        return (int)CP_OK;
    });
}

// Order-2 generators (genRectangle / genRectWithHole / genRing with order = 2):
// the order-1 mesh of doubled resolution regrouped by remapOrder.
int cp_mesh_rect2_sizes(int N1, int N2, int* nn, int* ne, int* nb, int* blen)
{
    int n1, e1;
    if (int s = cp_mesh_rect_sizes(2 * N1, 2 * N2, &n1, &e1, nb, blen)) return s;
    *ne = e1 / 4;
    *nn = n1 - *ne;
    return CP_OK;
}

static int remap_into(int nn1, int ne1, const std::vector<double>& x1, const std::vector<double>& y1,
                      const std::vector<int>& c1, int width, std::vector<int>& b1, int* bord, double* x, double* y,
                      int* conn, unsigned char* sec)
{
    int nn8 = 0;
    if (int s = cp_mesh_remap_order(nn1, ne1, x1.data(), y1.data(), c1.data(), width, (int)b1.size(), b1.data(), x, y,
                                    conn, sec, &nn8))
        return s;
    std::copy(b1.begin(), b1.end(), bord);
    return CP_OK;
}

int cp_mesh_rect2(double x1, double x2, double y1, double y2, int N1, int N2, double* x, double* y, int* conn,
                  unsigned char* sec, int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_rect_sizes(2 * N1, 2 * N2, &nn, &ne, &nb, blen)) return s;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(2 * (blen[0] + blen[1]));
    if (int s = cp_mesh_rect(x1, x2, y1, y2, 2 * N1, 2 * N2, X.data(), Y.data(), C.data(), B.data(), blen)) return s;
    return remap_into(nn, ne, X, Y, C, 2 * N2, B, bord, x, y, conn, sec);
}

int cp_mesh_hole2_sizes(double x1, double x2, double y1, double y2, double r, int N1, int N2, int* nn, int* ne,
                        int* nb, int* blen)
{
    if (N1 < 1 || N2 < 1 || !(r > 0)) return cp::fail(CP_ERR_INVALID, "cp_mesh_hole2_sizes: bad arguments");
    const int radN = 2 * radial_segments(x1, x2, y1, y2, r, N1, N2), borderN = 4 * (N1 + N2);
    if (radN < 2) return cp::fail(CP_ERR_INVALID, "cp_mesh_hole2_sizes: hole too large for the plate");
    const int n1 = borderN * (radN + 1), e1 = borderN * radN;
    *ne = e1 / 4;
    *nn = n1 - *ne;
    *nb = 5;
    if (blen) {
        blen[0] = blen[2] = 2 * N1 + 1;
        blen[1] = blen[3] = 2 * N2 + 1;
        blen[4] = borderN + 1;
    }
    return CP_OK;
}

int cp_mesh_hole2(double x1, double x2, double y1, double y2, double r, int N1, int N2, double* x, double* y,
                  int* conn, unsigned char* sec, int* bord, int* blen)
{
    int nn8, ne8, nb;
    if (int s = cp_mesh_hole2_sizes(x1, x2, y1, y2, r, N1, N2, &nn8, &ne8, &nb, blen)) return s;
    const int radN = 2 * radial_segments(x1, x2, y1, y2, r, N1, N2), borderN = 4 * (N1 + N2);
    const int nn = borderN * (radN + 1), ne = borderN * radN;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(2 * (blen[0] + blen[1]) + blen[4]);
    if (int s = hole_mesh(x1, x2, y1, y2, r, 2 * N1, 2 * N2, radN, X.data(), Y.data(), C.data(), B.data(), blen))
        return s;
    return remap_into(nn, ne, X, Y, C, radN, B, bord, x, y, conn, sec);
}

// The corner nodes of the 8-node mesh remapOrder builds from an order-1
// mesh of lattice n1 x n2 (node_of1: lattice point -> order-1 node; quads in
// strips of `width` along direction 2, one strip per direction-1 step): the
// lattice points with both coordinates even, as order-2 node numbers (the
// patch centres dropped and the rest renumbered in their order, as
// cp_mesh_remap_order does).
static int corner_lattice(int nn1, int ne1, const int* conn1, int width, int n1, int n2, const int* node_of1,
                          bool per1, int* m1, int* m2, int* corner_of)
{
    if (width < 2 || (width & 1) || ne1 % width || (per1 ? (n1 & 1) : !(n1 & 1)) || !(n2 & 1))
        return cp::fail(CP_ERR_INVALID, "corner lattice: the order-1 lattice must have even cell counts");
    std::vector<unsigned char> keep(nn1, 1);
    const int strips = ne1 / width;
    for (int i = 0; i < strips; i += 2)
        for (int j = 0; j < width; j += 2) keep[conn1[4 * ((size_t)i * width + j) + 0]] = 0;
    std::vector<int> id(nn1, -1);
    for (int n = 0, m = 0; n < nn1; ++n)
        if (keep[n]) id[n] = m++;
    *m1 = per1 ? n1 / 2 : n1 / 2 + 1;
    *m2 = n2 / 2 + 1;
    if (corner_of)
        for (int I1 = 0; I1 < *m1; ++I1)
            for (int I2 = 0; I2 < *m2; ++I2) {
                const int c = id[node_of1[(size_t)(2 * I1) * n2 + 2 * I2]];
                if (c < 0) return cp::fail(CP_ERR_INVALID, "corner lattice: a corner is a patch centre");
                corner_of[(size_t)I1 * *m2 + I2] = c;
            }
    return CP_OK;
}

int cp_mesh_rect2_lattice(int N1, int N2, int* m1, int* m2, int* corner_of)
{
    if (!m1 || !m2) return cp::fail(CP_ERR_INVALID, "cp_mesh_rect2_lattice: null argument");
    int nn, ne, nb, blen[8];
    if (int s = cp_mesh_rect_sizes(2 * N1, 2 * N2, &nn, &ne, &nb, blen)) return s;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(2 * (blen[0] + blen[1])), node_of(nn);
    if (int s = cp_mesh_rect(0., 1., 0., 1., 2 * N1, 2 * N2, X.data(), Y.data(), C.data(), B.data(), blen)) return s;
    for (int n = 0; n < nn; ++n) node_of[n] = n;
    return corner_lattice(nn, ne, C.data(), 2 * N2, 2 * N1 + 1, 2 * N2 + 1, node_of.data(), false, m1, m2, corner_of);
}

int cp_mesh_hole2_lattice(double x1, double x2, double y1, double y2, double r, int N1, int N2, int* m1, int* m2,
                          int* corner_of)
{
    if (!m1 || !m2) return cp::fail(CP_ERR_INVALID, "cp_mesh_hole2_lattice: null argument");
    int nn8, ne8, nb, blen[8];
    if (int s = cp_mesh_hole2_sizes(x1, x2, y1, y2, r, N1, N2, &nn8, &ne8, &nb, blen)) return s;
    const int radN = 2 * radial_segments(x1, x2, y1, y2, r, N1, N2), borderN = 4 * (N1 + N2);
    const int nn = borderN * (radN + 1), ne = borderN * radN;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(2 * (blen[0] + blen[1]) + blen[4]), node_of(nn);
    if (int s = hole_mesh(x1, x2, y1, y2, r, 2 * N1, 2 * N2, radN, X.data(), Y.data(), C.data(), B.data(), blen))
        return s;
    int l1 = 0, l2 = 0;
    if (int s = cp_mesh_ring_lattice(nn, borderN, &l1, &l2, node_of.data())) return s;
    return corner_lattice(nn, ne, C.data(), radN, l1, l2, node_of.data(), true, m1, m2, corner_of);
}

int cp_mesh_ring2_lattice(int Nphi, int Nr, int* m1, int* m2, int* corner_of)
{
    if (!m1 || !m2) return cp::fail(CP_ERR_INVALID, "cp_mesh_ring2_lattice: null argument");
    int nn, ne, nb, blen[8];
    if (int s = cp_mesh_ring_sizes(2 * Nphi, 2 * Nr, &nn, &ne, &nb, blen)) return s;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(blen[0] + blen[1]), node_of(nn);
    if (int s = cp_mesh_ring(0.5, 1.0, 2 * Nphi, 2 * Nr, X.data(), Y.data(), C.data(), B.data(), blen)) return s;
    int l1 = 0, l2 = 0;
    if (int s = cp_mesh_ring_lattice(nn, 2 * Nphi, &l1, &l2, node_of.data())) return s;
    return corner_lattice(nn, ne, C.data(), 2 * Nr, l1, l2, node_of.data(), true, m1, m2, corner_of);
}

int cp_mesh_ring2_sizes(int Nphi, int Nr, int* nn, int* ne, int* nb, int* blen)
{
    int n1, e1;
    if (int s = cp_mesh_ring_sizes(2 * Nphi, 2 * Nr, &n1, &e1, nb, blen)) return s;
    *ne = e1 / 4;
    *nn = n1 - *ne;
    return CP_OK;
}

int cp_mesh_ring2(double a, double b, int Nphi, int Nr, double* x, double* y, int* conn, unsigned char* sec,
                  int* bord, int* blen)
{
    int nn, ne, nb;
    if (int s = cp_mesh_ring_sizes(2 * Nphi, 2 * Nr, &nn, &ne, &nb, blen)) return s;
    std::vector<double> X(nn), Y(nn);
    std::vector<int> C(4 * (size_t)ne), B(blen[0] + blen[1]);
    if (int s = cp_mesh_ring(a, b, 2 * Nphi, 2 * Nr, X.data(), Y.data(), C.data(), B.data(), blen)) return s;
    return remap_into(nn, ne, X, Y, C, 2 * Nr, B, bord, x, y, conn, sec);
}

int cp_mesh_renumerate_ring(int nn, int ne, double* x, double* y, int* conn, int nb, int* bord, const int* blen,
                            int borderN, int* new_of_old)
{
    if (nn < 1 || ne < 0 || !x || !y || (ne > 0 && !conn) || nb < 0 || (nb > 0 && (!bord || !blen)) || borderN < 1 ||
        nn % borderN != 0)
        return cp::fail(CP_ERR_INVALID, "cp_mesh_renumerate_ring: the node count must be a multiple of borderN");
    for (int k = 0; k < 4 * ne; ++k)
        if (conn[k] < 0 || conn[k] >= nn) return cp::fail(CP_ERR_INVALID, "cp_mesh_renumerate_ring: node out of range");
    renumber_ring(nn, ne, x, y, conn, nb, bord, blen, borderN);
    if (new_of_old) {
        const std::vector<int> perm = ring_permutation(nn, borderN);
        std::copy(perm.begin(), perm.end(), new_of_old);
    }
    return CP_OK;
}

// Lattice of a mesh numbered by renumerateRing: columns (the rays of a ring
// or of a plate with a hole) along direction 1, periodic, rows along
// direction 2; node_of[j * (nn / borderN) + i] = number of ray j's row-i node.
int cp_mesh_ring_lattice(int nn, int borderN, int* n1, int* n2, int* node_of)
{
    if (nn < 1 || borderN < 3 || nn % borderN != 0 || !n1 || !n2)
        return cp::fail(CP_ERR_INVALID, "cp_mesh_ring_lattice: the node count must be a multiple of borderN >= 3");
    const int rows = nn / borderN;
    *n1 = borderN;
    *n2 = rows;
    if (node_of) {
        const std::vector<int> perm = ring_permutation(nn, borderN);
        for (int j = 0; j < borderN; ++j)
            for (int i = 0; i < rows; ++i) node_of[j * rows + i] = perm[i * borderN + j];
    }
    return CP_OK;
}

// Mesh::renumByDirection (Mesh.cu:537-630): nodes renumbered in increasing
// order of their projection x dx + y dy onto the direction.  The reference
// sorts with a top-down merge sort that takes the element of the right run
// on ties (Mesh.cu:505-535); that order is the order of (projection, original
// number descending), which is what ties get here.
int cp_mesh_renum_by_direction(int nn, double* x, double* y, int ne, int npe, int* conn, int nbord, int* bord,
                               double dx, double dy, int* new_of_old)
{
    if (nn < 1 || !x || !y || ne < 0 || (npe != 4 && npe != 8) || (ne > 0 && !conn) || nbord < 0 ||
        (nbord > 0 && !bord))
        return cp::fail(CP_ERR_INVALID, "cp_mesh_renum_by_direction: bad arguments");
    std::vector<double> key(nn);
    for (int i = 0; i < nn; ++i) {
        key[i] = x[i] * dx + y[i] * dy;
        if (key[i] != key[i]) return cp::fail(CP_ERR_INVALID, "cp_mesh_renum_by_direction: NaN coordinate or direction");
    }
    for (long long k = 0; k < (long long)npe * ne; ++k)
        if (conn[k] < 0 || conn[k] >= nn) return cp::fail(CP_ERR_INVALID, "cp_mesh_renum_by_direction: node out of range");
    for (int k = 0; k < nbord; ++k)
        if (bord[k] < 0 || bord[k] >= nn) return cp::fail(CP_ERR_INVALID, "cp_mesh_renum_by_direction: node out of range");
    std::vector<int> old_at(nn);   // old node at each new position
    for (int i = 0; i < nn; ++i) old_at[i] = i;
    std::sort(old_at.begin(), old_at.end(), [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a > b); });
    std::vector<int> perm(nn);
    for (int i = 0; i < nn; ++i) perm[old_at[i]] = i;
    for (long long k = 0; k < (long long)npe * ne; ++k) conn[k] = perm[conn[k]];
    for (int k = 0; k < nbord; ++k) bord[k] = perm[bord[k]];
    std::vector<double> nx(nn), ny(nn);
    for (int i = 0; i < nn; ++i) {
        nx[perm[i]] = x[i];
        ny[perm[i]] = y[i];
    }
    std::memcpy(x, nx.data(), sizeof(double) * nn);
    std::memcpy(y, ny.data(), sizeof(double) * nn);
    if (new_of_old) std::copy(perm.begin(), perm.end(), new_of_old);
    return CP_OK;
}

}  // extern "C"

namespace cp {

// Node adjacency in the reference's order (CudaSparseSLAE.cuh:339-391):
// the node itself, then the other nodes of its elements in the order they
// are met scanning elements in storage order (local order inside an
// element), duplicates dropped.  Built per node from an element-ordered
// incidence list, so the per-node work is independent (OpenMP).
int build_adjacency(int nn, int ne, int npe, const int* conn, Adjacency& adj)
{
    if (npe != 4 && npe != 8) return fail(CP_ERR_INVALID, "elements must have 4 or 8 nodes");
    if ((long long)ne * npe >= (1LL << 31)) return fail(CP_ERR_INVALID, "too many elements");
    const long long nc = (long long)npe * ne;
    for (long long k = 0; k < nc; ++k)
        if (conn[k] < 0 || conn[k] >= nn) return fail(CP_ERR_INVALID, "connectivity index out of range");
    for (int e = 0; e < ne; ++e) {
        const int* c = conn + (size_t)npe * e;
        for (int i = 0; i < npe; ++i)
            for (int j = i + 1; j < npe; ++j)
                if (c[i] == c[j]) return fail(CP_ERR_INVALID, "degenerate element (repeated node)");
    }
    std::vector<int> incPtr(nn + 1, 0);
    for (long long k = 0; k < nc; ++k) ++incPtr[conn[k] + 1];
    for (int n = 0; n < nn; ++n) incPtr[n + 1] += incPtr[n];
    std::vector<int> inc(incPtr[nn]);  // element * npe + local index, element order per node
    {
        std::vector<int> fill(incPtr.begin(), incPtr.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int i = 0; i < npe; ++i) inc[fill[conn[(size_t)npe * e + i]]++] = npe * e + i;
    }
    std::vector<int> cnt(nn + 1, 0);
    const int others = npe - 1;
    std::vector<int> scratch((size_t)incPtr[nn] * others + nn + 1);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nn; ++n) {
        int* list = scratch.data() + (size_t)others * incPtr[n] + n;
        int u = 0;
        list[u++] = n;
        for (int q = incPtr[n]; q < incPtr[n + 1]; ++q) {
            const int e = inc[q] / npe, i = inc[q] % npe;
            for (int j = 0; j < npe; ++j) {
                if (j == i) continue;
                const int v = conn[(size_t)npe * e + j];
                bool seen = false;
                for (int t = 1; t < u; ++t)
                    if (list[t] == v) { seen = true; break; }
                if (!seen) list[u++] = v;
            }
        }
        cnt[n + 1] = u;
    }
    adj.ptr.assign(nn + 1, 0);
    for (int n = 0; n < nn; ++n) adj.ptr[n + 1] = adj.ptr[n] + cnt[n + 1];
    adj.node.resize(adj.ptr[nn]);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nn; ++n)
        std::memcpy(adj.node.data() + adj.ptr[n], scratch.data() + (size_t)others * incPtr[n] + n,
                    sizeof(int) * (adj.ptr[n + 1] - adj.ptr[n]));
    adj.npe = npe;
    adj.incPtr.swap(incPtr);
    adj.inc.swap(inc);
    return CP_OK;
}

// Reference CSR layout (CudaSparseSLAE.cuh:414-440): rows 2n and 2n+1 both
// list the columns 2m, 2m+1 for every m in adj(n) (self first).
void export_csr(const Adjacency& adj, int nn, int* rows, int* cols)
{
    rows[0] = 0;
    for (int n = 0; n < nn; ++n) {
        const int len = 2 * (adj.ptr[n + 1] - adj.ptr[n]);
        rows[2 * n + 1] = rows[2 * n] + len;
        rows[2 * n + 2] = rows[2 * n + 1] + len;
        for (int i = 0; i < 2; ++i) {
            int* c = cols + rows[2 * n + i];
            for (int k = adj.ptr[n]; k < adj.ptr[n + 1]; ++k) {
                *c++ = 2 * adj.node[k];
                *c++ = 2 * adj.node[k] + 1;
            }
        }
    }
}

}  // namespace cp
