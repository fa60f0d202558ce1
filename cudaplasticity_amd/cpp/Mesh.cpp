// Mesh.cpp -- the drop-in Mesh class over the C-ABI generators.
#include "cudaplasticity/Mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cp_abi.h"

namespace {

void check(int status, const char* what)
{
    if (status != CP_OK) throw std::runtime_error(std::string(what) + ": " + cp_last_error());
}

int as_int(size_t n, const char* what)
{
    if (n > 0x3fffffff) throw std::invalid_argument(std::string(what) + ": size out of range");
    return (int)n;
}

void check_order(int order)
{
    if (order != 1 && order != 2) throw std::invalid_argument("Mesh: order must be 1 (4-node) or 2 (8-node quads)");
}

}  // namespace

Mesh::Mesh()
{
    cp_ctx* ctx = nullptr;
    if (cp_ctx_create(0, &ctx) == CP_OK) {
        char name[256] = {};
        size_t mem = 0;
        int cus = 0;
        cp_device_info(ctx, name, sizeof(name), &mem, &cus);
        cp_ctx_destroy(ctx);
        useCuda = true;
        std::cout << "Detected HIP device" << name << " " << std::llround((double)mem / 1048576.) << " MB, " << cus
                  << " CUs\n\n";
    } else {
        useCuda = false;
        std::cout << "No HIP device: meshes can be built, solving needs a GPU\n\n";
    }
}

Mesh::Mesh(const Mesh& o) { *this = o; }

Mesh& Mesh::operator=(const Mesh& o)
{
    if (this == &o) return *this;
    clear();
    useCuda = o.useCuda;
    ramSaved = o.ramSaved;
    analysed = o.analysed;
    count3 = o.count3;
    count4 = o.count4;
    count8 = o.count8;
    std::copy(o.elemPos, o.elemPos + 3, elemPos);
    nodeCount = o.nodeCount;
    if (o.node) {
        node = new vec2[nodeCount];
        std::copy(o.node, o.node + nodeCount, node);
        secOrdNodes = new bool[nodeCount];
        std::copy(o.secOrdNodes, o.secOrdNodes + nodeCount, secOrdNodes);
    }
    if (o.elem4) {
        elem4 = new int[4 * (size_t)count4];
        std::copy(o.elem4, o.elem4 + 4 * (size_t)count4, elem4);
    }
    if (o.elem8) {
        elem8 = new int[8 * (size_t)count8];
        std::copy(o.elem8, o.elem8 + 8 * (size_t)count8, elem8);
    }
    double* const* src[3] = {&o.spaces, &o.aspects, &o.skewAngles};
    double** dst[3] = {&spaces, &aspects, &skewAngles};
    for (int k = 0; k < 3; ++k)
        if (*src[k]) {
            *dst[k] = new double[count4];
            std::copy(*src[k], *src[k] + count4, *dst[k]);
        }
    if (!spaces) analysed = false;
    bordersCount = o.bordersCount;
    if (bordersCount) {
        borderLength = new int[bordersCount];
        borders = new int*[bordersCount];
        for (int b = 0; b < bordersCount; ++b) {
            borderLength[b] = o.borderLength[b];
            borders[b] = new int[borderLength[b]];
            std::copy(o.borders[b], o.borders[b] + borderLength[b], borders[b]);
        }
    }
    latticeN1 = o.latticeN1;
    latticeN2 = o.latticeN2;
    latticePeriodic = o.latticePeriodic;
    latticeNodes = o.latticeNodes;
    return *this;
}

Mesh::~Mesh() { clear(); }

void Mesh::clear()
{
    delete[] spaces;
    delete[] aspects;
    delete[] skewAngles;
    spaces = aspects = skewAngles = nullptr;
    delete[] node;
    delete[] secOrdNodes;
    delete[] elem3;
    delete[] elem4;
    delete[] elem8;
    for (int b = 0; b < bordersCount; ++b) delete[] borders[b];
    delete[] borders;
    delete[] borderLength;
    node = nullptr;
    secOrdNodes = nullptr;
    elem3 = elem4 = elem8 = nullptr;
    borders = nullptr;
    borderLength = nullptr;
    bordersCount = 0;
    nodeCount = 0;
    count3 = count4 = count8 = 0;
    elemPos[0] = elemPos[1] = elemPos[2] = 0;
    analysed = false;
    latticeN1 = latticeN2 = 0;
    latticePeriodic = false;
    latticeNodes.clear();
}

// sizes(nn, ne, nb, blen) then gen(x, y, conn, bord, blen) as the C-ABI
// generators define them; fills the reference's public arrays.
template <typename Sizes, typename Gen>
void Mesh::generate(Sizes sizes, Gen gen)
{
    int nn = 0, ne = 0, nb = 0;
    int blen[16] = {};
    check(sizes(&nn, &ne, &nb, blen), "mesh sizes");
    if (nb > 16) throw std::runtime_error("mesh generator: too many borders");
    std::vector<double> x(nn), y(nn);
    std::vector<int> conn(4 * (size_t)ne);
    size_t nbord = 0;
    for (int b = 0; b < nb; ++b) nbord += (size_t)blen[b];
    std::vector<int> bord(nbord);
    check(gen(x.data(), y.data(), conn.data(), bord.data(), blen), "mesh generator");

    clear();
    nodeCount = nn;
    node = new vec2[nn];
    secOrdNodes = new bool[nn];
    for (int i = 0; i < nn; ++i) {
        node[i] = vec2(x[i], y[i]);
        secOrdNodes[i] = false;
    }
    count4 = ne;
    elem4 = new int[4 * (size_t)ne];
    std::copy(conn.begin(), conn.end(), elem4);
    elemPos[0] = 0;
    elemPos[1] = ne;
    elemPos[2] = ne;
    setBorders(nb, blen, bord);
}

// The same for the order-2 generators: gen(x, y, conn8, sec, bord, blen).
template <typename Sizes, typename Gen>
void Mesh::generate2(Sizes sizes, Gen gen)
{
    int nn = 0, ne = 0, nb = 0;
    int blen[16] = {};
    check(sizes(&nn, &ne, &nb, blen), "mesh sizes");
    if (nb > 16) throw std::runtime_error("mesh generator: too many borders");
    std::vector<double> x(nn), y(nn);
    std::vector<int> conn(8 * (size_t)ne);
    std::vector<unsigned char> sec(nn);
    size_t nbord = 0;
    for (int b = 0; b < nb; ++b) nbord += (size_t)blen[b];
    std::vector<int> bord(nbord);
    check(gen(x.data(), y.data(), conn.data(), sec.data(), bord.data(), blen), "mesh generator");

    clear();
    nodeCount = nn;
    node = new vec2[nn];
    secOrdNodes = new bool[nn];
    for (int i = 0; i < nn; ++i) {
        node[i] = vec2(x[i], y[i]);
        secOrdNodes[i] = sec[i] != 0;
    }
    count8 = ne;
    elem8 = new int[8 * (size_t)ne];
    std::copy(conn.begin(), conn.end(), elem8);
    elemPos[0] = 0;
    elemPos[1] = 0;
    elemPos[2] = ne;
    setBorders(nb, blen, bord);
}

void Mesh::setBorders(int nb, const int* blen, const std::vector<int>& bord)
{
    bordersCount = nb;
    borderLength = new int[nb];
    borders = new int*[nb];
    size_t off = 0;
    for (int b = 0; b < nb; ++b) {
        borderLength[b] = blen[b];
        borders[b] = new int[blen[b]];
        std::copy(bord.begin() + off, bord.begin() + off + blen[b], borders[b]);
        off += blen[b];
    }
    ramSaved = true;
}

// Rectangles and quarter meshes number their nodes lattice point by lattice
// point: node i1 * n2 + i2.
void Mesh::setLattice(int n1, int n2)
{
    if ((long long)n1 * n2 != nodeCount) throw std::logic_error("mesh lattice does not match the node count");
    latticeN1 = n1;
    latticeN2 = n2;
    latticePeriodic = false;
    latticeNodes.resize(nodeCount);
    for (int n = 0; n < nodeCount; ++n) latticeNodes[n] = n;
}

// Plates with a hole and rings: borderN rays around the hole, numbered by
// renumerateRing.
void Mesh::setRingLattice(int borderN)
{
    int n1 = 0, n2 = 0;
    latticeNodes.resize(nodeCount);
    check(cp_mesh_ring_lattice(nodeCount, borderN, &n1, &n2, latticeNodes.data()), "mesh lattice");
    latticeN1 = n1;
    latticeN2 = n2;
    latticePeriodic = true;
}

// 8-node meshes: the lattice of their corner nodes, from the generator's
// lattice call lat(m1, m2, corner_of).
template <typename Lat>
void Mesh::setCornerLattice(Lat lat, bool periodic)
{
    int m1 = 0, m2 = 0;
    check(lat(&m1, &m2, nullptr), "mesh lattice");
    latticeNodes.resize((size_t)m1 * m2);
    check(lat(&m1, &m2, latticeNodes.data()), "mesh lattice");
    latticeN1 = m1;
    latticeN2 = m2;
    latticePeriodic = periodic;
}

void Mesh::renumberLattice(const std::vector<int>& newOfOld)
{
    for (int& n : latticeNodes) n = newOfOld[n];
}

void Mesh::genRectangle(double x1, double x2, double y1, double y2, size_t N1, size_t N2, int order)
{
    check_order(order);
    const int n1 = as_int(N1, "genRectangle"), n2 = as_int(N2, "genRectangle");
    if (order == 2) {
        generate2([&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_rect2_sizes(n1, n2, nn, ne, nb, bl); },
                  [&](double* x, double* y, int* c, unsigned char* sc, int* b, int* bl) {
                      return cp_mesh_rect2(x1, x2, y1, y2, n1, n2, x, y, c, sc, b, bl);
                  });
        return setCornerLattice(
            [&](int* m1, int* m2, int* co) { return cp_mesh_rect2_lattice(n1, n2, m1, m2, co); }, false);
    }
    generate([&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_rect_sizes(n1, n2, nn, ne, nb, bl); },
             [&](double* x, double* y, int* c, int* b, int* bl) {
                 return cp_mesh_rect(x1, x2, y1, y2, n1, n2, x, y, c, b, bl);
             });
    setLattice(n1 + 1, n2 + 1);
}

void Mesh::genRectWithHole(double x1, double x2, double y1, double y2, double holeRad, size_t N1, size_t N2,
                           int order)
{
    check_order(order);
    const int n1 = as_int(N1, "genRectWithHole"), n2 = as_int(N2, "genRectWithHole");
    if (order == 2) {
        generate2(
            [&](int* nn, int* ne, int* nb, int* bl) {
                return cp_mesh_hole2_sizes(x1, x2, y1, y2, holeRad, n1, n2, nn, ne, nb, bl);
            },
            [&](double* x, double* y, int* c, unsigned char* sc, int* b, int* bl) {
                return cp_mesh_hole2(x1, x2, y1, y2, holeRad, n1, n2, x, y, c, sc, b, bl);
            });
        return setCornerLattice(
            [&](int* m1, int* m2, int* co) {
                return cp_mesh_hole2_lattice(x1, x2, y1, y2, holeRad, n1, n2, m1, m2, co);
            },
            true);
    }
    generate(
        [&](int* nn, int* ne, int* nb, int* bl) {
            return cp_mesh_hole_sizes(x1, x2, y1, y2, holeRad, n1, n2, nn, ne, nb, bl);
        },
        [&](double* x, double* y, int* c, int* b, int* bl) {
            return cp_mesh_hole(x1, x2, y1, y2, holeRad, n1, n2, x, y, c, b, bl);
        });
    setRingLattice(2 * (n1 + n2));
}

void Mesh::genRing(double a, double b, size_t N_phi, size_t N_r, int order)
{
    check_order(order);
    const int np = as_int(N_phi, "genRing"), nr = as_int(N_r, "genRing");
    if (order == 2) {
        generate2([&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_ring2_sizes(np, nr, nn, ne, nb, bl); },
                  [&](double* x, double* y, int* c, unsigned char* sc, int* bd, int* bl) {
                      return cp_mesh_ring2(a, b, np, nr, x, y, c, sc, bd, bl);
                  });
        return setCornerLattice([&](int* m1, int* m2, int* co) { return cp_mesh_ring2_lattice(np, nr, m1, m2, co); },
                                true);
    }
    generate([&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_ring_sizes(np, nr, nn, ne, nb, bl); },
             [&](double* x, double* y, int* c, int* bd, int* bl) {
                 return cp_mesh_ring(a, b, np, nr, x, y, c, bd, bl);
             });
    setRingLattice(np);
}

void Mesh::genQuarterPlate(double L, double r, size_t N_theta, size_t N_r)
{
    const int nt = as_int(N_theta, "genQuarterPlate"), nr = as_int(N_r, "genQuarterPlate");
    generate(
        [&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_quarter_plate_sizes(nt, nr, nn, ne, nb, bl); },
        [&](double* x, double* y, int* c, int* b, int* bl) {
            return cp_mesh_quarter_plate(L, r, nt, nr, x, y, c, b, bl);
        });
    setLattice(nt + 1, nr + 1);
}

void Mesh::genQuarterRing(double a, double b, size_t N_theta, size_t N_r, double ratio)
{
    const int nt = as_int(N_theta, "genQuarterRing"), nr = as_int(N_r, "genQuarterRing");
    generate(
        [&](int* nn, int* ne, int* nb, int* bl) { return cp_mesh_quarter_ring_sizes(nt, nr, nn, ne, nb, bl); },
        [&](double* x, double* y, int* c, int* bd, int* bl) {
            return cp_mesh_quarter_ring(a, b, nt, nr, ratio, x, y, c, bd, bl);
        });
    setLattice(nt + 1, nr + 1);
}

bool Mesh::loadFromFile(const std::string& fileName)
{
    if (fileName.size() < 4 || fileName.compare(fileName.size() - 4, 4, ".inp") != 0)
        throw "Unknown mesh file format";
    if (!std::ifstream(fileName)) throw "There is no such file";
    int nn = 0, n3 = 0, n4 = 0, n8 = 0, nb = 0;
    check(cp_mesh_inp_sizes(fileName.c_str(), &nn, &n3, &n4, &n8, &nb, nullptr, 0), "loadFromFile");
    std::vector<int> blen(nb);
    check(cp_mesh_inp_sizes(fileName.c_str(), &nn, &n3, &n4, &n8, &nb, blen.data(), nb), "loadFromFile");
    std::vector<double> x(nn), y(nn);
    std::vector<int> c3(3 * (size_t)n3), c4(4 * (size_t)n4), c8(8 * (size_t)n8);
    std::vector<unsigned char> sec(nn);
    size_t nbord = 0;
    for (int b : blen) nbord += (size_t)b;
    std::vector<int> bord(nbord);
    check(cp_mesh_load_inp(fileName.c_str(), x.data(), y.data(), c3.data(), c4.data(), c8.data(), sec.data(),
                           bord.data()),
          "loadFromFile");
    clear();
    nodeCount = nn;
    node = new vec2[nn];
    secOrdNodes = new bool[nn];
    for (int i = 0; i < nn; ++i) {
        node[i] = vec2(x[i], y[i]);
        secOrdNodes[i] = sec[i] != 0;
    }
    count3 = n3;
    count4 = n4;
    count8 = n8;
    elem3 = new int[c3.size()];
    elem4 = new int[c4.size()];
    elem8 = new int[c8.size()];
    std::copy(c3.begin(), c3.end(), elem3);
    std::copy(c4.begin(), c4.end(), elem4);
    std::copy(c8.begin(), c8.end(), elem8);
    elemPos[0] = n3;   // fillPos: triangles, then quads, then 8-node quads
    elemPos[1] = n3 + n4;
    elemPos[2] = n3 + n4 + n8;
    setBorders(nb, blen.data(), bord);
    analysed = false;
    return true;
}

int Mesh::findMaxIndexDiff() const
{
    int d = 0;
    if (count4) check(cp_mesh_max_index_diff(count4, elem4, &d), "findMaxIndexDiff");
    for (int e = 0; e < count8; ++e)   // Mesh.cu:433-439
        for (int i = 0; i < 8; ++i)
            for (int j = i + 1; j < 8; ++j) d = std::max(d, std::abs(elem8[8 * e + i] - elem8[8 * e + j]));
    return d;
}

double Mesh::elemSpace8(int e) const
{
    const int* c = elem8 + 8 * (size_t)e;
    double s1 = 0., s2 = 0.;
    for (int i = 0; i < 4; ++i) {   // corner i, mid-side i + 4, next corner
        const vec2& p = node[c[i]];
        const vec2& m = node[c[i + 4]];
        const vec2& q = node[c[(i + 1) & 3]];
        s1 += p.x * m.y + m.x * q.y;
        s2 += p.y * m.x + m.y * q.x;
    }
    return std::fabs(s1 - s2) * 0.5;
}

double Mesh::elemSpace4(int e) const
{
    const int* c = elem4 + 4 * (size_t)e;
    double s1 = 0., s2 = 0.;
    for (int i = 0; i < 4; ++i) {
        const vec2& p = node[c[i]];
        const vec2& q = node[c[(i + 1) & 3]];
        s1 += p.x * q.y;
        s2 += p.y * q.x;
    }
    return std::fabs(s1 - s2) * 0.5;
}

double Mesh::aspectRatio(int e) const
{
    const int* c = elem4 + 4 * (size_t)e;
    double lo = 0., hi = 0.;
    for (int i = 0; i < 4; ++i) {
        const double len = (node[c[(i + 1) & 3]] - node[c[i]]).norm();
        lo = i == 0 ? len : std::min(lo, len);
        hi = i == 0 ? len : std::max(hi, len);
    }
    return hi / lo;
}

double Mesh::skewAngleSin(int e) const
{
    const int* c = elem4 + 4 * (size_t)e;
    const vec2 v1 = (node[c[0]] + node[c[1]] - node[c[2]] - node[c[3]]) * 0.5;
    const vec2 v2 = (node[c[1]] + node[c[2]] - node[c[3]] - node[c[0]]) * 0.5;
    const double d = v1 * v2;
    return 1. - d * d / ((v1 * v1) * (v2 * v2));
}

double Mesh::avrElemSize() const
{
    if (count4 + count8 == 0) return 0.;
    double s = 0.;
    for (int e = 0; e < count4; ++e) s += elemSpace4(e);
    for (int e = 0; e < count8; ++e) s += elemSpace8(e);
    return s / (count4 + count8);
}

void Mesh::printAnalysis()
{
    // the quality measures are those of 4-node quads (Mesh.cu:1032-1080)
    if (count8) throw std::invalid_argument("Mesh::printAnalysis: provided for 4-node meshes");
    delete[] spaces;
    delete[] aspects;
    delete[] skewAngles;
    spaces = new double[count4];
    aspects = new double[count4];
    skewAngles = new double[count4];
    double sa = 0., smin = 1e300, smax = 0., aa = 0., amin = 1e300, amax = 0., ka = 0., kmin = 1e300, kmax = 0.;
    for (int e = 0; e < count4; ++e) {
        const double s = elemSpace4(e), a = aspectRatio(e), k = skewAngleSin(e);
        spaces[e] = s;
        aspects[e] = a;
        skewAngles[e] = std::acos(k);   // the file's Element_skew_angle column
        sa += s;
        aa += a;
        ka += k;
        smin = std::min(smin, s);
        smax = std::max(smax, s);
        amin = std::min(amin, a);
        amax = std::max(amax, a);
        kmin = std::min(kmin, k);
        kmax = std::max(kmax, k);
    }
    if (count4) {
        sa /= count4;
        aa /= count4;
        ka /= count4;
    }
    std::cout << "Mesh analysis:\nNodes max index difference in element: " << findMaxIndexDiff()
              << "\nSpace: avr = " << sa << ", min = " << smin << ", max = " << smax << "\nAspect ratio "
              << 100. / aa << "%: avr = " << aa << ", min = " << amin << ", max = " << amax
              << "\nSkew angle sin " << 100 * ka << "%: avr = " << ka << ", min = " << kmin << ", max = " << kmax
              << "\n"
              << std::endl;
    analysed = true;
}

// Legacy VTK polydata: points (6 significant digits), quads, element ids,
// the quality fields when analysed, node ids.
void Mesh::saveAsVtk(const std::string& fileName) const
{
    std::ofstream f(fileName);
    if (!f) throw std::runtime_error("Mesh::saveAsVtk: cannot open " + fileName);
    f << "# vtk DataFile Version 2.0\nMesh\nASCII\nDATASET POLYDATA\nPOINTS " << nodeCount << " float\n";
    for (int i = 0; i < nodeCount; ++i) f << node[i].x << " " << node[i].y << " " << 0 << "\n";
    f << "POLYGONS " << elemCount() << " " << 5 * count4 + 9 * count8;
    for (int e = 0; e < count4; ++e) {
        f << "\n4 ";
        for (int j = 0; j < 4; ++j) f << elem4[4 * e + j] << " ";
    }
    for (int e = 0; e < count8; ++e) {   // corner, mid-side, corner, ... (Mesh.cu:959-963)
        f << "\n8 ";
        for (int j = 0; j < 4; ++j) f << elem8[8 * e + j] << " " << elem8[8 * e + j + 4] << " ";
    }
    f << "\nCELL_DATA " << elemCount() << "\nSCALARS ElementID int 1\nLOOKUP_TABLE default\n";
    for (int e = 0; e < elemCount(); ++e) f << e << "\n";
    if (analysed) {
        const double* cols[3] = {spaces, aspects, skewAngles};
        const char* names[3] = {"Element_space", "Element_aspect_ratio", "Element_skew_angle"};
        f << "FIELD FieldData 3\n";
        for (int c = 0; c < 3; ++c) {
            if (c) f << "\n";
            f << names[c] << " 1 " << elemCount() << " float\n";
            for (int e = 0; e < elemCount(); ++e) f << cols[c][e] << " ";
        }
    }
    f << "\nPOINT_DATA " << nodeCount << "\nSCALARS NodeID int 1\nLOOKUP_TABLE my_table";
    for (int i = 0; i < nodeCount; ++i) f << "\n" << i;
}

namespace {

// Flat copies of the mesh arrays the C-ABI renumberings take.
struct FlatMesh {
    std::vector<double> x, y;
    std::vector<int> bord;
    void load(const Mesh& m)
    {
        x.resize(m.nodeCount);
        y.resize(m.nodeCount);
        for (int i = 0; i < m.nodeCount; ++i) {
            x[i] = m.node[i].x;
            y[i] = m.node[i].y;
        }
        bord.clear();
        for (int b = 0; b < m.bordersCount; ++b) bord.insert(bord.end(), m.borders[b], m.borders[b] + m.borderLength[b]);
    }
    void store(Mesh& m) const
    {
        for (int i = 0; i < m.nodeCount; ++i) m.node[i] = vec2(x[i], y[i]);
        size_t off = 0;
        for (int b = 0; b < m.bordersCount; ++b) {
            std::copy(bord.begin() + off, bord.begin() + off + m.borderLength[b], m.borders[b]);
            off += m.borderLength[b];
        }
    }
};

}  // namespace

void Mesh::renumerateRing(int borderN)
{
    if (count8 > 0 || count3 > 0) throw std::invalid_argument("renumerateRing: 4-node meshes only");
    FlatMesh f;
    f.load(*this);
    std::vector<int> perm(nodeCount);
    check(cp_mesh_renumerate_ring(nodeCount, count4, f.x.data(), f.y.data(), elem4, bordersCount,
                                  f.bord.data(), borderLength, borderN, perm.data()),
          "renumerateRing");
    f.store(*this);
    renumberLattice(perm);
}

void Mesh::renumByDirection(vec2 direction)
{
    if (count3 > 0) throw std::invalid_argument("renumByDirection: triangle meshes are not supported");
    FlatMesh f;
    f.load(*this);
    const bool q8 = count8 > 0;
    std::vector<int> perm(nodeCount);
    check(cp_mesh_renum_by_direction(nodeCount, f.x.data(), f.y.data(), q8 ? count8 : count4, q8 ? 8 : 4,
                                     q8 ? elem8 : elem4, (int)f.bord.size(), f.bord.data(), direction.x,
                                     direction.y, perm.data()),
          "renumByDirection");
    f.store(*this);
    renumberLattice(perm);
    // mid-side nodes follow their elements (Mesh.cu:619-623)
    if (secOrdNodes) {
        std::fill(secOrdNodes, secOrdNodes + nodeCount, false);
        for (int e = 0; e < count8; ++e)
            for (int i = 4; i < 8; ++i) secOrdNodes[elem8[8 * e + i]] = true;
    }
}

void Mesh::print() const
{
    std::cout << "Nodes: " << nodeCount << "\n";
    for (int i = 0; i < nodeCount; ++i) std::cout << i << ": " << node[i].x << " " << node[i].y << "\n";
    std::cout << "Quads: " << count4 << "\n";
    for (int e = 0; e < count4; ++e)
        std::cout << e << ": " << elem4[4 * e] << " " << elem4[4 * e + 1] << " " << elem4[4 * e + 2] << " "
                  << elem4[4 * e + 3] << "\n";
    if (count8) std::cout << "8-node quads: " << count8 << "\n";
    for (int e = 0; e < count8; ++e) {
        std::cout << e << ":";
        for (int j = 0; j < 8; ++j) std::cout << " " << elem8[8 * e + j];
        std::cout << "\n";
    }
    for (int b = 0; b < bordersCount; ++b) {
        std::cout << "Border " << b << ":";
        for (int i = 0; i < borderLength[b]; ++i) std::cout << " " << borders[b][i];
        std::cout << "\n";
    }
}
