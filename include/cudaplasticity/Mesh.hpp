// Mesh.hpp -- mesh of the drop-in C++ API (Mesh.cuh:21-244).
//
// Same public members (node, elem4, borders, counts, flags) and generator
// signatures as the reference; the generators run in libcp_hip (host code)
// and produce the reference's node numbering and element order.  Meshes are
// host data here: the solver copies them to the device when it is built, so
// meshToGPU / meshToRAM are no-ops kept for source compatibility.  order = 2
// builds 8-node quads (elem8, count8) as the reference's remapOrder does;
// a mesh holds one element type.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "vec2.hpp"

class Mesh {
public:
    bool useCuda = false;   // a HIP device was found at construction
    bool ramSaved = true;   // host arrays are current (always, here)
    bool analysed = false;

    int count3 = 0;
    int count4 = 0;
    int count8 = 0;

    int* elem3 = nullptr;
    int* elem4 = nullptr;   // 4 node indices per quad, reference local order
    int* elem8 = nullptr;   // 8 node indices: corners, then mid-sides of edges 1-2, 2-3, 3-4, 4-1

    int elemPos[3] = {};    // first index of triangles / quads / 8-node elements

    int nodeCount = 0;
    vec2* node = nullptr;
    bool* secOrdNodes = nullptr;   // true on the mid-side nodes of 8-node quads

    int bordersCount = 0;
    int* borderLength = nullptr;
    int** borders = nullptr;

    // New: the logical lattice of a generated mesh, the input of the
    // multigrid preconditioner (SolverOptions::multigrid): latticeN1 x
    // latticeN2 points, latticeNodes[i1 * latticeN2 + i2] the node at
    // (i1, i2) (8-node meshes: the lattice of their corner nodes); direction
    // 1 closes on itself (latticePeriodic) around the hole of
    // genRectWithHole and around genRing.  Renumberings carry it along;
    // latticeN1 = 0 for meshes read from files.
    int latticeN1 = 0;
    int latticeN2 = 0;
    bool latticePeriodic = false;
    std::vector<int> latticeNodes;

    Mesh();
    Mesh(const Mesh& other);
    Mesh& operator=(const Mesh& other);
    ~Mesh();

    int elemCount() const { return elemPos[2]; }

    // Mesh.cu:17-109
    void genRectangle(double x1, double x2, double y1, double y2, size_t N1, size_t N2, int order = 1);
    // Mesh.cu:112-245 (+ renumerateRing)
    void genRectWithHole(double x1, double x2, double y1, double y2, double holeRad, size_t N1, size_t N2,
                         int order = 1);
    // Mesh.cu:249-329 (+ renumerateRing)
    void genRing(double a, double b, size_t N_phi, size_t N_r, int order = 1);
    // New: quarter plate [0,L]^2 minus a quarter disc of radius r at the origin.
    // Borders: 0 y=0, 1 x=L, 2 y=L, 3 x=0, 4 hole arc.
    void genQuarterPlate(double L, double r, size_t N_theta, size_t N_r);
    // New: quarter ring a <= rho <= b, radial sizes growing by `ratio` from the
    // inner radius.  Borders: 0 inner arc, 1 theta=pi/2, 2 outer arc, 3 theta=0.
    void genQuarterRing(double a, double b, size_t N_theta, size_t N_r, double ratio);

    // Mesh.cu:725-937: *NODE, *ELEMENT TYPE=S3/S4/S8R and *MPC (borders)
    // blocks of an .inp file.  As the reference does, throws a const char*
    // ("Unknown mesh file format", "There is no such file").
    bool loadFromFile(const std::string& fileName);

    // Mesh.cu:332-384: renumbers a 4-node mesh of nodeCount / borderN rows of
    // borderN columns column by column, alternately from both ends
    // (genRectWithHole and genRing already do).
    void renumerateRing(int borderN);
    // Mesh.cu:537-630: renumbers the nodes by increasing projection onto
    // `direction` (ties in decreasing original number, as the reference's
    // merge sort leaves them); elements, borders and secOrdNodes follow.
    void renumByDirection(vec2 direction = {1., 1.});

    // Mesh.cu:417-441
    int findMaxIndexDiff() const;

    // Mesh.cuh:258-267 (shoelace area of quad e)
    double elemSpace4(int e) const;
    // Mesh.cuh:269-278 (shoelace area of the 8-node quad's octagon)
    double elemSpace8(int e) const;
    // Mesh.cuh:281-293
    double aspectRatio(int e) const;
    // Mesh.cuh:296-305
    double skewAngleSin(int e) const;
    // mean element area (Mesh.cuh:308-318)
    double avrElemSize() const;

    // Element quality arrays (area, aspect ratio, skew sine per quad) filled
    // by printAnalysis (Mesh.cu:1032-1080), which also prints their ranges
    // and sets `analysed`.
    double* spaces = nullptr;
    double* aspects = nullptr;
    double* skewAngles = nullptr;
    void printAnalysis();

    // Legacy-VTK polydata of the mesh, the layout of Mesh.cu:939-989.
    void saveAsVtk(const std::string& fileName) const;

    void meshToRAM() {}
    void meshToGPU() {}
    void print() const;

private:
    void clear();
    template <typename Sizes, typename Gen>
    void generate(Sizes sizes, Gen gen);
    template <typename Sizes, typename Gen>
    void generate2(Sizes sizes, Gen gen);
    void setBorders(int nb, const int* blen, const std::vector<int>& bord);
    void setLattice(int n1, int n2);
    template <typename Lat>
    void setCornerLattice(Lat lat, bool periodic);
    void setRingLattice(int borderN);
    void renumberLattice(const std::vector<int>& newOfOld);
};
