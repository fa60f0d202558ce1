// api_test.cpp -- drives the drop-in C++ API (include/cudaplasticity/) the
// way the reference's experiments drive its classes, and writes raw arrays
// into OUTDIR for tests/test_cpp_api.py to check.
//
//   api_test cpu OUTDIR [INDIR]  meshes, material laws, load conditions and
//                                (with INDIR/ring_points.f64, ring_polys.i32) the
//                                mesh analysis and VTK writer; no device needed
//   api_test gpu OUTDIR          solves through PlasticitySolver on the GPU
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cudaplasticity/cudaplasticity.hpp"

namespace {

std::string outdir, indir;

template <typename T>
void dump(const std::string& name, const T* data, size_t n)
{
    std::ofstream f(outdir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char*>(data), sizeof(T) * n);
}

template <typename T>
void dump(const std::string& name, const std::vector<T>& v)
{
    dump(name, v.data(), v.size());
}

void dump_mesh(const std::string& tag, const Mesh& m)
{
    std::vector<double> xy(2 * (size_t)m.nodeCount);
    for (int i = 0; i < m.nodeCount; ++i) {
        xy[2 * i] = m.node[i].x;
        xy[2 * i + 1] = m.node[i].y;
    }
    dump(tag + "_xy.f64", xy);
    dump(tag + "_elem4.i32", m.elem4, 4 * (size_t)m.count4);
    std::vector<double> q(3 * (size_t)m.count4);
    for (int e = 0; e < m.count4; ++e) {
        q[3 * e] = m.elemSpace4(e);
        q[3 * e + 1] = m.aspectRatio(e);
        q[3 * e + 2] = m.skewAngleSin(e);
    }
    dump(tag + "_quality.f64", q);
    std::vector<int> meta = {m.nodeCount, m.count4, m.bordersCount, m.findMaxIndexDiff()};
    dump(tag + "_meta.i32", meta);
}

void dump_lattice(const std::string& tag, const Mesh& m)
{
    std::vector<int> meta = {m.latticeN1, m.latticeN2, m.latticePeriodic ? 1 : 0};
    dump(tag + "_lattice_meta.i32", meta);
    dump(tag + "_lattice.i32", m.latticeNodes);
}

void dump_conditions(const std::string& tag, const PlasticitySolver& ps)
{
    std::vector<double> rp, u0;
    std::vector<unsigned char> mask;
    ps.initConditions(rp, u0, mask);
    dump(tag + "_rp.f64", rp);
    dump(tag + "_u0.f64", u0);
    dump(tag + "_mask.u8", mask);
}

// plasticityTest (Main.cpp:252-296): 304L steel plate with a hole.
const double E_304L = 193e9, S_T_304L = 170e6, NU_304L = 0.29;

void hole_plate(Mesh& mesh, LoadConditions& lc, Material& m)
{
    mesh.genRectWithHole(-3., 3., -2., 2., 0.5, 48, 32, 1);
    const double S = 170;
    lc.setForce(1, vec2(S, 0.));
    lc.setForce(3, vec2(-S, 0.));
    lc.fixVertAxis(0.);
    lc.fixHorAxis(0.);
    m.h = 1.;
    m.setPowerPlast(E_304L * 1e-6, S_T_304L * 1e-6, E_304L * 0.01 * 1e-6, 0.015);
    m.nu = NU_304L;
}

int run_cpu()
{
    {
        Mesh mesh;
        LoadConditions lc;
        Material m;
        hole_plate(mesh, lc, m);
        dump_mesh("hole", mesh);
        PlasticitySolver ps(m, mesh, lc);
        dump_conditions("hole", ps);
        // the solver deep-copies the mesh
        Mesh copy(mesh);
        dump_mesh("holecopy", copy);
    }
    {
        // contactTest's renumbering (Main.cpp:1029: mesh.renumByDirection()
        // with the default direction {1, 1}) on the fixture plate, and
        // renumerateRing applied by hand to a rectangle
        Mesh mesh;
        mesh.genRectWithHole(-3., 3., -2., 2., 0.5, 12, 8, 1);
        mesh.renumByDirection();
        dump_mesh("renum", mesh);
        Mesh rect;
        rect.genRectangle(0., 1., 0., 1., 7, 5);
        rect.renumerateRing(8);
        dump_mesh("ringrenum", rect);
        // vec2::crossZ (vec2.cuh:81-85)
        vec2 a(3., -2.), b(0.5, 4.);
        std::vector<double> cz = {a.crossZ(b), b.crossZ(a), a.crossZ(a)};
        dump("crossz.f64", cz);
        dump_lattice("renum", mesh);
        dump_lattice("ringrenum", rect);
    }
    {
        // the lattices the generators record (multigrid's input)
        Mesh mesh;
        mesh.genRectWithHole(-3., 3., -2., 2., 0.5, 48, 32, 1);
        dump_mesh("lat_hole", mesh);
        dump_lattice("lat_hole", Mesh(mesh));
        mesh.genRing(0.1, 0.2, 40, 5);
        dump_mesh("lat_ring", mesh);
        dump_lattice("lat_ring", mesh);
        mesh.genRectangle(0., 2., 0., 1., 9, 4);
        dump_mesh("lat_rect", mesh);
        dump_lattice("lat_rect", mesh);
        mesh.genQuarterPlate(1.0, 0.1, 16, 7);
        dump_mesh("lat_qplate", mesh);
        dump_lattice("lat_qplate", mesh);
        mesh.genQuarterRing(0.5, 1.0, 12, 6, 1.1);
        dump_mesh("lat_qring", mesh);
        dump_lattice("lat_qring", mesh);
        mesh.genRectWithHole(-3., 3., -2., 2., 0.5, 6, 4, 2);
        dump_lattice("lat_hole2", mesh);
    }
    {
        // testRing (Main.cpp:75-106): position-dependent border force, h = 0.1
        Mesh mesh;
        mesh.genRing(0.1, 0.2, 40, 5);
        LoadConditions lc;
        lc.setForce(0, [](vec2 r) { return vec2(1e6 * std::fabs(r.x), 0); });
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.E = 1.1e10;
        m.nu = 0.3;
        m.h = 0.1;
        dump_mesh("ring", mesh);
        PlasticitySolver ps(m, mesh, lc);
        dump_conditions("ring", ps);
    }
    {
        // diagramTest (Main.cpp:929-994) level: gravity and an end displacement
        Mesh mesh;
        mesh.genRectangle(-1, 1, -1, 1, 2, 2);
        LoadConditions lc;
        lc.setGravity();
        lc.setDisplacement(2, vec2(0, 0), false, true);
        lc.setDisplacement(0, vec2(0.0015, 0), false, true);
        Material m;
        m.setLinearPlast(E_304L, S_T_304L, 0.1 * E_304L);
        m.nu = NU_304L;
        dump_mesh("diag", mesh);
        PlasticitySolver ps(m, mesh, lc);
        dump_conditions("diag", ps);
    }
    {
        // graded quarter ring, normal-oriented internal pressure
        Mesh mesh;
        mesh.genQuarterRing(0.5, 1.0, 48, 12, 1.05);
        LoadConditions lc;
        lc.setForce(0, vec2(-180., 0.), true);
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.setSwiftPlast(2.1e5, 240., 0.2);
        dump_mesh("qring", mesh);
        PlasticitySolver ps(m, mesh, lc);
        dump_conditions("qring", ps);
    }
    {
        // point conditions, fixed border, rotation body force, displacement function
        Mesh mesh;
        mesh.genRectangle(0., 1., 0., 1., 4, 4);
        LoadConditions lc;
        lc.setForcePoint(vec2(1., 1.), vec2(3., 4.));
        lc.setDisplPoint(vec2(0., 0.), vec2(0.1, 0.2));
        lc.fixPoint(vec2(1., 0.));
        lc.fixBorderVert(3);
        lc.setRotation(2.0, vec2(0.5, 0.));
        lc.setDisplacement(2, [](vec2 r) { return vec2(0.01 * r.x, -0.02 * r.x * r.x); });
        Material m;
        m.rho = 3.;
        m.h = 0.5;
        dump_mesh("pts", mesh);
        PlasticitySolver ps(m, mesh, lc);
        dump_conditions("pts", ps);
    }
    {
        // material laws on the host (the device's own law code)
        Material lin, pw, sw;
        lin.setLinearPlast(2.1e5, 240., 2.1e5 / 101);
        pw.setPowerPlast(193e3, 170., 1930., 0.015);
        sw.setSwiftPlast(2.1e5, 240., 0.2);
        std::vector<double> out;
        for (double eps : {1e-4, 1e-3, 2e-3, 5e-3, 2e-2, 0.1}) {
            out.push_back(lin.f(eps));
            out.push_back(pw.f(eps));
            out.push_back(sw.f(eps));
        }
        dump("laws.f64", out);
    }
    {
        // order-2 meshes of the reference's mesh files (Main.cpp:1173-1190):
        // generators, the mesh VTK writer, the 8-node conditions
        Mesh rect, hole, ring;
        rect.genRectangle(-2, 2, -1, 1, 12, 6, 2);
        hole.genRectWithHole(-3, 3, -2, 2, 0.5, 12, 8, 2);
        ring.genRing(1, 2, 40, 6, 2);
        rect.saveAsVtk(outdir + "/rect1.vtk");
        hole.saveAsVtk(outdir + "/hole1.vtk");
        ring.saveAsVtk(outdir + "/ring1.vtk");
        std::vector<unsigned char> sec(hole.nodeCount);
        for (int i = 0; i < hole.nodeCount; ++i) sec[i] = hole.secOrdNodes[i] ? 1 : 0;
        dump("hole1_sec.u8", sec);
        std::vector<double> q8 = {hole.avrElemSize(), (double)hole.findMaxIndexDiff(), hole.elemSpace8(0)};
        dump("hole1_q8.f64", q8);
        LoadConditions lc;
        lc.setForce(1, vec2(170., 3.));
        lc.setForce(3, vec2(-170., 0.));
        lc.setForce(4, vec2(5., 1.), true);
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.h = 0.7;
        PlasticitySolver ps(m, hole, lc);
        dump_conditions("hole1", ps);
        Mesh copy(hole);
        std::vector<int> c8(copy.elem8, copy.elem8 + 8 * (size_t)copy.count8);
        dump("hole1copy_elem8.i32", c8);
    }
    {
        // the CPU solver and the 4-node mesh analysis of an 8-node mesh are refused, loudly
        int refused = 0;
        Mesh mesh;
        mesh.genRectangle(0, 1, 0, 1, 2, 2, 2);
        try {
            mesh.printAnalysis();
        } catch (const std::invalid_argument&) {
            ++refused;
        }
        {
            // multigrid needs a lattice: none on a mesh that was not
            // generated (one read from a file has latticeN1 = 0)
            Mesh plain(mesh);
            plain.latticeN1 = 0;
            PlasticitySolver ps(Material(), plain, LoadConditions());
            ps.options.multigrid = true;
            try {
                ps.solve();
            } catch (const std::invalid_argument&) {
                ++refused;
            }
        }
        mesh.genRectangle(0, 1, 0, 1, 2, 2);
        PlasticitySolver ps(Material(), mesh, LoadConditions());
        try {
            ps.solveCPU();
        } catch (const std::runtime_error&) {
            ++refused;
        }
        std::vector<int> r = {refused};
        dump("refused.i32", r);
    }
    if (!indir.empty()) {
        // Mesh::loadFromFile on an .inp deck of 8-node quads; the reference's
        // errors are thrown as const char*
        Mesh mesh;
        mesh.loadFromFile(indir + "/plate8.inp");
        std::vector<double> xy(2 * (size_t)mesh.nodeCount);
        std::vector<unsigned char> sec(mesh.nodeCount);
        for (int i = 0; i < mesh.nodeCount; ++i) {
            xy[2 * i] = mesh.node[i].x;
            xy[2 * i + 1] = mesh.node[i].y;
            sec[i] = mesh.secOrdNodes[i] ? 1 : 0;
        }
        dump("inp_xy.f64", xy);
        dump("inp_sec.u8", sec);
        dump("inp_elem8.i32", mesh.elem8, 8 * (size_t)mesh.count8);
        std::vector<int> bl(mesh.borderLength, mesh.borderLength + mesh.bordersCount);
        dump("inp_blen.i32", bl);
        std::vector<int> b1(mesh.borders[1], mesh.borders[1] + mesh.borderLength[1]);
        dump("inp_border1.i32", b1);
        std::vector<int> counts = {mesh.count3, mesh.count4, mesh.count8, mesh.elemCount()};
        int thrown = 0;
        for (const char* bad : {"mesh.msh", "no_such_mesh.inp"}) {
            try {
                Mesh other;
                other.loadFromFile(indir + "/" + bad);
            } catch (const char* msg) {
                ++thrown;
                std::cout << "loadFromFile(" << bad << "): " << msg << "\n";
            }
        }
        counts.push_back(thrown);
        dump("inp_counts.i32", counts);
    }
    if (!indir.empty()) {
        // a mesh given through the public arrays (the reference's data/ring.vtk
        // points and quads): printAnalysis and the mesh VTK writer
        std::ifstream fp(indir + "/ring_points.f64", std::ios::binary), fq(indir + "/ring_polys.i32", std::ios::binary);
        fp.seekg(0, std::ios::end);
        const size_t np = (size_t)fp.tellg() / (2 * sizeof(double));
        fp.seekg(0);
        fq.seekg(0, std::ios::end);
        const size_t nq = (size_t)fq.tellg() / (4 * sizeof(int));
        fq.seekg(0);
        Mesh mesh;
        mesh.nodeCount = (int)np;
        mesh.node = new vec2[np];
        for (size_t i = 0; i < np; ++i) {
            double p[2];
            fp.read(reinterpret_cast<char*>(p), sizeof(p));
            mesh.node[i] = vec2(p[0], p[1]);
        }
        mesh.count4 = (int)nq;
        mesh.elem4 = new int[4 * nq];
        fq.read(reinterpret_cast<char*>(mesh.elem4), 4 * nq * sizeof(int));
        mesh.elemPos[1] = mesh.elemPos[2] = (int)nq;
        mesh.printAnalysis();
        mesh.saveAsVtk(outdir + "/ring_mesh.vtk");
    }
    std::cout << "cpu cases written to " << outdir << "\n";
    return 0;
}

int run_gpu()
{
    {
        // the reference's plasticityTest through the drop-in API, reference defaults
        Mesh mesh;
        LoadConditions lc;
        Material m;
        hole_plate(mesh, lc, m);
        PlasticitySolver ps(m, mesh, lc);
        const double t = ps.solve();
        const cp_stats& st = ps.lastStats();
        dump("plast_u.f64", ps.displacements());
        std::vector<double> F;
        for (int f = 0; f < CP_NFIELDS; ++f) F.insert(F.end(), ps.field(f).begin(), ps.field(f).end());
        dump("plast_fields.f64", F);
        std::vector<double> s = {(double)st.outer_iters, st.rel_err, (double)st.status, ps.deformEnergy(),
                                 ps.getSxx0(), ps.getExx0(), t};
        dump("plast_stats.f64", s);
        ps.saveDisplsToFile(outdir + "/plast_displs.txt");
        ps.saveAsVtk(outdir + "/holePlatePlastPow.vtk");
        ps.polarCoord = true;
        ps.saveAsVtk(outdir + "/holePlatePlastPow_polar.vtk");
    }
    {
        // the same with the multigrid preconditioner on the plate's periodic
        // lattice, then on the plate renumbered by direction (contactTest's
        // numbering), FP64 and mixed
        for (int k = 0; k < 3; ++k) {
            Mesh mesh;
            LoadConditions lc;
            Material m;
            hole_plate(mesh, lc, m);
            if (k > 0) mesh.renumByDirection();
            PlasticitySolver ps(m, mesh, lc);
            ps.options.multigrid = true;
            ps.options.mixedPrecision = k == 2;
            ps.iterOutput = false;
            const double t = ps.solve();
            const cp_stats& st = ps.lastStats();
            const std::string tag = "plastmg" + std::to_string(k);
            dump(tag + "_u.f64", ps.displacements());
            std::vector<double> F;
            for (int f = 0; f < CP_NFIELDS; ++f) F.insert(F.end(), ps.field(f).begin(), ps.field(f).end());
            dump(tag + "_fields.f64", F);
            std::vector<double> s = {(double)st.outer_iters, st.rel_err, (double)st.status, (double)st.cg_iters, t};
            dump(tag + "_stats.f64", s);
        }
    }
    {
        // solverTest (Main.cpp:14-38): elastic 4x4 block, end traction, mid fixings
        Mesh mesh;
        mesh.genRectangle(-1., 1., -1., 1., 4, 4);
        LoadConditions lc;
        lc.setDisplacement(2, vec2(0, 0.), false, true);
        lc.setForce(0, vec2(1, 0));
        Material m;
        m.setLinearPlast(E_304L * 1e-6, S_T_304L, E_304L * 1e-6);
        m.E = E_304L * 1e-6;
        m.nu = NU_304L * 1e-6;
        m.h = 1.;
        PlasticitySolver ps(m, mesh, lc);
        ps.solveElast();
        ps.saveAsVtk(outdir + "/solverTest.vtk");
    }
    {
        // config-1 style elastic solve with the Kirsch check of Main.cpp:65-68
        Mesh mesh;
        mesh.genQuarterPlate(1.0, 0.1, 100, 49);
        LoadConditions lc;
        lc.setForce(1, vec2(100., 0.));
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.E = 2.1e5;
        m.nu = 0.3;
        PlasticitySolver ps(m, mesh, lc);
        ps.options.cgTol = 1e-10;
        ps.iterOutput = false;
        ps.solveElast();
        dump("elast_u.f64", ps.displacements());
        dump("elast_sxx.f64", ps.field(CP_F_SXX));
        const double a = 0.1, S = 100.;
        auto kirsch_r = [&](vec2 p) {
            const double r = p.x, th = p.y, q = a * a / (r * r);
            return 0.5 * S * (1 - q) + 0.5 * S * (1 - 4 * q + 3 * q * q) * std::cos(2 * th);
        };
        auto kirsch_t = [&](vec2 p) {
            const double r = p.x, th = p.y, q = a * a / (r * r);
            return 0.5 * S * (1 + q) - 0.5 * S * (1 + 3 * q * q) * std::cos(2 * th);
        };
        auto kirsch_rt = [&](vec2 p) {
            const double r = p.x, th = p.y, q = a * a / (r * r);
            return -0.5 * S * (1 + 2 * q - 3 * q * q) * std::sin(2 * th);
        };
        std::vector<double> e = {ps.polarSigmaError(kirsch_r, kirsch_t, kirsch_rt, 0.3)};
        dump("elast_polar_err.f64", e);
        // the two error measures with fields of every component, for the
        // restatement of their definitions in the test
        dump("elast_syy.f64", ps.field(CP_F_SYY));
        dump("elast_tau.f64", ps.field(CP_F_TAU));
        auto fx = [](vec2 p) { return 100. + 40. * p.x - 7. * p.y; };
        auto fy = [](vec2 p) { return 30. + 12. * p.y * p.y; };
        auto fxy = [](vec2 p) { return -5. + 9. * p.x * p.y; };
        std::vector<double> e2 = {ps.sigmaError(fx, fy, fxy), ps.polarSigmaError(fx, fy, fxy, 0.6)};
        dump("elast_sigma_errors.f64", e2);
    }
    {
        // additive API: load steps, mixed precision, block-Jacobi, predictor, forcing
        Mesh mesh;
        mesh.genQuarterPlate(1.0, 0.1, 40, 20);
        LoadConditions lc;
        lc.setForce(1, vec2(300., 0.));
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.setLinearPlast(2.1e5, 240., 2.1e5 / 101);
        PlasticitySolver ps(m, mesh, lc);
        ps.iterOutput = false;
        ps.options.cgTol = 1e-9;
        ps.options.tol = 1e-6;
        ps.options.aitken = true;
        ps.options.mixedPrecision = true;
        ps.options.blockJacobi = true;
        ps.options.predictor = 1;
        ps.options.cgForcing = 0.1;
        ps.options.innerTol = 1e-4;
        ps.solveSteps(4, 1.0);
        dump("steps_u.f64", ps.displacements());
        dump("steps_sigi.f64", ps.field(CP_F_SIGI));
    }
    {
        // energyConvTest (Main.cpp:105-178) on the 660-element 8-node mesh,
        // power law, the reference's defaults: data/aitken__plast2.txt
        Mesh mesh;
        mesh.genRectWithHole(-3., 3., -2., 2., 0.5, 18, 12, 2);
        LoadConditions lc;
        lc.setForce(1, vec2(170e6, 0));
        lc.setForce(3, vec2(-170e6, 0));
        lc.fixVertAxis(0.);
        lc.fixHorAxis(0.);
        Material m;
        m.setPowerPlast(E_304L, S_T_304L, E_304L * 0.01, 0.015);
        m.nu = NU_304L;
        m.h = 1.;
        PlasticitySolver ps(m, mesh, lc);
        ps.iterOutput = false;
        ps.solve();
        std::vector<double> e = {ps.deformEnergy(), (double)mesh.elemCount(), (double)ps.lastStats().status};
        dump("q8_energy.f64", e);
        ps.saveAsVtk(outdir + "/q8_plate.vtk");
        // the same with multigrid on the mesh's corner lattice
        PlasticitySolver pm(m, mesh, lc);
        pm.iterOutput = false;
        pm.options.multigrid = true;
        pm.solve();
        std::vector<double> em = {pm.deformEnergy(), (double)pm.lastStats().status, (double)pm.lastStats().cg_iters,
                                  (double)ps.lastStats().cg_iters};
        dump("q8mg_energy.f64", em);
    }
    {
        // fixBorderVert / fixBorderHor through solveElastCUDA: a rectangle in
        // uniaxial tension with the symmetry borders has the exact uniform
        // state u = (S x / E, -nu S y / E), which bilinear quads reproduce
        Mesh mesh;
        mesh.genRectangle(0., 2., 0., 1., 8, 4);
        auto find_border = [&](auto on) {
            for (int b = 0; b < mesh.bordersCount; ++b) {
                bool all = true;
                for (int i = 0; i < mesh.borderLength[b]; ++i) all = all && on(mesh.node[mesh.borders[b][i]]);
                if (all) return b;
            }
            return -1;
        };
        const int right = find_border([](vec2 p) { return p.x == 2.; });
        const int left = find_border([](vec2 p) { return p.x == 0.; });
        const int bottom = find_border([](vec2 p) { return p.y == 0.; });
        LoadConditions lc;
        lc.setForce(right, vec2(50., 0.));
        lc.fixBorderVert(left);     // u_x = 0 on x = 0
        lc.fixBorderHor(bottom);    // u_y = 0 on y = 0
        Material m;
        m.E = 2.1e5;
        m.nu = 0.3;
        PlasticitySolver ps(m, mesh, lc);
        ps.options.cgTol = 1e-12;
        ps.iterOutput = false;
        ps.solveElastCUDA();
        std::vector<double> out = {(double)right, (double)left, (double)bottom};
        const std::vector<double>& u = ps.displacements();
        for (int i = 0; i < mesh.nodeCount; ++i) {
            out.push_back(mesh.node[i].x);
            out.push_back(mesh.node[i].y);
            out.push_back(u[2 * (size_t)i]);
            out.push_back(u[2 * (size_t)i + 1]);
        }
        dump("patch_u.f64", out);
        // the same in plane strain (setPlainCondition)
        PlasticitySolver pe(m, mesh, lc);
        pe.setPlainCondition(strain);
        pe.options.cgTol = 1e-12;
        pe.iterOutput = false;
        pe.solveElastCUDA();
        dump("patch_u_strain.f64", pe.displacements());
    }
    {
        // fixBorder + setVolumeForce of a constant field equals setGravity
        // (the same body-force load), bit for bit, through solveCUDA; and
        // solveCUDA_FD (FP32 CG with FP64 refinement) agrees with it
        std::vector<double> res[3];
        for (int variant = 0; variant < 3; ++variant) {
            Mesh mesh;
            mesh.genRectangle(0., 1., 0., 4., 3, 8);
            int top = -1;
            for (int b = 0; b < mesh.bordersCount && top < 0; ++b) {
                bool all = true;
                for (int i = 0; i < mesh.borderLength[b]; ++i) all = all && mesh.node[mesh.borders[b][i]].y == 4.;
                if (all) top = b;
            }
            LoadConditions lc;
            lc.fixBorder(top);
            if (variant == 0) lc.setGravity(9.81);
            else lc.setVolumeForce([](vec2) { return vec2(0., -9.81); });
            Material m;
            m.setLinearPlast(2.1e5, 10., 2.1e5 / 101);   // the column's own weight (up to 39) exceeds it
            PlasticitySolver ps(m, mesh, lc);
            ps.iterOutput = false;
            ps.options.cgTol = 1e-10;
            ps.options.tol = 1e-8;
            if (variant < 2) ps.solveCUDA();
            else ps.solveCUDA_FD();
            res[variant] = ps.displacements();
            const std::vector<double>& si = ps.field(CP_F_SIGI);
            res[variant].insert(res[variant].end(), si.begin(), si.end());
        }
        dump("vol_gravity.f64", res[0]);
        dump("vol_volume.f64", res[1]);
        dump("vol_fd.f64", res[2]);
    }
    std::cout << "gpu cases written to " << outdir << "\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 4) {
        std::cerr << "usage: api_test cpu|gpu OUTDIR [INDIR]\n";
        return 2;
    }
    outdir = argv[2];
    if (argc == 4) indir = argv[3];
    const std::string mode = argv[1];
    try {
        if (mode == "cpu") return run_cpu();
        if (mode == "gpu") return run_gpu();
    } catch (const std::exception& e) {
        std::cerr << "api_test: " << e.what() << "\n";
        return 1;
    }
    std::cerr << "unknown mode " << mode << "\n";
    return 2;
}
