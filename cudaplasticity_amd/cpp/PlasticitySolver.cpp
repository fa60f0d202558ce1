// PlasticitySolver.cpp -- the drop-in PlasticitySolver class over the C-ABI.
#include "cudaplasticity/PlasticitySolver.hpp"

#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

namespace {

void check(int status, const char* what)
{
    if (status != CP_OK) throw std::runtime_error(std::string(what) + ": " + cp_last_error());
}

// One device context per process (device 0), created on first use.
cp_ctx* shared_context()
{
    static std::once_flag once;
    static std::unique_ptr<cp_ctx, int (*)(cp_ctx*)> ctx(nullptr, cp_ctx_destroy);
    static int status = CP_OK;
    std::call_once(once, [] {
        cp_ctx* c = nullptr;
        status = cp_ctx_create(0, &c);
        ctx.reset(c);
    });
    check(status, "PlasticitySolver: device context");
    return ctx.get();
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool point_matches(const vec2& a, const vec2& b) { return (a - b).norm() < 1e-14; }   // :989, :997

}  // namespace

PlasticitySolver::PlasticitySolver(const Material& material, Mesh& mesh, const LoadConditions& loadConditions)
    : m_(material), mesh_(mesh), cond_(loadConditions)
{
}

PlasticitySolver::~PlasticitySolver()
{
    if (prob_) cp_problem_destroy(prob_);
}

void PlasticitySolver::setPlainCondition(plainCondition pc)
{
    if (pc != pc_ && prob_) {
        cp_problem_destroy(prob_);
        prob_ = nullptr;
        gridSet_ = false;
    }
    pc_ = pc;
}

// The load vector, prescribed displacements and constraint mask of the
// reference's initConditions (PlasticitySolver.cu:812-1005), in its order:
// border forces, body forces, border displacements, fixed axes, fixed
// borders, point forces, point displacements.
void PlasticitySolver::initConditions(std::vector<double>& rp, std::vector<double>& u0,
                                      std::vector<unsigned char>& mask) const
{
    const int nn = mesh_.nodeCount, ne = mesh_.count4 + mesh_.count8;
    if (nn <= 0 || ne <= 0 || (mesh_.count4 && !mesh_.elem4) || (mesh_.count8 && !mesh_.elem8))
        throw std::invalid_argument("PlasticitySolver: empty mesh");
    if (mesh_.count3 || (mesh_.count4 && mesh_.count8))
        throw std::invalid_argument("PlasticitySolver: the mesh must hold 4-node or 8-node quads only");
    std::vector<unsigned char> sec;
    if (mesh_.count8) {
        sec.assign(nn, 0);
        for (int i = 0; i < nn; ++i) sec[i] = mesh_.secOrdNodes && mesh_.secOrdNodes[i] ? 1 : 0;
    }
    std::vector<double> x(nn), y(nn);
    rp.assign(2 * (size_t)nn, 0.);
    u0.assign(2 * (size_t)nn, 0.);
    mask.assign(2 * (size_t)nn, 1);
    for (int i = 0; i < nn; ++i) {
        x[i] = mesh_.node[i].x;
        y[i] = mesh_.node[i].y;
    }
    auto border = [&](int id, const char* what) {
        if (id < 0 || id >= mesh_.bordersCount) throw std::invalid_argument(std::string(what) + ": no such border");
        return std::make_pair(mesh_.borders[id], mesh_.borderLength[id]);
    };
    for (const auto& [id, fc] : cond_.forceCond) {
        auto [b, len] = border(id, "setForce");
        std::vector<double> fval(2 * (size_t)len);
        for (int k = 0; k < len; ++k) {
            const vec2 f = fc.forceFunc(mesh_.node[b[k]]);
            fval[2 * k] = f.x;
            fval[2 * k + 1] = f.y;
        }
        check(cp_rhs_border_force_q(x.data(), y.data(), b, len, sec.empty() ? nullptr : sec.data(), fval.data(),
                                    fc.normOrient ? 1 : 0, m_.h, rp.data()),
              "border force");
    }
    if (cond_.Rset) {
        std::vector<double> rnode(2 * (size_t)nn);
        for (int i = 0; i < nn; ++i) {
            const vec2 r = cond_.R(mesh_.node[i]);
            rnode[2 * i] = r.x;
            rnode[2 * i + 1] = r.y;
        }
        if (mesh_.count8)
            check(cp_rhs_body_force_nodal8(ne, x.data(), y.data(), mesh_.elem8, rnode.data(), m_.rho, m_.h,
                                           rp.data()),
                  "body force");
        else
            check(cp_rhs_body_force_nodal(ne, x.data(), y.data(), mesh_.elem4, rnode.data(), m_.rho, m_.h,
                                          rp.data()),
                  "body force");
    }
    for (const auto& [id, dc] : cond_.displCond) {
        auto [b, len] = border(id, "setDisplacement");
        std::vector<double> uval(2 * (size_t)len);
        for (int k = 0; k < len; ++k) {
            const vec2 u = dc.displFunc(mesh_.node[b[k]]);
            uval[2 * k] = u.x;
            uval[2 * k + 1] = u.y;
        }
        check(cp_displacement_border_values(b, len, uval.data(), dc.fixMidX ? 1 : 0, dc.fixMidY ? 1 : 0,
                                            mask.data(), u0.data()),
              "border displacement");
    }
    for (const FixedAxis& fa : cond_.fixedAxis)
        check(cp_fix_axis(nn, x.data(), y.data(), fa.vertical ? 1 : 0, fa.coord, mask.data()), "fixed axis");
    for (const auto& [id, fb] : cond_.fixedBorder) {
        auto [b, len] = border(id, "fixBorderVert/Hor");
        const int end = len - (b[0] == b[len - 1] ? 1 : 0);
        for (int i = 0; i < end; ++i) mask[2 * (size_t)b[i] + (fb.vertical ? 0 : 1)] = 0;
    }
    for (const PointCondition& pcnd : cond_.forcePoint)
        for (int i = 0; i < nn; ++i)
            if (point_matches(mesh_.node[i], pcnd.point)) {
                rp[2 * (size_t)i] += pcnd.value.x;
                rp[2 * (size_t)i + 1] += pcnd.value.y;
                break;
            }
    for (const PointCondition& pcnd : cond_.displPoint)
        for (int i = 0; i < nn; ++i)
            if (point_matches(mesh_.node[i], pcnd.point)) {
                u0[2 * (size_t)i] = pcnd.value.x;
                u0[2 * (size_t)i + 1] = pcnd.value.y;
                mask[2 * (size_t)i] = mask[2 * (size_t)i + 1] = 0;
                break;
            }
}

void PlasticitySolver::buildProblem()
{
    std::vector<double> rp, u0;
    std::vector<unsigned char> mask;
    initConditions(rp, u0, mask);
    const int nn = mesh_.nodeCount;
    std::vector<double> x(nn), y(nn);
    for (int i = 0; i < nn; ++i) {
        x[i] = mesh_.node[i].x;
        y[i] = mesh_.node[i].y;
    }
    const cp_material mat = m_.toAbi(pc_ == strain ? CP_PLANE_STRAIN : CP_PLANE_STRESS);
    const bool q8 = mesh_.count8 > 0;
    check(cp_problem_create_n(shared_context(), q8 ? 8 : 4, nn, q8 ? mesh_.count8 : mesh_.count4, x.data(), y.data(),
                              q8 ? mesh_.elem8 : mesh_.elem4, rp.data(), u0.data(), mask.data(), &mat, &prob_),
          "problem setup");
}

// The problem, and with options.multigrid its lattice (once per problem).
void PlasticitySolver::prepare()
{
    if (options.multigrid && mesh_.latticeN1 == 0)
        throw std::invalid_argument("PlasticitySolver: multigrid needs a generated mesh (no lattice)");
    if (!prob_) buildProblem();
    if (!options.multigrid || gridSet_) return;
    const auto set = mesh_.latticePeriodic ? cp_problem_set_grid_periodic : cp_problem_set_grid;
    check(set(prob_, mesh_.latticeN1, mesh_.latticeN2, mesh_.latticeNodes.data()), "multigrid lattice");
    gridSet_ = true;
}

cp_options PlasticitySolver::abiOptions(bool mixed) const
{
    cp_options o{};
    o.cg_tol = options.cgTol;
    o.cg_maxit = options.cgMaxIt;
    o.tol = options.tol;
    o.max_outer = options.maxOuter;
    o.fixed_outer = options.fixedOuter;
    o.aitken = options.aitken ? 1 : 0;
    o.precision = mixed ? CP_MIXED : CP_FP64;
    o.warm = 0;
    o.inner_tol = options.innerTol;
    o.precond = options.multigrid ? 2 : options.blockJacobi ? 1 : 0;
    o.predictor = options.predictor;
    o.cg_forcing = options.cgForcing;
    return o;
}

double PlasticitySolver::run(bool elastic, bool mixed)
{
    const double t0 = now_s();
    prepare();
    check(cp_set_load(prob_, 1.0), "set load");
    const cp_options o = abiOptions(mixed);
    const int s = elastic ? cp_solve_elastic(prob_, &o, &stats_) : cp_solve(prob_, &o, &stats_);
    ramSaved_ = false;
    plastSolved_ = !elastic;
    stepStats_.assign(1, stats_);
    if (s != CP_OK && s != CP_ERR_NOT_CONVERGED) check(s, elastic ? "solveElast" : "solve");
    if (iterOutput || s == CP_ERR_NOT_CONVERGED)
        std::cout << (s == CP_ERR_NOT_CONVERGED ? "Not converged: " : "Solved: ") << stats_.outer_iters
                  << " iterations, inner iterations: " << stats_.cg_iters << ", error: " << stats_.rel_err << "\n";
    return now_s() - t0;
}

double PlasticitySolver::solve() { return run(false, floatBoost || options.mixedPrecision); }
double PlasticitySolver::solveElast() { return run(true, floatBoost || options.mixedPrecision); }
double PlasticitySolver::solveCUDA() { return run(false, options.mixedPrecision); }
double PlasticitySolver::solveCUDA_FD() { return run(false, true); }
double PlasticitySolver::solveElastCUDA() { return run(true, options.mixedPrecision); }

double PlasticitySolver::solveCPU()
{
    throw std::runtime_error("PlasticitySolver::solveCPU: this library has no CPU solver; use solve()");
}

double PlasticitySolver::solveElastCPU()
{
    throw std::runtime_error("PlasticitySolver::solveElastCPU: this library has no CPU solver; use solveElast()");
}

double PlasticitySolver::solveSteps(int nsteps, double finalFactor)
{
    if (nsteps < 1) throw std::invalid_argument("solveSteps: nsteps < 1");
    const double t0 = now_s();
    prepare();
    const cp_options o = abiOptions(floatBoost || options.mixedPrecision);
    stepStats_.assign(nsteps, cp_stats{});
    const int s = cp_solve_steps(prob_, nsteps, finalFactor, &o, stepStats_.data());
    ramSaved_ = false;
    plastSolved_ = true;
    stats_ = stepStats_.back();
    if (s != CP_OK && s != CP_ERR_NOT_CONVERGED) check(s, "solveSteps");
    if (iterOutput || s == CP_ERR_NOT_CONVERGED)
        for (int n = 0; n < nsteps; ++n)
            std::cout << "Load step " << n + 1 << "/" << nsteps << ": " << stepStats_[n].outer_iters
                      << " iterations, inner iterations: " << stepStats_[n].cg_iters
                      << ", error: " << stepStats_[n].rel_err << "\n";
    return now_s() - t0;
}

void PlasticitySolver::dataToRAM()
{
    if (ramSaved_) return;
    if (!prob_) throw std::runtime_error("PlasticitySolver: nothing solved yet");
    fields_.assign(CP_NFIELDS, std::vector<double>(mesh_.elemCount()));
    for (int f = 0; f < CP_NFIELDS; ++f) check(cp_download(prob_, f, fields_[f].data()), "download");
    u_.assign(2 * (size_t)mesh_.nodeCount, 0.);
    check(cp_download(prob_, CP_F_U, u_.data()), "download");
    ramSaved_ = true;
}

const std::vector<double>& PlasticitySolver::field(int f)
{
    if (f < 0 || f >= CP_NFIELDS) throw std::invalid_argument("field: unknown field");
    dataToRAM();
    return fields_[f];
}

const std::vector<double>& PlasticitySolver::displacements()
{
    dataToRAM();
    return u_;
}

double PlasticitySolver::getSxx0() const
{
    double v = 0.;
    if (!prob_) return 0.;
    std::vector<double> s(mesh_.elemCount());
    check(cp_download(prob_, CP_F_SXX, s.data()), "download");
    v = s[0];
    return v;
}

double PlasticitySolver::getExx0() const
{
    if (!prob_) return 0.;
    std::vector<double> s(mesh_.elemCount());
    check(cp_download(prob_, CP_F_EXX, s.data()), "download");
    return s[0];
}

// PlasticitySolver.cuh:415-445: elastic elements add eps_i sigma_i V / 2;
// yielded ones sigma_T eps_T V / 2 plus the 100-panel trapezoid rule of f
// from eps_T to eps_i.  8-node elements add eps_i sigma_i V / 2 in every
// case (:437-443).
double PlasticitySolver::deformEnergy()
{
    dataToRAM();
    const std::vector<double>& sI = fields_[CP_F_SIGI];
    const std::vector<double>& eI = fields_[CP_F_EPSI];
    const double eT = m_.sigmaT / m_.E;
    double energy = 0.;
    for (int e = 0; e < mesh_.count4; ++e) {
        const double V = m_.h * mesh_.elemSpace4(e);
        if (sI[e] > m_.sigmaT) {
            energy += 0.5 * m_.sigmaT * eT * V;
            const double he = (eI[e] - eT) * 0.01;
            for (int j = 0; j < 100; ++j) energy += he * 0.5 * (m_.f(eT + he * j) + m_.f(eT + he * (j + 1))) * V;
        } else {
            energy += 0.5 * eI[e] * sI[e] * V;
        }
    }
    for (int e = 0; e < mesh_.count8; ++e) energy += eI[e] * sI[e] * 0.5 * (m_.h * mesh_.elemSpace8(e));
    return energy;
}

void PlasticitySolver::saveDisplsToFile(const std::string& fileName)
{
    dataToRAM();
    std::ofstream out(fileName);
    if (!out) throw std::runtime_error("saveDisplsToFile: cannot open " + fileName);
    for (int i = 0; i < mesh_.nodeCount; ++i)
        out << mesh_.node[i].x << " " << mesh_.node[i].y << " " << u_[2 * (size_t)i] << " " << u_[2 * (size_t)i + 1]
            << "\n";
}

void PlasticitySolver::saveAsVtk(const std::string& fileName)
{
    dataToRAM();
    const int nn = mesh_.nodeCount, ne = mesh_.elemCount();
    std::ofstream f(fileName);
    if (!f) throw std::runtime_error("saveAsVtk: cannot open " + fileName);
    f << "# vtk DataFile Version 2.0\nSolution\nASCII\nDATASET POLYDATA\nPOINTS " << nn << " float\n";
    for (int i = 0; i < nn; ++i) f << mesh_.node[i].x << " " << mesh_.node[i].y << " " << 0 << "\n";
    f << "POLYGONS " << ne << " " << 5 * mesh_.count4 + 9 * mesh_.count8;
    for (int e = 0; e < mesh_.count4; ++e) {
        f << "\n4 ";
        for (int j = 0; j < 4; ++j) f << mesh_.elem4[4 * (size_t)e + j] << " ";
    }
    for (int e = 0; e < mesh_.count8; ++e) {   // corner, mid-side, ... (PlasticitySolver.cu:1809-1813)
        f << "\n8 ";
        for (int j = 0; j < 4; ++j) f << mesh_.elem8[8 * (size_t)e + j] << " " << mesh_.elem8[8 * (size_t)e + j + 4] << " ";
    }
    f << "\nCELL_DATA " << ne << "\nSCALARS ElementID int 1\nLOOKUP_TABLE default\n";
    for (int e = 0; e < ne; ++e) f << e << "\n";
    f.precision(16);
    f << "FIELD FieldData " << 8 + (plastSolved_ ? 7 : 0) + (mesh_.analysed ? 3 : 0);

    auto column = [&](const char* name, auto value) {
        f << "\n" << name << " 1 " << ne << " float\n";
        for (int e = 0; e < ne; ++e) f << value(e) << " ";
    };
    // polar components of a symmetric tensor (x, y, xy) at the element centre
    auto polar = [&](int e, double x, double y, double xy, int comp) {
        const vec2 c = elemCenter(e);
        const double r = c.norm(), co = c.x / r, si = c.y / r;
        if (comp == 0) return co * co * x + si * si * y + 2. * co * si * xy;
        if (comp == 1) return si * si * x + co * co * y - 2. * co * si * xy;
        return co * si * (y - x) + xy * (co * co - si * si);
    };
    auto tensor = [&](const char* const (&names)[3], int fx, int fy, int fxy, double xy_scale) {
        const std::vector<double>&X = fields_[fx], &Y = fields_[fy], &XY = fields_[fxy];
        for (int comp = 0; comp < 3; ++comp) {
            if (polarCoord)
                column(names[comp], [&](int e) { return polar(e, X[e], Y[e], xy_scale * XY[e], comp); });
            else
                column(names[comp], [&](int e) { return (comp == 0 ? X : comp == 1 ? Y : XY)[e]; });
        }
    };
    const std::vector<std::vector<double>>& F = fields_;
    static const char* const strain_xy[3] = {"Eps_xx", "Eps_yy", "Gamma_xy"};
    static const char* const strain_rp[3] = {"Eps_rr", "Eps_phiphi", "Gamma_rphi"};
    static const char* const pstrain_xy[3] = {"Eps_p_xx", "Eps_p_yy", "Gamma_p_xy"};
    static const char* const pstrain_rp[3] = {"Eps_p_rr", "Eps_p_phiphi", "Gamma_p_rphi"};
    static const char* const stress_xy[3] = {"Sigma_xx", "Sigma_yy", "Tau_xy"};
    static const char* const stress_rp[3] = {"Sigma_rr", "Sigma_phiphi", "Tau_rphi"};
    // engineering shear strain: its tensor component is half of it
    tensor(polarCoord ? strain_rp : strain_xy, CP_F_EXX, CP_F_EYY, CP_F_GXY, 0.5);
    column("Eps_intensity", [&](int e) { return F[CP_F_EPSI][e]; });
    if (plastSolved_) {
        tensor(polarCoord ? pstrain_rp : pstrain_xy, CP_F_EXXP, CP_F_EYYP, CP_F_GXYP, 0.5);
        column("Eps_p_intensity", [&](int e) { return F[CP_F_EPSIP][e]; });
        column("Psi", [&](int e) { return F[CP_F_PSI][e]; });
        column("E_c", [&](int e) { return F[CP_F_EC][e]; });
        column("nu_c", [&](int e) { return F[CP_F_NUC][e]; });
    }
    tensor(polarCoord ? stress_rp : stress_xy, CP_F_SXX, CP_F_SYY, CP_F_TAU, 1.0);
    column("Sigma_intensity", [&](int e) { return F[CP_F_SIGI][e]; });
    if (mesh_.analysed) {
        column("Element_space", [&](int e) { return mesh_.spaces[e]; });
        column("Element_aspect_ratio", [&](int e) { return mesh_.aspects[e]; });
        column("Element_skew_angle", [&](int e) { return mesh_.skewAngles[e]; });
    }
    f << "\nPOINT_DATA " << nn << "\nSCALARS NodeID int 1\nLOOKUP_TABLE my_table";
    for (int i = 0; i < nn; ++i) f << "\n" << i;
    f << "\nFIELD FieldData2 1\nDisplacement 3 " << nn << " float\n";
    for (int i = 0; i < nn; ++i) f << u_[2 * (size_t)i] << " " << u_[2 * (size_t)i + 1] << " 0 ";
}

// Mean of the element's corner nodes.
vec2 PlasticitySolver::elemCenter(int e) const
{
    vec2 c;
    const bool q8 = mesh_.count8 > 0;
    for (int i = 0; i < 4; ++i) c += mesh_.node[q8 ? mesh_.elem8[8 * (size_t)e + i] : mesh_.elem4[4 * (size_t)e + i]];
    return c / 4.;
}

double PlasticitySolver::elemArea(int e) const
{
    return mesh_.count8 > 0 ? mesh_.elemSpace8(e) : mesh_.elemSpace4(e);
}

// PlasticitySolver.cuh:464-494: largest nodal-centre error of each stress
// component, relative to its area-weighted mean; the worst of the three.
double PlasticitySolver::sigmaError(std::function<double(vec2)> s_x, std::function<double(vec2)> s_y,
                                    std::function<double(vec2)> s_xy)
{
    dataToRAM();
    const std::vector<double>&sx = fields_[CP_F_SXX], &sy = fields_[CP_F_SYY], &txy = fields_[CP_F_TAU];
    double mean[3] = {0., 0., 0.}, worst[3] = {0., 0., 0.}, vol = 0.;
    for (int e = 0; e < mesh_.elemCount(); ++e) {
        const double a = elemArea(e);
        mean[0] += sx[e] * a;
        mean[1] += sy[e] * a;
        mean[2] += txy[e] * a;
        vol += a;
        const vec2 c = elemCenter(e);
        worst[0] = std::max(worst[0], std::fabs(s_x(c) - sx[e]));
        worst[1] = std::max(worst[1], std::fabs(s_y(c) - sy[e]));
        worst[2] = std::max(worst[2], std::fabs(s_xy(c) - txy[e]));
    }
    for (int k = 0; k < 3; ++k) worst[k] *= vol / mean[k];
    return std::fmax(std::fmax(worst[0], worst[1]), worst[2]);
}

// PlasticitySolver.cuh:497-541: the same in polar components (r, phi) of
// the element centre, for elements with r <= maxR.
double PlasticitySolver::polarSigmaError(std::function<double(vec2)> s_r, std::function<double(vec2)> s_phi,
                                         std::function<double(vec2)> s_rphi, double maxR)
{
    dataToRAM();
    const std::vector<double>&sx = fields_[CP_F_SXX], &sy = fields_[CP_F_SYY], &txy = fields_[CP_F_TAU];
    const double pi = 3.141592653589793;
    double mean[3] = {0., 0., 0.}, worst[3] = {0., 0., 0.}, vol = 0.;
    for (int e = 0; e < mesh_.elemCount(); ++e) {
        const vec2 c = elemCenter(e);
        const double r = c.norm();
        const vec2 rp(r, std::atan(c.y / c.x) + (c.x < 0. ? pi : 0.));
        if (rp.x > maxR) continue;
        const double co = c.x / r, si = c.y / r;
        const double srr = co * co * sx[e] + si * si * sy[e] + 2. * co * si * txy[e];
        const double spp = si * si * sx[e] + co * co * sy[e] - 2. * co * si * txy[e];
        const double srp = co * si * (sy[e] - sx[e]) + txy[e] * (co * co - si * si);
        const double a = elemArea(e);
        mean[0] += std::fabs(srr) * a;
        mean[1] += std::fabs(spp) * a;
        mean[2] += std::fabs(srp) * a;
        vol += a;
        worst[0] = std::max(worst[0], std::fabs(s_r(rp) - srr));
        worst[1] = std::max(worst[1], std::fabs(s_phi(rp) - spp));
        worst[2] = std::max(worst[2], std::fabs(s_rphi(rp) - srp));
    }
    for (int k = 0; k < 3; ++k) worst[k] *= vol / mean[k];
    return std::fmax(std::fmax(worst[0], worst[1]), worst[2]);
}
