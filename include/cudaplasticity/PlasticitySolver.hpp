// PlasticitySolver.hpp -- solver of the drop-in C++ API
// (PlasticitySolver.cuh:21-520).
//
// Same constructor, solve / solveElast entry points (returning elapsed
// seconds), flags (polarCoord, iterOutput, floatBoost), setPlainCondition,
// deformEnergy, getSxx0 / getExx0, saveDisplsToFile, sigmaError and
// polarSigmaError.  The work runs on the GPU through the C-ABI of
// include/cp_abi.h: the constructor copies mesh, material and load
// conditions and the first solve builds the device problem.
//
// Additive API: SolverOptions (tolerances, Aitken, precision, block-Jacobi,
// load predictor, inexact secant iterations), solveSteps for incremental
// loading, lastStats, field access and displacements.
//
// There is no CPU solver in this library: solveCPU / solveElastCPU throw
// std::runtime_error.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "LoadConditions.hpp"
#include "Material.hpp"
#include "Mesh.hpp"
#include "cp_abi.h"

enum plainCondition { stress, strain };   // PlasticitySolver.cuh:15

// Options of one solve; the defaults are the reference's algorithm
// (unpreconditioned CG to 1e-7, plain secant update, exit at 1e-5).
struct SolverOptions {
    double cgTol = 1e-7;
    int cgMaxIt = 0;            // 0: no cap beyond the library's own
    double tol = 1e-5;
    int maxOuter = 1000;
    int fixedOuter = 0;         // > 0: exactly this many secant iterations
    bool aitken = false;
    bool mixedPrecision = false;  // FP32 CG + FP64 refinement (also chosen by floatBoost)
    double innerTol = 1e-4;
    bool blockJacobi = false;
    int predictor = 0;          // solveSteps: 0 none, 1 extrapolate u, 2 u and log(psi)
    double cgForcing = 0.;      // > 0: inexact secant iterations
    // New: geometric multigrid preconditioner on the mesh's lattice
    // (Mesh::latticeN1; generated meshes, 8-node ones through their corner
    // lattice), before blockJacobi.
    bool multigrid = false;
};

class PlasticitySolver {
public:
    bool polarCoord = false;
    bool iterOutput = true;
    bool floatBoost = false;
    SolverOptions options;

    PlasticitySolver(const Material& material, Mesh& mesh, const LoadConditions& loadConditions);
    ~PlasticitySolver();
    PlasticitySolver(const PlasticitySolver&) = delete;
    PlasticitySolver& operator=(const PlasticitySolver&) = delete;

    void setPlainCondition(plainCondition pc);

    double solve();
    double solveElast();
    double solveCUDA();
    double solveCUDA_FD();
    double solveElastCUDA();
    double solveCPU();
    double solveElastCPU();

    // Incremental loading: levels finalFactor * n / nsteps, n = 1..nsteps,
    // each warm-started from the previous one.  Returns elapsed seconds.
    double solveSteps(int nsteps, double finalFactor = 1.);

    double deformEnergy();
    double getSxx0() const;
    double getExx0() const;
    void saveDisplsToFile(const std::string& fileName);
    // Legacy-VTK solution file in the reference's layout
    // (PlasticitySolver.cu:1784-1945): mesh, element ids, strain, plastic
    // (after solve) and stress fields, Cartesian or, with polarCoord, polar
    // components about the origin, the mesh quality fields when the mesh was
    // analysed, node ids and displacements.
    void saveAsVtk(const std::string& fileName);
    double sigmaError(std::function<double(vec2)> s_x, std::function<double(vec2)> s_y,
                      std::function<double(vec2)> s_xy);
    double polarSigmaError(std::function<double(vec2)> s_r, std::function<double(vec2)> s_phi,
                           std::function<double(vec2)> s_rphi, double maxR);

    // Load vector at full load (2 per node), prescribed displacements and
    // constraint mask (1 free, 0 fixed) as the solver builds them from the
    // load conditions (initConditions, PlasticitySolver.cu:812-1005); needs
    // no device.
    void initConditions(std::vector<double>& rp, std::vector<double>& u0, std::vector<unsigned char>& mask) const;

    // Additive accessors.
    const cp_stats& lastStats() const { return stats_; }
    const std::vector<cp_stats>& stepStats() const { return stepStats_; }
    const std::vector<double>& field(int cpField);   // CP_F_* per element
    const std::vector<double>& displacements();      // u0, v0, u1, v1, ...
    const Mesh& mesh() const { return mesh_; }

private:
    Material m_;
    Mesh mesh_;
    LoadConditions cond_;
    plainCondition pc_ = stress;
    cp_problem* prob_ = nullptr;
    cp_stats stats_{};
    std::vector<cp_stats> stepStats_;
    std::vector<std::vector<double>> fields_;
    std::vector<double> u_;
    bool ramSaved_ = false;
    bool plastSolved_ = false;

    bool gridSet_ = false;

    void buildProblem();
    void prepare();
    cp_options abiOptions(bool mixed) const;
    double run(bool elastic, bool mixed);
    void dataToRAM();
    vec2 elemCenter(int e) const;
    double elemArea(int e) const;
};
