// Material.hpp -- material of the drop-in C++ API (Material.cuh:9-41).
//
// Public E, nu, h, rho, sigmaT and the setters of the reference; the device
// function pointer of the reference is replaced by a law tag that the
// device code switches on.  setSwiftPlast is new.
#pragma once

#include <limits>

#include "cp_abi.h"

class Material {
public:
    double E = 2.1e5;                                      // elastic modulus
    double nu = 0.3;                                       // Poisson ratio
    double h = 1.;                                         // thickness
    double rho = 1.;                                       // density (body forces)
    double sigmaT = std::numeric_limits<double>::max();    // yield stress (no yield until a setter is called)

    // sigma = E eps below eps_T = sigmaT / E, sigmaT + (eps - eps_T) K_T above (Material.cu:3-6).
    void setLinearPlast(double E, double sigmaT, double K_T);
    // Reference power law with A, m (Material.cu:8-12).
    void setPowerPlast(double E, double sigmaT, double A, double m);
    // Swift hardening sigma = sigmaT (1 + eps_p / eps0)^n, eps0 = sigmaT / E (new).
    void setSwiftPlast(double E, double sigmaT, double n);

    // Stress intensity for strain intensity eps (host side of the same law
    // the device evaluates).
    double f(double eps) const;

    // C-ABI view; plane = CP_PLANE_STRESS / CP_PLANE_STRAIN.
    cp_material toAbi(int plane) const;

    int law() const { return law_; }

private:
    int law_ = CP_LAW_LINEAR;
    double K_T_ = 0., A_ = 0., m_ = 0., n_ = 0.;
};
