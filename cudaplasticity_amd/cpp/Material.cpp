// Material.cpp -- the drop-in Material class; f() evaluates the same law
// code the device kernels use (cp_material_f).
#include "cudaplasticity/Material.hpp"

#include <stdexcept>
#include <string>

void Material::setLinearPlast(double E_, double sigmaT_, double K_T)
{
    E = E_;
    sigmaT = sigmaT_;
    K_T_ = K_T;
    law_ = CP_LAW_LINEAR;
}

void Material::setPowerPlast(double E_, double sigmaT_, double A, double m)
{
    E = E_;
    sigmaT = sigmaT_;
    A_ = A;
    m_ = m;
    law_ = CP_LAW_POWER;
}

void Material::setSwiftPlast(double E_, double sigmaT_, double n)
{
    E = E_;
    sigmaT = sigmaT_;
    n_ = n;
    law_ = CP_LAW_SWIFT;
}

cp_material Material::toAbi(int plane) const
{
    cp_material c{};
    c.E = E;
    c.nu = nu;
    c.h = h;
    c.sigmaT = sigmaT;
    c.law = law_;
    c.plane = plane;
    c.K_T = law_ == CP_LAW_LINEAR && sigmaT < std::numeric_limits<double>::max() ? K_T_ : E;
    c.A = A_;
    c.m = m_;
    c.n = n_;
    return c;
}

double Material::f(double eps) const
{
    const cp_material c = toAbi(CP_PLANE_STRESS);
    double s = 0.;
    if (cp_material_f(&c, eps, &s) != CP_OK) throw std::runtime_error(std::string("Material::f: ") + cp_last_error());
    return s;
}
