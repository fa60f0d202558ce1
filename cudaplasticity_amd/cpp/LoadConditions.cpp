// LoadConditions.cpp -- setters of the drop-in LoadConditions class
// (LoadConditions.cuh:69-179): they only record conditions; the solver turns
// them into the load vector and constraint mask when it builds the problem.
#include "cudaplasticity/LoadConditions.hpp"

void LoadConditions::setDisplacement(size_t borderId, vec2 u, bool fixOnlyMiddleX, bool fixOnlyMiddleY)
{
    DisplacementCondition& d = displCond[(int)borderId];
    d.displFunc = [u](vec2) { return u; };
    d.fixMidX = fixOnlyMiddleX;
    d.fixMidY = fixOnlyMiddleY;
}

void LoadConditions::setDisplacement(size_t borderId, std::function<vec2(vec2)> displFunc)
{
    DisplacementCondition& d = displCond[(int)borderId];
    d.displFunc = std::move(displFunc);
    d.fixMidX = d.fixMidY = false;
}

void LoadConditions::setForce(size_t borderId, vec2 f, bool normalOriented)
{
    ForceCondition& c = forceCond[(int)borderId];
    c.forceFunc = [f](vec2) { return f; };
    c.value = f;
    c.normOrient = normalOriented;
}

void LoadConditions::setForce(size_t borderId, std::function<vec2(vec2)> forceFunc, bool normalOriented)
{
    ForceCondition& c = forceCond[(int)borderId];
    c.forceFunc = std::move(forceFunc);
    c.normOrient = normalOriented;
}

void LoadConditions::fixVertAxis(double x) { fixedAxis.push_back(FixedAxis{true, x}); }

void LoadConditions::fixHorAxis(double y) { fixedAxis.push_back(FixedAxis{false, y}); }

void LoadConditions::fixBorder(size_t borderId) { setDisplacement(borderId, vec2(0., 0.)); }

void LoadConditions::fixBorderVert(size_t borderId) { fixedBorder[(int)borderId].vertical = true; }

void LoadConditions::fixBorderHor(size_t borderId) { fixedBorder[(int)borderId].vertical = false; }

void LoadConditions::fixPoint(vec2 point) { displPoint.push_back(PointCondition{point, vec2(0., 0.)}); }

void LoadConditions::setDisplPoint(vec2 point, vec2 displ) { displPoint.push_back(PointCondition{point, displ}); }

void LoadConditions::setForcePoint(vec2 point, vec2 force) { forcePoint.push_back(PointCondition{point, force}); }

void LoadConditions::setGravity(double g, vec2 direction)
{
    const vec2 force = g * direction.normalize();
    R = [force](vec2) { return force; };
    Rset = true;
}

void LoadConditions::setRotation(double omega, vec2 rotationAxis)
{
    const double w2 = omega * omega;
    R = [w2, rotationAxis](vec2 r) { return w2 * (r - rotationAxis); };
    Rset = true;
}

void LoadConditions::setVolumeForce(std::function<vec2(vec2)> R_)
{
    R = std::move(R_);
    Rset = true;
}
