// LoadConditions.hpp -- loads and constraints of the drop-in C++ API
// (LoadConditions.cuh:12-180).  Same containers and setters; the CUDA
// nvstd::function members become std::function, so host lambdas passed by
// reference callers compile unchanged.
#pragma once

#include <cstddef>
#include <functional>
#include <list>
#include <map>

#include "vec2.hpp"

struct ForceCondition {
    std::function<vec2(vec2)> forceFunc;
    vec2 value;
    bool normOrient = false;   // value.x normal, value.y tangential component
};

struct DisplacementCondition {
    std::function<vec2(vec2)> displFunc;
    bool fixMidX = false;
    bool fixMidY = false;
};

struct FixedAxis {
    bool vertical = true;
    double coord = 0.;
};

struct FixedBorder {
    bool vertical = true;
};

struct PointCondition {
    vec2 point;
    vec2 value;
};

class LoadConditions {
public:
    bool Rset = false;
    std::function<vec2(vec2)> R = [](vec2) { return vec2(0., 0.); };   // body force per unit mass

    std::map<int, DisplacementCondition> displCond;
    std::map<int, ForceCondition> forceCond;
    std::map<int, FixedBorder> fixedBorder;
    std::list<FixedAxis> fixedAxis;
    std::list<PointCondition> forcePoint;
    std::list<PointCondition> displPoint;

    void setDisplacement(size_t borderId, vec2 u, bool fixOnlyMiddleX = false, bool fixOnlyMiddleY = false);
    void setDisplacement(size_t borderId, std::function<vec2(vec2)> displFunc);
    void setForce(size_t borderId, vec2 f, bool normalOriented = false);
    void setForce(size_t borderId, std::function<vec2(vec2)> forceFunc, bool normalOriented = false);
    void fixVertAxis(double x);
    void fixHorAxis(double y);
    void fixBorder(size_t borderId);
    void fixBorderVert(size_t borderId);
    void fixBorderHor(size_t borderId);
    void fixPoint(vec2 point);
    void setDisplPoint(vec2 point, vec2 displ);
    void setForcePoint(vec2 point, vec2 force);
    void setGravity(double g = 9.81, vec2 direction = vec2(0., -1.));
    void setRotation(double omega, vec2 rotationAxis = vec2(0., 0.));
    void setVolumeForce(std::function<vec2(vec2)> R);
};
