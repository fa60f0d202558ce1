// vec2.hpp -- 2-D point/vector of the drop-in C++ API (vec2.cuh:7-104:
// same members x, y and the same operators; host only).
#pragma once

#include <cmath>

struct vec2 {
    double x = 0., y = 0.;

    vec2() = default;
    vec2(double x_, double y_) : x(x_), y(y_) {}

    double operator*(const vec2& v) const { return x * v.x + y * v.y; }  // dot product
    vec2 operator*(double k) const { return {x * k, y * k}; }
    vec2 operator/(double k) const { return {x / k, y / k}; }
    vec2 operator+(const vec2& v) const { return {x + v.x, y + v.y}; }
    vec2 operator-(const vec2& v) const { return {x - v.x, y - v.y}; }
    vec2 operator-() const { return {-x, -y}; }
    vec2& operator+=(const vec2& v) { x += v.x; y += v.y; return *this; }
    vec2& operator-=(const vec2& v) { x -= v.x; y -= v.y; return *this; }
    vec2& operator*=(double k) { x *= k; y *= k; return *this; }
    vec2& operator/=(double k) { x /= k; y /= k; return *this; }

    // z component of the cross product (vec2.cuh:81-85)
    double crossZ(const vec2& v) const { return x * v.y - y * v.x; }
    double norm() const { return std::sqrt(x * x + y * y); }
    vec2 normalize() const { return *this * (1. / norm()); }
};

inline vec2 operator*(double k, const vec2& v) { return v * k; }
