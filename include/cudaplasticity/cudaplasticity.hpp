// cudaplasticity.hpp -- the drop-in C++ API in one include: vec2, Mesh,
// Material, LoadConditions, PlasticitySolver (link libcudaplasticity.so).
#pragma once

#include "LoadConditions.hpp"
#include "Material.hpp"
#include "Mesh.hpp"
#include "PlasticitySolver.hpp"
#include "vec2.hpp"
