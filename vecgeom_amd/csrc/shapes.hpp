// All device shape kernels.
#pragma once
#include "shapes_box_tube.hpp"
#include "shapes_simple.hpp"
#include "shapes_trap.hpp"
#include "shapes_sphere.hpp"
#include "shapes_cone.hpp"
#include "shapes_polycone.hpp"
#include "shapes_torus.hpp"
#include "shapes_polyhedron.hpp"
