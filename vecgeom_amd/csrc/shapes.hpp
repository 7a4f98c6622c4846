// All device shape kernels and the per-type dispatch helpers.
#pragma once
#include "shapes_box_tube.hpp"
