// Host-side parameter blocks of the zoo shapes beyond Box and Tube.
#include "shapes.hpp"
#include "vg_internal.hpp"

namespace vga {

std::vector<double> BuildZooParams(int type, const double *, int, std::string *err) {
  if (err) *err = "shape type " + std::to_string(type) + " has no device kernels";
  return {};
}

bool ZooExtent(int, const double *, double *, double *) { return false; }

}  // namespace vga
