// Test infrastructure: round trip of the facade's flattening
// (vecgeom_amd_facade.h).  A geometry given as vg_flat_geometry arrays is
// built with the reference's own classes (vgref_geom_create of
// oracle/ref/refnav.cpp: LogicalVolume, PlaceDaughter, GeoManager), then
// flattened back by the facade from the reference's world volume; the
// arrays are returned for comparison with the original
// (tests/test_facade_flatten.py).
#include <cstring>

#include "management/GeoManager.h"
#include "vecgeom_amd_facade.h"

extern "C" void *vgref_geom_create(int n_shapes, const int *shape_type, const int *shape_param_begin,
                                   const double *params, int n_logical, const int *logical_shape,
                                   const int *logical_dbegin, const int *daughters, int n_placed,
                                   const int *placed_logical, const double *placed_xf, int world);

namespace {
vecgeom_amd_facade::FlatGeometry g_flat;
}

extern "C" {

// Builds the reference tree and flattens it; returns the counts
// (n_shapes, n_params, n_logical, n_daughters, n_placed, world) in out6, or
// -1 if the reference tree could not be built or flattened.
int facade_roundtrip(int n_shapes, const int *shape_type, const int *shape_param_begin, const double *params,
                     int n_logical, const int *logical_shape, const int *logical_dbegin, const int *daughters,
                     int n_placed, const int *placed_logical, const double *placed_xf, int world, int *out6) {
  if (!vgref_geom_create(n_shapes, shape_type, shape_param_begin, params, n_logical, logical_shape, logical_dbegin,
                         daughters, n_placed, placed_logical, placed_xf, world))
    return -1;
  g_flat = vecgeom_amd_facade::FlatGeometry();
  if (!vecgeom_amd_facade::Flatten(vecgeom::GeoManager::Instance().GetWorld(), g_flat)) return -1;
  out6[0] = (int)g_flat.shape_type.size();
  out6[1] = (int)g_flat.params.size();
  out6[2] = (int)g_flat.logical_shape.size();
  out6[3] = (int)g_flat.daughters.size();
  out6[4] = (int)g_flat.placed_logical.size();
  out6[5] = g_flat.world;
  return 0;
}

// Copies the flattened arrays (sizes from facade_roundtrip).
void facade_arrays(int *shape_type, int *shape_param_begin, double *params, int *logical_shape, int *logical_dbegin,
                   int *daughters, int *placed_logical, double *placed_xf) {
  auto cp = [](auto *dst, const auto &v) { std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
  cp(shape_type, g_flat.shape_type);
  cp(shape_param_begin, g_flat.shape_param_begin);
  cp(params, g_flat.params);
  cp(logical_shape, g_flat.logical_shape);
  cp(logical_dbegin, g_flat.logical_dbegin);
  cp(daughters, g_flat.daughters);
  cp(placed_logical, g_flat.placed_logical);
  cp(placed_xf, g_flat.placed_transform);
}

}  // extern "C"
