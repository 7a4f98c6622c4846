// The reference-side binding a VecGeom maintainer would add to hand a
// geometry built with VecGeom's own classes to libvecgeom_amd: it flattens
// the placed-volume tree under a world volume into the C-ABI's
// vg_flat_geometry (include/vecgeom_amd.h) and uploads it -- the part of
// CudaManager::LoadGeometry + Synchronize (management/CudaManager.h:97-160)
// that walks the tree.  Compiled against the reference headers by
// oracle/Makefile (target facade) and checked by tests/test_facade_flatten.py.
//
// Shapes are read through the unplaced classes' public getters; the types
// covered here are those with getters for every constructor parameter
// (Box, Tube, Orb, Trd); AddShape returns -1 for the others.
#pragma once
#include <map>
#include <vector>

#include "../../include/vecgeom_amd.h"
#include "base/Transformation3D.h"
#include "volumes/LogicalVolume.h"
#include "volumes/PlacedVolume.h"
#include "volumes/UnplacedBox.h"
#include "volumes/UnplacedOrb.h"
#include "volumes/UnplacedTrd.h"
#include "volumes/UnplacedTube.h"

namespace vecgeom_amd_facade {

struct FlatGeometry {
  std::vector<int32_t> shape_type, shape_param_begin{0}, logical_shape, logical_dbegin{0}, daughters, placed_logical;
  std::vector<double> params, placed_transform;
  int32_t world = -1;
  std::map<vecgeom::VUnplacedVolume const *, int> shape_id;
  std::map<vecgeom::LogicalVolume const *, int> logical_id;
  std::map<vecgeom::VPlacedVolume const *, int> placed_id;

  vg_flat_geometry View() const {
    return vg_flat_geometry{(int32_t)shape_type.size(), shape_type.data(), shape_param_begin.data(), params.data(),
                            (int32_t)logical_shape.size(), logical_shape.data(), logical_dbegin.data(),
                            daughters.data(), (int32_t)placed_logical.size(), placed_logical.data(),
                            placed_transform.data(), world};
  }
};

// Appends shape u (once per object); returns its id, or -1 if the shape type
// is not covered.
inline int AddShape(vecgeom::VUnplacedVolume const *u, FlatGeometry &f) {
  using namespace vecgeom;
  auto it = f.shape_id.find(u);
  if (it != f.shape_id.end()) return it->second;
  int type;
  std::vector<double> p;
  if (auto *b = dynamic_cast<UnplacedBox const *>(u)) {
    type = VG_BOX;
    p = {b->x(), b->y(), b->z()};
  } else if (auto *t = dynamic_cast<UnplacedTube const *>(u)) {
    type = VG_TUBE;
    p = {t->rmin(), t->rmax(), t->z(), t->sphi(), t->dphi()};
  } else if (auto *o = dynamic_cast<UnplacedOrb const *>(u)) {
    type = VG_ORB;
    p = {o->GetRadius()};
  } else if (auto *d = dynamic_cast<UnplacedTrd const *>(u)) {
    type = VG_TRD;
    p = {d->dx1(), d->dx2(), d->dy1(), d->dy2(), d->dz()};
  } else {
    return -1;
  }
  const int id = (int)f.shape_type.size();
  f.shape_type.push_back(type);
  f.params.insert(f.params.end(), p.begin(), p.end());
  f.shape_param_begin.push_back((int32_t)f.params.size());
  f.shape_id[u] = id;
  return id;
}

// Flattens the tree under `world`: logical volumes and placed volumes get
// ids in the order a depth-first walk first meets them (a placed volume of
// a logical volume placed several times is one placed volume), daughters
// keep their PlaceDaughter order.  Returns false if some shape is not
// covered.
inline bool Flatten(vecgeom::VPlacedVolume const *world, FlatGeometry &f) {
  using namespace vecgeom;
  // Ids first (depth-first, daughters in order), then the tables.
  std::vector<VPlacedVolume const *> order;
  std::vector<VPlacedVolume const *> stack{world};
  while (!stack.empty()) {
    VPlacedVolume const *pv = stack.back();
    stack.pop_back();
    if (f.placed_id.count(pv)) continue;  // placed again through a shared logical volume
    f.placed_id[pv] = (int)order.size();
    order.push_back(pv);
    Vector<Daughter> const &ds = pv->GetLogicalVolume()->daughters();
    for (int k = (int)ds.size() - 1; k >= 0; --k) stack.push_back(ds[k]);
  }
  std::vector<LogicalVolume const *> logicals;
  for (VPlacedVolume const *pv : order) {
    LogicalVolume const *lv = pv->GetLogicalVolume();
    if (f.logical_id.count(lv)) continue;
    f.logical_id[lv] = (int)logicals.size();
    logicals.push_back(lv);
  }
  for (LogicalVolume const *lv : logicals) {
    const int s = AddShape(lv->unplaced_volume(), f);
    if (s < 0) return false;
    f.logical_shape.push_back(s);
    for (Daughter d : lv->daughters()) f.daughters.push_back(f.placed_id.at(d));
    f.logical_dbegin.push_back((int32_t)f.daughters.size());
  }
  for (VPlacedVolume const *pv : order) {
    f.placed_logical.push_back(f.logical_id.at(pv->GetLogicalVolume()));
    Transformation3D const *t = pv->GetTransformation();
    for (int i = 0; i < 3; ++i) f.placed_transform.push_back(t->Translation(i));
    for (int i = 0; i < 9; ++i) f.placed_transform.push_back(t->Rotation()[i]);
  }
  f.world = f.placed_id.at(world);
  return true;
}

// CudaManager::LoadGeometry + Synchronize in one call.
inline int UploadWorld(vecgeom::VPlacedVolume const *world, vg_geometry **out) {
  FlatGeometry f;
  if (!Flatten(world, f)) return VG_ERR_UNSUPPORTED;
  const vg_flat_geometry v = f.View();
  return vg_geometry_upload(&v, out);
}

}  // namespace vecgeom_amd_facade
