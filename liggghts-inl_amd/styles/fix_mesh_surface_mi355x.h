/* fix mesh/surface/mi355x and mesh/surface/stress/mi355x -- fix mesh/surface
 * whose triangles also go to the device.  It runs the reference
 * FixMeshSurface unchanged (file input, scale/move/rotate, element checks,
 * the mesh modules named in the style, fix_mesh_surface.cpp; the
 * constructor sees the unsuffixed style, modify.cpp:837-844) and keeps the
 * values the device needs that FixMeshSurface holds privately: the wall's
 * atom type, the curvature tolerance (`curvature <deg>`, stored as its
 * cosine as fix_mesh_surface.cpp:193-200 does) and, with the stress module,
 * its reference point (`reference_point`, mesh_module_stress.cpp:106-118).
 * With the stress module the force and torque sums are made on the device
 * (dem_mesh_stress) and the fix's vector is read from it
 * (mesh_module_stress.cpp:479-488: force, torque about p_ref, p_ref).
 */
#ifdef FIX_CLASS

FixStyle(mesh/surface/mi355x,FixMeshSurfaceMI355X)
FixStyle(mesh/surface/stress/mi355x,FixMeshSurfaceMI355X)

#else

#ifndef LMP_FIX_MESH_SURFACE_MI355X_H
#define LMP_FIX_MESH_SURFACE_MI355X_H

#include "fix_mesh_surface.h"
#include "liggghts_dem.h"

#include <vector>

namespace LAMMPS_NS {

class FixMeshSurfaceMI355X : public FixMeshSurface {
 public:
  FixMeshSurfaceMI355X(class LAMMPS *, int, char **);
  virtual ~FixMeshSurfaceMI355X() {}

  // triangle nodes, 9 doubles per triangle, in mesh element order
  void nodes(std::vector<double> &out);
  int wall_type() const { return atom_type_mesh_; }
  // cosine of the curvature angle; 0 = the default (1 - 1e-5)
  double curvature() const { return cos_curvature_; }
  bool stress() const { return stress_; }
  const double *reference_point() const { return pref_; }
  // the device mesh this fix became (dem_add_mesh's index); with the
  // stress module, turns the device's sums on
  int attach(dem_ctx *ctx, int mesh);
  virtual double compute_vector(int n);

 private:
  double cos_curvature_;
  bool stress_;
  double pref_[3];
  dem_ctx *ctx_;
  int mesh_;
};

}

#endif
#endif
