/* fix mesh/surface/mi355x -- fix mesh/surface whose triangles also go to
 * the device.  It runs the reference FixMeshSurface unchanged (file input,
 * scale/move/rotate, element checks; fix_mesh_surface.cpp) and keeps the
 * two values the device needs that FixMeshSurface holds privately: the
 * wall's atom type and the curvature tolerance (`curvature <deg>`, stored
 * as its cosine as fix_mesh_surface.cpp:193-200 does).
 */
#ifdef FIX_CLASS

FixStyle(mesh/surface/mi355x,FixMeshSurfaceMI355X)

#else

#ifndef LMP_FIX_MESH_SURFACE_MI355X_H
#define LMP_FIX_MESH_SURFACE_MI355X_H

#include "fix_mesh_surface.h"

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

 private:
  double cos_curvature_;
};

}

#endif
#endif
