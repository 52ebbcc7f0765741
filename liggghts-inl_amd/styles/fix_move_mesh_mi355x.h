/* fix move/mesh/mi355x -- fix move/mesh whose linear motion runs on the
 * device.  `fix ID group move/mesh mesh <mesh-id> linear <vx> <vy> <vz>`
 * (fix_move_mesh.cpp:70-110, mesh_mover_linear.cpp:66-77); other movement
 * types have no device form and are refused.  The reference mover is still
 * constructed (it registers with the mesh), but it sets no integrate mask:
 * the device translates the nodes each step (MeshMoverLinear::
 * initial_integrate, mesh_mover_linear.cpp:94-112).
 */
#ifdef FIX_CLASS

FixStyle(move/mesh/mi355x,FixMoveMeshMI355X)

#else

#ifndef LMP_FIX_MOVE_MESH_MI355X_H
#define LMP_FIX_MOVE_MESH_MI355X_H

#include "fix_move_mesh.h"

namespace LAMMPS_NS {

class FixMoveMeshMI355X : public FixMoveMesh {
 public:
  FixMoveMeshMI355X(class LAMMPS *, int, char **);
  virtual ~FixMoveMeshMI355X() {}
  int setmask();

  bool moves(const class FixMesh *m) { return fixMeshCompare(m); }
  const double *velocity() const { return vel_; }

 private:
  double vel_[3];
};

}

#endif
#endif
