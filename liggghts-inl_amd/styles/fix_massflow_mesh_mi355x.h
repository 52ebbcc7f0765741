/* fix massflow/mesh/mi355x -- fix massflow/mesh counted on the device
 * (dem_add_massflow): particles of the fix's group that pass through the
 * face mesh to the vec_side side, once or at every crossing
 * (FixMassflowMesh::post_integrate, fix_massflow_mesh.cpp:366-530).  The
 * reference constructor parses the arguments and finds the face mesh; the
 * fix sets no host mask, and its vector (mass, count, and their rates over
 * the time between two reads, fix_massflow_mesh.cpp:637-662) comes from
 * the device, read once per output step.  point_at_outlet, delete_atoms
 * and sum_property have no device form.
 */
#ifdef FIX_CLASS

FixStyle(massflow/mesh/mi355x,FixMassflowMeshMI355X)

#else

#ifndef LMP_FIX_MASSFLOW_MESH_MI355X_H
#define LMP_FIX_MASSFLOW_MESH_MI355X_H

#include "fix_massflow_mesh.h"
#include "liggghts_dem.h"

namespace LAMMPS_NS {

class FixMassflowMeshMI355X : public FixMassflowMesh {
 public:
  FixMassflowMeshMI355X(class LAMMPS *, int, char **);
  int setmask() { return 0; }
  double compute_vector(int index);
  // group_bit: the fix's group (-1 for all), set before with dem_set_group
  int register_with(dem_ctx *ctx, int group_bit);

 private:
  double vec_side_[3];
  dem_ctx *ctx_;
  bigint read_step_;
  double vec_[4];
};

}

#endif
#endif
