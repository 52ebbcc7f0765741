/* fix massflow/mesh/mi355x: the face mesh and the counting rule go to the
 * device; the vector is read from it. */
#include "fix_massflow_mesh_mi355x.h"

#include "error.h"
#include "fix_mesh_surface.h"
#include "force.h"
#include "tri_mesh.h"
#include "update.h"

#include <cstring>
#include <vector>

using namespace LAMMPS_NS;

FixMassflowMeshMI355X::FixMassflowMeshMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixMassflowMesh(lmp, narg, arg), ctx_(NULL), read_step_(-1)
{
  vec_side_[0] = vec_side_[1] = vec_side_[2] = 0.0;
  for (int k = 0; k < 4; k++) vec_[k] = 0.0;
  for (int i = 3; i < narg; i++) {
    if (strcmp(arg[i], "vec_side") == 0 && i + 3 < narg) {
      for (int d = 0; d < 3; d++) vec_side_[d] = force->numeric(FLERR, arg[i + 1 + d]);
      i += 3;
    } else if (strcmp(arg[i], "point_at_outlet") == 0 || strcmp(arg[i], "sum_property") == 0) {
      error->fix_error(FLERR, this, "point_at_outlet and sum_property have no device form");
    }
  }
  if (delete_atoms_) error->fix_error(FLERR, this, "delete_atoms has no device form");
}

int FixMassflowMeshMI355X::register_with(dem_ctx *ctx, int group_bit)
{
  ctx_ = ctx;
  TriMesh *m = fix_mesh_->triMesh();
  const int n = m->sizeLocal();
  std::vector<double> nodes(9 * (size_t)n);
  for (int t = 0; t < n; t++)
    for (int k = 0; k < 3; k++) m->node(t, k, &nodes[9 * (size_t)t + 3 * k]);
  return dem_add_massflow(ctx, n, nodes.data(), vec_side_, once_ ? 1 : 0, group_bit);
}

// one device read per step: the rates' time window restarts at each read,
// and an output step asks for the components one by one
double FixMassflowMeshMI355X::compute_vector(int index)
{
  if (!ctx_ || index < 0) return 0.0;
  if (read_step_ != update->ntimestep) {
    if (dem_massflow_vector(ctx_, vec_) != 0) error->one(FLERR, dem_last_error(ctx_));
    read_step_ = update->ntimestep;
  }
  return index < 4 ? vec_[index] : 0.0;
}
