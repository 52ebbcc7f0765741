/* fix move/mesh/mi355x: linear mesh motion on the device. */
#include "fix_move_mesh_mi355x.h"

#include "error.h"
#include "force.h"

#include <cstring>

using namespace LAMMPS_NS;

FixMoveMeshMI355X::FixMoveMeshMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixMoveMesh(lmp, narg, arg)
{
  // move/mesh mesh <id> linear vx vy vz
  if (narg < 9 || strcmp(arg[3], "mesh") != 0 || strcmp(arg[5], "linear") != 0)
    error->fix_error(FLERR, this, "the device moves meshes with 'mesh <id> linear <vx> <vy> <vz>' only");
  for (int d = 0; d < 3; d++) vel_[d] = force->numeric(FLERR, arg[6 + d]);
}

int FixMoveMeshMI355X::setmask()
{
  return 0;
}
