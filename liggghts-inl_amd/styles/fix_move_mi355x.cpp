/* fix move/mi355x: argument parsing and hand-over to the device. */
#include "fix_move_mi355x.h"

#include "error.h"
#include "force.h"

#include <cstring>

using namespace LAMMPS_NS;

FixMoveMI355X::FixMoveMI355X(LAMMPS *lmp, int narg, char **arg) : FixMove(lmp, narg, arg)
{
  // fix ID group move linear Vx Vy Vz (fix_move.cpp:62-90)
  if (narg < 7 || strcmp(arg[3], "linear") != 0)
    error->fix_error(FLERR, this, "the device moves atoms with style linear only");
  for (int d = 0; d < 3; d++) {
    if (strcmp(arg[4 + d], "NULL") == 0) {
      style_[d] = 0;
      value_[d] = 0.0;
    } else {
      style_[d] = 1;
      value_[d] = force->numeric(FLERR, arg[4 + d]);
    }
  }
}

int FixMoveMI355X::register_with(dem_ctx *ctx, int bit, int *id) const
{
  return dem_add_move_linear(ctx, bit, style_, value_, id);
}
