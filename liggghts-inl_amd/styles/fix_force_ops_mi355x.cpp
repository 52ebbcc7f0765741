/* fix setforce/mi355x, fix aveforce/mi355x: constant or time-linear group
 * forces handed to the device. */
#include "fix_force_ops_mi355x.h"

#include "error.h"
#include "force.h"
#include "input.h"
#include "lammps.h"
#include "update.h"
#include "variable.h"

#include <cmath>
#include <cstring>
#include <vector>

using namespace LAMMPS_NS;

void ForceComponentMI355X::parse(LAMMPS *lmp, const char *arg)
{
  value = 0.0;
  if (strstr(arg, "v_") == arg) {
    style = 2;
    var = arg + 2;
  } else if (strcmp(arg, "NULL") == 0) {
    style = 0;
  } else {
    style = 1;
    value = lmp->force->numeric(FLERR, arg);
  }
}

FixSetForceMI355X::FixSetForceMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixSetForce(lmp, narg, arg)
{
  for (int d = 0; d < 3; d++) {
    c_[d].parse(lmp, arg[3 + d]);
    if (c_[d].style == 2) error->fix_error(FLERR, this, "the device takes constant setforce values only");
  }
}

int FixSetForceMI355X::register_with(dem_ctx *ctx, int bit, int *id) const
{
  int style[3];
  double value[3];
  for (int d = 0; d < 3; d++) {
    style[d] = c_[d].style != 0;
    value[d] = c_[d].value;
  }
  return dem_add_setforce(ctx, bit, style, value, id);
}

FixAveForceMI355X::FixAveForceMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixAveForce(lmp, narg, arg)
{
  for (int d = 0; d < 3; d++) c_[d].parse(lmp, arg[3 + d]);
  if (c_[0].style == 2 || c_[1].style == 2)
    error->fix_error(FLERR, this, "the device takes constant x and y aveforce values only");
}

// the z expression at a step: `time` is atime + (step - atimestep) * dt
double FixAveForceMI355X::z_at(bigint step)
{
  std::vector<char> name(c_[2].var.begin(), c_[2].var.end());
  name.push_back('\0');
  const int ivar = input->variable->find(name.data());
  if (ivar < 0 || !input->variable->equalstyle(ivar))
    error->fix_error(FLERR, this, "aveforce z variable must exist and be equal-style");
  const bigint keep = update->ntimestep;
  update->ntimestep = step;
  const double v = input->variable->compute_equal(ivar);
  update->ntimestep = keep;
  return v;
}

int FixAveForceMI355X::register_with(dem_ctx *ctx, int bit, int *id)
{
  int style[3];
  for (int d = 0; d < 3; d++) style[d] = c_[d].style != 0;
  double c0 = c_[2].value, t0 = 0.0, c1 = 0.0;
  if (c_[2].style == 2) {
    // c0 - (time - t0)*c1 from the variable at three steps; a variable that
    // is not linear in time has no device form
    const bigint s = update->ntimestep;
    const double v0 = z_at(s), v1 = z_at(s + 1), v2 = z_at(s + 2);
    const double dt = update->dt;
    t0 = update->atime + (s - update->atimestep) * dt;
    c0 = v0;
    c1 = -(v1 - v0) / dt;
    if (fabs((v2 - v1) - (v1 - v0)) > 1e-9 * (fabs(v0) + fabs(v1) + fabs(v2) + 1e-300))
      error->fix_error(FLERR, this, "aveforce z variable is not linear in time: no device form");
  }
  return dem_add_aveforce(ctx, bit, style, c_[0].value, c_[1].value, c0, t0, c1, id);
}
