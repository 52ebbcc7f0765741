/* fix setforce/mi355x and fix aveforce/mi355x -- the group force fixes of
 * the INL scripts, applied on the device after the walls, gravity and
 * freeze (dem_add_setforce, dem_add_aveforce).  Both run the reference
 * constructor unchanged (argument checks, region keyword) and keep the
 * component values that FixSetForce/FixAveForce hold privately
 * (fix_setforce.cpp:68-110, fix_aveforce.cpp:68-110): NULL, a constant,
 * or v_<name>.  setforce takes constants only.  aveforce takes constant x
 * and y, and a z that is a constant or an equal-style variable linear in
 * simulation time (c0 - (time - t0)*c1, as the INL examples write it).
 */
#ifdef FIX_CLASS

FixStyle(setforce/mi355x,FixSetForceMI355X)
FixStyle(aveforce/mi355x,FixAveForceMI355X)

#else

#ifndef LMP_FIX_FORCE_OPS_MI355X_H
#define LMP_FIX_FORCE_OPS_MI355X_H

#include "fix_aveforce.h"
#include "fix_setforce.h"
#include "liggghts_dem.h"

#include <string>

namespace LAMMPS_NS {

// one component argument of setforce/aveforce
struct ForceComponentMI355X {
  int style;            // 0 NULL, 1 constant, 2 v_<name>
  double value;
  std::string var;
  void parse(class LAMMPS *lmp, const char *arg);
};

class FixSetForceMI355X : public FixSetForce {
 public:
  FixSetForceMI355X(class LAMMPS *, int, char **);
  int setmask() { return 0; }
  // dem_set_group(bit) with the group's members must precede this
  int register_with(dem_ctx *ctx, int bit, int *id) const;
 private:
  ForceComponentMI355X c_[3];
};

class FixAveForceMI355X : public FixAveForce {
 public:
  FixAveForceMI355X(class LAMMPS *, int, char **);
  int setmask() { return 0; }
  int register_with(dem_ctx *ctx, int bit, int *id);
 private:
  ForceComponentMI355X c_[3];
  double z_at(bigint step);
};

}

#endif
#endif
