/* fix nve/sphere/mi355x -- fix nve/sphere integrated on the device: the
 * velocity-Verlet halves of FixNVESphere (fix_nve_sphere.cpp:134-243) run
 * in the device's last force pass.  The reference constructor still checks
 * the arguments; the fix sets no host integrate mask.  Only the plain form
 * (no `update dipole`, no added mass) has a device form.
 */
#ifdef FIX_CLASS

FixStyle(nve/sphere/mi355x,FixNVESphereMI355X)

#else

#ifndef LMP_FIX_NVE_SPHERE_MI355X_H
#define LMP_FIX_NVE_SPHERE_MI355X_H

#include "error.h"
#include "fix_nve_sphere.h"

namespace LAMMPS_NS {

class FixNVESphereMI355X : public FixNVESphere {
 public:
  FixNVESphereMI355X(class LAMMPS *lmp, int narg, char **arg) : FixNVESphere(lmp, narg, arg)
  {
    if (extra || useAM_) error->fix_error(FLERR, this, "the device integrates plain nve/sphere only");
  }
  int setmask() { return 0; }
};

}

#endif
#endif
