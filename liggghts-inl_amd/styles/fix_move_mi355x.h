/* fix move/mi355x -- fix move with style linear on the device
 * (dem_add_move_linear): each component given as a number follows
 * x = x_origin + v (t - t_origin) at velocity v, a NULL component is
 * integrated from the force (FixMove, fix_move.cpp:430-490, :900-940).
 * The reference constructor still checks the arguments and keeps its
 * per-atom arrays; the fix sets no host integrate mask.  The component
 * values, which FixMove holds privately, are parsed here from the fix's
 * own arguments.
 */
#ifdef FIX_CLASS

FixStyle(move/mi355x,FixMoveMI355X)

#else

#ifndef LMP_FIX_MOVE_MI355X_H
#define LMP_FIX_MOVE_MI355X_H

#include "fix_move.h"
#include "liggghts_dem.h"

namespace LAMMPS_NS {

class FixMoveMI355X : public FixMove {
 public:
  FixMoveMI355X(class LAMMPS *, int, char **);
  int setmask() { return 0; }
  // dem_set_group(bit) with the group's members must precede this
  int register_with(dem_ctx *ctx, int bit, int *id) const;

 private:
  int style_[3];
  double value_[3];
};

}

#endif
#endif
