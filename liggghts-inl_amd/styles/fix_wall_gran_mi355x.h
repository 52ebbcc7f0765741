/* fix wall/gran/mi355x -- the device form of `fix wall/gran`.
 *
 * Accepts the same arguments as FixWallGran (fix_wall_gran.cpp:93-360) for
 * the walls the device evaluates:
 *   fix ID group wall/gran model <normal> tangential history
 *       [rolling_friction <r>] [<setting> <value> ...]
 *       primitive type <T> xplane|yplane|zplane <pos>
 *   fix ID group wall/gran model ... mesh n_meshes <k> meshes <id1> ...
 * The meshes are fix mesh/surface(/mi355x) instances.  Planes are handed to
 * dem_add_plane, meshes to dem_add_mesh (with the velocity of a fix
 * move/mesh/mi355x linear on them) by run_style verlet/mi355x, in fix
 * order.  The fix itself does no host work.
 */
#ifdef FIX_CLASS

FixStyle(wall/gran/mi355x,FixWallGranMI355X)

#else

#ifndef LMP_FIX_WALL_GRAN_MI355X_H
#define LMP_FIX_WALL_GRAN_MI355X_H

#include "fix.h"
#include "liggghts_dem.h"

#include <string>
#include <vector>

namespace LAMMPS_NS {

class FixWallGranMI355X : public Fix {
 public:
  FixWallGranMI355X(class LAMMPS *, int, char **);
  virtual ~FixWallGranMI355X() {}
  int setmask();
  void init();

  // adds this fix's plane or meshes to the context (dem_add_plane and
  // dem_add_mesh index order follow the fix order)
  int register_with(dem_ctx *ctx) const;
  int normal_model() const { return normal_; }
  int limit_force() const { return limit_force_; }
  bool is_mesh() const { return !mesh_ids_.empty(); }
  int n_meshes() const { return (int)mesh_ids_.size(); }

 private:
  int normal_, limit_force_;
  int wtype_, dim_;
  double pos_;
  std::vector<std::string> mesh_ids_;
};

}

#endif
#endif
