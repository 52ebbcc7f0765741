/* fix wall/gran/mi355x: argument parsing and hand-over of planes and
 * meshes to the device context. */
#include "fix_wall_gran_mi355x.h"

#include "error.h"
#include "fix_mesh_surface_mi355x.h"
#include "fix_move_mesh_mi355x.h"
#include "force.h"
#include "modify.h"

#include <cstring>

using namespace LAMMPS_NS;
using namespace FixConst;

FixWallGranMI355X::FixWallGranMI355X(LAMMPS *lmp, int narg, char **arg) :
  Fix(lmp, narg, arg), normal_(DEM_NORMAL_HERTZ), limit_force_(0), wtype_(1), dim_(-1), pos_(0.0)
{
  bool have_model = false;
  int iarg = 3;
  while (iarg < narg) {
    const char *kw = arg[iarg];
    if (strcmp(kw, "model") == 0 && iarg + 1 < narg) {
      const char *v = arg[iarg + 1];
      if (strcmp(v, "hertz") == 0) normal_ = DEM_NORMAL_HERTZ, limit_force_ = 0;
      else if (strcmp(v, "hysteretic/nonlinear1") == 0) normal_ = DEM_NORMAL_HYST1, limit_force_ = 1;
      else if (strcmp(v, "hysteretic/nonlinear2") == 0) normal_ = DEM_NORMAL_HYST2, limit_force_ = 1;
      else error->fix_error(FLERR, this, "wall normal model not on the device");
      have_model = true;
      iarg += 2;
    } else if (strcmp(kw, "tangential") == 0 && iarg + 1 < narg) {
      if (strcmp(arg[iarg + 1], "history") != 0) error->fix_error(FLERR, this, "tangential history only");
      iarg += 2;
    } else if ((strcmp(kw, "rolling_friction") == 0 || strcmp(kw, "surface") == 0
                || strcmp(kw, "tangential_damping") == 0 || strcmp(kw, "torsionTorque") == 0)
               && iarg + 1 < narg) {
      // the wall chain's rolling and tangential settings are the pair
      // chain's on the device (one model per context)
      iarg += 2;
    } else if (strcmp(kw, "limitForce") == 0 && iarg + 1 < narg) {
      limit_force_ = strcmp(arg[iarg + 1], "on") == 0 || strcmp(arg[iarg + 1], "yes") == 0;
      iarg += 2;
    } else if (strcmp(kw, "primitive") == 0) {
      // primitive type <T> xplane|yplane|zplane <pos> (fix_wall_gran.cpp:230-290)
      if (iarg + 4 >= narg || strcmp(arg[iarg + 1], "type") != 0)
        error->fix_error(FLERR, this, "expecting 'primitive type <T> <plane> <pos>'");
      wtype_ = force->inumeric(FLERR, arg[iarg + 2]);
      const char *w = arg[iarg + 3];
      if (strcmp(w, "xplane") == 0) dim_ = 0;
      else if (strcmp(w, "yplane") == 0) dim_ = 1;
      else if (strcmp(w, "zplane") == 0) dim_ = 2;
      else error->fix_error(FLERR, this, "primitive walls on the device: xplane, yplane, zplane");
      pos_ = force->numeric(FLERR, arg[iarg + 4]);
      iarg += 5;
    } else if (strcmp(kw, "mesh") == 0) {
      // mesh n_meshes <k> meshes <id> ... (fix_wall_gran.cpp:295-340)
      if (iarg + 3 >= narg || strcmp(arg[iarg + 1], "n_meshes") != 0 || strcmp(arg[iarg + 3], "meshes") != 0)
        error->fix_error(FLERR, this, "expecting 'mesh n_meshes <k> meshes <id> ...'");
      const int k = force->inumeric(FLERR, arg[iarg + 2]);
      if (iarg + 4 + k > narg) error->fix_error(FLERR, this, "not enough mesh ids");
      for (int m = 0; m < k; m++) mesh_ids_.push_back(arg[iarg + 4 + m]);
      iarg += 4 + k;
    } else {
      error->fix_error(FLERR, this, "keyword not on the device path");
    }
  }
  if (!have_model) error->fix_error(FLERR, this, "needs 'model <normal>'");
  if (dim_ < 0 && mesh_ids_.empty()) error->fix_error(FLERR, this, "needs a primitive plane or meshes");
}

int FixWallGranMI355X::setmask()
{
  return 0;   // the device evaluates the wall inside its force passes
}

void FixWallGranMI355X::init()
{
  for (size_t m = 0; m < mesh_ids_.size(); m++) {
    Fix *f = modify->find_fix_id(mesh_ids_[m].c_str());
    if (!f || !dynamic_cast<FixMeshSurfaceMI355X *>(f))
      error->fix_error(FLERR, this, "mesh ids must name fix mesh/surface (run with -sf mi355x)");
  }
}

int FixWallGranMI355X::register_with(dem_ctx *ctx) const
{
  if (dim_ >= 0) {
    const int rc = dem_add_plane(ctx, dim_, pos_, wtype_);
    if (rc) return rc;
  }
  for (size_t m = 0; m < mesh_ids_.size(); m++) {
    FixMeshSurfaceMI355X *fm = dynamic_cast<FixMeshSurfaceMI355X *>(modify->find_fix_id(mesh_ids_[m].c_str()));
    if (!fm) return DEM_ERR_ARG;
    std::vector<double> nodes;
    fm->nodes(nodes);
    // a fix move/mesh/mi355x on this mesh gives its linear velocity
    const double *vel = NULL;
    for (int i = 0; i < modify->nfix; i++) {
      FixMoveMeshMI355X *mv = dynamic_cast<FixMoveMeshMI355X *>(modify->fix[i]);
      if (mv && mv->moves(fm)) vel = mv->velocity();
    }
    const int id = dem_add_mesh(ctx, (int)(nodes.size() / 9), nodes.data(), fm->wall_type(), fm->curvature(), vel);
    if (id < 0) return id;
    const int rc = fm->attach(ctx, id);
    if (rc) return rc;
  }
  return 0;
}
