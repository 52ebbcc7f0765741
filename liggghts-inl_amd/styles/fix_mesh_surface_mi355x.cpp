/* fix mesh/surface/mi355x: the reference mesh fix plus the values the
 * device mesh needs. */
#include "fix_mesh_surface_mi355x.h"

#include "error.h"
#include "force.h"
#include "tri_mesh.h"

#include <cmath>
#include <cstring>

using namespace LAMMPS_NS;

FixMeshSurfaceMI355X::FixMeshSurfaceMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixMeshSurface(lmp, narg, arg), cos_curvature_(0.0), stress_(false), ctx_(NULL), mesh_(-1)
{
  pref_[0] = pref_[1] = pref_[2] = 0.0;
  // the module names are the parts of the style after mesh/surface
  stress_ = strstr(arg[2], "/stress") != NULL;
  for (int i = 3; i + 1 < narg; i++) {
    if (strcmp(arg[i], "curvature") == 0) cos_curvature_ = cos(force->numeric(FLERR, arg[i + 1]) * M_PI / 180.);
    if (strcmp(arg[i], "reference_point") == 0 && i + 3 < narg)
      for (int d = 0; d < 3; d++) pref_[d] = force->numeric(FLERR, arg[i + 1 + d]);
  }
}

void FixMeshSurfaceMI355X::nodes(std::vector<double> &out)
{
  TriMesh *m = triMesh();
  const int n = m->sizeLocal();
  out.assign(9 * (size_t)n, 0.0);
  for (int t = 0; t < n; t++)
    for (int k = 0; k < 3; k++) m->node(t, k, &out[9 * (size_t)t + 3 * k]);
}

int FixMeshSurfaceMI355X::attach(dem_ctx *ctx, int mesh)
{
  ctx_ = ctx;
  mesh_ = mesh;
  return stress_ ? dem_mesh_stress(ctx, mesh, pref_) : 0;
}

double FixMeshSurfaceMI355X::compute_vector(int n)
{
  if (!stress_ || !ctx_) return FixMeshSurface::compute_vector(n);
  if (n < 0 || n >= 9) return 0.0;
  double v[9];
  if (dem_mesh_stress_vector(ctx_, mesh_, v) != 0) error->one(FLERR, dem_last_error(ctx_));
  return v[n];
}
