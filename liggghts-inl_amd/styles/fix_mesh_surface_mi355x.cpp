/* fix mesh/surface/mi355x: the reference mesh fix plus the values the
 * device mesh needs. */
#include "fix_mesh_surface_mi355x.h"

#include "force.h"
#include "tri_mesh.h"

#include <cmath>
#include <cstring>

using namespace LAMMPS_NS;

FixMeshSurfaceMI355X::FixMeshSurfaceMI355X(LAMMPS *lmp, int narg, char **arg) :
  FixMeshSurface(lmp, narg, arg), cos_curvature_(0.0)
{
  for (int i = 3; i + 1 < narg; i++)
    if (strcmp(arg[i], "curvature") == 0) cos_curvature_ = cos(force->numeric(FLERR, arg[i + 1]) * M_PI / 180.);
}

void FixMeshSurfaceMI355X::nodes(std::vector<double> &out)
{
  TriMesh *m = triMesh();
  const int n = m->sizeLocal();
  out.assign(9 * (size_t)n, 0.0);
  for (int t = 0; t < n; t++)
    for (int k = 0; k < 3; k++) m->node(t, k, &out[9 * (size_t)t + 3 * k]);
}
