// mesh_host.h -- host-side triangle mesh (fix mesh/surface) for mesh walls:
// element geometry, topology (active edges and corners, obtuse index, edge
// neighbours) and the spatial grid used to find candidate triangles.
#pragma once
#include <stdint.h>

#include <vector>

namespace dem {

constexpr int kMeshNeighMax = 5;   // SurfaceMesh<3,5> (tri_mesh.h:67)

struct MeshHost {
    int ntri = 0;
    double curvature = 1. - 0.00001;   // SurfaceMesh default (surface_mesh.h:61)
    double precision = 1e-8;           // MultiNodeMesh default (multi_node_mesh.h:59)
    std::vector<double> node, center, rbound, evec, elen, enorm, snorm;
    std::vector<int> obtuse, nneigh, neigh;
    std::vector<uint8_t> eact, cact;
};

// Element properties of every triangle (center, bounding radius, edge
// vectors and lengths, surface and edge normals, obtuse index), then the
// topology: SurfaceMesh::buildNeighbours + handleCorner
// (surface_mesh_I.h:474-560, 874-1236).  boxlo/boxhi: simulation box (a
// node outside it keeps its corner active, as in one process).
void mesh_build(MeshHost &m, int ntri, const double *nodes, double curvature, const double boxlo[3],
                const double boxhi[3]);

// Uniform grid of triangles for the candidate search: cell c lists (in
// ascending id) every triangle that a sphere centred in c may reach, i.e.
// that passes resolveTriSphereNeighbuild with maxDist = reach + the cell's
// half diagonal at the cell centre (a superset of the exact per-sphere
// test).  reach = max radius * contactDistanceFactor + treshold.
struct MeshGrid {
    double lo[3] = {0, 0, 0}, inv[3] = {1, 1, 1};
    int n[3] = {1, 1, 1};
    std::vector<int> start, list;
};
void mesh_grid_build(const MeshHost &m, double reach, MeshGrid &g);

// resolveTriSphereNeighbuild (tri_mesh_I.h:275-307) on the host
bool mesh_tri_neighbuild(const MeshHost &m, int t, double r, const double *cs, double treshold);

}  // namespace dem
