// mesh_geom.h -- sphere/triangle geometry of mesh walls on the device.
//
// TriMesh::resolveTriSphereContactBary (tri_mesh_I.h:65-273) with the
// region split of its barycentric signs, resolveTriSphereNeighbuild
// (tri_mesh_I.h:275-307) and SurfaceMesh::areCoplanarNodeNeighs
// (surface_mesh_I.h:922-962).  Same operation order as the reference, so
// FP64 results are bit-identical.
#pragma once
#include "dem_device.h"

namespace dem {

constexpr double kSmallTri = 1.e-10;     // SMALL_TRIMESH (tri_mesh.h)
constexpr double kLargeTri = 1000000;    // LARGE_TRIMESH

__device__ inline double gdot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// TriMesh::calcDist: delta = closest point - centre
__device__ inline double tri_dist(const double *cs, const double *cp, double *delta)
{
    delta[0] = cp[0] - cs[0];
    delta[1] = cp[1] - cs[1];
    delta[2] = cp[2] - cs[2];
    const double d0 = cs[0] - cp[0], d1 = cs[1] - cp[1], d2 = cs[2] - cp[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

__device__ inline bool edge_on(const TriGeo &t, int k) { return (t.flags >> k) & 1; }
__device__ inline bool corner_on(const TriGeo &t, int k) { return (t.flags >> (3 + k)) & 1; }

// TriMesh::resolveEdgeContactBary
__device__ inline double tri_edge(const TriGeo &t, int ie, const double *p, double *delta, double *bary)
{
    const int ip = (ie + 1) % 3, ipp = (ie + 2) % 3;
    const double *n = &t.node[3 * ie];
    const double *ev = &t.evec[3 * ie];
    const double ntp[3] = {p[0] - n[0], p[1] - n[1], p[2] - n[2]};
    const double dfn = gdot(ntp, ev);
    if (dfn < -kSmallTri) {
        if (!corner_on(t, ie)) return kLargeTri;
        bary[ie] = 1.; bary[ip] = 0.; bary[ipp] = 0.;
        return tri_dist(p, n, delta);
    }
    if (dfn > t.elen[ie] + kSmallTri) {
        if (!corner_on(t, ip)) return kLargeTri;
        bary[ie] = 0.; bary[ip] = 1.; bary[ipp] = 0.;
        return tri_dist(p, &t.node[3 * ip], delta);
    }
    if (!edge_on(t, ie)) return kLargeTri;
    const double cp[3] = {n[0] + dfn * ev[0], n[1] + dfn * ev[1], n[2] + dfn * ev[2]};
    const double d = tri_dist(p, cp, delta);
    bary[ipp] = 0.;
    bary[ie] = 1. - dfn / t.elen[ie];
    bary[ip] = 1. - bary[ie];
    return d;
}

// TriMesh::resolveCornerContactBary
__device__ inline double tri_corner(const TriGeo &t, int in, bool obtuse, const double *p, double *delta,
                                    double *bary)
{
    const int ip = (in + 1) % 3, ipp = (in + 2) % 3;
    const double *n = &t.node[3 * in];
    if (obtuse) {
        const double ntp[3] = {p[0] - n[0], p[1] - n[1], p[2] - n[2]};
        const double *evp = &t.evec[3 * ipp];
        double dfn = gdot(ntp, evp);
        if (dfn < kSmallTri) {
            if (dfn > -t.elen[ipp]) {
                if (!edge_on(t, ipp)) return kLargeTri;
                const double cp[3] = {n[0] + dfn * evp[0], n[1] + dfn * evp[1], n[2] + dfn * evp[2]};
                bary[ip] = 0.;
                bary[in] = 1. + dfn / t.elen[ipp];
                bary[ipp] = 1. - bary[in];
                return tri_dist(p, cp, delta);
            }
            if (!corner_on(t, ipp)) return kLargeTri;
            bary[ipp] = 1.; bary[in] = bary[ip] = 0.;
            return tri_dist(p, &t.node[3 * ipp], delta);
        }
        const double *ev = &t.evec[3 * in];
        dfn = gdot(ntp, ev);
        if (dfn > -kSmallTri) {
            if (dfn < t.elen[in]) {
                if (!edge_on(t, in)) return kLargeTri;
                const double cp[3] = {n[0] + dfn * ev[0], n[1] + dfn * ev[1], n[2] + dfn * ev[2]};
                bary[ipp] = 0.;
                bary[in] = 1. - dfn / t.elen[in];
                bary[ip] = 1. - bary[in];
                return tri_dist(p, cp, delta);
            }
            if (!corner_on(t, ip)) return kLargeTri;
            bary[ip] = 1.; bary[in] = bary[ipp] = 0.;
            return tri_dist(p, &t.node[3 * ip], delta);
        }
    }
    if (!corner_on(t, in)) return kLargeTri;
    bary[in] = 1.; bary[ip] = bary[ipp] = 0.;
    return tri_dist(p, n, delta);
}

// Signed gap (distance - r), delta (closest point - centre), barycentric
// weights and region code (7 face, 3/5/6 edge, 1/2/4 corner).
__device__ inline double tri_contact(const TriGeo &t, double r, const double *cs, double precision,
                                     double *delta, double *bary, int &region)
{
    const double n0c[3] = {cs[0] - t.node[0], cs[1] - t.node[1], cs[2] - t.node[2]};
    // MathExtraLiggghts::calcBaryTriCoords (math_extra_liggghts.h:583-594)
    const double a = gdot(n0c, &t.evec[0]);
    const double b = gdot(n0c, &t.evec[6]);
    const double c = gdot(&t.evec[0], &t.evec[6]);
    const double oneMinCSqr = 1 - c * c;
    bary[1] = (a - b * c) / (t.elen[0] * oneMinCSqr);
    bary[2] = (a * c - b) / (t.elen[2] * oneMinCSqr);
    bary[0] = 1. - bary[1] - bary[2];
    const double invlen = 1. / (2. * t.rbound);
    const double lim = -precision * invlen;
    const int bs = (bary[0] > lim) + 2 * (bary[1] > lim) + 4 * (bary[2] > lim);
    region = bs;
    const int ob = t.obtuse;
    double d;
    switch (bs) {
    case 1: d = tri_corner(t, 0, ob == 0, cs, delta, bary); break;
    case 2: d = tri_corner(t, 1, ob == 1, cs, delta, bary); break;
    case 3: d = tri_edge(t, 0, cs, delta, bary); break;
    case 4: d = tri_corner(t, 2, ob == 2, cs, delta, bary); break;
    case 5: d = tri_edge(t, 2, cs, delta, bary); break;
    case 6: d = tri_edge(t, 1, cs, delta, bary); break;
    case 7: {
        const double dNorm = gdot(t.snorm, n0c);
        const double tmp[3] = {dNorm * t.snorm[0], dNorm * t.snorm[1], dNorm * t.snorm[2]};
        const double csp[3] = {cs[0] - tmp[0], cs[1] - tmp[1], cs[2] - tmp[2]};
        d = tri_dist(cs, csp, delta);
        break;
    }
    default: d = kLargeTri; break;
    }
    return d - r;
}

// resolveTriSphereNeighbuild: may the sphere come within treshold?
__device__ inline bool tri_neighbuild(const TriGeo &t, const double *enorm, double r, const double *cs,
                                      double treshold)
{
    const double maxDist = r + treshold;
    double v[3] = {cs[0] - t.center[0], cs[1] - t.center[1], cs[2] - t.center[2]};
    const double dNorm = fabs(gdot(t.snorm, v));
    if (dNorm > maxDist) return false;
    const double dParaMax = maxDist * maxDist;
    for (int k = 0; k < 3; k++) {
        v[0] = cs[0] - t.node[3 * k];
        v[1] = cs[1] - t.node[3 * k + 1];
        v[2] = cs[2] - t.node[3 * k + 2];
        const double d = gdot(&enorm[3 * k], v);
        if (d > 0 && d * d > dParaMax) return false;
    }
    return true;
}

// MathExtraLiggghts::compDouble
__device__ inline bool gcomp(double a, double b, double prec)
{
    if (a == b) return true;
    if (b == 0) return a < prec && a > -prec;
    const double x = a - b;
    return x < prec && x > -prec;
}

// MultiNodeMesh::nSharedNodes (multi_node_mesh_I.h:328-346)
__device__ inline int tri_nshared(const TriGeo &a, const TriGeo &b, double prec)
{
    const double d[3] = {a.center[0] - b.center[0], a.center[1] - b.center[1], a.center[2] - b.center[2]};
    const double rs = a.rbound + b.rbound;
    if (gdot(d, d) > rs * rs) return 0;
    int ns = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (gcomp(a.node[3 * i], b.node[3 * j], prec) && gcomp(a.node[3 * i + 1], b.node[3 * j + 1], prec)
                && gcomp(a.node[3 * i + 2], b.node[3 * j + 2], prec))
                ns++;
    return ns;
}

// SurfaceMesh::areCoplanarNodeNeighs(a, b): b is an edge neighbour of a or
// shares a node with it, and the normals are parallel within the curvature
__device__ inline bool tri_coplanar_node_neighs(const TriGeo &a, int bid, const TriGeo &b, double curvature,
                                                double prec)
{
    bool neighs = false;
    const int nn = a.nneigh < kMeshNeigh ? a.nneigh : kMeshNeigh;
    for (int i = 0; i < nn; i++)
        if (a.neigh[i] == bid) neighs = true;
    if (!neighs && tri_nshared(a, b, prec) == 0) return false;
    return fabs(gdot(a.snorm, b.snorm)) > curvature;
}

}  // namespace dem
