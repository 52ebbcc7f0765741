// mesh_host.cpp -- triangle mesh geometry and topology on the host (once per
// mesh, at set-up).  Same per-element arithmetic as the reference so the
// device sees bit-identical normals, edge vectors and active flags.
#include "mesh_host.h"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace dem {

namespace {

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double mag3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
// vectorScalarDiv3D: multiply by the reciprocal (vector_liggghts.h)
inline void sdiv3(double *v, double s)
{
    const double sinv = 1. / s;
    v[0] = sinv * v[0];
    v[1] = sinv * v[1];
    v[2] = sinv * v[2];
}

// MathExtraLiggghts::compDouble (math_extra_liggghts.h:560-571)
inline bool comp_double(double a, double b, double prec)
{
    if (a == b) return true;
    if (b == 0) return a < prec && a > -prec;
    const double x = a - b;
    return x < prec && x > -prec;
}

inline bool nodes_equal(const double *a, const double *b, double prec)
{
    for (int i = 0; i < 3; i++)
        if (!comp_double(a[i], b[i], prec)) return false;
    return true;
}

struct Topo {
    MeshHost &m;
    const double *N(int t, int k) const { return &m.node[9 * t + 3 * k]; }
    const double *EV(int t, int k) const { return &m.evec[9 * t + 3 * k]; }
    const double *EN(int t, int k) const { return &m.enorm[9 * t + 3 * k]; }

    bool spheres_touch(int a, int b) const
    {
        double d[3];
        for (int q = 0; q < 3; q++) d[q] = m.center[3 * a + q] - m.center[3 * b + q];
        const double rs = m.rbound[a] + m.rbound[b];
        return dot3(d, d) <= rs * rs;
    }

    // MultiNodeMesh::share2Nodes (multi_node_mesh_I.h:281-325)
    bool share2(int a, int b, int &ia1, int &ja1, int &ia2, int &ja2) const
    {
        if (!spheres_touch(a, b)) return false;
        int ns = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (nodes_equal(N(a, i), N(b, j), m.precision)) {
                    if (ns == 0) {
                        ia1 = i;
                        ja1 = j;
                    } else {
                        ia2 = i;
                        ja2 = j;
                        return true;
                    }
                    ns++;
                }
        return false;
    }

    // SurfaceMesh::handleSharedEdge (surface_mesh_I.h:1064-1130), a > b
    void shared_edge(int i, int ie, int j, int je, bool coplanar)
    {
        if (m.nneigh[i] < kMeshNeighMax) m.neigh[kMeshNeighMax * i + m.nneigh[i]] = j;
        if (m.nneigh[j] < kMeshNeighMax) m.neigh[kMeshNeighMax * j + m.nneigh[j]] = i;
        m.nneigh[i]++;
        m.nneigh[j]++;
        bool overlap = false;
        if (coplanar) {
            // coplanarNeighsOverlap (surface_mesh_I.h:965-1004)
            double vi[3], vj[3];
            const double *pref = N(i, ie);
            const double *en = EN(i, ie);
            for (int q = 0; q < 3; q++) {
                vi[q] = N(i, (ie + 2) % 3)[q] - pref[q];
                vj[q] = N(j, (je + 2) % 3)[q] - pref[q];
            }
            overlap = dot3(vi, en) * dot3(vj, en) > 0.;
        }
        if (!coplanar || overlap) {
            if (i < j) {
                m.eact[3 * i + ie] = 0;
                m.eact[3 * j + je] = 1;
            } else {
                m.eact[3 * i + ie] = 1;
                m.eact[3 * j + je] = 0;
            }
        } else {
            m.eact[3 * i + ie] = 0;
            m.eact[3 * j + je] = 0;
        }
    }

    int contains_node(int t, const double *p) const
    {
        for (int k = 0; k < 3; k++) {
            const double *n = N(t, k);
            if (comp_double(n[0], p[0], m.precision) && comp_double(n[1], p[1], m.precision)
                && comp_double(n[2], p[2], m.precision))
                return k;
        }
        return -1;
    }

    // checkNodeRecursive (surface_mesh_I.h:1188-1236), iterative with the
    // same visiting rule (each element once, neighbours of elements that hold
    // the node are expanded)
    std::vector<int> vis_mark;
    int vis_epoch = 0;
    std::vector<int> stack, has;
    std::vector<double> elist, eend;

    void corner(int t, int k, const double boxlo[3], const double boxhi[3])
    {
        double p[3];
        for (int q = 0; q < 3; q++) p[q] = N(t, k)[q];
        vis_epoch++;
        has.clear();
        elist.clear();
        eend.clear();
        bool any_active = false;
        stack.clear();
        stack.push_back(t);
        while (!stack.empty()) {
            const int e = stack.back();
            stack.pop_back();
            if (vis_mark[e] == vis_epoch) continue;
            vis_mark[e] = vis_epoch;
            const int kk = contains_node(e, p);
            if (kk < 0) continue;
            has.push_back(e);
            const int km = (kk - 1 + 3) % 3;
            for (int q = 0; q < 3; q++) elist.push_back(EV(e, kk)[q]);
            for (int q = 0; q < 3; q++) elist.push_back(EV(e, km)[q]);
            for (int q = 0; q < 3; q++) eend.push_back(N(e, (kk + 1) % 3)[q]);
            for (int q = 0; q < 3; q++) eend.push_back(N(e, km)[q]);
            if (m.eact[3 * e + kk] || m.eact[3 * e + km]) any_active = true;
            const int nn = std::min(m.nneigh[e], kMeshNeighMax);
            for (int in = nn - 1; in >= 0; in--) {
                const int nb = m.neigh[kMeshNeighMax * e + in];
                if (nb >= 0) stack.push_back(nb);
            }
        }
        // Domain::is_in_subdomain with the 1e-8 padding of box faces
        // (domain_I.h:65-86): nodes outside keep their initial state
        for (int d = 0; d < 3; d++)
            if (!(p[d] >= boxlo[d] - 1.0e-8 && p[d] < boxhi[d] + 1.0e-8)) return;
        const int ne = 2 * (int)has.size();
        int maxid = -1;
        for (int h : has) maxid = std::max(maxid, h);
        bool colinear = false;
        for (int i = 0; i < ne && !colinear; i++)
            for (int j = i + 1; j < ne; j++)
                if (std::fabs(dot3(&elist[3 * i], &elist[3 * j])) > m.curvature
                    && !nodes_equal(&eend[3 * i], &eend[3 * j], m.precision)) {
                    colinear = true;
                    break;
                }
        if (colinear || !any_active) m.cact[3 * t + k] = 0;
        else m.cact[3 * t + k] = (t == maxid) ? 1 : 0;
    }
};

// element properties: MultiNodeMesh center/rBound, SurfaceMesh
// calcEdgeVecLen, calcSurfaceNorm, calcEdgeNormals, calcObtuseAngleIndex
// (surface_mesh_I.h:300-336, 400-467)
void element_props(MeshHost &m, int t)
{
    const double *nd = &m.node[9 * t];
    double avg[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
        for (int q = 0; q < 3; q++) avg[q] = nd[3 * k + q] + avg[q];
    sdiv3(avg, 3.0);
    for (int q = 0; q < 3; q++) m.center[3 * t + q] = avg[q];
    double rb = 0.;
    for (int k = 0; k < 3; k++) {
        double v[3];
        for (int q = 0; q < 3; q++) v[q] = avg[q] - nd[3 * k + q];
        rb = std::max(rb, mag3(v));
    }
    m.rbound[t] = rb;
    for (int k = 0; k < 3; k++) {
        double *e = &m.evec[9 * t + 3 * k];
        for (int q = 0; q < 3; q++) e[q] = nd[3 * ((k + 1) % 3) + q] - nd[3 * k + q];
        m.elen[3 * t + k] = mag3(e);
        sdiv3(e, m.elen[3 * t + k]);
    }
    const double *ev0 = &m.evec[9 * t], *ev1 = ev0 + 3, *ev2 = ev0 + 6;
    double *sn = &m.snorm[3 * t];
    cross3(ev0, ev1, sn);
    if (mag3(sn) < 1e-15) {
        cross3(ev1, ev2, sn);
        if (mag3(sn) < 1e-15) {
            cross3(ev2, ev0, sn);
            if (mag3(sn) < 1e-15) {
                const double tmp[3] = {1.1233, 2.123231, -3.3343434};
                cross3(ev0, tmp, sn);
            }
        }
    }
    sdiv3(sn, mag3(sn));
    for (int k = 0; k < 3; k++) {
        double *en = &m.enorm[9 * t + 3 * k];
        cross3(&m.evec[9 * t + 3 * k], sn, en);
        if (mag3(en) < 1e-15) {
            for (int q = 0; q < 3; q++) en[q] = m.evec[9 * t + 3 * ((k + 1) % 3) + q];
        } else {
            sdiv3(en, mag3(en));
        }
    }
    // only the last of the three calls decides (node 2)
    m.obtuse[t] = -1;
    for (int k = 0; k < 3; k++) {
        const double dd = dot3(&m.evec[9 * t + 3 * k], &m.evec[9 * t + 3 * ((k + 2) % 3)]);
        m.obtuse[t] = dd > 0. ? k : -1;
    }
}

}  // namespace

void mesh_build(MeshHost &m, int ntri, const double *nodes, double curvature, const double boxlo[3],
                const double boxhi[3])
{
    if (curvature > 0) m.curvature = curvature;
    // MultiNodeMesh::addElement (multi_node_mesh_I.h:138-146) refuses an
    // element with two equal nodes; ids count the accepted ones
    m.node.clear();
    for (int t = 0; t < ntri; t++) {
        const double *a = nodes + 9 * (size_t)t;
        if (nodes_equal(a, a + 3, m.precision) || nodes_equal(a + 3, a + 6, m.precision)
            || nodes_equal(a, a + 6, m.precision))
            continue;
        m.node.insert(m.node.end(), a, a + 9);
    }
    ntri = (int)(m.node.size() / 9);
    m.ntri = ntri;
    m.center.assign(3 * (size_t)ntri, 0.0);
    m.rbound.assign(ntri, 0.0);
    m.evec.assign(9 * (size_t)ntri, 0.0);
    m.elen.assign(3 * (size_t)ntri, 0.0);
    m.enorm.assign(9 * (size_t)ntri, 0.0);
    m.snorm.assign(3 * (size_t)ntri, 0.0);
    m.obtuse.assign(ntri, -1);
    m.nneigh.assign(ntri, 0);
    m.neigh.assign(kMeshNeighMax * (size_t)ntri, -1);
    m.eact.assign(3 * (size_t)ntri, 1);
    m.cact.assign(3 * (size_t)ntri, 1);
    for (int t = 0; t < ntri; t++) element_props(m, t);

    // candidate pairs: bounding spheres that touch, found through a hash grid
    // of centres with cell >= 2 * max bounding radius; each element is
    // compared with all earlier ones in ascending order
    double rmax = 0.0;
    for (int t = 0; t < ntri; t++) rmax = std::max(rmax, m.rbound[t]);
    const double h = std::max(2.0 * rmax, 1e-300) * (1.0 + 1e-9);
    auto cell_of = [&](int t, int64_t c[3]) {
        for (int d = 0; d < 3; d++) c[d] = (int64_t)std::floor(m.center[3 * t + d] / h);
    };
    auto hkey = [](int64_t x, int64_t y, int64_t z) {
        return (uint64_t)((x * 73856093) ^ (y * 19349663) ^ (z * 83492791));
    };
    std::unordered_map<uint64_t, std::vector<int>> grid;
    grid.reserve(2 * (size_t)ntri + 1);
    Topo T{m};
    std::vector<int> cand;
    for (int i = 0; i < ntri; i++) {
        int64_t c[3];
        cell_of(i, c);
        cand.clear();
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    auto it = grid.find(hkey(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        int64_t cj[3];
                        cell_of(j, cj);
                        if (cj[0] == c[0] + dx && cj[1] == c[1] + dy && cj[2] == c[2] + dz) cand.push_back(j);
                    }
                }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (int j : cand) {
            int a1, b1, a2, b2;
            if (!T.share2(i, j, a1, b1, a2, b2)) continue;
            const int ie = (a1 + a2 == 2) ? 2 : std::min(a1, a2);
            const int je = (b1 + b2 == 2) ? 2 : std::min(b1, b2);
            const bool cop = std::fabs(dot3(&m.snorm[3 * i], &m.snorm[3 * j])) >= m.curvature;
            T.shared_edge(i, ie, j, je, cop);
        }
        grid[hkey(c[0], c[1], c[2])].push_back(i);
    }
    T.vis_mark.assign(ntri, 0);
    for (int t = 0; t < ntri; t++)
        for (int k = 0; k < 3; k++) T.corner(t, k, boxlo, boxhi);
}

bool mesh_tri_neighbuild(const MeshHost &m, int t, double r, const double *cs, double treshold)
{
    const double maxDist = r + treshold;
    double v[3];
    for (int q = 0; q < 3; q++) v[q] = cs[q] - m.center[3 * t + q];
    const double dNorm = std::fabs(dot3(&m.snorm[3 * t], v));
    if (dNorm > maxDist) return false;
    const double dParaMax = maxDist * maxDist;
    for (int k = 0; k < 3; k++) {
        for (int q = 0; q < 3; q++) v[q] = cs[q] - m.node[9 * t + 3 * k + q];
        const double d = dot3(&m.enorm[9 * t + 3 * k], v);
        if (d > 0 && d * d > dParaMax) return false;
    }
    return true;
}

void mesh_grid_build(const MeshHost &m, double reach, MeshGrid &g)
{
    const int nt = m.ntri;
    // the passing region of a triangle is its prism offset by reach, whose
    // corners lie reach / sin(half angle) from the nodes
    std::vector<double> margin(nt);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    double rsum = 0.0;
    for (int t = 0; t < nt; t++) {
        double smin = 1.0;
        for (int k = 0; k < 3; k++) {
            const double c = -dot3(&m.evec[9 * t + 3 * k], &m.evec[9 * t + 3 * ((k + 2) % 3)]);
            const double ang = std::acos(std::max(-1.0, std::min(1.0, c)));
            smin = std::min(smin, std::sin(0.5 * ang));
        }
        smin = std::max(smin, 1e-3);
        margin[t] = reach / smin + reach;
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], m.node[9 * t + 3 * k + d] - margin[t]);
                hi[d] = std::max(hi[d], m.node[9 * t + 3 * k + d] + margin[t]);
            }
        rsum += m.rbound[t];
    }
    double h = std::max(reach, nt ? rsum / nt : reach);
    if (!(h > 0)) h = 1.0;
    auto ncell = [&](double hh, int d) { return std::max(1, (int)std::ceil((hi[d] - lo[d]) / hh)); };
    while ((double)ncell(h, 0) * ncell(h, 1) * ncell(h, 2) > 4.0e6) h *= 1.5;
    for (int d = 0; d < 3; d++) {
        g.lo[d] = lo[d];
        g.n[d] = ncell(h, d);
        g.inv[d] = 1.0 / h;
    }
    const double hd = 0.5 * h * std::sqrt(3.0) * (1.0 + 1e-9);
    const size_t ncells = (size_t)g.n[0] * g.n[1] * g.n[2];
    std::vector<std::vector<int>> cells(ncells);
    for (int t = 0; t < nt; t++) {
        int c0[3], c1[3];
        for (int d = 0; d < 3; d++) {
            double a = 1e300, b = -1e300;
            for (int k = 0; k < 3; k++) {
                a = std::min(a, m.node[9 * t + 3 * k + d]);
                b = std::max(b, m.node[9 * t + 3 * k + d]);
            }
            c0[d] = std::max(0, (int)std::floor((a - margin[t] - lo[d]) / h));
            c1[d] = std::min(g.n[d] - 1, (int)std::floor((b + margin[t] - lo[d]) / h));
        }
        for (int z = c0[2]; z <= c1[2]; z++)
            for (int y = c0[1]; y <= c1[1]; y++)
                for (int x = c0[0]; x <= c1[0]; x++) {
                    const double cc[3] = {lo[0] + (x + 0.5) * h, lo[1] + (y + 0.5) * h, lo[2] + (z + 0.5) * h};
                    if (mesh_tri_neighbuild(m, t, hd, cc, reach))
                        cells[((size_t)z * g.n[1] + y) * g.n[0] + x].push_back(t);
                }
    }
    g.start.assign(ncells + 1, 0);
    size_t tot = 0;
    for (size_t c = 0; c < ncells; c++) {
        g.start[c] = (int)tot;
        tot += cells[c].size();
    }
    g.start[ncells] = (int)tot;
    g.list.resize(std::max<size_t>(tot, 1));
    for (size_t c = 0; c < ncells; c++)
        std::copy(cells[c].begin(), cells[c].end(), g.list.begin() + g.start[c]);
}

}  // namespace dem
