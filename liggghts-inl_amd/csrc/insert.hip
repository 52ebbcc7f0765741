// insert.hip -- device side of particle insertion (fix insert/pack).
//
// The reference tests every candidate position of an insertion against the
// particles near the insertion region through a host-side bin list
// (FixInsert::load_xnear, fix_insert.cpp:980-1018; RegionNeighborList::
// hasOverlap, region_neighbor_list_I.h:72-98).  Here the resident particles
// stay in HBM: they are binned into a uniform grid over the box (cells at
// least as wide as the largest overlap distance), and one thread per
// candidate walks its 27 cells, periodic images included, for every
// distinct template radius at once.
//
// k_ins_bin      cell of each resident particle, cell counts
// k_ins_place    particle indices into their cells (order within a cell
//                does not matter: the test is an any-of)
// k_ins_check    rsq <= (r_t + r_j)^2 per candidate and radius, FP64 without
//                contraction, the reference's expression
// k_region_sums  count, volume and mass of the particles inside a block
//                (FixInsertPack::calc_ninsert_this), fixed-order partials
// k_tmpl_inside  inside tests of a multisphere template's Monte Carlo points
// k_extend_rows  empty previous-list rows for appended atoms
// k_stream_release  fix insert/stream's end_of_step: atoms awaiting release
//                ride the extrusion at the normal velocity
#include "kernels.h"

namespace dem {

__device__ inline int ins_cell_dim(double x, double lo, double inv, int n)
{
    int c = (int)floor((x - lo) * inv);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// TriMesh::isInElement (tri_mesh_I.h:345-368): barycentric coordinates of
// pos in triangle t, 1e-10 slack
__device__ inline bool in_element(const double *nd, double p0, double p1, double p2)
{
#pragma clang fp contract(off)
    const double v00 = nd[6] - nd[0], v01 = nd[7] - nd[1], v02 = nd[8] - nd[2];
    const double v10 = nd[3] - nd[0], v11 = nd[4] - nd[1], v12 = nd[5] - nd[2];
    const double v20 = p0 - nd[0], v21 = p1 - nd[1], v22 = p2 - nd[2];
    const double dot00 = v00 * v00 + v01 * v01 + v02 * v02;
    const double dot01 = v00 * v10 + v01 * v11 + v02 * v12;
    const double dot02 = v00 * v20 + v01 * v21 + v02 * v22;
    const double dot11 = v10 * v10 + v11 * v11 + v12 * v12;
    const double dot12 = v10 * v20 + v11 * v21 + v12 * v22;
    const double inv = 1. / (dot00 * dot11 - dot01 * dot01);
    const double u = (dot11 * dot02 - dot01 * dot12) * inv;
    const double v = (dot00 * dot12 - dot01 * dot02) * inv;
    return u > -1e-10 && v > -1e-10 && u + v < 1. + 1e-10;
}

// FixInsertStream::is_nearby (fix_insert_stream.cpp:568-593) and
// RegionNeighborList::isInBoundingBox (region_neighbor_list_I.h:696-709):
// the point x + d n, d the signed normal distance, tested on the face
__device__ inline bool stream_near(const StreamNear &f, double x0, double x1, double x2)
{
#pragma clang fp contract(off)
    const double r0 = x0 - f.pref[0], r1 = x1 - f.pref[1], r2 = x2 - f.pref[2];
    const double d = r0 * f.nrm[0] + r1 * f.nrm[1] + r2 * f.nrm[2];
    if (d > f.dmax) return false;
    if (d < f.dmin) return false;
    const double q0 = x0 + f.nrm[0] * d, q1 = x1 + f.nrm[1] * d, q2 = x2 + f.nrm[2] * d;
    bool on = false;
    for (int t = 0; t < f.ntri && !on; t++) on = in_element(f.face + 9 * t, q0, q1, q2);
    if (!on) return false;
    return x0 >= f.bblo[0] && x0 <= f.bbhi[0] && x1 >= f.bblo[1] && x1 <= f.bbhi[1] && x2 >= f.bblo[2]
           && x2 <= f.bbhi[2];
}

__global__ void __launch_bounds__(kBlock) k_ins_bin(const double4 *__restrict__ xr, int n, InsGrid g, StreamNear f,
                                                    int *cell, int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = xr[i];
    if (f.on && !(g.periodic[0] || g.periodic[1] || g.periodic[2]) && !stream_near(f, p.x, p.y, p.z)) {
        cell[i] = -1;
        return;
    }
    const int cx = ins_cell_dim(p.x, g.lo[0], g.inv[0], g.n[0]);
    const int cy = ins_cell_dim(p.y, g.lo[1], g.inv[1], g.n[1]);
    const int cz = ins_cell_dim(p.z, g.lo[2], g.inv[2], g.n[2]);
    const int c = (cz * g.n[1] + cy) * g.n[0] + cx;
    cell[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void __launch_bounds__(kBlock) k_ins_place(const int *__restrict__ cell, int n, int *cursor, int *idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || cell[i] < 0) return;
    idx[atomicAdd(&cursor[cell[i]], 1)] = i;
}

__global__ void __launch_bounds__(kBlock) k_ins_check(const double *__restrict__ cand, int ncand,
                                                      const double *__restrict__ rad, int nrad,
                                                      const double4 *__restrict__ xr, const int *__restrict__ start,
                                                      const int *__restrict__ idx, InsGrid g, StreamNear f,
                                                      uint32_t *hit)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncand) return;
    // with a periodic dimension the stream filter applies to the position
    // tested, the image as the reference's ghost atom holds it
    const bool fimg = f.on && (g.periodic[0] || g.periodic[1] || g.periodic[2]);
    const double x0 = cand[3 * k], x1 = cand[3 * k + 1], x2 = cand[3 * k + 2];
    const int c0 = ins_cell_dim(x0, g.lo[0], g.inv[0], g.n[0]);
    const int c1 = ins_cell_dim(x1, g.lo[1], g.inv[1], g.n[1]);
    const int c2 = ins_cell_dim(x2, g.lo[2], g.inv[2], g.n[2]);
    const uint32_t all = nrad >= 32 ? 0xffffffffu : ((1u << nrad) - 1u);
    uint32_t h = 0u;
    for (int oz = -1; oz <= 1; oz++) {
        int cz = c2 + oz, sz = 0;
        if (cz < 0 || cz >= g.n[2]) {
            if (!g.periodic[2]) continue;
            sz = cz < 0 ? -1 : 1;
            cz -= sz * g.n[2];
        }
        for (int oy = -1; oy <= 1; oy++) {
            int cy = c1 + oy, sy = 0;
            if (cy < 0 || cy >= g.n[1]) {
                if (!g.periodic[1]) continue;
                sy = cy < 0 ? -1 : 1;
                cy -= sy * g.n[1];
            }
            for (int ox = -1; ox <= 1; ox++) {
                int cx = c0 + ox, sx = 0;
                if (cx < 0 || cx >= g.n[0]) {
                    if (!g.periodic[0]) continue;
                    sx = cx < 0 ? -1 : 1;
                    cx -= sx * g.n[0];
                }
                const int c = (cz * g.n[1] + cy) * g.n[0] + cx;
                for (int q = start[c]; q < start[c + 1]; q++) {
                    const double4 p = xr[idx[q]];
                    // a periodic image as the reference's ghost atom holds
                    // it: x + pbc * prd
                    const double px = sx ? p.x + sx * g.prd[0] : p.x;
                    const double py = sy ? p.y + sy * g.prd[1] : p.y;
                    const double pz = sz ? p.z + sz * g.prd[2] : p.z;
                    if (fimg && !stream_near(f, px, py, pz)) continue;
                    const double d0 = x0 - px, d1 = x1 - py, d2 = x2 - pz;
                    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
                    for (int t = 0; t < nrad; t++) {
                        const double radsum = rad[t] + p.w;
                        if (rsq <= radsum * radsum) h |= 1u << t;
                    }
                    if (h == all) {
                        hit[k] = h;
                        return;
                    }
                }
            }
        }
    }
    hit[k] = h;
}

void launch_ins_bin(const double4 *xr, int n, const InsGrid &g, const StreamNear &f, int *cell, int *count,
                    hipStream_t s)
{
    if (n == 0) return;
    k_ins_bin<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(xr, n, g, f, cell, count);
}

void launch_ins_place(const int *cell, int n, int *cursor, int *idx, hipStream_t s)
{
    if (n == 0) return;
    k_ins_place<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(cell, n, cursor, idx);
}

void launch_ins_check(const double *cand, int ncand, const double *rad, int nrad, const double4 *xr,
                      const int *start, const int *idx, const InsGrid &g, const StreamNear &f, uint32_t *hit,
                      hipStream_t s)
{
    if (ncand == 0) return;
    k_ins_check<<<(ncand + kBlock - 1) / kBlock, kBlock, 0, s>>>(cand, ncand, rad, nrad, xr, start, idx, g, f, hit);
}

// FixInsertStream::end_of_step (fix_insert_stream.cpp:828-921) per atom:
// x = x_ins + v_integrate * ((step - i_step) dt), force and torque zeroed;
// before the release step v = v_integrate and no spin, at it the release
// velocity and omega
__global__ void __launch_bounds__(kBlock) k_stream_release(ReleaseArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    if (a.mask && !(a.gm[i] & a.mask)) return;
    const double4 r0 = a.rel[4 * (size_t)i];
    // MathExtraLiggghts::compDouble(i_step, 0., 1e-13)
    if (r0.w < 1e-13 && r0.w > -1e-13) return;
    const double4 r1 = a.rel[4 * (size_t)i + 1];
    const int i_step = (int)(r0.w + 1e-14);
    const int r_step = (int)(r1.w + 1e-14);
    const int step = (int)a.step;
    if (step > r_step) return;
    const double te = (double)(step - i_step) * a.dt;
    double4 x = a.xr[i];
    x.x = r0.x + r1.x * te;
    x.y = r0.y + r1.y * te;
    x.z = r0.z + r1.z * te;
    a.xr[i] = x;
    a.f4[i] = make_double4(0.0, 0.0, 0.0, a.f4[i].w);
    a.t4[i] = make_double4(0.0, 0.0, 0.0, a.t4[i].w);
    double4 v = a.vm[i], w = a.wt[i];
    if (r_step == step) {
        const double4 vr = a.rel[4 * (size_t)i + 2], wr = a.rel[4 * (size_t)i + 3];
        v.x = vr.x; v.y = vr.y; v.z = vr.z;
        w.x = wr.x; w.y = wr.y; w.z = wr.z;
    } else {
        v.x = r1.x; v.y = r1.y; v.z = r1.z;
        w.x = 0.0; w.y = 0.0; w.z = 0.0;
    }
    a.vm[i] = v;
    a.wt[i] = w;
}

void launch_stream_release(const ReleaseArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    k_stream_release<<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
}

// per block b of kRegionBlocks: part[4b..4b+2] = count, sum of 4/3 pi r^3,
// sum of mass of its grid-stride share, summed in a fixed tree order
__global__ void __launch_bounds__(kBlock) k_region_sums(const double4 *__restrict__ xr,
                                                        const double4 *__restrict__ vm, int n, double lx, double ly,
                                                        double lz, double hx, double hy, double hz, double *part)
{
#pragma clang fp contract(off)
    __shared__ double s[3][kBlock];
    double c = 0.0, vol = 0.0, mass = 0.0;
    const double f4pi3 = 4. * M_PI / 3.;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double4 p = xr[i];
        if (p.x >= lx && p.x <= hx && p.y >= ly && p.y <= hy && p.z >= lz && p.z <= hz) {
            c += 1.0;
            vol += f4pi3 * p.w * p.w * p.w;
            mass += vm[i].w;
        }
    }
    s[0][threadIdx.x] = c;
    s[1][threadIdx.x] = vol;
    s[2][threadIdx.x] = mass;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] += s[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) part[4 * blockIdx.x + q] = s[q][0];
}

void launch_region_sums(const double4 *xr, const double4 *vm, int n, const double *lo, const double *hi,
                        double *part, hipStream_t s)
{
    k_region_sums<<<kRegionBlocks, kBlock, 0, s>>>(xr, vm, n, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], part);
}

// FixTemplateMultiplespheres::calc_center_of_mass's inside test
// (fix_template_multiplespheres.cpp:420-435): point k inside a sphere j when
// ((0 + dx^2) + dy^2) + dz^2 < r_j^2; the spheres pass through LDS in tiles
__global__ void __launch_bounds__(kBlock) k_tmpl_inside(const double *__restrict__ pts, int npts,
                                                        const double *__restrict__ xs, const double *__restrict__ rs,
                                                        int nsph, uint8_t *inside)
{
#pragma clang fp contract(off)
    __shared__ double sx[kBlock], sy[kBlock], sz[kBlock], sr[kBlock];
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    if (k < npts) {
        p0 = pts[3 * (size_t)k];
        p1 = pts[3 * (size_t)k + 1];
        p2 = pts[3 * (size_t)k + 2];
    }
    bool in = false;
    for (int base = 0; base < nsph; base += kBlock) {
        const int j = base + (int)threadIdx.x;
        if (j < nsph) {
            sx[threadIdx.x] = xs[3 * (size_t)j];
            sy[threadIdx.x] = xs[3 * (size_t)j + 1];
            sz[threadIdx.x] = xs[3 * (size_t)j + 2];
            sr[threadIdx.x] = rs[j];
        }
        __syncthreads();
        const int m = min(kBlock, nsph - base);
        if (!in && k < npts)
            for (int t = 0; t < m; t++) {
                double d = 0.;
                d += (p0 - sx[t]) * (p0 - sx[t]);
                d += (p1 - sy[t]) * (p1 - sy[t]);
                d += (p2 - sz[t]) * (p2 - sz[t]);
                if (d < sr[t] * sr[t]) {
                    in = true;
                    break;
                }
            }
        __syncthreads();
    }
    if (k < npts) inside[k] = in ? 1 : 0;
}

void launch_tmpl_inside(const double *pts, int npts, const double *xs, const double *rs, int nsph, uint8_t *inside,
                        hipStream_t s)
{
    if (npts == 0) return;
    k_tmpl_inside<<<(npts + kBlock - 1) / kBlock, kBlock, 0, s>>>(pts, npts, xs, rs, nsph, inside);
}

__global__ void __launch_bounds__(kBlock) k_extend_rows(int *start, int from, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    start[from + 1 + t] = start[from];
}

void launch_extend_rows(int *start, int from, int count, hipStream_t s)
{
    if (count <= 0) return;
    k_extend_rows<<<(count + kBlock - 1) / kBlock, kBlock, 0, s>>>(start, from, count);
}

}  // namespace dem
