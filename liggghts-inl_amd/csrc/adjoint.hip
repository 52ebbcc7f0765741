// adjoint.hip -- reverse-mode gradients of the timestep over a window with a
// fixed neighbour list.
//
// Forward (in the engine): the list built at set-up is kept for the whole
// window; where a rebuild would fall, k_adj_mask does to the contact history
// what the rebuild's carry-over does (keep it only for pairs that touch and
// carry a contact flag, neigh_gran.cpp:592-612), and the distance check
// proves every particle stayed within half the skin, so the list still holds
// every pair that can touch.  The physics is that of normal stepping; only
// the order in which a particle's contact forces are summed differs.
//
// Backward, one step n at a time, from the step's force-pass inputs kept on a
// tape (x_n, v'_n, omega'_n, pair and wall history, list flags):
//   v_n = v'_n + dt/2 f_n / m                 final_integrate
//   f_n = Force(x_n, v'_n, omega'_n, H_{n-1}) contact pass
//   x_n = x_{n-1} + dt v'_n, v'_n = v_{n-1} + dt/2 f_{n-1} / m
// k_adj_step gathers, per particle, the vector-Jacobian products of the
// contact law (contact_adj.h) of its own entries (seeded with its force and
// torque adjoints) and of the reverse entries (seeded with the partner's), so
// every adjoint is written by the thread that owns it: no atomics, the same
// order on every run.  The parameter adjoints (Yeff, Geff, betaeff, mu) are
// summed per particle over the window, then per instance.
#define DEM_HD __host__ __device__
#include "contact_adj.h"
#include "contact_law.h"
#include "kernels.h"

namespace dem {

__global__ void __launch_bounds__(kBlock) k_adj_mask(AdjMaskArgs a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const double4 pi = a.xr[i];
    for (int e = a.nstart[i]; e < a.nstart[i + 1]; e++) {
        const uint32_t w = a.nj[e];
        const int j = (int)(w & kJMask);
        const int code = (int)((w >> 26) & 31u);
        const double4 pj = a.xr[j];
        double xj0 = pj.x, xj1 = pj.y, xj2 = pj.z;
        if (code != kNoShift) {
            const double3 im = image_of(pj, code, a.prd);
            xj0 = im.x; xj1 = im.y; xj2 = im.z;
        }
        const double dx = pi.x - xj0, dy = pi.y - xj1, dz = pi.z - xj2;
        const double rsq = dx * dx + dy * dy + dz * dz;
        const double radsum = (pi.w + pj.w) * a.cdf;
        const bool keep = (w & kFlagBit) && rsq < radsum * radsum;
        if (!keep) {
            if (w & kFlagBit) a.nj[e] = w & ~kFlagBit;
            for (int k = 0; k < a.dnum; k++) a.hist[(size_t)e * a.dnum + k] = 0.0;
        }
    }
}

// rev[e] = the entry (j -> i) of an entry (i -> j); an image entry's reverse
// carries the opposite image code
__global__ void __launch_bounds__(kBlock) k_adj_rev(const int *nstart, const uint32_t *nj, int *rev, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int e = nstart[i]; e < nstart[i + 1]; e++) {
        const uint32_t w = nj[e];
        const int j = (int)(w & kJMask);
        const int mirror = 26 - (int)((w >> 26) & 31u);
        int r = -1;
        for (int f = nstart[j]; f < nstart[j + 1]; f++) {
            const uint32_t u = nj[f];
            if ((int)(u & kJMask) == i && (int)((u >> 26) & 31u) == mirror) {
                r = f;
                break;
            }
        }
        rev[e] = r;
    }
}

__device__ inline LawP law_params(const double *mat, int it, int jt)
{
    LawP p;
    p.Y = tbl(mat, M_YEFF, it, jt);
    p.G = tbl(mat, M_GEFF, it, jt);
    p.B = tbl(mat, M_BETA, it, jt);
    p.MU = tbl(mat, M_MU, it, jt);
    return p;
}

__global__ void __launch_bounds__(kBlock) k_adj_step(AdjArgs a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const Model &M = a.M;
    const double4 pi = a.xr[i], vmi = a.vm[i], wti = a.wt[i];
    const int metai = (int)wti.w;
    const int it = meta_type(metai);
    const int inst = meta_inst(metai);
    const double *mat = a.mat + (size_t)inst * kMatStride;
    const double mi = vmi.w, radi = pi.w;
    const double4 lxi = a.lx[i], lvi = a.lv[i], lwi = a.lw[i];
    // force and torque adjoints of this particle (dtf = dt/2, I = 0.4 m r^2)
    const double cf = a.ckick / mi, ct = a.ckick / (0.4 * radi * radi * mi);
    const double gFi[3] = {cf * lvi.x, cf * lvi.y, cf * lvi.z};
    const double gTi[3] = {ct * lwi.x, ct * lwi.y, ct * lwi.z};
    double ax[3] = {0, 0, 0}, av[3] = {0, 0, 0}, aw[3] = {0, 0, 0};
    LawP gq{0, 0, 0, 0};
    const int dn = a.dnum;
    for (int e = a.nstart[i]; e < a.nstart[i + 1]; e++) {
        const uint32_t w = a.nj[e];
        const int j = (int)(w & kJMask);
        const int code = (int)((w >> 26) & 31u);
        const double4 pj = a.xr[j];
        double xj0 = pj.x, xj1 = pj.y, xj2 = pj.z;
        if (code != kNoShift) {
            const double3 im = image_of(pj, code, M.prd);
            xj0 = im.x; xj1 = im.y; xj2 = im.z;
        }
        LawIn in;
        in.d[0] = pi.x - xj0;
        in.d[1] = pi.y - xj1;
        in.d[2] = pi.z - xj2;
        const double rsq = in.d[0] * in.d[0] + in.d[1] * in.d[1] + in.d[2] * in.d[2];
        const double radj = pj.w;
        const double radsum = radi + radj;
        double *lh1 = a.lh1 + (size_t)e * dn;
        const double *lh = a.lh + (size_t)e * dn;
        if (!(rsq < radsum * radsum)) {
            // no contact: the history passes through unchanged
            const bool keep = !a.mask_step || (w & kFlagBit);
            for (int k = 0; k < dn; k++) lh1[k] = keep ? lh[k] : 0.0;
            continue;
        }
        const double4 vmj = a.vm[j], wtj = a.wt[j];
        const int jt = meta_type((int)wtj.w);
        const double mj = vmj.w;
        const double vi[3] = {vmi.x, vmi.y, vmi.z}, vj[3] = {vmj.x, vmj.y, vmj.z};
        const double wi[3] = {wti.x, wti.y, wti.z}, wj[3] = {wtj.x, wtj.y, wtj.z};
        for (int k = 0; k < 3; k++) {
            in.vi[k] = vi[k]; in.vj[k] = vj[k]; in.wi[k] = wi[k]; in.wj[k] = wj[k];
        }
        in.radi = radi;
        in.radj = radj;
        in.radsum = radsum;
        in.reff = radi * radj / (radi + radj);
        in.meff = mi * mj / (mi + mj);
        in.wall = 0;
        // own entry: seeded with this particle's adjoints
        {
            const bool known = (w & kFlagBit) != 0;
            const double *h = a.hist + (size_t)e * dn;
            for (int k = 0; k < 3; k++) in.sh[k] = known ? h[k] : 0.0;
            LawIn g = {};
            law_vjp(in, law_params(mat, it, jt), a.dt, a.update, gFi, gTi, lh, g, gq);
            for (int k = 0; k < 3; k++) {
                ax[k] += g.d[k];
                av[k] += g.vi[k];
                aw[k] += g.wi[k];
            }
            // an entry read as zero history (no flag) does not depend on the
            // stored values; a rebuild's mask zeroes the unflagged ones
            for (int k = 0; k < dn; k++) lh1[k] = known ? g.sh[k] : 0.0;
        }
        // reverse entry: the partner's force, seeded with its adjoints
        const int r = a.rev[e];
        if (r >= 0) {
            const double4 lvj = a.lv[j], lwj = a.lw[j];
            const double cfj = a.ckick / mj, ctj = a.ckick / (0.4 * radj * radj * mj);
            const double gFj[3] = {cfj * lvj.x, cfj * lvj.y, cfj * lvj.z};
            const double gTj[3] = {ctj * lwj.x, ctj * lwj.y, ctj * lwj.z};
            const uint32_t wr = a.nj[r];
            const bool known = (wr & kFlagBit) != 0;
            const double *h = a.hist + (size_t)r * dn;
            LawIn in2 = in;
            for (int k = 0; k < 3; k++) {
                in2.d[k] = -in.d[k];
                in2.vi[k] = vj[k]; in2.vj[k] = vi[k];
                in2.wi[k] = wj[k]; in2.wj[k] = wi[k];
                in2.sh[k] = known ? h[k] : 0.0;
            }
            in2.radi = radj;
            in2.radj = radi;
            LawIn g = {};
            LawP gdummy{0, 0, 0, 0};
            const double *matj = a.mat + (size_t)meta_inst((int)wtj.w) * kMatStride;
            law_vjp(in2, law_params(matj, jt, it), a.dt, a.update, gFj, gTj, a.lh + (size_t)r * dn, g, gdummy);
            for (int k = 0; k < 3; k++) {
                ax[k] -= g.d[k];
                av[k] += g.vj[k];
                aw[k] += g.wj[k];
            }
        }
    }

    // primitive planes (the contact pass's plane loop, backwards)
    if (a.P.n) {
        const uint32_t wm = a.wmask[i];
        const double *xc = &pi.x;
        const double obs = (a.wseed && a.obs_plane >= 0) ? a.wseed[inst] : 0.0;
        for (int k = 0; k < a.P.n; k++) {
            const double *lhw = a.lhw + ((size_t)i * a.P.n + k) * 3;
            double *lhw1 = a.lhw1 + ((size_t)i * a.P.n + k) * 3;
            const int d = a.P.dim[k];
            const double p = a.P.pos[k];
            const double pos = xc[d];
            const double deltan = (pos > p) ? pos - p - radi : p - pos - radi;
            if (!(wm & (1u << k)) || deltan > M.cutneighmax) {
                for (int q = 0; q < 3; q++) lhw1[q] = lhw[q];
                continue;
            }
            if (deltan > 0.0) {   // history zeroed by the pass
                for (int q = 0; q < 3; q++) lhw1[q] = 0.0;
                continue;
            }
            LawIn in = {};
            in.d[d] = pos - p;
            for (int q = 0; q < 3; q++) {
                in.vi[q] = (&vmi.x)[q];
                in.wi[q] = (&wti.x)[q];
                in.sh[q] = a.whist[((size_t)i * a.P.n + k) * a.wdnum + q];
            }
            in.radi = radi;
            in.radj = 0.0;
            in.radsum = radi;
            in.reff = radi;
            in.meff = mi;
            in.wall = 1;
            double gF[3] = {gFi[0], gFi[1], gFi[2]};
            if (k == a.obs_plane) gF[d] += obs;
            LawIn g = {};
            law_vjp(in, law_params(mat, it, a.P.wtype[k]), a.dt, a.update, gF, gTi, lhw, g, gq);
            ax[d] += g.d[d];
            for (int q = 0; q < 3; q++) {
                av[q] += g.vi[q];
                aw[q] += g.wi[q];
                // a set-up pass leaves the plane history as it was
                lhw1[q] = g.sh[q];
            }
        }
    }

    const double lx1[3] = {lxi.x + ax[0], lxi.y + ax[1], lxi.z + ax[2]};
    a.lx1[i] = make_double4(lx1[0], lx1[1], lx1[2], 0.0);
    a.lv1[i] = make_double4(lvi.x + av[0] + a.dt * lx1[0], lvi.y + av[1] + a.dt * lx1[1],
                            lvi.z + av[2] + a.dt * lx1[2], 0.0);
    a.lw1[i] = make_double4(lwi.x + aw[0], lwi.y + aw[1], lwi.z + aw[2], 0.0);
    double *g4 = a.gp + (size_t)i * 4;
    g4[0] += gq.Y;
    g4[1] += gq.G;
    g4[2] += gq.B;
    g4[3] += gq.MU;
}

// seed: lambda_x = w_height[b] / n_b along z for the mean height of instance b
__global__ void __launch_bounds__(kBlock) k_adj_seed(const double4 *wt, const double *wh, const double *count,
                                                    double4 *lx, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int b = meta_inst((int)wt[i].w);
    const double c = count[(size_t)b * 5];
    lx[i] = make_double4(0.0, 0.0, (wh && c > 0.0) ? wh[b] / c : 0.0, 0.0);
}

// per-instance sums of the per-particle parameter adjoints, one block per
// instance over its contiguous range, fixed order
constexpr int kRedThreads = 256;

__global__ void __launch_bounds__(kRedThreads) k_adj_reduce(const int *cstart, int nkl, const double *gp,
                                                           double *out)
{
    __shared__ double s_v[4][kRedThreads];
    const int b = blockIdx.x;
    const int p0 = cstart[(size_t)b * nkl], p1 = cstart[(size_t)(b + 1) * nkl];
    double acc[4] = {0, 0, 0, 0};
    for (int p = p0 + threadIdx.x; p < p1; p += kRedThreads)
        for (int k = 0; k < 4; k++) acc[k] += gp[(size_t)p * 4 + k];
    for (int k = 0; k < 4; k++) s_v[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 4; k++) s_v[k][threadIdx.x] += s_v[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[(size_t)b * 4 + threadIdx.x] = s_v[threadIdx.x][0];
}

static int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

void launch_adj_mask(const AdjMaskArgs &a, hipStream_t s)
{
    if (a.n) k_adj_mask<<<blocks(a.n), kBlock, 0, s>>>(a);
}

void launch_adj_rev(const int *nstart, const uint32_t *nj, int *rev, int n, hipStream_t s)
{
    if (n) k_adj_rev<<<blocks(n), kBlock, 0, s>>>(nstart, nj, rev, n);
}

void launch_adj_step(const AdjArgs &a, hipStream_t s)
{
    if (a.n) k_adj_step<<<blocks(a.n), kBlock, 0, s>>>(a);
}

void launch_adj_seed(const double4 *wt, const double *wh, const double *count, double4 *lx, int n, hipStream_t s)
{
    if (n) k_adj_seed<<<blocks(n), kBlock, 0, s>>>(wt, wh, count, lx, n);
}

void launch_adj_reduce(const int *cstart, int nkl, int ninst, const double *gp, double *out, hipStream_t s)
{
    if (ninst) k_adj_reduce<<<ninst, kRedThreads, 0, s>>>(cstart, nkl, gp, out);
}

}  // namespace dem
