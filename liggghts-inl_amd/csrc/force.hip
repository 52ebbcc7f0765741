// force.hip -- contact force pass and velocity-Verlet integrator.
//
// k_force: a group of G lanes per particle gathers over the particle's full
// neighbour row (lane g takes entries g, g+G, ...), then the group sums its
// partial forces and torques with a fixed xor-butterfly of shuffles, so every
// force and torque is summed by its owner in a fixed order: no atomics, no
// scatter, results independent of scheduling.  Replaces, per step,
// Verlet::force_clear (verlet.cpp:406-430), PairGran::compute
// (pair_gran_base.h:187-503), FixWallGran::post_force_primitive
// (fix_wall_gran.cpp:988-1120), FixGravity::post_force
// (fix_gravity.cpp:301-352) and FixFreeze::post_force (fix_freeze.cpp:121-144).
// The list is full (both directions), so a pair is evaluated once from each
// side; in the reference the j-side gets -F of the i-side evaluation, which
// for these laws is the same value (the per-side arithmetic is sign
// symmetric), and periodic-image pairs are evaluated from both sides there
// too (neigh_gran.cpp:570-640).
//
// An entry whose contact flag is still clear has never touched, so its
// history is zero; it is neither read nor, unless it touches now, written.
//
// k_integrate: FixNVESphere::initial_integrate / final_integrate
// (fix_nve_sphere.cpp:134-243) with dtf = 0.5*dt, INERTIA = 0.4; mode 3 runs
// final(n) and initial(n+1) in one pass over the state.
#include "contact_law.h"
#include "kernels.h"

#include <stdlib.h>

namespace dem {

// Blocks are dealt to the 8 XCDs round robin; remap so each XCD walks a
// contiguous range of Morton-ordered particles and their neighbours stay in
// that XCD's L2.
__device__ inline int xcd_block(int b, int nb_padded)
{
    const int per = nb_padded >> 3;
    return (b & 7) * per + (b >> 3);
}

template <int G>
__device__ inline double group_sum(double v)
{
    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

template <int G>
__device__ inline int group_sum_int(int v)
{
    for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

template <bool MIXED, int G, int NM, int WN, bool CH = false>
__global__ void __launch_bounds__(kBlock) k_force(ForceArgs a)
{
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tid = blk * blockDim.x + threadIdx.x;
    const int g = threadIdx.x & (G - 1);
    const int i = tid / G;
    const bool live = i < a.n;      // dead groups still join the shuffles
    const int ii = live ? i : a.n - 1;
    const Model &M = a.M;
    const double4 pi = a.xr[ii];
    const double4 vmi = a.vm[ii];
    const double4 wi = a.wt[ii];
    const int metai = (int)wi.w;
    const int itype = meta_type(metai);
    const bool fzi = meta_frozen(metai);
    // material tables of the particle's instance (partners share it)
    const double *mat = a.mat + (size_t)meta_inst(metai) * kMatStride;
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    const int dnum = M.dnum;
    // slots of the contact chain; the cohesion slots after them belong to
    // the bond pass
    const int hdn = (CH && M.lp.off_coh >= 0) ? M.lp.off_coh : dnum;
    const double cdm = M.cdf * M.cdf;
    int nc = 0;

    ContactIn c;
    c.is_wall = false;
    c.itype = itype;
    c.radi = pi.w;
    c.mi = vmi.w;
    c.vi[0] = vmi.x; c.vi[1] = vmi.y; c.vi[2] = vmi.z;
    c.wi[0] = wi.x; c.wi[1] = wi.y; c.wi[2] = wi.z;
    c.deltan_in = 0.0;

    const int e0 = live ? a.nstart[i] : 0;
    const int e1 = live ? a.nstart[i + 1] : 0;
    for (int e = e0 + g; e < e1; e += G) {
        const uint32_t w = a.nj[e];
        const int j = (int)(w & kJMask);
        const int code = (int)((w >> 26) & 31u);
        const double4 pj = a.xr[j];
        double xj0 = pj.x, xj1 = pj.y, xj2 = pj.z;
        if (code != kNoShift) {
            const double3 im = image_of(pj, code, M.prd);
            xj0 = im.x; xj1 = im.y; xj2 = im.z;
        }
        const double delx = pi.x - xj0;
        const double dely = pi.y - xj1;
        const double delz = pi.z - xj2;
        const double rsq = delx * delx + dely * dely + delz * delz;
        const double radj = pj.w;
        const double radsum = pi.w + radj;
        // cohesion force of a bonded entry, computed by the bond pass
        bool bonded = false;
        double coh[6];
        if (CH && (w & kBondBit) && a.bvalid[e]) {
            bonded = true;
            const double *bo = a.bout + (size_t)e * 6;
            for (int q = 0; q < 6; q++) coh[q] = bo[q];
        }
        if (rsq < radsum * radsum) {
            const double r = sqrt(rsq);
            if (r < 1e-20) atomicOr(a.err, 1);
            const double rinv = 1.0 / r;
            const double4 vmj = a.vm[j];
            const double4 wj = a.wt[j];
            const int metaj = (int)wj.w;
            c.jtype = meta_type(metaj);
            c.radj = radj;
            c.mj = vmj.w;
            double meff = c.mi * c.mj / (c.mi + c.mj);
            if (M.freeze_on && fzi) meff = c.mj;
            if (M.freeze_on && meta_frozen(metaj)) meff = c.mi;
            c.meff = meff;
            c.delta[0] = delx; c.delta[1] = dely; c.delta[2] = delz;
            c.r = r; c.rinv = rinv; c.radsum = radsum;
            c.vj[0] = vmj.x; c.vj[1] = vmj.y; c.vj[2] = vmj.z;
            c.wj[0] = wj.x; c.wj[1] = wj.y; c.wj[2] = wj.z;
            constexpr int HN = chain_max_dnum(NM);
            double h[HN];
            double *hp = a.hist + (size_t)e * dnum;
            const bool known = (w & (CH ? (kFlagBit | kBondBit) : kFlagBit)) != 0;
#pragma unroll
            for (int k = 0; k < HN; k++) h[k] = (known && k < hdn) ? hp[k] : 0.0;
            double dF[3], dT[3];
            if (MIXED && NM == NM_HERTZ) contact_mixed(M, mat, c, h, a.shearupdate, dF, dT);
            else if (CH) contact_fp64<NM, true>(M, M.lp, mat, c, h, a.shearupdate, dF, dT, coh,
                                                bonded ? M.cohesion : CM_OFF);
            else contact_fp64<NM>(M, M.lp, mat, c, h, a.shearupdate, dF, dT);
            // the hysteretic models write their state even in set-up passes
            if (a.shearupdate || !known || NM != NM_HERTZ) {
#pragma unroll
                for (int k = 0; k < HN; k++)
                    if (k < hdn) hp[k] = h[k];
            }
            if (!(w & kFlagBit)) a.nj[e] = w | kFlagBit;
            F[0] += dF[0]; F[1] += dF[1]; F[2] += dF[2];
            T[0] += dT[0]; T[1] += dT[1]; T[2] += dT[2];
            nc++;
        } else if (rsq < cdm * radsum * radsum) {
            // surfacesClose: history reset; a hysteretic chain also clears
            // its contact flag, so the entry no longer carries at a rebuild
            double *hp = a.hist + (size_t)e * dnum;
            if (!contact_close(M.lp, hp)) a.nj[e] = w & ~kFlagBit;
            if (CH && bonded) {
                F[0] += coh[0]; F[1] += coh[1]; F[2] += coh[2];
                T[0] += coh[3]; T[1] += coh[4]; T[2] += coh[5];
            }
        }
    }
    if (G > 1) {
        for (int k = 0; k < 3; k++) {
            F[k] = group_sum<G>(F[k]);
            T[k] = group_sum<G>(T[k]);
        }
        nc = group_sum_int<G>(nc);
    }
    if (!live || g != 0) return;

    // primitive planes, in the order the fixes were defined
    double pw = 0.0;    // force of the observed plane along its axis
    if (a.P.n) {
        const uint32_t wm = a.wmask[i];
        const double *xc = &pi.x;
        const double cdmw = M.cdf - 1.0;
        for (int k = 0; k < a.P.n; k++) {
            if (!(wm & (1u << k))) continue;
            const int d = a.P.dim[k];
            const double p = a.P.pos[k];
            const double radi = pi.w;
            const double pos = xc[d];
            double delta[3] = {0.0, 0.0, 0.0};
            double deltan;
            if (pos > p) {
                delta[d] = p - pos;
                deltan = pos - p - radi;
            } else {
                delta[d] = p - pos;
                deltan = p - pos - radi;
            }
            const int wdn = M.wdnum;
            double *hp = a.whist + ((size_t)i * a.P.n + k) * wdn;
            if (deltan > M.cutneighmax) continue;
            if (deltan <= 0 || deltan < cdmw * radi) {
                if (deltan <= 0) {
                    ContactIn cw;
                    cw.is_wall = true;
                    cw.itype = itype;
                    cw.jtype = a.P.wtype[k];
                    cw.radi = radi;
                    cw.radj = 0.0;
                    cw.deltan_in = -deltan;
                    cw.delta[0] = -delta[0]; cw.delta[1] = -delta[1]; cw.delta[2] = -delta[2];
                    cw.r = cw.radi - cw.deltan_in;
                    cw.rinv = 1.0 / cw.r;
                    cw.radsum = cw.radi;
                    cw.mi = vmi.w; cw.mj = 0.0; cw.meff = vmi.w;
                    cw.vi[0] = vmi.x; cw.vi[1] = vmi.y; cw.vi[2] = vmi.z;
                    cw.vj[0] = 0.0; cw.vj[1] = 0.0; cw.vj[2] = 0.0;
                    cw.wi[0] = wi.x; cw.wi[1] = wi.y; cw.wi[2] = wi.z;
                    cw.wj[0] = 0.0; cw.wj[1] = 0.0; cw.wj[2] = 0.0;
                    constexpr int WHN = chain_max_dnum(WN);
                    double h[WHN];
#pragma unroll
                    for (int q = 0; q < WHN; q++) h[q] = q < wdn ? hp[q] : 0.0;
                    double dF[3], dT[3];
                    if (MIXED && WN == NM_HERTZ) contact_mixed(M, mat, cw, h, a.shearupdate, dF, dT);
                    else contact_fp64<WN>(M, M.lw, mat, cw, h, a.shearupdate, dF, dT);
                    if (a.shearupdate || WN != NM_HERTZ) {
#pragma unroll
                        for (int q = 0; q < WHN; q++)
                            if (q < wdn) hp[q] = h[q];
                    }
                    F[0] += dF[0]; F[1] += dF[1]; F[2] += dF[2];
                    T[0] += dT[0]; T[1] += dT[1]; T[2] += dT[2];
                    if (k == a.obs_plane) pw = dF[d];
                } else {
                    contact_close(M.lw, hp);
                }
            } else {
                for (int q = 0; q < wdn; q++) hp[q] = 0.0;
            }
        }
    }
    if (M.gravity_on) {
        const double massone = vmi.w;
        F[0] += massone * M.gacc[0];
        F[1] += massone * M.gacc[1];
        F[2] += massone * M.gacc[2];
    }
    if (M.freeze_on && fzi && !a.defer_freeze) {
        F[0] = F[1] = F[2] = 0.0;
        T[0] = T[1] = T[2] = 0.0;
    }
    a.f4[i] = make_double4(F[0], F[1], F[2], 0.0);
    a.t4[i] = make_double4(T[0], T[1], T[2], 0.0);
    if (a.ncontact) a.ncontact[i] = nc;
    if (a.pwf) a.pwf[i] = pw;
}

static int padded_blocks(long long threads, int block = kBlock)
{
    long long nb = (threads + block - 1) / block;
    return (int)((nb + 7) & ~7LL);
}

// lanes per particle and block size: the ones the autotuner chose
// (ForceArgs::lanes/block), else DEM_CONTACT_LANES (1, 2, 4, 8 or 16),
// default 4 lanes (the fastest on MI355X for rows of ~8 neighbours: 0.53 ms
// vs 0.58 (8), 0.64 (1), 0.67 (16) per 1M-particle pass) and 256 threads
static int lanes_per_particle(const ForceArgs &a)
{
    if (a.lanes) return a.lanes;
    static int g = -1;
    if (g < 0) {
        const char *s = getenv("DEM_CONTACT_LANES");
        g = s ? atoi(s) : 4;
        if (g != 1 && g != 2 && g != 4 && g != 8 && g != 16) g = 4;
    }
    return g;
}

// Hertz pairs with Hertz walls (the common case) take any group width;
// other model combinations run with 4 lanes per particle.
template <bool MIXED, int NM, int WN>
static void launch_nm(const ForceArgs &a, hipStream_t s)
{
    const int G = (NM == NM_HERTZ && WN == NM_HERTZ) ? lanes_per_particle(a) : 4;
    const int B = a.block ? a.block : kBlock;
    const int nb = padded_blocks((long long)a.n * G, B);
    if constexpr (NM == NM_HERTZ && WN == NM_HERTZ) {
        switch (G) {
        case 1: k_force<MIXED, 1, NM, WN><<<nb, B, 0, s>>>(a); return;
        case 2: k_force<MIXED, 2, NM, WN><<<nb, B, 0, s>>>(a); return;
        case 8: k_force<MIXED, 8, NM, WN><<<nb, B, 0, s>>>(a); return;
        case 16: k_force<MIXED, 16, NM, WN><<<nb, B, 0, s>>>(a); return;
        default: break;
        }
    }
    k_force<MIXED, 4, NM, WN><<<nb, B, 0, s>>>(a);
}

// with a cohesion model: FP64, 4 lanes per particle
template <int NM, int WN>
static void launch_ch(const ForceArgs &a, hipStream_t s)
{
    const int nb = padded_blocks((long long)a.n * 4);
    k_force<false, 4, NM, WN, true><<<nb, kBlock, 0, s>>>(a);
}

template <int NM>
static void launch_ch_wn(const ForceArgs &a, hipStream_t s)
{
    switch (a.M.lw.normal) {
    case NM_HYST1: launch_ch<NM, NM_HYST1>(a, s); break;
    case NM_HYST2: launch_ch<NM, NM_HYST2>(a, s); break;
    default: launch_ch<NM, NM_HERTZ>(a, s); break;
    }
}

template <bool MIXED, int NM>
static void launch_wn(const ForceArgs &a, hipStream_t s)
{
    switch (a.M.lw.normal) {
    case NM_HYST1: launch_nm<MIXED, NM, NM_HYST1>(a, s); break;
    case NM_HYST2: launch_nm<MIXED, NM, NM_HYST2>(a, s); break;
    default: launch_nm<MIXED, NM, NM_HERTZ>(a, s); break;
    }
}

// The mixed form applies to Hertz chains with Hertz-family rolling
// (epsd*); hysteretic chains and cdtnonlinear2 rolling run in FP64.
template <bool MIXED>
static void launch_force(const ForceArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    if (a.M.cohesion != CM_OFF) {
        switch (a.M.lp.normal) {
        case NM_HYST1: launch_ch_wn<NM_HYST1>(a, s); break;
        case NM_HYST2: launch_ch_wn<NM_HYST2>(a, s); break;
        default: launch_ch_wn<NM_HERTZ>(a, s); break;
        }
        return;
    }
    switch (a.M.lp.normal) {
    case NM_HYST1: launch_wn<false, NM_HYST1>(a, s); break;
    case NM_HYST2: launch_wn<false, NM_HYST2>(a, s); break;
    default:
        if (MIXED && a.M.rolling != 4) launch_wn<true, NM_HERTZ>(a, s);
        else launch_wn<false, NM_HERTZ>(a, s);
        break;
    }
}

void launch_force_fp64(const ForceArgs &a, hipStream_t s) { launch_force<false>(a, s); }
void launch_force_mixed(const ForceArgs &a, hipStream_t s) { launch_force<true>(a, s); }

// ---------------------------------------------------------------------------

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_integrate(IntegArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double4 x = a.xr[i];
    double4 v = a.vm[i];
    double4 w = a.wt[i];
    const double4 f = a.f4[i];
    const double4 t = a.t4[i];
    // dtfm = dtf / (rmass*(1.+CAddRhoFluid/density)), CAddRhoFluid = 0
    const double dtfm = a.dtf / (v.w * (1. + 0.0 / 1.0));
    const double dtirotate = a.dtfrotate / (x.w * x.w * v.w);
    if (MODE & 2) {  // final_integrate(n)
        v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
        w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
    }
    if (MODE & 1) {  // initial_integrate(n+1)
        v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
        x.x += a.dtv * v.x; x.y += a.dtv * v.y; x.z += a.dtv * v.z;
        w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
        a.xr[i] = x;
    }
    a.vm[i] = v;
    a.wt[i] = w;
}

void launch_integrate(const IntegArgs &a, int mode, hipStream_t s)
{
    if (a.n == 0) return;
    const int nb = (a.n + kBlock - 1) / kBlock;
    if (mode == 1) k_integrate<1><<<nb, kBlock, 0, s>>>(a);
    else if (mode == 2) k_integrate<2><<<nb, kBlock, 0, s>>>(a);
    else k_integrate<3><<<nb, kBlock, 0, s>>>(a);
}

__global__ void __launch_bounds__(kBlock) k_check_distance(const double4 *xr, const double4 *xhold, int n,
                                                          double deltasq, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 a = xr[i], b = xhold[i];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double rsq = dx * dx + dy * dy + dz * dz;
    if (rsq > deltasq) atomicOr(flag, 1);
}

void launch_check_distance(const double4 *xr, const double4 *xhold, int n, double deltasq, int *flag,
                           hipStream_t s)
{
    if (n == 0) return;
    k_check_distance<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(xr, xhold, n, deltasq, flag);
}

}  // namespace dem
