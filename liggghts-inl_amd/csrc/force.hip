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
// k_force_q: the same pass with a wave-level contact queue (Hertz chains,
// the default).  Both kernels end in one per-particle epilogue (planes,
// gravity, freeze, the force and torque written) that also integrates when
// nothing adds to the forces after the pass, writing the new state to the
// other state buffers.
//
// k_integrate: FixNVESphere::initial_integrate / final_integrate
// (fix_nve_sphere.cpp:134-243) with dtf = 0.5*dt, INERTIA = 0.4; mode 3 runs
// final(n) and initial(n+1) in one pass over the state.
#include "contact_law.h"
#include "kernels.h"

#include <stdlib.h>

#include <algorithm>

#include <type_traits>

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

// Slots [0, lim) of one entry's history into h, zeros past lim (all zeros
// when !known).  vec: the entry starts 16-byte aligned and lim is even, so
// the slots move as 16-byte pairs and a lane's history reaches the L2 in
// half the requests (the entries a wave touches are scattered, so every
// lane's access is a request of its own).
template <int HN>
__device__ inline void hist_load(const double *hp, int lim, bool vec, bool known, double (&h)[HN])
{
    static_assert(HN % 2 == 0, "pairs");
    if (vec) {
        const double2 *h2 = reinterpret_cast<const double2 *>(hp);
#pragma unroll
        for (int q = 0; q < HN / 2; q++) {
            double2 v = make_double2(0.0, 0.0);
            if (known && 2 * q < lim) v = h2[q];
            h[2 * q] = v.x;
            h[2 * q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < HN; q++) h[q] = (known && q < lim) ? hp[q] : 0.0;
    }
}

template <int HN>
__device__ inline void hist_store(double *hp, int lim, bool vec, const double (&h)[HN])
{
    static_assert(HN % 2 == 0, "pairs");
    if (vec) {
        double2 *h2 = reinterpret_cast<double2 *>(hp);
#pragma unroll
        for (int q = 0; q < HN / 2; q++)
            if (2 * q < lim) h2[q] = make_double2(h[2 * q], h[2 * q + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < HN; q++)
            if (q < lim) hp[q] = h[q];
    }
}

// What the owner of particle i adds after its pair sum: the primitive
// planes in the order the fixes were defined, gravity, fix freeze; then the
// force, torque and optional per-atom outputs are written.
template <bool MIXED, int WN>
__device__ inline void particle_epilogue(const ForceArgs &a, int i, const double4 pi, const double4 vmi,
                                         const double4 wi, int itype, bool fzi, const double *mat, double F[3],
                                         double T[3], int nc)
{
    const Model &M = a.M;
    // primitive planes, in the order the fixes were defined
    double pw = 0.0;    // force of the observed plane along its axis
    if (a.P.n) {
        const uint32_t wm = a.wmask[i];
        const double cdmw = M.cdf - 1.0;
        for (int k = 0; k < a.P.n; k++) {
            if (!(wm & (1u << k))) continue;
            const int d = a.P.dim[k];
            const double p = a.P.pos[k];
            const double radi = pi.w;
            // a blend rather than an index, so the coordinate stays in
            // registers (x*1 + y*0 + z*0 is x exactly)
            const double pos = pi.x * (double)(d == 0) + pi.y * (double)(d == 1) + pi.z * (double)(d == 2);
            const double dd = p - pos;
            const double delta[3] = {d == 0 ? dd : 0.0, d == 1 ? dd : 0.0, d == 2 ? dd : 0.0};
            const double deltan = (pos > p) ? pos - p - radi : p - pos - radi;
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
                    const bool vec = (wdn & 1) == 0;
                    double h[WHN];
                    hist_load(hp, wdn, vec, true, h);
                    double dF[3], dT[3];
                    if (MIXED && WN == NM_HERTZ) contact_mixed(M, mat, cw, h, a.shearupdate, dF, dT);
                    else contact_fp64<WN>(M, M.lw, mat, cw, h, a.shearupdate, dF, dT);
                    if (a.shearupdate || WN != NM_HERTZ) hist_store(hp, wdn, vec, h);
                    F[0] += dF[0]; F[1] += dF[1]; F[2] += dF[2];
                    T[0] += dT[0]; T[1] += dT[1]; T[2] += dT[2];
                    if (k == a.obs_plane) pw = d == 0 ? dF[0] : (d == 1 ? dF[1] : dF[2]);
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
    // a pass that also runs the next step's initial_integrate is never the
    // last of a dem_step call, and nothing after it reads the force and
    // torque (the next pass recomputes them): they are written only by the
    // other passes
    if (a.fuse != 3) {
        a.f4[i] = make_double4(F[0], F[1], F[2], 0.0);
        a.t4[i] = make_double4(T[0], T[1], T[2], 0.0);
    }
    if (a.ncontact) a.ncontact[i] = nc;
    if (a.pwf) a.pwf[i] = pw;
}

// final_integrate(n) and, with fuse = 3, initial_integrate(n+1) of one
// particle from its state and its final force and torque: the operations of
// k_integrate, written to the other state buffers
__device__ inline void integrate_into(const ForceArgs &a, int i, double4 x, double4 v, double4 w, const double F[3],
                                      const double T[3])
{
    // dtfm = dtf / (rmass*(1.+CAddRhoFluid/density)), CAddRhoFluid = 0
    const double dtfm = a.dtf / (v.w * (1. + 0.0 / 1.0));
    const double dtirotate = a.dtfrotate / (x.w * x.w * v.w);
    if (a.fuse & 2) {
        v.x += dtfm * F[0]; v.y += dtfm * F[1]; v.z += dtfm * F[2];
        w.x += dtirotate * T[0]; w.y += dtirotate * T[1]; w.z += dtirotate * T[2];
    }
    if (a.fuse & 1) {
        v.x += dtfm * F[0]; v.y += dtfm * F[1]; v.z += dtfm * F[2];
        x.x += a.dtv * v.x; x.y += a.dtv * v.y; x.z += a.dtv * v.z;
        w.x += dtirotate * T[0]; w.y += dtirotate * T[1]; w.z += dtirotate * T[2];
    }
    a.xn[i] = x;
    a.vn[i] = v;
    a.wn[i] = w;
}

template <bool MIXED, int G, int NM, int WN, bool CH = false>
__global__ void __launch_bounds__(kBlock) k_force(ForceArgs a)
{
    if (a.halt && *a.halt) return;
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
            const bool vec = ((dnum | hdn) & 1) == 0;
            hist_load(hp, hdn, vec, known, h);
            double dF[3], dT[3];
            if (MIXED && NM == NM_HERTZ) contact_mixed(M, mat, c, h, a.shearupdate, dF, dT);
            else if (CH) contact_fp64<NM, true>(M, M.lp, mat, c, h, a.shearupdate, dF, dT, coh,
                                                bonded ? M.cohesion : CM_OFF);
            else contact_fp64<NM>(M, M.lp, mat, c, h, a.shearupdate, dF, dT);
            // the hysteretic models write their state even in set-up passes
            if (a.shearupdate || !known || NM != NM_HERTZ) hist_store(hp, hdn, vec, h);
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

    particle_epilogue<MIXED, WN>(a, i, pi, vmi, wi, itype, fzi, mat, F, T, nc);
    if (a.fuse) integrate_into(a, i, pi, vmi, wi, F, T);
}

// ---------------------------------------------------------------------------
// k_force_q: the same pass for Hertz pair and wall chains with a wave-level
// contact queue.  A wave owns a tile of P consecutive particles, whose rows
// form one contiguous range of the list.  The 64 lanes walk that range 64
// entries at a time, test overlap, and append the touching entries (ballot
// and popcount, in entry order) to a queue in LDS; every 64 queued contacts
// are evaluated together, one per lane, so no lane idles on an entry that
// does not touch and the partner gathers of a whole batch are in flight at
// once.  Results go to LDS and each particle of the tile adds its own in
// queue order, which is its list order: a sequential sum per particle,
// deterministic, no atomics.
constexpr int kQWaves = kBlock / 64;

__device__ inline void wave_sync()
{
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool MIXED, int P>
__global__ void __launch_bounds__(kBlock) k_force_q(ForceArgs a)
{
    static_assert(P <= 64, "a tile fits one wave");
    __shared__ int s_rs[kQWaves][P + 1];
    __shared__ double4 s_x[kQWaves][P];
    __shared__ double4 s_v[kQWaves][P], s_w[kQWaves][P];   // the tile's velocity/mass, spin/meta
    __shared__ int s_q[kQWaves][128], s_qo[kQWaves][128];
    __shared__ uint32_t s_qw[kQWaves][128];          // list word of the queued entry
    __shared__ double s_qd[kQWaves][4][128];         // its separation x_i - x_j, partner radius
    // per-contact results; the mixed law's are FP32 values, kept exactly
    using OutT = typename std::conditional<MIXED, float, double>::type;
    __shared__ OutT s_out[kQWaves][6][64];
    __shared__ int s_run0[kQWaves][P], s_run1[kQWaves][P];

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int p0 = (blk * (int)(blockDim.x >> 6) + wv) * P;
    if (p0 >= a.n) return;   // uniform over the wave
    if (a.halt && *a.halt) return;
    const int np = min(P, a.n - p0);
    const Model &M = a.M;
    const int dnum = M.dnum;
    if (lane < np) {
        s_x[wv][lane] = a.xr[p0 + lane];
        s_v[wv][lane] = a.vm[p0 + lane];
        s_w[wv][lane] = a.wt[p0 + lane];
    }
    // row starts of the tile and, from lane 0, the end of its last row (a
    // tile of 64 has no 65th lane)
    if (lane < np) s_rs[wv][lane] = a.nstart[p0 + lane];
    if (lane == 0) s_rs[wv][np] = a.nstart[p0 + np];
    wave_sync();
    const int E0 = s_rs[wv][0], E1 = s_rs[wv][np];
    const double cdm = M.cdf * M.cdf;
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    int nc = 0;
    int qlen = 0;

    // evaluate queue items [0, m): one contact per lane, then each owner adds
    // its run of results
    auto eval_batch = [&](int m) {
        if (lane < np) {
            s_run0[wv][lane] = 0;
            s_run1[wv][lane] = 0;
        }
        wave_sync();
        double dF[3] = {0.0, 0.0, 0.0}, dT[3] = {0.0, 0.0, 0.0};
        if (lane < m) {
            const int e = s_q[wv][lane];
            const int k = s_qo[wv][lane];
            const uint32_t w = s_qw[wv][lane];
            const int j = (int)(w & kJMask);
            // the partner's velocity, spin and the entry's history are
            // independent loads, all in flight together
            double4 vmj, wj;
            if (MIXED && a.rec) {
                const float4 r0 = a.rec[2 * (size_t)j], r1 = a.rec[2 * (size_t)j + 1];
                vmj = make_double4(r0.x, r0.y, r0.z, r0.w);
                wj = make_double4(r1.x, r1.y, r1.z, (double)__float_as_int(r1.w));
            } else {
                vmj = a.vm[j];
                wj = a.wt[j];
            }
            const double radj = s_qd[wv][3][lane];
            const double4 pi = s_x[wv][k], vmi = s_v[wv][k], wi = s_w[wv][k];
            const int metai = (int)wi.w;
            const double *mat = a.mat + (size_t)meta_inst(metai) * kMatStride;
            ContactIn c;
            c.is_wall = false;
            c.itype = meta_type(metai);
            c.radi = pi.w;
            c.mi = vmi.w;
            c.vi[0] = vmi.x; c.vi[1] = vmi.y; c.vi[2] = vmi.z;
            c.wi[0] = wi.x; c.wi[1] = wi.y; c.wi[2] = wi.z;
            c.deltan_in = 0.0;
            c.delta[0] = s_qd[wv][0][lane];
            c.delta[1] = s_qd[wv][1][lane];
            c.delta[2] = s_qd[wv][2][lane];
            const double rsq = c.delta[0] * c.delta[0] + c.delta[1] * c.delta[1] + c.delta[2] * c.delta[2];
            const double r = sqrt(rsq);
            if (r < 1e-20) atomicOr(a.err, 1);
            const int metaj = (int)wj.w;
            c.jtype = meta_type(metaj);
            c.radj = radj;
            c.mj = vmj.w;
            double meff = c.mi * c.mj / (c.mi + c.mj);
            if (M.freeze_on && meta_frozen(metai)) meff = c.mj;
            if (M.freeze_on && meta_frozen(metaj)) meff = c.mi;
            c.meff = meff;
            c.r = r;
            c.rinv = 1.0 / r;
            c.radsum = pi.w + radj;
            c.vj[0] = vmj.x; c.vj[1] = vmj.y; c.vj[2] = vmj.z;
            c.wj[0] = wj.x; c.wj[1] = wj.y; c.wj[2] = wj.z;
            constexpr int HN = chain_max_dnum(NM_HERTZ);
            double h[HN];
            double *hp = a.hist + (size_t)e * dnum;
            const bool known = (w & kFlagBit) != 0;
            const bool vec = (dnum & 1) == 0;
            hist_load(hp, dnum, vec, known, h);
            if (MIXED) contact_mixed(M, mat, c, h, a.shearupdate, dF, dT);
            else contact_fp64<NM_HERTZ>(M, M.lp, mat, c, h, a.shearupdate, dF, dT);
            if (a.shearupdate || !known) hist_store(hp, dnum, vec, h);
            if (!known) a.nj[e] = w | kFlagBit;
            if (lane == 0 || s_qo[wv][lane - 1] != k) s_run0[wv][k] = lane;
            if (lane == m - 1 || s_qo[wv][lane + 1] != k) s_run1[wv][k] = lane + 1;
        }
        s_out[wv][0][lane] = (OutT)dF[0]; s_out[wv][1][lane] = (OutT)dF[1]; s_out[wv][2][lane] = (OutT)dF[2];
        s_out[wv][3][lane] = (OutT)dT[0]; s_out[wv][4][lane] = (OutT)dT[1]; s_out[wv][5][lane] = (OutT)dT[2];
        wave_sync();
        if (lane < np) {
            const int l1 = s_run1[wv][lane];
            for (int l = s_run0[wv][lane]; l < l1; l++) {
                F[0] += (double)s_out[wv][0][l]; F[1] += (double)s_out[wv][1][l]; F[2] += (double)s_out[wv][2][l];
                T[0] += (double)s_out[wv][3][l]; T[1] += (double)s_out[wv][4][l]; T[2] += (double)s_out[wv][5][l];
                nc++;
            }
        }
        wave_sync();
    };

    // software pipeline: the next 64 entries' list words and partner
    // positions are in flight while this group is tested and queued (and
    // while a batch of contacts is evaluated)
    uint32_t w_next = 0;
    double4 pj_next = make_double4(0.0, 0.0, 0.0, 0.0);
    if (E0 + lane < E1) {
        w_next = a.nj[E0 + lane];
        pj_next = a.xr[w_next & kJMask];
    }
    for (int base = E0; base < E1; base += 64) {
        const int e = base + lane;
        bool touch = false;
        int owner = 0;
        uint32_t w = w_next;
        const double4 pj = pj_next;
        if (e + 64 < E1) {
            w_next = a.nj[e + 64];
            pj_next = a.xr[w_next & kJMask];
        }
        double dx = 0.0, dy = 0.0, dz = 0.0, rj = 0.0;
        if (e < E1) {
            int lo = 0, hi = np - 1;   // owner: the last row starting at or before e
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_rs[wv][mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            owner = lo;
            const int code = (int)((w >> 26) & 31u);
            const double4 pi = s_x[wv][owner];
            double xj0 = pj.x, xj1 = pj.y, xj2 = pj.z;
            if (code != kNoShift) {
                const double3 im = image_of(pj, code, M.prd);
                xj0 = im.x; xj1 = im.y; xj2 = im.z;
            }
            dx = pi.x - xj0;
            dy = pi.y - xj1;
            dz = pi.z - xj2;
            const double rsq = dx * dx + dy * dy + dz * dz;
            rj = pj.w;
            const double radsum = pi.w + pj.w;
            touch = rsq < radsum * radsum;
            if (!touch && rsq < cdm * radsum * radsum) {
                // surfacesClose: history reset (Hertz keeps its contact flag)
                contact_close(M.lp, a.hist + (size_t)e * dnum);
            }
        }
        const uint64_t bm = __ballot(touch);
        if (touch) {
            const int pos = qlen + __popcll(bm & ((1ull << lane) - 1ull));
            s_q[wv][pos] = e;
            s_qo[wv][pos] = owner;
            s_qw[wv][pos] = w;
            s_qd[wv][0][pos] = dx;
            s_qd[wv][1][pos] = dy;
            s_qd[wv][2][pos] = dz;
            s_qd[wv][3][pos] = rj;
        }
        qlen += __popcll(bm);
        wave_sync();
        if (qlen >= 64) {
            eval_batch(64);
            const int rest = qlen - 64;
            int tq = 0, to = 0;
            uint32_t tw = 0;
            double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
            if (lane < rest) {
                tq = s_q[wv][64 + lane];
                to = s_qo[wv][64 + lane];
                tw = s_qw[wv][64 + lane];
                t0 = s_qd[wv][0][64 + lane];
                t1 = s_qd[wv][1][64 + lane];
                t2 = s_qd[wv][2][64 + lane];
                t3 = s_qd[wv][3][64 + lane];
            }
            wave_sync();
            if (lane < rest) {
                s_q[wv][lane] = tq;
                s_qo[wv][lane] = to;
                s_qw[wv][lane] = tw;
                s_qd[wv][0][lane] = t0;
                s_qd[wv][1][lane] = t1;
                s_qd[wv][2][lane] = t2;
                s_qd[wv][3][lane] = t3;
            }
            wave_sync();
            qlen = rest;
        }
    }
    if (qlen > 0) eval_batch(qlen);
    if (lane < np) {
        const int i = p0 + lane;
        const double4 pi = s_x[wv][lane], vmi = s_v[wv][lane], wi = s_w[wv][lane];
        const int metai = (int)wi.w;
        const double *mat = a.mat + (size_t)meta_inst(metai) * kMatStride;
        particle_epilogue<MIXED, NM_HERTZ>(a, i, pi, vmi, wi, meta_type(metai), meta_frozen(metai), mat, F, T, nc);
        if (a.fuse) integrate_into(a, i, pi, vmi, wi, F, T);
    }
}

static int padded_blocks(long long threads, int block = kBlock)
{
    long long nb = (threads + block - 1) / block;
    return (int)((nb + 7) & ~7LL);
}

// Shape of the pass: ForceArgs::lanes/block (set by the autotuner or by
// hand), else DEM_CONTACT_LANES, else the default.  lanes > 0: lane groups
// of that width per particle (k_force); lanes < 0: contact queue with tiles
// of -lanes particles per wave (k_force_q, Hertz chains).  Default: queue
// tiles of 32 with 256-thread blocks, the fastest measured on MI355X for
// the 1M-sphere column (0.33 ms per pass against 0.41 ms for 4-lane groups,
// FP64 0.33 against 0.44).
static int contact_shape(const ForceArgs &a)
{
    if (a.lanes) return a.lanes;
    static int g = 0;
    if (g == 0) {
        const char *s = getenv("DEM_CONTACT_LANES");
        g = s ? atoi(s) : -32;
        if (g != 1 && g != 2 && g != 4 && g != 8 && g != 16 && g != -8 && g != -16 && g != -32 && g != -64) g = -32;
    }
    return g;
}

bool force_can_fuse(const ForceArgs &a) { return a.M.cohesion == CM_OFF; }

// Hertz pairs with Hertz walls (the common case) take any group width;
// other model combinations run with 4 lanes per particle.
template <bool MIXED, int NM, int WN>
static void launch_nm(const ForceArgs &a, hipStream_t s)
{
    const int B = a.block ? a.block : kBlock;
    const int shape = contact_shape(a);
    if constexpr (NM == NM_HERTZ && WN == NM_HERTZ) {
        if (shape < 0) {
            const int P = -shape;
            const long long tiles = ((long long)a.n + P - 1) / P;
            const int nb = padded_blocks(tiles * 64, B);
            switch (P) {
            case 8: k_force_q<MIXED, 8><<<nb, B, 0, s>>>(a); return;
            case 32: k_force_q<MIXED, 32><<<nb, B, 0, s>>>(a); return;
            case 64: k_force_q<MIXED, 64><<<nb, B, 0, s>>>(a); return;
            default: k_force_q<MIXED, 16><<<nb, B, 0, s>>>(a); return;
            }
        }
    }
    const int G = (NM == NM_HERTZ && WN == NM_HERTZ && shape > 0) ? shape : 4;
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

// FixMove::initial_integrate / final_integrate, style linear
// (fix_move.cpp:430-490, :900-940): prescribed components follow
// x = x_origin + v * (step - time_origin) * dt with v fixed; the others are
// integrated from the force like fix nve.  It runs before nve/sphere, as
// the fix that the INL scripts define first.
__device__ inline void move_final(const IntegArgs &a, double dtfm, const double4 &f, double4 &v)
{
    if (!(a.mv_style & 1)) v.x += dtfm * f.x;
    if (!(a.mv_style & 2)) v.y += dtfm * f.y;
    if (!(a.mv_style & 4)) v.z += dtfm * f.z;
}

__device__ inline void move_initial(const IntegArgs &a, int i, double dtfm, const double4 &f, double4 &x, double4 &v)
{
    const double4 xo = a.xo[i];
    if (a.mv_style & 1) { v.x = a.mv_val[0]; x.x = xo.x + a.mv_val[0] * a.mv_delta; }
    else { v.x += dtfm * f.x; x.x += a.dtv * v.x; }
    if (a.mv_style & 2) { v.y = a.mv_val[1]; x.y = xo.y + a.mv_val[1] * a.mv_delta; }
    else { v.y += dtfm * f.y; x.y += a.dtv * v.y; }
    if (a.mv_style & 4) { v.z = a.mv_val[2]; x.z = xo.z + a.mv_val[2] * a.mv_delta; }
    else { v.z += dtfm * f.z; x.z += a.dtv * v.z; }
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_integrate(IntegArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long step = 0;
    if (a.ctl) {
        // a checked period: nothing runs once it halted at or before this
        // step (a later halt is the one this pass may record itself)
        step = a.ctl[a.par];
        const long long h = a.ctl[2];
        if (h != 0 && h <= step) {
            // the counter still advances, so that a later recorded step
            // reads a number past the halt
            if (i == 0) a.ctl[a.par ^ 1] = step + 1;
            return;
        }
        for (int m = 0; m < a.nmove; m++)
            if (i < a.mntri[m]) move_tri(a.mtri[m][i], a.mdx[m][0], a.mdx[m][1], a.mdx[m][2]);
        if (i == 0) a.ctl[a.par ^ 1] = step + 1;
    } else if (a.halt && *a.halt) {
        return;
    }
    if (i >= a.n) return;
    double4 x = a.xr[i];
    double4 v = a.vm[i];
    double4 w = a.wt[i];
    const double4 f = a.f4[i];
    const double4 t = a.t4[i];
    // dtfm = dtf / (rmass*(1.+CAddRhoFluid/density)), CAddRhoFluid = 0
    const double dtfm = a.dtf / (v.w * (1. + 0.0 / 1.0));
    const double dtirotate = a.dtfrotate / (x.w * x.w * v.w);
    const int gm = a.gm ? a.gm[i] : 0;
    const bool moved = a.mv_bit >= 0 && ((gm >> a.mv_bit) & 1);
    const bool integ = a.integ_bit < 0 || ((gm >> a.integ_bit) & 1);
    if (MODE & 2) {  // final_integrate(n)
        if (moved) move_final(a, dtfm, f, v);
        if (integ) {
            v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
            w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
        }
    }
    if (MODE & 1) {  // initial_integrate(n+1)
        if (moved) move_initial(a, i, dtfm, f, x, v);
        if (integ) {
            v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
            x.x += a.dtv * v.x; x.y += a.dtv * v.y; x.z += a.dtv * v.z;
            w.x += dtirotate * t.x; w.y += dtirotate * t.y; w.z += dtirotate * t.z;
        }
        a.xr[i] = x;
        if (a.ctl) {   // Neighbor::check_distance of the next step
            const double4 b = a.xhold[i];
            const double dx = x.x - b.x, dy = x.y - b.y, dz = x.z - b.z;
            if (dx * dx + dy * dy + dz * dz > a.deltasq)
                atomicCAS((unsigned long long *)&a.ctl[2], 0ull, (unsigned long long)(step + 1));
        }
    }
    a.vm[i] = v;
    a.wt[i] = w;
}

__global__ void __launch_bounds__(kBlock) k_pack_rec(const double4 *vm, const double4 *wt, float4 *rec, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 v = vm[i], w = wt[i];
    rec[2 * (size_t)i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    rec[2 * (size_t)i + 1] = make_float4((float)w.x, (float)w.y, (float)w.z, __int_as_float((int)w.w));
}

void launch_pack_rec(const double4 *vm, const double4 *wt, float4 *rec, int n, hipStream_t s)
{
    if (n == 0) return;
    k_pack_rec<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(vm, wt, rec, n);
}

void launch_integrate(const IntegArgs &a, int mode, hipStream_t s)
{
    if (a.n == 0) return;
    int nt = a.n;
    for (int m = 0; m < a.nmove; m++) nt = nt > a.mntri[m] ? nt : a.mntri[m];
    const int nb = (nt + kBlock - 1) / kBlock;
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

__global__ void __launch_bounds__(kBlock) k_check_halt(const double4 *xr, const double4 *xhold, int n,
                                                      double deltasq, long long *ctl, int par)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long *cc = ctl;
    if (cc[2]) return;
    const long long step = cc[par];
    if (i == 0) ctl[par ^ 1] = step + 1;
    const double4 a = xr[i], b = xhold[i];
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double rsq = dx * dx + dy * dy + dz * dz;
    if (rsq > deltasq) atomicCAS((unsigned long long *)&ctl[2], 0ull, (unsigned long long)step);
}

void launch_check_halt(const double4 *xr, const double4 *xhold, int n, double deltasq, long long *ctl, int par,
                       hipStream_t s)
{
    if (n == 0) return;
    k_check_halt<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(xr, xhold, n, deltasq, ctl, par);
}

void launch_check_distance(const double4 *xr, const double4 *xhold, int n, double deltasq, int *flag,
                           hipStream_t s)
{
    if (n == 0) return;
    k_check_distance<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(xr, xhold, n, deltasq, flag);
}


// ---------------------------------------------------------------------------
// fix setforce / fix aveforce.  k_group_sum: each thread walks particles i,
// i + stride, ... in order, a fixed tree sums the block, one partial per
// block; setforce sets its components in the same pass (each particle is read
// and written by one thread).  k_group_total: the partials in block order.
// k_group_apply (aveforce): the average plus the added value.

constexpr int kGroupThreads = 256;

__global__ void __launch_bounds__(kGroupThreads) k_group_sum(GroupForceArgs a)
{
    __shared__ double s_p[4][kGroupThreads];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const unsigned bit = 1u << a.bit;
    for (int i = blockIdx.x * kGroupThreads + threadIdx.x; i < a.n; i += gridDim.x * kGroupThreads) {
        if (!((unsigned)a.gmask[i] & bit)) continue;
        double4 f = a.f4[i];
        acc[0] += f.x;
        acc[1] += f.y;
        acc[2] += f.z;
        acc[3] += 1.0;
        if (!a.aveforce) {
            if (a.style[0]) f.x = a.val[0];
            if (a.style[1]) f.y = a.val[1];
            if (a.style[2]) f.z = a.val[2];
            a.f4[i] = f;
        }
    }
    for (int c = 0; c < 4; c++) s_p[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kGroupThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; c++) s_p[c][threadIdx.x] += s_p[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) a.part[blockIdx.x * 4 + threadIdx.x] = s_p[threadIdx.x][0];
}

__global__ void k_group_total(GroupForceArgs a, int nb)
{
    if (threadIdx.x >= 4) return;
    double t = 0.0;
    for (int b = 0; b < nb; b++) t += a.part[b * 4 + threadIdx.x];
    a.total[threadIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock) k_group_apply(GroupForceArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    if (!((unsigned)a.gmask[i] & (1u << a.bit))) return;
    const double cnt = a.total[3];
    if (cnt == 0.0) return;
    const int ncount = (int)cnt;
    double4 f = a.f4[i];
    if (a.style[0]) f.x = a.total[0] / ncount + a.val[0];
    if (a.style[1]) f.y = a.total[1] / ncount + a.val[1];
    if (a.style[2]) f.z = a.total[2] / ncount + a.val[2];
    a.f4[i] = f;
}

void launch_group_force(const GroupForceArgs &a, hipStream_t s)
{
    if (a.n == 0) return;
    const int nb = (int)std::min<long long>(kGroupBlocks, ((long long)a.n + kGroupThreads - 1) / kGroupThreads);
    k_group_sum<<<nb, kGroupThreads, 0, s>>>(a);
    k_group_total<<<1, 64, 0, s>>>(a, nb);
    if (a.aveforce) k_group_apply<<<(a.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
}

}  // namespace dem
