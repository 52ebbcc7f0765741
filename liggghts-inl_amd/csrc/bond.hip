// bond.hip -- cohesion bond / bond/nonlinear of the pair chain on the device.
//
// CohesionModel<COHESION_BOND>::surfacesClose (cohesion_model_bond.h:490-935)
// and CohesionModel<COHESION_BOND_NONLINEAR>::surfacesClose
// (cohesion_model_bond_nonlinear.h:396-960).  The reference evaluates a pair
// once, from its half-list side i < j, and hands j the mirrored result; the
// bond/nonlinear history is not sign-symmetric (its extreme angles keep
// their sign for both partners), so here too one side evaluates: k_bond
// visits each particle's canonical entries (partner with a larger tag, or
// any periodic-image entry, which the reference also evaluates from both
// sides), updates the history kept on that entry and writes the force and
// torque of both sides into per-entry slots.  The contact pass adds them
// where the chain puts the cohesion model (between the normal and the
// tangential model).  No atomics: every entry is written by one thread.
//
// k_bond_rev finds, for each entry (i -> j, no image), the entry (j -> i).
#include "contact_law.h"
#include "kernels.h"

namespace dem {

__device__ inline int bsgn(double v) { return (0.0 < v) - (v < 0.0); }

// The reference's pow(z, 2.0), pow(z, 1.0) and pow(|d|, 0.5) of the
// bond/nonlinear normal law, as z*z, z and a correctly rounded square root.
// The host libm (glibc 2.35) gives z*z and z bit for bit, and sqrt on all
// but about 1 in 1,200 arguments; the device libm's pow is not correctly
// rounded, and the law's stiffness is unbounded at zero stretch, where last
// bits decide a ringing sphere's trajectory.
__device__ inline double sq(double z) { return z * z; }

// vectorProject3D (vector_liggghts.h:437-442)
__device__ inline void bproject(const double *v, const double *on, double *out)
{
    double nrm[3] = {on[0], on[1], on[2]};
    const double norm = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const double invnorm = (norm == 0.) ? 0. : 1. / norm;
    nrm[0] *= invnorm; nrm[1] *= invnorm; nrm[2] *= invnorm;
    const double d = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
    out[0] = d * nrm[0]; out[1] = d * nrm[1]; out[2] = d * nrm[2];
}

__device__ inline double bdamped(double x, double c, double rate, int smooth, double minvel, double dt)
{
    double m;
    if (smooth) {
        const double dv = 0.01 * x * dt;
        m = fmin(1.0, fmax(-1.0, rate / fmax(minvel, dv)));
    } else {
        m = bsgn(rate);
    }
    return x - c * fabs(x) * m;
}

// one component of the bond/nonlinear hysteretic torque
__device__ inline double bnl_torque(double th, double *tmax, double *tmin, double k, double ku, double kc, double Jl)
{
    const double c1 = (ku - k) / (ku + kc) * (*tmax);
    const double c2 = (ku - k) / (ku + kc) * (*tmin);
    double tq;
    if ((th >= *tmax) || (th <= *tmin)) tq = -k * Jl * th;
    else if (th > c1) tq = -ku * Jl * th + (ku - k) * Jl * (*tmax);
    else if (th >= c2) tq = kc * Jl * th;
    else tq = -ku * Jl * th + (ku - k) * Jl * (*tmin);
    if (th > *tmax) *tmax = th;
    if (th < *tmin) *tmin = th;
    return tq;
}

enum { B_NONE = 0, B_FORCE = 1, B_BROKEN = 2 };

// Returns B_FORCE with out = {F, dTorque_i, dTorque_j}, B_BROKEN, or B_NONE
// (no bond and none created).  created: set when the bond forms now.
__device__ inline int bond_eval(const Model &M, const double *mat, int it, int jt, double *hb, const double delta[3],
                                double rsq, double radi, double radj, const double xi[3], const double vi[3],
                                const double vj[3], const double omegai[3], const double omegaj[3], bool upd,
                                long long ntimestep, const double *brid, int i, int j, bool &created,
                                double out[9])
{
    const bool nl = M.cohesion == CM_BOND_NONLINEAR;
    const BondSet &bs = M.bs;
    const double lambda = tbl(mat, M_BLAMBDA, it, jt);
    const double small = nl ? 1.e-15 : 1.e-6;
    created = false;
    if (lambda < small) return B_NONE;
    const double r = sqrt(rsq);
    if (upd) {
        if (hb[0] < small) {
            // create_bond_A, else create_bond_B with a bond_random_id
            // (cohesion_model_bond.h:522-540; compDouble: |a - b| < 1e-13)
            const double rc = tbl(mat, M_BCREATEDIST, it, jt);
            bool create = (bs.create_always || ntimestep == bs.ts_create) && (r < rc);
            if (!create && brid) {
                const double bi = brid[i], bj = brid[j], db = bi - bj;
                create = (double)ntimestep < bi && (bi == bj || (db < 1e-13 && db > -1e-13)) && (r < rc);
            }
            if (!create) return B_NONE;
            created = true;
            hb[0] = 1.0;
            hb[1] = r;
            for (int d = 0; d < 3; d++) hb[2 + d] = xi[d] - delta[d];
            for (int d = 5; d < 14; d++) hb[d] = 0.0;
        }
    } else if (hb[0] < small) {
        return B_NONE;
    }
    double ft[3], tn[3], tt[3];
    for (int d = 0; d < 3; d++) {
        ft[d] = hb[5 + d];
        tn[d] = hb[8 + d];
        tt[d] = hb[11 + d];
    }
    if (!bs.stress_break && r > tbl(mat, M_BMAXDIST, it, jt) && upd) {
        hb[0] = 0.;
        hb[1] = 0.;
        return B_BROKEN;
    }
    const double rinv = 1. / r;
    const double en[3] = {rinv * delta[0], rinv * delta[1], rinv * delta[2]};
    const double rb = lambda * (radi < radj ? radi : radj);
    const double A = M_PI * rb * rb;
    const double J = 0.5 * A * rb * rb;
    const double I = 0.5 * J;
    const double dt = M.dt;
    const double radsuminv = 1. / (radi + radj);
    const double cri = r * radi * radsuminv;
    const double crj = r * radj * radsuminv;
    double vr[3], vn[3], vt[3], wr[3], vtr[3], wn[3], wt[3], tmp1[3], tmp2[3];
    for (int d = 0; d < 3; d++) vr[d] = vi[d] - vj[d];
    bproject(vr, en, vn);
    for (int d = 0; d < 3; d++) vt[d] = vr[d] - vn[d];
    const double fi = radi * radsuminv, fj = radj * radsuminv;
    for (int d = 0; d < 3; d++) {
        tmp1[d] = fi * omegai[d];
        tmp2[d] = fj * omegaj[d];
    }
    for (int d = 0; d < 3; d++) wr[d] = tmp1[d] + tmp2[d];
    tmp1[0] = delta[1] * wr[2] - delta[2] * wr[1];
    tmp1[1] = delta[2] * wr[0] - delta[0] * wr[2];
    tmp1[2] = delta[0] * wr[1] - delta[1] * wr[0];
    for (int d = 0; d < 3; d++) vtr[d] = vt[d] + tmp1[d];
    for (int d = 0; d < 3; d++) wr[d] = omegai[d] - omegaj[d];
    bproject(wr, en, wn);
    for (int d = 0; d < 3; d++) wt[d] = wr[d] - wn[d];

    double nforce[3] = {0., 0., 0.}, nfd[3] = {0., 0., 0.}, tfd[3] = {0., 0., 0.};
    double ntd[3] = {0., 0., 0.}, ttd[3] = {0., 0., 0.};
    const double displacement = hb[1] - r;
    const double minvel = 1e-5 * fmin(radi, radj) / dt;
    const int damp = bs.damping, smooth = bs.damping_smooth;
    double dmax = 0.0, dmin = 0.0;
    if (nl) {
        if (displacement > hb[14]) { hb[14] = displacement; dmax = displacement; }
        else dmax = hb[14];
        if (displacement < hb[27]) { hb[27] = displacement; dmin = displacement; }
        else dmin = hb[27];
        if (displacement < 0.0) hb[14] = 0.0;
        if (displacement > 0.0) hb[27] = 0.0;
    }
    if (displacement < -1.e-15 || displacement > 1.e-15) {
        double frcmag;
        if (!nl) {
            frcmag = tbl(mat, M_BKN, it, jt) * A * displacement;
        } else {
            const double k1 = tbl(mat, M_BKFN1, it, jt), ku1 = tbl(mat, M_BKUFN1, it, jt), kc1 = tbl(mat, M_BKCFN1, it, jt);
            const double k2 = tbl(mat, M_BKFN2, it, jt), ku2 = tbl(mat, M_BKUFN2, it, jt), kc2 = tbl(mat, M_BKCFN2, it, jt);
            const double d = displacement;
            if (!bs.tension_variant) {
                const double c1 = sq((ku1 - k1) / (ku1 + kc1)) * dmax;
                const double c2 = ((ku2 - k2) / (ku2 + kc2)) * dmin;
                if (d < dmin) frcmag = k2 * A * d;
                else if (d < c2) frcmag = ku2 * A * d + (k2 - ku2) * A * dmin;
                else if (d < 0.0) frcmag = -kc2 * A * d;
                else if (d < c1) frcmag = -kc1 * sqrt(fabs(d));
                else if (d < dmax) frcmag = ku1 * sqrt(fabs(d)) + (k1 - ku1) * sqrt(fabs(dmax));
                else frcmag = k1 * sqrt(fabs(d));
            } else {
                // the normal law of examples/LIGGGHTS/INL/cohesive_bond_nonlinear_
                // tension/src/cohesion_model_bond_nonlinear_tension.h:590-613
                const double c1 = ((ku1 - k1) / (ku1 + kc1)) * dmax;
                const double c2 = sq((ku2 - k2) / (ku2 + kc2)) * dmin;
                if (d <= dmin) frcmag = -k2 * sqrt(fabs(d));
                else if (d < c2) frcmag = -(ku2 * sqrt(fabs(d)) + (k2 - ku2) * sqrt(fabs(dmin)));
                else if (d < 0.0) frcmag = kc2 * sqrt(fabs(d));
                else if (d < c1) frcmag = -kc1 * A * d;
                else if (d < dmax) frcmag = ku1 * A * d + (k1 - ku1) * A * dmax;
                else frcmag = k1 * A * d;
            }
        }
        for (int d = 0; d < 3; d++) nforce[d] = frcmag * en[d];
        const double cfn = tbl(mat, M_BDFN, it, jt);
        for (int d = 0; d < 3; d++) nfd[d] = damp ? bdamped(nforce[d], cfn, vn[d], smooth, minvel, dt) : nforce[d];
    }
    {
        const double kft = nl ? tbl(mat, M_BKFT, it, jt) : tbl(mat, M_BKT, it, jt);
        const double c = -kft * A * dt;
        const double dtforce[3] = {c * vtr[0], c * vtr[1], c * vtr[2]};
        bproject(ft, en, tmp1);
        for (int d = 0; d < 3; d++) ft[d] = ft[d] - tmp1[d];
        for (int d = 0; d < 3; d++) ft[d] = ft[d] + dtforce[d];
        const double cft = tbl(mat, M_BDFT, it, jt);
        for (int d = 0; d < 3; d++) tfd[d] = damp ? bdamped(ft[d], cft, vtr[d], smooth, minvel, dt) : ft[d];
    }
    double torque_n[3], torque_t[3];
    const double ctn = tbl(mat, M_BDTN, it, jt), ctt = tbl(mat, M_BDTT, it, jt);
    if (!nl) {
        const double ktp = tbl(mat, M_BKT, it, jt), knp = tbl(mat, M_BKN, it, jt);
        const double cn = -ktp * J * dt;
        const double dnt[3] = {cn * wn[0], cn * wn[1], cn * wn[2]};
        bproject(tn, en, tn);
        for (int d = 0; d < 3; d++) tn[d] = tn[d] + dnt[d];
        for (int d = 0; d < 3; d++) ntd[d] = damp ? tn[d] - ctn * fabs(tn[d]) * bsgn(wn[d]) : tn[d];
        const double wtsq = wt[0] * wt[0] + wt[1] * wt[1] + wt[2] * wt[2];
        if (wtsq > 0) {
            const double ct = -knp * I * dt;
            const double dtt[3] = {ct * wt[0], ct * wt[1], ct * wt[2]};
            bproject(tt, wt, tt);
            for (int d = 0; d < 3; d++) tt[d] = tt[d] + dtt[d];
            for (int d = 0; d < 3; d++) ttd[d] = damp ? tt[d] - ctt * fabs(tt[d]) * bsgn(wt[d]) : tt[d];
        }
        for (int d = 0; d < 3; d++) {
            torque_n[d] = tn[d];
            torque_t[d] = tt[d];
        }
    } else {
        const double ktn = tbl(mat, M_BKTN, it, jt), kutn = tbl(mat, M_BKUTN, it, jt), kctn = tbl(mat, M_BKCTN, it, jt);
        const double ktt = tbl(mat, M_BKTT, it, jt), kutt = tbl(mat, M_BKUTT, it, jt), kctt = tbl(mat, M_BKCTT, it, jt);
        const double dth[3] = {dt * wn[0], dt * wn[1], dt * wn[2]};
        for (int d = 0; d < 3; d++) tn[d] = tn[d] + dth[d];
        const double cn = -ktn * J;
        for (int d = 0; d < 3; d++) torque_n[d] = cn * tn[d];
        bproject(torque_n, wn, torque_n);
        for (int d = 0; d < 3; d++) {
            if (tn[d] > 0.0) hb[21 + d] = 0.0;
            else if (tn[d] < 0.0) hb[15 + d] = 0.0;
        }
        for (int d = 0; d < 3; d++) torque_n[d] = bnl_torque(tn[d], &hb[15 + d], &hb[21 + d], ktn, kutn, kctn, J);
        for (int d = 0; d < 3; d++)
            ntd[d] = damp ? torque_n[d] - ctn * fabs(torque_n[d]) * bsgn(wn[d]) : torque_n[d];
        const double dtt[3] = {dt * wt[0], dt * wt[1], dt * wt[2]};
        for (int d = 0; d < 3; d++) tt[d] = tt[d] + dtt[d];
        const double ct = -ktt * I;
        for (int d = 0; d < 3; d++) torque_t[d] = ct * tt[d];
        bproject(torque_t, wt, torque_t);
        for (int d = 0; d < 3; d++) {
            if (tt[d] > 0.0) hb[24 + d] = 0.0;
            else if (tt[d] < 0.0) hb[18 + d] = 0.0;
        }
        for (int d = 0; d < 3; d++) torque_t[d] = bnl_torque(tt[d], &hb[18 + d], &hb[24 + d], ktt, kutt, kctt, I);
        for (int d = 0; d < 3; d++)
            ttd[d] = damp ? torque_t[d] - ctt * fabs(torque_t[d]) * bsgn(wt[d]) : torque_t[d];
    }
    if (bs.stress_break) {
        const double nforce_mag = sqrt(nforce[0] * nforce[0] + nforce[1] * nforce[1] + nforce[2] * nforce[2]);
        const double tforce_mag = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
        const double ntorque_mag = sqrt(torque_n[0] * torque_n[0] + torque_n[1] * torque_n[1] + torque_n[2] * torque_n[2]);
        const double ttorque_mag = sqrt(torque_t[0] * torque_t[0] + torque_t[1] * torque_t[1] + torque_t[2] * torque_t[2]);
        const double rtc = tbl(mat, M_BRATIOTC, it, jt);
        bool nstress, tstress;
        if (!nl) {
            const double maxSigma = tbl(mat, M_BMAXSIGMA, it, jt) * ((bs.ratio_tc && displacement < 1e-16) ? rtc : 1.0);
            nstress = (1. * maxSigma) < (nforce_mag / A + ttorque_mag * rb / I);
            tstress = (1. * tbl(mat, M_BMAXTAU, it, jt)) < (tforce_mag / A + ntorque_mag * rb / J);
        } else {
            const double maxSigma = tbl(mat, M_BMAXSIGMA, it, jt) * ((bs.ratio_tc && displacement < -1.e-15) ? rtc : 1.0);
            nstress = maxSigma < (nforce_mag / A + ttorque_mag * rb / I);
            tstress = tbl(mat, M_BMAXTAU, it, jt) < (tforce_mag / A + ntorque_mag * rb / J);
        }
        if ((nstress || tstress) && upd) {
            hb[0] = 0.;
            hb[1] = 0.;
            return B_BROKEN;
        }
    }
    double tor[3];
    tor[0] = tfd[1] * en[2] - tfd[2] * en[1];
    tor[1] = tfd[2] * en[0] - tfd[0] * en[2];
    tor[2] = tfd[0] * en[1] - tfd[1] * en[0];
    for (int d = 0; d < 3; d++) {
        out[d] = nfd[d] + tfd[d];
        out[3 + d] = cri * tor[d] + ntd[d] + ttd[d];
        out[6 + d] = crj * tor[d] - ntd[d] - ttd[d];
    }
    if (upd)
        for (int d = 0; d < 3; d++) {
            hb[5 + d] = ft[d];
            hb[8 + d] = tn[d];
            hb[11 + d] = tt[d];
        }
    return B_FORCE;
}

__global__ void __launch_bounds__(kBlock) k_bond(BondArgs a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.halt && *a.halt) return;
    const long long ntimestep = a.step_dev ? *a.step_dev : a.ntimestep;
    const Model &M = a.M;
    const double4 pi = a.xr[i];
    const double4 vmi = a.vm[i];
    const double4 wi = a.wt[i];
    const int tagi = a.tag[i];
    const int it = meta_type((int)wi.w);
    const double cdm = M.cdf * M.cdf;
    const int dnum = M.dnum;
    const int off = M.lp.off_coh;
    const double xi[3] = {pi.x, pi.y, pi.z};
    const double vi[3] = {vmi.x, vmi.y, vmi.z};
    const double omi[3] = {wi.x, wi.y, wi.z};
    for (int e = a.nstart[i]; e < a.nstart[i + 1]; e++) {
        const uint32_t w = a.nj[e];
        const int j = (int)(w & kJMask);
        const int code = (int)((w >> 26) & 31u);
        if (code == kNoShift && a.tag[j] < tagi) continue;   // the partner's entry evaluates
        const double4 pj = a.xr[j];
        double xj[3] = {pj.x, pj.y, pj.z};
        if (code != kNoShift) {
            const double3 im = image_of(pj, code, M.prd);
            xj[0] = im.x; xj[1] = im.y; xj[2] = im.z;
        }
        const double delta[3] = {pi.x - xj[0], pi.y - xj[1], pi.z - xj[2]};
        const double rsq = delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
        const double radsum = pi.w + pj.w;
        const int rv = code == kNoShift ? a.rev[e] : -1;
        if (!(rsq < cdm * radsum * radsum)) {
            // out of the cohesion model's reach this step: no bond force
            if (w & kBondBit) {
                a.bvalid[e] = 0;
                if (rv >= 0) a.bvalid[rv] = 0;
            }
            continue;
        }
        const bool bonded = (w & kBondBit) != 0;
        if (!bonded && !(a.shearupdate && (M.bs.create_always || ntimestep == M.bs.ts_create
                                           || (a.brid && (double)ntimestep < a.brid[i]))))
            continue;
        const double4 vmj = a.vm[j];
        const double4 wj = a.wt[j];
        const double vj[3] = {vmj.x, vmj.y, vmj.z};
        const double omj[3] = {wj.x, wj.y, wj.z};
        double *hp = a.hist + (size_t)e * dnum;
        double out[9];
        bool created = false;
        const int res = bond_eval(M, a.mat, it, meta_type((int)wj.w), hp + off, delta, rsq, pi.w, pj.w, xi, vi, vj, omi, omj,
                                  a.shearupdate != 0, ntimestep, a.brid, i, j, created, out);
        if (created) {
            // an entry that carried no flag holds no valid contact history
            if (!(w & (kFlagBit | kBondBit)))
                for (int q = 0; q < off; q++) hp[q] = 0.0;
            a.nj[e] = w | kBondBit;
            if (rv >= 0) {
                const uint32_t wr = a.nj[rv];
                if (!(wr & (kFlagBit | kBondBit))) {
                    double *hq = a.hist + (size_t)rv * dnum;
                    for (int q = 0; q < off; q++) hq[q] = 0.0;
                }
                a.nj[rv] = wr | kBondBit;
            }
        }
        if (res == B_BROKEN) {
            a.nj[e] = w & ~kBondBit;
            a.bvalid[e] = 0;
            if (rv >= 0) {
                a.nj[rv] = a.nj[rv] & ~kBondBit;
                a.bvalid[rv] = 0;
            }
            continue;
        }
        if (res != B_FORCE) continue;
        double *bo = a.bout + (size_t)e * 6;
        for (int q = 0; q < 6; q++) bo[q] = out[q];
        a.bvalid[e] = 1;
        if (rv >= 0) {
            double *br = a.bout + (size_t)rv * 6;
            br[0] = -out[0]; br[1] = -out[1]; br[2] = -out[2];
            br[3] = out[6]; br[4] = out[7]; br[5] = out[8];
            a.bvalid[rv] = 1;
        }
    }
}

// rev[e] = index of the entry (j -> i) for an entry (i -> j) without image
__global__ void __launch_bounds__(kBlock) k_bond_rev(const int *nstart, const uint32_t *nj, int *rev, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int e = nstart[i]; e < nstart[i + 1]; e++) {
        const uint32_t w = nj[e];
        const int j = (int)(w & kJMask);
        const int code = (int)((w >> 26) & 31u);
        int r = -1;
        if (code == kNoShift)
            for (int f = nstart[j]; f < nstart[j + 1]; f++) {
                const uint32_t u = nj[f];
                if ((int)(u & kJMask) == i && (int)((u >> 26) & 31u) == kNoShift) {
                    r = f;
                    break;
                }
            }
        rev[e] = r;
    }
}

static int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

void launch_bond(const BondArgs &a, hipStream_t s)
{
    if (a.n) k_bond<<<blocks(a.n), kBlock, 0, s>>>(a);
}

void launch_bond_rev(const int *nstart, const uint32_t *nj, int *rev, int n, hipStream_t s)
{
    if (n) k_bond_rev<<<blocks(n), kBlock, 0, s>>>(nstart, nj, rev, n);
}

}  // namespace dem
