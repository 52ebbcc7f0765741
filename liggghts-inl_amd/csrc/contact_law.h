// contact_law.h -- the sphere contact-model chain of pair_style gran
// (surface default -> normal hertz | hysteretic/nonlinear1/2 -> tangential
// history -> rolling epsd*) for one contact, as seen from atom i.
//
// FP64 form: same operations in the same order as the reference, so that
// compiled without contraction (-ffp-contract=off) every per-contact value
// rounds identically (transcendentals aside: pow/exp/log come from the
// device math library):
//   surface_model_default.h:146-208, normal_model_hertz.h:205-405,
//   normal_model_hysteretic_nonlinear{1,2}.h:149-420,
//   tangential_model_history.h:135-425, rolling_model_epsd.h:96-335,
//   rolling_model_epsd2.h:161-230, rolling_model_epsd3.h:178-262,
//   fix_wall_gran_base.h:159-378 (wall variant).
//
// Mixed form (contact_law_mixed, Hertz chains): overlap, separation and the
// history integration stay FP64; the contact-law arithmetic runs in FP32.
#pragma once
#include "dem_device.h"

namespace dem {

struct ContactIn {
    int itype, jtype;
    double radi, radj;
    double mi, mj, meff;
    double delta[3];     // x_i - x_j (pair) or particle - wall point (wall)
    double r, rinv;
    double radsum;
    double deltan_in;    // wall only: overlap handed over by the wall code
    double vi[3], vj[3], wi[3], wj[3];
    bool is_wall;
};

__device__ inline double tbl(const double *mat, int slot, int it, int jt)
{
    return mat[slot * kT1 * kT1 + it * kT1 + jt];
}

struct NormalOut {
    double Fn, kn, kt, gammat;
};

// NormalModel<HYSTERETIC_NONLINEAR1/2>::surfacesIntersect; writes its
// twelve history slots whatever shearupdate says, and sqrt(5/4) is the
// reference's integer expression, sqrt(1) = 1.
template <int NM>
__device__ inline NormalOut normal_hysteretic(const double *mat, const ContactIn &c, double deltan, double vn,
                                              int limitf, double *history)
{
    const int it = c.itype, jt = c.jtype;
    const double meff = c.meff;
    const double kn0 = tbl(mat, M_HK, it, jt);
    const double Alpha_in = tbl(mat, M_HALPHA, it, jt);
    const double Cin_in = tbl(mat, M_HCIN, it, jt);
    const double A1_in = tbl(mat, M_HA1, it, jt);
    const double A2_in = tbl(mat, M_HA2, it, jt);
    const double A3_in = tbl(mat, M_HA3, it, jt);
    const double kcin_in = tbl(mat, M_HKCIN, it, jt);
    const double crl = tbl(mat, M_RESTLOG, it, jt);
    const double k_c = kcin_in * A2_in;
    const double kc = tbl(mat, M_HKN2KC, it, jt) * kn0;
    const double f_0 = tbl(mat, M_HFADH, it, jt);
    const double sqrt54 = 1.0;
    const double fcc = 1.0;
    double deltaMax;
    if (deltan > history[0]) {
        history[0] = deltan;
        deltaMax = deltan;
    } else {
        deltaMax = history[0];
    }
    double deltaZero = history[1];
    double k1 = history[2];
    double deltaZero_old = history[3];
    const double k1_old = history[4];
    const double delta_old = history[5];
    double deltaMin = history[6];
    double betan = history[7];
    double f0 = history[8];
    double f0_old = history[9];
    double k2 = A3_in * k1;
    int tag_status;
    if (deltan >= delta_old) {
        if (delta_old == 0) tag_status = 1;
        else tag_status = (vn > 0) ? 2 : 1;
    } else {
        tag_status = 2;
    }
    const double dampfac = 1. + (M_PI / crl) * (M_PI / crl);
    double fHys, gamman, gammat;
    if (tag_status == 1) {
        if (deltaZero == 0) k1 = A2_in;
        else k1 = history[2];
        deltaMin = history[6];
        f0_old = f0;
        if (deltan <= deltaZero) {
            if (NM == NM_HYST2) fHys = Cin_in * k2 * (exp(betan * (deltan - deltaZero)) - 1);
            else if (deltan <= deltaMin) fHys = -k_c * deltan;
            else fHys = betan * (deltan - deltaZero);
        } else {
            fHys = Alpha_in * k1 * pow(deltan - deltaZero, 2) + f0;
        }
        const double pw = (NM == NM_HYST2) ? 0.5 : 0.25;
        gamman = sqrt54 * sqrt(4. * meff * Alpha_in * k1 / dampfac) * (pow(deltan, pw) + pow(deltaZero, pw));
        gammat = gamman;
        history[1] = deltaZero;
        history[2] = k1;
        history[3] = deltaZero;
        history[4] = k1;
        history[5] = deltan;
        history[6] = deltaMin;
        history[7] = betan;
        history[8] = f0;
        history[9] = f0_old;
    } else {
        if (NM == NM_HYST1) {
            k2 = A3_in * k1_old;
            deltaZero = (1 - k1_old / k2) * deltaMax;
            const double beta = Alpha_in * k1_old * pow(deltaMax - deltaZero_old, 2) / k2 / (deltaMax - deltaZero);
            deltaMin = beta * (k2 - k1_old) / (beta * k2 + k_c) * deltaMax;
            k1 = deltaMax * A1_in + A2_in;
            if (deltan >= deltaMin) {
                betan = beta * k2;
                if (deltan >= deltaZero) {
                    fHys = beta * k2 * (deltan - deltaZero) + (deltan - deltaZero) * f0_old / (deltaMax - deltaZero);
                    f0 = fHys - Alpha_in * k1 * pow(deltan - deltaZero, 2);
                } else {
                    fHys = beta * k2 * (deltan - deltaZero);
                    f0 = 0;
                }
            } else {
                fHys = -k_c * deltan;
                f0 = 0;
            }
            gamman = sqrt54 * sqrt(4. * meff * Alpha_in * k1_old / dampfac) * (pow(deltan, 0.25));
        } else {
            k2 = A3_in * (A1_in * deltaMax + A2_in);
            deltaZero = fcc * (1 - k1_old / k2) * deltaMax;
            deltaZero_old = history[3];
            betan = log(Alpha_in * k1_old / Cin_in / k2 * pow(deltaMax - deltaZero_old, 2) + 1)
                    / (deltaMax - fcc * (1 - k1_old / k2) * deltaMax);
            deltaMin = betan * (k2 - k1_old) / (betan * k2 + k_c) * deltaMax;
            k1 = deltaMax * A1_in + A2_in;
            if (deltan >= deltaZero) {
                fHys = Cin_in * k2 * (exp(betan * (deltan - deltaZero)) - 1)
                       + (deltan - deltaZero) * f0_old / (deltaMax - deltaZero);
                f0 = fHys - Alpha_in * k1 * pow(deltan - deltaZero, 2);
            } else {
                fHys = Cin_in * k2 * (exp(betan * (deltan - deltaZero)) - 1);
                f0 = 0;
            }
            gamman = 1 * 0.001 * sqrt54 * sqrt(4. * meff * Alpha_in * k1_old / dampfac) * (pow(deltan, -0.25));
        }
        gammat = gamman;
        history[1] = deltaZero;
        history[2] = k1;
        history[3] = deltaZero_old;
        history[4] = k1_old;
        history[5] = deltan;
        history[6] = deltaMin;
        history[7] = betan;
        history[8] = f0;
    }
    const double Fn_damping = -gamman * vn;
    double Fn = fHys + Fn_damping + f_0;
    if (limitf && (Fn < 0.0) && kc == 0 && f_0 == 0.0) Fn = 0.0;
    history[10] = kc;
    history[11] = f_0;
    NormalOut o;
    o.Fn = Fn;
    o.kn = k1;
    o.kt = k1;
    o.gammat = gammat;
    return o;
}

// History layout of a chain with normal model NM: hysteretic slots 0-11,
// then tangential (3), then rolling (3).  Offsets are compile-time so the
// per-contact history lives in registers.
__host__ __device__ constexpr int chain_off_shear(int nm) { return nm == NM_HERTZ ? 0 : 12; }
__host__ __device__ constexpr int chain_max_dnum(int nm) { return chain_off_shear(nm) + 6; }

// Returns the i-side force and torque increments (ForceData i_forces).
// h: L.dnum history doubles in the chain's layout, updated in place.
// CH: a cohesion model sits between the normal and tangential models; coh =
// its force and torque on this side, cmode CM_BOND adds them to the chain,
// CM_BOND_NONLINEAR replaces the normal force with them, CM_OFF skips.
template <int NM, bool CH = false>
__device__ inline void contact_fp64(const Model &M, const Layout &L, const double *mat, const ContactIn &c,
                                    double *h, bool update, double F[3], double T[3],
                                    const double *coh = nullptr, int cmode = CM_OFF)
{
    const double SQRT_FIVE_OVER_SIX = 0.91287092917527685576161630466800355658790782499663875;
    const double enx = c.delta[0] * c.rinv;
    const double eny = c.delta[1] * c.rinv;
    const double enz = c.delta[2] * c.rinv;

    // ---- surface default
    const double vr1 = c.vi[0] - c.vj[0];
    const double vr2 = c.vi[1] - c.vj[1];
    const double vr3 = c.vi[2] - c.vj[2];
    const double vn = vr1 * enx + vr2 * eny + vr3 * enz;
    const double vn1 = vn * enx, vn2 = vn * eny, vn3 = vn * enz;
    const double vt1 = vr1 - vn1, vt2 = vr2 - vn2, vt3 = vr3 - vn3;
    const double deltan = c.radsum - c.r;
    const double dx = c.delta[0], dy = c.delta[1], dz = c.delta[2];
    double wr1, wr2, wr3, cri, crj = 0.0;
    if (c.is_wall) {
        const double cr = c.radi - 0.5 * c.deltan_in;
        wr1 = cr * c.wi[0] * c.rinv;
        wr2 = cr * c.wi[1] * c.rinv;
        wr3 = cr * c.wi[2] * c.rinv;
        cri = cr;
    } else {
        cri = c.radi - 0.5 * deltan;
        crj = c.radj - 0.5 * deltan;
        wr1 = (cri * c.wi[0] + crj * c.wj[0]) * c.rinv;
        wr2 = (cri * c.wi[1] + crj * c.wj[1]) * c.rinv;
        wr3 = (cri * c.wi[2] + crj * c.wj[2]) * c.rinv;
    }
    const double vtr1 = vt1 - (dz * wr2 - dy * wr3);
    const double vtr2 = vt2 - (dx * wr3 - dz * wr1);
    const double vtr3 = vt3 - (dy * wr1 - dx * wr2);

    // ---- normal model
    const int it = c.itype, jt = c.jtype;
    const double reff = c.is_wall ? c.radi : (c.radi * c.radj / (c.radi + c.radj));
    double Fn, kn, kt, gammat;
    if (NM == NM_HERTZ) {
        const double sqrtval = sqrt(reff * deltan);
        const double Yeff = tbl(mat, M_YEFF, it, jt);
        const double Geff = tbl(mat, M_GEFF, it, jt);
        const double beta = tbl(mat, M_BETA, it, jt);
        const double Sn = 2. * Yeff * sqrtval;
        const double St = 8. * Geff * sqrtval;
        kn = 4. / 3. * Yeff * sqrtval;
        kt = St;
        const double gamman = -2. * SQRT_FIVE_OVER_SIX * beta * sqrt(Sn * c.meff);
        gammat = M.tdamp ? -2. * SQRT_FIVE_OVER_SIX * beta * sqrt(St * c.meff) : 0.0;
        const double Fn_damping = -gamman * vn;
        const double Fn_contact = kn * deltan;
        Fn = Fn_damping + Fn_contact;
        if (L.limitf && (Fn < 0.0)) Fn = 0.0;
    } else {
        const NormalOut o = normal_hysteretic<NM>(mat, c, deltan, vn, L.limitf, h);
        Fn = o.Fn;
        kn = o.kn;
        kt = o.kt;
        gammat = o.gammat;
    }
    F[0] = Fn * enx;
    F[1] = Fn * eny;
    F[2] = Fn * enz;
    double T0[3] = {0.0, 0.0, 0.0};
    if (CH && cmode != CM_OFF) {
        if (cmode == CM_BOND_NONLINEAR) {
            F[0] = coh[0]; F[1] = coh[1]; F[2] = coh[2];
        } else {
            F[0] += coh[0]; F[1] += coh[1]; F[2] += coh[2];
        }
        T0[0] = coh[3]; T0[1] = coh[4]; T0[2] = coh[5];
    }

    // ---- tangential history
    double *shear = h + chain_off_shear(NM);
    if (update) {
        const double dt = M.dt;
        shear[0] += vtr1 * dt;
        shear[1] += vtr2 * dt;
        shear[2] += vtr3 * dt;
        const double rsht = shear[0] * enx + shear[1] * eny + shear[2] * enz;
        shear[0] -= rsht * enx;
        shear[1] -= rsht * eny;
        shear[2] -= rsht * enz;
    }
    const double shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
    const double xmu = tbl(mat, M_MU, it, jt);
    double Ft1 = -(kt * shear[0]);
    double Ft2 = -(kt * shear[1]);
    double Ft3 = -(kt * shear[2]);
    const double Ft_shear = kt * shrmag;
    const double Ft_friction = xmu * fabs(Fn);
    if (Ft_shear > Ft_friction) {
        if (shrmag != 0.0) {
            const double ratio = Ft_friction / Ft_shear;
            Ft1 *= ratio;
            Ft2 *= ratio;
            Ft3 *= ratio;
            if (update) {
                shear[0] = -Ft1 / kt;
                shear[1] = -Ft2 / kt;
                shear[2] = -Ft3 / kt;
            }
        } else {
            Ft1 = Ft2 = Ft3 = 0.0;
        }
    } else {
        Ft1 -= (gammat * vtr1);
        Ft2 -= (gammat * vtr2);
        Ft3 -= (gammat * vtr3);
    }
    const double tor1 = eny * Ft3 - enz * Ft2;
    const double tor2 = enz * Ft1 - enx * Ft3;
    const double tor3 = enx * Ft2 - eny * Ft1;
    F[0] += Ft1;
    F[1] += Ft2;
    F[2] += Ft3;
    if (CH && cmode != CM_OFF) {
        T[0] = T0[0] + -cri * tor1;
        T[1] = T0[1] + -cri * tor2;
        T[2] = T0[2] + -cri * tor3;
    } else {
        T[0] = -cri * tor1;
        T[1] = -cri * tor2;
        T[2] = -cri * tor3;
    }

    // ---- rolling friction
    if (M.rolling && L.off_roll >= 0) {
        double w1, w2, w3, r_inertia;
        if (c.is_wall) {
            w1 = wr1; w2 = wr2; w3 = wr3;
            r_inertia = 1.4 * c.mi * reff * reff;
        } else {
            w1 = c.wi[0] - c.wj[0];
            w2 = c.wi[1] - c.wj[1];
            w3 = c.wi[2] - c.wj[2];
            const double ri = c.mi * c.radi * c.radi;
            const double rj = c.mj * c.radj * c.radj;
            r_inertia = 1.4 * ri * rj / (ri + rj);
        }
        double wt0, wt1, wt2;
        if (M.torsion) {
            wt0 = w1; wt1 = w2; wt2 = w3;
        } else {
            const double wdd = w1 * enx + w2 * eny + w3 * enz;
            const double wn0 = enx * wdd, wn1 = eny * wdd, wn2 = enz * wdd;
            wt0 = w1 - wn0; wt1 = w2 - wn1; wt2 = w3 - wn2;
        }
        const double rmu = tbl(mat, M_MUR, it, jt);
        double kr;
        if (M.rolling == 1) kr = 2.25 * kn * rmu * rmu * reff * reff;
        else if (M.rolling == 2) kr = kt * reff * reff;
        else kr = M.krs * kn * rmu * rmu * reff * reff;
        const double dtkr = M.dt * kr;
        double *ch = h + chain_off_shear(NM) + 3;
        double rt0 = ch[0] + dtkr * wt0;
        double rt1 = ch[1] + dtkr * wt1;
        double rt2 = ch[2] + dtkr * wt2;
        const double mag = sqrt(rt0 * rt0 + rt1 * rt1 + rt2 * rt2);
        const double rmax = fabs(Fn) * reff * rmu;
        if (mag > rmax) {
            const double factor = rmax / mag;
            rt0 *= factor; rt1 *= factor; rt2 *= factor;
            if (update) { ch[0] = rt0; ch[1] = rt1; ch[2] = rt2; }
        } else {
            if (update) { ch[0] = rt0; ch[1] = rt1; ch[2] = rt2; }
            if (M.rolling != 2) {
                const double r_coef = tbl(mat, M_ETAR, it, jt) * 2 * sqrt(r_inertia * kr);
                rt0 += r_coef * wt0;
                rt1 += r_coef * wt1;
                rt2 += r_coef * wt2;
            }
        }
        T[0] -= rt0;
        T[1] -= rt1;
        T[2] -= rt2;
    } else if (M.rolling == 4) {
        // RollingModel<ROLLING_CDTNONLINEAR2> (rolling_model_cdtnonlinear2.h:
        // 96-172): mu_r Fn reff along the relative rolling velocity, no history
        const double rmu = tbl(mat, M_MUR, it, jt);
        double w1, w2, w3;
        if (c.is_wall) {
            w1 = wr1; w2 = wr2; w3 = wr3;
        } else {
            w1 = c.wi[0] - c.wj[0];
            w2 = c.wi[1] - c.wj[1];
            w3 = c.wi[2] - c.wj[2];
        }
        const double mag = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
        if (mag > 0.) {
            double r0, r1, r2;
            if (c.is_wall) {
                r0 = rmu * Fn * w1 / mag * reff;
                r1 = rmu * Fn * w2 / mag * reff;
                r2 = rmu * Fn * w3 / mag * reff;
            } else {
                const double sc = rmu * Fn * reff / mag;
                r0 = w1 * sc; r1 = w2 * sc; r2 = w3 * sc;
            }
            if (!M.torsion) {
                const double d = r0 * enx + r1 * eny + r2 * enz;
                const double n0 = enx * d, n1 = eny * d, n2 = enz * d;
                r0 = r0 - n0; r1 = r1 - n1; r2 = r2 - n2;
            }
            T[0] -= r0;
            T[1] -= r1;
            T[2] -= r2;
        }
    }
}

// surfacesClose of a chain: hysteretic zeroes its ten state slots and clears
// its flag, tangential and rolling zero their history
// (normal_model_hysteretic_nonlinear1.h:444-458,
// tangential_model_history.h:428-440, rolling_model_epsd.h:222-229).
// Returns true if the chain's contact flag stays set (hertz keeps it).
__device__ inline bool contact_close(const Layout &L, double *hp)
{
    if (L.off_norm >= 0)
        for (int q = 0; q < 10; q++) hp[L.off_norm + q] = 0.0;
    for (int q = 0; q < 3; q++) hp[L.off_shear + q] = 0.0;
    if (L.off_roll >= 0)
        for (int q = 0; q < 3; q++) hp[L.off_roll + q] = 0.0;
    return L.normal == NM_HERTZ;
}

// Mixed precision: the overlap deltan (a small difference of large numbers)
// and the history integration are FP64; everything else of the law is FP32.
// Hertz chains only (shear at slot 0, rolling torque at slot 3).
__device__ inline void contact_mixed(const Model &M, const double *mat, const ContactIn &c,
                                     double *h, bool update, double F[3], double T[3])
{
#pragma clang fp contract(fast)
    const float SQRT_FIVE_OVER_SIX = 0.912870929175276856f;
    const float rinv = (float)c.rinv;
    const float enx = (float)c.delta[0] * rinv;
    const float eny = (float)c.delta[1] * rinv;
    const float enz = (float)c.delta[2] * rinv;

    const float vr1 = (float)(c.vi[0] - c.vj[0]);
    const float vr2 = (float)(c.vi[1] - c.vj[1]);
    const float vr3 = (float)(c.vi[2] - c.vj[2]);
    const float vn = vr1 * enx + vr2 * eny + vr3 * enz;
    const float vt1 = vr1 - vn * enx, vt2 = vr2 - vn * eny, vt3 = vr3 - vn * enz;
    const double deltan_d = c.radsum - c.r;
    const float deltan = (float)deltan_d;
    const float dx = (float)c.delta[0], dy = (float)c.delta[1], dz = (float)c.delta[2];
    float wr1, wr2, wr3, cri;
    if (c.is_wall) {
        const float cr = (float)(c.radi - 0.5 * c.deltan_in);
        wr1 = cr * (float)c.wi[0] * rinv;
        wr2 = cr * (float)c.wi[1] * rinv;
        wr3 = cr * (float)c.wi[2] * rinv;
        cri = cr;
    } else {
        cri = (float)(c.radi - 0.5 * deltan_d);
        const float crj = (float)(c.radj - 0.5 * deltan_d);
        wr1 = (cri * (float)c.wi[0] + crj * (float)c.wj[0]) * rinv;
        wr2 = (cri * (float)c.wi[1] + crj * (float)c.wj[1]) * rinv;
        wr3 = (cri * (float)c.wi[2] + crj * (float)c.wj[2]) * rinv;
    }
    const float vtr1 = vt1 - (dz * wr2 - dy * wr3);
    const float vtr2 = vt2 - (dx * wr3 - dz * wr1);
    const float vtr3 = vt3 - (dy * wr1 - dx * wr2);

    const int it = c.itype, jt = c.jtype;
    const float radi = (float)c.radi, radj = (float)c.radj;
    const float reff = c.is_wall ? radi : (radi * radj / (radi + radj));
    const float sqrtval = sqrtf(reff * deltan);
    const float Yeff = (float)tbl(mat, M_YEFF, it, jt);
    const float Geff = (float)tbl(mat, M_GEFF, it, jt);
    const float beta = (float)tbl(mat, M_BETA, it, jt);
    const float meff = (float)c.meff;
    const float Sn = 2.f * Yeff * sqrtval;
    const float St = 8.f * Geff * sqrtval;
    const float kn = (4.f / 3.f) * Yeff * sqrtval;
    const float kt = St;
    const float gamman = -2.f * SQRT_FIVE_OVER_SIX * beta * sqrtf(Sn * meff);
    const float gammat = M.tdamp ? -2.f * SQRT_FIVE_OVER_SIX * beta * sqrtf(St * meff) : 0.f;
    float Fn = -gamman * vn + kn * deltan;
    if ((c.is_wall ? M.lw.limitf : M.lp.limitf) && (Fn < 0.f)) Fn = 0.f;

    double *shear = h;
    if (update) {
        const double dt = M.dt;
        shear[0] += (double)vtr1 * dt;
        shear[1] += (double)vtr2 * dt;
        shear[2] += (double)vtr3 * dt;
        const double rsht = shear[0] * enx + shear[1] * eny + shear[2] * enz;
        shear[0] -= rsht * enx;
        shear[1] -= rsht * eny;
        shear[2] -= rsht * enz;
    }
    const float s0 = (float)shear[0], s1 = (float)shear[1], s2 = (float)shear[2];
    const float shrmag = sqrtf(s0 * s0 + s1 * s1 + s2 * s2);
    const float xmu = (float)tbl(mat, M_MU, it, jt);
    float Ft1 = -(kt * s0), Ft2 = -(kt * s1), Ft3 = -(kt * s2);
    const float Ft_shear = kt * shrmag;
    const float Ft_friction = xmu * fabsf(Fn);
    if (Ft_shear > Ft_friction) {
        if (shrmag != 0.f) {
            const float ratio = Ft_friction / Ft_shear;
            Ft1 *= ratio; Ft2 *= ratio; Ft3 *= ratio;
            if (update) {
                const double dratio = (double)ratio;
                shear[0] *= dratio;
                shear[1] *= dratio;
                shear[2] *= dratio;
            }
        } else {
            Ft1 = Ft2 = Ft3 = 0.f;
        }
    } else {
        Ft1 -= gammat * vtr1;
        Ft2 -= gammat * vtr2;
        Ft3 -= gammat * vtr3;
    }
    const float tor1 = eny * Ft3 - enz * Ft2;
    const float tor2 = enz * Ft1 - enx * Ft3;
    const float tor3 = enx * Ft2 - eny * Ft1;
    F[0] = (double)(Fn * enx + Ft1);
    F[1] = (double)(Fn * eny + Ft2);
    F[2] = (double)(Fn * enz + Ft3);
    float T0 = -cri * tor1, T1 = -cri * tor2, T2 = -cri * tor3;

    if (M.rolling) {
        float w1, w2, w3, r_inertia;
        if (c.is_wall) {
            w1 = wr1; w2 = wr2; w3 = wr3;
            r_inertia = 1.4f * (float)c.mi * reff * reff;
        } else {
            w1 = (float)(c.wi[0] - c.wj[0]);
            w2 = (float)(c.wi[1] - c.wj[1]);
            w3 = (float)(c.wi[2] - c.wj[2]);
            const float ri = (float)c.mi * radi * radi;
            const float rj = (float)c.mj * radj * radj;
            r_inertia = 1.4f * ri * rj / (ri + rj);
        }
        float wt0, wt1, wt2;
        if (M.torsion) {
            wt0 = w1; wt1 = w2; wt2 = w3;
        } else {
            const float wdd = w1 * enx + w2 * eny + w3 * enz;
            wt0 = w1 - enx * wdd; wt1 = w2 - eny * wdd; wt2 = w3 - enz * wdd;
        }
        const float rmu = (float)tbl(mat, M_MUR, it, jt);
        float kr;
        if (M.rolling == 1) kr = 2.25f * kn * rmu * rmu * reff * reff;
        else if (M.rolling == 2) kr = kt * reff * reff;
        else kr = (float)M.krs * kn * rmu * rmu * reff * reff;
        const double dtkr = M.dt * (double)kr;
        double *ch = h + 3;
        double rt0 = ch[0] + dtkr * wt0;
        double rt1 = ch[1] + dtkr * wt1;
        double rt2 = ch[2] + dtkr * wt2;
        const float mag = sqrtf((float)(rt0 * rt0 + rt1 * rt1 + rt2 * rt2));
        const float rmax = fabsf(Fn) * reff * rmu;
        if (mag > rmax) {
            const double factor = (double)(rmax / mag);
            rt0 *= factor; rt1 *= factor; rt2 *= factor;
            if (update) { ch[0] = rt0; ch[1] = rt1; ch[2] = rt2; }
        } else {
            if (update) { ch[0] = rt0; ch[1] = rt1; ch[2] = rt2; }
            if (M.rolling != 2) {
                const float r_coef = (float)tbl(mat, M_ETAR, it, jt) * 2.f * sqrtf(r_inertia * kr);
                rt0 += (double)(r_coef * wt0);
                rt1 += (double)(r_coef * wt1);
                rt2 += (double)(r_coef * wt2);
            }
        }
        T0 -= (float)rt0;
        T1 -= (float)rt1;
        T2 -= (float)rt2;
    }
    T[0] = (double)T0;
    T[1] = (double)T1;
    T[2] = (double)T2;
}

}  // namespace dem
