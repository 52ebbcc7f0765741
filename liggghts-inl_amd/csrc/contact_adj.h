// contact_adj.h -- the Hertz + tangential-history contact law of one side of
// a pair (or of a particle against a primitive plane) as a function of its
// inputs, and its hand-written reverse sweep (vector-Jacobian product).
//
// The law is the FP64 chain of the contact pass: SurfaceModel<DEFAULT>
// (surface_model_default.h:146-208), NormalModel<HERTZ>
// (normal_model_hertz.h:205-405, tangential damping on, limitForce off) and
// TangentialModel<TANGENTIAL_HISTORY> (tangential_model_history.h:135-425),
// with the history update of a normal step (shearupdate = 1) or of a set-up
// pass (shearupdate = 0).  Inputs: separation d = x_i - x_j (for a plane:
// from the plane to the particle along its axis), velocities, angular
// velocities, old shear, and the instance's Yeff, Geff, betaeff, mu.
// Outputs: force and torque on side i, new shear.
//
// law_vjp re-runs the forward arithmetic for its intermediates and walks it
// backwards statement by statement; given the adjoints of (F, T, shear_new)
// it accumulates the adjoints of every input, parameters included.
//
// The file is compiled by hipcc for the adjoint kernels and by the host
// compiler for the unit test that checks law_vjp against finite
// differences of law_fwd; the includer defines DEM_HD as the function
// qualifier (__host__ __device__ for hipcc, empty for the host compiler).
#pragma once
#include <math.h>

namespace dem {

struct LawIn {
    double d[3];               // x_i - x_j (plane: particle minus plane point)
    double vi[3], vj[3], wi[3], wj[3];
    double sh[3];              // shear history before the pass
    double radi, radj;         // plane: radj = 0
    double radsum, reff, meff; // plane: radsum = reff = radi, meff = m_i
    int wall;                  // plane: no partner angular velocity term
};

struct LawP {
    double Y, G, B, MU;        // Yeff, Geff, betaeff, coeffFrict
};

struct LawOut {
    double F[3], T[3], sh[3];
};

// c = sqrt(5/6), as the reference spells it (normal_model_hertz.h)
constexpr double kSqrt56 = 0.91287092917527685576161630466800355658790782499663875;

DEM_HD inline void law_fwd(const LawIn &a, const LawP &p, double dt, int update, LawOut &o)
{
    const double rsq = a.d[0] * a.d[0] + a.d[1] * a.d[1] + a.d[2] * a.d[2];
    const double r = sqrt(rsq);
    const double rinv = 1.0 / r;
    const double en[3] = {a.d[0] * rinv, a.d[1] * rinv, a.d[2] * rinv};
    double vr[3], vt[3], wr[3], vtr[3];
    for (int k = 0; k < 3; k++) vr[k] = a.vi[k] - a.vj[k];
    const double vn = vr[0] * en[0] + vr[1] * en[1] + vr[2] * en[2];
    for (int k = 0; k < 3; k++) vt[k] = vr[k] - vn * en[k];
    const double dn = a.radsum - r;
    const double cri = a.radi - 0.5 * dn;
    const double crj = a.wall ? 0.0 : a.radj - 0.5 * dn;
    for (int k = 0; k < 3; k++) wr[k] = (cri * a.wi[k] + crj * a.wj[k]) * rinv;
    vtr[0] = vt[0] + (a.d[1] * wr[2] - a.d[2] * wr[1]);
    vtr[1] = vt[1] + (a.d[2] * wr[0] - a.d[0] * wr[2]);
    vtr[2] = vt[2] + (a.d[0] * wr[1] - a.d[1] * wr[0]);
    const double s = sqrt(a.reff * dn);
    const double kn = 4. / 3. * p.Y * s;
    const double kt = 8. * p.G * s;
    const double Sn = 2. * p.Y * s;
    const double gn = -2. * kSqrt56 * p.B * sqrt(Sn * a.meff);
    const double gt = -2. * kSqrt56 * p.B * sqrt(kt * a.meff);
    const double Fn = -gn * vn + kn * dn;
    double sh2[3];
    if (update) {
        double sh1[3];
        for (int k = 0; k < 3; k++) sh1[k] = a.sh[k] + vtr[k] * dt;
        const double rs = sh1[0] * en[0] + sh1[1] * en[1] + sh1[2] * en[2];
        for (int k = 0; k < 3; k++) sh2[k] = sh1[k] - rs * en[k];
    } else {
        for (int k = 0; k < 3; k++) sh2[k] = a.sh[k];
    }
    const double shm = sqrt(sh2[0] * sh2[0] + sh2[1] * sh2[1] + sh2[2] * sh2[2]);
    double Ft[3];
    for (int k = 0; k < 3; k++) Ft[k] = -(kt * sh2[k]);
    const double Fs = kt * shm;
    const double Ff = p.MU * fabs(Fn);
    for (int k = 0; k < 3; k++) o.sh[k] = sh2[k];
    if (Fs > Ff) {
        if (shm != 0.0) {
            const double ratio = Ff / Fs;
            for (int k = 0; k < 3; k++) Ft[k] *= ratio;
            if (update)
                for (int k = 0; k < 3; k++) o.sh[k] = -Ft[k] / kt;
        } else {
            Ft[0] = Ft[1] = Ft[2] = 0.0;
        }
    } else {
        for (int k = 0; k < 3; k++) Ft[k] -= gt * vtr[k];
    }
    const double tor[3] = {en[1] * Ft[2] - en[2] * Ft[1], en[2] * Ft[0] - en[0] * Ft[2],
                           en[0] * Ft[1] - en[1] * Ft[0]};
    for (int k = 0; k < 3; k++) {
        o.F[k] = Fn * en[k] + Ft[k];
        o.T[k] = -cri * tor[k];
    }
}

// a x b
DEM_HD inline void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Adds to g (vector inputs) and gp (parameters) the adjoints of the inputs,
// given gF, gT, gS, the adjoints of F, T and the new shear.
DEM_HD inline void law_vjp(const LawIn &a, const LawP &p, double dt, int update, const double gF[3],
                           const double gT[3], const double gS[3], LawIn &g, LawP &gp)
{
    // ---- forward, keeping the intermediates
    const double rsq = a.d[0] * a.d[0] + a.d[1] * a.d[1] + a.d[2] * a.d[2];
    const double r = sqrt(rsq);
    const double rinv = 1.0 / r;
    const double en[3] = {a.d[0] * rinv, a.d[1] * rinv, a.d[2] * rinv};
    double vr[3], vt[3], wr[3], vtr[3];
    for (int k = 0; k < 3; k++) vr[k] = a.vi[k] - a.vj[k];
    const double vn = vr[0] * en[0] + vr[1] * en[1] + vr[2] * en[2];
    for (int k = 0; k < 3; k++) vt[k] = vr[k] - vn * en[k];
    const double dn = a.radsum - r;
    const double cri = a.radi - 0.5 * dn;
    const double crj = a.wall ? 0.0 : a.radj - 0.5 * dn;
    for (int k = 0; k < 3; k++) wr[k] = (cri * a.wi[k] + crj * a.wj[k]) * rinv;
    double dxw[3];
    cross3(a.d, wr, dxw);
    for (int k = 0; k < 3; k++) vtr[k] = vt[k] + dxw[k];
    const double s = sqrt(a.reff * dn);
    const double kn = 4. / 3. * p.Y * s;
    const double kt = 8. * p.G * s;
    const double Sn = 2. * p.Y * s;
    const double rootn = sqrt(Sn * a.meff), roott = sqrt(kt * a.meff);
    const double gn = -2. * kSqrt56 * p.B * rootn;
    const double gt = -2. * kSqrt56 * p.B * roott;
    const double Fn = -gn * vn + kn * dn;
    double sh1[3], sh2[3], rs = 0.0;
    if (update) {
        for (int k = 0; k < 3; k++) sh1[k] = a.sh[k] + vtr[k] * dt;
        rs = sh1[0] * en[0] + sh1[1] * en[1] + sh1[2] * en[2];
        for (int k = 0; k < 3; k++) sh2[k] = sh1[k] - rs * en[k];
    } else {
        for (int k = 0; k < 3; k++) sh1[k] = sh2[k] = a.sh[k];
    }
    const double shm = sqrt(sh2[0] * sh2[0] + sh2[1] * sh2[1] + sh2[2] * sh2[2]);
    double Ftv[3], Ft[3];
    for (int k = 0; k < 3; k++) Ftv[k] = -(kt * sh2[k]);
    const double Fs = kt * shm;
    const double Ff = p.MU * fabs(Fn);
    const bool capped = Fs > Ff;
    const bool zero_shear = capped && shm == 0.0;
    double ratio = 1.0;
    if (capped && !zero_shear) {
        ratio = Ff / Fs;
        for (int k = 0; k < 3; k++) Ft[k] = Ftv[k] * ratio;
    } else if (zero_shear) {
        Ft[0] = Ft[1] = Ft[2] = 0.0;
    } else {
        for (int k = 0; k < 3; k++) Ft[k] = Ftv[k] - gt * vtr[k];
    }
    double tor[3];
    cross3(en, Ft, tor);

    // ---- reverse
    double g_en[3] = {0, 0, 0}, g_Ft[3] = {0, 0, 0}, g_tor[3], g_vtr[3] = {0, 0, 0}, g_sh2[3] = {0, 0, 0};
    double g_Fn = 0, g_cri = 0, g_crj = 0, g_kt = 0, g_kn = 0, g_gn = 0, g_gt = 0, g_s = 0, g_dn = 0;
    double g_vn = 0, g_rinv = 0, g_r = 0, g_shm = 0, g_Ftv[3] = {0, 0, 0};
    // T = -cri tor
    for (int k = 0; k < 3; k++) {
        g_cri -= gT[k] * tor[k];
        g_tor[k] = -cri * gT[k];
    }
    // tor = en x Ft
    {
        double t1[3], t2[3];
        cross3(Ft, g_tor, t1);   // d/d en
        cross3(g_tor, en, t2);   // d/d Ft
        for (int k = 0; k < 3; k++) {
            g_en[k] += t1[k];
            g_Ft[k] += t2[k];
        }
    }
    // F = Fn en + Ft
    for (int k = 0; k < 3; k++) {
        g_Fn += gF[k] * en[k];
        g_en[k] += Fn * gF[k];
        g_Ft[k] += gF[k];
    }
    // new shear and the tangential force
    if (capped && !zero_shear) {
        if (update) {
            // sh_new = -Ft / kt
            for (int k = 0; k < 3; k++) {
                g_Ft[k] -= gS[k] / kt;
                g_kt += Ft[k] * gS[k] / (kt * kt);
            }
        } else {
            for (int k = 0; k < 3; k++) g_sh2[k] += gS[k];
        }
        // Ft = Ftv ratio, ratio = Ff / Fs
        double g_ratio = 0.0;
        for (int k = 0; k < 3; k++) {
            g_Ftv[k] += ratio * g_Ft[k];
            g_ratio += Ftv[k] * g_Ft[k];
        }
        const double g_Ff = g_ratio / Fs;
        const double g_Fs = -g_ratio * Ff / (Fs * Fs);
        // Ff = mu |Fn|
        gp.MU += fabs(Fn) * g_Ff;
        g_Fn += p.MU * ((Fn > 0.0) - (Fn < 0.0)) * g_Ff;
        // Fs = kt shm
        g_kt += shm * g_Fs;
        g_shm += kt * g_Fs;
    } else if (zero_shear) {
        for (int k = 0; k < 3; k++) g_sh2[k] += gS[k];
    } else {
        for (int k = 0; k < 3; k++) {
            g_sh2[k] += gS[k];
            g_Ftv[k] += g_Ft[k];
            g_gt -= vtr[k] * g_Ft[k];
            g_vtr[k] -= gt * g_Ft[k];
        }
    }
    // Ftv = -kt sh2
    for (int k = 0; k < 3; k++) {
        g_kt -= sh2[k] * g_Ftv[k];
        g_sh2[k] -= kt * g_Ftv[k];
    }
    // shm = |sh2|
    if (shm > 0.0)
        for (int k = 0; k < 3; k++) g_sh2[k] += sh2[k] / shm * g_shm;
    // sh2 = sh1 - (sh1.en) en ; sh1 = sh + vtr dt
    if (update) {
        double g_sh1[3];
        double g_rs = 0.0;
        for (int k = 0; k < 3; k++) {
            g_sh1[k] = g_sh2[k];
            g_rs -= en[k] * g_sh2[k];
            g_en[k] -= rs * g_sh2[k];
        }
        for (int k = 0; k < 3; k++) {
            g_sh1[k] += g_rs * en[k];
            g_en[k] += g_rs * sh1[k];
        }
        for (int k = 0; k < 3; k++) {
            g.sh[k] += g_sh1[k];
            g_vtr[k] += dt * g_sh1[k];
        }
    } else {
        for (int k = 0; k < 3; k++) g.sh[k] += g_sh2[k];
    }
    // Fn = -gn vn + kn dn
    g_gn -= vn * g_Fn;
    g_vn -= gn * g_Fn;
    g_kn += dn * g_Fn;
    g_dn += kn * g_Fn;
    // gt = -2c B sqrt(kt meff), gn = -2c B sqrt(Sn meff), Sn = 2 Y s
    gp.B += -2. * kSqrt56 * roott * g_gt;
    if (roott > 0.0) g_kt += -2. * kSqrt56 * p.B * 0.5 * a.meff / roott * g_gt;
    gp.B += -2. * kSqrt56 * rootn * g_gn;
    double g_Sn = 0.0;
    if (rootn > 0.0) g_Sn = -2. * kSqrt56 * p.B * 0.5 * a.meff / rootn * g_gn;
    gp.Y += 2. * s * g_Sn;
    g_s += 2. * p.Y * g_Sn;
    // kt = 8 G s, kn = 4/3 Y s
    gp.G += 8. * s * g_kt;
    g_s += 8. * p.G * g_kt;
    gp.Y += 4. / 3. * s * g_kn;
    g_s += 4. / 3. * p.Y * g_kn;
    // s = sqrt(reff dn)
    if (s > 0.0) g_dn += 0.5 * a.reff / s * g_s;
    // vtr = vt + d x wr
    double g_vt[3], g_wr[3];
    {
        double t1[3], t2[3];
        cross3(wr, g_vtr, t1);   // d/d d
        cross3(g_vtr, a.d, t2);  // d/d wr
        for (int k = 0; k < 3; k++) {
            g_vt[k] = g_vtr[k];
            g.d[k] += t1[k];
            g_wr[k] = t2[k];
        }
    }
    // wr = (cri wi + crj wj) rinv
    for (int k = 0; k < 3; k++) {
        g_cri += a.wi[k] * rinv * g_wr[k];
        g_crj += a.wj[k] * rinv * g_wr[k];
        g.wi[k] += cri * rinv * g_wr[k];
        if (!a.wall) g.wj[k] += crj * rinv * g_wr[k];
        g_rinv += (cri * a.wi[k] + crj * a.wj[k]) * g_wr[k];
    }
    if (a.wall) g_crj = 0.0;
    // cri = radi - dn/2, crj = radj - dn/2
    g_dn -= 0.5 * (g_cri + g_crj);
    // dn = radsum - r
    g_r -= g_dn;
    // vt = vr - vn en ; vn = vr.en ; vr = vi - vj
    double g_vr[3];
    for (int k = 0; k < 3; k++) {
        g_vr[k] = g_vt[k];
        g_vn -= en[k] * g_vt[k];
        g_en[k] -= vn * g_vt[k];
    }
    for (int k = 0; k < 3; k++) {
        g_vr[k] += g_vn * en[k];
        g_en[k] += g_vn * vr[k];
    }
    for (int k = 0; k < 3; k++) {
        g.vi[k] += g_vr[k];
        g.vj[k] -= g_vr[k];
    }
    // en = d rinv ; rinv = 1/r ; r = |d|
    for (int k = 0; k < 3; k++) {
        g.d[k] += rinv * g_en[k];
        g_rinv += a.d[k] * g_en[k];
    }
    g_r -= rinv * rinv * g_rinv;
    for (int k = 0; k < 3; k++) g.d[k] += a.d[k] / r * g_r;
}

}  // namespace dem
