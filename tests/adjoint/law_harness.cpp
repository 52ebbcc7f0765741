// Host build of the adjoint's contact law (liggghts-inl_amd/csrc/contact_adj.h)
// for the CPU tests: flat C entry points over law_fwd and law_vjp.
#define DEM_HD
#include "contact_adj.h"

using dem::LawIn;
using dem::LawOut;
using dem::LawP;

// in: d[3] vi[3] vj[3] wi[3] wj[3] sh[3] radi radj radsum reff meff wall (24)
static LawIn unpack(const double *in)
{
    LawIn a;
    for (int k = 0; k < 3; k++) {
        a.d[k] = in[k];
        a.vi[k] = in[3 + k];
        a.vj[k] = in[6 + k];
        a.wi[k] = in[9 + k];
        a.wj[k] = in[12 + k];
        a.sh[k] = in[15 + k];
    }
    a.radi = in[18];
    a.radj = in[19];
    a.radsum = in[20];
    a.reff = in[21];
    a.meff = in[22];
    a.wall = in[23] != 0.0;
    return a;
}

extern "C" void law_fwd_c(const double *in, const double *p, double dt, int update, double *out)
{
    const LawIn a = unpack(in);
    const LawP q{p[0], p[1], p[2], p[3]};
    LawOut o;
    dem::law_fwd(a, q, dt, update, o);
    for (int k = 0; k < 3; k++) {
        out[k] = o.F[k];
        out[3 + k] = o.T[k];
        out[6 + k] = o.sh[k];
    }
}

// seed[9] = adjoints of (F, T, sh); gin[18] = adjoints of d, vi, vj, wi, wj, sh;
// gp[4] = adjoints of Yeff, Geff, betaeff, mu
extern "C" void law_vjp_c(const double *in, const double *p, double dt, int update, const double *seed,
                          double *gin, double *gp)
{
    const LawIn a = unpack(in);
    const LawP q{p[0], p[1], p[2], p[3]};
    LawIn g = {};
    LawP h{0, 0, 0, 0};
    dem::law_vjp(a, q, dt, update, seed, seed + 3, seed + 6, g, h);
    for (int k = 0; k < 3; k++) {
        gin[k] = g.d[k];
        gin[3 + k] = g.vi[k];
        gin[6 + k] = g.vj[k];
        gin[9 + k] = g.wi[k];
        gin[12 + k] = g.wj[k];
        gin[15 + k] = g.sh[k];
    }
    gp[0] = h.Y;
    gp[1] = h.G;
    gp[2] = h.B;
    gp[3] = h.MU;
}
