/*
 * insert_oracle.c -- CPU restatement of fix insert/pack (test infrastructure
 * only; see insert_oracle.h).  Reference files are LIGGGHTS-INL src/.
 */
#include "insert_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[256];

const char *oi_error(void) { return g_err; }

/* ---- RanPark (random_park.cpp:44-103) ---- */
typedef struct {
    int seed, save;
    double second;
} park;

static void park_reset(park *p, int s)
{
    p->seed = s;
    p->save = 0;
}

static double park_uniform(park *p)
{
    /* Schrage: IA 16807, IM 2^31-1, IQ = IM / IA, IR = IM % IA */
    const int k = p->seed / 127773;
    p->seed = 16807 * (p->seed - k * 127773) - 2836 * k;
    if (p->seed < 0) p->seed += 2147483647;
    return (1.0 / 2147483647) * p->seed;
}

static double park_gaussian(park *p)
{
    double first, v1, v2, rsq, fac;
    if (!p->save) {
        int again = 1;
        while (again) {
            v1 = 2.0 * park_uniform(p) - 1.0;
            v2 = 2.0 * park_uniform(p) - 1.0;
            rsq = v1 * v1 + v2 * v2;
            if (rsq < 1.0 && rsq != 0.0) again = 0;
        }
        fac = sqrt(-2.0 * log(rsq) / rsq);
        p->second = v1 * fac;
        first = v2 * fac;
        p->save = 1;
    } else {
        first = p->second;
        p->save = 0;
    }
    return first;
}

void oi_park_states(int seed, int n, int *states, double *values)
{
    park p;
    park_reset(&p, seed);
    for (int i = 0; i < n; i++) {
        const double u = park_uniform(&p);
        if (states) states[i] = p.seed;
        if (values) values[i] = u;
    }
}

/* ---- fix particletemplate/multiplespheres (fix_template_multiplespheres.cpp) ---- */
struct oi_tmpl {
    int n;
    double *x, *r;
    int *types, atom_type, bonded;
    double density;
    double x_min[3], x_max[3], x_bound[3], r_bound, xcm[3], volume;
    park rins;   /* random_insertion: the seed plus the process shift 0 */
};

oi_tmpl *oi_tmpl_create(int seed, int nspheres, const double *x, const double *r, const int *types, int atom_type,
                        double density, double scale, int ntry, int bonded)
{
    oi_tmpl *t = (oi_tmpl *)calloc(1, sizeof(oi_tmpl));
    const int n = nspheres;
    t->n = n;
    t->atom_type = atom_type;
    t->density = density;
    t->bonded = bonded;
    t->types = NULL;
    if (types) {
        t->types = (int *)malloc(sizeof(int) * n);
        memcpy(t->types, types, sizeof(int) * n);
    }
    park_reset(&t->rins, seed);
    t->x = (double *)malloc(sizeof(double) * 3 * n);
    t->r = (double *)malloc(sizeof(double) * n);
    /* constructor, file branch (:118-150): r *= scale*cg, x = (scale*cg) x,
     * then the bounding box; cg = 1 */
    for (int d = 0; d < 3; d++) {
        t->x_min[d] = 1e8;
        t->x_max[d] = -1e8;
    }
    for (int i = 0; i < n; i++) {
        t->r[i] = r[i] * (scale * 1.0);
        for (int j = 0; j < 3; j++) t->x[3 * i + j] = (scale * 1.0) * x[3 * i + j];
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 3; j++) {
            if (t->x[3 * i + j] - t->r[i] < t->x_min[j]) t->x_min[j] = t->x[3 * i + j] - t->r[i];
            if (t->x[3 * i + j] + t->r[i] > t->x_max[j]) t->x_max[j] = t->x[3 * i + j] + t->r[i];
        }
    park mc;
    park_reset(&mc, seed);   /* random_mc: the seed, no process shift */

    /* calc_bounding_sphere (:321-375) */
    t->r_bound = 1e8;
    int *visited = (int *)malloc(sizeof(int) * n);
    for (int shuffle = 0; shuffle < 200; shuffle++) {
        for (int i = 0; i < n; i++) visited[i] = 0;
        int isphere = -1, nvisited = 0;
        double xbt[3], rbt, d[3], dist;
        while (isphere < 0 || visited[isphere] || isphere >= n) isphere = (int)(park_uniform(&mc) * n);
        nvisited++;
        visited[isphere] = 1;
        xbt[0] = t->x[3 * isphere];
        xbt[1] = t->x[3 * isphere + 1];
        xbt[2] = t->x[3 * isphere + 2];
        rbt = t->r[isphere];
        while (nvisited < n) {
            while (isphere < 0 || visited[isphere] || isphere >= n) isphere = (int)(park_uniform(&mc) * n);
            nvisited++;
            visited[isphere] = 1;
            d[0] = t->x[3 * isphere] - xbt[0];
            d[1] = t->x[3 * isphere + 1] - xbt[1];
            d[2] = t->x[3 * isphere + 2] - xbt[2];
            dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (dist + t->r[isphere] > rbt) {
                const double fact = (dist + t->r[isphere] - rbt) / (2. * dist);
                d[0] = fact * d[0];
                d[1] = fact * d[1];
                d[2] = fact * d[2];
                xbt[0] = xbt[0] + d[0];
                xbt[1] = xbt[1] + d[1];
                xbt[2] = xbt[2] + d[2];
                rbt += sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            }
        }
        if (rbt < t->r_bound) {
            t->r_bound = rbt;
            t->x_bound[0] = xbt[0];
            t->x_bound[1] = xbt[1];
            t->x_bound[2] = xbt[2];
        }
    }
    free(visited);

    /* calc_center_of_mass (:410-450), brute force over the spheres */
    int nsuccess = 0;
    double xcm[3] = {0., 0., 0.};
    for (int i = 0; i < ntry; i++) {
        double xt[3];
        xt[0] = t->x_min[0] + (t->x_max[0] - t->x_min[0]) * park_uniform(&mc);
        xt[1] = t->x_min[1] + (t->x_max[1] - t->x_min[1]) * park_uniform(&mc);
        xt[2] = t->x_min[2] + (t->x_max[2] - t->x_min[2]) * park_uniform(&mc);
        for (int j = 0; j < n; j++) {
            double dsq = 0.;
            dsq += (xt[0] - t->x[3 * j]) * (xt[0] - t->x[3 * j]);
            dsq += (xt[1] - t->x[3 * j + 1]) * (xt[1] - t->x[3 * j + 1]);
            dsq += (xt[2] - t->x[3 * j + 2]) * (xt[2] - t->x[3 * j + 2]);
            if (dsq < t->r[j] * t->r[j]) {
                xcm[0] = (xcm[0] * (double)nsuccess + xt[0]) / (double)(nsuccess + 1);
                xcm[1] = (xcm[1] * (double)nsuccess + xt[1]) / (double)(nsuccess + 1);
                xcm[2] = (xcm[2] * (double)nsuccess + xt[2]) / (double)(nsuccess + 1);
                nsuccess++;
                break;
            }
        }
    }
    t->volume = (double)nsuccess / (double)ntry * (t->x_max[0] - t->x_min[0]) * (t->x_max[1] - t->x_min[1])
                * (t->x_max[2] - t->x_min[2]);
    for (int d = 0; d < 3; d++) t->xcm[d] = xcm[d];
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) t->x[3 * i + d] = t->x[3 * i + d] - xcm[d];
    for (int d = 0; d < 3; d++) {
        t->x_min[d] = t->x_min[d] - xcm[d];
        t->x_max[d] = t->x_max[d] - xcm[d];
        t->x_bound[d] = t->x_bound[d] - xcm[d];
    }
    return t;
}

void oi_tmpl_destroy(oi_tmpl *t)
{
    if (!t) return;
    free(t->x);
    free(t->r);
    free(t->types);
    free(t);
}

void oi_tmpl_info(const oi_tmpl *t, double out[9])
{
    out[0] = t->r_bound;
    for (int d = 0; d < 3; d++) out[1 + d] = t->x_bound[d];
    out[4] = t->volume;
    out[5] = t->volume * t->density;
    for (int d = 0; d < 3; d++) out[6 + d] = t->xcm[d];
}

/* MathExtra::quatquat (math_extra.h:596-602) */
static void oi_quatquat(const double *a, const double *b, double *c)
{
    c[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    c[1] = a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2];
    c[2] = a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3];
    c[3] = a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1];
}

/* ParticleToInsert::set_x_v_omega (particleToInsert.cpp:330-350) with
 * MathExtraLiggghts::vec_quat_rotate (math_extra_liggghts.h:457-473), the
 * orientation as given */
static void tmpl_place_q(const oi_tmpl *t, const double x[3], const double q[4], double *out)
{
    for (int j = 0; j < t->n; j++) {
        double rel[3], vq[4], qc[4], tmp[4], res[4];
        for (int d = 0; d < 3; d++) rel[d] = t->x[3 * j + d] - t->x_bound[d];
        vq[0] = 0.;
        vq[1] = rel[0];
        vq[2] = rel[1];
        vq[3] = rel[2];
        qc[0] = q[0];
        qc[1] = -q[1];
        qc[2] = -q[2];
        qc[3] = -q[3];
        oi_quatquat(q, vq, tmp);
        oi_quatquat(tmp, qc, res);
        for (int d = 0; d < 3; d++) out[3 * j + d] = res[1 + d] + x[d];
    }
}

void oi_tmpl_place(const oi_tmpl *t, const double x[3], const double *quat, double *out)
{
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    if (quat) {
        /* MathExtraLiggghts::quat_normalize (math_extra_liggghts.h:513-520) */
        const double norm = 1.0 / sqrt(quat[0] * quat[0] + quat[1] * quat[1] + quat[2] * quat[2] + quat[3] * quat[3]);
        for (int k = 0; k < 4; k++) q[k] = quat[k] * norm;
    }
    tmpl_place_q(t, x, q, out);
}

/* ---- the fix ---- */
struct oi_fix {
    oi_spec s;
    park random, region_random, dist_random;
    double distweight[OI_MAXT], volexp[OI_MAXT], massexp[OI_MAXT], rbound[OI_MAXT];
    int distorder[OI_MAXT];
    double volexpect, massexpect, max_r_bound;
    double region_volume, region_volume_local, insertion_ratio;
    long long next_reneighbor, most_recent;
    long long ninserted;
    double massinserted;
    int last_inserted, last_aimed, last_tries;
    /* fix insert/stream */
    park mesh_random;               /* the face mesh's stream */
    double *area_acc, area;
    double nrm[3], vn[3], pref[3];
    double extrude, ext_min, ext_max;
    double ninsert_per;
    int ninsert, ninsert_exists, do_frac;
    double ins_fraction;
    double *rel;                    /* release data of the last insertion */
    int nrel;
};

static double o_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double o_mag(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* fix insert/stream's set-up: the face's element properties (SurfaceMesh::
 * calcEdgeVecLen, calcSurfaceNorm, TriMesh::calcArea; surface_mesh_I.h:
 * 187-217, tri_mesh_I.h:327-338; no degenerate triangles) and
 * FixInsertStream::calc_insertion_properties (fix_insert_stream.cpp:287-423) */
static int stream_setup(oi_fix *f)
{
    const oi_spec *s = &f->s;
    const int nt = s->face_ntri;
    if (nt < 1 || !s->face) {
        snprintf(g_err, sizeof g_err, "must define an insertion face");
        return -1;
    }
    f->area_acc = (double *)malloc(sizeof(double) * nt);
    f->area = 0.;
    double n0[3] = {0, 0, 0};
    for (int t = 0; t < nt; t++) {
        const double *nd = s->face + 9 * t;
        double ev[2][3], len[2];
        for (int k = 0; k < 2; k++) {
            for (int q = 0; q < 3; q++) ev[k][q] = nd[3 * (k + 1) + q] - nd[3 * k + q];
            len[k] = o_mag(ev[k]);
            const double inv = 1. / len[k];
            for (int q = 0; q < 3; q++) ev[k][q] = inv * ev[k][q];
        }
        double c[3];
        c[0] = ev[0][1] * ev[1][2] - ev[0][2] * ev[1][1];
        c[1] = ev[0][2] * ev[1][0] - ev[0][0] * ev[1][2];
        c[2] = ev[0][0] * ev[1][1] - ev[0][1] * ev[1][0];
        if (t == 0) {
            const double inv = 1. / o_mag(c);
            for (int q = 0; q < 3; q++) n0[q] = inv * c[q];
        }
        const double a = 0.5 * o_mag(c) * len[0] * len[1];
        f->area_acc[t] = t > 0 ? a + f->area_acc[t - 1] : a;
        f->area += a;
    }
    park_reset(&f->mesh_random, 179424799);   /* MultiNodeMesh (multi_node_mesh_I.h:61) */
    /* the normal of face 0 turned toward the insertion velocity */
    memcpy(f->nrm, n0, sizeof n0);
    if (o_dot(s->vel, f->nrm) < 0)
        for (int q = 0; q < 3; q++) f->nrm[q] = f->nrm[q] * -1.;
    const double dot = o_dot(s->vel, f->nrm);
    for (int q = 0; q < 3; q++) f->vn[q] = f->nrm[q] * dot;
    memcpy(f->pref, s->face, 3 * sizeof(double));
    const double vnm = o_mag(f->vn);
    if (vnm < 1.e-3 || o_mag(s->vel) < 1e-5) {
        snprintf(g_err, sizeof g_err, "insertion velocity too low");
        return -1;
    }
    f->extrude = s->extrude_length;
    if (f->s.insert_every == -1) {
        f->s.insert_every = (int)((f->extrude + 1e-14) / (s->dt * vnm));
        if (f->s.insert_every == 0) {
            snprintf(g_err, sizeof g_err, "insertion velocity too high or extrude_length too low");
            return -1;
        }
    } else {
        const int duration = s->duration > 0 ? s->duration : f->s.insert_every;
        f->extrude = (double)duration * s->dt * vnm;
    }
    f->ninsert_exists = !(s->nparticles < 0 || s->mass < 0.);
    f->ninsert = s->nparticles > 0 ? s->nparticles : 0;
    if (f->ninsert == 0 && f->ninsert_exists) {
        if (!(s->mass > 0.)) {
            snprintf(g_err, sizeof g_err, "must define either 'nparticles' or 'mass'");
            return -1;
        }
        f->ninsert = (int)((s->mass + 1e-14) / f->massexpect);
    }
    double nflowrate = s->particlerate;
    if (nflowrate == 0.) nflowrate = s->massrate / f->massexpect;
    f->ninsert_per = nflowrate * ((double)f->s.insert_every * s->dt);
    f->ext_min = 0.;
    f->ext_max = f->extrude;
    f->do_frac = 1;
    return 0;
}

/* TriMesh::generateRandomOwnedGhost (tri_mesh_I.h:395-428) with
 * SurfaceMesh::randomOwnedGhostElement and searchElementByAreaAcc
 * (surface_mesh_I.h:264-295) */
static int area_search(const double *acc, double a, int lo, int hi)
{
    if ((lo < 1 || a > acc[lo - 1]) && (a <= acc[lo])) return lo;
    if ((hi < 1 || a > acc[hi - 1]) && (a <= acc[hi])) return hi;
    const int mid = (lo + hi) / 2;
    if (a > acc[mid]) return area_search(acc, a, mid, hi);
    return area_search(acc, a, lo, mid);
}

static void face_point(oi_fix *f, double *pos)
{
    const double a = park_uniform(&f->mesh_random) * (f->area + 0.);
    const int t = area_search(f->area_acc, a, 0, f->s.face_ntri - 1);
    double u, v;
    do {
        u = park_uniform(&f->mesh_random);
        v = park_uniform(&f->mesh_random);
    } while (u + v > 1);
    const double b0 = 1. - u - v, b1 = v, b2 = u;
    const double *nd = f->s.face + 9 * t;
    for (int q = 0; q < 3; q++) pos[q] = b0 * nd[q] + b1 * nd[3 + q] + b2 * nd[6 + q];
}

/* Domain::is_in_subdomain (domain_I.h:65-88), one sub-box padded by
 * SMALL_DMBRDR (1e-8) on every face */
static int in_subdomain(const oi_fix *f, const double *p)
{
    for (int d = 0; d < 3; d++)
        if (!(p[d] >= f->s.boxlo[d] - 1.0e-8 && p[d] < f->s.boxhi[d] + 1.0e-8)) return 0;
    return 1;
}

/* FixInsertStream::generate_random (fix_insert_stream.cpp:617-641) */
static void stream_random(oi_fix *f, double *pos, double rad)
{
    face_point(f, pos);
    double r;
    if (f->s.overlapcheck) r = -1. * (park_uniform(&f->random) * (f->ext_max) + rad + f->ext_min);
    else r = -1. * (park_uniform(&f->random) * (f->ext_max - 2. * rad) + rad + f->ext_min);
    for (int q = 0; q < 3; q++) pos[q] = pos[q] + f->nrm[q] * r;
}

/* FixInsertStream::calc_ins_fraction (fix_insert_stream.cpp:493-556) */
static int stream_fraction(oi_fix *f)
{
    f->do_frac = 0;
    int n_in = 0;
    for (int i = 0; i < f->s.ntry_mc; i++) {
        double pos[3];
        face_point(f, pos);
        const double r = -1. * (park_uniform(&f->random) * f->extrude);
        for (int q = 0; q < 3; q++) pos[q] = pos[q] + f->nrm[q] * r;
        if (in_subdomain(f, pos)) n_in++;
    }
    f->ins_fraction = (double)n_in / (double)f->s.ntry_mc;
    if (f->s.parallel) {
        f->ext_min = f->extrude;
        f->ext_max = 0.;
        for (int c = 0; c < 8; c++) {
            double be[3];
            for (int d = 0; d < 3; d++) be[d] = (((c >> d) & 1) ? f->s.boxhi[d] : f->s.boxlo[d]) - f->pref[d];
            const double dot = -o_dot(be, f->nrm);
            if (dot > 0. && dot < f->extrude) {
                if (dot > f->ext_max) f->ext_max = dot;
                if (dot < f->ext_min) f->ext_min = dot;
            } else if (dot < 0.) {
                f->ext_min = 0.;
            } else if (dot >= f->extrude) {
                f->ext_max = f->extrude;
            }
        }
        if (f->ext_min == f->extrude) f->ext_min = 0.;
        if (f->ext_max == 0.) f->ext_max = f->extrude;
    }
    if (f->ins_fraction < 0.9 || f->ins_fraction > 1.1) {
        snprintf(g_err, sizeof g_err, "insertion volume could not be distributed properly");
        return -1;
    }
    return 0;
}

/* TriMesh::isInElement (tri_mesh_I.h:345-368) */
static int in_element(const double *nd, const double *p)
{
    double v0[3], v1[3], v2[3];
    for (int q = 0; q < 3; q++) {
        v0[q] = nd[6 + q] - nd[q];
        v1[q] = nd[3 + q] - nd[q];
        v2[q] = p[q] - nd[q];
    }
    const double d00 = o_dot(v0, v0), d01 = o_dot(v0, v1), d02 = o_dot(v0, v2);
    const double d11 = o_dot(v1, v1), d12 = o_dot(v1, v2);
    const double inv = 1. / (d00 * d11 - d01 * d01);
    const double u = (d11 * d02 - d01 * d12) * inv;
    const double v = (d00 * d12 - d01 * d02) * inv;
    return u > -1e-10 && v > -1e-10 && u + v < 1. + 1e-10;
}

/* FixInsertStream::is_nearby (fix_insert_stream.cpp:568-593) and the
 * bounding box of FixInsert::load_xnear / FixInsertStream::getBoundingBox
 * (fix_insert.cpp:980-1018, fix_insert_stream.cpp:597-610) */
static int stream_nearby(const oi_fix *f, const double *x, double maxrad)
{
    double rel[3];
    for (int q = 0; q < 3; q++) rel[q] = x[q] - f->pref[q];
    const double d = o_dot(rel, f->nrm);
    if (d > maxrad) return 0;
    if (d < -(f->extrude + 3. * maxrad)) return 0;
    double pr[3];
    for (int q = 0; q < 3; q++) pr[q] = x[q] + f->nrm[q] * d;
    int on = 0;
    for (int t = 0; t < f->s.face_ntri && !on; t++) on = in_element(f->s.face + 9 * t, pr);
    if (!on) return 0;
    double lo[3], hi[3];
    for (int q = 0; q < 3; q++) lo[q] = hi[q] = f->s.face[q];
    for (int k = 0; k < 9 * f->s.face_ntri; k++) {
        if (f->s.face[k] < lo[k % 3]) lo[k % 3] = f->s.face[k];
        if (f->s.face[k] > hi[k % 3]) hi[k % 3] = f->s.face[k];
    }
    const double cut = 3 * maxrad, delta = -(f->extrude + 2 * cut);
    const double extend = 3 * maxrad + 2. * f->max_r_bound;
    for (int q = 0; q < 3; q++) {
        const double a = lo[q] + delta * f->nrm[q], b = hi[q] + delta * f->nrm[q];
        if (a < lo[q]) lo[q] = a;
        if (b > hi[q]) hi[q] = b;
        if (lo[q] < f->s.boxlo[q]) lo[q] = f->s.boxlo[q];
        if (hi[q] > f->s.boxhi[q]) hi[q] = f->s.boxhi[q];
        if (!(x[q] >= lo[q] - extend && x[q] <= hi[q] + extend)) return 0;
    }
    return 1;
}

int oi_release(const oi_fix *f, int cap, double *out)
{
    const int n = f->nrel < cap ? f->nrel : cap;
    if (n > 0) memcpy(out, f->rel, sizeof(double) * 14 * n);
    return f->nrel;
}

void oi_stream_info(const oi_fix *f, double out[9])
{
    out[0] = f->s.insert_every;
    out[1] = f->ninsert_per;
    out[2] = f->ins_fraction;
    for (int q = 0; q < 3; q++) {
        out[3 + q] = f->nrm[q];
        out[6 + q] = f->vn[q];
    }
}

/* RegBlock::inside (region_block.cpp:272-277) */
static int block_inside(const oi_fix *f, const double *x)
{
    const double *lo = f->s.region_lo, *hi = f->s.region_hi;
    return x[0] >= lo[0] && x[0] <= hi[0] && x[1] >= lo[1] && x[1] <= hi[1] && x[2] >= lo[2] && x[2] <= hi[2];
}

/* Region::match_cut -> RegBlock::surface_interior (region_block.cpp:286-345):
 * the number of faces closer than cut to a point inside */
static int block_surface_interior(const oi_fix *f, const double *x, double cut)
{
    const double *lo = f->s.region_lo, *hi = f->s.region_hi;
    if (x[0] < lo[0] || x[0] > hi[0] || x[1] < lo[1] || x[1] > hi[1] || x[2] < lo[2] || x[2] > hi[2]) return 0;
    int n = 0;
    for (int d = 0; d < 3; d++) {
        if (x[d] - lo[d] < cut) n++;
        if (hi[d] - x[d] < cut) n++;
    }
    return n;
}

oi_fix *oi_create(const oi_spec *s, long long ntimestep)
{
    g_err[0] = 0;
    if (s->ntemplates < 1 || s->ntemplates > OI_MAXT) {
        snprintf(g_err, sizeof g_err, "illegal number of templates");
        return NULL;
    }
    if (s->insert_every < 0 && !(s->stream && s->insert_every == -1)) {
        snprintf(g_err, sizeof g_err, "must define 'insert_every'");
        return NULL;
    }
    oi_fix *f = (oi_fix *)calloc(1, sizeof(oi_fix));
    f->s = *s;
    const int T = s->ntemplates;
    if (s->stream) f->s.check_border = 1;
    /* FixTemplateSphere (fix_template_sphere.cpp:243-245), constant radius
     * and density: volume_expect = r^3 * 4 pi / 3, mass_expect = rho * that */
    f->max_r_bound = 0.0;
    for (int i = 0; i < T; i++) {
        const oi_tmpl *ms = (const oi_tmpl *)s->ms[i];
        if (ms) {
            /* FixTemplateMultiplespheres: Monte Carlo volume and mass, bounding radius */
            f->volexp[i] = ms->volume;
            f->massexp[i] = ms->volume * ms->density;
            f->rbound[i] = ms->r_bound;
        } else {
            const double r = s->radius[i];
            f->volexp[i] = r * r * r * 4. * M_PI / 3.;
            f->massexp[i] = s->density[i] * f->volexp[i];
            f->rbound[i] = r;
        }
        if (f->rbound[i] > f->max_r_bound) f->max_r_bound = f->rbound[i];
    }
    /* FixParticledistributionDiscrete constructor
     * (fix_particledistribution_discrete.cpp:126-233) */
    double weightsum = 0;
    for (int i = 0; i < T; i++) {
        f->distweight[i] = s->weight[i];
        weightsum += f->distweight[i];
    }
    for (int i = 0; i < T; i++) f->distweight[i] /= weightsum;
    if (!s->number_based) {
        for (int i = 0; i < T; i++) f->distweight[i] = f->distweight[i] / f->massexp[i];
        weightsum = 0;
        for (int i = 0; i < T; i++) weightsum += f->distweight[i];
        for (int i = 0; i < T; i++) f->distweight[i] /= weightsum;
    }
    f->volexpect = f->massexpect = 0.;
    for (int i = 0; i < T; i++) {
        f->volexpect += f->volexp[i] * f->distweight[i];
        f->massexpect += f->massexp[i] * f->distweight[i];
    }
    for (int i = 0; i < T; i++) f->distorder[i] = i;
    int swapped, n = T;
    do {
        swapped = 0;
        for (int i = 0; i < T - 1; i++)
            if (f->volexp[f->distorder[i]] < f->volexp[f->distorder[i + 1]]) {
                const int tmp = f->distorder[i];
                f->distorder[i] = f->distorder[i + 1];
                f->distorder[i + 1] = tmp;
                swapped = 1;
            }
        n--;
    } while (swapped && n > 0);

    /* FixInsert constructor (fix_insert.cpp:101, :369-370, :458) */
    park_reset(&f->random, s->seed);
    park_reset(&f->dist_random, s->dist_seed);
    f->next_reneighbor = s->start > 0 ? s->start : ntimestep + 1;
    f->most_recent = -1;
    if (s->stream) {
        if (stream_setup(f)) {
            oi_destroy(f);
            return NULL;
        }
        return f;
    }

    /* FixInsert::setup -> FixInsertPack::calc_insertion_properties
     * (fix_insert_pack.cpp:281-322): reset the region's stream to
     * seed + SEED_OFFSET (12), then Region::volume_mc (region.cpp:633-700) */
    park_reset(&f->region_random, s->seed + 12);
    int n_in_global = 0, n_in_local = 0;
    const double *lo = s->region_lo, *hi = s->region_hi;
    for (int i = 0; i < s->ntry_mc; i++) {
        double pos[3];
        pos[0] = lo[0] + park_uniform(&f->region_random) * (hi[0] - lo[0]);
        pos[1] = lo[1] + park_uniform(&f->region_random) * (hi[1] - lo[1]);
        pos[2] = lo[2] + park_uniform(&f->region_random) * (hi[2] - lo[2]);
        /* Domain::is_in_domain (domain_I.h:51-63) */
        if (!(pos[0] >= s->boxlo[0] && pos[0] <= s->boxhi[0] && pos[1] >= s->boxlo[1] && pos[1] <= s->boxhi[1]
              && pos[2] >= s->boxlo[2] && pos[2] <= s->boxhi[2]))
            continue;
        if (!s->all_in) {
            if (block_inside(f, pos)) {
                n_in_global++;
                n_in_local++;   /* one process: the sub-box is the box */
            }
        } else {
            if (block_inside(f, pos)) {
                n_in_global++;
                if (!block_surface_interior(f, pos, f->max_r_bound)) n_in_local++;
            }
        }
    }
    if (n_in_global == 0) {
        snprintf(g_err, sizeof g_err, "Unable to calculate region volume");
        free(f);
        return NULL;
    }
    const double vol_bbox = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    f->region_volume = (double)n_in_global / (double)(s->ntry_mc * 1) * vol_bbox;
    double vol_local = (double)n_in_local / (double)s->ntry_mc * vol_bbox;
    const double vol_local_all = vol_local;
    vol_local *= (f->region_volume / vol_local_all);
    f->region_volume_local = vol_local;
    return f;
}

void oi_destroy(oi_fix *f)
{
    if (!f) return;
    free(f->area_acc);
    free(f->rel);
    free(f);
}

long long oi_next_step(const oi_fix *f) { return f->next_reneighbor; }

double oi_region_volume(const oi_fix *f) { return f->region_volume; }

void oi_vector(const oi_fix *f, double out[6])
{
    out[0] = (double)f->ninserted;
    out[1] = f->massinserted;
    out[2] = f->last_inserted;
    out[3] = f->last_aimed;
    out[4] = f->last_tries;
    out[5] = f->next_reneighbor > 0 ? (double)f->next_reneighbor : 0.0;
}

/* Region::generate_random / generate_random_shrinkby_cut
 * (region.cpp:506-571), sub-domain = box */
static void region_random_point(oi_fix *f, double *pos, double cut)
{
    double lo[3], hi[3], diff[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = f->s.region_lo[d] > f->s.boxlo[d] ? f->s.region_lo[d] : f->s.boxlo[d];
        hi[d] = f->s.region_hi[d] < f->s.boxhi[d] ? f->s.region_hi[d] : f->s.boxhi[d];
        diff[d] = hi[d] - lo[d];
    }
    for (;;) {
        pos[0] = lo[0] + park_uniform(&f->region_random) * diff[0];
        pos[1] = lo[1] + park_uniform(&f->region_random) * diff[1];
        pos[2] = lo[2] + park_uniform(&f->region_random) * diff[2];
        if (!block_inside(f, pos)) continue;
        if (f->s.all_in && block_surface_interior(f, pos, cut)) continue;
        return;
    }
}

/* FixInsert::generate_random_velocity (fix_insert.cpp:1025-1040) */
static void random_velocity(oi_fix *f, double *v)
{
    if (f->s.vel_style == 1) {
        v[0] = f->s.vel[0] + f->s.vel_fluct[0] * 2.0 * (park_uniform(&f->random) - 0.50);
        v[1] = f->s.vel[1] + f->s.vel_fluct[1] * 2.0 * (park_uniform(&f->random) - 0.50);
        v[2] = f->s.vel[2] + f->s.vel_fluct[2] * 2.0 * (park_uniform(&f->random) - 0.50);
    } else if (f->s.vel_style == 2) {
        v[0] = f->s.vel[0] + f->s.vel_fluct[0] * park_gaussian(&f->random);
        v[1] = f->s.vel[1] + f->s.vel_fluct[1] * park_gaussian(&f->random);
        v[2] = f->s.vel[2] + f->s.vel_fluct[2] * park_gaussian(&f->random);
    }
}

/* rsq <= radsum^2 against every resident particle and its periodic images
 * (ghost atoms, x + pbc * prd, comm) and every particle placed so far
 * (RegionNeighborList::hasOverlap, region_neighbor_list_I.h:72-98) */
static int overlaps(const oi_fix *f, const double *x, double radius, int n, const double *xr, const double *rr,
                    int nnew, const double *xn, const double *rn)
{
    double prd[3];
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        prd[d] = f->s.boxhi[d] - f->s.boxlo[d];
        lo[d] = f->s.periodic[d] ? -1 : 0;
        hi[d] = f->s.periodic[d] ? 1 : 0;
    }
    for (int j = 0; j < n; j++)
        for (int a = lo[0]; a <= hi[0]; a++)
            for (int b = lo[1]; b <= hi[1]; b++)
                for (int c = lo[2]; c <= hi[2]; c++) {
                    double p[3], del[3];
                    p[0] = a ? xr[3 * j] + a * prd[0] : xr[3 * j];
                    p[1] = b ? xr[3 * j + 1] + b * prd[1] : xr[3 * j + 1];
                    p[2] = c ? xr[3 * j + 2] + c * prd[2] : xr[3 * j + 2];
                    del[0] = x[0] - p[0];
                    del[1] = x[1] - p[1];
                    del[2] = x[2] - p[2];
                    const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
                    const double radsum = radius + rr[j];
                    if (rsq <= radsum * radsum) return 1;
                }
    for (int j = 0; j < nnew; j++) {
        double del[3];
        del[0] = x[0] - xn[3 * j];
        del[1] = x[1] - xn[3 * j + 1];
        del[2] = x[2] - xn[3 * j + 2];
        const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        const double radsum = radius + rn[j];
        if (rsq <= radsum * radsum) return 1;
    }
    return 0;
}

/* Domain::check_dist_subbox_borders (domain_I.h:127-149), sub-box = box */
static int subbox_dist_ok(const oi_fix *f, const double *pos, double dist)
{
    for (int d = 0; d < 3; d++) {
        const double dlo = fabs(f->s.boxlo[d] - pos[d]);
        const double dhi = fabs(f->s.boxhi[d] - pos[d]);
        if (dlo < dist || dhi < dist) return 0;
    }
    return 1;
}

/* MathExtraLiggghts::random_unit_quat (math_extra_liggghts.h:616-629) */
static void unit_quat(park *p, double *quat)
{
    const double u1 = park_uniform(p);
    const double u2 = park_uniform(p);
    const double u3 = park_uniform(p);
    const double h1 = sqrt(1. - u1);
    const double h2 = sqrt(u1);
    quat[0] = h1 * sin(2. * M_PI * u2);
    quat[1] = h1 * cos(2. * M_PI * u2);
    quat[2] = h2 * sin(2. * M_PI * u3);
    quat[3] = h2 * cos(2. * M_PI * u3);
}

/* the body's spheres: a sphere at pos, a clump through set_x_v_omega's
 * rotation; returns their number */
static int body_spheres(const oi_fix *f, int t, const double *pos, const double *q, double *xs, double *rs)
{
    const oi_tmpl *ms = (const oi_tmpl *)f->s.ms[t];
    if (!ms) {
        memcpy(xs, pos, 3 * sizeof(double));
        rs[0] = f->s.radius[t];
        return 1;
    }
    tmpl_place_q(ms, pos, q, xs);
    memcpy(rs, ms->r, ms->n * sizeof(double));
    return ms->n;
}

static double maxrad_extra = 0.0;

void oi_set_maxrad(double r) { maxrad_extra = r; }

int oi_insert(oi_fix *f, long long step, int n, const double *x, const double *r, const double *m, int cap,
              double *x_out, double *v_out, double *omega_out, double *r_out, double *m_out, int *type_out,
              double *brid_out)
{
    const oi_spec *s = &f->s;
    /* FixInsert::pre_exchange (fix_insert.cpp:676-888) */
    if (f->next_reneighbor != step || f->most_recent == step) return 0;
    f->most_recent = step;

    int ninsert_this, ninsert_this_local;
    if (s->stream) {
        /* FixInsert::calc_ninsert_this (fix_insert.cpp:658-668), the cap in
         * pre_exchange (:693-700), distribute_ninsert_this (:893-960) with
         * FixInsertStream::insertion_fraction (fix_insert_stream.cpp:475-486) */
        ninsert_this = (int)(f->ninsert_per + park_uniform(&f->random));
        if (f->ninsert_exists && f->ninserted + ninsert_this > f->ninsert)
            ninsert_this = (int)(f->ninsert - f->ninserted);
        if (f->ninsert_exists && f->ninserted + ninsert_this >= f->ninsert) {
            ninsert_this = (int)(f->ninsert - f->ninserted);
            if (ninsert_this < 0) ninsert_this = 0;
            f->s.exact_number = 1;
        }
        if (f->do_frac && stream_fraction(f)) return -1;
        ninsert_this_local = ninsert_this;
        if (!s->exact_number) ninsert_this_local = (int)(f->ins_fraction * (double)ninsert_this + park_uniform(&f->random));
        goto distribute;
    }
    /* FixInsertPack::calc_ninsert_this (fix_insert_pack.cpp:333-432); the
     * sums run over the resident particles in the order given */
    int np_region = 0;
    double vol_region = 0., mass_region = 0.;
    const double _4Pi3 = 4. * M_PI / 3.;
    for (int i = 0; i < n; i++)
        if (block_inside(f, x + 3 * i)) {
            np_region++;
            vol_region += _4Pi3 * r[i] * r[i] * r[i];
            mass_region += m[i];
        }
    if (s->target == 0) {
        ninsert_this = (int)((s->target_value * f->region_volume - vol_region) / f->volexpect
                             + park_uniform(&f->random));
        f->insertion_ratio = vol_region / (s->target_value * f->region_volume);
    } else if (s->target == 1) {
        const int ntotal_region = (int)s->target_value;
        ninsert_this = ntotal_region - np_region;
        f->insertion_ratio = (double)np_region / (double)ntotal_region;
    } else {
        ninsert_this = (int)((s->target_value - mass_region) / f->massexpect + park_uniform(&f->random));
        f->insertion_ratio = mass_region / s->target_value;
    }
    if (ninsert_this < 0) ninsert_this = 0;
    if (f->insertion_ratio < 0.) f->insertion_ratio = 0.;
    if (f->insertion_ratio > 1.) f->insertion_ratio = 1.;

    /* distribute_ninsert_this (fix_insert.cpp:893-970), one process: the
     * exact form keeps the number; the uncorrelated one draws */
    ninsert_this_local = ninsert_this;
    if (!s->exact_number)
        ninsert_this_local = (int)(f->region_volume_local / f->region_volume * (double)ninsert_this
                                   + park_uniform(&f->random));
distribute:;

    /* FixParticledistributionDiscrete::randomize_list
     * (fix_particledistribution_discrete.cpp:388-470) */
    const int T = s->ntemplates;
    int parttogen[OI_MAXT];
    if (!s->exact_number) {
        for (int i = 0; i < T; i++)
            parttogen[i] = (int)((double)ninsert_this_local * f->distweight[i] + park_uniform(&f->dist_random));
    } else {
        int ninsert_truncated = 0;
        double remainder[OI_MAXT];
        for (int i = 0; i < T; i++) {
            parttogen[i] = (int)((double)ninsert_this_local * f->distweight[i]);
            ninsert_truncated += parttogen[i];
            remainder[i] = (double)ninsert_this_local * f->distweight[i] - (double)parttogen[i];
        }
        const int ninsert_gap = ninsert_this_local - ninsert_truncated;
        for (int i = 0; i < ninsert_gap; i++) {
            const double rr = park_uniform(&f->dist_random) * (double)ninsert_gap;
            int j = 0;
            double rsum = remainder[0];
            while (rsum < rr && j < T - 1) {
                j++;
                rsum += remainder[j];
            }
            parttogen[j]++;
        }
    }
    int npti = 0;
    for (int i = 0; i < T; i++) npti += parttogen[i];
    /* each template's randomize_ptilist in template order: a bonded clump
     * template draws a bond_random_id per body (fix_template_multiplespheres.cpp:650-672) */
    double *ids[OI_MAXT];
    for (int i = 0; i < T; i++) {
        ids[i] = NULL;
        oi_tmpl *ms = (oi_tmpl *)s->ms[i];
        if (ms && ms->bonded) {
            ids[i] = (double *)malloc(sizeof(double) * (parttogen[i] > 0 ? parttogen[i] : 1));
            for (int k = 0; k < parttogen[i]; k++) ids[i][k] = (double)step + park_uniform(&ms->rins);
        }
    }
    int *pti_t = (int *)malloc(sizeof(int) * (npti > 0 ? npti : 1));
    double *pti_b = (double *)malloc(sizeof(double) * (npti > 0 ? npti : 1));
    {
        int k = 0;
        for (int i = 0; i < T; i++) {
            const int c = f->distorder[i];
            for (int j = 0; j < parttogen[c]; j++) {
                pti_b[k] = ids[c] ? ids[c][j] : -1.0;
                pti_t[k++] = c;
            }
        }
    }
    for (int i = 0; i < T; i++) free(ids[i]);
    ninsert_this_local = npti;
    f->last_aimed = npti;
    f->last_inserted = 0;
    f->last_tries = 0;
    /* FixInsert's `!ninsert_exists || ninserted < ninsert` */
    const int more = !s->stream || !f->ninsert_exists || f->ninserted < f->ninsert;
    f->nrel = 0;
    if (npti == 0) {
        if (s->insert_every && more) f->next_reneighbor += s->insert_every;
        else if (s->insert_every == 0) f->next_reneighbor = -1;
        free(pti_t);
        free(pti_b);
        return 0;
    }

    /* FixInsertPack::calc_maxtry (fix_insert_pack.cpp:485-490); the stream
     * tries ninsert * maxattempt (fix_insert_stream.cpp:680) */
    int maxtry;
    if (s->stream || f->insertion_ratio >= 1.) maxtry = ninsert_this_local * s->maxattempt;
    else maxtry = (int)((double)(ninsert_this_local * s->maxattempt) / (1. - f->insertion_ratio));

    int maxsph = 1;
    for (int i = 0; i < T; i++)
        if (s->ms[i] && ((oi_tmpl *)s->ms[i])->n > maxsph) maxsph = ((oi_tmpl *)s->ms[i])->n;
    /* the accepted bodies: position, velocity, orientation */
    double *bx = (double *)malloc(sizeof(double) * 3 * npti);
    double *bv = (double *)malloc(sizeof(double) * 3 * npti);
    double *bq = (double *)malloc(sizeof(double) * 4 * npti);
    /* the spheres placed so far (neighList) */
    double *xn = (double *)malloc(sizeof(double) * 3 * (size_t)npti * maxsph);
    double *rn = (double *)malloc(sizeof(double) * (size_t)npti * maxsph);
    double *xs = (double *)malloc(sizeof(double) * 3 * maxsph);
    double *rs = (double *)malloc(sizeof(double) * maxsph);
    int nnew = 0;
    /* quat_insert: identity, or the constant orientation (fix_insert.cpp:468, :316-330) */
    double quat_insert[4] = {1.0, 0.0, 0.0, 0.0};
    if (s->orientation == 2)
        for (int k = 0; k < 4; k++) quat_insert[k] = s->quat[k];

    /* FixInsertPack::x_v_omega (fix_insert_pack.cpp:498-605) */
    int ninserted_this_local = 0, ntry = 0;
    double mass_inserted_this_local = 0.;
    /* the largest radius of any atom or template (Modify::get_max_radius,
     * modify_liggghts.cpp:575-620) for FixInsertStream::is_nearby */
    double maxrad = 0.0;
    if (s->stream) {
        /* a template's max_rad is its largest sphere
         * (fix_template_multiplespheres.cpp:556-564, fix_template_sphere.cpp:420-424) */
        for (int i = 0; i < T; i++) {
            const oi_tmpl *ms = (const oi_tmpl *)s->ms[i];
            if (ms)
                for (int j = 0; j < ms->n; j++) maxrad = ms->r[j] > maxrad ? ms->r[j] : maxrad;
            else
                maxrad = s->radius[i] > maxrad ? s->radius[i] : maxrad;
        }
        for (int i = 0; i < n; i++) maxrad = r[i] > maxrad ? r[i] : maxrad;
        if (maxrad_extra > maxrad) maxrad = maxrad_extra;
    }
    if (s->stream && !s->overlapcheck) {
        /* FixInsertStream::x_v_omega without overlap check
         * (fix_insert_stream.cpp:684-716) */
        for (int itotal = 0; itotal < ninsert_this_local; itotal++) {
            const int t = pti_t[ninserted_this_local];
            double pos[3];
            do {
                stream_random(f, pos, f->rbound[t]);
                ntry++;
            } while (ntry < maxtry && !in_subdomain(f, pos));
            if (ntry < maxtry) {
                if (s->orientation == 1) unit_quat(&f->random, quat_insert);
                const int k = ninserted_this_local;
                memcpy(bx + 3 * k, pos, sizeof pos);
                memcpy(bv + 3 * k, f->vn, 3 * sizeof(double));
                memcpy(bq + 4 * k, quat_insert, sizeof quat_insert);
                mass_inserted_this_local += s->ms[t] ? ((oi_tmpl *)s->ms[t])->volume * ((oi_tmpl *)s->ms[t])->density
                                                     : s->density[t] * (s->radius[t] * s->radius[t] * s->radius[t] * 4. * M_PI / 3.);
                ninserted_this_local++;
            }
        }
    } else if (s->stream) {
        /* with overlap check (fix_insert_stream.cpp:717-758): residents that
         * is_nearby keeps, periodic images as ghosts hold them */
        double prd[3];
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            prd[d] = s->boxhi[d] - s->boxlo[d];
            lo[d] = s->periodic[d] ? -1 : 0;
            hi[d] = s->periodic[d] ? 1 : 0;
        }
        int nnear = 0;
        double *xnear = (double *)malloc(sizeof(double) * 3 * (size_t)(27 * n + 1));
        double *rnear = (double *)malloc(sizeof(double) * (size_t)(27 * n + 1));
        for (int j = 0; j < n; j++)
            for (int a = lo[0]; a <= hi[0]; a++)
                for (int b = lo[1]; b <= hi[1]; b++)
                    for (int c = lo[2]; c <= hi[2]; c++) {
                        double p[3];
                        p[0] = a ? x[3 * j] + a * prd[0] : x[3 * j];
                        p[1] = b ? x[3 * j + 1] + b * prd[1] : x[3 * j + 1];
                        p[2] = c ? x[3 * j + 2] + c * prd[2] : x[3 * j + 2];
                        if (!stream_nearby(f, p, maxrad)) continue;
                        memcpy(xnear + 3 * nnear, p, sizeof p);
                        rnear[nnear++] = r[j];
                    }
        oi_spec flat = *s;
        for (int d = 0; d < 3; d++) flat.periodic[d] = 0;
        oi_fix g = *f;
        g.s = flat;
        while (ntry < maxtry && ninserted_this_local < ninsert_this_local) {
            const int t = pti_t[ninserted_this_local];
            const double rbound = f->rbound[t];
            int nins = 0;
            double pos[3];
            while (nins == 0 && ntry < maxtry) {
                do {
                    stream_random(f, pos, rbound);
                    ntry++;
                } while (ntry < maxtry && (!in_subdomain(f, pos) || !subbox_dist_ok(f, pos, rbound)));
                if (ntry >= maxtry) break;
                if (s->orientation == 1) unit_quat(&f->random, quat_insert);
                const int ns = body_spheres(f, t, pos, quat_insert, xs, rs);
                int free_ = 1;
                for (int j = 0; j < ns && free_; j++)
                    if (overlaps(&g, xs + 3 * j, rs[j], nnear, xnear, rnear, nnew, xn, rn)) free_ = 0;
                if (free_) {
                    for (int j = 0; j < ns; j++) {
                        memcpy(xn + 3 * nnew, xs + 3 * j, 3 * sizeof(double));
                        rn[nnew++] = rs[j];
                    }
                    nins = ns;
                }
            }
            if (nins > 0) {
                const int k = ninserted_this_local;
                memcpy(bx + 3 * k, pos, sizeof pos);
                memcpy(bv + 3 * k, f->vn, 3 * sizeof(double));
                memcpy(bq + 4 * k, quat_insert, sizeof quat_insert);
                mass_inserted_this_local += s->ms[t] ? ((oi_tmpl *)s->ms[t])->volume * ((oi_tmpl *)s->ms[t])->density
                                                     : s->density[t] * (rbound * rbound * rbound * 4. * M_PI / 3.);
                ninserted_this_local++;
            }
        }
        free(xnear);
        free(rnear);
    } else if (!s->overlapcheck) {
        for (int itotal = 0; itotal < ninsert_this_local; itotal++) {
            const int t = pti_t[ninserted_this_local];
            double pos[3], v[3];
            region_random_point(f, pos, f->rbound[t]);
            v[0] = s->vel[0];
            v[1] = s->vel[1];
            v[2] = s->vel[2];
            random_velocity(f, v);
            if (s->orientation == 1) unit_quat(&f->random, quat_insert);
            const int k = ninserted_this_local;
            memcpy(bx + 3 * k, pos, sizeof pos);
            memcpy(bv + 3 * k, v, sizeof v);
            memcpy(bq + 4 * k, quat_insert, sizeof quat_insert);
            mass_inserted_this_local += s->ms[t] ? ((oi_tmpl *)s->ms[t])->volume * ((oi_tmpl *)s->ms[t])->density
                                                 : s->density[t] * (s->radius[t] * s->radius[t] * s->radius[t] * 4. * M_PI / 3.);
            ninserted_this_local++;
        }
    } else {
        while (ntry < maxtry && ninserted_this_local < ninsert_this_local) {
            const int t = pti_t[ninserted_this_local];
            const double rbound = f->rbound[t];
            int nins = 0;
            double pos[3], v[3];
            while (nins == 0 && ntry < maxtry) {
                do {
                    region_random_point(f, pos, rbound);
                    ntry++;
                } while (s->check_border && ntry < maxtry && !subbox_dist_ok(f, pos, rbound));
                if (ntry == maxtry) break;
                v[0] = s->vel[0];
                v[1] = s->vel[1];
                v[2] = s->vel[2];
                random_velocity(f, v);
                if (s->orientation == 1) unit_quat(&f->random, quat_insert);
                /* ParticleToInsert::check_near_set_x_v_omega[_ms]
                 * (particleToInsert.cpp:246-323): every sphere free */
                const int ns = body_spheres(f, t, pos, quat_insert, xs, rs);
                int free_ = 1;
                for (int j = 0; j < ns && free_; j++)
                    if (overlaps(f, xs + 3 * j, rs[j], n, x, r, nnew, xn, rn)) free_ = 0;
                if (free_) {
                    for (int j = 0; j < ns; j++) {
                        memcpy(xn + 3 * nnew, xs + 3 * j, 3 * sizeof(double));
                        rn[nnew++] = rs[j];
                    }
                    nins = ns;
                }
            }
            if (nins > 0) {
                const int k = ninserted_this_local;
                memcpy(bx + 3 * k, pos, sizeof pos);
                memcpy(bv + 3 * k, v, sizeof v);
                memcpy(bq + 4 * k, quat_insert, sizeof quat_insert);
                mass_inserted_this_local += s->ms[t] ? ((oi_tmpl *)s->ms[t])->volume * ((oi_tmpl *)s->ms[t])->density
                                                     : s->density[t] * (rbound * rbound * rbound * 4. * M_PI / 3.);
                ninserted_this_local++;
            }
        }
    }
    /* ParticleToInsert::insert (particleToInsert.cpp:97-150): each body's
     * spheres; a single sphere has mass_ins, a clump's sphere
     * 4.18879020479 r^3 rho */
    int natoms = 0;
    for (int k = 0; k < ninserted_this_local; k++) {
        const int t = pti_t[k];
        const oi_tmpl *ms = (const oi_tmpl *)s->ms[t];
        const int ns = body_spheres(f, t, bx + 3 * k, bq + 4 * k, xs, rs);
        for (int j = 0; j < ns; j++, natoms++) {
            if (natoms >= cap) continue;
            memcpy(x_out + 3 * natoms, xs + 3 * j, 3 * sizeof(double));
            memcpy(v_out + 3 * natoms, bv + 3 * k, 3 * sizeof(double));
            /* insert/stream: set_x_v_omega gets no spin (fix_insert_stream.cpp:677) */
            omega_out[3 * natoms] = s->stream ? 0. : s->omega[0];
            omega_out[3 * natoms + 1] = s->stream ? 0. : s->omega[1];
            omega_out[3 * natoms + 2] = s->stream ? 0. : s->omega[2];
            r_out[natoms] = rs[j];
            if (ms) {
                m_out[natoms] = 4.18879020479 * rs[j] * rs[j] * rs[j] * ms->density;
                type_out[natoms] = ms->types ? ms->types[j] : ms->atom_type;
            } else {
                m_out[natoms] = s->density[t] * (rs[j] * rs[j] * rs[j] * 4. * M_PI / 3.);
                type_out[natoms] = s->type[t];
            }
            if (brid_out) brid_out[natoms] = pti_b[k];
        }
    }
    if (s->stream) {
        /* FixInsertStream::finalize_insertion (fix_insert_stream.cpp:763-826):
         * release data per atom, its velocity drawn in atom order */
        f->rel = (double *)realloc(f->rel, sizeof(double) * 14 * (natoms > 0 ? natoms : 1));
        f->nrel = natoms;
        const double vnm = o_mag(f->vn);
        const int istep = (int)step;
        for (int i = 0; i < natoms && i < cap; i++) {
            double pr[3];
            for (int q = 0; q < 3; q++) pr[q] = f->pref[q] - x_out[3 * i + q];
            const double dist = o_dot(pr, f->nrm);
            const int n_steps = (int)((dist + 1e-14) / (vnm * s->dt));
            double vi[3] = {s->vel[0], s->vel[1], s->vel[2]};
            random_velocity(f, vi);
            double *o = f->rel + 14 * i;
            memcpy(o, x_out + 3 * i, 3 * sizeof(double));
            o[3] = (double)istep;
            o[4] = (double)(istep + n_steps);
            memcpy(o + 5, f->vn, 3 * sizeof(double));
            memcpy(o + 8, vi, 3 * sizeof(double));
            memcpy(o + 11, s->omega, 3 * sizeof(double));
        }
    }
    f->ninserted += ninserted_this_local;
    f->massinserted += mass_inserted_this_local;
    f->last_inserted = ninserted_this_local;
    f->last_tries = ntry;
    const int more_after = !s->stream || !f->ninsert_exists || f->ninserted < f->ninsert;
    if (s->insert_every && more_after) f->next_reneighbor += s->insert_every;
    else f->next_reneighbor = 0;
    free(pti_t);
    free(pti_b);
    free(bx);
    free(bv);
    free(bq);
    free(xn);
    free(rn);
    free(xs);
    free(rs);
    return natoms;
}
