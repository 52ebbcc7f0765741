/*
 * dem_oracle.c -- CPU restatement of the LIGGGHTS-INL sphere DEM timestep.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * dem_oracle.h.  Each function cites the reference source it restates;
 * arithmetic is written in the same operation order so that per-contact
 * results round identically.
 *
 * Deliberate differences from the reference, none of which changes a
 * per-contact value:
 *   - atoms are never re-sorted (atom.cpp:1326); only summation order moves;
 *   - the neighbour list visits 27 cells of edge >= cutneighmax instead of
 *     the reference's half-size bins and stencil; membership is decided by
 *     the same predicate, only entry order differs;
 *   - periodic images are not materialised as ghost atoms; a list entry
 *     carries the image offset and the image position is x_j + s*prd, the
 *     same expression comm->borders() uses (atom_vec_sphere.cpp:625-637).
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXDNUM 8
#define T1 (OR_MAXTYPES + 1)

enum { OP_WALL = 1, OP_GRAVITY = 2, OP_FREEZE = 3, OP_SETFORCE = 4, OP_AVEFORCE = 5 };

/* contact_model_constants.h:64-67 */
enum { CONTACT_NORMAL_MODEL = 1, CONTACT_COHESION_MODEL = 2, CONTACT_TANGENTIAL_MODEL = 4,
       CONTACT_ROLLING_MODEL = 8 };

typedef struct {
    int j;       /* owned atom index */
    int shift;   /* (sx+1)+3(sy+1)+9(sz+1), 13 = no image */
    int flag;    /* contact flags (CONTACT_* bits) */
} entry_t;

/* history layout of one model chain; slots are registered in constructor
 * order surface, normal, cohesion, tangential, rolling (contact_models.h:
 * 150-157) */
typedef struct {
    int normal;      /* OR_NORMAL_* */
    int limitf;
    int off_norm;    /* hysteretic: 12 slots */
    int off_shear;   /* 3 slots */
    int off_roll;    /* 3 slots or -1 */
    int dnum;
    int newton[32];  /* newtonflag per slot (fix_contact_history.cpp:404-408) */
} model_t;

typedef struct {
    int id, kind, arg, bit;
    int style[3];
    double val[3];
    double c0, t0, c1;
    double foriginal[3];
} op_t;

struct or_sim {
    or_params p;
    int ntypes;
    double Yeff[T1 * T1], Geff[T1 * T1], betaeff[T1 * T1];
    double mu[T1 * T1], mur[T1 * T1], etar[T1 * T1];
    double restlog[T1 * T1];
    double hyst[OR_NHYST][T1 * T1];
    double krs;
    model_t mp, mw;   /* pair and wall model chains */
    int *gmask;       /* group bits per atom */
    double atime;
    long long atimestep;

    int n;
    double *x, *v, *omega, *f, *torque, *radius, *rmass, *xhold;
    int *type, *tag, *freeze, *image;
    double *fscale, *tscale;

    int dnum;          /* pair history doubles: 3 shear (+3 rolling) */
    int dnum_w;        /* wall history doubles */

    /* half neighbour list (granular_bin_no_newton) */
    int *nstart, *ncount;
    entry_t *ent;
    double *hist;
    int nent, cap;
    int have_list;

    /* per-atom partner storage filled by pre_exchange */
    int *npartner, *pstart, *ptag;
    double *phist;

    /* primitive planes */
    int nplanes;
    int pdim[OR_MAXPLANES], ptype[OR_MAXPLANES];
    double ppos[OR_MAXPLANES];
    unsigned char *pin;     /* n * nplanes: atom in the wall's neighbour list */
    double *phist_w;        /* n * nplanes * dnum_w */

    int nops;
    op_t ops[32];
    int next_op_id;
    int has_freeze;
    double gacc[3];

    long long ntimestep;
    int ago;
    int setupflag;
    int wall_rebuild;
    double cutneighmax;
};

static const double SQRT_FIVE_OVER_SIX = 0.91287092917527685576161630466800355658790782499663875;

/* ------------------------------------------------------------------------ */

/* history slots in registration order: hysteretic normal (12: deltaMax,
 * deltaZero, k1, deltaZero_old, k1_old, delta_old, deltaMin, betan, f0,
 * f0_old with newtonflag 0; kc, fo with newtonflag 1 --
 * normal_model_hysteretic_nonlinear1.h:84-97), shear x/y/z
 * (tangential_model_history.h:79-81), rolling torque x/y/z
 * (rolling_model_epsd.h:71-73), all with newtonflag 1. */
static void model_layout(model_t *m, int normal, int limitf, int rolling)
{
    int k = 0;
    memset(m, 0, sizeof(*m));
    m->normal = normal;
    m->limitf = limitf;
    m->off_norm = -1;
    if (normal != OR_NORMAL_HERTZ) {
        m->off_norm = k;
        for (int q = 0; q < 10; q++) m->newton[k++] = 0;
        m->newton[k++] = 1;
        m->newton[k++] = 1;
    }
    m->off_shear = k;
    for (int q = 0; q < 3; q++) m->newton[k++] = 1;
    m->off_roll = -1;
    if (rolling != OR_ROLL_OFF) {
        m->off_roll = k;
        for (int q = 0; q < 3; q++) m->newton[k++] = 1;
    }
    m->dnum = k;
}

or_sim *or_create(const or_params *p)
{
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->p = *p;
    model_layout(&s->mp, p->normal_model, p->limit_force, p->rolling_model);
    model_layout(&s->mw, p->wall_normal_model, p->wall_limit_force, p->rolling_model);
    s->dnum = s->mp.dnum;
    s->dnum_w = s->mw.dnum;
    return s;
}

static void free_atoms(or_sim *s)
{
    free(s->x); free(s->v); free(s->omega); free(s->f); free(s->torque);
    free(s->radius); free(s->rmass); free(s->xhold); free(s->type); free(s->tag);
    free(s->freeze); free(s->image); free(s->fscale); free(s->tscale);
    free(s->nstart); free(s->ncount); free(s->npartner); free(s->pstart);
    free(s->ptag); free(s->phist); free(s->pin); free(s->phist_w); free(s->gmask);
}

void or_destroy(or_sim *s)
{
    if (!s) return;
    free_atoms(s);
    free(s->ent); free(s->hist);
    free(s);
}

/* global_properties.cpp:428-562 (createYeff, createGeff, createCoeffRestLog,
 * createBetaEff) and :565-584 (per-type-pair pass-through). */
int or_set_materials(or_sim *s, int ntypes, const double *Y, const double *nu,
                     const double *e, const double *mu, const double *mur,
                     const double *etar, double krs)
{
    if (ntypes < 1 || ntypes > OR_MAXTYPES) return -1;
    s->ntypes = ntypes;
    const int w = ntypes + 1;
    for (int i = 1; i <= ntypes; i++)
        for (int j = 1; j <= ntypes; j++) {
            const double Yi = Y[i], Yj = Y[j], vi = nu[i], vj = nu[j];
            s->Yeff[i * T1 + j] = 1. / ((1. - pow(vi, 2.)) / Yi + (1. - pow(vj, 2.)) / Yj);
            s->Geff[i * T1 + j] = 1. / (2. * (2. - vi) * (1. + vi) / Yi + 2. * (2. - vj) * (1. + vj) / Yj);
            const double lg = log(e[i * w + j]);
            s->betaeff[i * T1 + j] = lg / sqrt(pow(lg, 2.) + pow(M_PI, 2.));
            s->restlog[i * T1 + j] = lg;
            s->mu[i * T1 + j] = mu[i * w + j];
            s->mur[i * T1 + j] = mur ? mur[i * w + j] : 0.0;
            s->etar[i * T1 + j] = etar ? etar[i * w + j] : 0.0;
        }
    s->krs = krs;
    return 0;
}

int or_set_hysteretic(or_sim *s, const double *tables)
{
    const int w = s->ntypes + 1;
    if (s->ntypes < 1) return -1;
    for (int t = 0; t < OR_NHYST; t++)
        for (int i = 1; i <= s->ntypes; i++)
            for (int j = 1; j <= s->ntypes; j++)
                s->hyst[t][i * T1 + j] = tables[(size_t)t * w * w + i * w + j];
    return 0;
}

void or_get_material_tables(const or_sim *s, double *Yeff, double *Geff, double *betaeff)
{
    const int w = s->ntypes + 1;
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) {
            Yeff[i * w + j] = s->Yeff[i * T1 + j];
            Geff[i * w + j] = s->Geff[i * T1 + j];
            betaeff[i * w + j] = s->betaeff[i * T1 + j];
        }
}

int or_set_atoms(or_sim *s, int n, const double *x, const double *v, const double *omega,
                 const double *radius, const double *rmass, const int *type,
                 const int *tag, const int *freeze)
{
    free_atoms(s);
    s->n = n;
    size_t n3 = (size_t)n * 3 * sizeof(double), n1 = (size_t)n * sizeof(double);
    size_t ni = (size_t)n * sizeof(int);
    s->x = malloc(n3); s->v = malloc(n3); s->omega = malloc(n3);
    s->f = calloc(1, n3); s->torque = calloc(1, n3); s->xhold = malloc(n3);
    s->radius = malloc(n1); s->rmass = malloc(n1);
    s->fscale = calloc(1, n1); s->tscale = calloc(1, n1);
    s->type = malloc(ni); s->tag = malloc(ni); s->freeze = calloc(1, ni);
    s->image = calloc(3, ni);
    s->gmask = calloc(1, ni);
    s->nstart = calloc(1, ni); s->ncount = calloc(1, ni);
    s->npartner = calloc(1, ni); s->pstart = calloc(1, ni);
    s->ptag = NULL; s->phist = NULL;
    memcpy(s->x, x, n3); memcpy(s->v, v, n3); memcpy(s->omega, omega, n3);
    memcpy(s->radius, radius, n1); memcpy(s->rmass, rmass, n1);
    memcpy(s->type, type, ni); memcpy(s->tag, tag, ni);
    if (freeze) memcpy(s->freeze, freeze, ni);
    s->pin = calloc((size_t)n * OR_MAXPLANES, 1);
    s->phist_w = calloc((size_t)n * OR_MAXPLANES * s->dnum_w, sizeof(double));
    s->nent = 0;
    s->have_list = 0;
    return 0;
}

static op_t *new_op(or_sim *s, int kind)
{
    op_t *o = &s->ops[s->nops++];
    memset(o, 0, sizeof(*o));
    o->id = s->next_op_id++;
    o->kind = kind;
    return o;
}

int or_add_plane(or_sim *s, int dim, double pos, int wtype)
{
    if (s->nplanes >= OR_MAXPLANES || dim < 0 || dim > 2) return -1;
    const int k = s->nplanes++;
    s->pdim[k] = dim; s->ppos[k] = pos; s->ptype[k] = wtype;
    op_t *o = new_op(s, OP_WALL);
    o->arg = k;
    return o->id;
}

void or_add_gravity(or_sim *s, double ax, double ay, double az)
{
    s->gacc[0] = ax; s->gacc[1] = ay; s->gacc[2] = az;
    new_op(s, OP_GRAVITY);
}

void or_add_freeze(or_sim *s)
{
    new_op(s, OP_FREEZE);
    s->has_freeze = 1;
}

void or_set_group(or_sim *s, int bit, const int *members)
{
    for (int i = 0; i < s->n; i++) {
        if (members[i]) s->gmask[i] |= 1 << bit;
        else s->gmask[i] &= ~(1 << bit);
    }
}

int or_add_setforce(or_sim *s, int bit, const int style[3], const double value[3])
{
    op_t *o = new_op(s, OP_SETFORCE);
    o->bit = bit;
    for (int d = 0; d < 3; d++) { o->style[d] = style[d]; o->val[d] = value[d]; }
    return o->id;
}

int or_add_aveforce(or_sim *s, int bit, const int style[3], double xv, double yv,
                    double c0, double t0, double c1)
{
    op_t *o = new_op(s, OP_AVEFORCE);
    o->bit = bit;
    for (int d = 0; d < 3; d++) o->style[d] = style[d];
    o->val[0] = xv; o->val[1] = yv;
    o->c0 = c0; o->t0 = t0; o->c1 = c1;
    return o->id;
}

void or_unfix(or_sim *s, int id)
{
    int k = 0;
    for (int q = 0; q < s->nops; q++)
        if (s->ops[q].id != id) s->ops[k++] = s->ops[q];
    s->nops = k;
}

void or_aveforce_vector(const or_sim *s, int id, double out[3])
{
    for (int q = 0; q < s->nops; q++)
        if (s->ops[q].id == id)
            for (int d = 0; d < 3; d++) out[d] = s->ops[q].foriginal[d];
}

/* update.cpp: atime + (ntimestep - atimestep)*dt */
double or_time(const or_sim *s)
{
    return s->atime + (s->ntimestep - s->atimestep) * s->p.dt;
}

/* ------------------------------------------------------------------------ */
/* Geometry helpers                                                          */

static inline void shift_decode(int code, int *sx, int *sy, int *sz)
{
    *sx = code % 3 - 1;
    *sy = (code / 3) % 3 - 1;
    *sz = code / 9 - 1;
}

/* position of atom j's image: x[j] + dx with dx = pbc*prd
 * (atom_vec_sphere.cpp:625-637) */
static inline void image_pos(const or_sim *s, int j, int code, double xj[3])
{
    int sh[3];
    shift_decode(code, &sh[0], &sh[1], &sh[2]);
    for (int d = 0; d < 3; d++) {
        const double prd = s->p.boxhi[d] - s->p.boxlo[d];
        const double dx = sh[d] * prd;
        xj[d] = s->x[3 * j + d] + dx;
    }
}

/* Domain::pbc (domain.cpp:550-620), orthogonal box */
static void domain_pbc(or_sim *s)
{
    for (int i = 0; i < s->n; i++)
        for (int d = 0; d < 3; d++) {
            if (!s->p.periodic[d]) continue;
            const double lo = s->p.boxlo[d], hi = s->p.boxhi[d];
            const double prd = hi - lo;
            double *xi = &s->x[3 * i + d];
            if (*xi < lo) {
                *xi += prd;
                s->image[3 * i + d]--;
            }
            if (*xi >= hi) {
                *xi -= prd;
                *xi = (*xi > lo) ? *xi : lo;
                s->image[3 * i + d]++;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Contact history: FixContactHistory::pre_exchange                          */
/* (fix_contact_history.cpp:305-426).  The partner's copy negates the slots */
/* registered with newtonflag 1 (model_layout).                              */

static void pre_exchange(or_sim *s)
{
    const int n = s->n, dnum = s->dnum;
    for (int i = 0; i < n; i++) s->npartner[i] = 0;
    if (!s->have_list) return;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < s->ncount[i]; k++) {
            const entry_t *e = &s->ent[s->nstart[i] + k];
            if (e->flag) {
                s->npartner[i]++;
                if (e->shift == 13) s->npartner[e->j]++;
            }
        }
    int tot = 0;
    for (int i = 0; i < n; i++) { s->pstart[i] = tot; tot += s->npartner[i]; }
    free(s->ptag); free(s->phist);
    s->ptag = malloc((size_t)(tot + 1) * sizeof(int));
    s->phist = malloc((size_t)(tot + 1) * dnum * sizeof(double));
    for (int i = 0; i < n; i++) s->npartner[i] = 0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < s->ncount[i]; k++) {
            const int idx = s->nstart[i] + k;
            const entry_t *e = &s->ent[idx];
            if (!e->flag) continue;
            const double *h = &s->hist[(size_t)idx * dnum];
            int m = s->pstart[i] + s->npartner[i];
            s->ptag[m] = s->tag[e->j];
            for (int d = 0; d < dnum; d++) s->phist[(size_t)m * dnum + d] = h[d];
            s->npartner[i]++;
            if (e->shift == 13) {
                const int j = e->j;
                m = s->pstart[j] + s->npartner[j];
                s->ptag[m] = s->tag[i];
                for (int d = 0; d < dnum; d++)
                    s->phist[(size_t)m * dnum + d] = s->mp.newton[d] ? -h[d] : h[d];
                s->npartner[j]++;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Neighbour list: Neighbor::granular_bin_no_newton (neigh_gran.cpp:485-640) */

static void push_entry(or_sim *s, int j, int code, int flag, const double *h)
{
    if (s->nent == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->ent = realloc(s->ent, (size_t)s->cap * sizeof(entry_t));
        s->hist = realloc(s->hist, (size_t)s->cap * s->dnum * sizeof(double));
    }
    entry_t *e = &s->ent[s->nent];
    e->j = j; e->shift = code; e->flag = flag;
    double *dst = &s->hist[(size_t)s->nent * s->dnum];
    for (int d = 0; d < s->dnum; d++) dst[d] = h ? h[d] : 0.0;
    s->nent++;
}

/* candidate (i, j-image) test and history re-seed, neigh_gran.cpp:570-612 */
static void try_pair(or_sim *s, int i, int j, int code)
{
    double xj[3];
    image_pos(s, j, code, xj);
    const double delx = s->x[3 * i] - xj[0];
    const double dely = s->x[3 * i + 1] - xj[1];
    const double delz = s->x[3 * i + 2] - xj[2];
    const double rsq = delx * delx + dely * dely + delz * delz;
    const double radsum = (s->radius[i] + s->radius[j]) * s->p.cdf;
    const double cutsq = (radsum + s->p.skin) * (radsum + s->p.skin);
    if (!(rsq <= cutsq)) return;
    if (rsq < radsum * radsum) {
        int m;
        const int np = s->npartner[i], ps = s->pstart[i];
        for (m = 0; m < np; m++)
            if (s->ptag[ps + m] == s->tag[j]) break;
        if (m < np) {
            push_entry(s, j, code, 1, &s->phist[(size_t)(ps + m) * s->dnum]);
            return;
        }
    }
    push_entry(s, j, code, 0, NULL);
}

static int cmp_pair(const void *a, const void *b)
{
    const int *x = (const int *)a, *y = (const int *)b;
    if (x[0] != y[0]) return (x[0] > y[0]) - (x[0] < y[0]);
    return (x[1] > y[1]) - (x[1] < y[1]);
}

static void build_list(or_sim *s)
{
    const int n = s->n;
    double maxrad = 0.0;
    for (int i = 0; i < n; i++) if (s->radius[i] > maxrad) maxrad = s->radius[i];
    /* PairGran::init_one (pair_gran.cpp:591-603) + skin */
    s->cutneighmax = (maxrad + maxrad) * s->p.cdf + s->p.skin;
    s->nent = 0;

    /* cell grid: periodic dims span the box, others the atoms' extent */
    double lo[3], hi[3];
    int nc[3];
    for (int d = 0; d < 3; d++) {
        if (s->p.periodic[d]) { lo[d] = s->p.boxlo[d]; hi[d] = s->p.boxhi[d]; }
        else {
            lo[d] = 1e300; hi[d] = -1e300;
            for (int i = 0; i < n; i++) {
                const double xi = s->x[3 * i + d];
                if (xi < lo[d]) lo[d] = xi;
                if (xi > hi[d]) hi[d] = xi;
            }
            if (n == 0) { lo[d] = 0; hi[d] = 1; }
        }
        nc[d] = (int)((hi[d] - lo[d]) / s->cutneighmax);
        if (nc[d] < 1) nc[d] = 1;
        if (!s->p.periodic[d] && nc[d] > 4096) nc[d] = 4096;
    }
    int brute = 0;
    for (int d = 0; d < 3; d++)
        if (s->p.periodic[d] && nc[d] < 3) brute = 1;

    if (brute || n < 64) {
        /* every image offset the periodic dims allow */
        for (int i = 0; i < n; i++) {
            s->nstart[i] = s->nent;
            for (int j = 0; j < n; j++)
                for (int code = 0; code < 27; code++) {
                    int sh[3];
                    shift_decode(code, &sh[0], &sh[1], &sh[2]);
                    int ok = 1;
                    for (int d = 0; d < 3; d++)
                        if (sh[d] != 0 && !s->p.periodic[d]) ok = 0;
                    if (!ok) continue;
                    if (code == 13 && j <= i) continue;
                    try_pair(s, i, j, code);
                }
            s->ncount[i] = s->nent - s->nstart[i];
        }
        s->have_list = 1;
        return;
    }

    const long long ncell = (long long)nc[0] * nc[1] * nc[2];
    int *cstart = calloc((size_t)ncell + 1, sizeof(int));
    int *cell = malloc((size_t)n * sizeof(int));
    int *sorted = malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) {
        int c[3];
        for (int d = 0; d < 3; d++) {
            int ci = (int)floor((s->x[3 * i + d] - lo[d]) / (hi[d] - lo[d]) * nc[d]);
            if (ci < 0) ci = 0;
            if (ci >= nc[d]) ci = nc[d] - 1;
            c[d] = ci;
        }
        cell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        cstart[cell[i] + 1]++;
    }
    for (long long c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
    {
        int *fill = malloc((size_t)ncell * sizeof(int));
        for (long long c = 0; c < ncell; c++) fill[c] = cstart[c];
        for (int i = 0; i < n; i++) sorted[fill[cell[i]]++] = i; /* ascending index per cell */
        free(fill);
    }
    int *cand = malloc(sizeof(int) * 2 * 4096);
    int ccap = 4096;
    for (int i = 0; i < n; i++) {
        s->nstart[i] = s->nent;
        const int ci = cell[i];
        const int cx = ci % nc[0], cy = (ci / nc[0]) % nc[1], cz = ci / (nc[0] * nc[1]);
        int ncand = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int c[3] = {cx + dx, cy + dy, cz + dz}, sh[3] = {0, 0, 0}, ok = 1;
                    for (int d = 0; d < 3; d++) {
                        if (c[d] < 0) {
                            if (s->p.periodic[d]) { c[d] += nc[d]; sh[d] = -1; } else ok = 0;
                        } else if (c[d] >= nc[d]) {
                            if (s->p.periodic[d]) { c[d] -= nc[d]; sh[d] = 1; } else ok = 0;
                        }
                    }
                    if (!ok) continue;
                    const int code = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
                    const int cc = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
                    for (int q = cstart[cc]; q < cstart[cc + 1]; q++) {
                        const int j = sorted[q];
                        if (code == 13 && j <= i) continue;
                        if (ncand == ccap) { ccap *= 2; cand = realloc(cand, sizeof(int) * 2 * ccap); }
                        cand[2 * ncand] = j; cand[2 * ncand + 1] = code;
                        ncand++;
                    }
                }
        /* canonical order: by (j, shift) */
        qsort(cand, ncand, 2 * sizeof(int), cmp_pair);
        for (int q = 0; q < ncand; q++) try_pair(s, i, cand[2 * q], cand[2 * q + 1]);
        s->ncount[i] = s->nent - s->nstart[i];
    }
    free(cand); free(cstart); free(cell); free(sorted);
    s->have_list = 1;
}

/* PrimitiveWall::buildNeighList with treshold = skin
 * (fix_wall_gran.cpp:688-706, primitive_wall.h:139-148,
 *  primitive_wall_definitions.h:142-148) */
static void build_wall_lists(or_sim *s)
{
    for (int k = 0; k < s->nplanes; k++) {
        const int d = s->pdim[k];
        for (int i = 0; i < s->n; i++) {
            const double dMax = s->radius[i] + s->p.skin;
            const double dist = s->x[3 * i + d] - s->ppos[k];
            const double absdist = (dist > 0.0) ? dist : -dist;
            s->pin[(size_t)i * OR_MAXPLANES + k] = (absdist <= dMax);
        }
    }
}

void or_rebuild(or_sim *s)
{
    pre_exchange(s);
    domain_pbc(s);
    build_list(s);
    memcpy(s->xhold, s->x, (size_t)s->n * 3 * sizeof(double));
    s->ago = 0;
    s->wall_rebuild = 1;
}

/* ------------------------------------------------------------------------ */
/* Contact model chain: ContactModel::surfacesIntersect (contact_models.h:   */
/* 228-238) = surface default -> normal hertz -> (cohesion off) ->           */
/* tangential history -> rolling.                                            */

typedef struct {
    int is_wall;
    int itype, jtype;
    double radi, radj, radsum, r, rinv, rsq;
    double delta[3], en[3];
    double meff, mi, mj;
    const double *v_i, *v_j, *omega_i, *omega_j;
    double deltan;      /* wall: set by caller before the chain */
    double vn, wr[3], vtr[3], cri, crj;
    double Fn, kn, kt, gamman, gammat;
    int update_history;
    double *hist;       /* dnum doubles */
    const model_t *model;
    int *flags;         /* contact flags, NULL for walls */
} sidata_t;

/* surface_model_default.h:146-208 */
static void surface_default(sidata_t *sd)
{
    const double enx = sd->en[0], eny = sd->en[1], enz = sd->en[2];
    const double vr1 = sd->v_i[0] - sd->v_j[0];
    const double vr2 = sd->v_i[1] - sd->v_j[1];
    const double vr3 = sd->v_i[2] - sd->v_j[2];
    const double vn = vr1 * enx + vr2 * eny + vr3 * enz;
    const double vn1 = vn * enx, vn2 = vn * eny, vn3 = vn * enz;
    const double vt1 = vr1 - vn1, vt2 = vr2 - vn2, vt3 = vr3 - vn3;
    const double deltan = sd->radsum - sd->r;
    const double dx = sd->delta[0], dy = sd->delta[1], dz = sd->delta[2];
    const double rinv = sd->rinv;
    double wr1, wr2, wr3;
    if (sd->is_wall) {
        const double cr = sd->radi - 0.5 * sd->deltan;   /* caller's deltan */
        wr1 = cr * sd->omega_i[0] * rinv;
        wr2 = cr * sd->omega_i[1] * rinv;
        wr3 = cr * sd->omega_i[2] * rinv;
        sd->cri = cr;
    } else {
        const double cri = sd->radi - 0.5 * deltan;
        const double crj = sd->radj - 0.5 * deltan;
        wr1 = (cri * sd->omega_i[0] + crj * sd->omega_j[0]) * rinv;
        wr2 = (cri * sd->omega_i[1] + crj * sd->omega_j[1]) * rinv;
        wr3 = (cri * sd->omega_i[2] + crj * sd->omega_j[2]) * rinv;
        sd->cri = cri;
        sd->crj = crj;
    }
    sd->vtr[0] = vt1 - (dz * wr2 - dy * wr3);
    sd->vtr[1] = vt2 - (dx * wr3 - dz * wr1);
    sd->vtr[2] = vt3 - (dy * wr1 - dx * wr2);
    sd->vn = vn;
    sd->deltan = deltan;
    sd->wr[0] = wr1; sd->wr[1] = wr2; sd->wr[2] = wr3;
}

/* normal_model_hertz.h:205-405 (no heating, energy or bond options) */
static void normal_hertz(const or_sim *s, sidata_t *sd, double Fi[3], double Fj[3])
{
    const int t = sd->itype * T1 + sd->jtype;
    const double reff = sd->is_wall ? sd->radi : (sd->radi * sd->radj / (sd->radi + sd->radj));
    const double meff = sd->meff;
    const double sqrtval = sqrt(reff * sd->deltan);
    const double deltan = sd->deltan;
    const double Sn = 2. * s->Yeff[t] * sqrtval;
    const double St = 8. * s->Geff[t] * sqrtval;
    double kn = 4. / 3. * s->Yeff[t] * sqrtval;
    double kt = St;
    const double gamman = -2. * SQRT_FIVE_OVER_SIX * s->betaeff[t] * sqrt(Sn * meff);
    const double gammat = s->p.tangential_damping ? -2. * SQRT_FIVE_OVER_SIX * s->betaeff[t] * sqrt(St * meff) : 0.0;
    /* kn /= nktv2p, kt /= nktv2p with nktv2p = 1 for units si */
    const double Fn_damping = -gamman * sd->vn;
    const double Fn_contact = kn * deltan;
    double Fn = Fn_damping + Fn_contact;
    if (sd->model->limitf && (Fn < 0.0)) Fn = 0.0;
    sd->Fn = Fn; sd->kn = kn; sd->kt = kt; sd->gamman = gamman; sd->gammat = gammat;
    if (sd->is_wall) {
        const double Fn_ = Fn * 1.0;  /* area_ratio = 1 */
        Fi[0] += Fn_ * sd->en[0];
        Fi[1] += Fn_ * sd->en[1];
        Fi[2] += Fn_ * sd->en[2];
    } else {
        Fi[0] += Fn * sd->en[0];
        Fi[1] += Fn * sd->en[1];
        Fi[2] += Fn * sd->en[2];
        Fj[0] += -Fi[0];
        Fj[1] += -Fi[1];
        Fj[2] += -Fi[2];
    }
}

/* tangential_model_history.h:135-425 (torsion off, no energy terms) */
static void tangential_history(const or_sim *s, sidata_t *sd, double Fi[3], double Ti[3],
                               double Fj[3], double Tj[3])
{
    const double enx = sd->en[0], eny = sd->en[1], enz = sd->en[2];
    double *shear = sd->hist + sd->model->off_shear;
    if (sd->update_history) {
        const double dt = s->p.dt;
        shear[0] += sd->vtr[0] * dt;
        shear[1] += sd->vtr[1] * dt;
        shear[2] += sd->vtr[2] * dt;
        const double rsht = shear[0] * enx + shear[1] * eny + shear[2] * enz;
        shear[0] -= rsht * enx;
        shear[1] -= rsht * eny;
        shear[2] -= rsht * enz;
    }
    const double shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
    const double kt = sd->kt;
    const double xmu = s->mu[sd->itype * T1 + sd->jtype];
    double Ft1 = -(kt * shear[0]);
    double Ft2 = -(kt * shear[1]);
    double Ft3 = -(kt * shear[2]);
    const double Ft_shear = kt * shrmag;
    const double Ft_friction = xmu * fabs(sd->Fn);
    if (Ft_shear > Ft_friction) {
        if (shrmag != 0.0) {
            const double ratio = Ft_friction / Ft_shear;
            Ft1 *= ratio; Ft2 *= ratio; Ft3 *= ratio;
            if (sd->update_history) {
                shear[0] = -Ft1 / kt;
                shear[1] = -Ft2 / kt;
                shear[2] = -Ft3 / kt;
            }
        } else
            Ft1 = Ft2 = Ft3 = 0.0;
    } else {
        const double gammat = sd->gammat;
        Ft1 -= (gammat * sd->vtr[0]);
        Ft2 -= (gammat * sd->vtr[1]);
        Ft3 -= (gammat * sd->vtr[2]);
    }
    const double tor1 = eny * Ft3 - enz * Ft2;
    const double tor2 = enz * Ft1 - enx * Ft3;
    const double tor3 = enx * Ft2 - eny * Ft1;
    /* Tn_shear = 0 without torsion: x + 0.0*en == x */
    const double ti0 = -sd->cri * tor1, ti1 = -sd->cri * tor2, ti2 = -sd->cri * tor3;
    if (sd->is_wall) {
        Fi[0] += Ft1 * 1.0; Fi[1] += Ft2 * 1.0; Fi[2] += Ft3 * 1.0;
        Ti[0] += ti0 * 1.0; Ti[1] += ti1 * 1.0; Ti[2] += ti2 * 1.0;
    } else {
        const double tj0 = -sd->crj * tor1, tj1 = -sd->crj * tor2, tj2 = -sd->crj * tor3;
        Fi[0] += Ft1; Fi[1] += Ft2; Fi[2] += Ft3;
        Fj[0] += -Ft1; Fj[1] += -Ft2; Fj[2] += -Ft3;
        Ti[0] += ti0; Ti[1] += ti1; Ti[2] += ti2;
        Tj[0] += tj0; Tj[1] += tj1; Tj[2] += tj2;
    }
}

/* rolling_model_epsd.h:230-330, rolling_model_epsd2.h:161-230,
 * rolling_model_epsd3.h:178-262 (calcRollTorque) */
static void roll_torque(const or_sim *s, const sidata_t *sd, double reff,
                        double wr1, double wr2, double wr3, double r_inertia, double rt[3])
{
    const int t = sd->itype * T1 + sd->jtype;
    const double enx = sd->en[0], eny = sd->en[1], enz = sd->en[2];
    const double dt = s->p.dt;
    double *c_history = sd->hist + sd->model->off_roll;
    const double rmu = s->mur[t];
    double wr_tot[3];
    if (s->p.torsion_torque) {
        wr_tot[0] = wr1; wr_tot[1] = wr2; wr_tot[2] = wr3;
    } else {
        const double wr_dot_delta = wr1 * enx + wr2 * eny + wr3 * enz;
        const double wn0 = enx * wr_dot_delta, wn1 = eny * wr_dot_delta, wn2 = enz * wr_dot_delta;
        wr_tot[0] = wr1 - wn0; wr_tot[1] = wr2 - wn1; wr_tot[2] = wr3 - wn2;
    }
    double kr;
    if (s->p.rolling_model == OR_ROLL_EPSD) kr = 2.25 * sd->kn * rmu * rmu * reff * reff;
    else if (s->p.rolling_model == OR_ROLL_EPSD2) kr = sd->kt * reff * reff;
    else kr = s->krs * sd->kn * rmu * rmu * reff * reff;
    const double dtkr = dt * kr;
    const double dr0 = dtkr * wr_tot[0], dr1 = dtkr * wr_tot[1], dr2 = dtkr * wr_tot[2];
    rt[0] = c_history[0] + dr0;
    rt[1] = c_history[1] + dr1;
    rt[2] = c_history[2] + dr2;
    const double mag = sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    const double rmax = fabs(sd->Fn) * reff * rmu;
    if (s->p.rolling_model == OR_ROLL_EPSD2) {
        if (mag > rmax) {
            const double factor = rmax / mag;
            rt[0] *= factor; rt[1] *= factor; rt[2] *= factor;
        }
        if (sd->update_history) {
            c_history[0] = rt[0]; c_history[1] = rt[1]; c_history[2] = rt[2];
        }
        return;
    }
    if (mag > rmax) {
        const double factor = rmax / mag;
        rt[0] *= factor; rt[1] *= factor; rt[2] *= factor;
        if (sd->update_history) {
            c_history[0] = rt[0]; c_history[1] = rt[1]; c_history[2] = rt[2];
        }
    } else {
        if (sd->update_history) {
            c_history[0] = rt[0]; c_history[1] = rt[1]; c_history[2] = rt[2];
        }
        const double r_coef = s->etar[t] * 2 * sqrt(r_inertia * kr);
        rt[0] += r_coef * wr_tot[0];
        rt[1] += r_coef * wr_tot[1];
        rt[2] += r_coef * wr_tot[2];
    }
}

/* RollingModel<EPSD*>::surfacesIntersect (rolling_model_epsd.h:96-220 etc.) */
static void rolling(const or_sim *s, sidata_t *sd, double Ti[3], double Tj[3])
{
    double rt[3] = {0, 0, 0};
    const double radi = sd->radi, radj = sd->radj;
    const double reff = sd->is_wall ? radi : (radi * radj / (radi + radj));
    if (sd->is_wall) {
        /* epsd: 1.4*mi*reff*reff; epsd3: 1.4*mi*radius*radius, radius = radi */
        const double r_inertia = 1.4 * sd->mi * reff * reff;
        roll_torque(s, sd, reff, sd->wr[0], sd->wr[1], sd->wr[2], r_inertia, rt);
    } else {
        const double w0 = sd->omega_i[0] - sd->omega_j[0];
        const double w1 = sd->omega_i[1] - sd->omega_j[1];
        const double w2 = sd->omega_i[2] - sd->omega_j[2];
        const double ri = sd->mi * radi * radi;
        const double rj = sd->mj * radj * radj;
        const double r_inertia = 1.4 * ri * rj / (ri + rj);
        roll_torque(s, sd, reff, w0, w1, w2, r_inertia, rt);
    }
    Ti[0] -= rt[0]; Ti[1] -= rt[1]; Ti[2] -= rt[2];
    Tj[0] += rt[0]; Tj[1] += rt[1]; Tj[2] += rt[2];
}

/* NormalModel<HYSTERETIC_NONLINEAR1/2>::surfacesIntersect
 * (normal_model_hysteretic_nonlinear1.h:149-420, ..._nonlinear2.h:149-420).
 * The history is written whatever shearupdate says, and sqrt(5/4) is the
 * integer expression of the reference, i.e. sqrt(1) = 1. */
static void normal_hysteretic(const or_sim *s, sidata_t *sd, double Fi[3], double Fj[3])
{
    const int t = sd->itype * T1 + sd->jtype;
    const int nl2 = sd->model->normal == OR_NORMAL_HYST2;
    const double deltan = sd->deltan;
    const double ri = sd->radi, rj = sd->radj;
    const double reff = sd->is_wall ? sd->radi : (ri * rj / (ri + rj));
    const double meff = sd->meff;
    double kn = s->hyst[0][t];
    const double Alpha_in = s->hyst[5][t];
    const double Cin_in = s->hyst[6][t];
    const double A1_in = s->hyst[7][t];
    const double A2_in = s->hyst[8][t];
    const double A3_in = s->hyst[9][t];
    const double kcin_in = s->hyst[10][t];
    const double k_c = kcin_in * A2_in;
    const double k2Max = kn * s->hyst[1][t];
    const double kc = s->hyst[2][t] * kn;
    const double f_0 = s->hyst[4][t];
    const double crl = s->restlog[t];
    const double sqrt54 = sqrt((double)(5 / 4));
    const double fcc = 1.0;
    double gamman, gammat;
    if (sd->flags) *sd->flags |= CONTACT_NORMAL_MODEL;
    double *const history = sd->hist + sd->model->off_norm;
    double *const kc_history = history + 10;
    double *const fo_history = history + 11;
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
    double k1_old = history[4];
    const double delta_old = history[5];
    double deltaMin = history[6];
    double betan = history[7];
    double f0 = history[8];
    double f0_old = history[9];
    (void)k2Max; (void)reff;  /* deltaMaxLim is computed but unused there */
    double k2 = A3_in * k1;
    double beta;
    int tag_status;
    if (deltan >= delta_old) {
        if (delta_old == 0) tag_status = 1;
        else tag_status = (sd->vn > 0) ? 2 : 1;
    } else {
        tag_status = 2;
    }
    double fHys;
    const double dampfac = 1. + (M_PI / crl) * (M_PI / crl);
    if (tag_status == 1) {
        if (deltaZero == 0) k1 = A2_in;
        else k1 = history[2];
        deltaMin = history[6];
        f0_old = f0;
        if (deltan <= deltaZero) {
            if (nl2)
                fHys = Cin_in * k2 * (exp(betan * (deltan - deltaZero)) - 1);
            else if (deltan <= deltaMin)
                fHys = -k_c * deltan;
            else
                fHys = betan * (deltan - deltaZero);
        } else {
            fHys = Alpha_in * k1 * pow(deltan - deltaZero, 2) + f0;
        }
        const double pw = nl2 ? 0.5 : 0.25;
        gamman = sqrt54 * sqrt(4. * meff * Alpha_in * k1 / dampfac) * (pow(deltan, pw) + pow(deltaZero, pw));
        gammat = sqrt54 * sqrt(4. * meff * Alpha_in * k1 / dampfac) * (pow(deltan, pw) + pow(deltaZero, pw));
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
        if (!nl2) {
            k2 = A3_in * k1_old;
            deltaZero = (1 - k1_old / k2) * deltaMax;
            beta = Alpha_in * k1_old * pow(deltaMax - deltaZero_old, 2) / k2 / (deltaMax - deltaZero);
            deltaMin = beta * (k2 - k1_old) / (beta * k2 + k_c) * deltaMax;
            k1 = deltaMax * A1_in + A2_in;
            if (deltan >= deltaMin) {
                if (deltan >= deltaZero) {
                    betan = beta * k2;
                    fHys = beta * k2 * (deltan - deltaZero) + (deltan - deltaZero) * f0_old / (deltaMax - deltaZero);
                    f0 = fHys - Alpha_in * k1 * pow(deltan - deltaZero, 2);
                } else {
                    betan = beta * k2;
                    fHys = beta * k2 * (deltan - deltaZero);
                    f0 = 0;
                }
            } else {
                fHys = -k_c * deltan;
                f0 = 0;
            }
            gamman = sqrt54 * sqrt(4. * meff * Alpha_in * k1_old / dampfac) * (pow(deltan, 0.25));
            gammat = sqrt54 * sqrt(4. * meff * Alpha_in * k1_old / dampfac) * (pow(deltan, 0.25));
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
            gammat = 1 * 0.001 * sqrt54 * sqrt(4. * meff * Alpha_in * k1_old / dampfac) * (pow(deltan, -0.25));
        }
        history[1] = deltaZero;
        history[2] = k1;
        history[3] = deltaZero_old;
        history[4] = k1_old;
        history[5] = deltan;
        history[6] = deltaMin;
        history[7] = betan;
        history[8] = f0;
    }
    kn = k1;
    const double kt = kn;
    const double Fn_damping = -gamman * sd->vn;
    double Fn = fHys + Fn_damping + f_0;
    if (sd->model->limitf && (Fn < 0.0) && kc == 0 && f_0 == 0.0) Fn = 0.0;
    sd->Fn = Fn;
    sd->kn = kn;
    sd->kt = kt;
    kc_history[0] = kc;
    fo_history[0] = f_0;
    sd->gamman = gamman;
    sd->gammat = gammat;
    if (sd->is_wall) {
        const double Fn_ = Fn * 1.0;
        Fi[0] = Fn_ * sd->en[0];
        Fi[1] = Fn_ * sd->en[1];
        Fi[2] = Fn_ * sd->en[2];
    } else {
        Fi[0] = sd->Fn * sd->en[0];
        Fi[1] = sd->Fn * sd->en[1];
        Fi[2] = sd->Fn * sd->en[2];
        Fj[0] = -Fi[0];
        Fj[1] = -Fi[1];
        Fj[2] = -Fi[2];
    }
}

static void chain_intersect(const or_sim *s, sidata_t *sd, double Fi[3], double Ti[3],
                            double Fj[3], double Tj[3])
{
    surface_default(sd);
    if (sd->model->normal == OR_NORMAL_HERTZ) {
        if (sd->flags) *sd->flags |= CONTACT_NORMAL_MODEL;
        normal_hertz(s, sd, Fi, Fj);
    } else {
        normal_hysteretic(s, sd, Fi, Fj);
    }
    if (sd->flags) *sd->flags |= CONTACT_TANGENTIAL_MODEL;
    tangential_history(s, sd, Fi, Ti, Fj, Tj);
    if (sd->model->off_roll >= 0) {
        if (sd->flags) *sd->flags |= CONTACT_ROLLING_MODEL;
        rolling(s, sd, Ti, Tj);
    }
}

/* surfacesClose chain: hysteretic zeroes its ten state slots and clears its
 * flag (normal_model_hysteretic_nonlinear1.h:444-458; hertz keeps its flag),
 * tangential zeroes shear (tangential_model_history.h:428-440), rolling
 * zeroes its torque history (rolling_model_epsd.h:222-229). */
static void chain_close(const model_t *m, double *hist, int *flag)
{
    if (m->normal != OR_NORMAL_HERTZ) {
        if (flag) *flag &= ~CONTACT_NORMAL_MODEL;
        for (int q = 0; q < 10; q++) hist[m->off_norm + q] = 0.0;
    }
    if (flag) *flag &= ~CONTACT_TANGENTIAL_MODEL;
    for (int q = 0; q < 3; q++) hist[m->off_shear + q] = 0.0;
    if (m->off_roll >= 0) {
        if (flag) *flag &= ~CONTACT_ROLLING_MODEL;
        for (int q = 0; q < 3; q++) hist[m->off_roll + q] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Granular<>::compute_force (pair_gran_base.h:187-503)                      */

static void pair_compute(or_sim *s, int shearupdate)
{
    const double cdm = s->p.cdf * s->p.cdf;
    for (int i = 0; i < s->n; i++) {
        const double xtmp = s->x[3 * i], ytmp = s->x[3 * i + 1], ztmp = s->x[3 * i + 2];
        const double radi = s->radius[i];
        for (int k = 0; k < s->ncount[i]; k++) {
            const int idx = s->nstart[i] + k;
            entry_t *e = &s->ent[idx];
            const int j = e->j;
            double xj[3];
            image_pos(s, j, e->shift, xj);
            const double delx = xtmp - xj[0];
            const double dely = ytmp - xj[1];
            const double delz = ztmp - xj[2];
            const double rsq = delx * delx + dely * dely + delz * delz;
            const double radj = s->radius[j];
            const double radsum = radi + radj;
            double *h = &s->hist[(size_t)idx * s->dnum];
            double Fi[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, Fj[3] = {0, 0, 0}, Tj[3] = {0, 0, 0};
            if (rsq < radsum * radsum) {
                const double r = sqrt(rsq);
                const double rinv = 1.0 / r;
                sidata_t sd;
                memset(&sd, 0, sizeof(sd));
                sd.is_wall = 0;
                sd.itype = s->type[i]; sd.jtype = s->type[j];
                sd.radi = radi; sd.radj = radj; sd.radsum = radsum;
                sd.r = r; sd.rinv = rinv; sd.rsq = rsq;
                sd.delta[0] = delx; sd.delta[1] = dely; sd.delta[2] = delz;
                sd.en[0] = delx * rinv; sd.en[1] = dely * rinv; sd.en[2] = delz * rinv;
                const double mi = s->rmass[i], mj = s->rmass[j];
                double meff = mi * mj / (mi + mj);
                /* freeze_group_bit is nonzero only with a fix freeze
                 * (pair_gran.cpp:532-537) */
                if (s->has_freeze && s->freeze[i]) meff = mj;
                if (s->has_freeze && s->freeze[j]) meff = mi;
                sd.meff = meff; sd.mi = mi; sd.mj = mj;
                sd.v_i = &s->v[3 * i]; sd.v_j = &s->v[3 * j];
                sd.omega_i = &s->omega[3 * i]; sd.omega_j = &s->omega[3 * j];
                sd.update_history = shearupdate;
                sd.hist = h;
                sd.model = &s->mp;
                sd.flags = &e->flag;
                chain_intersect(s, &sd, Fi, Ti, Fj, Tj);
                for (int d = 0; d < 3; d++) {
                    s->f[3 * i + d] += Fi[d];
                    s->torque[3 * i + d] += Ti[d];
                }
                s->fscale[i] += sqrt(Fi[0] * Fi[0] + Fi[1] * Fi[1] + Fi[2] * Fi[2]);
                s->tscale[i] += sqrt(Ti[0] * Ti[0] + Ti[1] * Ti[1] + Ti[2] * Ti[2]);
                if (e->shift == 13) {
                    for (int d = 0; d < 3; d++) {
                        s->f[3 * j + d] += Fj[d];
                        s->torque[3 * j + d] += Tj[d];
                    }
                    s->fscale[j] += sqrt(Fj[0] * Fj[0] + Fj[1] * Fj[1] + Fj[2] * Fj[2]);
                    s->tscale[j] += sqrt(Tj[0] * Tj[0] + Tj[1] * Tj[1] + Tj[2] * Tj[2]);
                }
            } else if (rsq < cdm * radsum * radsum) {
                chain_close(&s->mp, h, &e->flag);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FixWallGran::post_force_primitive (fix_wall_gran.cpp:988-1120) with       */
/* Walls::Granular<>::compute_force (fix_wall_gran_base.h:159-378) and       */
/* Plane::resolveContact (primitive_wall_definitions.h:130-141).             */

static void wall_plane(or_sim *s, int k, int shearupdate)
{
    const int d = s->pdim[k];
    const double p = s->ppos[k];
    const double cdm = s->p.cdf - 1.0;
    static const double vwall[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < s->n; i++) {
        if (!s->pin[(size_t)i * OR_MAXPLANES + k]) continue;
        double *h = &s->phist_w[((size_t)i * OR_MAXPLANES + k) * s->dnum_w];
        const double radi = s->radius[i];
        const double pos = s->x[3 * i + d];
        double delta[3] = {0, 0, 0}, deltan;
        if (pos > p) {
            delta[d] = p - pos;
            deltan = pos - p - radi;
        } else {
            delta[d] = p - pos;
            deltan = p - pos - radi;
        }
        if (deltan > s->cutneighmax) continue;
        if (deltan <= 0 || deltan < cdm * radi) {
            const int intersect = (deltan <= 0);
            sidata_t sd;
            memset(&sd, 0, sizeof(sd));
            sd.is_wall = 1;
            sd.radi = radi;
            sd.deltan = -deltan;
            sd.delta[0] = -delta[0]; sd.delta[1] = -delta[1]; sd.delta[2] = -delta[2];
            sd.v_i = &s->v[3 * i]; sd.v_j = vwall;
            sd.omega_i = &s->omega[3 * i]; sd.omega_j = vwall;
            sd.r = sd.radi - sd.deltan;
            sd.rsq = sd.r * sd.r;
            const double rinv = 1.0 / sd.r;
            sd.rinv = rinv;
            sd.itype = s->type[i];
            sd.jtype = s->ptype[k];
            sd.meff = s->rmass[i];
            sd.mi = s->rmass[i];
            sd.en[0] = sd.delta[0] * rinv;
            sd.en[1] = sd.delta[1] * rinv;
            sd.en[2] = sd.delta[2] * rinv;
            sd.radsum = sd.radi;
            sd.update_history = shearupdate;
            sd.hist = h;
            sd.model = &s->mw;
            sd.flags = NULL;
            double Fi[3] = {0, 0, 0}, Ti[3] = {0, 0, 0}, Fj[3] = {0, 0, 0}, Tj[3] = {0, 0, 0};
            if (intersect) {
                chain_intersect(s, &sd, Fi, Ti, Fj, Tj);
                for (int q = 0; q < 3; q++) {
                    s->f[3 * i + q] += Fi[q];
                    s->torque[3 * i + q] += Ti[q];
                }
                s->fscale[i] += sqrt(Fi[0] * Fi[0] + Fi[1] * Fi[1] + Fi[2] * Fi[2]);
                s->tscale[i] += sqrt(Ti[0] * Ti[0] + Ti[1] * Ti[1] + Ti[2] * Ti[2]);
            } else {
                chain_close(&s->mw, h, NULL);
            }
        } else {
            for (int q = 0; q < s->dnum_w; q++) h[q] = 0.0;
        }
    }
}

/* FixGravity::post_force (fix_gravity.cpp:301-352), rmass branch */
static void gravity(or_sim *s)
{
    for (int i = 0; i < s->n; i++) {
        const double massone = s->rmass[i];
        s->f[3 * i] += massone * s->gacc[0];
        s->f[3 * i + 1] += massone * s->gacc[1];
        s->f[3 * i + 2] += massone * s->gacc[2];
        s->fscale[i] += fabs(massone * s->gacc[0]) + fabs(massone * s->gacc[1]) + fabs(massone * s->gacc[2]);
    }
}

static void freeze(or_sim *s)
{
    for (int i = 0; i < s->n; i++)
        if (s->freeze[i])
            for (int d = 0; d < 3; d++) s->f[3 * i + d] = s->torque[3 * i + d] = 0.0;
}

/* FixSetForce::post_force, constant values (fix_setforce.cpp:246-277) */
static void setforce(or_sim *s, op_t *o)
{
    o->foriginal[0] = o->foriginal[1] = o->foriginal[2] = 0.0;
    for (int i = 0; i < s->n; i++)
        if (s->gmask[i] & (1 << o->bit)) {
            for (int d = 0; d < 3; d++) o->foriginal[d] += s->f[3 * i + d];
            for (int d = 0; d < 3; d++)
                if (o->style[d]) s->f[3 * i + d] = o->val[d];
        }
}

/* FixAveForce::post_force (fix_aveforce.cpp:222-281); the z value is the
 * equal-style variable c0 - (time - t0)*c1 evaluated at this step */
static void aveforce(or_sim *s, op_t *o)
{
    double fo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < s->n; i++)
        if (s->gmask[i] & (1 << o->bit)) {
            fo[0] += s->f[3 * i];
            fo[1] += s->f[3 * i + 1];
            fo[2] += s->f[3 * i + 2];
            fo[3] += 1.0;
        }
    for (int d = 0; d < 3; d++) o->foriginal[d] = fo[d];
    const int ncount = (int)fo[3];
    if (ncount == 0) return;
    const double zvalue = o->c0 - (or_time(s) - o->t0) * o->c1;
    double fave[3];
    fave[0] = fo[0] / ncount + o->val[0];
    fave[1] = fo[1] / ncount + o->val[1];
    fave[2] = fo[2] / ncount + zvalue;
    for (int i = 0; i < s->n; i++)
        if (s->gmask[i] & (1 << o->bit))
            for (int d = 0; d < 3; d++)
                if (o->style[d]) s->f[3 * i + d] = fave[d];
}

void or_compute_forces(or_sim *s, int shearupdate)
{
    /* Verlet::force_clear (verlet.cpp:406-430) */
    memset(s->f, 0, (size_t)s->n * 3 * sizeof(double));
    memset(s->torque, 0, (size_t)s->n * 3 * sizeof(double));
    memset(s->fscale, 0, (size_t)s->n * sizeof(double));
    memset(s->tscale, 0, (size_t)s->n * sizeof(double));
    /* FixWallGran::pre_force (fix_wall_gran.cpp:688-706) */
    if (s->wall_rebuild) build_wall_lists(s);
    s->wall_rebuild = 0;
    pair_compute(s, shearupdate);
    for (int q = 0; q < s->nops; q++) {
        op_t *o = &s->ops[q];
        if (o->kind == OP_WALL) wall_plane(s, o->arg, shearupdate);
        else if (o->kind == OP_GRAVITY) gravity(s);
        else if (o->kind == OP_FREEZE) freeze(s);
        else if (o->kind == OP_SETFORCE) setforce(s, o);
        else if (o->kind == OP_AVEFORCE) aveforce(s, o);
    }
}

/* ------------------------------------------------------------------------ */
/* FixNVESphere (fix_nve_sphere.cpp:134-243), CAddRhoFluid = 0, 3d           */

static void initial_integrate(or_sim *s)
{
    const double dtf = 0.5 * s->p.dt;   /* dtf = 0.5*dt*ftm2v, ftm2v = 1 */
    const double dtv = s->p.dt;
    const double dtfrotate = dtf / 0.4;
    for (int i = 0; i < s->n; i++) {
        const double dtfm = dtf / (s->rmass[i] * (1. + 0.0 / 1.0));
        double *v = &s->v[3 * i], *x = &s->x[3 * i], *f = &s->f[3 * i];
        v[0] += dtfm * f[0]; v[1] += dtfm * f[1]; v[2] += dtfm * f[2];
        x[0] += dtv * v[0]; x[1] += dtv * v[1]; x[2] += dtv * v[2];
        const double dtirotate = dtfrotate / (s->radius[i] * s->radius[i] * s->rmass[i]);
        double *w = &s->omega[3 * i], *t = &s->torque[3 * i];
        w[0] += dtirotate * t[0]; w[1] += dtirotate * t[1]; w[2] += dtirotate * t[2];
    }
}

static void final_integrate(or_sim *s)
{
    const double dtf = 0.5 * s->p.dt;
    const double dtfrotate = dtf / 0.4;
    for (int i = 0; i < s->n; i++) {
        const double dtfm = dtf / (s->rmass[i] * (1. + 0.0 / 1.0));
        double *v = &s->v[3 * i], *f = &s->f[3 * i];
        v[0] += dtfm * f[0]; v[1] += dtfm * f[1]; v[2] += dtfm * f[2];
        const double dtirotate = dtfrotate / (s->radius[i] * s->radius[i] * s->rmass[i]);
        double *w = &s->omega[3 * i], *t = &s->torque[3 * i];
        w[0] += dtirotate * t[0]; w[1] += dtirotate * t[1]; w[2] += dtirotate * t[2];
    }
}

/* Neighbor::decide / check_distance (neighbor.cpp:1362-1465), no box change */
static int decide(or_sim *s)
{
    s->ago++;
    if (s->ago >= s->p.neigh_delay && s->ago % s->p.neigh_every == 0) {
        if (!s->p.neigh_check) return 1;
        const double trig = 0.5 * s->p.skin;
        const double deltasq = trig * trig;
        int flag = 0;
        for (int i = 0; i < s->n; i++) {
            const double dx = s->x[3 * i] - s->xhold[3 * i];
            const double dy = s->x[3 * i + 1] - s->xhold[3 * i + 1];
            const double dz = s->x[3 * i + 2] - s->xhold[3 * i + 2];
            const double rsq = dx * dx + dy * dy + dz * dz;
            if (rsq > deltasq) flag = 1;
        }
        return flag;
    }
    return 0;
}

void or_setup(or_sim *s)
{
    s->setupflag = 1;
    or_rebuild(s);
    or_compute_forces(s, 0);
    s->setupflag = 0;
}

void or_run(or_sim *s, int nsteps)
{
    for (int st = 0; st < nsteps; st++) {
        s->ntimestep++;
        initial_integrate(s);
        if (decide(s)) or_rebuild(s);
        or_compute_forces(s, 1);
        final_integrate(s);
    }
}

/* end of a `run` command: Verlet::cleanup -> Update::update_time */
void or_cleanup(or_sim *s)
{
    s->atime += (s->ntimestep - s->atimestep) * s->p.dt;
    s->atimestep = s->ntimestep;
}

/* ------------------------------------------------------------------------ */

void or_get_state(const or_sim *s, double *x, double *v, double *omega, double *f, double *torque)
{
    const size_t n3 = (size_t)s->n * 3 * sizeof(double);
    if (x) memcpy(x, s->x, n3);
    if (v) memcpy(v, s->v, n3);
    if (omega) memcpy(omega, s->omega, n3);
    if (f) memcpy(f, s->f, n3);
    if (torque) memcpy(torque, s->torque, n3);
}

long long or_timestep(const or_sim *s) { return s->ntimestep; }
int or_dnum(const or_sim *s) { return s->dnum; }
int or_dnum_wall(const or_sim *s) { return s->dnum_w; }

int or_get_pairs(const or_sim *s, int *tag_i, int *tag_j, int *shift, double *hist, int *flag)
{
    int c = 0;
    const int dnum = s->dnum;
    for (int i = 0; i < s->n; i++)
        for (int k = 0; k < s->ncount[i]; k++) {
            const int idx = s->nstart[i] + k;
            const entry_t *e = &s->ent[idx];
            const double *h = &s->hist[(size_t)idx * dnum];
            if (tag_i) {
                tag_i[c] = s->tag[i]; tag_j[c] = s->tag[e->j]; shift[c] = e->shift;
                if (hist) for (int d = 0; d < dnum; d++) hist[(size_t)c * dnum + d] = h[d];
                if (flag) flag[c] = e->flag;
            }
            c++;
            if (e->shift == 13) {
                if (tag_i) {
                    tag_i[c] = s->tag[e->j]; tag_j[c] = s->tag[i]; shift[c] = 13;
                    if (hist) for (int d = 0; d < dnum; d++) hist[(size_t)c * dnum + d] = -h[d];
                    if (flag) flag[c] = e->flag;
                }
                c++;
            }
        }
    return c;
}

void or_get_wall_history(const or_sim *s, double *hist)
{
    for (int i = 0; i < s->n; i++)
        for (int k = 0; k < s->nplanes; k++)
            for (int d = 0; d < s->dnum_w; d++)
                hist[((size_t)i * s->nplanes + k) * s->dnum_w + d] =
                    s->phist_w[((size_t)i * OR_MAXPLANES + k) * s->dnum_w + d];
}

void or_get_force_scale(const or_sim *s, double *fscale, double *tscale)
{
    memcpy(fscale, s->fscale, (size_t)s->n * sizeof(double));
    memcpy(tscale, s->tscale, (size_t)s->n * sizeof(double));
}
