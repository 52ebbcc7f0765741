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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXDNUM 8
#define T1 (OR_MAXTYPES + 1)

enum { OP_WALL = 1, OP_GRAVITY = 2, OP_FREEZE = 3, OP_SETFORCE = 4, OP_AVEFORCE = 5, OP_MESH = 6, OP_MOVE = 7 };

/* contact_model_constants.h:64-67 */
enum { CONTACT_NORMAL_MODEL = 1, CONTACT_COHESION_MODEL = 2, CONTACT_TANGENTIAL_MODEL = 4,
       CONTACT_ROLLING_MODEL = 8 };

typedef struct {
    int j;       /* owned atom index */
    int shift;   /* (sx+1)+3(sy+1)+9(sz+1), 13 = no image */
    int flag;    /* contact flags (CONTACT_* bits) */
} entry_t;

#define OR_MAX_SLOTS 64

/* history layout of one model chain; slots are registered in constructor
 * order surface, normal, cohesion, tangential, rolling (contact_models.h:
 * 150-157) */
typedef struct {
    int normal;      /* OR_NORMAL_* */
    int limitf;
    int cohesion;    /* OR_COH_* */
    int off_coh;     /* bond: 14 slots, bond/nonlinear: 28, else -1 */
    int off_norm;    /* hysteretic: 12 slots */
    int off_shear;   /* 3 slots */
    int off_roll;    /* 3 slots or -1 */
    int dnum;
    /* newtonflag per slot (fix_contact_history.cpp:404-408); the longest
     * chain, hysteretic + bond/nonlinear + shear + rolling, has 46 slots */
    int newton[OR_MAX_SLOTS];
} model_t;

typedef struct {
    int id, kind, arg, bit;
    int style[3];
    double val[3];
    double c0, t0, c1;
    double foriginal[3];
    double *xorig;          /* fix move: x at creation, n*3 */
    long long time_origin;
} op_t;

#define MESH_NEIGH_MAX 5   /* SurfaceMesh<3,5> (tri_mesh.h:67) */

/* one triangle-mesh wall: fix mesh/surface + fix wall/gran ... mesh */
typedef struct {
    int ntri;
    double *node, *center, *rbound, *evec, *elen, *enorm, *snorm;
    int *obtuse, *nneigh, *neigh;
    unsigned char *eact, *cact;
    double curvature, precision;
    int wtype;
    int moving;
    double vel[3];
    double *node_re;   /* node positions at the last rebuild (moving mesh) */
    /* triangle -> candidate particles, rebuilt with the neighbour list */
    int *tstart, *tlist;
    /* per-particle contact slots (FixContactHistoryMesh) */
    int *nneighs, *npartner, *sstart;
    int scap;
    int *partner;
    unsigned char *keep, *inter;
    double *hist;
    double force[3];
    /* fix mesh/surface/stress (MeshModuleStress): total force and torque
     * on the mesh about p_ref, which moves with the mesh */
    int stress;
    double pref[3], stress_f[3], stress_t[3];
} mesh_t;

struct or_sim {
    or_params p;
    int ntypes;
    double Yeff[T1 * T1], Geff[T1 * T1], betaeff[T1 * T1];
    double mu[T1 * T1], mur[T1 * T1], etar[T1 * T1];
    double restlog[T1 * T1];
    double hyst[OR_NHYST][T1 * T1];
    double bond[OR_NBOND][T1 * T1];
    or_bond_settings bs;
    int integ_bit;    /* nve/sphere group bit, -1 = all */
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

    int nmesh;
    mesh_t mesh[4];

    int nops;
    op_t ops[32];
    int next_op_id;
    int has_freeze;
    double gacc[3];

    long long ntimestep;
    int ago;
    int setupflag;
    int wall_rebuild;
    long long next_reneighbor;   /* FixMesh::next_reneighbor, -1 none */
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
static void model_layout(model_t *m, int normal, int limitf, int rolling, int cohesion)
{
    int k = 0;
    memset(m, 0, sizeof(*m));
    m->normal = normal;
    m->limitf = limitf;
    m->cohesion = cohesion;
    m->off_norm = -1;
    if (normal != OR_NORMAL_HERTZ) {
        m->off_norm = k;
        for (int q = 0; q < 10; q++) m->newton[k++] = 0;
        m->newton[k++] = 1;
        m->newton[k++] = 1;
    }
    /* cohesion bond (cohesion_model_bond.h:237-250): contflag, initial_dist,
     * contactPos x3 with newtonflag 0, ft x3, torquen x3, torquet x3 with 1;
     * bond/nonlinear (cohesion_model_bond_nonlinear.h:204-230): the same 14
     * with theta_n/theta_t in place of the torques, then deltan_max,
     * thetan_max x3, thetat_max x3, thetan_min x3, thetat_min x3,
     * deltan_min, all newtonflag 0 */
    m->off_coh = -1;
    if (cohesion != OR_COH_OFF) {
        m->off_coh = k;
        for (int q = 0; q < 5; q++) m->newton[k++] = 0;
        for (int q = 0; q < 9; q++) m->newton[k++] = 1;
        if (cohesion == OR_COH_BOND_NONLINEAR)
            for (int q = 0; q < 14; q++) m->newton[k++] = 0;
    }
    m->off_shear = k;
    for (int q = 0; q < 3; q++) m->newton[k++] = 1;
    m->off_roll = -1;
    if (rolling == OR_ROLL_EPSD || rolling == OR_ROLL_EPSD2 || rolling == OR_ROLL_EPSD3) {
        m->off_roll = k;
        for (int q = 0; q < 3; q++) m->newton[k++] = 1;
    }
    m->dnum = k;
}

or_sim *or_create(const or_params *p)
{
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->p = *p;
    model_layout(&s->mp, p->normal_model, p->limit_force, p->rolling_model, p->cohesion_model);
    model_layout(&s->mw, p->wall_normal_model, p->wall_limit_force, p->rolling_model, OR_COH_OFF);
    s->integ_bit = -1;
    s->bs.damping = 1;
    s->dnum = s->mp.dnum;
    s->dnum_w = s->mw.dnum;
    s->next_reneighbor = -1;
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
    for (int k = 0; k < s->nmesh; k++) {
        mesh_t *m = &s->mesh[k];
        free(m->node); free(m->node_re); free(m->center); free(m->rbound); free(m->evec); free(m->elen);
        free(m->enorm); free(m->snorm); free(m->obtuse); free(m->nneigh); free(m->neigh);
        free(m->eact); free(m->cact); free(m->tstart); free(m->tlist); free(m->nneighs);
        free(m->npartner); free(m->sstart); free(m->partner); free(m->keep); free(m->inter);
        free(m->hist);
    }
    for (int q = 0; q < s->nops; q++) free(s->ops[q].xorig);
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
    if (s->nops >= (int)(sizeof(s->ops) / sizeof(s->ops[0]))) {
        fprintf(stderr, "dem_oracle: more than %d fixes\n", (int)(sizeof(s->ops) / sizeof(s->ops[0])));
        abort();
    }
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

int or_add_gravity(or_sim *s, double ax, double ay, double az)
{
    s->gacc[0] = ax; s->gacc[1] = ay; s->gacc[2] = az;
    return new_op(s, OP_GRAVITY)->id;
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
    for (int q = 0; q < s->nops; q++) {
        if (s->ops[q].id != id) s->ops[k++] = s->ops[q];
        else free(s->ops[q].xorig);
    }
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

static void mesh_sort_contacts(or_sim *s, mesh_t *m);
static void mesh_rebuild(or_sim *s, mesh_t *m);
static void mesh_move(or_sim *s, mesh_t *m);
static void mesh_post_force(or_sim *s, mesh_t *m, int shearupdate);

void or_rebuild(or_sim *s)
{
    pre_exchange(s);
    for (int k = 0; k < s->nmesh; k++) mesh_sort_contacts(s, &s->mesh[k]);
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
    int i, j;           /* atom indices (cohesion models read x, omega) */
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

/* RollingModel<ROLLING_CDTNONLINEAR2>::surfacesIntersect
 * (rolling_model_cdtnonlinear2.h:96-172): constant torque mu_r |..| Fn reff
 * along the relative rolling velocity, no history, no contact flag */
static void rolling_cdt2(const or_sim *s, sidata_t *sd, double Ti[3], double Tj[3])
{
    const double rmu = s->mur[sd->itype * T1 + sd->jtype];
    double rt[3] = {0, 0, 0};
    const double radi = sd->radi, radj = sd->radj;
    const double reff = sd->is_wall ? radi : (radi * radj / (radi + radj));
    const double enx = sd->en[0], eny = sd->en[1], enz = sd->en[2];
    double mag, w[3];
    if (sd->is_wall) {
        w[0] = sd->wr[0]; w[1] = sd->wr[1]; w[2] = sd->wr[2];
    } else {
        w[0] = sd->omega_i[0] - sd->omega_j[0];
        w[1] = sd->omega_i[1] - sd->omega_j[1];
        w[2] = sd->omega_i[2] - sd->omega_j[2];
    }
    mag = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (mag > 0.) {
        const double Fn = sd->Fn;
        if (sd->is_wall) {
            rt[0] = rmu * Fn * w[0] / mag * reff;
            rt[1] = rmu * Fn * w[1] / mag * reff;
            rt[2] = rmu * Fn * w[2] / mag * reff;
        } else {
            const double sc = rmu * Fn * reff / mag;
            rt[0] = w[0] * sc; rt[1] = w[1] * sc; rt[2] = w[2] * sc;
        }
        if (!s->p.torsion_torque) {
            const double d = rt[0] * enx + rt[1] * eny + rt[2] * enz;
            const double n0 = enx * d, n1 = eny * d, n2 = enz * d;
            rt[0] = rt[0] - n0; rt[1] = rt[1] - n1; rt[2] = rt[2] - n2;
        }
    }
    Ti[0] -= rt[0]; Ti[1] -= rt[1]; Ti[2] -= rt[2];
    Tj[0] += rt[0]; Tj[1] += rt[1]; Tj[2] += rt[2];
}

/* ------------------------------------------------------------------------ */
/* Cohesion models of the pair chain                                         */

enum { BT_LAMBDA = 0, BT_KN, BT_KT, BT_DFN, BT_DFT, BT_DTN, BT_DTT, BT_MAXDIST, BT_MAXSIGMA, BT_MAXTAU,
       BT_RATIOTC, BT_CREATEDIST, BT_KFN1, BT_KUFN1, BT_KCFN1, BT_KFN2, BT_KUFN2, BT_KCFN2, BT_KFT,
       BT_KTN, BT_KUTN, BT_KCTN, BT_KTT, BT_KUTT, BT_KCTT };

static inline int isgn(double v) { return (0.0 < v) - (v < 0.0); }

/* vectorProject3D (vector_liggghts.h:437-442) */
static void project3(const double *v, const double *on, double *out)
{
    double nrm[3] = {on[0], on[1], on[2]};
    const double norm = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const double invnorm = (norm == 0.) ? 0. : 1. / norm;
    nrm[0] *= invnorm; nrm[1] *= invnorm; nrm[2] *= invnorm;
    const double d = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
    out[0] = d * nrm[0]; out[1] = d * nrm[1]; out[2] = d * nrm[2];
}

/* damped component: x - c |x| m with m = sgn(rate) or, smoothed,
 * clamp(rate / max(minvel, 0.01 x dt), -1, 1) */
static double bond_damped(double x, double c, double rate, int smooth, double minvel, double dt)
{
    double m;
    if (smooth) {
        const double dv = 0.01 * x * dt;
        m = fmin(1.0, fmax(-1.0, rate / fmax(minvel, dv)));
    } else {
        m = isgn(rate);
    }
    return x - c * fabs(x) * m;
}

/* one component of the bond/nonlinear hysteretic torque
 * (cohesion_model_bond_nonlinear.h:660-760): returns the torque, records
 * the extreme angles */
static double bond_nl_torque(double th, double *tmax, double *tmin, double k, double ku, double kc, double Jl)
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

/* CohesionModel<COHESION_BOND>::surfacesClose (cohesion_model_bond.h:490-
 * 935) and CohesionModel<COHESION_BOND_NONLINEAR>::surfacesClose
 * (cohesion_model_bond_nonlinear.h:396-960) for a particle pair, called
 * from surfacesIntersect as well.  Dissipation, elastic potential,
 * strengthening factor, random-id creation and walls are not restated.
 * bond adds its force to the chain; bond/nonlinear replaces it.  Returns
 * has_force_update. */
static int cohesion_bond(const or_sim *s, sidata_t *sd, double Fi[3], double Ti[3], double Fj[3], double Tj[3])
{
    const model_t *m = sd->model;
    const int nl = m->cohesion == OR_COH_BOND_NONLINEAR;
    const int upd = sd->update_history;
    const int t = sd->itype * T1 + sd->jtype;
    const double lambda = s->bond[BT_LAMBDA][t];
    const double small = nl ? 1.e-15 : 1.e-6;   /* SMALL_COHESION_MODEL_BOND */
    if (lambda < small) return 0;
    double *hb = sd->hist + m->off_coh;
    const double r_create = sqrt(sd->rsq);
    const double r = r_create;
    const int i = sd->i;
    if (upd) {
        if (hb[0] < small) {
            const int create = (s->bs.create_always || (s->ntimestep == (long long)(int)s->bs.ts_create))
                               && (r_create < s->bond[BT_CREATEDIST][t]);
            if (!create) return 0;
            /* createBond */
            if (sd->flags) *sd->flags |= CONTACT_COHESION_MODEL;
            hb[0] = 1.0;
            hb[1] = r;
            for (int d = 0; d < 3; d++) hb[2 + d] = s->x[3 * i + d] - sd->delta[d];
            for (int d = 5; d < 14; d++) hb[d] = 0.0;
        }
    } else if (hb[0] < small) {
        return 0;
    }
    double ft[3], tn[3], tt[3];
    for (int d = 0; d < 3; d++) {
        ft[d] = hb[5 + d];
        tn[d] = hb[8 + d];
        tt[d] = hb[11 + d];
    }
    const double *delta = sd->delta;
    if (!s->bs.stress_break && r > s->bond[BT_MAXDIST][t] && upd) {
        /* breakBond */
        if (sd->flags) *sd->flags &= ~CONTACT_COHESION_MODEL;
        hb[0] = 0.;
        hb[1] = 0.;
        return 0;
    }
    if (sd->flags) *sd->flags |= CONTACT_COHESION_MODEL;
    const double rinv = 1. / r;
    double en[3] = {rinv * delta[0], rinv * delta[1], rinv * delta[2]};
    const double *vi = sd->v_i, *vj = sd->v_j;
    const double *omegai = sd->omega_i, *omegaj = sd->omega_j;
    const double radi = sd->radi, radj = sd->radj;
    const double rb = lambda * (radi < radj ? radi : radj);
    const double A = M_PI * rb * rb;
    const double J = 0.5 * A * rb * rb;
    const double I = 0.5 * J;
    const double dt = s->p.dt;
    const double radsuminv = 1. / (radi + radj);
    const double cri = r * radi * radsuminv;
    const double crj = r * radj * radsuminv;
    double vr[3], vn[3], vt[3], wr[3], vtr[3], wn[3], wt[3], tmp1[3], tmp2[3];
    for (int d = 0; d < 3; d++) vr[d] = vi[d] - vj[d];
    project3(vr, en, vn);
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
    project3(wr, en, wn);
    for (int d = 0; d < 3; d++) wt[d] = wr[d] - wn[d];

    double nforce[3] = {0., 0., 0.}, nfd[3] = {0., 0., 0.}, tfd[3] = {0., 0., 0.};
    double ntd[3] = {0., 0., 0.}, ttd[3] = {0., 0., 0.};
    const double displacement = hb[1] - r;
    const double minvel = 1e-5 * fmin(radi, radj) / dt;
    const int damp = s->bs.damping, smooth = s->bs.damping_smooth;
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
            frcmag = s->bond[BT_KN][t] * A * displacement;
        } else {
            const double k1 = s->bond[BT_KFN1][t], ku1 = s->bond[BT_KUFN1][t], kc1 = s->bond[BT_KCFN1][t];
            const double k2 = s->bond[BT_KFN2][t], ku2 = s->bond[BT_KUFN2][t], kc2 = s->bond[BT_KCFN2][t];
            const double d = displacement;
            if (!s->bs.tension_variant) {
                const double c1 = pow((ku1 - k1) / (ku1 + kc1), 2.0) * dmax;
                const double c2 = pow((ku2 - k2) / (ku2 + kc2), 1.0) * dmin;
                if (d < dmin) frcmag = k2 * A * d;
                else if (d < c2) frcmag = ku2 * A * d + (k2 - ku2) * A * dmin;
                else if (d < 0.0) frcmag = -kc2 * A * d;
                else if (d < c1) frcmag = -kc1 * pow(fabs(d), 0.5);
                else if (d < dmax) frcmag = ku1 * pow(fabs(d), 0.5) + (k1 - ku1) * pow(fabs(dmax), 0.5);
                else frcmag = k1 * pow(fabs(d), 0.5);
            } else {
                /* examples/LIGGGHTS/INL/cohesive_bond_nonlinear_tension/src/
                 * cohesion_model_bond_nonlinear_tension.h:590-613 */
                const double c1 = pow((ku1 - k1) / (ku1 + kc1), 1.0) * dmax;
                const double c2 = pow((ku2 - k2) / (ku2 + kc2), 2.0) * dmin;
                if (d <= dmin) frcmag = -k2 * pow(fabs(d), 0.5);
                else if (d < c2) frcmag = -(ku2 * pow(fabs(d), 0.5) + (k2 - ku2) * pow(fabs(dmin), 0.5));
                else if (d < 0.0) frcmag = kc2 * pow(fabs(d), 0.5);
                else if (d < c1) frcmag = -kc1 * A * d;
                else if (d < dmax) frcmag = ku1 * A * d + (k1 - ku1) * A * dmax;
                else frcmag = k1 * A * d;
            }
        }
        for (int d = 0; d < 3; d++) nforce[d] = frcmag * en[d];
        for (int d = 0; d < 3; d++)
            nfd[d] = damp ? bond_damped(nforce[d], s->bond[BT_DFN][t], vn[d], smooth, minvel, dt) : nforce[d];
    }
    /* shear: rotate the old force into the tangent plane, add the increment */
    {
        const double kft = nl ? s->bond[BT_KFT][t] : s->bond[BT_KT][t];
        const double c = -kft * A * dt;
        double dtforce[3] = {c * vtr[0], c * vtr[1], c * vtr[2]};
        project3(ft, en, tmp1);
        for (int d = 0; d < 3; d++) ft[d] = ft[d] - tmp1[d];
        for (int d = 0; d < 3; d++) ft[d] = ft[d] + dtforce[d];
        for (int d = 0; d < 3; d++)
            tfd[d] = damp ? bond_damped(ft[d], s->bond[BT_DFT][t], vtr[d], smooth, minvel, dt) : ft[d];
    }
    double torque_n[3], torque_t[3];
    if (!nl) {
        const double ktp = s->bond[BT_KT][t], knp = s->bond[BT_KN][t];
        const double cn = -ktp * J * dt;
        const double dnt[3] = {cn * wn[0], cn * wn[1], cn * wn[2]};
        project3(tn, en, tn);
        for (int d = 0; d < 3; d++) tn[d] = tn[d] + dnt[d];
        for (int d = 0; d < 3; d++)
            ntd[d] = damp ? tn[d] - s->bond[BT_DTN][t] * fabs(tn[d]) * isgn(wn[d]) : tn[d];
        const double wtsq = wt[0] * wt[0] + wt[1] * wt[1] + wt[2] * wt[2];
        if (wtsq > 0) {
            const double ct = -knp * I * dt;
            const double dtt[3] = {ct * wt[0], ct * wt[1], ct * wt[2]};
            project3(tt, wt, tt);
            for (int d = 0; d < 3; d++) tt[d] = tt[d] + dtt[d];
            for (int d = 0; d < 3; d++)
                ttd[d] = damp ? tt[d] - s->bond[BT_DTT][t] * fabs(tt[d]) * isgn(wt[d]) : tt[d];
        }
        for (int d = 0; d < 3; d++) { torque_n[d] = tn[d]; torque_t[d] = tt[d]; }
    } else {
        /* twisting: tn holds the twist angle theta_n */
        const double ktn = s->bond[BT_KTN][t], kutn = s->bond[BT_KUTN][t], kctn = s->bond[BT_KCTN][t];
        const double ktt = s->bond[BT_KTT][t], kutt = s->bond[BT_KUTT][t], kctt = s->bond[BT_KCTT][t];
        double dth[3] = {dt * wn[0], dt * wn[1], dt * wn[2]};
        for (int d = 0; d < 3; d++) tn[d] = tn[d] + dth[d];
        const double cn = -ktn * J;
        for (int d = 0; d < 3; d++) torque_n[d] = cn * tn[d];
        project3(torque_n, wn, torque_n);
        for (int d = 0; d < 3; d++) {
            if (tn[d] > 0.0) hb[21 + d] = 0.0;
            else if (tn[d] < 0.0) hb[15 + d] = 0.0;
        }
        for (int d = 0; d < 3; d++) torque_n[d] = bond_nl_torque(tn[d], &hb[15 + d], &hb[21 + d], ktn, kutn, kctn, J);
        for (int d = 0; d < 3; d++)
            ntd[d] = damp ? torque_n[d] - s->bond[BT_DTN][t] * fabs(torque_n[d]) * isgn(wn[d]) : torque_n[d];
        /* bending: tt holds the bending angle theta_t */
        double dtt[3] = {dt * wt[0], dt * wt[1], dt * wt[2]};
        for (int d = 0; d < 3; d++) tt[d] = tt[d] + dtt[d];
        const double ct = -ktt * I;
        for (int d = 0; d < 3; d++) torque_t[d] = ct * tt[d];
        project3(torque_t, wt, torque_t);
        for (int d = 0; d < 3; d++) {
            if (tt[d] > 0.0) hb[24 + d] = 0.0;
            else if (tt[d] < 0.0) hb[18 + d] = 0.0;
        }
        for (int d = 0; d < 3; d++) torque_t[d] = bond_nl_torque(tt[d], &hb[18 + d], &hb[24 + d], ktt, kutt, kctt, I);
        for (int d = 0; d < 3; d++)
            ttd[d] = damp ? torque_t[d] - s->bond[BT_DTT][t] * fabs(torque_t[d]) * isgn(wt[d]) : torque_t[d];
    }
    if (s->bs.stress_break) {
        const double nforce_mag = sqrt(nforce[0] * nforce[0] + nforce[1] * nforce[1] + nforce[2] * nforce[2]);
        const double tforce_mag = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
        const double ntorque_mag = sqrt(torque_n[0] * torque_n[0] + torque_n[1] * torque_n[1] + torque_n[2] * torque_n[2]);
        const double ttorque_mag = sqrt(torque_t[0] * torque_t[0] + torque_t[1] * torque_t[1] + torque_t[2] * torque_t[2]);
        double maxSigma;
        int nstress, tstress;
        if (!nl) {
            maxSigma = s->bond[BT_MAXSIGMA][t] * ((s->bs.ratio_tc && displacement < 1e-16) ? s->bond[BT_RATIOTC][t] : 1.0);
            nstress = (1. * maxSigma) < (nforce_mag / A + ttorque_mag * rb / I);
            tstress = (1. * s->bond[BT_MAXTAU][t]) < (tforce_mag / A + ntorque_mag * rb / J);
        } else {
            maxSigma = s->bond[BT_MAXSIGMA][t] * ((s->bs.ratio_tc && displacement < -1.e-15) ? s->bond[BT_RATIOTC][t] : 1.0);
            nstress = maxSigma < (nforce_mag / A + ttorque_mag * rb / I);
            tstress = s->bond[BT_MAXTAU][t] < (tforce_mag / A + ntorque_mag * rb / J);
        }
        if ((nstress || tstress) && upd) {
            if (sd->flags) *sd->flags &= ~CONTACT_COHESION_MODEL;
            hb[0] = 0.;
            hb[1] = 0.;
            return 0;
        }
    }
    double tor[3];
    tor[0] = tfd[1] * en[2] - tfd[2] * en[1];
    tor[1] = tfd[2] * en[0] - tfd[0] * en[2];
    tor[2] = tfd[0] * en[1] - tfd[1] * en[0];
    double F[3], dTi[3], dTj[3];
    for (int d = 0; d < 3; d++) {
        F[d] = nfd[d] + tfd[d];
        dTi[d] = cri * tor[d] + ntd[d] + ttd[d];
        dTj[d] = crj * tor[d] - ntd[d] - ttd[d];
    }
    if (upd)
        for (int d = 0; d < 3; d++) {
            hb[5 + d] = ft[d];
            hb[8 + d] = tn[d];
            hb[11 + d] = tt[d];
        }
    if (nl) {
        /* bond/nonlinear replaces the chain's force so far */
        for (int d = 0; d < 3; d++) {
            Fi[d] = F[d]; Ti[d] = dTi[d];
            Fj[d] = -F[d]; Tj[d] = dTj[d];
        }
    } else {
        for (int d = 0; d < 3; d++) {
            Fi[d] += F[d]; Ti[d] += dTi[d];
            Fj[d] -= F[d]; Tj[d] += dTj[d];
        }
    }
    return 1;
}

int or_set_bond(or_sim *s, const double *tables, const or_bond_settings *bs)
{
    const int w = s->ntypes + 1;
    for (int q = 0; q < OR_NBOND; q++)
        for (int i = 0; i < w; i++)
            for (int j = 0; j < w; j++) s->bond[q][i * T1 + j] = tables[(size_t)q * w * w + i * w + j];
    s->bs = *bs;
    /* the model's contact distance factor (cohesion_model_bond.h:415-470,
     * cohesion_model_bond_nonlinear.h:330-380), registered as a maximum */
    double minrad = 1e300;
    for (int i = 0; i < s->n; i++) minrad = s->radius[i] < minrad ? s->radius[i] : minrad;
    double cdf_all = 1.;
    const int nl = s->mp.cohesion == OR_COH_BOND_NONLINEAR;
    for (int i = 1; i < w; i++)
        for (int j = 1; j < w; j++) {
            const int t = i * T1 + j;
            double cdf_one;
            if (!bs->stress_break) {
                cdf_one = 1.1 * 0.5 * s->bond[BT_MAXDIST][t] / minrad;
            } else {
                double stress = s->bond[BT_MAXSIGMA][t];
                if (bs->ratio_tc) stress = fmax(stress, stress * s->bond[BT_RATIOTC][t]);
                if (!nl)
                    cdf_one = 0.5 * (1.1 * s->bond[BT_CREATEDIST][t] / minrad
                                     + 1.1 * stress / (s->bond[BT_KN][t] * minrad));
                else
                    cdf_one = 0.5 * 1.1 * s->bond[BT_CREATEDIST][t] / minrad
                              + 0.5 * 1.1 * stress / (s->bond[BT_KFN2][t] * minrad);
            }
            cdf_all = cdf_one > cdf_all ? cdf_one : cdf_all;
        }
    if (cdf_all > s->p.cdf) s->p.cdf = cdf_all;
    return 0;
}

int or_add_move_linear(or_sim *s, int bit, const int set[3], const double v[3])
{
    op_t *o = new_op(s, OP_MOVE);
    o->bit = bit;
    for (int d = 0; d < 3; d++) {
        o->style[d] = set[d];
        o->val[d] = v[d];
    }
    o->xorig = malloc(sizeof(double) * 3 * (s->n > 0 ? s->n : 1));
    memcpy(o->xorig, s->x, sizeof(double) * 3 * s->n);
    o->time_origin = s->ntimestep;
    return o->id;
}

void or_set_integrate_group(or_sim *s, int bit)
{
    s->integ_bit = bit;
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
    if (sd->model->cohesion != OR_COH_OFF) cohesion_bond(s, sd, Fi, Ti, Fj, Tj);
    if (sd->flags) *sd->flags |= CONTACT_TANGENTIAL_MODEL;
    tangential_history(s, sd, Fi, Ti, Fj, Tj);
    if (sd->model->off_roll >= 0) {
        if (sd->flags) *sd->flags |= CONTACT_ROLLING_MODEL;
        rolling(s, sd, Ti, Tj);
    } else if (s->p.rolling_model == OR_ROLL_CDT2) {
        rolling_cdt2(s, sd, Ti, Tj);
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
                sd.i = i;
                sd.j = j;
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
                if (s->mp.cohesion != OR_COH_OFF) {
                    /* the cohesion model's surfacesClose may apply a force
                     * (has_force_update, pair_gran_base.h:417-430) */
                    sidata_t sd;
                    memset(&sd, 0, sizeof(sd));
                    sd.itype = s->type[i]; sd.jtype = s->type[j];
                    sd.radi = radi; sd.radj = radj; sd.radsum = radsum;
                    sd.rsq = rsq;
                    sd.delta[0] = delx; sd.delta[1] = dely; sd.delta[2] = delz;
                    sd.v_i = &s->v[3 * i]; sd.v_j = &s->v[3 * j];
                    sd.omega_i = &s->omega[3 * i]; sd.omega_j = &s->omega[3 * j];
                    sd.update_history = shearupdate;
                    sd.hist = h;
                    sd.model = &s->mp;
                    sd.flags = &e->flag;
                    sd.i = i;
                    sd.j = j;
                    if (cohesion_bond(s, &sd, Fi, Ti, Fj, Tj)) {
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
                    }
                }
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

/* ------------------------------------------------------------------------ */
/* Triangle meshes.  Topology: SurfaceMesh (surface_mesh_I.h:302-470,       */
/* 474-580, 874-1240) and MultiNodeMesh (multi_node_mesh_I.h:140-175,       */
/* 246-350).  Contact: TriMesh::resolveTriSphereContactBary and friends     */
/* (tri_mesh_I.h:65-320).  History: FixContactHistoryMesh                   */
/* (fix_contact_history_mesh_I.h:51-200, fix_contact_history_mesh.cpp:      */
/* 250-520).  Force: FixWallGran::post_force_mesh (fix_wall_gran.cpp:803-980) */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double mag3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void sdiv3(double *v, double s)
{
    const double sinv = 1. / s;
    v[0] = sinv * v[0]; v[1] = sinv * v[1]; v[2] = sinv * v[2];
}

/* MathExtraLiggghts::compDouble (math_extra_liggghts.h:560-571) */
static int comp_double(double a, double b, double prec)
{
    if (a == b) return 1;
    if (b == 0) return a < prec && a > -prec;
    const double x = (a - b);
    return x < prec && x > -prec;
}

static int nodes_equal(const double *a, const double *b, double prec)
{
    for (int i = 0; i < 3; i++)
        if (!comp_double(a[i], b[i], prec)) return 0;
    return 1;
}

#define MN(m, t, k) (&(m)->node[9 * (t) + 3 * (k)])
#define EV(m, t, k) (&(m)->evec[9 * (t) + 3 * (k)])
#define EN(m, t, k) (&(m)->enorm[9 * (t) + 3 * (k)])

static int contains_node(const mesh_t *m, int t, const double *p)
{
    for (int k = 0; k < 3; k++) {
        const double *n = MN(m, t, k);
        if (comp_double(n[0], p[0], m->precision) && comp_double(n[1], p[1], m->precision)
            && comp_double(n[2], p[2], m->precision))
            return k;
    }
    return -1;
}

/* share2Nodes (multi_node_mesh_I.h:281-325) */
static int share2(const mesh_t *m, int a, int b, int *ia1, int *ja1, int *ia2, int *ja2)
{
    double d[3];
    for (int q = 0; q < 3; q++) d[q] = m->center[3 * a + q] - m->center[3 * b + q];
    const double radsum = m->rbound[a] + m->rbound[b];
    if (dot3(d, d) > radsum * radsum) return 0;
    int ns = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (nodes_equal(MN(m, a, i), MN(m, b, j), m->precision)) {
                if (ns == 0) { *ia1 = i; *ja1 = j; }
                else { *ia2 = i; *ja2 = j; return 1; }
                ns++;
            }
    return 0;
}

/* nSharedNodes (multi_node_mesh_I.h:328-346) */
static int nshared(const mesh_t *m, int a, int b)
{
    double d[3];
    for (int q = 0; q < 3; q++) d[q] = m->center[3 * a + q] - m->center[3 * b + q];
    const double radsum = m->rbound[a] + m->rbound[b];
    if (dot3(d, d) > radsum * radsum) return 0;
    int ns = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (nodes_equal(MN(m, a, i), MN(m, b, j), m->precision)) ns++;
    return ns;
}

/* areCoplanarNodeNeighs (surface_mesh_I.h:922-962), ids = indices */
static int coplanar_node_neighs(const mesh_t *m, int a, int b)
{
    int neighs = 0;
    for (int i = 0; i < m->nneigh[a] && i < MESH_NEIGH_MAX; i++)
        if (m->neigh[MESH_NEIGH_MAX * a + i] == b) neighs = 1;
    if (!neighs && nshared(m, a, b) == 0) return 0;
    const double dot = dot3(&m->snorm[3 * a], &m->snorm[3 * b]);
    return fabs(dot) > m->curvature;
}

static void handle_shared_edge(mesh_t *m, int i, int ie, int j, int je, int coplanar)
{
    if (m->nneigh[i] < MESH_NEIGH_MAX) m->neigh[MESH_NEIGH_MAX * i + m->nneigh[i]] = j;
    if (m->nneigh[j] < MESH_NEIGH_MAX) m->neigh[MESH_NEIGH_MAX * j + m->nneigh[j]] = i;
    m->nneigh[i]++;
    m->nneigh[j]++;
    int overlap = 0;
    if (coplanar) {
        /* coplanarNeighsOverlap (surface_mesh_I.h:965-1004) */
        double vi[3], vj[3], pref[3], en[3];
        for (int q = 0; q < 3; q++) {
            pref[q] = MN(m, i, ie)[q];
            en[q] = EN(m, i, ie)[q];
        }
        for (int q = 0; q < 3; q++) {
            vi[q] = MN(m, i, (ie + 2) % 3)[q] - pref[q];
            vj[q] = MN(m, j, (je + 2) % 3)[q] - pref[q];
        }
        overlap = dot3(vi, en) * dot3(vj, en) > 0.;
    }
    if (!coplanar || overlap) {
        if (i < j) { m->eact[3 * i + ie] = 0; m->eact[3 * j + je] = 1; }
        else { m->eact[3 * i + ie] = 1; m->eact[3 * j + je] = 0; }
    } else {
        m->eact[3 * i + ie] = 0;
        m->eact[3 * j + je] = 0;
    }
}

typedef struct {
    int nvis, nhas;
    int *vis, *has;
    double *elist, *eend;
    int any_active;
} corner_walk_t;

/* checkNodeRecursive (surface_mesh_I.h:1188-1236) */
static void check_node(const mesh_t *m, int t, const double *p, corner_walk_t *w)
{
    for (int i = 0; i < w->nvis; i++)
        if (w->vis[i] == t) return;
    w->vis[w->nvis++] = t;
    const int k = contains_node(m, t, p);
    if (k < 0) return;
    const int ne = 2 * w->nhas;
    w->has[w->nhas++] = t;
    const int km = (k - 1 + 3) % 3;
    for (int q = 0; q < 3; q++) {
        w->elist[3 * ne + q] = EV(m, t, k)[q];
        w->elist[3 * (ne + 1) + q] = EV(m, t, km)[q];
        w->eend[3 * ne + q] = MN(m, t, (k + 1) % 3)[q];
        w->eend[3 * (ne + 1) + q] = MN(m, t, km)[q];
    }
    if (m->eact[3 * t + k]) w->any_active = 1;
    else if (m->eact[3 * t + km]) w->any_active = 1;
    for (int in = 0; in < m->nneigh[t] && in < MESH_NEIGH_MAX; in++) {
        const int nb = m->neigh[MESH_NEIGH_MAX * t + in];
        if (nb < 0) return;
        check_node(m, nb, p, w);
    }
}

/* handleCorner (surface_mesh_I.h:1139-1185); a node outside the box keeps
 * its initial active state, as there.  Domain::is_in_subdomain pads a
 * subdomain face lying on the box face by SMALL_DMBRDR = 1e-8 (domain_I.h:
 * 65-86, domain.h:64), which in one process is every face. */
static void handle_corner(const or_sim *s, mesh_t *m, int t, int k, corner_walk_t *w)
{
    double p[3];
    for (int q = 0; q < 3; q++) p[q] = MN(m, t, k)[q];
    w->nvis = w->nhas = 0;
    w->any_active = 0;
    check_node(m, t, p, w);
    for (int d = 0; d < 3; d++)
        if (!(p[d] >= s->p.boxlo[d] - 1.0e-8 && p[d] < s->p.boxhi[d] + 1.0e-8)) return;
    const int ne = 2 * w->nhas;
    int maxid = -1;
    for (int i = 0; i < w->nhas; i++)
        if (w->has[i] > maxid) maxid = w->has[i];
    int colinear = 0;
    for (int i = 0; i < ne; i++)
        for (int j = i + 1; j < ne; j++)
            if (fabs(dot3(&w->elist[3 * i], &w->elist[3 * j])) > m->curvature
                && !nodes_equal(&w->eend[3 * i], &w->eend[3 * j], m->precision))
                colinear = 1;
    if (colinear || !w->any_active) m->cact[3 * t + k] = 0;
    else if (t == maxid) m->cact[3 * t + k] = 1;
    else m->cact[3 * t + k] = 0;
}

static void mesh_element_props(mesh_t *m, int t)
{
    double avg[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
        for (int q = 0; q < 3; q++) avg[q] = MN(m, t, k)[q] + avg[q];
    sdiv3(avg, 3.0);
    for (int q = 0; q < 3; q++) m->center[3 * t + q] = avg[q];
    double rb = 0.;
    for (int k = 0; k < 3; k++) {
        double v[3];
        for (int q = 0; q < 3; q++) v[q] = m->center[3 * t + q] - MN(m, t, k)[q];
        const double mg = mag3(v);
        rb = rb > mg ? rb : mg;
    }
    m->rbound[t] = rb;
    /* calcEdgeVecLen */
    for (int k = 0; k < 3; k++) {
        double *e = EV(m, t, k);
        for (int q = 0; q < 3; q++) e[q] = MN(m, t, (k + 1) % 3)[q] - MN(m, t, k)[q];
        m->elen[3 * t + k] = mag3(e);
        sdiv3(e, m->elen[3 * t + k]);
    }
    /* calcSurfaceNorm */
    double *sn = &m->snorm[3 * t];
    cross3(EV(m, t, 0), EV(m, t, 1), sn);
    if (mag3(sn) < 1e-15) {
        cross3(EV(m, t, 1), EV(m, t, 2), sn);
        if (mag3(sn) < 1e-15) {
            cross3(EV(m, t, 2), EV(m, t, 0), sn);
            if (mag3(sn) < 1e-15) {
                const double tmp[3] = {1.1233, 2.123231, -3.3343434};
                cross3(EV(m, t, 0), tmp, sn);
            }
        }
    }
    sdiv3(sn, mag3(sn));
    /* calcEdgeNormals */
    for (int k = 0; k < 3; k++) {
        double *en = EN(m, t, k);
        cross3(EV(m, t, k), sn, en);
        if (mag3(en) < 1e-15) {
            for (int q = 0; q < 3; q++) en[q] = EV(m, t, (k + 1) % 3)[q];
        } else {
            sdiv3(en, mag3(en));
        }
    }
    /* calcObtuseAngleIndex, called for nodes 0,1,2 in turn: the last call
     * decides (only an obtuse angle at node 2 is recorded) */
    m->obtuse[t] = -1;
    for (int k = 0; k < 3; k++) {
        const double dd = dot3(EV(m, t, k), EV(m, t, (k - 1 + 3) % 3));
        m->obtuse[t] = dd > 0. ? k : -1;
    }
}

int or_add_mesh(or_sim *s, int ntri, const double *nodes, int wtype, double curvature, const double *vel)
{
    if (s->nmesh >= 4 || ntri < 1) return -1;
    mesh_t *m = &s->mesh[s->nmesh];
    memset(m, 0, sizeof(*m));
    m->ntri = ntri;
    m->wtype = wtype;
    m->curvature = curvature > 0 ? curvature : 1. - 0.00001;
    m->precision = 1e-8;
    if (vel) {
        m->moving = 1;
        for (int q = 0; q < 3; q++) m->vel[q] = vel[q];
    }
    m->node = malloc(sizeof(double) * 9 * ntri);
    /* MultiNodeMesh::addElement (multi_node_mesh_I.h:138-146) refuses an
     * element with two equal nodes; ids count the accepted ones */
    int kept = 0;
    for (int t = 0; t < ntri; t++) {
        const double *a = nodes + 9 * t;
        if (nodes_equal(a, a + 3, 1e-8) || nodes_equal(a + 3, a + 6, 1e-8) || nodes_equal(a, a + 6, 1e-8))
            continue;
        memcpy(m->node + 9 * kept, a, sizeof(double) * 9);
        kept++;
    }
    if (kept == 0) { free(m->node); return -1; }
    ntri = kept;
    m->ntri = ntri;
    m->node_re = malloc(sizeof(double) * 9 * ntri);
    memcpy(m->node_re, m->node, sizeof(double) * 9 * ntri);
    m->center = malloc(sizeof(double) * 3 * ntri);
    m->rbound = malloc(sizeof(double) * ntri);
    m->evec = malloc(sizeof(double) * 9 * ntri);
    m->enorm = malloc(sizeof(double) * 9 * ntri);
    m->elen = malloc(sizeof(double) * 3 * ntri);
    m->snorm = malloc(sizeof(double) * 3 * ntri);
    m->obtuse = malloc(sizeof(int) * ntri);
    m->nneigh = calloc(ntri, sizeof(int));
    m->neigh = malloc(sizeof(int) * MESH_NEIGH_MAX * ntri);
    m->eact = malloc(3 * ntri);
    m->cact = malloc(3 * ntri);
    for (int t = 0; t < ntri; t++) {
        mesh_element_props(m, t);
        for (int k = 0; k < MESH_NEIGH_MAX; k++) m->neigh[MESH_NEIGH_MAX * t + k] = -1;
        for (int k = 0; k < 3; k++) m->eact[3 * t + k] = m->cact[3 * t + k] = 1;
    }
    /* buildNeighbours: every earlier element that shares an edge */
    for (int i = 0; i < ntri; i++)
        for (int j = 0; j < i; j++) {
            int a1, b1, a2, b2;
            if (!share2(m, i, j, &a1, &b1, &a2, &b2)) continue;
            const int ie = (a1 + a2 == 2) ? 2 : (a1 < a2 ? a1 : a2);
            const int je = (b1 + b2 == 2) ? 2 : (b1 < b2 ? b1 : b2);
            const int cop = fabs(dot3(&m->snorm[3 * i], &m->snorm[3 * j])) >= m->curvature;
            handle_shared_edge(m, i, ie, j, je, cop);
        }
    corner_walk_t w;
    w.vis = malloc(sizeof(int) * ntri);
    w.has = malloc(sizeof(int) * ntri);
    w.elist = malloc(sizeof(double) * 6 * ntri);
    w.eend = malloc(sizeof(double) * 6 * ntri);
    for (int t = 0; t < ntri; t++)
        for (int k = 0; k < 3; k++) handle_corner(s, m, t, k, &w);
    free(w.vis); free(w.has); free(w.elist); free(w.eend);
    const int id = s->nmesh++;
    op_t *o = new_op(s, OP_MESH);
    o->arg = id;
    return o->id;
}

void or_mesh_topology(const or_sim *s, int im, int *eact, int *cact, int *obtuse, int *nneigh)
{
    const mesh_t *m = &s->mesh[im];
    for (int t = 0; t < m->ntri; t++) {
        for (int k = 0; k < 3; k++) {
            eact[3 * t + k] = m->eact[3 * t + k];
            cact[3 * t + k] = m->cact[3 * t + k];
        }
        obtuse[t] = m->obtuse[t];
        nneigh[t] = m->nneigh[t];
    }
}

int or_mesh_size(const or_sim *s, int im)
{
    return (im >= 0 && im < s->nmesh) ? s->mesh[im].ntri : 0;
}

void or_mesh_stress(or_sim *s, int im, int on, const double pref[3])
{
    mesh_t *m = &s->mesh[im];
    m->stress = on;
    for (int q = 0; q < 3; q++) m->pref[q] = pref ? pref[q] : 0.0;
}

/* FixMeshSurfaceStress::compute_vector (mesh_module_stress.cpp:479-488):
 * total force, torque about p_ref, p_ref */
void or_mesh_stress_vector(const or_sim *s, int im, double out[9])
{
    const mesh_t *m = &s->mesh[im];
    for (int q = 0; q < 3; q++) {
        out[q] = m->stress_f[q];
        out[3 + q] = m->stress_t[q];
        out[6 + q] = m->pref[q];
    }
}

void or_mesh_force(const or_sim *s, int im, double f[3])
{
    for (int q = 0; q < 3; q++) f[q] = s->mesh[im].force[q];
}

/* TriMesh::calcDist (tri_mesh_I.h:308-312) */
static double tri_calc_dist(const double *cs, const double *cp, double *delta)
{
    for (int q = 0; q < 3; q++) delta[q] = cp[q] - cs[q];
    double d[3];
    for (int q = 0; q < 3; q++) d[q] = cs[q] - cp[q];
    return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

#define SMALL_TRIMESH 1.e-10
#define LARGE_TRIMESH 1000000

static double tri_edge(const mesh_t *m, int t, int ie, const double *p, double *delta, double *bary)
{
    const int ip = (ie + 1) % 3, ipp = (ie + 2) % 3;
    double ntp[3];
    for (int q = 0; q < 3; q++) ntp[q] = p[q] - MN(m, t, ie)[q];
    const double dfn = dot3(ntp, EV(m, t, ie));
    double d;
    if (dfn < -SMALL_TRIMESH) {
        if (!m->cact[3 * t + ie]) return LARGE_TRIMESH;
        d = tri_calc_dist(p, MN(m, t, ie), delta);
        bary[ie] = 1.; bary[ip] = 0.; bary[ipp] = 0.;
    } else if (dfn > m->elen[3 * t + ie] + SMALL_TRIMESH) {
        if (!m->cact[3 * t + ip]) return LARGE_TRIMESH;
        d = tri_calc_dist(p, MN(m, t, ip), delta);
        bary[ie] = 0.; bary[ip] = 1.; bary[ipp] = 0.;
    } else {
        if (!m->eact[3 * t + ie]) return LARGE_TRIMESH;
        double cp[3];
        for (int q = 0; q < 3; q++) cp[q] = MN(m, t, ie)[q] + dfn * EV(m, t, ie)[q];
        d = tri_calc_dist(p, cp, delta);
        bary[ipp] = 0.;
        bary[ie] = 1. - dfn / m->elen[3 * t + ie];
        bary[ip] = 1. - bary[ie];
    }
    return d;
}

static double tri_corner(const mesh_t *m, int t, int in, int obtuse, const double *p, double *delta, double *bary)
{
    const int ip = (in + 1) % 3, ipp = (in + 2) % 3;
    const double *n = MN(m, t, in);
    if (obtuse) {
        double ntp[3], cp[3];
        for (int q = 0; q < 3; q++) ntp[q] = p[q] - n[q];
        double dfn = dot3(ntp, EV(m, t, ipp));
        if (dfn < SMALL_TRIMESH) {
            if (dfn > -m->elen[3 * t + ipp]) {
                if (!m->eact[3 * t + ipp]) return LARGE_TRIMESH;
                for (int q = 0; q < 3; q++) cp[q] = n[q] + dfn * EV(m, t, ipp)[q];
                bary[ip] = 0.;
                bary[in] = 1. + dfn / m->elen[3 * t + ipp];
                bary[ipp] = 1. - bary[in];
                return tri_calc_dist(p, cp, delta);
            } else {
                if (!m->cact[3 * t + ipp]) return LARGE_TRIMESH;
                bary[ipp] = 1.; bary[in] = bary[ip] = 0.;
                return tri_calc_dist(p, MN(m, t, ipp), delta);
            }
        }
        dfn = dot3(ntp, EV(m, t, in));
        if (dfn > -SMALL_TRIMESH) {
            if (dfn < m->elen[3 * t + in]) {
                if (!m->eact[3 * t + in]) return LARGE_TRIMESH;
                for (int q = 0; q < 3; q++) cp[q] = n[q] + dfn * EV(m, t, in)[q];
                bary[ipp] = 0.;
                bary[in] = 1. - dfn / m->elen[3 * t + in];
                bary[ip] = 1. - bary[in];
                return tri_calc_dist(p, cp, delta);
            } else {
                if (!m->cact[3 * t + ip]) return LARGE_TRIMESH;
                bary[ip] = 1.; bary[in] = bary[ipp] = 0.;
                return tri_calc_dist(p, MN(m, t, ip), delta);
            }
        }
    }
    if (!m->cact[3 * t + in]) return LARGE_TRIMESH;
    bary[in] = 1.; bary[ip] = bary[ipp] = 0.;
    return tri_calc_dist(p, MN(m, t, in), delta);
}

/* resolveTriSphereContactBary: distance - radius, delta from sphere centre
 * to the closest point, barycentric coordinates, region code */
static double tri_contact(const mesh_t *m, int t, double r, const double *cs, double *delta,
                          double *bary, int *barysign)
{
    double n0c[3];
    for (int q = 0; q < 3; q++) n0c[q] = cs[q] - MN(m, t, 0)[q];
    /* MathExtraLiggghts::calcBaryTriCoords (math_extra_liggghts.h:583-594) */
    const double a = dot3(n0c, EV(m, t, 0));
    const double b = dot3(n0c, EV(m, t, 2));
    const double c = dot3(EV(m, t, 0), EV(m, t, 2));
    const double oneMinCSqr = 1 - c * c;
    bary[1] = (a - b * c) / (m->elen[3 * t + 0] * oneMinCSqr);
    bary[2] = (a * c - b) / (m->elen[3 * t + 2] * oneMinCSqr);
    bary[0] = 1. - bary[1] - bary[2];
    const double invlen = 1. / (2. * m->rbound[t]);
    const int bs = (bary[0] > -m->precision * invlen) + 2 * (bary[1] > -m->precision * invlen)
                   + 4 * (bary[2] > -m->precision * invlen);
    *barysign = bs;
    const int ob = m->obtuse[t];
    double d = 1.;
    switch (bs) {
    case 1: d = tri_corner(m, t, 0, ob == 0, cs, delta, bary); break;
    case 2: d = tri_corner(m, t, 1, ob == 1, cs, delta, bary); break;
    case 3: d = tri_edge(m, t, 0, cs, delta, bary); break;
    case 4: d = tri_corner(m, t, 2, ob == 2, cs, delta, bary); break;
    case 5: d = tri_edge(m, t, 2, cs, delta, bary); break;
    case 6: d = tri_edge(m, t, 1, cs, delta, bary); break;
    case 7: {
        const double *sn = &m->snorm[3 * t];
        const double dNorm = dot3(sn, n0c);
        double csp[3], tmp[3];
        for (int q = 0; q < 3; q++) tmp[q] = dNorm * sn[q];
        for (int q = 0; q < 3; q++) csp[q] = cs[q] - tmp[q];
        d = tri_calc_dist(cs, csp, delta);
        break;
    }
    default: d = LARGE_TRIMESH; break;
    }
    return d - r;
}

/* resolveTriSphereNeighbuild (tri_mesh_I.h:275-307) */
static int tri_neighbuild(const mesh_t *m, int t, double r, const double *cs, double treshold)
{
    const double maxDist = r + treshold;
    double v[3];
    for (int q = 0; q < 3; q++) v[q] = cs[q] - m->center[3 * t + q];
    const double dNorm = fabs(dot3(&m->snorm[3 * t], v));
    if (dNorm > maxDist) return 0;
    const double dParaMax = maxDist * maxDist;
    for (int k = 0; k < 3; k++) {
        for (int q = 0; q < 3; q++) v[q] = cs[q] - MN(m, t, k)[q];
        const double d = dot3(EN(m, t, k), v);
        if (d > 0 && d * d > dParaMax) return 0;
    }
    return 1;
}

/* FixContactHistoryMesh::sort_contacts (fix_contact_history_mesh.cpp:
 * 375-402): move filled slots to the front, first empty <-> last filled */
static void mesh_slot_swap(mesh_t *m, int dnum, int i, int a, int b)
{
    int *pt = &m->partner[m->sstart[i]];
    double *h = &m->hist[(size_t)m->sstart[i] * dnum];
    const int t = pt[a]; pt[a] = pt[b]; pt[b] = t;
    for (int d = 0; d < dnum; d++) {
        const double x = h[a * dnum + d]; h[a * dnum + d] = h[b * dnum + d]; h[b * dnum + d] = x;
    }
}

static void mesh_sort_contacts(or_sim *s, mesh_t *m)
{
    if (!m->partner) return;
    for (int i = 0; i < s->n; i++) {
        const int nn = m->nneighs[i];
        int *pt = &m->partner[m->sstart[i]];
        int fe, lf;
        do {
            fe = lf = -1;
            for (int j = 0; j < nn; j++) {
                if (fe == -1 && pt[j] == -1) fe = j;
                if (pt[j] >= 0) lf = j;
            }
            if (fe > -1 && lf > -1 && fe < lf) mesh_slot_swap(m, s->dnum_w, i, fe, lf);
        } while (fe > -1 && lf > -1 && fe < lf);
    }
}

static int tri_has_candidate(const mesh_t *m, int t, int i)
{
    for (int q = m->tstart[t]; q < m->tstart[t + 1]; q++)
        if (m->tlist[q] == i) return 1;
    return 0;
}

/* FixNeighlistMesh::pre_force (fix_neighlist_mesh.cpp:230-300) and
 * FixContactHistoryMesh::pre_force (fix_contact_history_mesh.cpp:314-365)
 * with cleanUpContactJumps (:466-500) */
static void mesh_rebuild(or_sim *s, mesh_t *m)
{
    const int dnum = s->dnum_w;
    const double treshold = m->moving ? s->p.skin : 0.5 * s->p.skin;
    /* MultiNodeMesh::storeNodePosRebuild (multi_node_mesh_I.h:833) */
    if (m->moving) memcpy(m->node_re, m->node, sizeof(double) * 9 * m->ntri);
    free(m->tstart); free(m->tlist);
    m->tstart = malloc(sizeof(int) * (m->ntri + 1));
    int cap = 1024, cnt = 0;
    m->tlist = malloc(sizeof(int) * cap);
    int *nn = calloc(s->n, sizeof(int));
    for (int t = 0; t < m->ntri; t++) {
        m->tstart[t] = cnt;
        for (int i = 0; i < s->n; i++)
            if (tri_neighbuild(m, t, s->radius[i] * s->p.cdf, &s->x[3 * i], treshold)) {
                if (cnt == cap) { cap *= 2; m->tlist = realloc(m->tlist, sizeof(int) * cap); }
                m->tlist[cnt++] = i;
                nn[i]++;
            }
    }
    m->tstart[m->ntri] = cnt;
    /* carry the first npartner slots into slot arrays of the new size */
    int *ostart = m->sstart, *opart = m->partner, *onp = m->npartner;
    double *ohist = m->hist;
    if (opart) {
        for (int i = 0; i < s->n; i++) {
            int ip = 0;
            while (ip < onp[i]) {
                const int tri = opart[ostart[i] + ip];
                if (!tri_has_candidate(m, tri, i)) {
                    opart[ostart[i] + ip] = -1;
                    for (int d = 0; d < dnum; d++) ohist[(size_t)(ostart[i] + ip) * dnum + d] = 0.0;
                    /* swap(i, ip, npartner-1) on the old arrays */
                    const int last = onp[i] - 1;
                    const int tt = opart[ostart[i] + ip];
                    opart[ostart[i] + ip] = opart[ostart[i] + last];
                    opart[ostart[i] + last] = tt;
                    for (int d = 0; d < dnum; d++) {
                        const double x = ohist[(size_t)(ostart[i] + ip) * dnum + d];
                        ohist[(size_t)(ostart[i] + ip) * dnum + d] = ohist[(size_t)(ostart[i] + last) * dnum + d];
                        ohist[(size_t)(ostart[i] + last) * dnum + d] = x;
                    }
                    onp[i]--;
                } else {
                    ip++;
                }
            }
        }
    }
    int tot = 0;
    int *nstart = malloc(sizeof(int) * (s->n + 1));
    for (int i = 0; i < s->n; i++) { nstart[i] = tot; tot += nn[i]; }
    nstart[s->n] = tot;
    int *npart = malloc(sizeof(int) * (tot + 1));
    double *nhist = calloc((size_t)(tot + 1) * dnum, sizeof(double));
    for (int q = 0; q < tot; q++) npart[q] = -1;
    int *newnp = calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(int));
    for (int i = 0; i < s->n; i++) {
        if (!opart) continue;
        for (int ip = 0; ip < onp[i] && ip < nn[i]; ip++) {
            npart[nstart[i] + ip] = opart[ostart[i] + ip];
            for (int d = 0; d < dnum; d++)
                nhist[(size_t)(nstart[i] + ip) * dnum + d] = ohist[(size_t)(ostart[i] + ip) * dnum + d];
        }
        newnp[i] = onp[i];
    }
    free(ostart); free(opart); free(onp); free(ohist); free(m->keep); free(m->inter); free(m->nneighs);
    m->sstart = nstart;
    m->partner = npart;
    m->hist = nhist;
    m->npartner = newnp;
    m->nneighs = nn;
    m->keep = calloc(tot + 1, 1);
    m->inter = calloc(tot + 1, 1);
}

/* MultiNodeMesh::decideRebuild (multi_node_mesh_I.h:792-826): any node
 * moved more than half the skin since the last rebuild */
static int mesh_decide_rebuild(const or_sim *s, const mesh_t *m)
{
    if (!m->moving) return 0;
    const double triggersq = 0.25 * s->p.skin * s->p.skin;
    for (int q = 0; q < 3 * m->ntri; q++) {
        const double dx = m->node[3 * q] - m->node_re[3 * q];
        const double dy = m->node[3 * q + 1] - m->node_re[3 * q + 1];
        const double dz = m->node[3 * q + 2] - m->node_re[3 * q + 2];
        if (dx * dx + dy * dy + dz * dz > triggersq) return 1;
    }
    return 0;
}

static void mesh_move(or_sim *s, mesh_t *m)
{
    if (!m->moving) return;
    double dx[3];
    for (int q = 0; q < 3; q++) dx[q] = s->p.dt * m->vel[q];
    for (int t = 0; t < m->ntri; t++) {
        for (int k = 0; k < 3; k++)
            for (int q = 0; q < 3; q++) MN(m, t, k)[q] += dx[q];
        for (int q = 0; q < 3; q++) m->center[3 * t + q] += dx[q];
    }
    /* p_ref is a mesh property of frame_general: it moves with the nodes
     * (mesh_module_stress.cpp:71) */
    for (int q = 0; q < 3; q++) m->pref[q] += dx[q];
}

/* FixWallGran::post_force_mesh (fix_wall_gran.cpp:803-980) */
static void mesh_post_force(or_sim *s, mesh_t *m, int shearupdate)
{
    const int dnum = s->dnum_w;
    const double cdm = s->p.cdf - 1.0;
    double vwall[3];
    m->force[0] = m->force[1] = m->force[2] = 0.0;
    /* MeshModuleStress::pre_force (mesh_module_stress.cpp:288-300) */
    for (int q = 0; q < 3; q++) m->stress_f[q] = m->stress_t[q] = 0.0;
    /* markAllContacts */
    for (int q = 0; q < m->sstart[s->n]; q++) { m->keep[q] = 0; m->inter[q] = 0; }
    for (int t = 0; t < m->ntri; t++) {
        for (int qq = m->tstart[t]; qq < m->tstart[t + 1]; qq++) {
            const int i = m->tlist[qq];
            double delta[3], bary[3];
            int bs = -1;
            const double radi = s->radius[i];
            const double deltan = tri_contact(m, t, radi, &s->x[3 * i], delta, bary, &bs);
            if (deltan > s->cutneighmax) continue;
            const int intersect = deltan <= 0;
            if (!(deltan <= 0 || deltan < cdm * radi)) continue;
            /* FixContactHistoryMesh::handleContact */
            const int base = m->sstart[i], ns = m->nneighs[i];
            int *pt = &m->partner[base];
            int slot = -1;
            for (int k = 0; k < ns; k++)
                if (pt[k] == t) { slot = k; break; }
            if (slot >= 0) {
                m->keep[base + slot] = 1;
                m->inter[base + slot] = intersect;
            } else {
                const int face = (bs == 7);
                if (face) {
                    int already = 0;
                    for (int k = 0; k < ns; k++)
                        if (pt[k] >= 0 && pt[k] != t && coplanar_node_neighs(m, pt[k], t) && m->keep[base + k])
                            already = 1;
                    if (already) continue;
                }
                for (int k = 0; k < ns; k++)
                    if (pt[k] == -1) { slot = k; break; }
                if (slot < 0) continue;  /* the reference stops with an internal error */
                pt[slot] = t;
                m->keep[base + slot] = 1;
                m->inter[base + slot] = intersect;
                double *h = &m->hist[(size_t)(base + slot) * dnum];
                for (int d = 0; d < dnum; d++) h[d] = 0.0;
                m->npartner[i]++;
                if (face)
                    for (int k = 0; k < ns; k++)
                        if (pt[k] >= 0 && pt[k] != t && coplanar_node_neighs(m, pt[k], t))
                            for (int d = 0; d < dnum; d++) h[d] = m->hist[(size_t)(base + k) * dnum + d];
            }
            double *h = &m->hist[(size_t)(base + slot) * dnum];
            for (int q = 0; q < 3; q++)
                vwall[q] = m->moving ? (bary[0] * m->vel[q] + bary[1] * m->vel[q] + bary[2] * m->vel[q]) : 0.0;
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
            sd.jtype = m->wtype;
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
                    m->force[q] += Fi[q];
                }
                if (m->stress) {
                    /* MeshModuleStress::add_particle_contribution
                     * (mesh_module_stress.cpp:318-344), called with the force
                     * the contact added to the particle and -sidata.delta
                     * (fix_wall_gran_base.h:349-362): frc = -F at the contact
                     * point x + delta, torque (cp - p_ref) x frc */
                    const double frc[3] = {-Fi[0], -Fi[1], -Fi[2]};
                    double arm[3];
                    for (int q = 0; q < 3; q++) arm[q] = s->x[3 * i + q] + delta[q] - m->pref[q];
                    m->stress_f[0] += frc[0]; m->stress_f[1] += frc[1]; m->stress_f[2] += frc[2];
                    m->stress_t[0] += arm[1] * frc[2] - arm[2] * frc[1];
                    m->stress_t[1] += arm[2] * frc[0] - arm[0] * frc[2];
                    m->stress_t[2] += arm[0] * frc[1] - arm[1] * frc[0];
                }
                s->fscale[i] += sqrt(Fi[0] * Fi[0] + Fi[1] * Fi[1] + Fi[2] * Fi[2]);
                s->tscale[i] += sqrt(Ti[0] * Ti[0] + Ti[1] * Ti[1] + Ti[2] * Ti[2]);
            } else {
                chain_close(&s->mw, h, NULL);
            }
        }
    }
    /* cleanUpContacts */
    for (int i = 0; i < s->n; i++)
        for (int k = 0; k < m->nneighs[i]; k++) {
            const int q = m->sstart[i] + k;
            if (!m->keep[q]) {
                if (m->partner[q] > -1) m->npartner[i]--;
                m->partner[q] = -1;
                m->inter[q] = 0;
                for (int d = 0; d < dnum; d++) m->hist[(size_t)q * dnum + d] = 0.0;
            }
        }
}

double or_mesh_distance(const or_sim *s, int im, int tri, double r, const double x[3], double delta[3],
                        double bary[3], int *region)
{
    return tri_contact(&s->mesh[im], tri, r, x, delta, bary, region);
}

int or_mesh_contacts(const or_sim *s, int im, int *tag, int *tri, int *slot, double *hist)
{
    const mesh_t *m = &s->mesh[im];
    int c = 0;
    if (!m->partner) return 0;
    for (int i = 0; i < s->n; i++)
        for (int k = 0; k < m->nneighs[i]; k++) {
            const int q = m->sstart[i] + k;
            if (m->partner[q] < 0) continue;
            if (tag) {
                tag[c] = s->tag[i];
                tri[c] = m->partner[q];
                if (slot) slot[c] = k;
                for (int d = 0; d < s->dnum_w; d++) hist[(size_t)c * s->dnum_w + d] = m->hist[(size_t)q * s->dnum_w + d];
            }
            c++;
        }
    return c;
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
    if (s->wall_rebuild) {
        build_wall_lists(s);
        for (int k = 0; k < s->nmesh; k++) mesh_rebuild(s, &s->mesh[k]);
    } else {
        /* FixMesh::pre_force on a regular step (fix_mesh.cpp:553-575) */
        for (int k = 0; k < s->nmesh; k++)
            if (mesh_decide_rebuild(s, &s->mesh[k])) s->next_reneighbor = s->ntimestep + 1;
    }
    s->wall_rebuild = 0;
    pair_compute(s, shearupdate);
    for (int q = 0; q < s->nops; q++) {
        op_t *o = &s->ops[q];
        if (o->kind == OP_WALL) wall_plane(s, o->arg, shearupdate);
        else if (o->kind == OP_MESH) mesh_post_force(s, &s->mesh[o->arg], shearupdate);
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
    /* FixMoveMesh::initial_integrate -> MeshMoverLinear (mesh_mover_linear.cpp:94-112) */
    for (int k = 0; k < s->nmesh; k++) mesh_move(s, &s->mesh[k]);
    /* FixMove::initial_integrate, style linear (fix_move.cpp:430-490) */
    for (int q = 0; q < s->nops; q++) {
        op_t *o = &s->ops[q];
        if (o->kind != OP_MOVE) continue;
        const double delta = (s->ntimestep - o->time_origin) * s->p.dt;
        for (int i = 0; i < s->n; i++) {
            if (!(s->gmask[i] & (1 << o->bit))) continue;
            const double dtfm = dtf / s->rmass[i];
            for (int d = 0; d < 3; d++) {
                if (o->style[d]) {
                    s->v[3 * i + d] = o->val[d];
                    s->x[3 * i + d] = o->xorig[3 * i + d] + o->val[d] * delta;
                } else {
                    s->v[3 * i + d] += dtfm * s->f[3 * i + d];
                    s->x[3 * i + d] += dtv * s->v[3 * i + d];
                }
            }
        }
    }
    for (int i = 0; i < s->n; i++) {
        if (s->integ_bit >= 0 && !(s->gmask[i] & (1 << s->integ_bit))) continue;
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
    for (int q = 0; q < s->nops; q++) {
        op_t *o = &s->ops[q];
        if (o->kind != OP_MOVE) continue;
        for (int i = 0; i < s->n; i++) {
            if (!(s->gmask[i] & (1 << o->bit))) continue;
            const double dtfm = dtf / s->rmass[i];
            for (int d = 0; d < 3; d++)
                if (!o->style[d]) s->v[3 * i + d] += dtfm * s->f[3 * i + d];
        }
    }
    for (int i = 0; i < s->n; i++) {
        if (s->integ_bit >= 0 && !(s->gmask[i] & (1 << s->integ_bit))) continue;
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
    if (s->ntimestep == s->next_reneighbor) return 1;
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
                    if (hist)
                        for (int d = 0; d < dnum; d++)
                            hist[(size_t)c * dnum + d] = s->mp.newton[d] ? -h[d] : h[d];
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
