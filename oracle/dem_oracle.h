/*
 * dem_oracle.h -- CPU restatement of the LIGGGHTS-INL sphere DEM timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never links or calls it.
 *
 * Everything is IEEE FP64 and single threaded, in the operation order of the
 * reference C++ sources cited per function in dem_oracle.c.  Compiled with
 * -O2 -ffp-contract=off so no fused multiply-add changes a rounding.
 *
 * Indexing: atoms are kept in the order they were given (the reference sorts
 * atoms spatially every 1000 steps, atom.cpp:98; that only changes summation
 * order, so results are reported per tag).  Material tables are
 * (ntypes+1)x(ntypes+1) row-major with row/column 0 unused, like the
 * reference's double** matrices indexed [itype][jtype].
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXTYPES 8
#define OR_MAXPLANES 8

/* rolling_friction keyword of pair_style gran */
enum { OR_ROLL_OFF = 0, OR_ROLL_EPSD = 1, OR_ROLL_EPSD2 = 2, OR_ROLL_EPSD3 = 3, OR_ROLL_CDT2 = 4 };

typedef struct {
    double dt;
    double boxlo[3], boxhi[3];
    int periodic[3];
    double skin;                 /* neighbor <skin> bin                      */
    double cdf;                  /* neighbor->contactDistanceFactor (1.0)    */
    int neigh_every, neigh_delay, neigh_check;   /* neigh_modify            */
    int rolling_model;           /* OR_ROLL_*                               */
    int tangential_damping;      /* hertz setting, reference default on     */
    int limit_force;             /* model setting: hertz default off,
                                    hysteretic default on                   */
    int torsion_torque;          /* rolling setting, default off            */
    int normal_model;            /* OR_NORMAL_* of pair_style gran          */
    int wall_normal_model;       /* OR_NORMAL_* of the fix wall/gran model  */
    int wall_limit_force;        /* limitForce of the wall model            */
    int cohesion_model;          /* OR_COH_* of pair_style gran (pairs only) */
} or_params;

/* cohesion keyword of pair_style gran */
enum { OR_COH_OFF = 0, OR_COH_BOND = 1, OR_COH_BOND_NONLINEAR = 2 };

/* bond tables, each (ntypes+1)^2, in this order (a table a model does not
 * use is ignored):
 *  0 radiusMultiplierBond[nonlinear]   1 normalBondStiffnessPerUnitArea
 *  2 tangentialBondStiffnessPerUnitArea
 *  3-6 damping{Normal,Tangential}{Force,Torque}Bond[nonlinear]
 *  7 maxDistanceBond  8 maxSigmaBond  9 maxTauBond
 * 10 ratioTensionCompression  11 createDistanceBond
 * 12-24 stiffnessPerUnitArea K_fn1 Ku_fn1 Kc_fn1 K_fn2 Ku_fn2 Kc_fn2 K_ft
 *       K_tn Ku_tn Kc_tn K_tt Ku_tt Kc_tt  (bond/nonlinear) */
#define OR_NBOND 25
/* bond settings (registerSettings of the cohesion model) */
typedef struct {
    double ts_create;        /* tsCreateBond[nonlinear]                      */
    int stress_break;        /* stressBreak                                  */
    int create_always;       /* createBondAlways                             */
    int damping;             /* dampingBond (default on)                     */
    int damping_smooth;      /* dampingBondSmooth                            */
    int ratio_tc;            /* ratioTensionCompression[Bond]                */
    int tension_variant;     /* bond/nonlinear: the normal law of the INL
                                example header cohesion_model_bond_nonlinear_
                                tension.h instead of the in-tree one         */
} or_bond_settings;

/* model keyword of pair_style gran / fix wall/gran */
enum { OR_NORMAL_HERTZ = 0, OR_NORMAL_HYST1 = 1, OR_NORMAL_HYST2 = 2 };

/* hysteretic/nonlinear* tables, in this order, each (ntypes+1)^2:
 * LoadingStiffness, UnloadingStiffness, coefficientAdhesionStiffness,
 * coefficientPlasticityDepth, pullOffForce, alphaCustom, cinCustom,
 * aoneCustom, atwoCustom, athreeCustom, kcinCustom */
#define OR_NHYST 11

typedef struct or_sim or_sim;

or_sim *or_create(const or_params *p);
void or_destroy(or_sim *s);

/* Effective material tables exactly as global_properties.cpp builds them.
 * Y, nu: per type (index 1..ntypes, length ntypes+1).  e, mu, mur, etar:
 * per type pair, (ntypes+1)^2 row-major.  krs: coefficientRollingStiffness. */
int or_set_materials(or_sim *s, int ntypes, const double *Y, const double *nu,
                     const double *e, const double *mu, const double *mur,
                     const double *etar, double krs);
/* Read back the derived tables (Yeff, Geff, betaeff), (ntypes+1)^2 each. */
void or_get_material_tables(const or_sim *s, double *Yeff, double *Geff, double *betaeff);

/* atoms: x,v,omega are n*3 row-major; freeze marks members of fix freeze. */
int or_set_atoms(or_sim *s, int n, const double *x, const double *v, const double *omega,
                 const double *radius, const double *rmass, const int *type,
                 const int *tag, const int *freeze);

/* post_force fixes run in the order they are added, like Modify's fix list
 * (modify.cpp:1673-1750).
 * fix wall/gran ... primitive type <wtype> {x,y,z}plane <pos> */
int or_add_plane(or_sim *s, int dim, double pos, int wtype);
/* fix gravity: acceleration vector xacc,yacc,zacc (fix_gravity.cpp:395-397).
 * Returns the fix id. */
int or_add_gravity(or_sim *s, double ax, double ay, double az);
/* fix freeze on the atoms flagged in or_set_atoms (fix_freeze.cpp:121-144) */
void or_add_freeze(or_sim *s);
/* fix mesh/surface (+ fix move/mesh linear when vel != NULL) and
 * fix wall/gran ... mesh: triangle wall, nodes ntri*9 in file order
 * (element ids 0..ntri-1), curvature = cos(curvature angle), <= 0 for the
 * default 1-1e-5.  Must be added after or_set_atoms.  Returns the fix id. */
int or_add_mesh(or_sim *s, int ntri, const double *nodes, int wtype, double curvature, const double *vel);
/* number of triangles kept (elements with two equal nodes are refused) */
int or_mesh_size(const or_sim *s, int mesh);
/* topology as SurfaceMesh computes it: edge/corner active flags (ntri*3),
 * obtuse angle index (-1 none) and number of edge neighbours per element */
void or_mesh_topology(const or_sim *s, int mesh, int *edge_active, int *corner_active, int *obtuse,
                      int *nneigh);
/* TriMesh::resolveTriSphereContactBary of one triangle: returns the signed
 * gap (distance - r), delta = closest point - centre, barycentric weights and
 * the region code (7 face, 3/5/6 edge, 1/2/4 corner); LARGE (1e6) - r when
 * the region's edge or corner is inactive */
double or_mesh_distance(const or_sim *s, int mesh, int tri, double r, const double x[3], double delta[3],
                        double bary[3], int *region);
/* sum of the mesh's contact forces on the particles in the last pass */
void or_mesh_force(const or_sim *s, int mesh, double f[3]);
/* fix mesh/surface/stress on a mesh: force and torque sums about p_ref */
void or_mesh_stress(or_sim *s, int mesh, int on, const double pref[3]);
void or_mesh_stress_vector(const or_sim *s, int mesh, double out[9]);
/* current mesh contacts: (atom tag, triangle id, slot index, history);
 * NULL to count */
int or_mesh_contacts(const or_sim *s, int mesh, int *tag, int *tri, int *slot, double *hist);
/* cohesion bond / bond/nonlinear tables (OR_NBOND) and settings; registers
 * the model's contact distance factor (neighbor.h:141-142) from the tables
 * and the smallest atom radius, so call after or_set_atoms.  Returns 0. */
int or_set_bond(or_sim *s, const double *tables, const or_bond_settings *bs);
/* fix move linear on the atoms of group bit (fix_move.cpp:430-490):
 * x = x_original + v * (ntimestep - origin) * dt each step; components with
 * set[d] == 0 are integrated like nve.  Returns the fix id. */
int or_add_move_linear(or_sim *s, int bit, const int set[3], const double v[3]);
/* fix nve/sphere only on the atoms of group bit (default: all atoms) */
void or_set_integrate_group(or_sim *s, int bit);
/* hysteretic/nonlinear* parameter tables, see OR_NHYST */
int or_set_hysteretic(or_sim *s, const double *tables);
/* group membership: bit 'bit' of atom i's mask (group command) */
void or_set_group(or_sim *s, int bit, const int *members);
/* fix setforce <group> with constant values; style 0 = NULL (untouched)
 * (fix_setforce.cpp:246-277).  Returns the fix id. */
int or_add_setforce(or_sim *s, int bit, const int style[3], const double value[3]);
/* fix aveforce <group> x y z (fix_aveforce.cpp:222-281); x, y constant, z an
 * equal-style variable of the form c0 - (time - t0)*c1.  Returns the id. */
int or_add_aveforce(or_sim *s, int bit, const int style[3], double xv, double yv,
                    double c0, double t0, double c1);
/* unfix <id> (modify.cpp delete_fix): removes it from the post_force order */
void or_unfix(or_sim *s, int id);
/* fix aveforce compute_vector: group force before it was replaced */
void or_aveforce_vector(const or_sim *s, int id, double out[3]);
/* simulation time (the `time` thermo keyword, update.cpp) */
double or_time(const or_sim *s);

/* Verlet::setup (verlet.cpp:134-199): pbc, build list, force pass with
 * shearupdate = 0. */
void or_setup(or_sim *s);
/* Verlet::run (verlet.cpp:264-390) for nsteps. */
void or_run(or_sim *s, int nsteps);
/* end of a run command: Update::update_time (the `time` bookkeeping) */
void or_cleanup(or_sim *s);
/* One force evaluation on the current list (force_clear + pair + walls +
 * gravity), shearupdate as given.  No list rebuild. */
void or_compute_forces(or_sim *s, int shearupdate);
/* Force a neighbour rebuild now (pre_exchange + pbc + build + wall lists). */
void or_rebuild(or_sim *s);

void or_get_state(const or_sim *s, double *x, double *v, double *omega,
                  double *f, double *torque);
long long or_timestep(const or_sim *s);
int or_dnum(const or_sim *s);
int or_dnum_wall(const or_sim *s);

/* Directed pair set of the current list: every half-list entry (i,j) with j
 * an owned atom yields (tag_i,tag_j,13) and (tag_j,tag_i,13); an entry with a
 * periodic image of j yields (tag_i,tag_j,shift).  shift packs the image
 * offsets (sx,sy,sz) in {-1,0,1} as (sx+1) + 3(sy+1) + 9(sz+1) (13 = none).
 * hist gets dnum doubles per directed entry (the reversed entry of an owned
 * pair carries the negated newton slots), flag the contact flag.
 * Returns the count; pass NULL arrays to query the count. */
int or_get_pairs(const or_sim *s, int *tag_i, int *tag_j, int *shift,
                 double *hist, int *flag);

/* Per-atom primitive-wall history, n * nplanes * dnum_wall. */
void or_get_wall_history(const or_sim *s, double *hist);

/* Per-atom per-contact contributions of the last force pass, for scaling
 * tolerances: sum over contributions of |dF| and |dT| per atom (n each). */
void or_get_force_scale(const or_sim *s, double *fscale, double *tscale);

#ifdef __cplusplus
}
#endif
#endif
