/*
 * liggghts_dem.h -- C-ABI of the MI355X DEM timestep engine.
 *
 * The engine owns one simulation per context: particle state in HBM, the
 * neighbour list with its pair-keyed contact history, primitive planes and
 * material tables, and one HIP stream of default priority.  Calls on one
 * context are not thread safe; they are serialised by the caller (the
 * LIGGGHTS main thread in the style plug-in).  Every entry point returns 0
 * on success or a negative DEM_ERR_* code; dem_last_error() gives the text.
 *
 * Each entry point names the reference interface whose work it takes over
 * (paths relative to the LIGGGHTS-INL src/ directory).
 */
#ifndef LIGGGHTS_DEM_H
#define LIGGGHTS_DEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DEM_OK = 0,
    DEM_ERR_ARG = -1,      /* bad argument or call order                    */
    DEM_ERR_HIP = -2,      /* HIP runtime failure                           */
    DEM_ERR_OVERFLOW = -3, /* "Neighbor list overflow, boost neigh_modify one"
                              (neigh_gran.cpp:553,633)                       */
    DEM_ERR_STATE = -4,    /* "ilegal situation, particle positions
                              identical" (pair_gran_base.h:361-362), NaN      */
    DEM_ERR_NOMEM = -5
};

enum { DEM_FP64 = 0, DEM_MIXED = 1 };
enum { DEM_ROLL_OFF = 0, DEM_ROLL_EPSD = 1, DEM_ROLL_EPSD2 = 2, DEM_ROLL_EPSD3 = 3, DEM_ROLL_CDTNONLINEAR2 = 4 };

#define DEM_MAXTYPES 8
#define DEM_MAXPLANES 8

typedef struct {
    double dt;                    /* timestep (update.cpp timestep command)  */
    double boxlo[3], boxhi[3];    /* orthogonal box (domain.cpp)             */
    int periodic[3];              /* boundary p / non-periodic per dim       */
    double skin;                  /* neighbor <skin> bin                     */
    double cdf;                   /* contactDistanceFactor (neighbor.h:141)  */
    int neigh_every, neigh_delay, neigh_check;  /* neigh_modify             */
    int neigh_one;                /* neigh_modify one (default 2000)         */
    int rolling_model;            /* DEM_ROLL_*  (rolling_friction keyword)  */
    int tangential_damping;       /* hertz setting, default 1                */
    int limit_force;              /* hertz setting, default 0                */
    int torsion_torque;           /* rolling setting, default 0              */
    int precision;                /* DEM_FP64 or DEM_MIXED (Hertz chains)    */
    int device;                   /* HIP device ordinal                      */
    int normal_model;             /* DEM_NORMAL_* of pair_style gran model   */
    int wall_normal_model;        /* DEM_NORMAL_* of the fix wall/gran model */
    int wall_limit_force;         /* limitForce of the wall model            */
    int cohesion_model;           /* DEM_COHESION_* of pair_style gran       */
} dem_params;

/* `cohesion` keyword of the pair chain: off, bond (cohesion_model_bond.h),
 * bond/nonlinear (cohesion_model_bond_nonlinear.h).  Particle pairs only. */
enum { DEM_COHESION_OFF = 0, DEM_COHESION_BOND = 1, DEM_COHESION_BOND_NONLINEAR = 2 };
/* bond tables, each (ntypes+1)^2 row-major, in this order:
 *  0 radiusMultiplierBond[nonlinear]  1 normalBondStiffnessPerUnitArea
 *  2 tangentialBondStiffnessPerUnitArea
 *  3-6 damping{Normal,Tangential}{Force,Torque}Bond[nonlinear]
 *  7 maxDistanceBond  8 maxSigmaBond  9 maxTauBond  10 ratioTensionCompression
 * 11 createDistanceBond  12-24 stiffnessPerUnitArea K_fn1 Ku_fn1 Kc_fn1 K_fn2
 *    Ku_fn2 Kc_fn2 K_ft K_tn Ku_tn Kc_tn K_tt Ku_tt Kc_tt (bond/nonlinear) */
#define DEM_NBOND 25
/* settings of the cohesion model (registerSettings) */
typedef struct {
    double ts_create;             /* tsCreateBond[nonlinear]                 */
    int stress_break;             /* stressBreak                             */
    int create_always;            /* createBondAlways                        */
    int damping;                  /* dampingBond (default 1)                 */
    int damping_smooth;           /* dampingBondSmooth                       */
    int ratio_tc;                 /* ratioTensionCompression[Bond]           */
    int tension_variant;          /* bond/nonlinear: normal law of the INL
                                     example header cohesion_model_bond_
                                     nonlinear_tension.h                     */
} dem_bond_settings;

/* `model` keyword: hertz (normal_model_hertz.h), hysteretic/nonlinear1
 * (normal_model_hysteretic_nonlinear1.h), hysteretic/nonlinear2 */
enum { DEM_NORMAL_HERTZ = 0, DEM_NORMAL_HYST1 = 1, DEM_NORMAL_HYST2 = 2 };
#define DEM_NHYST 11

typedef struct dem_ctx dem_ctx;

/* Creates a context.  Replaces the device-side state that Atom, Neighbor,
 * PairGran and FixWallGran hold for one run (atom.cpp, neighbor.cpp,
 * pair_gran.cpp:312-560).  Returns NULL on failure (dem_last_error(NULL)). */
dem_ctx *dem_ctx_create(const dem_params *p);
void dem_ctx_destroy(dem_ctx *ctx);
const char *dem_last_error(const dem_ctx *ctx);

/* Host helper: effective material tables as global_properties.cpp:428-562
 * builds them (createYeff, createGeff, createBetaEff).  Y, nu per type
 * (length ntypes+1, index 0 unused); e per type pair, (ntypes+1)^2
 * row-major.  Outputs (ntypes+1)^2 each. */
int dem_material_tables(int ntypes, const double *Y, const double *nu, const double *e,
                        double *Yeff, double *Geff, double *betaeff);

/* Material tables of the contact models, (ntypes+1)^2 row-major each:
 * Yeff, Geff, betaeff (normal_model_hertz.h:150-160), coeffFrict
 * (tangential_model_history.h:126-131), coeffRollFrict, coeffRollVisc
 * (rolling_model_epsd.h:86-90; NULL when off) and the scalar
 * coefficientRollingStiffness (rolling_model_epsd3.h:86-91). */
int dem_set_materials(dem_ctx *ctx, int ntypes, const double *Yeff, const double *Geff,
                      const double *betaeff, const double *mu, const double *mur,
                      const double *etar, double krs);

/* hysteretic/nonlinear1/2 parameters (normal_model_hysteretic_nonlinear1.h
 * :105-133): CoeffRestLog = log(coefficientRestitution)
 * (global_properties.cpp createCoeffRestLog) and DEM_NHYST tables, each
 * (ntypes+1)^2, in the order LoadingStiffness, UnloadingStiffness,
 * coefficientAdhesionStiffness, coefficientPlasticityDepth, pullOffForce,
 * alphaCustom, cinCustom, aoneCustom, atwoCustom, athreeCustom, kcinCustom.
 * Call after dem_set_materials. */
int dem_set_hysteretic(dem_ctx *ctx, const double *coeff_rest_log, const double *tables);

/* Particle state upload (AtomVecSphere arrays, atom_vec_sphere.cpp:125-160):
 * x, v, omega are n*3 row-major; radius, rmass per atom; type 1..ntypes;
 * tag > 0 unique; freeze != 0 marks members of fix freeze (may be NULL).
 * Replaces any previous state; contact history is cleared. */
int dem_upload_state(dem_ctx *ctx, int n, const double *x, const double *v,
                     const double *omega, const double *radius, const double *rmass,
                     const int *type, const int *tag, const int *freeze);

/* Download in the order of the last upload (NULL skips an array).
 * Replaces the host copies of x, v, omega, f, torque. */
int dem_download(dem_ctx *ctx, double *x, double *v, double *omega, double *f, double *torque);

/* fix wall/gran ... primitive type <wtype> {x,y,z}plane <pos>
 * (fix_wall_gran.cpp:93-360, primitive_wall_definitions.h:130-148). */
int dem_add_plane(dem_ctx *ctx, int dim, double pos, int wtype);
/* Triangle-mesh wall: fix mesh/surface file ... type <wtype> curvature
 * <cos> (fix_mesh.cpp:86-430, input_mesh_tri.cpp:305-603) + fix wall/gran
 * ... mesh (fix_wall_gran.cpp:803-980), and with vel != NULL fix
 * move/mesh linear vx vy vz (mesh_mover_linear.cpp:94-112).  nodes: ntri*9
 * doubles in file order; triangles with two equal nodes are skipped, the
 * remaining ones get ids 0.. in order.  curvature <= 0 takes the default
 * 1-1e-5.  Topology (active edges/corners) is built here on the host.
 * Returns the mesh index (>= 0) or a negative DEM_ERR_*.  Up to 4 meshes,
 * box without periodic dimensions. */
int dem_add_mesh(dem_ctx *ctx, int ntri, const double *nodes, int wtype, double curvature,
                 const double *vel);
/* Number of triangles the mesh kept. */
int dem_mesh_size(const dem_ctx *ctx, int mesh);
/* Host-only: the topology SurfaceMesh::buildNeighbours + handleCorner
 * (surface_mesh_I.h:474-560, 1139-1236) computes for a mesh in the given
 * box: per kept triangle edge_active[3], corner_active[3], obtuse index
 * (-1 none) and number of edge neighbours.  Output arrays may be NULL. */
int dem_mesh_topology(int ntri, const double *nodes, double curvature, const double *boxlo,
                      const double *boxhi, int *ntri_kept, int *edge_active, int *corner_active,
                      int *obtuse, int *nneigh);
/* Mesh contact history (FixContactHistoryMesh partner_/contacthistory_,
 * fix_contact_history_mesh.cpp): every filled slot as (atom tag, triangle
 * id, slot index, hist[wall dnum]).  NULL arrays: count only. */
int dem_get_mesh_contacts(dem_ctx *ctx, int mesh, int64_t *count, int *tag, int *tri, int *slot,
                          double *hist);
/* Restores mesh contact slots before dem_setup (the restart path of
 * FixContactHistoryMesh); the next rebuild carries them over. */
int dem_set_mesh_contacts(dem_ctx *ctx, int mesh, int64_t count, const int *tag, const int *tri,
                          const int *slot, const double *hist);
/* fix mesh/surface/stress on a mesh (MeshModuleStress, mesh_module_stress.cpp
 * :288-360): from the next force pass, the force the particles put on the
 * mesh and its torque about the reference point pref (NULL: the origin;
 * `reference_point`), which moves with the mesh.  Every particle counts
 * (the fix's group is all). */
int dem_mesh_stress(dem_ctx *ctx, int mesh, const double *pref);
/* FixMeshSurfaceStress's vector (mesh_module_stress.cpp:479-488): out[0..2]
 * total force, out[3..5] torque about p_ref, out[6..8] p_ref, of the last
 * force pass. */
int dem_mesh_stress_vector(dem_ctx *ctx, int mesh, double *out);
/* fix massflow/mesh ... mesh <face> vec_side vx vy vz count once|multiple
 * (fix_massflow_mesh.cpp:95-300, post_integrate :366-530): a planar face
 * mesh (nodes as for dem_add_mesh; it is no wall) through which particles
 * are counted when they pass to the vec_side side inside a triangle's face
 * region, once ever (once != 0) or at every crossing; group_bit restricts
 * the count to a group (-1: all).  One per context; no point_at_outlet or
 * delete_atoms. */
int dem_add_massflow(dem_ctx *ctx, int ntri, const double *nodes, const double *vec_side, int once,
                     int group_bit);
/* FixMassflowMesh::compute_vector (fix_massflow_mesh.cpp:637-662): out[0]
 * mass and out[1] number of particles counted so far, out[2] and out[3]
 * their rates over the time since the previous call. */
int dem_massflow_vector(dem_ctx *ctx, double *out);
/* Bond tables (DEM_NBOND) and settings of the cohesion model
 * (cohesion_model_bond.h:237-470, cohesion_model_bond_nonlinear.h:204-386).
 * Call after dem_upload_state and after registering the templates and
 * insertion fixes: registers the model's contact distance factor from the
 * tables and the smallest radius of the atoms and templates
 * (neighbor.h:141-142, modify_liggghts.cpp:515-567). */
int dem_set_bond(dem_ctx *ctx, const double *tables, const dem_bond_settings *bs);
/* fix gravity acceleration (fix_gravity.cpp:301-352, :395-397). */
int dem_set_gravity(dem_ctx *ctx, double ax, double ay, double az);
/* fix freeze present (fix_freeze.cpp:121-144; pair_gran.cpp:532-537). */
int dem_set_freeze(dem_ctx *ctx, int on);

/* Verlet::setup (verlet.cpp:134-199): pbc, neighbour build (with history
 * re-seed), one force pass with shearupdate = 0. */
int dem_setup(dem_ctx *ctx);
/* Verlet::run(n) (verlet.cpp:264-390) with fix nve/sphere
 * (fix_nve_sphere.cpp:134-243), the granular pair style, primitive walls
 * and gravity, all on the device. */
int dem_step(dem_ctx *ctx, int nsteps);
/* One force evaluation on the current list: force_clear + pair + walls +
 * gravity (+ freeze).  shearupdate as PairGran::compute sets it
 * (pair_gran.cpp:632-635).  For set-up passes and parity probes. */
int dem_compute_forces(dem_ctx *ctx, int shearupdate);
/* Neighbour rebuild now: FixContactHistory::pre_exchange
 * (fix_contact_history.cpp:305-426) + Domain::pbc + granular_bin_no_newton
 * (neigh_gran.cpp:485-640) + primitive-wall lists. */
int dem_rebuild(dem_ctx *ctx);

/* Binary restart of the device state (write_restart / read_restart,
 * write_restart.cpp:119-200, read_restart.cpp:106): a header (step, time
 * bookkeeping, history sizes), the atoms in upload order (tag, type, fix
 * freeze membership, group bits, x, v, omega, radius, rmass), each atom's
 * contact-history record in the layout of FixContactHistory::pack_restart
 * (fix_contact_history.cpp:588-602: record size, partner count, then per
 * partner its tag and the dnum history values, integers in the bits of a
 * double as LAMMPS' ubuf stores them), the primitive-wall history per atom
 * and the mesh contact slots.  dem_read_restart needs a context with the
 * same contact model, planes and meshes (added in their current positions);
 * it uploads the atoms, restores their groups and histories and the step,
 * and is followed by dem_setup.  fix move and massflow counters are not
 * carried (an upload clears them). */
int dem_write_restart(dem_ctx *ctx, const char *path);
int dem_read_restart(dem_ctx *ctx, const char *path);
/* Directed pair set of the current (full) list and its contact history:
 * entry (tag_i, tag_j, shift, hist[dnum], flag), shift packs periodic
 * image offsets as (sx+1)+3(sy+1)+9(sz+1), 13 = none.  This is the data
 * `compute pair/gran/local id history` reports (compute_pair_gran_local.cpp
 * :100-116).  Pass NULL arrays to get only the count in *count. */
int dem_get_pairs(dem_ctx *ctx, int64_t *count, int *tag_i, int *tag_j, int *shift,
                  double *hist, int *flag);
/* History doubles per pair entry: 12 hysteretic state (when that normal
 * model is on) + 3 shear + 3 rolling torque (when rolling is on). */
int dem_history_size(const dem_ctx *ctx);
/* History doubles per (atom, plane) of the wall model. */
int dem_wall_history_size(const dem_ctx *ctx);
/* Contact-history injection (FixContactHistory::unpack_restart,
 * fix_contact_history.cpp:588-656): per-atom partner data used by the next
 * rebuild's re-seed, keyed by (tag_i, tag_j). */
int dem_set_history(dem_ctx *ctx, int64_t count, const int *tag_i, const int *tag_j,
                    const double *hist);
/* Primitive-wall history per atom (property/atom history_<id>,
 * fix_wall_gran.cpp:476-505), n * nplanes * dnum, upload order. */
int dem_get_wall_history(dem_ctx *ctx, double *hist);
/* Restores it (the property/atom restart path, fix_property_atom.cpp
 * unpack_restart), same layout. */
int dem_set_wall_history(dem_ctx *ctx, const double *hist);

/* Kernel timing with HIP events on the context's stream. */
int dem_timing_enable(dem_ctx *ctx, int on);
/* name: "contact" | "integrate" | "rebuild".  Total ms and launch count
 * since enable. */
int dem_timing_get(dem_ctx *ctx, const char *name, double *ms, int64_t *launches);
/* Mean full-list neighbours and contacts per atom of the last force pass;
 * the contacts are counted only in passes with timing on (-1 otherwise). */
int dem_list_stats(dem_ctx *ctx, double *mean_neigh, double *mean_contacts, int64_t *entries);
int dem_sync(dem_ctx *ctx);
int64_t dem_timestep(const dem_ctx *ctx);

/* ---- Groups and per-group force fixes.  They run after every other
 * post_force contribution (pair, primitive and mesh walls, gravity, freeze),
 * in the order they were added, as in a script that defines them after its
 * walls; while any exists, dem_step does not use graph stepping. ---- */
/* Group bit 0..30: members[i] != 0 puts atom i (upload order) in the group
 * (Group::assign, group.cpp:119). */
int dem_set_group(dem_ctx *ctx, int bit, const int *members);
/* fix setforce with constant values: the group's force components with
 * style[d] != 0 are set to value[d], NULL components are left alone
 * (FixSetForce::post_force, fix_setforce.cpp:246-277).  *id names it for
 * dem_unfix and dem_force_op_vector. */
int dem_add_setforce(dem_ctx *ctx, int bit, const int *style, const double *value, int *id);
/* fix aveforce: each styled component becomes the group average plus the
 * given value (FixAveForce::post_force, fix_aveforce.cpp:222-281); the z
 * value is the equal-style expression c0 - (time - t0) * c1 evaluated at
 * the step's time (constant when c1 = 0). */
int dem_add_aveforce(dem_ctx *ctx, int bit, const int *style, double xvalue, double yvalue, double c0,
                     double t0, double c1, int *id);
/* fix nve/sphere on a group: only atoms with group bit `bit` are integrated
 * (FixNVESphere::initial_integrate/final_integrate skip atoms outside the
 * fix's group, fix_nve_sphere.cpp:140-150); -1 integrates every atom. */
int dem_set_integrate_group(dem_ctx *ctx, int bit);
/* fix move linear on group bit `bit` (FixMove, style linear,
 * fix_move.cpp:430-490, :900-940): components with style[d] != 0 follow
 * x = x_origin + value[d] * (step - time_origin) * dt at velocity value[d];
 * the others are integrated from the force as fix nve does.  x_origin and
 * time_origin are the positions and step at this call, made after
 * dem_upload_state (an upload clears group bits and with them this fix).
 * It acts before nve/sphere, as in scripts that define it first.  One per
 * context; *id names it for dem_unfix. */
int dem_add_move_linear(dem_ctx *ctx, int bit, const int *style, const double *value, int *id);
/* unfix (Modify::delete_fix, modify.cpp:985). */
int dem_unfix(dem_ctx *ctx, int id);
/* The fix's vector: the group's total force before the fix acted, at the
 * last force pass (FixSetForce::compute_vector, fix_setforce.cpp:348;
 * FixAveForce::compute_vector, fix_aveforce.cpp:346). */
int dem_force_op_vector(dem_ctx *ctx, int id, double *out3);
/* End of a run command: Update::update_time (update.cpp:511, called from
 * Verlet::cleanup, verlet.cpp:398).  dem_time is atime + (ntimestep -
 * atimestep) * dt, the time equal-style variables see. */
int dem_cleanup(dem_ctx *ctx);
double dem_time(const dem_ctx *ctx);

/* ---- Particle insertion: fix insert/pack with a particledistribution/
 * discrete of particletemplate/sphere templates, run inside dem_step at the
 * fix's insertion steps (FixInsert::pre_exchange, fix_insert.cpp:676-888),
 * before that step's neighbour build.  The placement loop and its random
 * streams are the reference's; each batch of candidate positions is tested
 * against the resident particles on the device, so their state never
 * leaves HBM.  A context may start empty (dem_upload_state with n = 0). ---- */
/* fix particletemplate/sphere <seed> atom_type T density constant rho
 * radius constant r (fix_template_sphere.cpp:86-240; constant radius and
 * density are the styles this fork parses). */
typedef struct {
    int atom_type;
    double radius;
    double density;
} dem_template_sphere;

enum { DEM_INSERT_VOLUMEFRACTION = 0, DEM_INSERT_PARTICLES = 1, DEM_INSERT_MASS = 2 };
enum { DEM_VEL_CONSTANT = 0, DEM_VEL_UNIFORM = 1, DEM_VEL_GAUSSIAN = 2 };

/* fix <id> <group> insert/pack seed S distributiontemplate D ... with its
 * region block (fix_insert.cpp:96-330, fix_insert_pack.cpp:72-200,
 * region_block.cpp) and fix D all particledistribution/discrete S2 n t1 w1 ...
 * (fix_particledistribution_discrete.cpp:68-233).  dem_insert_pack_defaults
 * fills the keyword defaults (FixInsert::init_defaults, fix_insert.cpp:446-490;
 * FixInsertPack::init_defaults, fix_insert_pack.cpp:252-266). */
typedef struct {
    int seed;                    /* insert/pack seed (prime > 10000)                  */
    int dist_seed;               /* particledistribution/discrete seed                */
    int ntemplates;
    const dem_template_sphere *templates;
    const double *weights;       /* per template: mass fraction, or number fraction
                                    with number_based (style .../numberbased)          */
    int number_based;
    double region_lo[3], region_hi[3];   /* region block, units box                  */
    int target;                  /* DEM_INSERT_*: volumefraction_region,
                                    particles_in_region, mass_in_region              */
    double target_value;
    int maxattempt;              /* default 50                                         */
    int insert_every;            /* 0: once (the default -1 is an error, as in
                                    FixInsertPack::calc_insertion_properties)          */
    long long start;             /* first insertion step; default: next step          */
    int overlapcheck;            /* default 1                                          */
    int all_in;                  /* default 0                                          */
    int check_dist_from_subdomain_border;   /* default 1                               */
    int ntry_mc;                 /* region volume samples, default 100000             */
    int exact_number;            /* random_distribute exact (1, default) or
                                    uncorrelated (0)                                   */
    int vel_style;               /* DEM_VEL_*                                          */
    double vel[3], vel_fluct[3];
    double omega[3];             /* omega constant                                     */
    int group_mask;              /* engine group bits the inserted atoms carry (the
                                    insert fix's group; ParticleToInsert::insert sets
                                    mask = 1 | groupbit, particleToInsert.cpp:113)    */
    const int *multisphere;      /* per template: the id dem_add_template_multisphere
                                    gave a particletemplate/multiplespheres, or 0 for
                                    the sphere of `templates` (NULL: all spheres)      */
    int orientation;             /* DEM_ORIENT_*, keyword orientation                 */
    double quat[4];              /* orientation constant q0 q1 q2 q3                  */
} dem_insert_pack_spec;

enum { DEM_ORIENT_TEMPLATE = 0, DEM_ORIENT_RANDOM = 1, DEM_ORIENT_CONSTANT = 2 };

void dem_insert_pack_defaults(dem_insert_pack_spec *spec);
/* Registers the fix: validates the spec, builds the distribution
 * (mass-to-number weights, large-to-small template order) and runs
 * FixInsert::setup's region volume estimate (FixInsertPack::
 * calc_insertion_properties, fix_insert_pack.cpp:281-322), which advances
 * the region's random stream as the reference does.  *id names it for
 * dem_unfix and dem_insert_vector. */
int dem_add_insert_pack(dem_ctx *ctx, const dem_insert_pack_spec *spec, int *id);
/* The fix's totals: out[0] particles inserted, out[1] mass inserted, out[2]
 * particles inserted by the last insertion, out[3] particles it aimed for,
 * out[4] attempts it made, out[5] the next insertion step (0: none). */
int dem_insert_vector(dem_ctx *ctx, int id, double *out6);
/* Appends atoms to a live context (ParticleToInsert::insert +
 * Atom::tag_extend, particleToInsert.cpp:97-150, atom.cpp:574-595): state as
 * for dem_upload_state, engine group bits per atom (NULL: none); the new
 * atoms follow the existing ones in download order.  The next dem_step step
 * rebuilds the neighbour list (dem_rebuild rebuilds now); every existing
 * contact history is carried.  Not in a batched run, an adjoint window or
 * with an observed plane. */
int dem_append_atoms(dem_ctx *ctx, int m, const double *x, const double *v, const double *omega,
                     const double *radius, const double *rmass, const int *type, const int *tag,
                     const int *gmask);
/* fix <id> all particletemplate/multiplespheres <seed> atom_type T density
 * constant rho nspheres N ntry M spheres[_different_types] file F scale s
 * [bonded yes] (fix_template_multiplespheres.cpp:78-300): x (3N) and r (N)
 * as the file lists them, types per sphere or NULL.  Registering runs the
 * template's set-up (post_create, :296-330): the bounding sphere and the
 * Monte Carlo centre of mass on the template's stream, the inside tests of
 * the ntry sample points on the device. */
typedef struct {
    int seed;
    int atom_type;
    double density;
    int nspheres;
    const double *x;
    const double *r;
    const int *types;
    double scale;
    int ntry;
    int bonded;
} dem_template_multisphere_spec;

int dem_add_template_multisphere(dem_ctx *ctx, const dem_template_multisphere_spec *spec, int *id);
/* out[0] bounding radius, out[1..3] bounding-sphere centre (centre-of-mass
 * frame), out[4] volume and out[5] mass expectancy, out[6..8] the centre of
 * mass in the file's frame. */
int dem_template_info(dem_ctx *ctx, int id, double *out9);
/* create_particles <template> single x y z [velocity v] [omega w]
 * [orientation q] (create_particles.cpp:100-330): one clump, its spheres at
 * x + rotate(x_j - x_bound, q) with rmass 4.18879020479 r^3 rho
 * (particleToInsert.cpp:129), tags continuing from the largest; nothing is
 * created when x lies outside the box.  quat NULL: identity (normalised as
 * the command does otherwise).  gmask: engine group bits of the new atoms.
 * Returns the number of spheres created, or a DEM_ERR_* code. */
int dem_create_particles_single(dem_ctx *ctx, int tmpl, const double *x, const double *v, const double *omega,
                                const double *quat, int gmask);
/* Atom count (atom->nlocal) and, in download order, each atom's tag, type,
 * radius and mass (NULL pointers skipped). */
int dem_natoms(const dem_ctx *ctx);
int dem_download_props(dem_ctx *ctx, int *tag, int *type, double *radius, double *rmass);

/* fix <id> <group> insert/stream seed S distributiontemplate D ...
 * insertion_face F extrude_length L | insert_every N [duration K]
 * (fix_insert.cpp:96-330, fix_insert_stream.cpp:76-151).  Bodies are placed
 * in the face extruded against the inflow direction: a point of the face
 * drawn on the face mesh's own stream (area-weighted triangle, barycentric
 * coordinates; MultiNodeMesh seeds it 179424799), pushed back along the
 * face normal by a random length.  Until its release step an inserted atom
 * moves with the velocity's normal component and no spin, forces and
 * torques zeroed (FixInsertStream::end_of_step, fix_insert_stream.cpp:
 * 828-921); at the release step it takes the insertion velocity and omega.
 * While any atom awaits release the integrator runs un-fused and steps run
 * without graphs.  dem_insert_stream_defaults fills the keyword defaults
 * (FixInsert::init_defaults, fix_insert.cpp:446-490; FixInsertStream::
 * init_defaults, fix_insert_stream.cpp:247-262). */
typedef struct {
    int seed, dist_seed;         /* as dem_insert_pack_spec                            */
    int ntemplates;
    const dem_template_sphere *templates;
    const double *weights;
    int number_based;
    const int *multisphere;
    int face_ntri;               /* the insertion face: a planar fix mesh/surface,   */
    const double *face_nodes;    /* 9 doubles per triangle, scaled and moved         */
    double extrude_length;       /* 0: not given                                       */
    int insert_every;            /* -1 (default): from extrude_length                 */
    int duration;                /* 0: insert_every                                    */
    long long start;             /* first insertion step; default: next step          */
    int nparticles;              /* bodies to insert; -1: INF; 0: from mass           */
    double mass;                 /* mass to insert when nparticles is 0; -1: INF      */
    double particlerate, massrate;   /* exactly one > 0                               */
    int maxattempt;              /* default 50                                         */
    int overlapcheck;            /* default 1                                          */
    int all_in;                  /* 1 needs fix mesh/surface/planar: refused here      */
    int parallel;                /* parallel yes: extrusion clipped to the sub-box    */
    int ntry_mc;                 /* insertion-fraction samples, default 100000        */
    int exact_number;            /* random_distribute exact (1) or uncorrelated (0)  */
    int vel_style;               /* DEM_VEL_*                                          */
    double vel[3], vel_fluct[3];
    double omega[3];
    int group_mask;              /* group bits of the inserted atoms; the release
                                    applies to atoms holding the first stream fix's  */
    int orientation;             /* DEM_ORIENT_*                                       */
    double quat[4];
} dem_insert_stream_spec;

void dem_insert_stream_defaults(dem_insert_stream_spec *spec);
/* Registers the fix and runs FixInsertStream::calc_insertion_properties
 * (fix_insert_stream.cpp:287-423): face normal flipped toward the velocity,
 * insert_every = (int)((L + 1e-14) / (dt |v_n|)) or L = duration dt |v_n|,
 * bodies per insertion = flow rate * insert_every * dt.  The timestep must
 * be set.  dem_insert_vector and dem_unfix take its id. */
int dem_add_insert_stream(dem_ctx *ctx, const dem_insert_stream_spec *spec, int *id);
/* The release data of atoms tag 1..ntags (fix property/atom
 * release_fix_insert_stream, fix_insert_stream.cpp:763-826): 14 doubles per
 * tag, x at insertion (3), insertion step, release step, integration
 * velocity (3), velocity and omega at release (3 + 3); zero for atoms no
 * stream created. */
int dem_stream_release_data(dem_ctx *ctx, int ntags, double *out14);

/* ---- Batched runs.  No counterpart in the reference, which runs one
 * simulation per process (one LAMMPS instance, lammps.cpp); these let many
 * independent simulations of the same box, walls and contact model share one
 * context, each kernel launch covering all of them. ---- */
/* Assigns every uploaded atom (upload order) to an instance 0..ninst-1.
 * Call after dem_upload_state and before dem_setup.  Particles of different
 * instances never interact; each instance may have its own materials. */
int dem_set_instances(dem_ctx *ctx, int ninst, const int *inst);
/* Yeff, Geff, betaeff, mu (and mur, etar, may be NULL) of one instance,
 * shapes as in dem_set_materials; the instance's other tables stay the
 * common ones. */
int dem_set_instance_materials(dem_ctx *ctx, int inst, const double *Yeff, const double *Geff,
                               const double *betaeff, const double *mu, const double *mur,
                               const double *etar);
/* Record, in every force pass, each atom's force from primitive plane
 * `plane` (index in dem_add_plane order) along the plane's axis; -1 off.
 * Call after dem_upload_state. */
int dem_observe_plane(dem_ctx *ctx, int plane);
/* Per instance b, out[5b..5b+4] = atom count, sum of z, sum of the
 * observed plane's force on the atoms (the wall's normal force; 0 if no
 * plane is observed), total mass, kinetic energy, at the current state and
 * last force pass. */
int dem_instance_observables(dem_ctx *ctx, double *out);
/* dem_instance_observables without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate (a calibration loop
 * polling every instance between steps, code that must not call hipMalloc).
 * The per-instance sums are staged in a fixed per-GPU device array of 4096
 * instances (160 KiB), held until the process ends; more instances are
 * rejected with an error.  Calls that overlap on one GPU (contexts on the
 * same device used from several host threads) race on that array. */
int dem_instance_observables_noalloc(dem_ctx *ctx, double *out);
/* Step with a recorded HIP graph of two rebuild periods when every rebuild
 * falls on a fixed step (neigh_modify check no, mesh walls only if none
 * moves, no bonds, no setforce/aveforce, timing off); other steps run as
 * usual.  The results are bitwise those of stepping without the graph. */
int dem_use_graph(dem_ctx *ctx, int on);
/* Autotuner: times the contact pass on the current state for every shape
 * (lane groups of 1..16 lanes per particle, contact-queue tiles of 8..64
 * particles per wave, given as -8..-64) and block size (64..256), in the
 * form the steps run (with the integrator when the pass integrates), and
 * keeps the fastest for this context.  State and history are left
 * unchanged. */
int dem_autotune(dem_ctx *ctx, int reps, int *lanes, int *block, double *ms);
/* Sets the contact pass's shape by hand (lanes per particle 1, 2, 4, 8 or
 * 16, or a contact-queue tile of -8, -16, -32 or -64; block 64, 128 or 256;
 * 0 keeps the default, the -32 queue), e.g. to reuse an earlier autotuner
 * result.  Hertz chains take any shape; others run 4 lanes. */
int dem_set_contact_shape(dem_ctx *ctx, int lanes, int block);

/* ---- Adjoint window.  No counterpart in the reference (which has no
 * derivatives); gradients of per-instance observables with respect to the
 * contact parameters, for calibration.  FP64 Hertz pairs and planes with
 * tangential history only. ---- */
/* Right after dem_setup: runs nsteps with the set-up neighbour list kept
 * (where a rebuild would fall, the contact history is carried over as the
 * rebuild does; stepping stops with an error if any particle moves more
 * than half the skin), keeping a checkpoint every checkpoint_every steps. */
int dem_adjoint_forward(dem_ctx *ctx, int nsteps, int checkpoint_every);
/* Reverse sweep over that window, segment by segment from the checkpoints.
 * J = sum_b w_height[b] * (mean z of instance b) + w_wall[b] * (force of the
 * observed plane on instance b), at the window's last step (either weight
 * array may be NULL).  grad[4b..4b+3] = dJ/d(Yeff, Geff, betaeff, mu) of
 * instance b, summed over its type pairs.  The context is left at the
 * window's final state. */
int dem_adjoint_backward(dem_ctx *ctx, const double *w_height, const double *w_wall, double *grad);

#ifdef __cplusplus
}
#endif
#endif
