/*
 * insert_oracle.h -- CPU restatement of LIGGGHTS-INL's fix insert/pack and
 * fix insert/stream with a particledistribution/discrete of
 * particletemplate/sphere and particletemplate/multiplespheres templates.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the engine's insertion.  Only
 * tests/ load it.  The placement loop is the reference's, serial and plain;
 * every overlap test is brute force over all resident particles (their
 * periodic images included) and all particles placed so far, where the
 * reference bins them (RegionNeighborList) and the engine tests batches on
 * the device.  Cited per function in insert_oracle.c.
 */
#ifndef INSERT_ORACLE_H
#define INSERT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OI_MAXT 32

typedef struct {
    int seed, dist_seed;
    int ntemplates, number_based;
    int type[OI_MAXT];
    double radius[OI_MAXT], density[OI_MAXT], weight[OI_MAXT];
    double region_lo[3], region_hi[3];
    int target;                 /* 0 volumefraction_region, 1 particles_in_region, 2 mass_in_region */
    double target_value;
    int maxattempt, insert_every;
    long long start;            /* <= 0: the step after creation */
    int overlapcheck, all_in, check_border, ntry_mc, exact_number;
    int vel_style;              /* 0 constant, 1 uniform, 2 gaussian */
    double vel[3], vel_fluct[3], omega[3];
    double boxlo[3], boxhi[3];
    int periodic[3];
    void *ms[OI_MAXT];          /* per template: an oi_tmpl (particletemplate/multiplespheres) or NULL */
    int orientation;            /* 0 template, 1 random, 2 constant */
    double quat[4];
    /* fix insert/stream (stream = 1; the region and target fields are then
     * unused): the insertion face, 9 doubles per triangle, the extrusion,
     * the totals and rates, the timestep.  insert_every -1: from
     * extrude_length; nparticles -1: INF, 0: from mass; mass -1: INF. */
    int stream;
    int face_ntri;
    const double *face;
    double extrude_length;
    int duration;
    int nparticles;
    double mass, particlerate, massrate;
    int parallel;
    double dt;
} oi_spec;

typedef struct oi_fix oi_fix;

/* The fix at timestep ntimestep; NULL on a bad spec (oi_error tells). */
oi_fix *oi_create(const oi_spec *s, long long ntimestep);
const char *oi_error(void);
void oi_destroy(oi_fix *f);
long long oi_next_step(const oi_fix *f);
/* FixInsert::pre_exchange at `step` against n resident particles; writes
 * the created particles (at most cap) and returns their number, -1 on an
 * error, 0 when the step is no insertion step. */
int oi_insert(oi_fix *f, long long step, int n, const double *x, const double *r, const double *m, int cap,
              double *x_out, double *v_out, double *omega_out, double *r_out, double *m_out, int *type_out,
              double *brid_out);
/* the created atoms (a clump's spheres in order) and, per atom, the clump's
 * bond_random_id (-1 for spheres and unbonded clumps).
 * ninserted, massinserted, last inserted, last aimed, last tries, next step */
void oi_vector(const oi_fix *f, double out[6]);
/* region volume estimate of the fix */
double oi_region_volume(const oi_fix *f);
/* fix insert/stream: radius of atoms or templates outside this fix that
 * Modify::get_max_radius would also see (0: none) */
void oi_set_maxrad(double r);
/* fix insert/stream: the release data of the atoms the last insertion
 * created, 14 doubles each (x_ins, i_step, r_step, v_integrate, v_release,
 * omega_release); returns their number */
int oi_release(const oi_fix *f, int cap, double *out14);
/* fix insert/stream: insert_every, ninsert_per, insertion fraction, face
 * normal (3), normal velocity (3) */
void oi_stream_info(const oi_fix *f, double out[9]);
/* RanPark from `seed`: the integer state after each of n draws */
void oi_park_states(int seed, int n, int *states, double *values);

/* fix particletemplate/multiplespheres (no coarse graining): spheres as in
 * the file, scaled; bounding sphere and Monte Carlo centre of mass on the
 * template's stream, every inside test brute force. */
typedef struct oi_tmpl oi_tmpl;
oi_tmpl *oi_tmpl_create(int seed, int nspheres, const double *x, const double *r, const int *types, int atom_type,
                        double density, double scale, int ntry, int bonded);
void oi_tmpl_destroy(oi_tmpl *t);
/* r_bound, x_bound[3] (centre-of-mass frame), volume, (unused), xcm[3] */
void oi_tmpl_info(const oi_tmpl *t, double out[9]);
/* create_particles single: sphere positions of a clump at x, orientation
 * quat (normalised as the command does; NULL = identity) */
void oi_tmpl_place(const oi_tmpl *t, const double x[3], const double *quat, double *out);

#ifdef __cplusplus
}
#endif
#endif
