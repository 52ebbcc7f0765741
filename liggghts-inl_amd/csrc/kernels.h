// kernels.h -- host-callable launchers of the engine's HIP kernels.
#pragma once
#include "dem_device.h"

namespace dem {

// Everything the force pass reads and writes (one particle per thread).
struct ForceArgs {
    const double4 *xr, *vm, *wt;
    const int *nstart;
    uint32_t *nj;
    double *hist;
    double4 *f4, *t4;
    const double *mat;
    const uint8_t *wmask;
    double *whist;
    int *ncontact;       // optional per-atom contact count (stats), may be null
    int *err;            // device error flags
    int n;
    int shearupdate;
    int defer_freeze;    // mesh walls follow: they apply fix freeze
    const double *bout;  // cohesion force/torque per entry (6), when bonded
    const uint8_t *bvalid;
    double *pwf;         // optional per-atom force of plane obs_plane along its axis
    int obs_plane;
    int lanes, block;    // launch shape (0: defaults)
    // fuse = 2 or 3: the pass also runs final_integrate (and, with 3, the
    // next step's initial_integrate) of each particle, writing the new x,
    // v, omega to xn/vn/wn (never the arrays it reads)
    int fuse;
    double4 *xn, *vn, *wn;
    double dtf, dtv, dtfrotate;
    // mixed law: the partners' velocity/mass and spin/meta as one 32-byte
    // FP32 record per particle (2 float4), or null to read vm/wt
    const float4 *rec;
    // a checked graph period: nonzero *halt (the step whose distance check
    // asked for a rebuild) makes the pass return at once; null: always run
    const long long *halt;
    Model M;
    Planes P;
};

// Cohesion bond pass (one thread per particle, canonical entries).
struct BondArgs {
    const double4 *xr, *vm, *wt;
    const int *tag;
    const int *nstart;
    uint32_t *nj;
    double *hist;
    const int *rev;
    double *bout;
    uint8_t *bvalid;
    const double *mat;
    // bond_random_id per atom, or null (FixPropertyAtom bond_random_id,
    // default -1: bonds of rule B form only between spheres of one clump at
    // the step it was inserted)
    const double *brid;
    int n, shearupdate;
    long long ntimestep;
    // in a checked graph period: the step counter on the device (read
    // instead of ntimestep) and the halt word, as ForceArgs::halt
    const long long *step_dev;
    const long long *halt;
    Model M;
};

struct BinArgs {
    double4 *xr;         // wrapped in place (Domain::pbc)
    const double4 *wt;   // instance in meta
    int4 *img;
    uint32_t *key;
    int *count;          // per Morton key
    int n;
    Grid G;
    double prd[3];
    double hi[3];
};

struct PermuteArgs {
    const int *sorted;   // new -> old
    const double4 *xr0, *vm0, *wt0;
    const int *tag0, *orig0;
    const int4 *img0;
    const double *whist0;
    double4 *xr1, *vm1, *wt1, *xhold;
    int *tag1, *orig1;
    int4 *img1;
    double *whist1;
    const int *gm0;      // group bits
    int *gm1;
    const double4 *xo0;  // fix move origins (may be null)
    double4 *xo1;
    const int *mf0;      // fix massflow/mesh side counters (may be null)
    int *mf1;
    const double *br0;   // bond_random_id per atom (may be null)
    double *br1;
    const double4 *rl0;  // fix insert/stream release data, 4 per atom (may be null)
    double4 *rl1;
    uint8_t *wmask;
    int n;
    int wdn;             // nplanes * dnum
    Planes P;
    double skin;
};

// entries per particle in the staging rows of the single-search rebuild;
// a longer row sends the rebuild to the count + fill passes
constexpr int kStageSlots = 32;

struct NeighArgs {
    const double4 *xr;   // new order
    const double4 *wt;   // new order (instance in meta)
    const int *tag;      // new order
    const int *cstart;   // per key, exclusive prefix + total at nkeys
    int *count;          // per atom (count pass)
    int *maxcount;
    const int *nstart;   // per atom (fill pass)
    uint32_t *nj;
    double *hist;
    // previous list, old index space
    const int *perm;     // new -> old
    const int *onstart;
    const uint32_t *onj;
    const double *ohist;
    const int *otag;
    int have_old;
    int n;
    int dnum;
    double cdf, skin;
    double rmax;         // largest radius (the cell grid's)
    double cw[3];        // cell widths, 1 / G.inv
    Grid G;
    double prd[3];
    int64_t ecap;        // fill pass: entries allocated
    int coh_off;         // first cohesion history slot, -1 without a cohesion model
    int *err;            // bit 3: list longer than ecap
    int2 *stage;         // n * kStageSlots staged entries (list word, carried entry or -1)
    int fill_if_over;    // fill pass: run only if the longest row exceeds kStageSlots
};

struct IntegArgs {
    double4 *xr, *vm, *wt;
    const double4 *f4, *t4;
    int n;
    double dtf, dtv, dtfrotate;
    // fix nve/sphere on a group (integ_bit >= 0) and fix move linear on a
    // group (mv_bit >= 0); both need the group bits
    const int *gm;
    int integ_bit;
    int mv_bit, mv_style;        // mv_style bit d: component d prescribed
    double mv_val[3];
    double mv_delta;             // (step - time_origin) * dt of the initial half
    const double4 *xo;           // positions at the fix's creation
    const long long *halt;       // as ForceArgs::halt
    // a checked graph period with the integrator in its own pass (mode 3):
    // ctl as in launch_check_halt; the pass runs unless the period halted
    // before this step, then checks the next step's positions (the next
    // step's decision) and moves up to kFoldMeshes meshes by their step
    long long *ctl;
    int par;
    const double4 *xhold;
    double deltasq;
    int nmove;
    TriGeo *mtri[2];
    int mntri[2];
    double mdx[2][3];
};
constexpr int kFoldMeshes = 2;

// Mesh wall neighbour build: candidate triangles per particle through a
// uniform grid of triangles (built on the host in the mesh's initial frame;
// disp = how far the mesh has moved since).
struct MeshNbArgs {
    const double4 *xr;
    const TriGeo *tri;
    const double *enorm;     // 9 per triangle
    const int *gstart;       // per grid cell, CSR into glist
    const int *glist;        // triangle ids, ascending per cell
    double glo[3], ginv[3];
    int gn[3];
    double disp[3];
    double cdf, threshold;
    int n;
    int *count;              // count pass
    const int *mstart;       // fill pass
    int *mcand;
    int *stage;              // n * kMeshStage candidates of the single-search pass
    int fill_if_over;        // fill pass: search only the rows longer than kMeshStage
};
constexpr int kMeshStage = 64;

// Carry each particle's contact slots from the previous build (old index
// space, perm = new -> old) into the new rows: sort_contacts then
// cleanUpContactJumps, then the first slots are copied.
struct MeshRemapArgs {
    const int *perm;
    const int *ostart;
    int *opart;
    double *ohist;
    int *onp;
    int have_old;
    const int *mstart, *mcand;
    int *mpart;
    double *mhist;
    int *mnp;
    int n, dnum;
};

struct MeshForceArgs {
    const double4 *xr, *vm, *wt;
    double4 *f4, *t4;
    const TriGeo *tri;
    const int *mstart, *mcand;
    int *mpart;
    double *mhist;
    int *mnp;
    const double *mat;
    int *err;                // bit 2: more contacts than slots
    int n, shearupdate;
    int wtype, moving;
    double vel[3];
    double curvature, precision;
    int apply_freeze;        // this pass ends the post_force chain
    // fuse = 2: final_integrate(n) of the particle after its last force
    // term; fuse = 3 also initial_integrate(n+1); the new state goes to
    // xn/vn/wn (the other state buffers)
    int fuse;
    double4 *xn, *vn, *wn;
    double dtf, dtv, dtfrotate;
    // fix mesh/surface/stress: per particle, the force its contacts put on
    // the mesh and that force's torque about pref (6 doubles), or null
    double *mstress;
    double pref[3];
    const long long *halt;   // as ForceArgs::halt
    Model M;
};

void launch_mesh_count(const MeshNbArgs &a, hipStream_t s);
void launch_mesh_fill(const MeshNbArgs &a, hipStream_t s);
// one search: counts and the first kMeshStage candidates of each row staged
void launch_mesh_stage(const MeshNbArgs &a, hipStream_t s);
// staged rows into mcand after the scan; rows longer than kMeshStage are left
// to the fill pass (fill_if_over)
void launch_mesh_compact(const MeshNbArgs &a, hipStream_t s);
void launch_mesh_remap(const MeshRemapArgs &a, hipStream_t s);
void launch_mesh_force(const MeshForceArgs &a, hipStream_t s);
// MeshMoverLinear: node += dx, center += dx for ntri triangles
void launch_mesh_move(TriGeo *tri, int ntri, double dx, double dy, double dz, hipStream_t s,
                      const long long *halt = nullptr);

// Adjoint window (adjoint.hip).
struct AdjMaskArgs {
    const double4 *xr;
    const int *nstart;
    uint32_t *nj;
    double *hist;
    int n, dnum;
    double cdf;
    double prd[3];
};

struct AdjArgs {
    // the step's force-pass inputs, from the tape
    const double4 *xr, *vm, *wt;
    const uint32_t *nj;
    const double *hist, *whist;
    const int *nstart, *rev;
    const uint8_t *wmask;
    const double *mat;
    // adjoints after the step (in) and before it (out)
    const double4 *lx, *lv, *lw;
    double4 *lx1, *lv1, *lw1;
    const double *lh, *lhw;
    double *lh1, *lhw1;
    double *gp;              // per particle: Yeff, Geff, betaeff, mu, summed over steps
    const double *wseed;     // per instance: adjoint of the observed plane's force, or null
    int obs_plane;
    int n, dnum, wdnum, update, mask_step;
    double ckick;            // lambda_f = ckick / m * lambda_v
    double dt;
    Model M;
    Planes P;
};

void launch_adj_mask(const AdjMaskArgs &a, hipStream_t s);
void launch_adj_rev(const int *nstart, const uint32_t *nj, int *rev, int n, hipStream_t s);
void launch_adj_step(const AdjArgs &a, hipStream_t s);
void launch_adj_seed(const double4 *wt, const double *wh, const double *count, double4 *lx, int n, hipStream_t s);
void launch_adj_reduce(const int *cstart, int nkl, int ninst, const double *gp, double *out, hipStream_t s);

void launch_bond(const BondArgs &a, hipStream_t s);
void launch_bond_rev(const int *nstart, const uint32_t *nj, int *rev, int n, hipStream_t s);

void launch_bin(const BinArgs &a, hipStream_t s);
void launch_scan(int *data, int n, int *tmp, hipStream_t s);   // exclusive, data[n] = total
void launch_scatter(const uint32_t *key, int *cursor, int *sorted, int n, hipStream_t s);
void launch_cellsort(const int *cstart, int *sorted, int nkeys, hipStream_t s);
void launch_permute(const PermuteArgs &a, hipStream_t s);
void launch_neigh_count(const NeighArgs &a, hipStream_t s);
// *maxcount = max(*maxcount, max of count[0..n))
void launch_row_max(const int *count, int n, int *maxcount, hipStream_t s);
void launch_neigh_fill(const NeighArgs &a, hipStream_t s);

// fix setforce / fix aveforce on a group (FixSetForce::post_force,
// fix_setforce.cpp:246-277; FixAveForce::post_force, fix_aveforce.cpp:
// 222-281): the group's force before the op and its count are summed in a
// fixed order (per-block partial sums, then an ordered total into total[4]);
// setforce sets the styled components to val, aveforce to the group average
// plus val.
constexpr int kGroupBlocks = 256;
struct GroupForceArgs {
    double4 *f4;
    const int *gmask;
    int n, bit;
    int aveforce;
    int style[3];        // 1: the component is set, 0: NULL
    double val[3];
    double *part;        // kGroupBlocks * 4
    double *total;       // 4: force sum and count of the group before the op
};
void launch_group_force(const GroupForceArgs &a, hipStream_t s);
// one search: counts (then the row max) and staging rows
void launch_neigh_stage(const NeighArgs &a, hipStream_t s);
// staging rows into the list (after the scan); no-op when a row outgrew kStageSlots
void launch_neigh_compact(const NeighArgs &a, hipStream_t s);
void launch_fill_int(int *p, int v, int n, hipStream_t s);
// err |= bit and all rows emptied when nstart[n] > ecap (flag: one int of scratch)
void launch_cap_guard(int *nstart, int n, int64_t ecap, int *flag, int *err, int bit, hipStream_t s);
void launch_row_max_check(const int *maxcount, int limit, int *err, int bit, hipStream_t s);
void launch_copy_int(int *dst, const int *src, int n, hipStream_t s);
// out[b*5 + {0..4}] = count, sum z, sum of pwf, sum mass, kinetic energy of
// instance b (particles in key order)
void launch_inst_obs(const int *cstart, int nkl, int ninst, const double4 *xr, const double4 *vm,
                     const double *pwf, double *out, hipStream_t s);
void launch_force_fp64(const ForceArgs &a, hipStream_t s);
// whether the contact pass of these arguments can integrate (every form
// without a cohesion model)
bool force_can_fuse(const ForceArgs &a);
void launch_force_mixed(const ForceArgs &a, hipStream_t s);
// mode: 1 = initial_integrate, 2 = final_integrate, 3 = final then initial
void launch_integrate(const IntegArgs &a, int mode, hipStream_t s);
void launch_pack_rec(const double4 *vm, const double4 *wt, float4 *rec, int n, hipStream_t s);
// fix massflow/mesh on a planar face mesh (FixMassflowMesh::post_integrate,
// fix_massflow_mesh.cpp:366-530): per particle in the mesh's candidate rows
struct MassflowArgs {
    const double4 *xr, *vm;
    const TriGeo *tri;
    const int *mstart, *mcand;
    int *counter;        // per particle: -1 not yet seen, 0 / 1 side, 2 counted once
    const int *gmask;    // group bits, null: every particle
    int gbit;
    int n, once;
    double pref[3], nvec[3];
    double precision;
    double *part;        // kMassflowBlocks * 2 partial sums (mass, count)
    double *total;       // running totals (mass, count)
};
constexpr int kMassflowBlocks = 256;
void launch_massflow(const MassflowArgs &a, hipStream_t s);
// out[0..5] = sum over particles of v[6 i + k], a fixed-order reduction
// (part: kSumBlocks * 6 doubles of partial sums)
constexpr int kSumBlocks = 256;
void launch_sum6(const double *v, int n, double *part, double *out, hipStream_t s);
int scan_tmp_size(int n);
// Neighbor::check_distance (neighbor.cpp:1390-1465): flag |= 1 if any atom
// moved more than sqrt(deltasq) since the last build.
// A checked graph period keeps the current step in ctl[par] (par alternates
// over the recorded steps) and in ctl[2] the step whose distance check asked
// for a rebuild (0: none yet).  k_check_halt, unless halted, writes the next
// step to ctl[par ^ 1] and records the step when an atom moved more than
// sqrt(deltasq).
void launch_check_halt(const double4 *xr, const double4 *xhold, int n, double deltasq, long long *ctl, int par,
                       hipStream_t s);
void launch_check_distance(const double4 *xr, const double4 *xhold, int n, double deltasq, int *flag,
                           hipStream_t s);

// Particle insertion (insert.hip): a uniform grid over the box, cells at
// least as wide as the largest overlap distance; periodic dimensions span
// the box exactly (n cells of prd / n).
struct InsGrid {
    double lo[3], inv[3], prd[3];
    int n[3], periodic[3];
};
// Which resident particles a fix insert/stream tests its candidates against
// (FixInsertStream::is_nearby, fix_insert_stream.cpp:568-593, and the
// bounding box of FixInsert::load_xnear, fix_insert.cpp:980-1018): normal
// distance d = (x - p_ref).n within [dmin, dmax], x + d n inside a face
// triangle, x inside [bblo, bbhi].  on = 0: every particle.
struct StreamNear {
    int on;
    double pref[3], nrm[3];
    double dmin, dmax;
    double bblo[3], bbhi[3];
    const double *face;     // 9 per triangle (device)
    int ntri;
};
// cell[i] = -1 (not binned) for a particle the filter rejects; with a
// periodic dimension the filter runs in the check instead, on the image
void launch_ins_bin(const double4 *xr, int n, const InsGrid &g, const StreamNear &f, int *cell, int *count,
                    hipStream_t s);
void launch_ins_place(const int *cell, int n, int *cursor, int *idx, hipStream_t s);
// hit[k] bit t: candidate k with radius rad[t] overlaps a binned particle
void launch_ins_check(const double *cand, int ncand, const double *rad, int nrad, const double4 *xr,
                      const int *start, const int *idx, const InsGrid &g, const StreamNear &f, uint32_t *hit,
                      hipStream_t s);
// fix insert/stream's end_of_step on every atom holding a bit of `mask`
// (every atom when mask = 0).  rel: 4 per atom, (x_ins, i_step),
// (v_integrate, r_step), (v_release, 0), (omega_release, 0); i_step 0 marks
// an atom no stream created, atoms past their release step are left alone.
struct ReleaseArgs {
    double4 *xr, *vm, *wt, *f4, *t4;
    const int *gm;
    const double4 *rel;
    int n, mask;
    long long step;
    double dt;
};
void launch_stream_release(const ReleaseArgs &a, hipStream_t s);
constexpr int kRegionBlocks = 256;
// part[4b..4b+2]: count, sum 4/3 pi r^3, sum mass of particles in [lo, hi]
void launch_region_sums(const double4 *xr, const double4 *vm, int n, const double *lo, const double *hi,
                        double *part, hipStream_t s);
// inside[k] = 1 when point k lies inside one of the nsph spheres (xs, rs)
void launch_tmpl_inside(const double *pts, int npts, const double *xs, const double *rs, int nsph, uint8_t *inside,
                        hipStream_t s);
// start[from + 1 .. from + count] = start[from]
void launch_extend_rows(int *start, int from, int count, hipStream_t s);

}  // namespace dem
