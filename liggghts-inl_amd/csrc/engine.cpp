// engine.cpp -- host side of the MI355X DEM engine: context, device memory,
// the Verlet step sequence and the C-ABI declared in include/liggghts_dem.h.
//
// Step sequence, Verlet::run (verlet.cpp:264-390):
//   initial_integrate -> [decide -> rebuild] -> force pass -> final_integrate
// final_integrate(n) and initial_integrate(n+1) are one kernel (mode 3);
// no force is computed between them, so both read the same f.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/liggghts_dem.h"
#include "insert_pack.h"
#include "kernels.h"
#include "mesh_host.h"

using namespace dem;

namespace {

std::string g_last_error;

struct EventPair {
    hipEvent_t a, b;
};

struct Timer {
    std::vector<EventPair> ev;
    std::vector<EventPair> pool;
    double ms = 0.0;
    int64_t launches = 0;
};

// One mesh wall (fix mesh/surface + fix wall/gran mesh [+ fix move/mesh
// linear]).  The host keeps the geometry and advances its copy of the nodes
// with the same additions as the device, so the re-neighbour trigger of a
// moving mesh needs no device read-back.
struct MeshDev {
    MeshHost h;
    int wtype = 1;
    bool moving = false;
    double vel[3] = {0, 0, 0};
    TriGeo *tri = nullptr;
    double *enorm = nullptr;
    int *gstart = nullptr, *glist = nullptr;
    MeshGrid grid;
    double disp[3] = {0, 0, 0};      // displacement since the grid was built
    std::vector<double> node_re;     // nodes at the last rebuild
    // per-particle rows (candidates and contact slots), double buffered
    int *mstart[2] = {}, *mcand[2] = {}, *mpart[2] = {}, *mnp[2] = {};
    double *mhist[2] = {};
    int ncap[2] = {0, 0};
    int64_t scap[2] = {0, 0};
    int cur = 0;
    bool have = false;
    // fix mesh/surface/stress: per-particle contributions, partial sums,
    // the total force and torque of the last force pass; p_ref moves with
    // the mesh
    bool stress = false;
    double pref[3] = {0, 0, 0};
    // a face mesh of fix massflow/mesh: neighbour rows only, no contacts
    bool passive = false;
    double *d_ms = nullptr, *d_spart = nullptr, *d_sout = nullptr;
    int ms_cap = 0;
};

}  // namespace

struct dem_ctx {
    dem_params p;
    hipStream_t stream = nullptr;
    std::string err;

    int ntypes = 0;
    double *d_mat = nullptr;
    bool mat_set = false;
    std::vector<double> hmat;          // host copy of the material tables (one block)
    int mat_blocks = 1;                // blocks allocated in d_mat

    // batched run: instances, their material overrides, observed plane
    int ninst = 1;
    std::vector<double> inst_mat;      // per instance: slots M_YEFF..M_ETAR
    std::vector<uint8_t> inst_set;
    int obs_plane = -1;
    double *pwf = nullptr;
    double *d_obs = nullptr;
    int obs_cap = 0;
    double *obs_scratch = nullptr;     // this device's copy of `scratch` (below), or null

    // launch shape of the contact pass (0: defaults), set by dem_autotune
    int tune_lanes = 0, tune_block = 0;

    // graph-captured stepping: one graph covers two rebuild periods, so the
    // double buffers are back where they started at its end
    bool use_graph = false;
    bool gexec_capturing = false;
    hipGraphExec_t gexec = nullptr;
    int g_period = 0;

    // adjoint window: fixed list, checkpoints, tape of force-pass inputs
    long long setup_ts = -1;           // timestep of the last dem_setup
    bool window = false;               // rebuild steps apply the history mask instead
    int64_t adj_E = 0;                 // list entries of the window
    int adj_steps = 0, adj_every = 0;
    std::vector<char *> ckpts;         // one per segment start, then the final state
    std::vector<long long> ck_ts;
    std::vector<int> ck_ago;
    char *tape = nullptr;
    int tape_cap = 0, tape_pos = -1;   // tape_pos >= 0: recording
    std::vector<uint8_t> tape_mask;
    int *adj_rev = nullptr;
    std::vector<const void *> g_key;   // buffers the graph was recorded with

    int n = 0, cap = 0;
    int cur = 0;
    double4 *xr[2] = {}, *vm[2] = {}, *wt[2] = {};
    double4 *f4 = nullptr, *t4 = nullptr, *xhold = nullptr;
    int *tag[2] = {}, *orig[2] = {};
    int *gmask[2] = {};      // group bits of each particle (fix setforce / aveforce)
    // x, v, omega of a contact pass that integrates go to the other buffer:
    // par = 0, the state is in xr/vm/wt[cur]; par = 1, in xn/vn/wn
    double4 *xn = nullptr, *vn = nullptr, *wn = nullptr;
    int par = 0;
    float4 *rec = nullptr;   // FP32 partner records of the mixed law (2 float4 per particle)
    int rec_cap = 0;
    int4 *img[2] = {};
    double *whist[2] = {};
    uint8_t *wmask = nullptr;
    int *ncontact = nullptr;
    bool nc_fresh = false;             // the last force pass counted contacts (timing on)
    double maxrad = 0.0;

    // counting sort
    uint32_t *key = nullptr;
    int *sorted = nullptr;
    int *cstart = nullptr, *cursor = nullptr;
    int nkeys_cap = 0;
    int *scan_tmp = nullptr;
    int scan_tmp_cap = 0;
    int2 *stage = nullptr;   // n * kStageSlots: staging rows of the single-search rebuild
    int *mstage = nullptr;   // n * kMeshStage: staged mesh candidate rows (shared by the meshes)
    int mstage_cap = 0;
    uint32_t *spread = nullptr;
    int spread_cap = 0;

    // neighbour list, double buffered for the history remap
    int lcur = 0;
    int *nstart[2] = {};
    uint32_t *nj[2] = {};
    double *hist[2] = {};
    int64_t ecap[2] = {0, 0};
    bool have_list = false;
    bool list_valid = false;
    int *d_small = nullptr;   // [0] err flag, [1] maxcount, [2] move flag

    Grid G{};
    Model M{};
    Planes P{};
    long long ntimestep = 0;
    // Update::atime / atimestep (update.cpp): the time at the end of the last run
    double atime = 0.0;
    long long atimestep = 0;
    // fix setforce / aveforce after the other post_force contributions, in
    // the order they were added
    struct ForceOp {
        int id, kind, bit, slot;   // kind 0 setforce, 1 aveforce
        int style[3];
        double val[3], c0, t0, c1;   // aveforce z: c0 - (time - t0) * c1
    };
    std::vector<ForceOp> fops;
    int next_op_id = 1;
    // fix nve/sphere on a group (-1: all atoms) and fix move linear
    int integ_bit = -1;
    struct MoveOp {
        int id, bit, style;
        double val[3];
        long long t0;        // time_origin (fix_move.cpp: update->ntimestep at creation)
    };
    std::vector<MoveOp> moves;
    double4 *xo[2] = {};     // positions at the fix's creation, per particle
    double *brid[2] = {};    // bond_random_id per particle (null: none set)
    // fix insert/stream release data, 4 per particle (null: no stream fix);
    // rel_until: the last release step of any atom, rel_mask: the group bits
    // of the first stream fix (FixInsertStream::i_am_integrator)
    double4 *rel[2] = {};
    long long rel_until = -1;
    int rel_mask = 0;
    // checked graph periods (neigh_modify check yes): ctl[0], ctl[1] the
    // current step, alternately; ctl[2] the step whose distance check asked
    // for a rebuild; while such a period is being recorded halt_dev = ctl + 2
    // and step_dev = the recorded step's counter
    long long *ctl = nullptr;
    long long ctl_h[3] = {0, 0, 0};
    const long long *halt_dev = nullptr;
    const long long *step_dev = nullptr;
    hipGraphExec_t cgexec = nullptr;
    std::vector<const void *> cg_key;
    int cg_period = 0;
    int cg_flip = 0;    // 1: each recorded step swaps the state buffers (fused integrator)
    // the integrator's own pass also checks the next step and moves the
    // meshes (fold_par: the recorded step's counter); cg_fold: the recorded
    // period does so, and its first check runs before the first replay
    bool fold = false;
    int fold_par = 0;
    bool cg_fold = false;
    double *d_gpart = nullptr, *d_gtot = nullptr;   // partial sums; 4 per op slot
    int ago = 0;

    bool timing = false;
    Timer t_contact, t_integrate, t_rebuild;

    // cohesion bond pass: per entry force/torque, validity, reverse entry
    double *bout = nullptr;
    uint8_t *bvalid = nullptr;
    int *brev = nullptr;
    int64_t bcap = 0;

    std::vector<MeshDev> meshes;
    // fix massflow/mesh (one per context): its face mesh, side counters
    // per particle, running totals on the device; the rate bookkeeping of
    // FixMassflowMesh::compute_vector on the host
    struct Massflow {
        int mesh = -1, once = 1, gbit = -1;
        double pref[3] = {0, 0, 0}, nvec[3] = {0, 0, 0};
        int *counter[2] = {};
        double *part = nullptr, *total = nullptr;
        double t_since = 0.0, delta_t = 0.0, mass_last = 0.0, n_last = 0.0, mass_prev = 0.0, n_prev = 0.0;
        bool stepped = false;
    } mf;
    bool mesh_grids_valid = false;    // built for the current radii
    long long next_reneighbor = -1;   // FixMesh::next_reneighbor

    // fix insert/pack (insert_pack.cpp): the fixes in the order they were added,
    // the largest tag (Atom::tag_extend continues from it), and the buffers
    // of the device overlap test (grid over the resident particles, built
    // once per insertion) and of the region sums
    struct Insert {
        int id;
        int gmask;                     // group bits of the atoms it creates
        std::unique_ptr<dem_ins::InsertPack> fix;   // insert/pack or insert/stream
        double *face_dev = nullptr;    // insert/stream: the face triangles, 9 each
        ~Insert()
        {
            if (face_dev) (void)hipFree(face_dev);
        }
    };
    std::vector<std::unique_ptr<Insert>> inserts;
    // fix particletemplate/multiplespheres templates (create_particles)
    struct Template {
        int id;
        dem_ins::TemplateMultisphere t;
    };
    std::vector<std::unique_ptr<Template>> templates;
    int tag_max = 0;
    int *ins_cell = nullptr, *ins_start = nullptr, *ins_idx = nullptr, *ins_tmp = nullptr;
    int ins_ncap = 0, ins_ccap = 0, ins_tcap = 0;
    double *ins_cand = nullptr, *ins_rad = nullptr, *ins_part = nullptr;
    uint32_t *ins_hit = nullptr;
    int ins_kcap = 0;
    bool ins_grid_ready = false;
    InsGrid ins_grid{};
};

static int massflow_step(dem_ctx *ctx);
static int massflow_counters(dem_ctx *ctx);

#define HIPCHK(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ctx->err = std::string("HIP error ") + hipGetErrorString(e_) + " at " \
                       + #call;                                                   \
            return DEM_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

#define KCHK()                                                                    \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);    \
            return DEM_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

static int fail(dem_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    else g_last_error = msg;
    return code;
}

// Scratch of dem_instance_observables_noalloc: the per-instance sums of
// k_inst_obs (5 doubles per instance) staged in a fixed device array, so the
// entry point calls no allocator.  Every GPU holds its own copy until the
// process ends; calls on one GPU share it, and each call writes every value
// it reads back, so nothing needs clearing.  In host code the name is a
// host-side stand-in: kernels and copies get the device address that
// hipGetSymbolAddress returns with the device current, resolved once per
// device when a context is created.
constexpr int kObsScratchInst = 4096;
static __device__ __align__(256) double scratch[kObsScratchInst * 5];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices];

// the device's scratch address (the device is current), null if unresolved
static double *resolve_scratch(int dev)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_dev[dev] = (double *)ptr;
        else hipGetLastError();   // leave no sticky error for the next call
    }
    return g_scratch_dev[dev];
}

// Host calls that read or write device data with the null stream first wait
// for the context's own stream: it is non-blocking, so the null stream does
// not order against it.
static void sync_enter(dem_ctx *ctx)
{
    hipSetDevice(ctx->p.device);
    hipStreamSynchronize(ctx->stream);
}

template <typename T>
static int dev_alloc(dem_ctx *ctx, T **ptr, size_t count)
{
    if (*ptr) {
        HIPCHK(hipFree(*ptr));
        *ptr = nullptr;
    }
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void **)ptr, count * sizeof(T)));
    return DEM_OK;
}

// the current x/v/omega buffers
static double4 *sx(const dem_ctx *ctx) { return ctx->par ? ctx->xn : ctx->xr[ctx->cur]; }
static double4 *sv(const dem_ctx *ctx) { return ctx->par ? ctx->vn : ctx->vm[ctx->cur]; }
static double4 *sw(const dem_ctx *ctx) { return ctx->par ? ctx->wn : ctx->wt[ctx->cur]; }

static void timer_begin(dem_ctx *ctx, Timer &t)
{
    if (!ctx->timing) return;
    EventPair e;
    if (!t.pool.empty()) {
        e = t.pool.back();
        t.pool.pop_back();
    } else {
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
    }
    hipEventRecord(e.a, ctx->stream);
    t.ev.push_back(e);
}

static void timer_end(dem_ctx *ctx, Timer &t)
{
    if (!ctx->timing || t.ev.empty()) return;
    hipEventRecord(t.ev.back().b, ctx->stream);
    t.launches++;
}

static void timer_collect(Timer &t, bool release = false)
{
    for (auto &e : t.ev) {
        float ms = 0.f;
        hipEventSynchronize(e.b);
        hipEventElapsedTime(&ms, e.a, e.b);
        t.ms += ms;
        t.pool.push_back(e);
    }
    t.ev.clear();
    if (release) {
        for (auto &e : t.pool) {
            hipEventDestroy(e.a);
            hipEventDestroy(e.b);
        }
        t.pool.clear();
    }
}

// ---------------------------------------------------------------------------

extern "C" const char *dem_last_error(const dem_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

// History slots in the reference's registration order (contact_models.h:
// 150-157): hysteretic normal state (12, normal_model_hysteretic_nonlinear1.h
// :84-97), shear (3, tangential_model_history.h:79-81), rolling torque (3,
// rolling_model_epsd.h:71-73).
static Layout chain_layout(int normal, int limitf, int rolling, int cohesion)
{
    Layout L;
    int k = 0;
    L.normal = normal;
    L.limitf = limitf;
    L.off_norm = -1;
    if (normal != NM_HERTZ) {
        L.off_norm = k;
        k += 12;
    }
    L.off_shear = k;
    k += 3;
    L.off_roll = -1;
    if (rolling >= 1 && rolling <= 3) {   // epsd*; cdtnonlinear2 keeps no history
        L.off_roll = k;
        k += 3;
    }
    // cohesion slots last on the device (the reference registers them after
    // the normal model; dem_get_pairs / dem_set_history convert)
    L.off_coh = -1;
    if (cohesion != CM_OFF) {
        L.off_coh = k;
        k += cohesion == CM_BOND ? 14 : 28;
    }
    L.dnum = k;
    return L;
}

extern "C" dem_ctx *dem_ctx_create(const dem_params *p)
{
    if (!p) {
        g_last_error = "dem_ctx_create: null params";
        return nullptr;
    }
    if (!(p->dt > 0.0) || !(p->skin >= 0.0) || !(p->cdf >= 1.0) || p->neigh_every < 1
        || p->rolling_model < 0 || p->rolling_model > 4 || p->precision < 0 || p->precision > 1
        || p->normal_model < 0 || p->normal_model > 2 || p->wall_normal_model < 0 || p->wall_normal_model > 2
        || p->cohesion_model < 0 || p->cohesion_model > 2) {
        g_last_error = "dem_ctx_create: invalid parameters";
        return nullptr;
    }
    for (int d = 0; d < 3; d++)
        if (!(p->boxhi[d] > p->boxlo[d])) {
            g_last_error = "dem_ctx_create: empty box";
            return nullptr;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0) {
        g_last_error = "dem_ctx_create: no HIP device";
        return nullptr;
    }
    if (hipSetDevice(p->device) != hipSuccess) {
        g_last_error = "dem_ctx_create: hipSetDevice failed";
        return nullptr;
    }
    dem_ctx *ctx = new dem_ctx();
    ctx->p = *p;
    if (ctx->p.neigh_one <= 0) ctx->p.neigh_one = 2000;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_last_error = "dem_ctx_create: stream creation failed";
        delete ctx;
        return nullptr;
    }
    if (hipMalloc((void **)&ctx->d_mat, sizeof(double) * M_COUNT * kT1 * kT1) != hipSuccess
        || hipMalloc((void **)&ctx->d_small, sizeof(int) * 8) != hipSuccess) {
        g_last_error = "dem_ctx_create: out of device memory";
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return nullptr;
    }
    hipMemset(ctx->d_small, 0, sizeof(int) * 8);
    ctx->obs_scratch = resolve_scratch(p->device);
    ctx->hmat.assign(kMatStride, 0.0);
    Model &M = ctx->M;
    M.dt = p->dt;
    M.cdf = p->cdf;
    M.skin = p->skin;
    for (int d = 0; d < 3; d++) M.prd[d] = p->boxhi[d] - p->boxlo[d];
    M.rolling = p->rolling_model;
    M.tdamp = p->tangential_damping;
    M.torsion = p->torsion_torque;
    M.lp = chain_layout(p->normal_model, p->limit_force, p->rolling_model, p->cohesion_model);
    M.lw = chain_layout(p->wall_normal_model, p->wall_limit_force, p->rolling_model, CM_OFF);
    M.cohesion = p->cohesion_model;
    M.bs = BondSet{0, 0, 0, 1, 0, 0, 0};
    M.dnum = M.lp.dnum;
    M.wdnum = M.lw.dnum;
    ctx->P.n = 0;
    return ctx;
}

static void free_all(dem_ctx *ctx)
{
    auto F = [](void *q) {
        if (q) hipFree(q);
    };
    for (int b = 0; b < 2; b++) {
        F(ctx->xr[b]); F(ctx->vm[b]); F(ctx->wt[b]); F(ctx->tag[b]); F(ctx->orig[b]);
        if (b == 0) { F(ctx->xn); F(ctx->vn); F(ctx->wn); }
        F(ctx->img[b]); F(ctx->whist[b]); F(ctx->nstart[b]); F(ctx->nj[b]); F(ctx->hist[b]);
        F(ctx->gmask[b]);
        F(ctx->xo[b]);
        F(ctx->brid[b]);
        F(ctx->rel[b]);
        F(ctx->mf.counter[b]);

    }
    F(ctx->f4); F(ctx->t4); F(ctx->rec); F(ctx->xhold); F(ctx->wmask); F(ctx->ncontact); F(ctx->key);
    F(ctx->sorted); F(ctx->cstart); F(ctx->cursor); F(ctx->scan_tmp); F(ctx->spread); F(ctx->stage); F(ctx->mstage);
    F(ctx->d_mat); F(ctx->d_small); F(ctx->mf.part); F(ctx->mf.total); F(ctx->ctl);
    F(ctx->bout); F(ctx->bvalid); F(ctx->brev);
    F(ctx->pwf); F(ctx->d_obs); F(ctx->d_gtot); F(ctx->d_gpart);
    F(ctx->ins_cell); F(ctx->ins_start); F(ctx->ins_idx); F(ctx->ins_tmp); F(ctx->ins_cand); F(ctx->ins_rad);
    F(ctx->ins_part); F(ctx->ins_hit);
    for (char *c : ctx->ckpts) F(c);
    F(ctx->tape); F(ctx->adj_rev);
    if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
    if (ctx->cgexec) hipGraphExecDestroy(ctx->cgexec);
    for (auto &m : ctx->meshes) {
        F(m.tri); F(m.enorm); F(m.gstart); F(m.glist);
        for (int b = 0; b < 2; b++) {
            F(m.mstart[b]); F(m.mcand[b]); F(m.mpart[b]); F(m.mnp[b]); F(m.mhist[b]); if (b == 0) { F(m.d_ms); F(m.d_spart); F(m.d_sout); }
        }
    }
}

extern "C" void dem_ctx_destroy(dem_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->p.device);
    hipStreamSynchronize(ctx->stream);
    timer_collect(ctx->t_contact, true);
    timer_collect(ctx->t_integrate, true);
    timer_collect(ctx->t_rebuild, true);
    free_all(ctx);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

// global_properties.cpp:428-562
extern "C" int dem_material_tables(int ntypes, const double *Y, const double *nu, const double *e,
                                   double *Yeff, double *Geff, double *betaeff)
{
    if (ntypes < 1 || ntypes > DEM_MAXTYPES || !Y || !nu || !e || !Yeff || !Geff || !betaeff)
        return fail(nullptr, DEM_ERR_ARG, "dem_material_tables: bad arguments");
    const int w = ntypes + 1;
    for (int k = 0; k < w * w; k++) Yeff[k] = Geff[k] = betaeff[k] = 0.0;
    for (int i = 1; i <= ntypes; i++)
        for (int j = 1; j <= ntypes; j++) {
            const double Yi = Y[i], Yj = Y[j], vi = nu[i], vj = nu[j];
            Yeff[i * w + j] = 1. / ((1. - std::pow(vi, 2.)) / Yi + (1. - std::pow(vj, 2.)) / Yj);
            Geff[i * w + j] = 1. / (2. * (2. - vi) * (1. + vi) / Yi + 2. * (2. - vj) * (1. + vj) / Yj);
            const double lg = std::log(e[i * w + j]);
            betaeff[i * w + j] = lg / std::sqrt(std::pow(lg, 2.) + std::pow(M_PI, 2.));
        }
    return DEM_OK;
}

// Material tables to the device: one block per instance, each the common
// tables with the instance's own Yeff..etar when it has them.
static int upload_mat(dem_ctx *ctx)
{
    const size_t S = kMatStride, T = (size_t)kT1 * kT1;
    const int B = ctx->ninst;
    std::vector<double> all((size_t)B * S);
    for (int b = 0; b < B; b++) {
        std::copy(ctx->hmat.begin(), ctx->hmat.end(), all.begin() + (size_t)b * S);
        if (!ctx->inst_set.empty() && ctx->inst_set[b])
            std::copy(ctx->inst_mat.begin() + (size_t)b * (M_ETAR + 1) * T,
                      ctx->inst_mat.begin() + (size_t)(b + 1) * (M_ETAR + 1) * T, all.begin() + (size_t)b * S);
    }
    if (B > ctx->mat_blocks) {
        if (dev_alloc(ctx, &ctx->d_mat, all.size())) return DEM_ERR_HIP;
        ctx->mat_blocks = B;
    }
    HIPCHK(hipMemcpy(ctx->d_mat, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    return DEM_OK;
}

extern "C" int dem_set_materials(dem_ctx *ctx, int ntypes, const double *Yeff, const double *Geff,
                                 const double *betaeff, const double *mu, const double *mur,
                                 const double *etar, double krs)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ntypes < 1 || ntypes > DEM_MAXTYPES || !Yeff || !Geff || !betaeff || !mu)
        return fail(ctx, DEM_ERR_ARG, "dem_set_materials: bad arguments");
    if (ctx->p.rolling_model && !mur) return fail(ctx, DEM_ERR_ARG, "rolling model needs coeffRollFrict");
    if ((ctx->p.rolling_model == 1 || ctx->p.rolling_model == 3) && !etar)
        return fail(ctx, DEM_ERR_ARG, "rolling model needs coeffRollVisc");
    sync_enter(ctx);
    // slots Yeff..etar; the hysteretic and bond tables set before are kept
    std::vector<double> &h = ctx->hmat;
    const int w = ntypes + 1;
    const double *src[M_ETAR + 1] = {Yeff, Geff, betaeff, mu, mur, etar};
    for (int s = 0; s <= M_ETAR; s++)
        for (int i = 0; i < kT1; i++)
            for (int j = 0; j < kT1; j++)
                h[s * kT1 * kT1 + i * kT1 + j] = (src[s] && i < w && j < w) ? src[s][i * w + j] : 0.0;
    const int rc = upload_mat(ctx);
    if (rc) return rc;
    ctx->ntypes = ntypes;
    ctx->M.krs = krs;
    ctx->mat_set = true;
    return DEM_OK;
}

extern "C" int dem_set_hysteretic(dem_ctx *ctx, const double *coeff_rest_log, const double *tables)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!ctx->mat_set || !coeff_rest_log || !tables)
        return fail(ctx, DEM_ERR_ARG, "dem_set_hysteretic: set materials first, non-null tables");
    sync_enter(ctx);
    const int w = ctx->ntypes + 1;
    std::vector<double> &h = ctx->hmat;
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) {
            h[M_RESTLOG * kT1 * kT1 + i * kT1 + j] = coeff_rest_log[i * w + j];
            for (int t = 0; t < kNHyst; t++)
                h[(M_HK + t) * kT1 * kT1 + i * kT1 + j] = tables[(size_t)t * w * w + i * w + j];
        }
    return upload_mat(ctx);
}

extern "C" int dem_add_plane(dem_ctx *ctx, int dim, double pos, int wtype)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ctx->P.n >= kMaxPlanes || dim < 0 || dim > 2 || wtype < 1 || wtype > DEM_MAXTYPES)
        return fail(ctx, DEM_ERR_ARG, "dem_add_plane: bad arguments");
    if (ctx->n > 0) return fail(ctx, DEM_ERR_ARG, "dem_add_plane: add walls before uploading state");
    const int k = ctx->P.n++;
    ctx->P.dim[k] = dim;
    ctx->P.pos[k] = pos;
    ctx->P.wtype[k] = wtype;
    return DEM_OK;
}

// TriGeo records of the mesh's current host geometry
static int upload_mesh(dem_ctx *ctx, MeshDev &m)
{
    const MeshHost &h = m.h;
    std::vector<TriGeo> g(h.ntri);
    for (int t = 0; t < h.ntri; t++) {
        TriGeo &r = g[t];
        for (int q = 0; q < 9; q++) {
            r.node[q] = h.node[9 * t + q];
            r.evec[q] = h.evec[9 * t + q];
        }
        for (int q = 0; q < 3; q++) {
            r.elen[q] = h.elen[3 * t + q];
            r.snorm[q] = h.snorm[3 * t + q];
            r.center[q] = h.center[3 * t + q];
        }
        r.rbound = h.rbound[t];
        r.obtuse = h.obtuse[t];
        r.flags = 0;
        for (int k = 0; k < 3; k++) {
            if (h.eact[3 * t + k]) r.flags |= 1 << k;
            if (h.cact[3 * t + k]) r.flags |= 1 << (3 + k);
        }
        r.nneigh = h.nneigh[t];
        for (int k = 0; k < kMeshNeigh; k++) r.neigh[k] = h.neigh[kMeshNeighMax * t + k];
    }
    if (dev_alloc(ctx, &m.tri, h.ntri) || dev_alloc(ctx, &m.enorm, 9 * (size_t)h.ntri)) return DEM_ERR_HIP;
    HIPCHK(hipMemcpy(m.tri, g.data(), sizeof(TriGeo) * h.ntri, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.enorm, h.enorm.data(), sizeof(double) * 9 * h.ntri, hipMemcpyHostToDevice));
    return DEM_OK;
}

extern "C" int dem_mesh_topology(int ntri, const double *nodes, double curvature, const double *boxlo,
                                 const double *boxhi, int *ntri_kept, int *edge_active, int *corner_active,
                                 int *obtuse, int *nneigh)
{
    if (ntri < 1 || !nodes || !boxlo || !boxhi || !ntri_kept)
        return fail(nullptr, DEM_ERR_ARG, "dem_mesh_topology: bad arguments");
    MeshHost h;
    mesh_build(h, ntri, nodes, curvature, boxlo, boxhi);
    *ntri_kept = h.ntri;
    for (int t = 0; t < h.ntri; t++) {
        for (int k = 0; k < 3; k++) {
            if (edge_active) edge_active[3 * t + k] = h.eact[3 * t + k];
            if (corner_active) corner_active[3 * t + k] = h.cact[3 * t + k];
        }
        if (obtuse) obtuse[t] = h.obtuse[t];
        if (nneigh) nneigh[t] = h.nneigh[t];
    }
    return DEM_OK;
}

extern "C" int dem_add_mesh(dem_ctx *ctx, int ntri, const double *nodes, int wtype, double curvature,
                            const double *vel)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ntri < 1 || !nodes || wtype < 1 || wtype > DEM_MAXTYPES || (int)ctx->meshes.size() >= kMaxMeshes)
        return fail(ctx, DEM_ERR_ARG, "dem_add_mesh: bad arguments or too many meshes");
    for (int d = 0; d < 3; d++)
        if (ctx->p.periodic[d])
            return fail(ctx, DEM_ERR_ARG, "dem_add_mesh: mesh walls need a box without periodic dimensions");
    sync_enter(ctx);
    ctx->meshes.emplace_back();
    MeshDev &m = ctx->meshes.back();
    mesh_build(m.h, ntri, nodes, curvature, ctx->p.boxlo, ctx->p.boxhi);
    if (m.h.ntri == 0) {
        ctx->meshes.pop_back();
        return fail(ctx, DEM_ERR_ARG, "dem_add_mesh: no valid triangle");
    }
    m.wtype = wtype;
    if (vel) {
        m.moving = true;
        for (int q = 0; q < 3; q++) m.vel[q] = vel[q];
    }
    m.node_re = m.h.node;
    int rc = upload_mesh(ctx, m);
    if (rc) return rc;
    return (int)ctx->meshes.size() - 1;
}

extern "C" int dem_mesh_size(const dem_ctx *ctx, int mesh)
{
    if (!ctx || mesh < 0 || mesh >= (int)ctx->meshes.size()) return 0;
    return ctx->meshes[mesh].h.ntri;
}

static int setup_grid(dem_ctx *ctx);

extern "C" int dem_set_bond(dem_ctx *ctx, const double *tables, const dem_bond_settings *bs)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!tables || !bs || ctx->M.cohesion == CM_OFF || !ctx->mat_set
        || (ctx->n == 0 && ctx->templates.empty() && ctx->inserts.empty()))
        return fail(ctx, DEM_ERR_ARG,
                    "dem_set_bond: needs a cohesion model, materials, atoms or templates, non-null tables");
    if (ctx->ninst > 1) return fail(ctx, DEM_ERR_ARG, "dem_set_bond: bonds are not available in a batched run");
    sync_enter(ctx);
    const int w = ctx->ntypes + 1;
    std::vector<double> &h = ctx->hmat;
    for (int q = 0; q < kNBond; q++)
        for (int i = 0; i < w; i++)
            for (int j = 0; j < w; j++)
                h[(size_t)(M_BLAMBDA + q) * kT1 * kT1 + i * kT1 + j] = tables[(size_t)q * w * w + i * w + j];
    {
        const int rc = upload_mat(ctx);
        if (rc) return rc;
    }
    ctx->M.bs = BondSet{(long long)(int)bs->ts_create, bs->stress_break, bs->create_always, bs->damping,
                        bs->damping_smooth, bs->ratio_tc, bs->tension_variant};
    // the model's contact distance factor (cohesion_model_bond.h:415-470,
    // cohesion_model_bond_nonlinear.h:330-380) with the smallest radius of
    // the atoms and of the templates that insert or create them
    // (Modify::get_min_radius, modify_liggghts.cpp:515-567)
    std::vector<double4> hx(ctx->n);
    if (ctx->n) HIPCHK(hipMemcpy(hx.data(), sx(ctx), sizeof(double4) * ctx->n, hipMemcpyDeviceToHost));
    double minrad = 1e300;
    for (const auto &q : hx) minrad = std::min(minrad, q.w);
    for (const auto &t : ctx->templates)
        for (int j = 0; j < t->t.nspheres(); j++) minrad = std::min(minrad, t->t.radius(j));
    for (const auto &in : ctx->inserts) minrad = std::min(minrad, in->fix->min_radius());
    const bool nl = ctx->M.cohesion == CM_BOND_NONLINEAR;
    double cdf_all = 1.;
    auto T = [&](int slot, int i, int j) { return tables[(size_t)slot * w * w + i * w + j]; };
    for (int i = 1; i < w; i++)
        for (int j = 1; j < w; j++) {
            double cdf_one;
            if (!bs->stress_break) {
                cdf_one = 1.1 * 0.5 * T(7, i, j) / minrad;
            } else {
                double stress = T(8, i, j);
                if (bs->ratio_tc) stress = std::fmax(stress, stress * T(10, i, j));
                const double k = nl ? T(15, i, j) : T(1, i, j);
                if (k <= 1e-15) return fail(ctx, DEM_ERR_ARG, "Bond settings: In case of stress breakage, the normal bond stiffness can't be <= 0!");
                if (!nl) cdf_one = 0.5 * (1.1 * T(11, i, j) / minrad + 1.1 * stress / (k * minrad));
                else cdf_one = 0.5 * 1.1 * T(11, i, j) / minrad + 0.5 * 1.1 * stress / (k * minrad);
            }
            cdf_all = std::max(cdf_all, cdf_one);
        }
    if (cdf_all > 10.)
        return fail(ctx, DEM_ERR_ARG, "Maximum bond distance exceeding 10 x particle diameter, please reduce "
                                      "maxDistanceBond or maxSigmaBond/maxTauBond");
    if (cdf_all > ctx->p.cdf) {
        ctx->p.cdf = cdf_all;
        ctx->M.cdf = cdf_all;
        ctx->mesh_grids_valid = false;
        const int rc = setup_grid(ctx);
        if (rc) return rc;
    }
    return DEM_OK;
}

extern "C" int dem_set_gravity(dem_ctx *ctx, double ax, double ay, double az)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    ctx->M.gacc[0] = ax;
    ctx->M.gacc[1] = ay;
    ctx->M.gacc[2] = az;
    ctx->M.gravity_on = 1;
    return DEM_OK;
}

extern "C" int dem_set_freeze(dem_ctx *ctx, int on)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    ctx->M.freeze_on = on ? 1 : 0;
    return DEM_OK;
}

// Cell grid + Morton spread tables for the current cut-off.
// Cells of a whole cutneighmax searched with the 3x3x3 stencil, pruned by
// distance.  DEM_NEIGH_REACH=2 selects cells of half cutneighmax, the
// reference's bin size (neighbor.cpp:1689 binsize_optimal = 0.5 *
// cutneighmax), with a 5x5x5 stencil: 40 % fewer candidates, but 125 cell
// look-ups per particle; on the 1M column its search pass took 805 us per
// rebuild against 480.  The pair sets are the same; the order of each row
// (and so each particle's force sum) follows the cells.
static int neigh_reach()
{
    static int r = 0;
    if (r == 0) {
        const char *s = getenv("DEM_NEIGH_REACH");
        r = (s && atoi(s) == 2) ? 2 : 1;
    }
    return r;
}

static int setup_grid(dem_ctx *ctx)
{
    const dem_params &p = ctx->p;
    const double cut = (ctx->maxrad + ctx->maxrad) * p.cdf + p.skin;
    ctx->M.cutneighmax = cut;
    Grid &G = ctx->G;
    G.reach = neigh_reach();
    int bits[3];
    int maxc = 1;
    for (int d = 0; d < 3; d++) {
        const double L = p.boxhi[d] - p.boxlo[d];
        int nc = (int)std::floor(L / (cut / G.reach * 1.0001));
        if (nc < 1) nc = 1;
        if (p.periodic[d] && nc < 2 * G.reach + 1)
            return fail(ctx, DEM_ERR_ARG,
                        "periodic box dimension must span at least 3 x cutneighmax (half-size neighbour "
                        "cells: 2.5 x cutneighmax)");
        if (nc > 4096) nc = 4096;
        G.nc[d] = nc;
        G.lo[d] = p.boxlo[d];
        G.inv[d] = nc / L;
        G.periodic[d] = p.periodic[d];
        int b = 0;
        while ((1 << b) < nc) b++;
        bits[d] = b;
        maxc = std::max(maxc, nc);
    }
    int total = bits[0] + bits[1] + bits[2];
    while (total > 26) {  // coarsen the largest non-periodic dimension
        int dm = -1;
        for (int d = 0; d < 3; d++)
            if (!p.periodic[d] && (dm < 0 || G.nc[d] > G.nc[dm])) dm = d;
        if (dm < 0) return fail(ctx, DEM_ERR_ARG, "cell grid too large");
        G.nc[dm] = (G.nc[dm] + 1) / 2;
        G.inv[dm] = G.nc[dm] / (p.boxhi[dm] - p.boxlo[dm]);
        int b = 0;
        while ((1 << b) < G.nc[dm]) b++;
        bits[dm] = b;
        total = bits[0] + bits[1] + bits[2];
    }
    maxc = std::max(G.nc[0], std::max(G.nc[1], G.nc[2]));
    // bit positions of the interleave, dimensions drop out as they run out
    int posmap[3][32];
    int pos = 0;
    for (int l = 0; l < 27; l++)
        for (int d = 0; d < 3; d++)
            if (l < bits[d]) posmap[d][l] = pos++;
    std::vector<uint32_t> h(3 * maxc, 0u);
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < G.nc[d]; c++) {
            uint32_t v = 0;
            for (int l = 0; l < bits[d]; l++)
                if (c & (1 << l)) v |= 1u << posmap[d][l];
            h[d * maxc + c] = v;
        }
    if (3 * maxc > ctx->spread_cap) {
        if (dev_alloc(ctx, &ctx->spread, 3 * maxc)) return DEM_ERR_HIP;
        ctx->spread_cap = 3 * maxc;
    }
    HIPCHK(hipMemcpy(ctx->spread, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    G.spread = ctx->spread;
    G.maxc = maxc;
    G.nkeys_local = 1 << total;
    if ((int64_t)G.nkeys_local * ctx->ninst >= (int64_t)1 << 30)
        return fail(ctx, DEM_ERR_ARG, "cell grid too large for this many instances");
    G.nkeys = G.nkeys_local * ctx->ninst;
    if (G.nkeys + 1 > ctx->nkeys_cap) {
        if (dev_alloc(ctx, &ctx->cstart, G.nkeys + 1) || dev_alloc(ctx, &ctx->cursor, G.nkeys + 1))
            return DEM_ERR_HIP;
        ctx->nkeys_cap = G.nkeys + 1;
    }
    const int need = std::max(scan_tmp_size(G.nkeys + 1), scan_tmp_size(ctx->n + 1));
    if (need > ctx->scan_tmp_cap) {
        if (dev_alloc(ctx, &ctx->scan_tmp, need)) return DEM_ERR_HIP;
        ctx->scan_tmp_cap = need;
    }
    return DEM_OK;
}

extern "C" int dem_upload_state(dem_ctx *ctx, int n, const double *x, const double *v,
                                const double *omega, const double *radius, const double *rmass,
                                const int *type, const int *tag, const int *freeze)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (n < 0 || n > (int)kJMask || (n > 0 && (!x || !v || !omega || !radius || !rmass || !type || !tag)))
        return fail(ctx, DEM_ERR_ARG, "dem_upload_state: bad arguments");
    if (!ctx->mat_set) return fail(ctx, DEM_ERR_ARG, "dem_upload_state: set materials first");
    sync_enter(ctx);
    const int wdn = ctx->P.n * ctx->M.wdnum;
    std::vector<double4> hx(n), hv(n), hw(n);
    std::vector<int> ho(n);
    std::vector<int4> hi(n, make_int4(0, 0, 0, 0));
    double maxrad = 0.0;
    for (int i = 0; i < n; i++) {
        if (type[i] < 1 || type[i] > ctx->ntypes)
            return fail(ctx, DEM_ERR_ARG, "dem_upload_state: atom type out of range");
        if (!(radius[i] > 0.0) || !(rmass[i] > 0.0))
            return fail(ctx, DEM_ERR_ARG, "Fix nve/sphere requires extended particles");
        hx[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], radius[i]);
        hv[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], rmass[i]);
        const int meta = type[i] + ((freeze && freeze[i]) ? 16 : 0);
        hw[i] = make_double4(omega[3 * i], omega[3 * i + 1], omega[3 * i + 2], (double)meta);
        ho[i] = i;
        maxrad = std::max(maxrad, radius[i]);
    }
    int tag_max = 0;
    for (int i = 0; i < n; i++) tag_max = std::max(tag_max, tag[i]);
    ctx->tag_max = tag_max;
    if (n > ctx->cap) {
        const size_t c = (size_t)n;
        for (int b = 0; b < 2; b++) {
            if (dev_alloc(ctx, &ctx->xr[b], c) || dev_alloc(ctx, &ctx->vm[b], c)
                || dev_alloc(ctx, &ctx->wt[b], c) || dev_alloc(ctx, &ctx->tag[b], c)
                || dev_alloc(ctx, &ctx->orig[b], c) || dev_alloc(ctx, &ctx->img[b], c)
                || dev_alloc(ctx, &ctx->whist[b], c * std::max(wdn, 1))
                || dev_alloc(ctx, &ctx->nstart[b], c + 1) || dev_alloc(ctx, &ctx->gmask[b], c))
                return DEM_ERR_HIP;
        }
        if (dev_alloc(ctx, &ctx->f4, c) || dev_alloc(ctx, &ctx->t4, c) || dev_alloc(ctx, &ctx->xhold, c)
            || dev_alloc(ctx, &ctx->wmask, c) || dev_alloc(ctx, &ctx->ncontact, c)
            || dev_alloc(ctx, &ctx->key, c) || dev_alloc(ctx, &ctx->sorted, c)
            || dev_alloc(ctx, &ctx->stage, c * kStageSlots) || dev_alloc(ctx, &ctx->xn, c)
            || dev_alloc(ctx, &ctx->vn, c) || dev_alloc(ctx, &ctx->wn, c))
            return DEM_ERR_HIP;
        ctx->cap = n;
    }
    ctx->n = n;
    ctx->cur = 0;
    ctx->par = 0;
    ctx->maxrad = maxrad;
    if (n > 0) {
        HIPCHK(hipMemcpy(ctx->xr[0], hx.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->vm[0], hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->wt[0], hw.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->tag[0], tag, n * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->orig[0], ho.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->img[0], hi.data(), n * sizeof(int4), hipMemcpyHostToDevice));
        if (wdn) HIPCHK(hipMemset(ctx->whist[0], 0, sizeof(double) * n * wdn));
        HIPCHK(hipMemset(ctx->f4, 0, n * sizeof(double4)));
        HIPCHK(hipMemset(ctx->t4, 0, n * sizeof(double4)));
        HIPCHK(hipMemset(ctx->wmask, 0, n));
        HIPCHK(hipMemset(ctx->gmask[0], 0, n * sizeof(int)));
    }
    // the group bits are cleared, and with them the groups fix move acts on
    ctx->moves.clear();
    // no atom carries a bond_random_id or awaits a stream's release
    for (int b = 0; b < 2; b++) {
        if (ctx->brid[b]) {
            HIPCHK(hipFree(ctx->brid[b]));
            ctx->brid[b] = nullptr;
        }
        if (ctx->rel[b]) {
            HIPCHK(hipFree(ctx->rel[b]));
            ctx->rel[b] = nullptr;
        }
    }
    ctx->rel_until = -1;
    // new atoms: massflow side counters not seen yet
    if (ctx->mf.mesh >= 0 && massflow_counters(ctx)) return DEM_ERR_HIP;
    ctx->integ_bit = -1;
    ctx->have_list = false;
    ctx->list_valid = false;
    ctx->mesh_grids_valid = false;
    for (auto &m : ctx->meshes) m.have = false;
    int rc = setup_grid(ctx);
    if (rc) return rc;
    return DEM_OK;
}

static int ensure_list_cap(dem_ctx *ctx, int b, int64_t entries)
{
    if (entries <= ctx->ecap[b]) return DEM_OK;
    const int64_t c = entries + entries / 4 + 1024;
    // keep the old contents of the other buffer; this one is about to be overwritten
    if (dev_alloc(ctx, &ctx->nj[b], (size_t)c) || dev_alloc(ctx, &ctx->hist[b], (size_t)c * ctx->M.dnum))
        return DEM_ERR_NOMEM;
    ctx->ecap[b] = c;
    return DEM_OK;
}

// Candidate triangles and contact slots of every particle for one mesh, after
// the particles were re-sorted (perm = new -> old): FixNeighlistMesh::
// pre_force + FixContactHistoryMesh::sort_contacts / pre_force.
// captured (graph stepping, static meshes): no host reads or allocations;
// the buffers were sized before the capture and guards flag what outgrows them
static int mesh_rebuild(dem_ctx *ctx, MeshDev &m, bool captured = false)
{
    const int n = ctx->n;
    hipStream_t s = ctx->stream;
    const int b = m.have ? 1 - m.cur : m.cur;
    const int ob = m.cur;
    if (n + 1 > m.ncap[b]) {
        if (dev_alloc(ctx, &m.mstart[b], n + 1) || dev_alloc(ctx, &m.mnp[b], n + 1)) return DEM_ERR_HIP;
        m.ncap[b] = n + 1;
    }
    MeshNbArgs na{};
    na.xr = sx(ctx);
    na.tri = m.tri;
    na.enorm = m.enorm;
    na.gstart = m.gstart;
    na.glist = m.glist;
    for (int d = 0; d < 3; d++) {
        na.glo[d] = m.grid.lo[d];
        na.ginv[d] = m.grid.inv[d];
        na.gn[d] = m.grid.n[d];
        na.disp[d] = m.disp[d];
    }
    na.cdf = ctx->p.cdf;
    // FixNeighlistMesh::pre_force: skin for a moving mesh, half of it else
    na.threshold = m.moving ? ctx->p.skin : 0.5 * ctx->p.skin;
    na.n = n;
    na.count = m.mstart[b];
    if (n > ctx->mstage_cap) {   // sized at the first (eager) rebuild, before any capture
        if (captured) return fail(ctx, DEM_ERR_STATE, "mesh staging rows not sized before graph capture");
        if (dev_alloc(ctx, &ctx->mstage, (size_t)n * kMeshStage)) return DEM_ERR_NOMEM;
        ctx->mstage_cap = n;
    }
    na.stage = ctx->mstage;
    launch_fill_int(m.mstart[b] + n, 0, 1, s);
    launch_fill_int(ctx->d_small + 3, 0, 1, s);
    launch_mesh_stage(na, s);   // one search: counts and staged rows
    launch_row_max(m.mstart[b], n, ctx->d_small + 3, s);
    launch_scan(m.mstart[b], n + 1, ctx->scan_tmp, s);
    KCHK();
    const int dw = ctx->M.wdnum;
    int hs1 = 0;   // longest row (eager path)
    if (captured) {
        launch_row_max_check(ctx->d_small + 3, kMaxMeshSlots, ctx->d_small, 64, s);
        launch_cap_guard(m.mstart[b], n, m.scap[b], ctx->d_small + 5, ctx->d_small, 32, s);
    } else {
        int hs[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&hs[0], m.mstart[b] + n, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&hs[1], ctx->d_small + 3, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (hs[1] > kMaxMeshSlots)
            return fail(ctx, DEM_ERR_OVERFLOW, "mesh contact history overflow: a particle has more than 4096 "
                                               "candidate triangles");
        hs1 = hs[1];
        const int64_t tot = hs[0];
        if (tot > m.scap[b]) {
            const int64_t c = tot + tot / 4 + 256;
            if (dev_alloc(ctx, &m.mcand[b], (size_t)c) || dev_alloc(ctx, &m.mpart[b], (size_t)c)
                || dev_alloc(ctx, &m.mhist[b], (size_t)c * std::max(dw, 1)))
                return DEM_ERR_NOMEM;
            m.scap[b] = c;
        }
    }
    na.mstart = m.mstart[b];
    na.mcand = m.mcand[b];
    // rows that fit their stage are copied; the rest are searched again
    launch_mesh_compact(na, s);
    if (captured || hs1 > kMeshStage) {
        na.fill_if_over = 1;
        launch_mesh_fill(na, s);
    }
    KCHK();
    MeshRemapArgs ra{};
    ra.perm = ctx->sorted;
    ra.have_old = m.have ? 1 : 0;
    ra.ostart = m.mstart[ob];
    ra.opart = m.mpart[ob];
    ra.ohist = m.mhist[ob];
    ra.onp = m.mnp[ob];
    ra.mstart = m.mstart[b];
    ra.mcand = m.mcand[b];
    ra.mpart = m.mpart[b];
    ra.mhist = m.mhist[b];
    ra.mnp = m.mnp[b];
    ra.n = n;
    ra.dnum = dw;
    launch_mesh_remap(ra, s);
    KCHK();
    m.cur = b;
    m.have = true;
    // MultiNodeMesh::storeNodePosRebuild
    m.node_re = m.h.node;
    return DEM_OK;
}

static int setup_mesh_grids(dem_ctx *ctx);

// Neighbour rebuild: pre_exchange + pbc + sort + list + wall lists.
// One neighbour rebuild.  With captured = true it is being recorded into a
// graph: no host round trip, the list buffers were sized beforehand and a
// longer list raises error bit 3 instead of being written.
// DEM_NEIGH_STAGE=0 selects the two-search rebuild (count pass, then a fill
// pass that repeats the search); the default searches once into staging rows
static bool neigh_staged()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("DEM_NEIGH_STAGE");
        v = (e && atoi(e) == 0) ? 0 : 1;
    }
    return v == 1;
}

static int rebuild(dem_ctx *ctx, bool captured = false)
{
    const int n = ctx->n;
    hipStream_t s = ctx->stream;
    if (!captured) timer_begin(ctx, ctx->t_rebuild);
    const Grid &G = ctx->G;
    const int c0 = ctx->cur, c1 = 1 - ctx->cur;
    launch_fill_int(ctx->cstart, 0, G.nkeys + 1, s);
    BinArgs ba{};
    ba.xr = sx(ctx);
    ba.wt = sw(ctx);
    ba.img = ctx->img[c0];
    ba.key = ctx->key;
    ba.count = ctx->cstart;
    ba.n = n;
    ba.G = G;
    for (int d = 0; d < 3; d++) {
        ba.prd[d] = ctx->M.prd[d];
        ba.hi[d] = ctx->p.boxhi[d];
    }
    launch_bin(ba, s);
    KCHK();
    launch_scan(ctx->cstart, G.nkeys + 1, ctx->scan_tmp, s);
    KCHK();
    launch_copy_int(ctx->cursor, ctx->cstart, G.nkeys + 1, s);
    launch_scatter(ctx->key, ctx->cursor, ctx->sorted, n, s);
    launch_cellsort(ctx->cstart, ctx->sorted, G.nkeys, s);
    KCHK();
    PermuteArgs pa{};
    pa.sorted = ctx->sorted;
    pa.xr0 = sx(ctx); pa.vm0 = sv(ctx); pa.wt0 = sw(ctx);
    pa.tag0 = ctx->tag[c0]; pa.orig0 = ctx->orig[c0]; pa.img0 = ctx->img[c0];
    pa.whist0 = ctx->whist[c0];
    pa.xr1 = ctx->xr[c1]; pa.vm1 = ctx->vm[c1]; pa.wt1 = ctx->wt[c1];
    pa.tag1 = ctx->tag[c1]; pa.orig1 = ctx->orig[c1]; pa.img1 = ctx->img[c1];
    pa.whist1 = ctx->whist[c1];
    pa.gm0 = ctx->gmask[c0];
    pa.gm1 = ctx->gmask[c1];
    pa.xo0 = ctx->moves.empty() ? nullptr : ctx->xo[c0];
    pa.xo1 = ctx->moves.empty() ? nullptr : ctx->xo[c1];
    pa.mf0 = ctx->mf.mesh >= 0 ? ctx->mf.counter[c0] : nullptr;
    pa.mf1 = ctx->mf.mesh >= 0 ? ctx->mf.counter[c1] : nullptr;
    pa.br0 = ctx->brid[c0];
    pa.br1 = ctx->brid[c1];
    pa.rl0 = ctx->rel[c0];
    pa.rl1 = ctx->rel[c1];
    pa.xhold = ctx->xhold;
    pa.wmask = ctx->wmask;
    pa.n = n;
    pa.wdn = ctx->P.n * ctx->M.wdnum;
    pa.P = ctx->P;
    pa.skin = ctx->p.skin;
    launch_permute(pa, s);
    KCHK();
    ctx->cur = c1;
    ctx->par = 0;

    const int l0 = ctx->lcur, l1 = 1 - ctx->lcur;
    NeighArgs na{};
    na.xr = ctx->xr[c1];
    na.wt = ctx->wt[c1];
    na.tag = ctx->tag[c1];
    na.cstart = ctx->cstart;
    na.count = ctx->nstart[l1];
    na.maxcount = ctx->d_small + 1;
    na.n = n;
    na.dnum = ctx->M.dnum;
    na.cdf = ctx->p.cdf;
    na.skin = ctx->p.skin;
    na.rmax = ctx->maxrad;
    for (int d = 0; d < 3; d++) na.cw[d] = 1.0 / G.inv[d];
    na.G = G;
    for (int d = 0; d < 3; d++) na.prd[d] = ctx->M.prd[d];
    na.perm = ctx->sorted;
    na.have_old = ctx->have_list ? 1 : 0;
    na.onstart = ctx->nstart[l0];
    na.onj = ctx->nj[l0];
    na.ohist = ctx->hist[l0];
    na.otag = ctx->tag[c0];
    na.stage = ctx->stage;
    const bool staged = neigh_staged();
    launch_fill_int(ctx->d_small + 1, 0, 1, s);
    launch_fill_int(ctx->nstart[l1] + n, 0, 1, s);
    if (staged) launch_neigh_stage(na, s);
    else launch_neigh_count(na, s);
    KCHK();
    launch_scan(ctx->nstart[l1], n + 1, ctx->scan_tmp, s);
    KCHK();
    int hsmall[2] = {0, 0};
    if (!captured) {
        HIPCHK(hipMemcpyAsync(&hsmall[0], ctx->nstart[l1] + n, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&hsmall[1], ctx->d_small + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (hsmall[1] > ctx->p.neigh_one)
            return fail(ctx, DEM_ERR_OVERFLOW, "Neighbor list overflow, boost neigh_modify one");
        if (ensure_list_cap(ctx, l1, hsmall[0])) return DEM_ERR_NOMEM;
    }
    if (captured) launch_cap_guard(ctx->nstart[l1], n, ctx->ecap[l1], ctx->d_small + 4, ctx->d_small, 8, s);
    na.ecap = ctx->ecap[l1];
    na.coh_off = ctx->M.cohesion != CM_OFF ? ctx->M.lp.off_coh : -1;
    na.err = ctx->d_small;
    na.nstart = ctx->nstart[l1];
    na.nj = ctx->nj[l1];
    na.hist = ctx->hist[l1];
    if (!staged) {
        launch_neigh_fill(na, s);
    } else if (captured) {   // the device picks: each kernel returns at once unless it is the one
        launch_neigh_compact(na, s);
        na.fill_if_over = 1;
        launch_neigh_fill(na, s);
    } else if (hsmall[1] <= kStageSlots) {
        launch_neigh_compact(na, s);
    } else {
        launch_neigh_fill(na, s);
    }
    KCHK();
    if (captured) launch_row_max_check(ctx->d_small + 1, ctx->p.neigh_one, ctx->d_small, 16, s);
    ctx->lcur = l1;
    ctx->have_list = true;
    ctx->list_valid = true;
    ctx->ago = 0;
    if (!ctx->meshes.empty() && !ctx->mesh_grids_valid) {
        const int rc = setup_mesh_grids(ctx);
        if (rc) return rc;
    }
    for (auto &m : ctx->meshes) {
        const int rc = mesh_rebuild(ctx, m, captured);
        if (rc) return rc;
    }
    if (ctx->M.cohesion != CM_OFF) {
        const int64_t E = hsmall[0];
        if (E > ctx->bcap) {
            const int64_t c = E + E / 4 + 1024;
            if (dev_alloc(ctx, &ctx->bout, (size_t)c * 6) || dev_alloc(ctx, &ctx->bvalid, (size_t)c)
                || dev_alloc(ctx, &ctx->brev, (size_t)c))
                return DEM_ERR_NOMEM;
            ctx->bcap = c;
        }
        launch_bond_rev(ctx->nstart[l1], ctx->nj[l1], ctx->brev, n, s);
        KCHK();
    }
    if (!captured) timer_end(ctx, ctx->t_rebuild);
    return DEM_OK;
}

static int integrate(dem_ctx *ctx, int mode);

// index of the last mesh that is a wall (not a massflow face), -1 if none
static int last_wall_mesh(const dem_ctx *ctx)
{
    int k = -1;
    for (size_t q = 0; q < ctx->meshes.size(); q++)
        if (!ctx->meshes[q].passive) k = (int)q;
    return k;
}

// nothing adds to the forces after the contact pass (or, with mesh walls,
// after the last mesh pass), so that pass may integrate
static bool fuse_possible(const dem_ctx *ctx)
{
    return ctx->fops.empty() && !ctx->window && ctx->tape_pos < 0 && ctx->M.cohesion == CM_OFF
           && ctx->moves.empty() && ctx->integ_bit < 0;
}

// The force pass; integ = 2 or 3 also runs final_integrate (and the next
// step's initial_integrate), inside the last pass that adds to the forces
// (the contact pass, or the last mesh pass) when nothing follows it (no
// setforce/aveforce, bonds, adjoint window or tape): the new state then goes
// to the other x/v/omega buffers and par flips.
static int force_pass(dem_ctx *ctx, int shearupdate, int integ = 0)
{
    ForceArgs fa{};
    const int c = ctx->cur, l = ctx->lcur;
    fa.xr = sx(ctx); fa.vm = sv(ctx); fa.wt = sw(ctx);
    fa.nstart = ctx->nstart[l];
    fa.nj = ctx->nj[l];
    fa.hist = ctx->hist[l];
    fa.f4 = ctx->f4; fa.t4 = ctx->t4;
    fa.mat = ctx->d_mat;
    fa.wmask = ctx->wmask;
    fa.whist = ctx->whist[c];
    fa.ncontact = ctx->timing ? ctx->ncontact : nullptr;
    ctx->nc_fresh = ctx->timing;
    fa.err = ctx->d_small;
    fa.n = ctx->n;
    fa.shearupdate = shearupdate;
    fa.M = ctx->M;
    fa.P = ctx->P;
    const int lastwall = last_wall_mesh(ctx);
    fa.defer_freeze = lastwall < 0 ? 0 : 1;
    fa.pwf = ctx->obs_plane >= 0 ? ctx->pwf : nullptr;
    fa.obs_plane = ctx->obs_plane;
    fa.lanes = ctx->tune_lanes;
    fa.block = ctx->tune_block;
    fa.halt = ctx->halt_dev;
    const bool fuse = integ && lastwall < 0 && fuse_possible(ctx) && force_can_fuse(fa);
    const bool mfuse = integ && lastwall >= 0 && fuse_possible(ctx);
    if (fuse) {
        fa.fuse = integ;
        fa.xn = ctx->par ? ctx->xr[c] : ctx->xn;
        fa.vn = ctx->par ? ctx->vm[c] : ctx->vn;
        fa.wn = ctx->par ? ctx->wt[c] : ctx->wn;
        fa.dtf = 0.5 * ctx->p.dt;   // as in integrate()
        fa.dtv = ctx->p.dt;
        fa.dtfrotate = fa.dtf / 0.4;
    }
    if (ctx->M.cohesion != CM_OFF) {
        BondArgs ba{};
        ba.xr = fa.xr; ba.vm = fa.vm; ba.wt = fa.wt;
        ba.tag = ctx->tag[c];
        ba.nstart = fa.nstart;
        ba.nj = fa.nj;
        ba.hist = fa.hist;
        ba.rev = ctx->brev;
        ba.bout = ctx->bout;
        ba.bvalid = ctx->bvalid;
        ba.mat = ctx->d_mat;
        ba.brid = ctx->brid[c];
        ba.n = ctx->n;
        ba.shearupdate = shearupdate;
        ba.ntimestep = ctx->ntimestep;
        ba.halt = ctx->halt_dev;
        ba.step_dev = ctx->step_dev;
        ba.M = ctx->M;
        launch_bond(ba, ctx->stream);
        KCHK();
        fa.bout = ctx->bout;
        fa.bvalid = ctx->bvalid;
    }
    static const bool use_rec = getenv("DEM_MIXED_REC") && atoi(getenv("DEM_MIXED_REC")) == 1;
    if (ctx->p.precision == DEM_MIXED && use_rec) {
        if (ctx->rec_cap < ctx->n) {
            if (dev_alloc(ctx, &ctx->rec, 2 * (size_t)ctx->cap)) return DEM_ERR_HIP;
            ctx->rec_cap = ctx->cap;
        }
        launch_pack_rec(fa.vm, fa.wt, ctx->rec, ctx->n, ctx->stream);
        KCHK();
        fa.rec = ctx->rec;
    }
    if (!ctx->gexec_capturing) timer_begin(ctx, ctx->t_contact);
    if (ctx->p.precision == DEM_MIXED) launch_force_mixed(fa, ctx->stream);
    else launch_force_fp64(fa, ctx->stream);
    if (!ctx->gexec_capturing) timer_end(ctx, ctx->t_contact);
    KCHK();
    if (fuse) ctx->par ^= 1;
    for (size_t k = 0; k < ctx->meshes.size(); k++) {
        MeshDev &m = ctx->meshes[k];
        if (m.passive) continue;
        MeshForceArgs ma{};
        ma.halt = ctx->halt_dev;
        ma.xr = fa.xr; ma.vm = fa.vm; ma.wt = fa.wt;
        ma.f4 = fa.f4; ma.t4 = fa.t4;
        ma.tri = m.tri;
        ma.mstart = m.mstart[m.cur];
        ma.mcand = m.mcand[m.cur];
        ma.mpart = m.mpart[m.cur];
        ma.mhist = m.mhist[m.cur];
        ma.mnp = m.mnp[m.cur];
        ma.mat = ctx->d_mat;
        ma.err = ctx->d_small;
        ma.n = ctx->n;
        ma.shearupdate = shearupdate;
        ma.wtype = m.wtype;
        ma.moving = m.moving ? 1 : 0;
        for (int q = 0; q < 3; q++) ma.vel[q] = m.vel[q];
        ma.curvature = m.h.curvature;
        ma.precision = m.h.precision;
        ma.apply_freeze = (int)k == lastwall;
        if (mfuse && ma.apply_freeze) {
            ma.fuse = integ;
            ma.xn = ctx->par ? ctx->xr[c] : ctx->xn;
            ma.vn = ctx->par ? ctx->vm[c] : ctx->vn;
            ma.wn = ctx->par ? ctx->wt[c] : ctx->wn;
            ma.dtf = 0.5 * ctx->p.dt;   // as in integrate()
            ma.dtv = ctx->p.dt;
            ma.dtfrotate = ma.dtf / 0.4;
        }
        if (m.stress) {
            if (m.ms_cap < ctx->n) {
                if (dev_alloc(ctx, &m.d_ms, 6 * (size_t)ctx->cap)) return DEM_ERR_HIP;
                m.ms_cap = ctx->cap;
            }
            ma.mstress = m.d_ms;
            for (int q = 0; q < 3; q++) ma.pref[q] = m.pref[q];
        }
        ma.M = ctx->M;
        launch_mesh_force(ma, ctx->stream);
        KCHK();
        if (m.stress) {
            launch_sum6(m.d_ms, ctx->n, m.d_spart, m.d_sout, ctx->stream);
            KCHK();
        }
    }
    if (!ctx->fops.empty()) {
        // Update::atime + (ntimestep - atimestep) * dt, as the reference's
        // equal-style variables see it
        const double time = ctx->atime + (double)(ctx->ntimestep - ctx->atimestep) * ctx->p.dt;
        for (const auto &o : ctx->fops) {
            GroupForceArgs ga{};
            ga.f4 = ctx->f4;
            ga.gmask = ctx->gmask[c];
            ga.n = ctx->n;
            ga.bit = o.bit;
            ga.aveforce = o.kind;
            for (int d = 0; d < 3; d++) {
                ga.style[d] = o.style[d];
                ga.val[d] = o.val[d];
            }
            if (o.kind == 1) ga.val[2] = o.c0 - (time - o.t0) * o.c1;
            ga.part = ctx->d_gpart;
            ga.total = ctx->d_gtot + 4 * o.slot;
            launch_group_force(ga, ctx->stream);
            KCHK();
        }
    }
    if (mfuse) ctx->par ^= 1;
    if (integ && !fuse && !mfuse) return integrate(ctx, integ);
    return DEM_OK;
}

// the host side of move_meshes: the host copy of the nodes, the
// displacement since the grid build, p_ref
static void move_meshes_host(dem_ctx *ctx)
{
    for (auto &m : ctx->meshes) {
        if (!m.moving) continue;
        double dx[3];
        for (int q = 0; q < 3; q++) dx[q] = m.vel[q] * ctx->p.dt;
        for (int t = 0; t < m.h.ntri; t++) {
            for (int k = 0; k < 9; k++) m.h.node[9 * t + k] = m.h.node[9 * t + k] + dx[k % 3];
            for (int q = 0; q < 3; q++) m.h.center[3 * t + q] = m.h.center[3 * t + q] + dx[q];
        }
        for (int q = 0; q < 3; q++) m.disp[q] += dx[q];
        // p_ref moves with the mesh (frame_general, mesh_module_stress.cpp:71)
        for (int q = 0; q < 3; q++) m.pref[q] += dx[q];
    }
}

// fix move/mesh linear, MeshMoverLinear::initial_integrate: the device nodes
// and the host copy get the same additions (a captured graph records the
// device part; the host part is applied per replayed step)
static int move_meshes(dem_ctx *ctx)
{
    for (auto &m : ctx->meshes) {
        if (!m.moving) continue;
        double dx[3];
        for (int q = 0; q < 3; q++) dx[q] = m.vel[q] * ctx->p.dt;
        launch_mesh_move(m.tri, m.h.ntri, dx[0], dx[1], dx[2], ctx->stream, ctx->halt_dev);
        KCHK();
    }
    if (!ctx->gexec_capturing) move_meshes_host(ctx);
    return DEM_OK;
}

// MultiNodeMesh::decideRebuild on a step without re-neighbouring: a node
// moved more than half the skin since the last build -> rebuild next step
static void mesh_trigger(dem_ctx *ctx)
{
    const double trig = 0.25 * ctx->p.skin * ctx->p.skin;
    for (auto &m : ctx->meshes) {
        if (!m.moving) continue;
        for (int q = 0; q < 3 * m.h.ntri; q++) {
            const double dx = m.h.node[3 * q] - m.node_re[3 * q];
            const double dy = m.h.node[3 * q + 1] - m.node_re[3 * q + 1];
            const double dz = m.h.node[3 * q + 2] - m.node_re[3 * q + 2];
            if (dx * dx + dy * dy + dz * dz > trig) {
                ctx->next_reneighbor = ctx->ntimestep + 1;
                return;
            }
        }
    }
}

// triangle grids for the candidate search, in each mesh's current frame
static int setup_mesh_grids(dem_ctx *ctx)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));   // the old grids may still be read
    double ext = 0.0;
    for (int d = 0; d < 3; d++) ext = std::max(ext, ctx->p.boxhi[d] - ctx->p.boxlo[d]);
    for (auto &m : ctx->meshes) {
        const double thr = m.moving ? ctx->p.skin : 0.5 * ctx->p.skin;
        const double reach = ctx->maxrad * ctx->p.cdf + thr + 1e-9 * ext;
        mesh_grid_build(m.h, reach, m.grid);
        if (dev_alloc(ctx, &m.gstart, m.grid.start.size()) || dev_alloc(ctx, &m.glist, m.grid.list.size()))
            return DEM_ERR_HIP;
        HIPCHK(hipMemcpy(m.gstart, m.grid.start.data(), sizeof(int) * m.grid.start.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(m.glist, m.grid.list.data(), sizeof(int) * m.grid.list.size(), hipMemcpyHostToDevice));
        for (int q = 0; q < 3; q++) m.disp[q] = 0.0;
    }
    ctx->mesh_grids_valid = true;
    return DEM_OK;
}

static int integrate(dem_ctx *ctx, int mode)
{
    IntegArgs ia{};
    const int c = ctx->cur;
    ia.xr = sx(ctx); ia.vm = sv(ctx); ia.wt = sw(ctx);
    ia.f4 = ctx->f4; ia.t4 = ctx->t4;
    ia.n = ctx->n;
    ia.dtf = 0.5 * ctx->p.dt;   // dtf = 0.5*dt*ftm2v (fix_nve.cpp:85), ftm2v = 1 (si)
    ia.dtv = ctx->p.dt;
    ia.dtfrotate = ia.dtf / 0.4;
    ia.integ_bit = ctx->integ_bit;
    ia.halt = ctx->halt_dev;
    if (ctx->fold && mode == 3) {
        ia.ctl = ctx->ctl;
        ia.par = ctx->fold_par;
        ia.xhold = ctx->xhold;
        const double trig = 0.5 * ctx->p.skin;   // as decide() squares it
        ia.deltasq = trig * trig;
        for (auto &m : ctx->meshes) {
            if (!m.moving) continue;
            ia.mtri[ia.nmove] = m.tri;
            ia.mntri[ia.nmove] = m.h.ntri;
            for (int q = 0; q < 3; q++) ia.mdx[ia.nmove][q] = m.vel[q] * ctx->p.dt;
            ia.nmove++;
        }
    }
    ia.mv_bit = -1;
    if (ctx->integ_bit >= 0 || !ctx->moves.empty()) ia.gm = ctx->gmask[c];
    if (!ctx->moves.empty()) {
        const auto &mv = ctx->moves[0];
        ia.mv_bit = mv.bit;
        ia.mv_style = mv.style;
        for (int d = 0; d < 3; d++) ia.mv_val[d] = mv.val[d];
        // the initial half belongs to the step being started: the current
        // one when a call begins (mode 1), the next one after a force pass
        const long long step = (mode == 1) ? ctx->ntimestep : ctx->ntimestep + 1;
        ia.mv_delta = (double)(step - mv.t0) * ctx->p.dt;
        ia.xo = ctx->xo[c];
    }
    if (!ctx->gexec_capturing) timer_begin(ctx, ctx->t_integrate);
    launch_integrate(ia, mode, ctx->stream);
    if (!ctx->gexec_capturing) timer_end(ctx, ctx->t_integrate);
    KCHK();
    return DEM_OK;
}

static int check_err_flag(dem_ctx *ctx)
{
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, ctx->d_small, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (h & 8)
        return fail(ctx, DEM_ERR_OVERFLOW,
                    "neighbour list longer than the buffers sized for graph stepping (step without dem_use_graph)");
    if (h & 1)
        return fail(ctx, DEM_ERR_STATE,
                    "ilegal situation, particle positions identical. Most probable reason: Simulation diverged");
    if (h & 4) return fail(ctx, DEM_ERR_OVERFLOW, "mesh contact history: no free contact slot (internal error)");
    if (h & 16) return fail(ctx, DEM_ERR_OVERFLOW, "Neighbor list overflow, boost neigh_modify one");
    if (h & 32)
        return fail(ctx, DEM_ERR_OVERFLOW,
                    "mesh candidate rows longer than the buffers sized for graph stepping (step without dem_use_graph)");
    if (h & 64)
        return fail(ctx, DEM_ERR_OVERFLOW, "mesh contact history overflow: a particle has more than 4096 candidate "
                                           "triangles");
    return DEM_OK;
}

extern "C" int dem_rebuild(dem_ctx *ctx)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    sync_enter(ctx);
    if (ctx->n == 0) return DEM_OK;
    return rebuild(ctx);
}

extern "C" int dem_compute_forces(dem_ctx *ctx, int shearupdate)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!ctx->list_valid) return fail(ctx, DEM_ERR_ARG, "dem_compute_forces: no neighbour list (call dem_setup)");
    sync_enter(ctx);
    if (ctx->n == 0) return DEM_OK;
    int rc = force_pass(ctx, shearupdate);
    if (rc) return rc;
    return check_err_flag(ctx);
}

extern "C" int dem_setup(dem_ctx *ctx)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!ctx->mat_set) return fail(ctx, DEM_ERR_ARG, "dem_setup: set materials first");
    sync_enter(ctx);
    if (ctx->n == 0) {   // an empty box (a pour yet to start): nothing to build
        ctx->list_valid = true;
        ctx->setup_ts = ctx->ntimestep;
        return DEM_OK;
    }
    int rc = setup_mesh_grids(ctx);
    if (rc) return rc;
    rc = rebuild(ctx);
    if (rc) return rc;
    rc = force_pass(ctx, 0);
    if (rc) return rc;
    ctx->setup_ts = ctx->ntimestep;
    return check_err_flag(ctx);
}

// ---- particle insertion (fix insert/pack) ----

// the next step at which an insertion fix runs (LLONG_MAX: none)
static long long next_insert_step(const dem_ctx *ctx)
{
    long long s = LLONG_MAX;
    for (const auto &in : ctx->inserts)
        if (in->fix->next_step() > 0) s = std::min(s, in->fix->next_step());
    return s;
}

static bool insert_due(const dem_ctx *ctx)
{
    for (const auto &in : ctx->inserts)
        if (in->fix->due(ctx->ntimestep)) return true;
    return false;
}

// A buffer of `count` elements keeping its first `keep` (a null buffer is
// allocated).
template <typename T>
static int dev_grow(dem_ctx *ctx, T **ptr, size_t keep, size_t count)
{
    T *q = nullptr;
    HIPCHK(hipMalloc((void **)&q, std::max<size_t>(count, 1) * sizeof(T)));
    if (*ptr) {
        if (keep) HIPCHK(hipMemcpyAsync(q, *ptr, keep * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(*ptr));
    }
    *ptr = q;
    return DEM_OK;
}

// Every per-particle buffer to capacity c, contents kept.
static int grow_particles(dem_ctx *ctx, int c)
{
    const size_t oc = ctx->cap, nc = c;
    const size_t wdn = std::max(ctx->P.n * ctx->M.wdnum, 1);
    for (int b = 0; b < 2; b++) {
        if (dev_grow(ctx, &ctx->xr[b], oc, nc) || dev_grow(ctx, &ctx->vm[b], oc, nc)
            || dev_grow(ctx, &ctx->wt[b], oc, nc) || dev_grow(ctx, &ctx->tag[b], oc, nc)
            || dev_grow(ctx, &ctx->orig[b], oc, nc) || dev_grow(ctx, &ctx->img[b], oc, nc)
            || dev_grow(ctx, &ctx->whist[b], oc * wdn, nc * wdn)
            || dev_grow(ctx, &ctx->nstart[b], oc ? oc + 1 : 0, nc + 1) || dev_grow(ctx, &ctx->gmask[b], oc, nc))
            return DEM_ERR_HIP;
        if (ctx->xo[b] && dev_grow(ctx, &ctx->xo[b], oc, nc)) return DEM_ERR_HIP;
        if (ctx->brid[b] && dev_grow(ctx, &ctx->brid[b], oc, nc)) return DEM_ERR_HIP;
        if (ctx->rel[b] && dev_grow(ctx, &ctx->rel[b], 4 * oc, 4 * nc)) return DEM_ERR_HIP;
        if (ctx->mf.counter[b] && dev_grow(ctx, &ctx->mf.counter[b], oc, nc)) return DEM_ERR_HIP;
    }
    if (dev_grow(ctx, &ctx->f4, oc, nc) || dev_grow(ctx, &ctx->t4, oc, nc) || dev_grow(ctx, &ctx->xhold, oc, nc)
        || dev_grow(ctx, &ctx->wmask, oc, nc) || dev_grow(ctx, &ctx->ncontact, oc, nc)
        || dev_grow(ctx, &ctx->key, 0, nc) || dev_grow(ctx, &ctx->sorted, 0, nc)
        || dev_grow(ctx, &ctx->stage, 0, nc * kStageSlots) || dev_grow(ctx, &ctx->xn, oc, nc)
        || dev_grow(ctx, &ctx->vn, oc, nc) || dev_grow(ctx, &ctx->wn, oc, nc))
        return DEM_ERR_HIP;
    ctx->cap = c;
    return DEM_OK;
}

// ParticleToInsert::insert for m atoms: appended after the resident ones,
// which keep their slots, so the previous list and the mesh rows need only
// empty rows for the new atoms to carry every history at the next rebuild.
static int append_atoms(dem_ctx *ctx, int m, const double *x, const double *v, const double *omega,
                        const double *radius, const double *rmass, const int *type, const int *tag, const int *gm,
                        const double *brid = nullptr, const double *rel = nullptr)
{
    if (m == 0) return DEM_OK;
    if (m < 0 || !x || !v || !omega || !radius || !rmass || !type || !tag)
        return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: bad arguments");
    if (!ctx->mat_set) return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: set materials first");
    if (ctx->ninst > 1 || ctx->obs_plane >= 0 || ctx->window || ctx->tape_pos >= 0)
        return fail(ctx, DEM_ERR_ARG,
                    "dem_append_atoms: not in a batched run, an adjoint window or with an observed plane");
    const int n0 = ctx->n;
    if ((int64_t)n0 + m > (int64_t)kJMask) return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: too many atoms");
    const int n1 = n0 + m;
    std::vector<double4> hx(m), hv(m), hw(m);
    std::vector<int> ho(m), hg(m, 0);
    std::vector<int4> hi(m, make_int4(0, 0, 0, 0));
    double maxrad = ctx->maxrad;
    int tag_max = ctx->tag_max;
    for (int k = 0; k < m; k++) {
        if (type[k] < 1 || type[k] > ctx->ntypes) return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: atom type out of range");
        if (!(radius[k] > 0.0) || !(rmass[k] > 0.0))
            return fail(ctx, DEM_ERR_ARG, "Fix nve/sphere requires extended particles");
        if (tag[k] <= 0) return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: tags must be > 0");
        hx[k] = make_double4(x[3 * k], x[3 * k + 1], x[3 * k + 2], radius[k]);
        hv[k] = make_double4(v[3 * k], v[3 * k + 1], v[3 * k + 2], rmass[k]);
        hw[k] = make_double4(omega[3 * k], omega[3 * k + 1], omega[3 * k + 2], (double)type[k]);
        ho[k] = n0 + k;
        if (gm) hg[k] = gm[k];
        for (const auto &mv : ctx->moves)
            if ((hg[k] >> mv.bit) & 1)
                return fail(ctx, DEM_ERR_ARG, "dem_append_atoms: a new atom in the group of fix move");
        maxrad = std::max(maxrad, radius[k]);
        tag_max = std::max(tag_max, tag[k]);
    }
    if (n1 > ctx->cap) {
        const int c = (int)std::min<int64_t>((int64_t)kJMask, std::max<int64_t>(n1, (int64_t)ctx->cap * 3 / 2 + 1024));
        if (grow_particles(ctx, c)) return DEM_ERR_HIP;
    }
    hipStream_t s = ctx->stream;
    const int c = ctx->cur;
    const size_t wdn = ctx->P.n * ctx->M.wdnum;
    HIPCHK(hipMemcpyAsync(sx(ctx) + n0, hx.data(), m * sizeof(double4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sv(ctx) + n0, hv.data(), m * sizeof(double4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(sw(ctx) + n0, hw.data(), m * sizeof(double4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->tag[c] + n0, tag, m * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->orig[c] + n0, ho.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->img[c] + n0, hi.data(), m * sizeof(int4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->gmask[c] + n0, hg.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->xhold + n0, hx.data(), m * sizeof(double4), hipMemcpyHostToDevice, s));
    if (wdn) HIPCHK(hipMemsetAsync(ctx->whist[c] + n0 * wdn, 0, sizeof(double) * m * wdn, s));
    HIPCHK(hipMemsetAsync(ctx->f4 + n0, 0, m * sizeof(double4), s));
    HIPCHK(hipMemsetAsync(ctx->t4 + n0, 0, m * sizeof(double4), s));
    HIPCHK(hipMemsetAsync(ctx->wmask + n0, 0, m, s));
    // the counter property's default for a new atom: not seen yet
    if (ctx->mf.mesh >= 0) HIPCHK(hipMemsetAsync(ctx->mf.counter[c] + n0, 0xff, sizeof(int) * m, s));
    // bond_random_id: the clump's, or the property's default -1
    if (brid && !ctx->brid[0]) {
        for (int b = 0; b < 2; b++)
            if (dev_alloc(ctx, &ctx->brid[b], (size_t)ctx->cap)) return DEM_ERR_HIP;
        std::vector<double> minus(std::max(n0, 1), -1.0);
        if (n0) HIPCHK(hipMemcpyAsync(ctx->brid[c], minus.data(), n0 * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (ctx->brid[c]) {
        std::vector<double> hb(m, -1.0);
        if (brid) hb.assign(brid, brid + m);
        HIPCHK(hipMemcpyAsync(ctx->brid[c] + n0, hb.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    // release data: the stream's, or zero (the property's default)
    if (rel && !ctx->rel[0]) {
        for (int b = 0; b < 2; b++)
            if (dev_alloc(ctx, &ctx->rel[b], 4 * (size_t)ctx->cap)) return DEM_ERR_HIP;
        if (n0) HIPCHK(hipMemsetAsync(ctx->rel[c], 0, 4 * (size_t)n0 * sizeof(double4), s));
    }
    if (ctx->rel[c]) {
        if (rel) HIPCHK(hipMemcpyAsync(ctx->rel[c] + 4 * (size_t)n0, rel, 4 * (size_t)m * sizeof(double4),
                                       hipMemcpyHostToDevice, s));
        else HIPCHK(hipMemsetAsync(ctx->rel[c] + 4 * (size_t)n0, 0, 4 * (size_t)m * sizeof(double4), s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (ctx->have_list) launch_extend_rows(ctx->nstart[ctx->lcur], n0, m, s);
    for (auto &mh : ctx->meshes) {
        if (!mh.have) continue;
        const int b = mh.cur;
        if (n1 + 1 > mh.ncap[b]) {
            if (dev_grow(ctx, &mh.mstart[b], (size_t)n0 + 1, (size_t)n1 + 1)
                || dev_grow(ctx, &mh.mnp[b], (size_t)n0 + 1, (size_t)n1 + 1))
                return DEM_ERR_HIP;
            mh.ncap[b] = n1 + 1;
        }
        launch_extend_rows(mh.mstart[b], n0, m, s);
        launch_fill_int(mh.mnp[b] + n0, 0, m, s);
    }
    KCHK();
    HIPCHK(hipStreamSynchronize(s));
    ctx->n = n1;
    ctx->tag_max = tag_max;
    if (maxrad > ctx->maxrad) ctx->mesh_grids_valid = false;
    ctx->maxrad = maxrad;
    ctx->nc_fresh = false;
    // cell grid for the new largest radius, scan scratch for n1 + 1 rows
    return setup_grid(ctx);
}

// Resident particles into the overlap-test grid: cells at least R wide
static int ins_build_grid(dem_ctx *ctx, double R, const StreamNear &f)
{
    InsGrid &g = ctx->ins_grid;
    int64_t cells = 1;
    for (int d = 0; d < 3; d++) {
        const double L = ctx->p.boxhi[d] - ctx->p.boxlo[d];
        int nc = (int)std::floor(L / (R * 1.0001));
        g.n[d] = std::max(1, std::min(nc, 1 << 12));
    }
    for (;;) {
        cells = (int64_t)g.n[0] * g.n[1] * g.n[2];
        if (cells <= (1 << 22)) break;
        int dm = 0;
        for (int d = 1; d < 3; d++)
            if (g.n[d] > g.n[dm]) dm = d;
        g.n[dm] = (g.n[dm] + 1) / 2;
    }
    for (int d = 0; d < 3; d++) {
        const double L = ctx->p.boxhi[d] - ctx->p.boxlo[d];
        g.lo[d] = ctx->p.boxlo[d];
        g.inv[d] = g.n[d] / L;
        g.prd[d] = L;
        g.periodic[d] = ctx->p.periodic[d];
    }
    const int n = ctx->n;
    if (n > ctx->ins_ncap) {
        if (dev_alloc(ctx, &ctx->ins_cell, n) || dev_alloc(ctx, &ctx->ins_idx, n)) return DEM_ERR_HIP;
        ctx->ins_ncap = n;
    }
    if (cells + 1 > ctx->ins_ccap) {
        if (dev_alloc(ctx, &ctx->ins_start, 2 * (cells + 1))) return DEM_ERR_HIP;
        ctx->ins_ccap = (int)(cells + 1);
    }
    const int need = scan_tmp_size((int)cells + 1);
    if (need > ctx->ins_tcap) {
        if (dev_alloc(ctx, &ctx->ins_tmp, need)) return DEM_ERR_HIP;
        ctx->ins_tcap = need;
    }
    hipStream_t s = ctx->stream;
    int *start = ctx->ins_start, *cursor = ctx->ins_start + (cells + 1);
    launch_fill_int(start, 0, (int)cells + 1, s);
    launch_ins_bin(sx(ctx), n, g, f, ctx->ins_cell, start, s);
    launch_scan(start, (int)cells + 1, ctx->ins_tmp, s);
    launch_copy_int(cursor, start, (int)cells + 1, s);
    launch_ins_place(ctx->ins_cell, n, cursor, ctx->ins_idx, s);
    KCHK();
    ctx->ins_grid_ready = true;
    return DEM_OK;
}

// Modify::get_max_radius (modify_liggghts.cpp:575-620): atoms, templates
// and insertion fixes
static double max_radius_sim(const dem_ctx *ctx)
{
    double m = ctx->maxrad;
    for (const auto &t : ctx->templates)
        for (int j = 0; j < t->t.nspheres(); j++) m = std::max(m, t->t.radius(j));
    for (const auto &in : ctx->inserts) m = std::max(m, in->fix->max_radius());
    return m;
}

static dem_ins::Hooks ins_hooks(dem_ctx *ctx, const dem_ctx::Insert *in)
{
    dem_ins::Hooks h;
    h.maxrad_sim = max_radius_sim(ctx);
    h.region_sums = [ctx](const double *lo, const double *hi, long long *np, double *vol, double *mass) -> int {
        *np = 0;
        *vol = *mass = 0.0;
        if (ctx->n == 0) return DEM_OK;
        if (!ctx->ins_part && dev_alloc(ctx, &ctx->ins_part, 4 * kRegionBlocks)) return DEM_ERR_HIP;
        launch_region_sums(sx(ctx), sv(ctx), ctx->n, lo, hi, ctx->ins_part, ctx->stream);
        KCHK();
        std::vector<double> part(4 * kRegionBlocks);
        HIPCHK(hipMemcpyAsync(part.data(), ctx->ins_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        double c = 0.0;
        for (int b = 0; b < kRegionBlocks; b++) {
            c += part[4 * b];
            *vol += part[4 * b + 1];
            *mass += part[4 * b + 2];
        }
        *np = (long long)c;
        return DEM_OK;
    };
    h.overlap = [ctx, in](const std::vector<double> &cand, const std::vector<double> &rad,
                          const dem_ins::FaceFilter *near, std::vector<uint32_t> &hit) -> int {
        const int K = (int)(cand.size() / 3);
        hit.assign(K, 0u);
        if (ctx->n == 0 || K == 0) return DEM_OK;
        double rmax = 0.0;
        for (double r : rad) rmax = std::max(rmax, r);
        StreamNear f{};
        if (near) {
            f.on = 1;
            for (int d = 0; d < 3; d++) {
                f.pref[d] = near->pref[d];
                f.nrm[d] = near->nrm[d];
                f.bblo[d] = near->bblo[d];
                f.bbhi[d] = near->bbhi[d];
            }
            f.dmin = near->dmin;
            f.dmax = near->dmax;
            f.face = in->face_dev;
            f.ntri = (int)(near->face->size() / 9);
        }
        if (!ctx->ins_grid_ready && ins_build_grid(ctx, rmax + ctx->maxrad, f)) return DEM_ERR_HIP;
        if (K > ctx->ins_kcap) {
            if (dev_alloc(ctx, &ctx->ins_cand, 3 * (size_t)K) || dev_alloc(ctx, &ctx->ins_hit, K)) return DEM_ERR_HIP;
            ctx->ins_kcap = K;
        }
        if (!ctx->ins_rad && dev_alloc(ctx, &ctx->ins_rad, 32)) return DEM_ERR_HIP;
        hipStream_t s = ctx->stream;
        HIPCHK(hipMemcpyAsync(ctx->ins_cand, cand.data(), cand.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(ctx->ins_rad, rad.data(), rad.size() * sizeof(double), hipMemcpyHostToDevice, s));
        launch_ins_check(ctx->ins_cand, K, ctx->ins_rad, (int)rad.size(), sx(ctx), ctx->ins_start, ctx->ins_idx,
                         ctx->ins_grid, f, ctx->ins_hit, s);
        KCHK();
        HIPCHK(hipMemcpyAsync(hit.data(), ctx->ins_hit, K * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return DEM_OK;
    };
    return h;
}

// FixInsert::pre_exchange of every insertion fix due at this step, in the
// order they were added (each sees the atoms the earlier ones created)
static int run_inserts(dem_ctx *ctx)
{
    for (auto &in : ctx->inserts) {
        if (!in->fix->due(ctx->ntimestep)) continue;
        ctx->ins_grid_ready = false;
        std::vector<dem_ins::Placed> out;
        std::string err;
        const int rc = in->fix->insert(ins_hooks(ctx, in.get()), ctx->ntimestep, out, err);
        if (rc) return fail(ctx, rc, std::string(in->fix->is_stream() ? "fix insert/stream: " : "fix insert/pack: ") + err);
        const int m = (int)out.size();
        if (m == 0) continue;
        std::vector<double> x(3 * m), v(3 * m), w(3 * m), r(m), ms(m), br(m), rel;
        std::vector<int> ty(m), tg(m), gm(m, in->gmask);
        bool any_brid = false;
        if (out[0].has_release) rel.resize(16 * (size_t)m);
        for (int k = 0; k < m; k++) {
            if (!rel.empty()) {
                std::copy(out[k].release, out[k].release + 16, rel.begin() + 16 * (size_t)k);
                ctx->rel_until = std::max(ctx->rel_until, (long long)(int)(out[k].release[7] + 1e-14));
            }
            for (int d = 0; d < 3; d++) {
                x[3 * k + d] = out[k].x[d];
                v[3 * k + d] = out[k].v[d];
                w[3 * k + d] = out[k].omega[d];
            }
            r[k] = out[k].radius;
            ms[k] = out[k].mass;
            ty[k] = out[k].type;
            br[k] = out[k].brid;
            any_brid = any_brid || out[k].brid >= 0.0;
            // Atom::tag_extend (atom.cpp:574-595): after the largest tag
            tg[k] = ctx->tag_max + 1 + k;
        }
        const int a = append_atoms(ctx, m, x.data(), v.data(), w.data(), r.data(), ms.data(), ty.data(), tg.data(),
                                   gm.data(), any_brid ? br.data() : nullptr, rel.empty() ? nullptr : rel.data());
        if (a) return a;
    }
    ctx->ins_grid_ready = false;
    return DEM_OK;
}

extern "C" void dem_insert_pack_defaults(dem_insert_pack_spec *s)
{
    if (!s) return;
    *s = dem_insert_pack_spec{};
    s->target = DEM_INSERT_VOLUMEFRACTION;
    s->maxattempt = 50;
    s->insert_every = -1;
    s->start = 0;
    s->overlapcheck = 1;
    s->all_in = 0;
    s->check_dist_from_subdomain_border = 1;
    s->ntry_mc = 100000;
    s->exact_number = 1;
    s->vel_style = DEM_VEL_CONSTANT;
    s->orientation = DEM_ORIENT_TEMPLATE;
    s->quat[0] = 1.0;
}

extern "C" int dem_add_insert_pack(dem_ctx *ctx, const dem_insert_pack_spec *spec, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!spec || !id) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: bad arguments");
    if (!ctx->mat_set) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: set materials first");
    if (ctx->ninst > 1) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: not in a batched run");
    std::vector<dem_ins::TemplateMultisphere *> ms(std::max(spec->ntemplates, 0), nullptr);
    for (int t = 0; t < spec->ntemplates && spec->templates; t++) {
        const int mid = spec->multisphere ? spec->multisphere[t] : 0;
        if (mid) {
            for (auto &tp : ctx->templates)
                if (tp->id == mid) ms[t] = &tp->t;
            if (!ms[t]) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: no such multisphere template");
            for (int j = 0; j < ms[t]->nspheres(); j++)
                if (ms[t]->type(j) < 1 || ms[t]->type(j) > ctx->ntypes)
                    return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: template atom type out of range");
        } else if (spec->templates[t].atom_type < 1 || spec->templates[t].atom_type > ctx->ntypes) {
            return fail(ctx, DEM_ERR_ARG, "dem_add_insert_pack: template atom type out of range");
        }
    }
    auto in = std::make_unique<dem_ctx::Insert>();
    auto fix = std::make_unique<dem_ins::InsertPack>();
    const std::string err = fix->init(*spec, ctx->p.boxlo, ctx->p.boxhi, ctx->ntimestep, ms);
    if (!err.empty()) return fail(ctx, DEM_ERR_ARG, "fix insert/pack: " + err);
    in->fix = std::move(fix);
    in->id = ctx->next_op_id++;
    in->gmask = spec->group_mask;
    *id = in->id;
    ctx->inserts.push_back(std::move(in));
    return DEM_OK;
}

// the group whose atoms the releases move: the first stream fix's
// (FixInsertStream::init, i_am_integrator = first of its style); no stream
// fix left: nothing is released (the property goes with the last one,
// FixInsertStream::pre_delete)
static void stream_integrator(dem_ctx *ctx)
{
    for (const auto &in : ctx->inserts)
        if (in->fix->is_stream()) {
            ctx->rel_mask = in->gmask;
            return;
        }
    ctx->rel_until = -1;
}

extern "C" void dem_insert_stream_defaults(dem_insert_stream_spec *s)
{
    if (!s) return;
    *s = dem_insert_stream_spec{};
    s->insert_every = -1;
    s->maxattempt = 50;
    s->overlapcheck = 1;
    s->ntry_mc = 100000;
    s->exact_number = 1;
    s->vel_style = DEM_VEL_CONSTANT;
    s->orientation = DEM_ORIENT_TEMPLATE;
    s->quat[0] = 1.0;
}

extern "C" int dem_add_insert_stream(dem_ctx *ctx, const dem_insert_stream_spec *spec, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!spec || !id) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: bad arguments");
    if (!ctx->mat_set) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: set materials first");
    if (ctx->ninst > 1) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: not in a batched run");
    std::vector<dem_ins::TemplateMultisphere *> ms(std::max(spec->ntemplates, 0), nullptr);
    for (int t = 0; t < spec->ntemplates && spec->templates; t++) {
        const int mid = spec->multisphere ? spec->multisphere[t] : 0;
        if (mid) {
            for (auto &tp : ctx->templates)
                if (tp->id == mid) ms[t] = &tp->t;
            if (!ms[t]) return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: no such multisphere template");
            for (int j = 0; j < ms[t]->nspheres(); j++)
                if (ms[t]->type(j) < 1 || ms[t]->type(j) > ctx->ntypes)
                    return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: template atom type out of range");
        } else if (spec->templates[t].atom_type < 1 || spec->templates[t].atom_type > ctx->ntypes) {
            return fail(ctx, DEM_ERR_ARG, "dem_add_insert_stream: template atom type out of range");
        }
    }
    sync_enter(ctx);
    auto in = std::make_unique<dem_ctx::Insert>();
    auto fix = std::make_unique<dem_ins::InsertStream>();
    const std::string err = fix->init(*spec, ctx->p.boxlo, ctx->p.boxhi, ctx->ntimestep, ctx->p.dt, ms);
    if (!err.empty()) return fail(ctx, DEM_ERR_ARG, "fix insert/stream: " + err);
    const std::vector<double> &face = fix->face();
    if (dev_alloc(ctx, &in->face_dev, face.size())) return DEM_ERR_HIP;
    HIPCHK(hipMemcpy(in->face_dev, face.data(), face.size() * sizeof(double), hipMemcpyHostToDevice));
    in->fix = std::move(fix);
    in->id = ctx->next_op_id++;
    in->gmask = spec->group_mask;
    *id = in->id;
    ctx->inserts.push_back(std::move(in));
    stream_integrator(ctx);
    return DEM_OK;
}

extern "C" int dem_stream_release_data(dem_ctx *ctx, int ntags, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ntags < 0 || (ntags > 0 && !out)) return fail(ctx, DEM_ERR_ARG, "dem_stream_release_data: bad arguments");
    sync_enter(ctx);
    std::fill(out, out + 14 * (size_t)ntags, 0.0);
    const int n = ctx->n, c = ctx->cur;
    if (n == 0 || !ctx->rel[c]) return DEM_OK;
    std::vector<double4> hr(4 * (size_t)n);
    std::vector<int> tg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(hr.data(), ctx->rel[c], hr.size() * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tg.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        if (tg[i] < 1 || tg[i] > ntags) continue;
        double *o = out + 14 * (size_t)(tg[i] - 1);
        const double4 *r = &hr[4 * (size_t)i];
        const double v[14] = {r[0].x, r[0].y, r[0].z, r[0].w, r[1].w, r[1].x, r[1].y,
                              r[1].z, r[2].x, r[2].y, r[2].z, r[3].x, r[3].y, r[3].z};
        std::copy(v, v + 14, o);
    }
    return DEM_OK;
}

extern "C" int dem_insert_vector(dem_ctx *ctx, int id, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!out) return fail(ctx, DEM_ERR_ARG, "dem_insert_vector: null output");
    for (const auto &in : ctx->inserts)
        if (in->id == id) {
            in->fix->vector(out);
            return DEM_OK;
        }
    return fail(ctx, DEM_ERR_ARG, "dem_insert_vector: no such fix");
}

extern "C" int dem_append_atoms(dem_ctx *ctx, int m, const double *x, const double *v, const double *omega,
                                const double *radius, const double *rmass, const int *type, const int *tag,
                                const int *gmask)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    sync_enter(ctx);
    const int rc = append_atoms(ctx, m, x, v, omega, radius, rmass, type, tag, gmask);
    // the next step re-neighbours (Neighbor::decide sees the forced step)
    if (rc == DEM_OK && m > 0) ctx->next_reneighbor = ctx->ntimestep + 1;
    return rc;
}

extern "C" int dem_add_template_multisphere(dem_ctx *ctx, const dem_template_multisphere_spec *spec, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!spec || !id) return fail(ctx, DEM_ERR_ARG, "dem_add_template_multisphere: bad arguments");
    if (ctx->ntypes > 0) {
        if (spec->atom_type < 1 || spec->atom_type > ctx->ntypes)
            return fail(ctx, DEM_ERR_ARG, "dem_add_template_multisphere: atom type out of range");
        for (int j = 0; spec->types && j < spec->nspheres; j++)
            if (spec->types[j] < 1 || spec->types[j] > ctx->ntypes)
                return fail(ctx, DEM_ERR_ARG, "dem_add_template_multisphere: sphere type out of range");
    }
    sync_enter(ctx);
    // the inside tests of the Monte Carlo points, on the device
    auto inside = [ctx](const std::vector<double> &pts, const std::vector<double> &xs, const std::vector<double> &rs,
                        std::vector<uint8_t> &in) -> int {
        const int np = (int)(pts.size() / 3), ns = (int)rs.size();
        in.assign(np, 0);
        double *dp = nullptr, *dx = nullptr, *dr = nullptr;
        uint8_t *din = nullptr;
        hipStream_t s = ctx->stream;
        int rc = DEM_OK;
        auto run = [&]() -> int {
            HIPCHK(hipMalloc((void **)&dp, sizeof(double) * std::max<size_t>(pts.size(), 1)));
            HIPCHK(hipMalloc((void **)&dx, sizeof(double) * std::max<size_t>(xs.size(), 1)));
            HIPCHK(hipMalloc((void **)&dr, sizeof(double) * std::max<size_t>(rs.size(), 1)));
            HIPCHK(hipMalloc((void **)&din, std::max(np, 1)));
            HIPCHK(hipMemcpyAsync(dp, pts.data(), sizeof(double) * pts.size(), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(dx, xs.data(), sizeof(double) * xs.size(), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(dr, rs.data(), sizeof(double) * rs.size(), hipMemcpyHostToDevice, s));
            launch_tmpl_inside(dp, np, dx, dr, ns, din, s);
            KCHK();
            HIPCHK(hipMemcpyAsync(in.data(), din, np, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            return DEM_OK;
        };
        rc = run();
        hipStreamSynchronize(s);
        if (dp) hipFree(dp);
        if (dx) hipFree(dx);
        if (dr) hipFree(dr);
        if (din) hipFree(din);
        return rc;
    };
    auto tp = std::make_unique<dem_ctx::Template>();
    const std::string err = tp->t.init(*spec, inside);
    if (!err.empty()) return fail(ctx, DEM_ERR_ARG, "fix particletemplate/multiplespheres: " + err);
    tp->id = ctx->next_op_id++;
    *id = tp->id;
    ctx->templates.push_back(std::move(tp));
    return DEM_OK;
}

extern "C" int dem_template_info(dem_ctx *ctx, int id, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!out) return fail(ctx, DEM_ERR_ARG, "dem_template_info: null output");
    for (const auto &t : ctx->templates)
        if (t->id == id) {
            t->t.info(out);
            return DEM_OK;
        }
    return fail(ctx, DEM_ERR_ARG, "dem_template_info: no such template");
}

extern "C" int dem_create_particles_single(dem_ctx *ctx, int tmpl, const double *x, const double *v,
                                           const double *omega, const double *quat, int gmask)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!x) return fail(ctx, DEM_ERR_ARG, "dem_create_particles_single: null position");
    dem_ctx::Template *tp = nullptr;
    for (auto &t : ctx->templates)
        if (t->id == tmpl) tp = t.get();
    if (!tp) return fail(ctx, DEM_ERR_ARG, "dem_create_particles_single: no such template");
    // CreateParticles::add_single (create_particles.cpp:307-331): only a
    // point of the (single) sub-box creates
    for (int d = 0; d < 3; d++)
        if (!(x[d] >= ctx->p.boxlo[d] && x[d] < ctx->p.boxhi[d])) return 0;
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    if (quat) {
        // MathExtraLiggghts::quat_normalize (math_extra_liggghts.h:513-520)
        const double norm = 1.0 / std::sqrt(quat[0] * quat[0] + quat[1] * quat[1] + quat[2] * quat[2] + quat[3] * quat[3]);
        for (int k = 0; k < 4; k++) q[k] = quat[k] * norm;
    }
    const auto &t = tp->t;
    const int m = t.nspheres();
    std::vector<double> xs, vs(3 * (size_t)m), ws(3 * (size_t)m), r(m), ms(m);
    std::vector<int> ty(m), tg(m), gm(m, gmask);
    t.place(x, q, xs);
    for (int j = 0; j < m; j++) {
        for (int d = 0; d < 3; d++) {
            vs[3 * j + d] = v ? v[d] : 0.0;
            ws[3 * j + d] = omega ? omega[d] : 0.0;
        }
        r[j] = t.radius(j);
        // ParticleToInsert::insert (particleToInsert.cpp:129)
        ms[j] = 4.18879020479 * r[j] * r[j] * r[j] * t.density();
        ty[j] = t.type(j);
        tg[j] = ctx->tag_max + 1 + j;
    }
    // FixTemplateMultiplespheres::randomize_single draws the clump's
    // bond_random_id (fix_template_multiplespheres.cpp:600-615)
    std::vector<double> bid;
    if (t.bonded()) bid.assign(m, tp->t.bond_id(ctx->ntimestep));
    sync_enter(ctx);
    const int rc = append_atoms(ctx, m, xs.data(), vs.data(), ws.data(), r.data(), ms.data(), ty.data(), tg.data(),
                                gm.data(), bid.empty() ? nullptr : bid.data());
    if (rc) return rc;
    ctx->next_reneighbor = ctx->ntimestep + 1;
    return m;
}

extern "C" int dem_natoms(const dem_ctx *ctx)
{
    return ctx ? ctx->n : 0;
}

extern "C" int dem_download_props(dem_ctx *ctx, int *tag, int *type, double *radius, double *rmass)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur;
    if (n == 0) return DEM_OK;
    std::vector<int> ho(n), ht(n);
    std::vector<double4> hx(n), hv(n), hw(n);
    HIPCHK(hipMemcpy(ho.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ht.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hx.data(), sx(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv.data(), sv(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hw.data(), sw(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        const int o = ho[i];
        if (tag) tag[o] = ht[i];
        if (type) type[o] = (int)hw[i].w & 15;
        if (radius) radius[o] = hx[i].w;
        if (rmass) rmass[o] = hv[i].w;
    }
    return DEM_OK;
}

// Neighbor::decide + check_distance (neighbor.cpp:1362-1465); a fix that
// inserts at this step forces the build (fix_insert.cpp:369-370)
static int decide(dem_ctx *ctx, int *out)
{
    *out = 0;
    if (ctx->ntimestep == ctx->next_reneighbor || insert_due(ctx)) {
        *out = 1;
        return DEM_OK;
    }
    ctx->ago++;
    if (!(ctx->ago >= ctx->p.neigh_delay && ctx->ago % ctx->p.neigh_every == 0)) return DEM_OK;
    if (!ctx->p.neigh_check) {
        *out = 1;
        return DEM_OK;
    }
    if (ctx->n == 0) return DEM_OK;   // no atom moved
    const double trig = 0.5 * ctx->p.skin;
    HIPCHK(hipMemsetAsync(ctx->d_small + 2, 0, sizeof(int), ctx->stream));
    launch_check_distance(sx(ctx), ctx->xhold, ctx->n, trig * trig, ctx->d_small + 2, ctx->stream);
    KCHK();
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, ctx->d_small + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = h;
    return DEM_OK;
}

// ---- adjoint window helpers (the entry points are at the end of the file)

// layout of a state record: xr vm wt [f4 t4] hist nj whist
struct StateView {
    double4 *xr, *vm, *wt, *f4, *t4;
    double *hist;
    uint32_t *nj;
    double *whist;
};

static StateView state_view(const dem_ctx *ctx, char *base, bool forces)
{
    const size_t n = ctx->n;
    StateView v{};
    char *q = base;
    auto take = [&](size_t b) {
        char *r = q;
        q += (b + 255) & ~(size_t)255;
        return r;
    };
    v.xr = (double4 *)take(n * sizeof(double4));
    v.vm = (double4 *)take(n * sizeof(double4));
    v.wt = (double4 *)take(n * sizeof(double4));
    if (forces) {
        v.f4 = (double4 *)take(n * sizeof(double4));
        v.t4 = (double4 *)take(n * sizeof(double4));
    }
    v.hist = (double *)take((size_t)ctx->adj_E * ctx->M.dnum * sizeof(double));
    v.nj = (uint32_t *)take((size_t)ctx->adj_E * sizeof(uint32_t));
    v.whist = (double *)take(n * ctx->P.n * ctx->M.wdnum * sizeof(double));
    return v;
}

static size_t state_record_bytes(const dem_ctx *ctx, bool forces)
{
    const size_t n = ctx->n;
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return 3 * r(n * sizeof(double4)) + (forces ? 2 * r(n * sizeof(double4)) : 0)
           + r((size_t)ctx->adj_E * ctx->M.dnum * sizeof(double)) + r((size_t)ctx->adj_E * sizeof(uint32_t))
           + r(n * ctx->P.n * ctx->M.wdnum * sizeof(double));
}

static int copy_state(dem_ctx *ctx, char *rec, bool forces, bool save)
{
    StateView v = state_view(ctx, rec, forces);
    const size_t n = ctx->n;
    const int c = ctx->cur, l = ctx->lcur;
    hipStream_t s = ctx->stream;
    auto cp = [&](void *dev, void *r, size_t b) {
        if (!b) return hipSuccess;
        return save ? hipMemcpyAsync(r, dev, b, hipMemcpyDeviceToDevice, s)
                    : hipMemcpyAsync(dev, r, b, hipMemcpyDeviceToDevice, s);
    };
    hipError_t e = hipSuccess;
    auto acc = [&](hipError_t x) {
        if (x != hipSuccess) e = x;
    };
    acc(cp(sx(ctx), v.xr, n * sizeof(double4)));
    acc(cp(sv(ctx), v.vm, n * sizeof(double4)));
    acc(cp(sw(ctx), v.wt, n * sizeof(double4)));
    if (forces) {
        acc(cp(ctx->f4, v.f4, n * sizeof(double4)));
        acc(cp(ctx->t4, v.t4, n * sizeof(double4)));
    }
    acc(cp(ctx->hist[l], v.hist, (size_t)ctx->adj_E * ctx->M.dnum * sizeof(double)));
    acc(cp(ctx->nj[l], v.nj, (size_t)ctx->adj_E * sizeof(uint32_t)));
    acc(cp(ctx->whist[c], v.whist, n * ctx->P.n * ctx->M.wdnum * sizeof(double)));
    if (e != hipSuccess) return fail(ctx, DEM_ERR_HIP, std::string("adjoint state copy: ") + hipGetErrorString(e));
    return DEM_OK;
}

// where a rebuild would fall inside the window: the history carry-over of
// the rebuild on the unchanged list, and the distance check that keeps the
// list valid (flag in d_small[2])
static int window_mask(dem_ctx *ctx)
{
    AdjMaskArgs m{};
    m.xr = sx(ctx);
    m.nstart = ctx->nstart[ctx->lcur];
    m.nj = ctx->nj[ctx->lcur];
    m.hist = ctx->hist[ctx->lcur];
    m.n = ctx->n;
    m.dnum = ctx->M.dnum;
    m.cdf = ctx->p.cdf;
    for (int d = 0; d < 3; d++) m.prd[d] = ctx->M.prd[d];
    launch_adj_mask(m, ctx->stream);
    KCHK();
    const double trig = 0.5 * ctx->p.skin;
    launch_check_distance(sx(ctx), ctx->xhold, ctx->n, trig * trig, ctx->d_small + 2, ctx->stream);
    KCHK();
    ctx->ago = 0;
    return DEM_OK;
}

// the force-pass inputs of this step into tape slot tape_pos
static int tape_store(dem_ctx *ctx, int mask_step)
{
    if (ctx->tape_pos >= ctx->tape_cap) return fail(ctx, DEM_ERR_STATE, "adjoint tape overflow (internal error)");
    char *rec = ctx->tape + (size_t)ctx->tape_pos * state_record_bytes(ctx, false);
    const int rc = copy_state(ctx, rec, false, true);
    if (rc) return rc;
    ctx->tape_mask[ctx->tape_pos] = (uint8_t)(mask_step != 0);
    ctx->tape_pos++;
    return DEM_OK;
}

// Graph stepping applies when every rebuild falls on a fixed step: no
// distance check, delay within one period, nothing that needs the host
// between steps (meshes, bonds, timers).
static bool graph_ok(const dem_ctx *ctx)
{
    bool static_meshes = true;   // a moving mesh's displacement is a per-step host value
    for (const auto &m : ctx->meshes) static_meshes = static_meshes && !m.moving;
    return ctx->use_graph && !ctx->window && ctx->fops.empty() && ctx->moves.empty() && !ctx->p.neigh_check && ctx->p.neigh_delay <= ctx->p.neigh_every
           && static_meshes && ctx->M.cohesion == CM_OFF && !ctx->timing && ctx->n > 0;
}

// Appends the bytes of a by-value kernel argument source to a graph key.
static void key_bytes(std::vector<const void *> &k, const void *src, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(src);
    for (size_t o = 0; o < n; o += sizeof(intptr_t)) {
        intptr_t v = 0;
        memcpy(&v, b + o, std::min(sizeof(intptr_t), n - o));
        k.push_back((const void *)v);
    }
}

// What a recorded graph depends on: the buffers its kernels were given and
// every value they received by value (model, cell grid, planes, largest
// radius, list capacities, mesh grids).  A setter that changes any of them
// (gravity, freeze, materials, a new upload with other radii) makes the key
// differ, so the next graph period records again.
static std::vector<const void *> graph_key(const dem_ctx *ctx)
{
    std::vector<const void *> k = {ctx->xr[0], ctx->xr[1], ctx->vm[0], ctx->vm[1], ctx->wt[0], ctx->wt[1],
                                   ctx->xn, ctx->vn, ctx->wn, (const void *)(intptr_t)ctx->par,
                                   ctx->tag[0], ctx->tag[1], ctx->orig[0], ctx->orig[1], ctx->img[0],
                                   ctx->img[1], ctx->whist[0], ctx->whist[1], ctx->nstart[0], ctx->nstart[1],
                                   ctx->nj[0], ctx->nj[1], ctx->hist[0], ctx->hist[1], ctx->cstart,
                                   ctx->cursor, ctx->scan_tmp, ctx->spread, ctx->d_mat, ctx->pwf,
                                   ctx->rel[0], ctx->rel[1]};
    k.push_back((const void *)(intptr_t)(ctx->cur + 2 * ctx->lcur));
    k.push_back((const void *)(intptr_t)ctx->n);
    k.push_back((const void *)(intptr_t)(ctx->tune_lanes * 4096 + ctx->tune_block));
    k.push_back((const void *)(intptr_t)ctx->obs_plane);
    for (const auto &m : ctx->meshes) {
        for (int b = 0; b < 2; b++) {
            k.push_back(m.mstart[b]);
            k.push_back(m.mnp[b]);
            k.push_back(m.mcand[b]);
            k.push_back(m.mpart[b]);
            k.push_back(m.mhist[b]);
        }
        k.push_back(m.tri);
        k.push_back(m.gstart);
        k.push_back(m.glist);
        k.push_back((const void *)(intptr_t)(m.cur + 2 * (m.have ? 1 : 0)));
        k.push_back((const void *)(intptr_t)m.wtype);
        k.push_back(m.d_ms);
        k.push_back(m.d_sout);
        key_bytes(k, m.pref, sizeof(m.pref));
        key_bytes(k, m.grid.lo, sizeof(m.grid.lo));
        key_bytes(k, m.grid.inv, sizeof(m.grid.inv));
        key_bytes(k, m.grid.n, sizeof(m.grid.n));
    }
    k.push_back(ctx->mf.counter[0]);
    k.push_back(ctx->mf.counter[1]);
    k.push_back(ctx->mf.total);
    key_bytes(k, &ctx->p, sizeof(ctx->p));
    key_bytes(k, &ctx->M, sizeof(ctx->M));
    key_bytes(k, &ctx->G, sizeof(ctx->G));
    key_bytes(k, &ctx->P, sizeof(ctx->P));
    key_bytes(k, &ctx->maxrad, sizeof(ctx->maxrad));
    key_bytes(k, ctx->ecap, sizeof(ctx->ecap));
    k.push_back((const void *)(intptr_t)(ctx->ntypes + 4096 * ctx->mat_blocks));
    return k;
}

// Record 2*neigh_every steps, each: [rebuild on the period's last step]
// force pass, final_integrate(n) + initial_integrate(n+1).  The list
// buffers are sized first for twice the current list (and at least 32
// entries per particle); a longer list at replay raises an error.
static int capture_graph(dem_ctx *ctx)
{
    if (ctx->gexec) {
        HIPCHK(hipGraphExecDestroy(ctx->gexec));
        ctx->gexec = nullptr;
    }
    int64_t total = 0;
    HIPCHK(hipStreamSynchronize(ctx->stream));   // the last rebuild may still be running
    {
        int h = 0;
        HIPCHK(hipMemcpy(&h, ctx->nstart[ctx->lcur] + ctx->n, sizeof(int), hipMemcpyDeviceToHost));
        total = h;
    }
    const int64_t want = std::max<int64_t>(2 * total, (int64_t)ctx->n * 32);
    for (int b = 0; b < 2; b++) {
        if (ctx->ecap[b] >= want) continue;
        if (b == ctx->lcur) {
            // the current list: grow without losing it
            uint32_t *nj = nullptr;
            double *hi = nullptr;
            const size_t dn = ctx->M.dnum;
            HIPCHK(hipMalloc((void **)&nj, sizeof(uint32_t) * want));
            HIPCHK(hipMalloc((void **)&hi, sizeof(double) * want * dn));
            HIPCHK(hipMemcpy(nj, ctx->nj[b], sizeof(uint32_t) * total, hipMemcpyDeviceToDevice));
            HIPCHK(hipMemcpy(hi, ctx->hist[b], sizeof(double) * total * dn, hipMemcpyDeviceToDevice));
            HIPCHK(hipFree(ctx->nj[b]));
            HIPCHK(hipFree(ctx->hist[b]));
            ctx->nj[b] = nj;
            ctx->hist[b] = hi;
            ctx->ecap[b] = want;
        } else if (ensure_list_cap(ctx, b, want)) {
            return DEM_ERR_NOMEM;
        }
    }
    // mesh candidate rows: both buffers sized for twice the current rows
    const size_t dw = std::max(ctx->M.wdnum, 1);
    for (auto &m : ctx->meshes) {
        int h = 0;
        HIPCHK(hipMemcpy(&h, m.mstart[m.cur] + ctx->n, sizeof(int), hipMemcpyDeviceToHost));
        const int64_t mt = h;
        const int64_t mwant = std::max<int64_t>(2 * mt, (int64_t)ctx->n + 1024);
        for (int b = 0; b < 2; b++) {
            if (ctx->n + 1 > m.ncap[b]) {
                if (dev_alloc(ctx, &m.mstart[b], ctx->n + 1) || dev_alloc(ctx, &m.mnp[b], ctx->n + 1))
                    return DEM_ERR_HIP;
                m.ncap[b] = ctx->n + 1;
            }
            if (m.scap[b] >= mwant) continue;
            // the current rows are kept: grow by copy
            int *nc = nullptr, *np = nullptr;
            double *nh = nullptr;
            HIPCHK(hipMalloc((void **)&nc, sizeof(int) * mwant));
            HIPCHK(hipMalloc((void **)&np, sizeof(int) * mwant));
            HIPCHK(hipMalloc((void **)&nh, sizeof(double) * mwant * dw));
            if (b == m.cur && m.have && mt > 0) {
                HIPCHK(hipMemcpy(nc, m.mcand[b], sizeof(int) * mt, hipMemcpyDeviceToDevice));
                HIPCHK(hipMemcpy(np, m.mpart[b], sizeof(int) * mt, hipMemcpyDeviceToDevice));
                HIPCHK(hipMemcpy(nh, m.mhist[b], sizeof(double) * mt * dw, hipMemcpyDeviceToDevice));
            }
            if (m.mcand[b]) HIPCHK(hipFree(m.mcand[b]));
            if (m.mpart[b]) HIPCHK(hipFree(m.mpart[b]));
            if (m.mhist[b]) HIPCHK(hipFree(m.mhist[b]));
            m.mcand[b] = nc;
            m.mpart[b] = np;
            m.mhist[b] = nh;
            m.scap[b] = mwant;
        }
    }
    const int every = ctx->p.neigh_every;
    const int cur0 = ctx->cur, lcur0 = ctx->lcur, par0 = ctx->par;
    std::vector<std::pair<int, bool>> mesh0;
    for (const auto &m : ctx->meshes) mesh0.emplace_back(m.cur, m.have);
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->gexec_capturing = true;
    int rc = DEM_OK;
    const auto mf_keep = ctx->mf;   // capture records kernels; the host bookkeeping is redone per replay
    for (int k = 1; k <= 2 * every && rc == DEM_OK; k++) {
        if (ctx->mf.mesh >= 0) rc = massflow_step(ctx);
        if (rc == DEM_OK && k % every == 0) rc = rebuild(ctx, true);
        if (rc == DEM_OK) rc = force_pass(ctx, 1, 3);
    }
    ctx->mf = mf_keep;
    ctx->gexec_capturing = false;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    // a replay must leave the state in the buffers it found it in
    const bool same_par = ctx->par == par0;
    ctx->cur = cur0;
    ctx->lcur = lcur0;
    ctx->par = par0;
    for (size_t k = 0; k < ctx->meshes.size(); k++) {
        ctx->meshes[k].cur = mesh0[k].first;
        ctx->meshes[k].have = mesh0[k].second;
    }
    if (rc == DEM_OK && ec == hipSuccess && !same_par) {
        if (graph) hipGraphDestroy(graph);
        return fail(ctx, DEM_ERR_STATE, "graph stepping: the recorded period changes the state buffers");
    }
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(ctx, DEM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        ctx->gexec = nullptr;
        return fail(ctx, DEM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
    ctx->g_period = 2 * every;
    ctx->g_key = graph_key(ctx);
    return DEM_OK;
}

// ---- checked graph periods (neigh_modify check yes, every 1, delay 0)
//
// Neighbor::decide's distance check runs on the device inside the recorded
// steps: the first step whose check asks for a rebuild halts the period (the
// step's own kernels and every later one return at once), and the host,
// reading the halt word once per batch of periods, goes on eagerly from that
// step, whose check then asks again.  Rebuilds therefore never run inside
// such a graph, and a moving mesh's trigger is foreseen on the host (its
// motion is known), so no host round trip is needed per step.

static bool checked_graph_ok(const dem_ctx *ctx)
{
    if (!ctx->use_graph || ctx->window || !ctx->fops.empty() || !ctx->moves.empty() || !ctx->p.neigh_check
        || ctx->p.neigh_every != 1 || ctx->p.neigh_delay != 0 || ctx->mf.mesh >= 0 || ctx->timing || ctx->n == 0
        || ctx->tape_pos >= 0)
        return false;
    // a moving stress mesh's p_ref is a per-step kernel argument
    for (const auto &m : ctx->meshes)
        if (m.passive || (m.moving && m.stress)) return false;
    return true;
}

static std::vector<const void *> checked_key(const dem_ctx *ctx)
{
    std::vector<const void *> k = graph_key(ctx);
    k.push_back(ctx->ctl);
    k.push_back(ctx->xhold);
    k.push_back(ctx->f4);
    k.push_back(ctx->t4);
    k.push_back(ctx->brev);
    k.push_back(ctx->bout);
    k.push_back(ctx->bvalid);
    k.push_back(ctx->brid[0]);
    k.push_back(ctx->brid[1]);
    k.push_back((const void *)(intptr_t)ctx->par);
    for (const auto &m : ctx->meshes) {
        k.push_back((const void *)(intptr_t)m.moving);
        key_bytes(k, m.vel, sizeof(m.vel));
    }
    return k;
}

// Record P steps (P even), each: the distance check (which also advances
// the step counter), the force pass with final_integrate(n) +
// initial_integrate(n+1), the mesh move.  When the integrator has its own
// pass (no fused integration) that pass does the next step's check and the
// mesh moves instead (two launches fewer per step).
static int capture_checked_graph(dem_ctx *ctx, int P)
{
    int nmoving = 0;
    for (const auto &m : ctx->meshes) nmoving += m.moving ? 1 : 0;
    const bool fold = !fuse_possible(ctx) && nmoving <= kFoldMeshes;
    if (ctx->cgexec) {
        HIPCHK(hipGraphExecDestroy(ctx->cgexec));
        ctx->cgexec = nullptr;
    }
    if (!ctx->ctl && dev_alloc(ctx, &ctx->ctl, 3)) return DEM_ERR_HIP;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int cur0 = ctx->cur, lcur0 = ctx->lcur, par0 = ctx->par;
    const double trig = 0.5 * ctx->p.skin;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->gexec_capturing = true;
    ctx->halt_dev = ctx->ctl + 2;
    int rc = DEM_OK, flip = 0;
    for (int k = 0; k < P && rc == DEM_OK; k++) {
        ctx->step_dev = ctx->ctl + (k & 1);
        if (!fold) launch_check_halt(sx(ctx), ctx->xhold, ctx->n, trig * trig, ctx->ctl, k & 1, ctx->stream);
        const int pb = ctx->par;
        ctx->fold = fold;
        ctx->fold_par = k & 1;
        rc = force_pass(ctx, 1, 3);
        ctx->fold = false;
        if (k == 0) flip = ctx->par != pb;
        if (rc == DEM_OK && !fold) rc = move_meshes(ctx);
    }
    ctx->halt_dev = nullptr;
    ctx->step_dev = nullptr;
    ctx->fold = false;
    ctx->gexec_capturing = false;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    const bool same_par = ctx->par == par0;
    ctx->cur = cur0;
    ctx->lcur = lcur0;
    ctx->par = par0;
    if (rc == DEM_OK && ec == hipSuccess && !same_par) {
        if (graph) hipGraphDestroy(graph);
        return fail(ctx, DEM_ERR_STATE, "checked graph: the recorded period changes the state buffers");
    }
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(ctx, DEM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&ctx->cgexec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        ctx->cgexec = nullptr;
        return fail(ctx, DEM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
    ctx->cg_period = P;
    ctx->cg_flip = flip;
    ctx->cg_fold = fold;
    ctx->cg_key = checked_key(ctx);
    return DEM_OK;
}

// Of the next `room` steps, how many may run before a moving mesh's trigger
// forces a rebuild: the steps through the first one at which a node has
// moved more than half the skin since the last rebuild (its trigger makes
// the step after it rebuild).  The nodes advance as the host's copy will.
static long long mesh_horizon(const dem_ctx *ctx, long long room)
{
    const double trig = 0.25 * ctx->p.skin * ctx->p.skin;
    for (const auto &m : ctx->meshes) {
        if (!m.moving) continue;
        double dx[3];
        for (int q = 0; q < 3; q++) dx[q] = m.vel[q] * ctx->p.dt;
        std::vector<double> nd = m.h.node;   // the nodes of the next step
        for (long long k = 1; k <= room; k++) {
            if (k > 1)
                for (size_t q = 0; q < nd.size(); q++) nd[q] = nd[q] + dx[q % 3];
            bool hit = false;
            for (size_t q = 0; q < nd.size() / 3 && !hit; q++) {
                const double ex = nd[3 * q] - m.node_re[3 * q];
                const double ey = nd[3 * q + 1] - m.node_re[3 * q + 1];
                const double ez = nd[3 * q + 2] - m.node_re[3 * q + 2];
                hit = ex * ex + ey * ey + ez * ez > trig;
            }
            if (hit) {
                room = k;
                break;
            }
        }
    }
    return room;
}

// Whole checked periods from iteration st (step ntimestep + 1) on, at most
// up to the step before the last, the next insertion or forced rebuild.
// Advances st and ntimestep by the steps that completed.
static int checked_periods(dem_ctx *ctx, int nsteps, int &st)
{
    constexpr int P = 16;
    constexpr long long kMaxRun = 64 * P;   // steps per host read-back
    long long room = (long long)(nsteps - 1) - st;
    room = std::min(room, next_insert_step(ctx) - 1 - ctx->ntimestep);
    if (ctx->next_reneighbor > ctx->ntimestep) room = std::min(room, ctx->next_reneighbor - 1 - ctx->ntimestep);
    room = std::min(room, kMaxRun);
    if (room < P) return DEM_OK;
    room = mesh_horizon(ctx, room);
    const long long R = room / P;
    if (R < 1) return DEM_OK;
    if (!ctx->cgexec || ctx->cg_key != checked_key(ctx)) {
        const int rc = capture_checked_graph(ctx, P);
        if (rc) return rc;
    }
    const long long t0 = ctx->ntimestep;
    ctx->ctl_h[0] = t0 + 1;   // the first recorded step (P is even: every period starts at ctl[0])
    ctx->ctl_h[1] = 0;
    ctx->ctl_h[2] = 0;
    HIPCHK(hipMemcpyAsync(ctx->ctl, ctx->ctl_h, sizeof(ctx->ctl_h), hipMemcpyHostToDevice, ctx->stream));
    // a folded period checks each next step at the end of the step before:
    // the first step's check runs here
    if (ctx->cg_fold) {
        const double trig = 0.5 * ctx->p.skin;
        launch_check_halt(sx(ctx), ctx->xhold, ctx->n, trig * trig, ctx->ctl, 0, ctx->stream);
        KCHK();
    }
    for (long long r = 0; r < R; r++) HIPCHK(hipGraphLaunch(ctx->cgexec, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->ctl_h, ctx->ctl, sizeof(ctx->ctl_h), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const long long halt = ctx->ctl_h[2];
    const long long done = halt ? halt - t0 - 1 : R * P;
    // the host side of each completed step: Neighbor::decide's count, the
    // mesh trigger (foreseen: it can fire on the last one only), the mesh move
    for (long long k = 0; k < done; k++) {
        ctx->ntimestep++;
        ctx->ago++;
        mesh_trigger(ctx);
        move_meshes_host(ctx);
    }
    if (ctx->cg_flip && (done & 1)) ctx->par ^= 1;
    st += (int)done;
    return DEM_OK;
}

extern "C" int dem_step(dem_ctx *ctx, int nsteps)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!ctx->list_valid) return fail(ctx, DEM_ERR_ARG, "dem_step: call dem_setup first");
    if (nsteps <= 0) return DEM_OK;
    sync_enter(ctx);
    int rc;
    const bool gok = graph_ok(ctx);
    const bool cgok = !gok && checked_graph_ok(ctx);
    bool split_prev = false;   // the last step ran a stream release: no fused initial half
    for (int st = 0; st < nsteps; st++) {
        if (cgok && st >= 1 && !split_prev && ctx->ntimestep + 1 != ctx->next_reneighbor
            && ctx->ntimestep > ctx->rel_until) {
            if ((rc = checked_periods(ctx, nsteps, st))) return rc;
        }
        // whole graph periods while they fit before the last step, starting
        // right after a rebuild
        // an insertion step runs eagerly: graph periods end before it, and
        // none starts while an atom awaits a stream's release
        if (gok && st >= 1 && !split_prev && ctx->ago == 0 && ctx->ntimestep + 1 != ctx->next_reneighbor
            && st + 2 * ctx->p.neigh_every <= nsteps - 1
            && ctx->ntimestep + 2 * ctx->p.neigh_every < next_insert_step(ctx) && ctx->ntimestep > ctx->rel_until) {
            if (!ctx->gexec || ctx->g_key != graph_key(ctx)) {
                if ((rc = capture_graph(ctx))) return rc;
            }
            while (st + ctx->g_period <= nsteps - 1 && ctx->ntimestep + ctx->g_period < next_insert_step(ctx)) {
                HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
                ctx->ntimestep += ctx->g_period;
                st += ctx->g_period;
                if (ctx->mf.mesh >= 0) {
                    if (!ctx->mf.stepped) {
                        ctx->mf.mass_last = ctx->mf.mass_prev;
                        ctx->mf.n_last = ctx->mf.n_prev;
                        ctx->mf.stepped = true;
                    }
                    ctx->mf.t_since += ctx->g_period * ctx->p.dt;
                }
            }
            if ((rc = check_err_flag(ctx))) return rc;
            if (st >= nsteps) break;
        }
        ctx->ntimestep++;
        if (st == 0 && ((rc = integrate(ctx, 1)) || (rc = move_meshes(ctx)))) return rc;
        if (st > 0 && split_prev && ctx->n > 0 && (rc = integrate(ctx, 1))) return rc;
        if (ctx->mf.mesh >= 0 && (rc = massflow_step(ctx))) return rc;
        int nflag = 0;
        if ((rc = decide(ctx, &nflag))) return rc;
        // pre_exchange (fix insert) before the build, as Verlet::run
        // (verlet.cpp:304-327)
        if (nflag && !ctx->inserts.empty() && (rc = run_inserts(ctx))) return rc;
        if (nflag && ctx->n > 0 && (rc = ctx->window ? window_mask(ctx) : rebuild(ctx))) return rc;
        if (!nflag) mesh_trigger(ctx);
        if (ctx->tape_pos >= 0 && (rc = tape_store(ctx, nflag))) return rc;
        const bool last = st == nsteps - 1;
        // end_of_step of fix insert/stream after the final half: the
        // integrator runs un-fused on such steps
        const bool release = ctx->ntimestep <= ctx->rel_until && ctx->rel[ctx->cur];
        if (ctx->n > 0 && (rc = force_pass(ctx, 1, (last || release) ? 2 : 3))) return rc;
        if (release && ctx->n > 0) {
            ReleaseArgs ra{};
            const int c = ctx->cur;
            ra.xr = sx(ctx); ra.vm = sv(ctx); ra.wt = sw(ctx);
            ra.f4 = ctx->f4; ra.t4 = ctx->t4;
            ra.gm = ctx->gmask[c];
            ra.rel = ctx->rel[c];
            ra.n = ctx->n;
            ra.mask = ctx->rel_mask;
            ra.step = ctx->ntimestep;
            ra.dt = ctx->p.dt;
            launch_stream_release(ra, ctx->stream);
            KCHK();
        }
        split_prev = release;
        if (!last && (rc = move_meshes(ctx))) return rc;
    }
    return check_err_flag(ctx);
}

extern "C" int dem_download(dem_ctx *ctx, double *x, double *v, double *omega, double *f, double *torque)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur;
    if (n == 0) return DEM_OK;
    std::vector<double4> hx(n), hv(n), hw(n), hf(n), ht(n);
    std::vector<int> ho(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ho.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hx.data(), sx(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv.data(), sv(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hw.data(), sw(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hf.data(), ctx->f4, n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ht.data(), ctx->t4, n * sizeof(double4), hipMemcpyDeviceToHost));
    auto put = [&](double *dst, const std::vector<double4> &src) {
        if (!dst) return;
        for (int i = 0; i < n; i++) {
            const int o = ho[i];
            dst[3 * o] = src[i].x;
            dst[3 * o + 1] = src[i].y;
            dst[3 * o + 2] = src[i].z;
        }
    };
    put(x, hx); put(v, hv); put(omega, hw); put(f, hf); put(torque, ht);
    return DEM_OK;
}

extern "C" int dem_history_size(const dem_ctx *ctx)
{
    return ctx ? ctx->M.dnum : 0;
}

extern "C" int dem_wall_history_size(const dem_ctx *ctx)
{
    return ctx ? ctx->M.wdnum : 0;
}

extern "C" int dem_get_pairs(dem_ctx *ctx, int64_t *count, int *tag_i, int *tag_j, int *shift,
                             double *hist, int *flag)
{
    if (!ctx || !count) return fail(ctx, DEM_ERR_ARG, "dem_get_pairs: bad arguments");
    sync_enter(ctx);
    const int n = ctx->n, l = ctx->lcur, c = ctx->cur, dnum = ctx->M.dnum;
    if (!ctx->have_list || n == 0) {
        *count = 0;
        return DEM_OK;
    }
    std::vector<int> ns(n + 1), tg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ns.data(), ctx->nstart[l], (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    const int64_t E = ns[n];
    if (!tag_i) {
        *count = E;
        return DEM_OK;
    }
    if (*count < E) return fail(ctx, DEM_ERR_ARG, "dem_get_pairs: output arrays too small");
    std::vector<uint32_t> nj(E);
    std::vector<double> hh((size_t)E * dnum);
    HIPCHK(hipMemcpy(tg.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    if (E) {
        HIPCHK(hipMemcpy(nj.data(), ctx->nj[l], E * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hh.data(), ctx->hist[l], E * dnum * sizeof(double), hipMemcpyDeviceToHost));
    }
    // reference slot order: [hysteretic][cohesion][shear][rolling]; the
    // device keeps cohesion last, and only on the evaluating entry of a pair
    const Layout &L = ctx->M.lp;
    const int ncoh = L.off_coh >= 0 ? dnum - L.off_coh : 0;
    const int off_ref_coh = L.off_norm >= 0 ? 12 : 0;
    std::unordered_map<int64_t, int64_t> entry_of;
    if (ncoh)
        for (int i = 0; i < n; i++)
            for (int e = ns[i]; e < ns[i + 1]; e++)
                if (((nj[e] >> 26) & 31u) == (uint32_t)kNoShift)
                    entry_of[(int64_t)i * (int64_t)n + (int64_t)(nj[e] & kJMask)] = e;
    for (int i = 0; i < n; i++)
        for (int e = ns[i]; e < ns[i + 1]; e++) {
            const int j = (int)(nj[e] & kJMask);
            const int code = (int)((nj[e] >> 26) & 31u);
            tag_i[e] = tg[i];
            tag_j[e] = tg[j];
            if (shift) shift[e] = code;
            const bool known = (nj[e] & (kFlagBit | kBondBit)) != 0;
            if (flag) flag[e] = known ? 1 : 0;
            if (!hist) continue;
            double *out = hist + (size_t)e * dnum;
            const double *src = &hh[(size_t)e * dnum];
            // an entry without contact flags holds no contact history (zero
            // by definition); its cohesion slots are kept (zero, or what a
            // broken bond left until the next rebuild)
            int k = 0;
            for (int q = 0; q < off_ref_coh; q++) out[k++] = known ? src[q] : 0.0;
            if (ncoh) {
                const double *cs = src + L.off_coh;
                double sign = 1.0;
                if (code == kNoShift && tg[j] < tg[i]) {
                    // the partner's entry evaluates the bond: its copy with the
                    // newtonflag slots (ft, torque/angle vectors) negated
                    auto it = entry_of.find((int64_t)j * (int64_t)n + i);
                    cs = it != entry_of.end() ? &hh[(size_t)it->second * dnum + L.off_coh] : cs;
                    sign = -1.0;
                }
                for (int q = 0; q < ncoh; q++) out[k++] = (q >= 5 && q < 14) ? sign * cs[q] : cs[q];
            }
            for (int q = off_ref_coh; q < (L.off_coh >= 0 ? L.off_coh : dnum); q++) out[k++] = known ? src[q] : 0.0;
        }
    *count = E;
    return DEM_OK;
}

extern "C" int dem_set_history(dem_ctx *ctx, int64_t count, const int *tag_i, const int *tag_j,
                               const double *hist)
{
    if (!ctx || count < 0 || (count > 0 && (!tag_i || !tag_j || !hist)))
        return fail(ctx, DEM_ERR_ARG, "dem_set_history: bad arguments");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur, dnum = ctx->M.dnum;
    std::vector<int> tg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n) HIPCHK(hipMemcpy(tg.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    std::unordered_map<int, int> idx;
    idx.reserve(n * 2 + 1);
    for (int i = 0; i < n; i++) idx[tg[i]] = i;
    std::vector<int> cnt(n + 1, 0);
    for (int64_t k = 0; k < count; k++) {
        auto it = idx.find(tag_i[k]);
        if (it == idx.end() || idx.find(tag_j[k]) == idx.end())
            return fail(ctx, DEM_ERR_ARG, "dem_set_history: unknown tag");
        cnt[it->second + 1]++;
    }
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    std::vector<uint32_t> nj(std::max<int64_t>(count, 1));
    std::vector<double> hh(std::max<int64_t>(count, 1) * dnum);
    for (int64_t k = 0; k < count; k++) {
        const int i = idx[tag_i[k]];
        const int e = fillp[i]++;
        nj[e] = (uint32_t)idx[tag_j[k]] | ((uint32_t)kNoShift << 26) | kFlagBit;
        // reference slot order [hysteretic][cohesion][shear][rolling] to the
        // device's [hysteretic][shear][rolling][cohesion]
        const double *src = hist + (size_t)k * dnum;
        double *dst = &hh[(size_t)e * dnum];
        const Layout &L = ctx->M.lp;
        if (L.off_coh < 0) {
            for (int d = 0; d < dnum; d++) dst[d] = src[d];
        } else {
            const int ncoh = dnum - L.off_coh;
            const int off_ref_coh = L.off_norm >= 0 ? 12 : 0;
            int q = 0;
            for (int d = 0; d < off_ref_coh; d++) dst[q++] = src[d];
            for (int d = off_ref_coh + ncoh; d < dnum; d++) dst[q++] = src[d];
            for (int d = 0; d < ncoh; d++) dst[q++] = src[off_ref_coh + d];
            const double small = ctx->M.cohesion == CM_BOND ? 1e-6 : 1e-15;
            if (src[off_ref_coh] >= small) nj[e] |= kBondBit;
        }
    }
    const int l = ctx->lcur;
    if (ensure_list_cap(ctx, l, count)) return DEM_ERR_NOMEM;
    HIPCHK(hipMemcpy(ctx->nstart[l], cnt.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (count) {
        HIPCHK(hipMemcpy(ctx->nj[l], nj.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->hist[l], hh.data(), count * dnum * sizeof(double), hipMemcpyHostToDevice));
    }
    ctx->have_list = true;
    ctx->list_valid = false;
    return DEM_OK;
}

extern "C" int dem_get_wall_history(dem_ctx *ctx, double *hist)
{
    if (!ctx || !hist) return fail(ctx, DEM_ERR_ARG, "dem_get_wall_history: bad arguments");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur, wdn = ctx->P.n * ctx->M.wdnum;
    if (n == 0 || wdn == 0) return DEM_OK;
    std::vector<double> h((size_t)n * wdn);
    std::vector<int> ho(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(h.data(), ctx->whist[c], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ho.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < wdn; k++) hist[(size_t)ho[i] * wdn + k] = h[(size_t)i * wdn + k];
    return DEM_OK;
}

extern "C" int dem_set_wall_history(dem_ctx *ctx, const double *hist)
{
    if (!ctx || !hist) return fail(ctx, DEM_ERR_ARG, "dem_set_wall_history: bad arguments");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur, wdn = ctx->P.n * ctx->M.wdnum;
    if (n == 0 || wdn == 0) return DEM_OK;
    std::vector<double> h((size_t)n * wdn);
    std::vector<int> ho(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ho.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < wdn; k++) h[(size_t)i * wdn + k] = hist[(size_t)ho[i] * wdn + k];
    HIPCHK(hipMemcpy(ctx->whist[c], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return DEM_OK;
}

extern "C" int dem_get_mesh_contacts(dem_ctx *ctx, int mesh, int64_t *count, int *tag, int *tri, int *slot,
                                     double *hist)
{
    if (!ctx || !count || mesh < 0 || mesh >= (int)ctx->meshes.size())
        return fail(ctx, DEM_ERR_ARG, "dem_get_mesh_contacts: bad arguments");
    sync_enter(ctx);
    MeshDev &m = ctx->meshes[mesh];
    const int n = ctx->n, dw = ctx->M.wdnum;
    if (!m.have || n == 0) {
        *count = 0;
        return DEM_OK;
    }
    std::vector<int> ms(n + 1), tg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ms.data(), m.mstart[m.cur], (n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tg.data(), ctx->tag[ctx->cur], n * sizeof(int), hipMemcpyDeviceToHost));
    const int64_t S = ms[n];
    std::vector<int> pt(std::max<int64_t>(S, 1));
    std::vector<double> hh(std::max<int64_t>(S, 1) * std::max(dw, 1));
    if (S) {
        HIPCHK(hipMemcpy(pt.data(), m.mpart[m.cur], S * sizeof(int), hipMemcpyDeviceToHost));
        if (dw) HIPCHK(hipMemcpy(hh.data(), m.mhist[m.cur], S * dw * sizeof(double), hipMemcpyDeviceToHost));
    }
    int64_t c = 0;
    for (int i = 0; i < n; i++)
        for (int k = ms[i]; k < ms[i + 1]; k++) {
            if (pt[k] < 0) continue;
            if (tag) {
                if (c >= *count) return fail(ctx, DEM_ERR_ARG, "dem_get_mesh_contacts: output arrays too small");
                tag[c] = tg[i];
                if (tri) tri[c] = pt[k];
                if (slot) slot[c] = k - ms[i];
                if (hist)
                    for (int d = 0; d < dw; d++) hist[(size_t)c * dw + d] = hh[(size_t)k * dw + d];
            }
            c++;
        }
    *count = c;
    return DEM_OK;
}

extern "C" int dem_set_mesh_contacts(dem_ctx *ctx, int mesh, int64_t count, const int *tag, const int *tri,
                                     const int *slot, const double *hist)
{
    if (!ctx || mesh < 0 || mesh >= (int)ctx->meshes.size() || count < 0
        || (count > 0 && (!tag || !tri || !slot || (!hist && ctx->M.wdnum))))
        return fail(ctx, DEM_ERR_ARG, "dem_set_mesh_contacts: bad arguments");
    sync_enter(ctx);
    MeshDev &m = ctx->meshes[mesh];
    const int n = ctx->n, c = ctx->cur, dw = ctx->M.wdnum;
    std::vector<int> tg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n) HIPCHK(hipMemcpy(tg.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    std::unordered_map<int, int> idx;
    idx.reserve(2 * n + 1);
    for (int i = 0; i < n; i++) idx[tg[i]] = i;
    std::vector<int> len(n, 0), np(n, 0);
    for (int64_t k = 0; k < count; k++) {
        auto it = idx.find(tag[k]);
        if (it == idx.end() || slot[k] < 0 || slot[k] >= kMaxMeshSlots || tri[k] < 0 || tri[k] >= m.h.ntri)
            return fail(ctx, DEM_ERR_ARG, "dem_set_mesh_contacts: unknown tag, bad slot or triangle");
        len[it->second] = std::max(len[it->second], slot[k] + 1);
        np[it->second]++;
    }
    std::vector<int> st(n + 1, 0);
    for (int i = 0; i < n; i++) st[i + 1] = st[i] + len[i];
    const int64_t S = st[n];
    std::vector<int> pt(std::max<int64_t>(S, 1), -1);
    std::vector<double> hh(std::max<int64_t>(S, 1) * std::max(dw, 1), 0.0);
    for (int64_t k = 0; k < count; k++) {
        const int i = idx[tag[k]];
        const int64_t e = st[i] + slot[k];
        if (pt[e] != -1) return fail(ctx, DEM_ERR_ARG, "dem_set_mesh_contacts: slot given twice");
        pt[e] = tri[k];
        for (int d = 0; d < dw; d++) hh[(size_t)e * dw + d] = hist[(size_t)k * dw + d];
    }
    const int b = m.cur;
    if (n + 1 > m.ncap[b]) {
        if (dev_alloc(ctx, &m.mstart[b], n + 1) || dev_alloc(ctx, &m.mnp[b], n + 1)) return DEM_ERR_HIP;
        m.ncap[b] = n + 1;
    }
    if (S > m.scap[b]) {
        if (dev_alloc(ctx, &m.mcand[b], (size_t)S) || dev_alloc(ctx, &m.mpart[b], (size_t)S)
            || dev_alloc(ctx, &m.mhist[b], (size_t)S * std::max(dw, 1)))
            return DEM_ERR_NOMEM;
        m.scap[b] = S;
    }
    HIPCHK(hipMemcpy(m.mstart[b], st.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (n) HIPCHK(hipMemcpy(m.mnp[b], np.data(), n * sizeof(int), hipMemcpyHostToDevice));
    if (S) {
        HIPCHK(hipMemcpy(m.mpart[b], pt.data(), S * sizeof(int), hipMemcpyHostToDevice));
        if (dw) HIPCHK(hipMemcpy(m.mhist[b], hh.data(), S * dw * sizeof(double), hipMemcpyHostToDevice));
    }
    m.have = true;
    ctx->list_valid = false;
    return DEM_OK;
}

extern "C" int dem_timing_enable(dem_ctx *ctx, int on)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    hipStreamSynchronize(ctx->stream);
    for (Timer *t : {&ctx->t_contact, &ctx->t_integrate, &ctx->t_rebuild}) {
        timer_collect(*t);
        t->ms = 0.0;
        t->launches = 0;
    }
    ctx->timing = on != 0;
    return DEM_OK;
}

extern "C" int dem_timing_get(dem_ctx *ctx, const char *name, double *ms, int64_t *launches)
{
    if (!ctx || !name || !ms || !launches) return fail(ctx, DEM_ERR_ARG, "dem_timing_get: bad arguments");
    Timer *t = nullptr;
    if (!strcmp(name, "contact")) t = &ctx->t_contact;
    else if (!strcmp(name, "integrate")) t = &ctx->t_integrate;
    else if (!strcmp(name, "rebuild")) t = &ctx->t_rebuild;
    else return fail(ctx, DEM_ERR_ARG, "dem_timing_get: unknown timer");
    timer_collect(*t);
    *ms = t->ms;
    *launches = t->launches;
    return DEM_OK;
}

extern "C" int dem_list_stats(dem_ctx *ctx, double *mean_neigh, double *mean_contacts, int64_t *entries)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    hipSetDevice(ctx->p.device);
    const int n = ctx->n;
    if (!ctx->have_list || n == 0) return fail(ctx, DEM_ERR_ARG, "dem_list_stats: no list");
    int E = 0;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(&E, ctx->nstart[ctx->lcur] + n, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> nc(n);
    HIPCHK(hipMemcpy(nc.data(), ctx->ncontact, n * sizeof(int), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (int i = 0; i < n; i++) s += nc[i];
    if (mean_neigh) *mean_neigh = (double)E / n;
    if (mean_contacts) *mean_contacts = ctx->nc_fresh ? s / n : -1.0;
    if (entries) *entries = E;
    return DEM_OK;
}

extern "C" int dem_sync(dem_ctx *ctx)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

extern "C" int64_t dem_timestep(const dem_ctx *ctx)
{
    return ctx ? ctx->ntimestep : 0;
}

// ---------------------------------------------------------------------------
// Batched runs: independent instances in one context.  Every kernel launch
// covers all instances; the key space gives each instance its own cells, so
// no neighbour search crosses instances, and the sort leaves each instance's
// particles contiguous.

// ---- groups, fix setforce / fix aveforce, run boundaries

constexpr int kMaxForceOps = 16;

extern "C" int dem_set_group(dem_ctx *ctx, int bit, const int *members)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (bit < 0 || bit > 30 || (ctx->n > 0 && !members)) return fail(ctx, DEM_ERR_ARG, "dem_set_group: bad arguments");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur;
    if (n == 0) return DEM_OK;
    // members are in upload order; orig maps each device slot to it
    std::vector<int> og(n), gm(n);
    HIPCHK(hipMemcpy(og.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gm.data(), ctx->gmask[c], n * sizeof(int), hipMemcpyDeviceToHost));
    for (int q = 0; q < n; q++) {
        if (members[og[q]]) gm[q] |= 1 << bit;
        else gm[q] &= ~(1 << bit);
    }
    HIPCHK(hipMemcpy(ctx->gmask[c], gm.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return DEM_OK;
}

static int add_force_op(dem_ctx *ctx, dem_ctx::ForceOp o, int *id)
{
    if ((int)ctx->fops.size() >= kMaxForceOps) return fail(ctx, DEM_ERR_ARG, "at most 16 setforce/aveforce fixes");
    if (o.bit < 0 || o.bit > 30) return fail(ctx, DEM_ERR_ARG, "group bit out of range");
    sync_enter(ctx);
    if (!ctx->d_gtot) {
        if (dev_alloc(ctx, &ctx->d_gtot, 4 * kMaxForceOps) || dev_alloc(ctx, &ctx->d_gpart, 4 * kGroupBlocks))
            return DEM_ERR_NOMEM;
        HIPCHK(hipMemset(ctx->d_gtot, 0, 4 * kMaxForceOps * sizeof(double)));
    }
    std::vector<bool> used(kMaxForceOps, false);
    for (const auto &q : ctx->fops) used[q.slot] = true;
    o.slot = 0;
    while (used[o.slot]) o.slot++;
    HIPCHK(hipMemset(ctx->d_gtot + 4 * o.slot, 0, 4 * sizeof(double)));
    o.id = ctx->next_op_id++;
    ctx->fops.push_back(o);
    if (id) *id = o.id;
    return DEM_OK;
}

extern "C" int dem_add_setforce(dem_ctx *ctx, int bit, const int *style, const double *value, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!style || !value) return fail(ctx, DEM_ERR_ARG, "dem_add_setforce: bad arguments");
    dem_ctx::ForceOp o{};
    o.kind = 0;
    o.bit = bit;
    for (int d = 0; d < 3; d++) {
        o.style[d] = style[d] ? 1 : 0;
        o.val[d] = value[d];
    }
    return add_force_op(ctx, o, id);
}

extern "C" int dem_add_aveforce(dem_ctx *ctx, int bit, const int *style, double xvalue, double yvalue, double c0,
                                double t0, double c1, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!style) return fail(ctx, DEM_ERR_ARG, "dem_add_aveforce: bad arguments");
    dem_ctx::ForceOp o{};
    o.kind = 1;
    o.bit = bit;
    for (int d = 0; d < 3; d++) o.style[d] = style[d] ? 1 : 0;
    o.val[0] = xvalue;
    o.val[1] = yvalue;
    o.c0 = c0;
    o.t0 = t0;
    o.c1 = c1;
    return add_force_op(ctx, o, id);
}

extern "C" int dem_unfix(dem_ctx *ctx, int id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    for (size_t k = 0; k < ctx->fops.size(); k++)
        if (ctx->fops[k].id == id) {
            ctx->fops.erase(ctx->fops.begin() + k);
            return DEM_OK;
        }
    for (size_t k = 0; k < ctx->moves.size(); k++)
        if (ctx->moves[k].id == id) {
            ctx->moves.erase(ctx->moves.begin() + k);
            return DEM_OK;
        }
    for (size_t k = 0; k < ctx->inserts.size(); k++)
        if (ctx->inserts[k]->id == id) {
            ctx->inserts.erase(ctx->inserts.begin() + k);
            stream_integrator(ctx);
            return DEM_OK;
        }
    return fail(ctx, DEM_ERR_ARG, "dem_unfix: no such fix");
}

extern "C" int dem_mesh_stress(dem_ctx *ctx, int mesh, const double *pref)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (mesh < 0 || mesh >= (int)ctx->meshes.size()) return fail(ctx, DEM_ERR_ARG, "dem_mesh_stress: no such mesh");
    MeshDev &m = ctx->meshes[mesh];
    sync_enter(ctx);
    if (!m.d_spart && (dev_alloc(ctx, &m.d_spart, 6 * kSumBlocks) || dev_alloc(ctx, &m.d_sout, 6)))
        return DEM_ERR_HIP;
    HIPCHK(hipMemset(m.d_sout, 0, 6 * sizeof(double)));
    m.stress = true;
    for (int q = 0; q < 3; q++) m.pref[q] = pref ? pref[q] : 0.0;
    return DEM_OK;
}

extern "C" int dem_mesh_stress_vector(dem_ctx *ctx, int mesh, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (mesh < 0 || mesh >= (int)ctx->meshes.size() || !out || !ctx->meshes[mesh].stress)
        return fail(ctx, DEM_ERR_ARG, "dem_mesh_stress_vector: no stress tracking on that mesh");
    MeshDev &m = ctx->meshes[mesh];
    sync_enter(ctx);
    HIPCHK(hipMemcpyAsync(out, m.d_sout, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 3; q++) out[6 + q] = m.pref[q];
    return DEM_OK;
}

// FixMassflowMesh::post_integrate of one step, on the face mesh's rows of
// the last rebuild (none before the first)
static int massflow_step(dem_ctx *ctx)
{
    auto &f = ctx->mf;
    const MeshDev &m = ctx->meshes[f.mesh];
    // the rate bookkeeping (fix_massflow_mesh.cpp:398-405): the first step
    // after a read of the vector starts the next interval
    f.t_since += ctx->p.dt;
    if (!f.stepped) {
        f.mass_last = f.mass_prev;
        f.n_last = f.n_prev;
        f.stepped = true;
    }
    if (!m.have) return DEM_OK;
    MassflowArgs a{};
    a.xr = sx(ctx);
    a.vm = sv(ctx);
    a.tri = m.tri;
    a.mstart = m.mstart[m.cur];
    a.mcand = m.mcand[m.cur];
    a.counter = f.counter[ctx->cur];
    a.gmask = f.gbit >= 0 ? ctx->gmask[ctx->cur] : nullptr;
    a.gbit = f.gbit;
    a.n = ctx->n;
    a.once = f.once;
    for (int q = 0; q < 3; q++) {
        a.pref[q] = f.pref[q];
        a.nvec[q] = f.nvec[q];
    }
    a.precision = m.h.precision;
    a.part = f.part;
    a.total = f.total;
    launch_massflow(a, ctx->stream);
    KCHK();
    return DEM_OK;
}

static int massflow_counters(dem_ctx *ctx)
{
    auto &f = ctx->mf;
    for (int b = 0; b < 2; b++)
        if (dev_alloc(ctx, &f.counter[b], std::max(ctx->cap, 1))) return DEM_ERR_HIP;
    // -1: not seen yet (the counter property's default, fix_massflow_mesh.cpp:240-250)
    if (ctx->n > 0) HIPCHK(hipMemset(f.counter[ctx->cur], 0xff, sizeof(int) * ctx->n));
    return DEM_OK;
}

extern "C" int dem_add_massflow(dem_ctx *ctx, int ntri, const double *nodes, const double *vec_side, int once,
                                int group_bit)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ctx->mf.mesh >= 0) return fail(ctx, DEM_ERR_ARG, "dem_add_massflow: one fix massflow/mesh per context");
    if (!vec_side || (vec_side[0] == 0.0 && vec_side[1] == 0.0 && vec_side[2] == 0.0))
        return fail(ctx, DEM_ERR_ARG, "vec_side > 0 required");
    if (group_bit < -1 || group_bit > 30) return fail(ctx, DEM_ERR_ARG, "dem_add_massflow: bad group bit");
    const int k = dem_add_mesh(ctx, ntri, nodes, 1, 0.0, nullptr);
    if (k < 0) return k;
    MeshDev &m = ctx->meshes[k];
    m.passive = true;
    auto &f = ctx->mf;
    // FixMassflowMesh::setup: a planar face (TriMesh::isPlanar); setRefPoint
    // (fix_massflow_mesh.cpp:669-681): node 0 and the normal of element 0,
    // turned towards vec_side
    const double *n0 = &m.h.snorm[0];
    for (int t = 1; t < m.h.ntri; t++) {
        const double *nt = &m.h.snorm[3 * t];
        const double d = n0[0] * nt[0] + n0[1] * nt[1] + n0[2] * nt[2];
        if (std::fabs(std::fabs(d) - 1.0) > 1e-8) {
            ctx->meshes.pop_back();
            return fail(ctx, DEM_ERR_ARG, "fix massflow/mesh requires a planar face mass flow measurement");
        }
    }
    for (int q = 0; q < 3; q++) {
        f.pref[q] = m.h.node[q];
        f.nvec[q] = n0[q];
    }
    const double dot = f.nvec[0] * vec_side[0] + f.nvec[1] * vec_side[1] + f.nvec[2] * vec_side[2];
    if (std::fabs(dot) < 1e-6) {
        ctx->meshes.pop_back();
        return fail(ctx, DEM_ERR_ARG, "need to change 'vec_side', it is currently in or to close to the mesh plane");
    }
    if (dot < 0.0)
        for (int q = 0; q < 3; q++) f.nvec[q] = -f.nvec[q];
    f.mesh = k;
    f.once = once ? 1 : 0;
    f.gbit = group_bit;
    if (dev_alloc(ctx, &f.part, 2 * kMassflowBlocks) || dev_alloc(ctx, &f.total, 2)) return DEM_ERR_HIP;
    HIPCHK(hipMemset(f.total, 0, 2 * sizeof(double)));
    return massflow_counters(ctx);
}

extern "C" int dem_massflow_vector(dem_ctx *ctx, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ctx->mf.mesh < 0 || !out) return fail(ctx, DEM_ERR_ARG, "dem_massflow_vector: no fix massflow/mesh");
    auto &f = ctx->mf;
    sync_enter(ctx);
    double tot[2];
    HIPCHK(hipMemcpyAsync(tot, f.total, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // FixMassflowMesh::compute_vector (fix_massflow_mesh.cpp:637-662)
    if (f.stepped) {
        f.delta_t = f.t_since;
        f.t_since = 0.0;
        f.stepped = false;
    }
    out[0] = tot[0];
    out[1] = tot[1];
    out[2] = f.delta_t == 0.0 ? 0.0 : (tot[0] - f.mass_last) / f.delta_t;
    out[3] = f.delta_t == 0.0 ? 0.0 : (tot[1] - f.n_last) / f.delta_t;
    f.mass_prev = tot[0];
    f.n_prev = tot[1];
    return DEM_OK;
}

// ---- restart files ----
// A binary restart of the device state: header, the atoms (upload order),
// each atom's contact-history record exactly as FixContactHistory::
// pack_restart writes it (fix_contact_history.cpp:588-602: record size,
// partner count, then per partner its tag and dnum values, the integers
// stored in the bits of a double as LAMMPS' ubuf does), the primitive-wall
// history per atom (property/atom history_<id>) and the mesh contact slots.
namespace {
const char kRestartMagic[8] = {'L', 'I', 'G', 'D', 'E', 'M', 'R', '1'};
struct RestartHeader {
    char magic[8];
    int32_t version, n, dnum, wdn, nmesh, dw;
    int64_t ntimestep, atimestep;
    double atime;
};
double ubuf_int(int v)
{
    union { double d; int64_t i; } u;
    u.i = v;
    return u.d;
}
int ubuf_to_int(double d)
{
    union { double d; int64_t i; } u;
    u.d = d;
    return (int)u.i;
}
}  // namespace

extern "C" int dem_write_restart(dem_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, DEM_ERR_ARG, "dem_write_restart: bad arguments");
    if (!ctx->list_valid) return fail(ctx, DEM_ERR_ARG, "dem_write_restart: call dem_setup first");
    sync_enter(ctx);
    const int n = ctx->n, c = ctx->cur, dnum = ctx->M.dnum;
    const int dw = ctx->M.wdnum, wdn = ctx->P.n * dw;
    std::vector<double4> hx(n), hv(n), hw(n);
    std::vector<int> ho(n), htag(n), hgm(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ho.data(), ctx->orig[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(htag.data(), ctx->tag[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hgm.data(), ctx->gmask[c], n * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hx.data(), sx(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hv.data(), sv(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hw.data(), sw(ctx), n * sizeof(double4), hipMemcpyDeviceToHost));
    // atoms in upload order
    std::vector<int32_t> ai(4 * (size_t)n);   // tag, type, freeze, group bits
    std::vector<double> ad(11 * (size_t)n);   // x, v, omega, radius, rmass
    for (int q = 0; q < n; q++) {
        const int o = ho[q];
        const int meta = (int)hw[q].w;
        ai[4 * o] = htag[q];
        ai[4 * o + 1] = meta_type(meta);
        ai[4 * o + 2] = meta_frozen(meta) ? 1 : 0;
        ai[4 * o + 3] = hgm[q];
        const double v[11] = {hx[q].x, hx[q].y, hx[q].z, hv[q].x, hv[q].y, hv[q].z, hw[q].x, hw[q].y, hw[q].z,
                              hx[q].w, hv[q].w};
        for (int k = 0; k < 11; k++) ad[11 * (size_t)o + k] = v[k];
    }
    // contact history, per atom in upload order
    int64_t E = 0;
    int rc = dem_get_pairs(ctx, &E, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<int> ti(E), tj(E), sh(E), fl(E);
    std::vector<double> hh((size_t)E * dnum);
    if (E && (rc = dem_get_pairs(ctx, &E, ti.data(), tj.data(), sh.data(), hh.data(), fl.data()))) return rc;
    std::unordered_map<int, int> slot_of_tag;
    for (int o = 0; o < n; o++) slot_of_tag[ai[4 * o]] = o;
    std::vector<std::vector<int64_t>> touch(n);
    for (int64_t e = 0; e < E; e++)
        if (fl[e]) touch[slot_of_tag[ti[e]]].push_back(e);
    std::vector<double> rec;
    for (int o = 0; o < n; o++) {
        const int np = (int)touch[o].size();
        rec.push_back((double)((dnum + 1) * np + 2));
        rec.push_back(ubuf_int(np));
        for (int64_t e : touch[o]) {
            rec.push_back(ubuf_int(tj[e]));
            for (int d = 0; d < dnum; d++) rec.push_back(hh[(size_t)e * dnum + d]);
        }
    }
    std::vector<double> whh((size_t)n * wdn);
    if (wdn && (rc = dem_get_wall_history(ctx, whh.data()))) return rc;
    RestartHeader h{};
    memcpy(h.magic, kRestartMagic, 8);
    h.version = 1;
    h.n = n;
    h.dnum = dnum;
    h.wdn = wdn;
    h.nmesh = (int32_t)ctx->meshes.size();
    h.dw = dw;
    h.ntimestep = ctx->ntimestep;
    h.atimestep = ctx->atimestep;
    h.atime = ctx->atime;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(ctx, DEM_ERR_ARG, std::string("dem_write_restart: cannot open ") + path);
    bool ok = fwrite(&h, sizeof(h), 1, fp) == 1;
    ok = ok && fwrite(ai.data(), sizeof(int32_t), ai.size(), fp) == ai.size();
    ok = ok && fwrite(ad.data(), sizeof(double), ad.size(), fp) == ad.size();
    const int64_t nrec = (int64_t)rec.size();
    ok = ok && fwrite(&nrec, sizeof(nrec), 1, fp) == 1;
    ok = ok && (nrec == 0 || fwrite(rec.data(), sizeof(double), rec.size(), fp) == rec.size());
    ok = ok && (whh.empty() || fwrite(whh.data(), sizeof(double), whh.size(), fp) == whh.size());
    for (int k = 0; k < h.nmesh && ok; k++) {
        int64_t cnt = 0;
        if ((rc = dem_get_mesh_contacts(ctx, k, &cnt, nullptr, nullptr, nullptr, nullptr))) {
            fclose(fp);
            return rc;
        }
        std::vector<int> mt(cnt), mtri(cnt), ms(cnt);
        std::vector<double> mh((size_t)cnt * std::max(dw, 1));
        if (cnt && (rc = dem_get_mesh_contacts(ctx, k, &cnt, mt.data(), mtri.data(), ms.data(), mh.data()))) {
            fclose(fp);
            return rc;
        }
        ok = fwrite(&cnt, sizeof(cnt), 1, fp) == 1;
        for (int64_t q = 0; q < cnt && ok; q++) {
            const int32_t tri3[3] = {mt[q], mtri[q], ms[q]};
            ok = fwrite(tri3, sizeof(int32_t), 3, fp) == 3;
        }
        ok = ok && (cnt == 0 || fwrite(mh.data(), sizeof(double), (size_t)cnt * dw, fp) == (size_t)cnt * dw);
    }
    ok = (fclose(fp) == 0) && ok;
    return ok ? DEM_OK : fail(ctx, DEM_ERR_ARG, std::string("dem_write_restart: write failed: ") + path);
}

extern "C" int dem_read_restart(dem_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, DEM_ERR_ARG, "dem_read_restart: bad arguments");
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(ctx, DEM_ERR_ARG, std::string("dem_read_restart: cannot open ") + path);
    RestartHeader h{};
    auto bad = [&](const char *what) {
        fclose(fp);
        return fail(ctx, DEM_ERR_ARG, std::string("dem_read_restart: ") + what);
    };
    if (fread(&h, sizeof(h), 1, fp) != 1 || memcmp(h.magic, kRestartMagic, 8) != 0 || h.version != 1)
        return bad("not a restart file of this engine");
    // "saved simulation state used different contact history model - can not
    // restart" (fix_contact_history.cpp:580-581)
    if (h.dnum != ctx->M.dnum || h.dw != ctx->M.wdnum || h.wdn != ctx->P.n * ctx->M.wdnum
        || h.nmesh != (int32_t)ctx->meshes.size())
        return bad("saved simulation state used different contact history model or walls - can not restart");
    const int n = h.n, dnum = h.dnum;
    std::vector<int32_t> ai(4 * (size_t)n);
    std::vector<double> ad(11 * (size_t)n);
    if (fread(ai.data(), sizeof(int32_t), ai.size(), fp) != ai.size()
        || fread(ad.data(), sizeof(double), ad.size(), fp) != ad.size())
        return bad("truncated atoms");
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), w(3 * (size_t)n), r(n), m(n);
    std::vector<int> type(n), tag(n), frz(n), grp(n);
    for (int o = 0; o < n; o++) {
        tag[o] = ai[4 * o];
        type[o] = ai[4 * o + 1];
        frz[o] = ai[4 * o + 2];
        grp[o] = ai[4 * o + 3];
        for (int k = 0; k < 3; k++) {
            x[3 * o + k] = ad[11 * (size_t)o + k];
            v[3 * o + k] = ad[11 * (size_t)o + 3 + k];
            w[3 * o + k] = ad[11 * (size_t)o + 6 + k];
        }
        r[o] = ad[11 * (size_t)o + 9];
        m[o] = ad[11 * (size_t)o + 10];
    }
    int64_t nrec = 0;
    if (fread(&nrec, sizeof(nrec), 1, fp) != 1 || nrec < 0) return bad("truncated history");
    std::vector<double> rec(nrec);
    if (nrec && fread(rec.data(), sizeof(double), nrec, fp) != (size_t)nrec) return bad("truncated history");
    std::vector<double> whh((size_t)n * h.wdn);
    if (!whh.empty() && fread(whh.data(), sizeof(double), whh.size(), fp) != whh.size())
        return bad("truncated wall history");
    struct MeshRec { std::vector<int> tag, tri, slot; std::vector<double> hist; };
    std::vector<MeshRec> mr(h.nmesh);
    for (int k = 0; k < h.nmesh; k++) {
        int64_t cnt = 0;
        if (fread(&cnt, sizeof(cnt), 1, fp) != 1 || cnt < 0) return bad("truncated mesh contacts");
        mr[k].tag.resize(cnt);
        mr[k].tri.resize(cnt);
        mr[k].slot.resize(cnt);
        mr[k].hist.resize((size_t)cnt * h.dw);
        for (int64_t q = 0; q < cnt; q++) {
            int32_t t3[3];
            if (fread(t3, sizeof(int32_t), 3, fp) != 3) return bad("truncated mesh contacts");
            mr[k].tag[q] = t3[0];
            mr[k].tri[q] = t3[1];
            mr[k].slot[q] = t3[2];
        }
        if (cnt && fread(mr[k].hist.data(), sizeof(double), mr[k].hist.size(), fp) != mr[k].hist.size())
            return bad("truncated mesh contacts");
    }
    fclose(fp);
    int rc = dem_upload_state(ctx, n, x.data(), v.data(), w.data(), r.data(), m.data(), type.data(), tag.data(),
                              frz.data());
    if (rc) return rc;
    // group bits, in upload order (the device permutes them with the atoms)
    for (int b = 0; b < 31; b++) {
        bool any = false;
        std::vector<int> mem(n);
        for (int o = 0; o < n; o++) any |= (mem[o] = (grp[o] >> b) & 1) != 0;
        if (any && (rc = dem_set_group(ctx, b, mem.data()))) return rc;
    }
    // FixContactHistory::unpack_restart (fix_contact_history.cpp:608-640)
    std::vector<int> ti, tj;
    std::vector<double> hh;
    size_t q = 0;
    for (int o = 0; o < n; o++) {
        if (q + 2 > rec.size()) return fail(ctx, DEM_ERR_ARG, "dem_read_restart: bad history record");
        const size_t size = (size_t)rec[q];
        const int np = ubuf_to_int(rec[q + 1]);
        if (np < 0 || size != (size_t)(dnum + 1) * np + 2 || q + size > rec.size())
            return fail(ctx, DEM_ERR_ARG, "dem_read_restart: bad history record");
        size_t p = q + 2;
        for (int k = 0; k < np; k++) {
            ti.push_back(tag[o]);
            tj.push_back(ubuf_to_int(rec[p++]));
            for (int d = 0; d < dnum; d++) hh.push_back(rec[p++]);
        }
        q += size;
    }
    if (!ti.empty() && (rc = dem_set_history(ctx, (int64_t)ti.size(), ti.data(), tj.data(), hh.data()))) return rc;
    if (!whh.empty() && (rc = dem_set_wall_history(ctx, whh.data()))) return rc;
    for (int k = 0; k < h.nmesh; k++)
        if (!mr[k].tag.empty()
            && (rc = dem_set_mesh_contacts(ctx, k, (int64_t)mr[k].tag.size(), mr[k].tag.data(), mr[k].tri.data(),
                                           mr[k].slot.data(), mr[k].hist.data())))
            return rc;
    ctx->ntimestep = h.ntimestep;
    ctx->atimestep = h.atimestep;
    ctx->atime = h.atime;
    return DEM_OK;
}

extern "C" int dem_set_integrate_group(dem_ctx *ctx, int bit)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (bit < -1 || bit > 30) return fail(ctx, DEM_ERR_ARG, "dem_set_integrate_group: bit out of range");
    ctx->integ_bit = bit;
    return DEM_OK;
}

extern "C" int dem_add_move_linear(dem_ctx *ctx, int bit, const int *style, const double *value, int *id)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!style || !value || bit < 0 || bit > 30)
        return fail(ctx, DEM_ERR_ARG, "dem_add_move_linear: bad arguments");
    if (!ctx->moves.empty()) return fail(ctx, DEM_ERR_ARG, "dem_add_move_linear: one fix move on the device");
    if (ctx->n <= 0) return fail(ctx, DEM_ERR_ARG, "dem_add_move_linear: upload the atoms first");
    sync_enter(ctx);
    const size_t c = (size_t)ctx->cap;
    for (int b = 0; b < 2; b++)
        if (dev_alloc(ctx, &ctx->xo[b], c)) return DEM_ERR_HIP;
    // FixMove keeps each atom's position at its creation (xoriginal)
    HIPCHK(hipMemcpyAsync(ctx->xo[ctx->cur], sx(ctx), sizeof(double4) * ctx->n, hipMemcpyDeviceToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    dem_ctx::MoveOp m{};
    m.id = ctx->next_op_id++;
    m.bit = bit;
    m.style = (style[0] ? 1 : 0) | (style[1] ? 2 : 0) | (style[2] ? 4 : 0);
    for (int d = 0; d < 3; d++) m.val[d] = value[d];
    m.t0 = ctx->ntimestep;
    ctx->moves.push_back(m);
    if (id) *id = m.id;
    return DEM_OK;
}

extern "C" int dem_force_op_vector(dem_ctx *ctx, int id, double *out)
{
    if (!ctx || !out) return fail(ctx, DEM_ERR_ARG, "dem_force_op_vector: bad arguments");
    for (const auto &o : ctx->fops)
        if (o.id == id) {
            sync_enter(ctx);
            double t[4];
            HIPCHK(hipMemcpy(t, ctx->d_gtot + 4 * o.slot, sizeof(t), hipMemcpyDeviceToHost));
            for (int d = 0; d < 3; d++) out[d] = t[d];
            return DEM_OK;
        }
    return fail(ctx, DEM_ERR_ARG, "dem_force_op_vector: no such fix");
}

extern "C" int dem_cleanup(dem_ctx *ctx)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    ctx->atime += (double)(ctx->ntimestep - ctx->atimestep) * ctx->p.dt;
    ctx->atimestep = ctx->ntimestep;
    return DEM_OK;
}

extern "C" double dem_time(const dem_ctx *ctx)
{
    return ctx ? ctx->atime + (double)(ctx->ntimestep - ctx->atimestep) * ctx->p.dt : 0.0;
}

extern "C" int dem_set_instances(dem_ctx *ctx, int ninst, const int *inst)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (ninst < 1 || (ninst > 1 && !inst) || ninst > (1 << 20))
        return fail(ctx, DEM_ERR_ARG, "dem_set_instances: bad arguments");
    if (ctx->n == 0 || ctx->have_list)
        return fail(ctx, DEM_ERR_ARG, "dem_set_instances: call after dem_upload_state and before dem_setup");
    if (ninst > 1 && (!ctx->meshes.empty() || ctx->M.cohesion != CM_OFF))
        return fail(ctx, DEM_ERR_ARG, "dem_set_instances: mesh walls and bonds are not available in a batched run");
    sync_enter(ctx);
    std::vector<double4> hw(ctx->n);
    HIPCHK(hipMemcpy(hw.data(), sw(ctx), sizeof(double4) * ctx->n, hipMemcpyDeviceToHost));
    for (int i = 0; i < ctx->n; i++) {
        const int b = ninst > 1 ? inst[i] : 0;
        if (b < 0 || b >= ninst) return fail(ctx, DEM_ERR_ARG, "dem_set_instances: instance out of range");
        const int m = (int)hw[i].w;
        hw[i].w = (double)((m & 31) | (b << 5));
    }
    HIPCHK(hipMemcpy(sw(ctx), hw.data(), sizeof(double4) * ctx->n, hipMemcpyHostToDevice));
    ctx->ninst = ninst;
    ctx->inst_mat.assign((size_t)ninst * (M_ETAR + 1) * kT1 * kT1, 0.0);
    ctx->inst_set.assign(ninst, 0);
    int rc = upload_mat(ctx);
    if (rc) return rc;
    return setup_grid(ctx);
}

extern "C" int dem_set_instance_materials(dem_ctx *ctx, int inst, const double *Yeff, const double *Geff,
                                          const double *betaeff, const double *mu, const double *mur,
                                          const double *etar)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (inst < 0 || inst >= ctx->ninst || !Yeff || !Geff || !betaeff || !mu || !ctx->mat_set)
        return fail(ctx, DEM_ERR_ARG, "dem_set_instance_materials: bad arguments");
    sync_enter(ctx);
    const int w = ctx->ntypes + 1;
    const size_t T = (size_t)kT1 * kT1;
    double *dst = ctx->inst_mat.data() + (size_t)inst * (M_ETAR + 1) * T;
    const double *src[M_ETAR + 1] = {Yeff, Geff, betaeff, mu, mur, etar};
    for (int sl = 0; sl <= M_ETAR; sl++)
        for (int i = 0; i < kT1; i++)
            for (int j = 0; j < kT1; j++) {
                const double base = ctx->hmat[sl * T + i * kT1 + j];
                dst[sl * T + i * kT1 + j] = (src[sl] && i < w && j < w) ? src[sl][i * w + j] : base;
            }
    ctx->inst_set[inst] = 1;
    return upload_mat(ctx);
}

extern "C" int dem_observe_plane(dem_ctx *ctx, int plane)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (plane >= ctx->P.n || plane < -1) return fail(ctx, DEM_ERR_ARG, "dem_observe_plane: no such plane");
    if (ctx->n == 0) return fail(ctx, DEM_ERR_ARG, "dem_observe_plane: upload the atoms first");
    sync_enter(ctx);
    ctx->obs_plane = plane;
    if (plane >= 0) {
        if (dev_alloc(ctx, &ctx->pwf, ctx->cap)) return DEM_ERR_HIP;
        HIPCHK(hipMemset(ctx->pwf, 0, sizeof(double) * ctx->cap));
    }
    return DEM_OK;
}

extern "C" int dem_instance_observables(dem_ctx *ctx, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!out || !ctx->list_valid) return fail(ctx, DEM_ERR_ARG, "dem_instance_observables: call dem_setup first");
    sync_enter(ctx);
    const int B = ctx->ninst;
    if (B * 5 > ctx->obs_cap) {
        if (dev_alloc(ctx, &ctx->d_obs, (size_t)B * 5)) return DEM_ERR_HIP;
        ctx->obs_cap = B * 5;
    }
    launch_inst_obs(ctx->cstart, ctx->G.nkeys_local, B, sx(ctx), sv(ctx),
                    ctx->obs_plane >= 0 ? ctx->pwf : nullptr, ctx->d_obs, ctx->stream);
    KCHK();
    HIPCHK(hipMemcpyAsync(out, ctx->d_obs, sizeof(double) * B * 5, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

extern "C" int dem_instance_observables_noalloc(dem_ctx *ctx, double *out)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!out || !ctx->list_valid)
        return fail(ctx, DEM_ERR_ARG, "dem_instance_observables_noalloc: call dem_setup first");
    const int B = ctx->ninst;
    if (B > kObsScratchInst)
        return fail(ctx, DEM_ERR_ARG, "dem_instance_observables_noalloc: " + std::to_string(B)
                                          + " instances exceed the scratch limit of "
                                          + std::to_string(kObsScratchInst));
    if (!ctx->obs_scratch)
        return fail(ctx, DEM_ERR_HIP, "dem_instance_observables_noalloc: no device address for the scratch array");
    sync_enter(ctx);
    launch_inst_obs(ctx->cstart, ctx->G.nkeys_local, B, sx(ctx), sv(ctx),
                    ctx->obs_plane >= 0 ? ctx->pwf : nullptr, ctx->obs_scratch, ctx->stream);
    KCHK();
    HIPCHK(hipMemcpyAsync(out, ctx->obs_scratch, sizeof(double) * B * 5, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

extern "C" int dem_use_graph(dem_ctx *ctx, int on)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    ctx->use_graph = on != 0;
    return DEM_OK;
}

// Time the contact pass for every lane-group width and block size on the
// current state (set-up passes, shearupdate = 0) and keep the fastest.  The
// pair and wall history, list flags and forces are restored afterwards.
extern "C" int dem_autotune(dem_ctx *ctx, int reps, int *lanes, int *block, double *ms)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!ctx->list_valid || reps < 1) return fail(ctx, DEM_ERR_ARG, "dem_autotune: call dem_setup first, reps >= 1");
    sync_enter(ctx);
    hipStream_t s = ctx->stream;
    const int n = ctx->n, l = ctx->lcur, c = ctx->cur;
    int64_t E = 0;
    {
        int h = 0;
        HIPCHK(hipMemcpy(&h, ctx->nstart[l] + n, sizeof(int), hipMemcpyDeviceToHost));
        E = h;
    }
    const size_t hb = sizeof(double) * (size_t)E * ctx->M.dnum, jb = sizeof(uint32_t) * (size_t)E;
    const size_t wb = sizeof(double) * (size_t)n * ctx->P.n * ctx->M.wdnum, fb = sizeof(double4) * (size_t)n;
    // mesh contact slots: the mesh passes create, fill and clean them (and
    // hysteretic walls write their history even without a shear update)
    struct MeshSave { int *part, *np; double *hist; size_t pb, nb, hb; };
    std::vector<MeshSave> msave;
    size_t mtot = 0;
    const size_t dw = std::max(ctx->M.wdnum, 1);
    for (const auto &m : ctx->meshes) {
        MeshSave ms{nullptr, nullptr, nullptr, 0, 0, 0};
        if (m.have) {
            int h = 0;
            HIPCHK(hipMemcpy(&h, m.mstart[m.cur] + n, sizeof(int), hipMemcpyDeviceToHost));
            ms = MeshSave{m.mpart[m.cur], m.mnp[m.cur], m.mhist[m.cur], sizeof(int) * (size_t)h,
                          sizeof(int) * (size_t)n, sizeof(double) * (size_t)h * dw};
        }
        msave.push_back(ms);
        mtot += ms.pb + ms.nb + ms.hb + 48;
    }
    char *save = nullptr;
    const size_t tot = hb + jb + wb + 2 * fb + 64 + mtot;
    HIPCHK(hipMalloc((void **)&save, tot));
    char *sh = save, *sj = sh + hb, *sw = sj + jb, *sf = sw + wb, *st = sf + fb;
    auto copy = [&](void *d, const void *q, size_t b) {
        return b ? hipMemcpyAsync(d, q, b, hipMemcpyDeviceToDevice, s) : hipSuccess;
    };
    auto align16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
    auto mesh_copy = [&](bool back) {
        hipError_t e = hipSuccess;
        char *q = st + fb;
        for (const auto &ms : msave) {
            char *qp = q, *qn = qp + align16(ms.pb), *qh = qn + align16(ms.nb);
            q = qh + align16(ms.hb);
            if (back)
                e = (hipError_t)(e | copy(ms.part, qp, ms.pb) | copy(ms.np, qn, ms.nb) | copy(ms.hist, qh, ms.hb));
            else
                e = (hipError_t)(e | copy(qp, ms.part, ms.pb) | copy(qn, ms.np, ms.nb) | copy(qh, ms.hist, ms.hb));
        }
        return e;
    };
    auto restore = [&]() {
        return (copy(ctx->hist[l], sh, hb) | copy(ctx->nj[l], sj, jb) | copy(ctx->whist[c], sw, wb)
                | copy(ctx->f4, sf, fb) | copy(ctx->t4, st, fb) | mesh_copy(true));
    };
    if (copy(sh, ctx->hist[l], hb) | copy(sj, ctx->nj[l], jb) | copy(sw, ctx->whist[c], wb) | copy(sf, ctx->f4, fb)
        | copy(st, ctx->t4, fb) | mesh_copy(false)) {
        hipFree(save);
        return fail(ctx, DEM_ERR_HIP, "dem_autotune: copy failed");
    }
    const bool any_width = ctx->p.normal_model == 0 && ctx->p.wall_normal_model == 0 && ctx->M.cohesion == CM_OFF;
    const int widths[] = {1, 2, 4, 8, 16, -8, -16, -32, -64};   // negative: contact-queue tiles
    const int blocks[] = {64, 128, 256};
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const bool was_timing = ctx->timing;
    ctx->timing = false;
    double best = 1e300;
    int bl = 4, bb = kBlock;
    int rc = DEM_OK;
    // time the form the steps run: with the integrator fused when it will
    // be; such a pass writes the new state to the other buffers only, so
    // undoing its parity flip leaves the state as it was
    const int integ = fuse_possible(ctx) ? 3 : 0;
    auto pass = [&]() {
        const int p0 = ctx->par;
        const int r = force_pass(ctx, 0, integ);
        ctx->par = p0;
        return r;
    };
    for (int g : widths) {
        if (!any_width && g != 4) continue;
        for (int b : blocks) {
            ctx->tune_lanes = g;
            ctx->tune_block = b;
            if ((rc = pass())) break;
            hipEventRecord(e0, s);
            for (int r = 0; r < reps && rc == DEM_OK; r++) rc = pass();
            hipEventRecord(e1, s);
            if (rc) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(ctx, DEM_ERR_HIP, "dem_autotune: event failed");
                break;
            }
            float t = 0.f;
            hipEventElapsedTime(&t, e0, e1);
            if (restore() != hipSuccess) {
                rc = fail(ctx, DEM_ERR_HIP, "dem_autotune: restore failed");
                break;
            }
            if (t / reps < best) {
                best = t / reps;
                bl = g;
                bb = b;
            }
        }
        if (rc) break;
    }
    ctx->timing = was_timing;
    hipStreamSynchronize(s);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(save);
    if (rc) return rc;
    ctx->tune_lanes = bl;
    ctx->tune_block = bb;
    if (lanes) *lanes = bl;
    if (block) *block = bb;
    if (ms) *ms = best;
    return check_err_flag(ctx);
}

// ---------------------------------------------------------------------------
// Adjoint window: gradients of per-instance observables with respect to each
// instance's Yeff, Geff, betaeff and mu.

static int adjoint_requirements(dem_ctx *ctx)
{
    const dem_params &p = ctx->p;
    if (p.normal_model != 0 || p.wall_normal_model != 0 || p.rolling_model != 0 || p.cohesion_model != 0
        || p.limit_force || p.wall_limit_force || !p.tangential_damping || p.cdf != 1.0 || p.neigh_check
        || p.precision != DEM_FP64 || !ctx->meshes.empty() || ctx->M.freeze_on || !ctx->fops.empty())
        return fail(ctx, DEM_ERR_ARG,
                    "adjoint window: needs FP64 Hertz pairs and planes with tangential history (tangential "
                    "damping on, limitForce off), no rolling, bonds, meshes, freeze, setforce or aveforce, "
                    "contactDistanceFactor 1 and neigh_modify check no");
    return DEM_OK;
}

extern "C" int dem_adjoint_forward(dem_ctx *ctx, int nsteps, int checkpoint_every)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    int rc = adjoint_requirements(ctx);
    if (rc) return rc;
    if (nsteps < 1 || checkpoint_every < 1) return fail(ctx, DEM_ERR_ARG, "dem_adjoint_forward: nsteps, checkpoint_every >= 1");
    if (!ctx->list_valid || ctx->ntimestep != ctx->setup_ts)
        return fail(ctx, DEM_ERR_ARG, "dem_adjoint_forward: call right after dem_setup");
    sync_enter(ctx);
    for (char *c : ctx->ckpts) hipFree(c);
    ctx->ckpts.clear();
    ctx->ck_ts.clear();
    ctx->ck_ago.clear();
    {
        int h = 0;
        HIPCHK(hipMemcpy(&h, ctx->nstart[ctx->lcur] + ctx->n, sizeof(int), hipMemcpyDeviceToHost));
        ctx->adj_E = h;
    }
    if (dev_alloc(ctx, &ctx->adj_rev, (size_t)std::max<int64_t>(ctx->adj_E, 1))) return DEM_ERR_NOMEM;
    launch_adj_rev(ctx->nstart[ctx->lcur], ctx->nj[ctx->lcur], ctx->adj_rev, ctx->n, ctx->stream);
    KCHK();
    launch_fill_int(ctx->d_small + 2, 0, 1, ctx->stream);
    const size_t rb = state_record_bytes(ctx, true);
    ctx->window = true;
    const int nseg = (nsteps + checkpoint_every - 1) / checkpoint_every;
    for (int sgi = 0; sgi <= nseg && rc == DEM_OK; sgi++) {
        char *c = nullptr;
        if (hipMalloc((void **)&c, rb) != hipSuccess) {
            rc = fail(ctx, DEM_ERR_NOMEM, "dem_adjoint_forward: out of device memory for checkpoints");
            break;
        }
        ctx->ckpts.push_back(c);
        ctx->ck_ts.push_back(ctx->ntimestep);
        ctx->ck_ago.push_back(ctx->ago);
        if ((rc = copy_state(ctx, c, true, true))) break;
        if (sgi == nseg) break;
        const int len = std::min(checkpoint_every, nsteps - sgi * checkpoint_every);
        rc = dem_step(ctx, len);
    }
    ctx->window = false;
    if (rc) return rc;
    int moved = 0;
    HIPCHK(hipMemcpy(&moved, ctx->d_small + 2, sizeof(int), hipMemcpyDeviceToHost));
    if (moved)
        return fail(ctx, DEM_ERR_STATE,
                    "adjoint window: a particle moved more than half the skin; the fixed neighbour list would miss "
                    "pairs (larger skin or a shorter window)");
    ctx->adj_steps = nsteps;
    ctx->adj_every = checkpoint_every;
    return DEM_OK;
}

static int restore_ckpt(dem_ctx *ctx, int k)
{
    const int rc = copy_state(ctx, ctx->ckpts[k], true, false);
    ctx->ntimestep = ctx->ck_ts[k];
    ctx->ago = ctx->ck_ago[k];
    return rc;
}

extern "C" int dem_adjoint_backward(dem_ctx *ctx, const double *w_height, const double *w_wall, double *grad)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!grad || ctx->adj_steps < 1 || ctx->ckpts.empty())
        return fail(ctx, DEM_ERR_ARG, "dem_adjoint_backward: run dem_adjoint_forward first");
    if (w_wall && ctx->obs_plane < 0)
        return fail(ctx, DEM_ERR_ARG, "dem_adjoint_backward: a wall-force seed needs dem_observe_plane");
    sync_enter(ctx);
    hipStream_t s = ctx->stream;
    const int n = ctx->n, B = ctx->ninst, N = ctx->adj_steps, K = ctx->adj_every;
    const int nseg = (N + K - 1) / K;
    const size_t E = (size_t)ctx->adj_E, dn = ctx->M.dnum, pw = (size_t)n * ctx->P.n * 3;
    // adjoint buffers, double buffered; parameter sums; seeds
    char *buf = nullptr;
    auto r256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b4 = r256(n * sizeof(double4)), bh = r256(std::max<size_t>(E * dn, 1) * sizeof(double)),
                 bw = r256(std::max<size_t>(pw, 1) * sizeof(double)), bg = r256((size_t)n * 4 * sizeof(double)),
                 bs = r256((size_t)B * sizeof(double));
    const size_t total = 6 * b4 + 2 * bh + 2 * bw + bg + 2 * bs + r256((size_t)B * 4 * sizeof(double))
                         + r256((size_t)B * 5 * sizeof(double));
    HIPCHK(hipMalloc((void **)&buf, total));
    char *q = buf;
    auto take = [&](size_t b) {
        char *r = q;
        q += b;
        return r;
    };
    double4 *lx[2] = {(double4 *)take(b4), (double4 *)take(b4)};
    double4 *lv[2] = {(double4 *)take(b4), (double4 *)take(b4)};
    double4 *lw[2] = {(double4 *)take(b4), (double4 *)take(b4)};
    double *lh[2] = {(double *)take(bh), (double *)take(bh)};
    double *lhw[2] = {(double *)take(bw), (double *)take(bw)};
    double *gp = (double *)take(bg);
    double *d_wh = (double *)take(bs), *d_ww = (double *)take(bs);
    double *d_grad = (double *)take(r256((size_t)B * 4 * sizeof(double)));
    double *d_cnt = (double *)take(r256((size_t)B * 5 * sizeof(double)));
    int rc = DEM_OK;
    auto done = [&](int code) {
        hipStreamSynchronize(s);
        hipFree(buf);
        ctx->tape_pos = -1;
        return code;
    };
    // a failed call from here on frees the buffers before it returns
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        fail(ctx, DEM_ERR_HIP, std::string("dem_adjoint_backward: ") + what + ": " + hipGetErrorString(e));
        return false;
    };
    if (!hip_ok(hipMemsetAsync(buf, 0, total, s), "clear")) return done(DEM_ERR_HIP);
    if (w_height && !hip_ok(hipMemcpyAsync(d_wh, w_height, sizeof(double) * B, hipMemcpyHostToDevice, s), "seed"))
        return done(DEM_ERR_HIP);
    if (w_wall && !hip_ok(hipMemcpyAsync(d_ww, w_wall, sizeof(double) * B, hipMemcpyHostToDevice, s), "seed"))
        return done(DEM_ERR_HIP);
    // seed at the final state: the mean height's gradient
    if ((rc = restore_ckpt(ctx, nseg))) return done(rc);
    launch_inst_obs(ctx->cstart, ctx->G.nkeys_local, B, sx(ctx), sv(ctx), nullptr, d_cnt, s);
    launch_adj_seed(sw(ctx), w_height ? d_wh : nullptr, d_cnt, lx[0], n, s);
    if (!hip_ok(hipGetLastError(), "seed kernel")) return done(DEM_ERR_HIP);
    // tape for the longest segment
    const size_t tb = state_record_bytes(ctx, false);
    if (K > ctx->tape_cap) {
        if (ctx->tape) hipFree(ctx->tape);
        ctx->tape = nullptr;
        if (hipMalloc((void **)&ctx->tape, tb * K) != hipSuccess) {
            ctx->tape_cap = 0;
            return done(fail(ctx, DEM_ERR_NOMEM, "dem_adjoint_backward: out of device memory for the tape"));
        }
        ctx->tape_cap = K;
    }
    ctx->tape_mask.assign(K, 0);
    int cur = 0;
    AdjArgs a{};
    a.nstart = ctx->nstart[ctx->lcur];
    a.rev = ctx->adj_rev;
    a.wmask = ctx->wmask;
    a.mat = ctx->d_mat;
    a.gp = gp;
    a.obs_plane = ctx->obs_plane;
    a.n = n;
    a.dnum = (int)dn;
    a.wdnum = ctx->M.wdnum;
    a.dt = ctx->p.dt;
    a.M = ctx->M;
    a.P = ctx->P;
    auto step_back = [&](const StateView &v, int update, int mask_step, double ckick, const double *wseed) {
        a.xr = v.xr; a.vm = v.vm; a.wt = v.wt;
        a.nj = v.nj; a.hist = v.hist; a.whist = v.whist;
        a.lx = lx[cur]; a.lv = lv[cur]; a.lw = lw[cur]; a.lh = lh[cur]; a.lhw = lhw[cur];
        a.lx1 = lx[1 - cur]; a.lv1 = lv[1 - cur]; a.lw1 = lw[1 - cur]; a.lh1 = lh[1 - cur]; a.lhw1 = lhw[1 - cur];
        a.update = update;
        a.mask_step = mask_step;
        a.ckick = ckick;
        a.wseed = wseed;
        launch_adj_step(a, s);
        cur = 1 - cur;
    };
    for (int sgi = nseg - 1; sgi >= 0 && rc == DEM_OK; sgi--) {
        const int len = std::min(K, N - sgi * K);
        if ((rc = restore_ckpt(ctx, sgi))) break;
        ctx->window = true;
        ctx->tape_pos = 0;
        rc = dem_step(ctx, len);
        ctx->window = false;
        ctx->tape_pos = -1;
        if (rc) break;
        for (int t = len - 1; t >= 0; t--) {
            const int nstep = sgi * K + t + 1;
            const StateView v = state_view(ctx, ctx->tape + (size_t)t * tb, false);
            step_back(v, 1, ctx->tape_mask[t], nstep == N ? 0.5 * ctx->p.dt : ctx->p.dt,
                      (nstep == N && w_wall) ? d_ww : nullptr);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(ctx, DEM_ERR_HIP, std::string("adjoint step: ") + hipGetErrorString(e));
    }
    if (rc) return done(rc);
    // the set-up pass (shearupdate = 0): its forces kick the first half step
    {
        const StateView v = state_view(ctx, ctx->ckpts[0], true);
        step_back(v, 0, 0, 0.5 * ctx->p.dt, nullptr);
    }
    launch_adj_reduce(ctx->cstart, ctx->G.nkeys_local, B, gp, d_grad, s);
    if (!hip_ok(hipGetLastError(), "reduce kernel")
        || !hip_ok(hipMemcpyAsync(grad, d_grad, sizeof(double) * B * 4, hipMemcpyDeviceToHost, s), "copy"))
        return done(DEM_ERR_HIP);
    // leave the context at the window's final state
    if ((rc = restore_ckpt(ctx, nseg))) return done(rc);
    return done(check_err_flag(ctx));
}

extern "C" int dem_set_contact_shape(dem_ctx *ctx, int lanes, int block)
{
    if (!ctx) return fail(nullptr, DEM_ERR_ARG, "null context");
    if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == -8
          || lanes == -16 || lanes == -32 || lanes == -64)
        || !(block == 0 || block == 64 || block == 128 || block == 256))
        return fail(ctx, DEM_ERR_ARG,
                    "dem_set_contact_shape: lanes 1, 2, 4, 8, 16 (or -8, -16, -32: contact queue with tiles of that "
                    "many particles) and block 64, 128, 256 (0: default)");
    ctx->tune_lanes = lanes;
    ctx->tune_block = block;
    return DEM_OK;
}
